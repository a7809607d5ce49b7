// Particle passes of one space-charge kick over SoA FP64 particle rows.
//
// Reference: ocelot/cpbd/sc.py:208-251 (apply), :170-205 (el_field) and
// ocelot/cpbd/coord_transform.py:16-96.  The kick is four passes over the
// particles, each recomputing the frame transform from the six FP64 rows
// instead of storing intermediates (recomputing costs ALU, storing costs HBM):
//   A  mean momentum                       reads x', y', delta     (24 B/particle)
//   B  rest-frame extent, charge centroid  reads 6 rows + q        (56 B)
//   C  cell index + deposit                reads 6 rows + q        (56 B)
//   D  field gather + kick + write-back    reads and writes 6 rows (96 B)
// Single-block reduction kernels between the passes turn the per-block
// partials into the frame, cell sizes and grid offset, in a fixed order so
// results are reproducible run to run.
//
// The translation unit is compiled with -ffp-contract=off: every product and
// sum below rounds once, in the order numpy evaluates the same expression.
// Fused multiply-adds appear only where numpy's BLAS call uses them (the 3x3
// frame rotation, fma(x2,T2k, fma(x1,T1k, x0*T0k))).
#pragma once
#include "common.hpp"

namespace ocsc {

constexpr int kPT = 256;          // threads per workgroup, particle passes
constexpr int kMaxRedBlocks = 8192;  // partial-sum blocks of passes A and B
constexpr int kNPartB = 11;          // partials per block in pass B
constexpr int kNumXcd = 8;           // grid copies of the per-XCD deposit

struct Rows {
  const double* r;  // row-major (6, ld), row i at r + i*ld
  int64_t ld;
};

// coord_transform.py:57-96, one particle.  xp = common-time positions (m) and
// momenta (eV/c).
__device__ __forceinline__ void to_xp(double r0, double r1, double r2, double r3, double r4, double r5,
                                      const KickScalars& s, double* xp) {
  const double gamma = (s.betaref * r5 + 1) * s.gamref;
  const double beta = __builtin_sqrt(1 - inv_sq(gamma));
  const double t = (gamma * beta) / s.gbref;
  const double pz2pref = __builtin_sqrt(t * t - r1 * r1 - r3 * r3);
  double u0 = r1 / pz2pref, u1 = r3 / pz2pref;
  const double nrm = __builtin_sqrt(u0 * u0 + u1 * u1 + 1.0);
  u0 = u0 / nrm;
  u1 = u1 / nrm;
  const double u2 = 1.0 / nrm;
  xp[0] = r0 - u0 * beta * r4;
  xp[1] = r2 - u1 * beta * r4;
  xp[2] = -u2 * beta * r4;
  xp[3] = u0 * gamma * beta * kMeEV;
  xp[4] = u1 * gamma * beta * kMeEV;
  xp[5] = u2 * gamma * beta * kMeEV;
}

// the momentum part only (pass A)
__device__ __forceinline__ void to_p(double r1, double r3, double r5, const KickScalars& s, double* p) {
  const double gamma = (s.betaref * r5 + 1) * s.gamref;
  const double beta = __builtin_sqrt(1 - inv_sq(gamma));
  const double t = (gamma * beta) / s.gbref;
  const double pz2pref = __builtin_sqrt(t * t - r1 * r1 - r3 * r3);
  double u0 = r1 / pz2pref, u1 = r3 / pz2pref;
  const double nrm = __builtin_sqrt(u0 * u0 + u1 * u1 + 1.0);
  u0 = u0 / nrm;
  u1 = u1 / nrm;
  const double u2 = 1.0 / nrm;
  p[0] = u0 * gamma * beta * kMeEV;
  p[1] = u1 * gamma * beta * kMeEV;
  p[2] = u2 * gamma * beta * kMeEV;
}

// v^T T, numpy.dot(v.T, T) on three components
__device__ __forceinline__ void rot(const double* v, const double* T, double* o) {
#pragma unroll
  for (int k = 0; k < 3; ++k) o[k] = __builtin_fma(v[2], T[6 + k], __builtin_fma(v[1], T[3 + k], v[0] * T[k]));
}
// v^T T^T
__device__ __forceinline__ void rot_back(const double* v, const double* T, double* o) {
#pragma unroll
  for (int k = 0; k < 3; ++k)
    o[k] = __builtin_fma(v[2], T[3 * k + 2], __builtin_fma(v[1], T[3 * k + 1], v[0] * T[3 * k]));
}

// coord_transform.py:16-54 for one particle: positions xp[0:3], momenta p
__device__ __forceinline__ void to_ocelot(const double* xp, const double* p, const KickScalars& s, double* o) {
  const double s2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
  const double gamma = __builtin_sqrt(1 + s2 / kMeEV2);
  const double beta = __builtin_sqrt(1 - inv_sq(gamma));
  const double nrm = __builtin_sqrt(s2);
  const double u0 = p[0] / nrm, u1 = p[1] / nrm, u2 = p[2] / nrm;
  const double cdt = -xp[2] / (beta * u2);
  o[0] = xp[0] + beta * u0 * cdt;
  o[2] = xp[1] + beta * u1 * cdt;
  o[5] = (gamma / s.gamref - 1) / s.betaref;
  o[4] = cdt;
  o[1] = p[0] / s.pref;
  o[3] = p[1] / s.pref;
}

// fractional grid coordinates (sc.py:172-186): z scaled by gamma0, divided by
// the cell size, minus the offset
__device__ __forceinline__ void grid_coords(const double* xyz, const KickScalars& s, double* X) {
#pragma unroll
  for (int k = 0; k < 3; ++k) X[k] = xyz[k] / s.steps[k] - s.X_off[k];
}

// ---------------------------------------------------------------------------
// deterministic block reduction of NV doubles per thread (sum)
template <int NV>
__device__ __forceinline__ void block_sum(double* v, double* red) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int c = 0; c < NV; ++c) red[c * kPT + tid] = v[c];
  __syncthreads();
  for (int w = kPT / 2; w > 0; w >>= 1) {
    if (tid < w) {
#pragma unroll
      for (int c = 0; c < NV; ++c) red[c * kPT + tid] += red[c * kPT + tid + w];
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < NV; ++c) v[c] = red[c * kPT];
}

// pass A: per-block sums of the momenta (sc.py:224 numerator)
__global__ void __launch_bounds__(kPT) k_passA(Rows R, KickScalars* __restrict__ scp, double* __restrict__ part) {
  __shared__ double red[3 * kPT];
  const KickScalars s = *scp;
  double acc[3] = {0, 0, 0};
  const int64_t n = s.n;
  for (int64_t p = (int64_t)blockIdx.x * kPT + threadIdx.x; p < n; p += (int64_t)gridDim.x * kPT) {
    double m[3];
    to_p(R.r[R.ld + p], R.r[3 * R.ld + p], R.r[5 * R.ld + p], s, m);
    acc[0] += m[0];
    acc[1] += m[1];
    acc[2] += m[2];
  }
  block_sum<3>(acc, red);
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) part[blockIdx.x * 3 + c] = acc[c];
}

// frame from the mean momentum (sc.py:224-239)
__global__ void __launch_bounds__(kPT) k_reduceA(KickScalars* __restrict__ scp, const double* __restrict__ part,
                                                 int nblk) {
  __shared__ double red[3 * kPT];
  double acc[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nblk; b += kPT)
    for (int c = 0; c < 3; ++c) acc[c] += part[b * 3 + c];
  block_sum<3>(acc, red);
  if (threadIdx.x != 0) return;
  KickScalars& s = *scp;
  const double n = (double)s.n;
  double t3[3] = {acc[0] / n, acc[1] / n, acc[2] / n};
  const double Pav = __builtin_sqrt(__builtin_fma(t3[2], t3[2], __builtin_fma(t3[1], t3[1], t3[0] * t3[0])));
  for (int k = 0; k < 3; ++k) t3[k] = t3[k] / Pav;
  // t1 = ey x t3 = (t3z, 0, -t3x)
  double t1[3] = {t3[2], 0.0, -t3[0]};
  const double n1 = __builtin_sqrt(__builtin_fma(t1[2], t1[2], t1[0] * t1[0]));
  for (int k = 0; k < 3; ++k) t1[k] = t1[k] / n1;
  const double t2[3] = {t3[1] * t1[2] - t3[2] * t1[1], t3[2] * t1[0] - t3[0] * t1[2], t3[0] * t1[1] - t3[1] * t1[0]};
  for (int j = 0; j < 3; ++j) {
    s.T[3 * j + 0] = t1[j];
    s.T[3 * j + 1] = t2[j];
    s.T[3 * j + 2] = t3[j];
  }
  s.Pav = Pav;
  const double q = Pav / kMeEV;
  s.gamma0 = __builtin_sqrt(q * q + 1);
  s.beta0 = __builtin_sqrt(1 - inv_sq(s.gamma0));
  s.kperp = s.cdT * (1 - s.beta0 * s.beta0);
  if (!(Pav > 0) || !__builtin_isfinite(Pav)) s.err |= kErrNonFinite;
}

// pass B: rest-frame min/max, total charge, charge-weighted sums
__global__ void __launch_bounds__(kPT) k_passB(Rows R, const double* __restrict__ q, KickScalars* __restrict__ scp,
                                               double* __restrict__ part) {
  __shared__ double red[kPT];
  const KickScalars s = *scp;
  const double inf = __builtin_inf();
  double mn[3] = {inf, inf, inf}, mx[3] = {-inf, -inf, -inf};
  double sq = 0, saq = 0, sqx[3] = {0, 0, 0};
  int bad = 0;
  const int64_t n = s.n;
  for (int64_t p = (int64_t)blockIdx.x * kPT + threadIdx.x; p < n; p += (int64_t)gridDim.x * kPT) {
    double xp[6], xyz[3];
    to_xp(R.r[p], R.r[R.ld + p], R.r[2 * R.ld + p], R.r[3 * R.ld + p], R.r[4 * R.ld + p], R.r[5 * R.ld + p], s, xp);
    rot(xp, s.T, xyz);
    xyz[2] = xyz[2] * s.gamma0;
    const double qq = q[p];
    bad |= !(__builtin_isfinite(xyz[0]) && __builtin_isfinite(xyz[1]) && __builtin_isfinite(xyz[2]) &&
             __builtin_isfinite(qq));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      mn[k] = xyz[k] < mn[k] ? xyz[k] : mn[k];
      mx[k] = xyz[k] > mx[k] ? xyz[k] : mx[k];
      sqx[k] += qq * xyz[k];
    }
    sq += qq;
    saq += __builtin_fabs(qq);
  }
  if (bad) atomicOr(&scp->err, (int)kErrNonFinite);
  // min / max / sums reduced one quantity at a time through LDS
  double vals[kNPartB] = {mn[0], mn[1], mn[2], mx[0], mx[1], mx[2], sq, saq, sqx[0], sqx[1], sqx[2]};
  const int tid = threadIdx.x;
  for (int c = 0; c < kNPartB; ++c) {
    red[tid] = vals[c];
    __syncthreads();
    for (int w = kPT / 2; w > 0; w >>= 1) {
      if (tid < w) {
        const double a = red[tid], b = red[tid + w];
        red[tid] = c < 3 ? (b < a ? b : a) : (c < 6 ? (b > a ? b : a) : a + b);
      }
      __syncthreads();
    }
    if (tid == 0) part[blockIdx.x * kNPartB + c] = red[0];
    __syncthreads();
  }
}

// cell sizes, grid offset, charge scale and kernel-cache decision
// (sc.py:172-186)
__global__ void __launch_bounds__(kPT) k_reduceB(KickScalars* __restrict__ scp, const double* __restrict__ part,
                                                 int nblk, int prec) {
  __shared__ double red[kPT];
  const int tid = threadIdx.x;
  double out[kNPartB];
  for (int c = 0; c < kNPartB; ++c) {
    double a = c < 3 ? __builtin_inf() : (c < 6 ? -__builtin_inf() : 0.0);
    for (int b = tid; b < nblk; b += kPT) {
      const double v = part[b * kNPartB + c];
      a = c < 3 ? (v < a ? v : a) : (c < 6 ? (v > a ? v : a) : a + v);
    }
    red[tid] = a;
    __syncthreads();
    for (int w = kPT / 2; w > 0; w >>= 1) {
      if (tid < w) {
        const double x = red[tid], y = red[tid + w];
        red[tid] = c < 3 ? (y < x ? y : x) : (c < 6 ? (y > x ? y : x) : x + y);
      }
      __syncthreads();
    }
    out[c] = red[0];
    __syncthreads();
  }
  if (tid != 0) return;
  KickScalars& s = *scp;
  int err = s.err;
  for (int k = 0; k < 3; ++k) {
    s.xmin[k] = out[k];
    s.xmax[k] = out[3 + k];
    s.sumqx[k] = out[8 + k];
  }
  s.sumq = out[6];
  s.sumabsq = out[7];
  bool same = s.key_prec == prec;
  for (int k = 0; k < 3; ++k) {
    double XX = s.xmax[k] - s.xmin[k];
    if (s.random_mesh) XX = XX * s.mesh_scale;
    const double h = XX / (double)(s.nmesh[k] - 3);
    s.steps[k] = h;
    if (!(h > 0) || !__builtin_isfinite(h)) err |= kErrSteps;
    const double Xmin = s.xmin[k] / h;
    const double Xmid = (s.sumqx[k] / h) / s.sumq;
    double off = __builtin_floor(Xmin - Xmid) + Xmid;
    if (s.random_mesh) off = off + s.mesh_shift;
    s.X_off[k] = off;
    if (!__builtin_isfinite(off)) err |= kErrZeroCharge;
    same = same && s.key_steps[k] == h && s.key_nmesh[k] == s.nmesh[k];
  }
  if (s.sumq == 0) err |= kErrZeroCharge;
  int e2 = 0;
  const double fr = __builtin_frexp(s.sumabsq, &e2);
  (void)fr;
  s.qscale = __builtin_ldexp(1.0, 62 - e2);
  s.inv_qscale = __builtin_ldexp(1.0, e2 - 62);
  s.rebuild = same ? 0 : 1;
  if (!same) {
    for (int k = 0; k < 3; ++k) {
      s.key_steps[k] = s.steps[k];
      s.key_nmesh[k] = s.nmesh[k];
    }
    s.key_prec = prec;
  }
  if (err) s.rebuild = 1;
  s.err = err;
}

// pass C: nearest-grid-point deposit as exact fixed-point sums (sc.py:187-193).
// NWORD = 1: one int64 word, q*qscale rounded (relative error <= N*2^-63);
// NWORD = 2: a second word carries the rounding residual scaled by 2^30, so
// the cell sum is exact to double rounding.  Integer addition is associative,
// so the grid is identical whatever order the atomics land in.
// XCD (accelerator complex die) executing this wave; gfx94x/gfx950 expose it
// as hardware register XCC_ID (id 20), bits [3:0]
__device__ __forceinline__ int xcc_id() {
  return (int)(__builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11)) & (kNumXcd - 1));
}

// one fixed-point atomic add; XCD mode adds into the copy of the grid owned
// by the executing XCD with workgroup-scope atomics, which complete in that
// XCD's L2 instead of at the memory side
template <bool XCD>
__device__ __forceinline__ void grid_add(unsigned long long* g, int64_t idx, long long v) {
  if constexpr (XCD)
    __hip_atomic_fetch_add(g + idx, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else
    atomicAdd(g + idx, (unsigned long long)v);
}

template <int NWORD, bool XCD = false>
__global__ void __launch_bounds__(kPT) k_deposit(Rows R, const double* __restrict__ q, KickScalars* __restrict__ scp,
                                                 unsigned long long* __restrict__ grid) {
  const KickScalars& s0 = *scp;
  if (s0.err) return;
  const KickScalars s = s0;
  const int nx = s.nmesh[0], ny = s.nmesh[1], nz = s.nmesh[2];
  const int64_t ncell = (int64_t)nx * ny * nz;
  const int64_t p = (int64_t)blockIdx.x * kPT + threadIdx.x;
  if (p >= s.n) return;
  double xp[6], xyz[3], X[3];
  to_xp(R.r[p], R.r[R.ld + p], R.r[2 * R.ld + p], R.r[3 * R.ld + p], R.r[4 * R.ld + p], R.r[5 * R.ld + p], s, xp);
  rot(xp, s.T, xyz);
  xyz[2] = xyz[2] * s.gamma0;
  grid_coords(xyz, s, X);
  const double f0 = __builtin_floor(X[0]) + 1, f1 = __builtin_floor(X[1]) + 1, f2 = __builtin_floor(X[2]) + 1;
  if (!(f0 >= 0 && f0 < nx && f1 >= 0 && f1 < ny && f2 >= 0 && f2 < nz)) {
    atomicOr(&scp->err, (int)kErrCell);
    return;
  }
  const int64_t idx = ((int64_t)f0 * ny + (int64_t)f1) * nz + (int64_t)f2;
  const double v = q[p] * s.qscale;  // exact: qscale is a power of two
  const double hi = __builtin_rint(v);
  unsigned long long* g = grid;
  if constexpr (XCD) g += (int64_t)xcc_id() * NWORD * ncell;
  grid_add<XCD>(g, idx, (long long)hi);
  if constexpr (NWORD == 2) {
    const double lo = __builtin_rint((v - hi) * 0x1p30);
    grid_add<XCD>(g, ncell + idx, (long long)lo);
  }
}

// Trilinear interpolation at fractional grid coordinate c of the staggered
// difference field E_AX = (phi[i] - phi[i+1]) / h along axis AX (zero on the
// last plane), with scipy.ndimage.map_coordinates(order=1, mode='constant')
// edge semantics: zero if any coordinate lies outside [0, n-1].
template <typename F, int AX>
__device__ __forceinline__ F gather_diff(const F* __restrict__ phi, int nx, int ny, int nz, const double* c, F h) {
  const int n[3] = {nx, ny, nz};
  if (!(c[0] >= 0 && c[0] <= nx - 1 && c[1] >= 0 && c[1] <= ny - 1 && c[2] >= 0 && c[2] <= nz - 1)) return (F)0;
  int i0[3];
  F t[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int f = (int)__builtin_floor(c[k]);
    f = f > n[k] - 2 ? n[k] - 2 : f;
    i0[k] = f;
    t[k] = (F)(c[k] - (double)f);
  }
  const int64_t sx = (int64_t)ny * nz, sy = nz;
  const int64_t st = AX == 0 ? sx : (AX == 1 ? sy : 1);
  const int nax = n[AX];
  F acc = 0;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int ii = i0[0] + a, jj = i0[1] + b, kk = i0[2] + e;
        const int iax = AX == 0 ? ii : (AX == 1 ? jj : kk);
        F ev = 0;
        if (iax < nax - 1) {
          const int64_t o = ii * sx + jj * sy + kk;
          if constexpr (sizeof(F) == 8)
            ev = (phi[o] - phi[o + st]) / h;
          else
            ev = (phi[o] - phi[o + st]) * h;  // h holds 1/step in FP32
        }
        const F w = (a ? t[0] : 1 - t[0]) * (b ? t[1] : 1 - t[1]) * (e ? t[2] : 1 - t[2]);
        acc += w * ev;
      }
    }
  }
  return acc;
}

// Mixed form: trilinear interpolation at c of a precomputed FP32 field
// component E_AX (spectrally differenced, so no cancellation here), with the
// same zero-plane and edge rules as gather_diff.
template <int AX>
__device__ __forceinline__ float gather_grid(const float* __restrict__ E, int nx, int ny, int nz, const double* c) {
  const int n[3] = {nx, ny, nz};
  if (!(c[0] >= 0 && c[0] <= nx - 1 && c[1] >= 0 && c[1] <= ny - 1 && c[2] >= 0 && c[2] <= nz - 1)) return 0.f;
  int i0[3];
  float t[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int f = (int)__builtin_floor(c[k]);
    f = f > n[k] - 2 ? n[k] - 2 : f;
    i0[k] = f;
    t[k] = (float)(c[k] - (double)f);
  }
  const int64_t sx = (int64_t)ny * nz, sy = nz;
  const int nax = n[AX];
  const int64_t o0 = i0[0] * sx + i0[1] * sy + i0[2];
  float v[8];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int iax = (AX == 0 ? i0[0] + a : (AX == 1 ? i0[1] + b : i0[2] + e));
        v[a * 4 + b * 2 + e] = iax < nax - 1 ? E[o0 + a * sx + b * sy + e] : 0.f;
      }
  float acc = 0.f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const float w = (a ? t[0] : 1 - t[0]) * (b ? t[1] : 1 - t[1]) * (e ? t[2] : 1 - t[2]);
        acc += w * v[a * 4 + b * 2 + e];
      }
  return acc;
}

// field at a particle from the mesh result: FP64 potential (differenced
// here, as the reference does) or, in the mixed form, three FP32 field grids
// stored one after the other
template <typename F>
__device__ __forceinline__ void gather_field(const F* __restrict__ phi, const KickScalars& s, const double* X,
                                             double* E) {
  if constexpr (sizeof(F) == 4) {
    const int nx = s.nmesh[0], ny = s.nmesh[1], nz = s.nmesh[2];
    const int64_t nc = (int64_t)nx * ny * nz;
    const double cx[3] = {X[0], X[1] + 0.5, X[2] + 0.5};
    const double cy[3] = {X[0] + 0.5, X[1], X[2] + 0.5};
    const double cz[3] = {X[0] + 0.5, X[1] + 0.5, X[2]};
    E[0] = (double)gather_grid<0>(phi, nx, ny, nz, cx) * s.gamma0;
    E[1] = (double)gather_grid<1>(phi + nc, nx, ny, nz, cy) * s.gamma0;
    E[2] = (double)gather_grid<2>(phi + 2 * nc, nx, ny, nz, cz);
    return;
  }
  const int nx = s.nmesh[0], ny = s.nmesh[1], nz = s.nmesh[2];
  F h[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) h[k] = sizeof(F) == 8 ? (F)s.steps[k] : (F)(1.0 / s.steps[k]);
  const double cx[3] = {X[0], X[1] + 0.5, X[2] + 0.5};
  const double cy[3] = {X[0] + 0.5, X[1], X[2] + 0.5};
  const double cz[3] = {X[0] + 0.5, X[1] + 0.5, X[2]};
  E[0] = (double)gather_diff<F, 0>(phi, nx, ny, nz, cx, h[0]) * s.gamma0;
  E[1] = (double)gather_diff<F, 1>(phi, nx, ny, nz, cy, h[1]) * s.gamma0;
  E[2] = (double)gather_diff<F, 2>(phi, nx, ny, nz, cz, h[2]);
}

// pass D: gather, kick in the rotated frame, rotate back, write all six rows
// in place (sc.py:241-251, coord_transform.py:16-54)
template <typename F>
__global__ void __launch_bounds__(kPT) k_kick(double* __restrict__ r, int64_t ld, const F* __restrict__ phi,
                                              const KickScalars* __restrict__ scp) {
  const KickScalars& s0 = *scp;
  if (s0.err) return;
  const KickScalars s = s0;
  const int64_t p = (int64_t)blockIdx.x * kPT + threadIdx.x;
  if (p >= s.n) return;
  double xp[6], xyz[3], X[3], pr[3], E[3], pb[3], o[6];
  to_xp(r[p], r[ld + p], r[2 * ld + p], r[3 * ld + p], r[4 * ld + p], r[5 * ld + p], s, xp);
  rot(xp, s.T, xyz);
  xyz[2] = xyz[2] * s.gamma0;
  grid_coords(xyz, s, X);
  rot(xp + 3, s.T, pr);
  gather_field<F>(phi, s, X, E);
  pr[0] = pr[0] + s.kperp * E[0];
  pr[1] = pr[1] + s.kperp * E[1];
  pr[2] = pr[2] + s.cdT * E[2];
  rot_back(pr, s.T, pb);
  to_ocelot(xp, pb, s, o);
#pragma unroll
  for (int i = 0; i < 6; ++i) r[i * ld + p] = o[i];
}

// ---------------------------------------------------------------------------
// el_field entry (sc.py:170-205): rest-frame positions given directly as
// three rows (x, y, z before the gamma scaling of z).

__global__ void __launch_bounds__(kPT) k_passB_xyz(Rows R, const double* __restrict__ q,
                                                   KickScalars* __restrict__ scp, double* __restrict__ part) {
  __shared__ double red[kPT];
  const KickScalars s = *scp;
  const double inf = __builtin_inf();
  double mn[3] = {inf, inf, inf}, mx[3] = {-inf, -inf, -inf};
  double sq = 0, saq = 0, sqx[3] = {0, 0, 0};
  int bad = 0;
  for (int64_t p = (int64_t)blockIdx.x * kPT + threadIdx.x; p < s.n; p += (int64_t)gridDim.x * kPT) {
    double xyz[3] = {R.r[p], R.r[R.ld + p], R.r[2 * R.ld + p] * s.gamma0};
    const double qq = q[p];
    bad |= !(__builtin_isfinite(xyz[0]) && __builtin_isfinite(xyz[1]) && __builtin_isfinite(xyz[2]) &&
             __builtin_isfinite(qq));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      mn[k] = xyz[k] < mn[k] ? xyz[k] : mn[k];
      mx[k] = xyz[k] > mx[k] ? xyz[k] : mx[k];
      sqx[k] += qq * xyz[k];
    }
    sq += qq;
    saq += __builtin_fabs(qq);
  }
  if (bad) atomicOr(&scp->err, (int)kErrNonFinite);
  double vals[kNPartB] = {mn[0], mn[1], mn[2], mx[0], mx[1], mx[2], sq, saq, sqx[0], sqx[1], sqx[2]};
  const int tid = threadIdx.x;
  for (int c = 0; c < kNPartB; ++c) {
    red[tid] = vals[c];
    __syncthreads();
    for (int w = kPT / 2; w > 0; w >>= 1) {
      if (tid < w) {
        const double a = red[tid], b = red[tid + w];
        red[tid] = c < 3 ? (b < a ? b : a) : (c < 6 ? (b > a ? b : a) : a + b);
      }
      __syncthreads();
    }
    if (tid == 0) part[blockIdx.x * kNPartB + c] = red[0];
    __syncthreads();
  }
}

template <int NWORD>
__global__ void __launch_bounds__(kPT) k_deposit_xyz(Rows R, const double* __restrict__ q,
                                                     KickScalars* __restrict__ scp,
                                                     unsigned long long* __restrict__ grid) {
  const KickScalars& s0 = *scp;
  if (s0.err) return;
  const KickScalars s = s0;
  const int nx = s.nmesh[0], ny = s.nmesh[1], nz = s.nmesh[2];
  const int64_t ncell = (int64_t)nx * ny * nz;
  const int64_t p = (int64_t)blockIdx.x * kPT + threadIdx.x;
  if (p >= s.n) return;
  double xyz[3] = {R.r[p], R.r[R.ld + p], R.r[2 * R.ld + p] * s.gamma0}, X[3];
  grid_coords(xyz, s, X);
  const double f0 = __builtin_floor(X[0]) + 1, f1 = __builtin_floor(X[1]) + 1, f2 = __builtin_floor(X[2]) + 1;
  if (!(f0 >= 0 && f0 < nx && f1 >= 0 && f1 < ny && f2 >= 0 && f2 < nz)) {
    atomicOr(&scp->err, (int)kErrCell);
    return;
  }
  const int64_t idx = ((int64_t)f0 * ny + (int64_t)f1) * nz + (int64_t)f2;
  const double v = q[p] * s.qscale;
  const double hi = __builtin_rint(v);
  atomicAdd(&grid[idx], (unsigned long long)(long long)hi);
  if constexpr (NWORD == 2) {
    const double lo = __builtin_rint((v - hi) * 0x1p30);
    atomicAdd(&grid[ncell + idx], (unsigned long long)(long long)lo);
  }
}

template <typename F>
__global__ void __launch_bounds__(kPT) k_gather_xyz(Rows R, const F* __restrict__ phi,
                                                    const KickScalars* __restrict__ scp, double* __restrict__ E,
                                                    int64_t lde, int32_t* __restrict__ cells) {
  const KickScalars& s0 = *scp;
  if (s0.err) return;
  const KickScalars s = s0;
  const int64_t p = (int64_t)blockIdx.x * kPT + threadIdx.x;
  if (p >= s.n) return;
  double xyz[3] = {R.r[p], R.r[R.ld + p], R.r[2 * R.ld + p] * s.gamma0}, X[3], e[3];
  grid_coords(xyz, s, X);
  gather_field<F>(phi, s, X, e);
  E[p] = e[0];
  E[lde + p] = e[1];
  E[2 * lde + p] = e[2];
  if (cells) {
    const int nx = s.nmesh[0], ny = s.nmesh[1], nz = s.nmesh[2];
    const int f0 = (int)__builtin_floor(X[0]) + 1, f1 = (int)__builtin_floor(X[1]) + 1,
              f2 = (int)__builtin_floor(X[2]) + 1;
    cells[p] = (f0 * ny + f1) * nz + f2;
    (void)nx;
  }
}

// zero an int64 buffer (graph-safe replacement for a memset)
__global__ void k_zero64(unsigned long long* __restrict__ g, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    g[i] = 0ull;
}

}  // namespace ocsc
