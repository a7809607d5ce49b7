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
constexpr int kMaxRedBlocks = 2048;  // partial-sum blocks of passes A and B
constexpr int kFoldBlocks = 128;     // first-level blocks folding the kick's per-block momentum sums
constexpr int kNPartB = 11;          // partials per block in pass B
constexpr int kNumXcd = 8;           // grid copies of the per-XCD deposit

struct Rows {
  const double* r;  // row-major (6, ld), row i at r + i*ld
  int64_t ld;
};

// first-order drift x += L x', y += L y', tau += r56 delta
struct DriftArgs {
  double L, r56;
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

// Particle rows are streamed once per pass: non-temporal loads and stores
// keep them from evicting the field and charge grids from L2 (measured: the
// tracking step is 5% faster than with ordinary accesses).
__device__ __forceinline__ double ldnt(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void stnt(double v, double* p) { __builtin_nontemporal_store(v, p); }

// Fast variants (mixed passes B and C, both forms' pass D, sort keys), from the
// same kinematics with the common factors cancelled.  With t = gamma beta /
// (gamma_ref beta_ref) and pz = sqrt(t^2 - x'^2 - y'^2), the direction
// (a0, a1, 1)/|.| of coord_transform.py:83-91 is (x', y', pz)/t, so the
// momentum is (x', y', pz) p_ref and the common-time shift beta u tau is
// (x', y', pz) tau gamma_ref beta_ref / gamma: one square root and one
// reciprocal per particle instead of three of each.  Going back
// (coord_transform.py:31-53), beta u0 cdt = -z p0/p2 and cdt = -z m gamma/p2.
// Results differ from the exact sequence in the last bits only.
__device__ __forceinline__ void to_xp_fast(double r0, double r1, double r2, double r3, double r4, double r5,
                                           const KickScalars& s, double* xp) {
  const double gamma = (s.betaref * r5 + 1) * s.gamref;
  const double t2 = (gamma * gamma - 1) * (s.inv_gbref * s.inv_gbref);
  const double pz = __builtin_sqrt(t2 - r1 * r1 - r3 * r3);
  const double w = r4 * (s.gbref * rcp_nr(gamma));
  xp[0] = r0 - r1 * w;
  xp[1] = r2 - r3 * w;
  xp[2] = -(pz * w);
  xp[3] = r1 * s.pref;
  xp[4] = r3 * s.pref;
  xp[5] = pz * s.pref;
}

// coord_transform.py:16-54 for one particle (positions xp[0:3], momenta p),
// in the cancelled form above
__device__ __forceinline__ void to_ocelot_fast(const double* xp, const double* p, const KickScalars& s, double* o) {
  const double s2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
  const double gamma = __builtin_sqrt(1 + s2 * (1.0 / kMeEV2));
  const double a = -xp[2] * rcp_nr(p[2]);
  o[0] = xp[0] + p[0] * a;
  o[2] = xp[1] + p[1] * a;
  o[4] = a * (kMeEV * gamma);
  o[5] = (gamma * s.inv_gamref - 1) * s.inv_betaref;
  o[1] = p[0] * s.inv_pref;
  o[3] = p[1] * s.inv_pref;
}

__device__ __forceinline__ void grid_coords_fast(const double* xyz, const KickScalars& s, double* X) {
#pragma unroll
  for (int k = 0; k < 3; ++k) X[k] = xyz[k] * s.inv_steps[k] - s.X_off[k];
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

// fractional grid coordinates (sc.py:172-186): z scaled by gamma0, divided by
// the cell size, minus the offset
__device__ __forceinline__ void grid_coords(const double* xyz, const KickScalars& s, double* X) {
#pragma unroll
  for (int k = 0; k < 3; ++k) X[k] = xyz[k] / s.steps[k] - s.X_off[k];
}

// ---------------------------------------------------------------------------
// Deterministic block reductions: wave shuffles, then the waves' results in
// a fixed order.  KIND 0: every value is a sum; KIND 1 (pass B layout):
// values 0-2 minima, 3-5 maxima, the rest sums.
template <int KIND>
__device__ __forceinline__ double red_op(int c, double a, double b) {
  if (KIND == 1 && c < 3) return b < a ? b : a;
  if (KIND == 1 && c < 6) return b > a ? b : a;
  return a + b;
}

// result valid in thread 0; lds holds (blockDim/64) * NV doubles
template <int NV, int KIND>
__device__ __forceinline__ void block_reduce(double* v, double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int c = 0; c < NV; ++c) v[c] = red_op<KIND>(c, v[c], __shfl_down(v[c], off, 64));
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < NV; ++c) lds[w * NV + c] = v[c];
  __syncthreads();
  // the waves' results combined by the same shuffle tree in wave 0
  if (w == 0) {
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const double id = (KIND == 1 && c < 3) ? __builtin_inf() : ((KIND == 1 && c < 6) ? -__builtin_inf() : 0.0);
      v[c] = lane < nw ? lds[lane * NV + c] : id;
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1)
#pragma unroll
      for (int c = 0; c < NV; ++c) v[c] = red_op<KIND>(c, v[c], __shfl_down(v[c], off, 64));
  }
}

// partials are stored transposed, part[c * nblk + b], so the final reduction
// reads them coalesced
template <int NV, int KIND>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ part, int nblk, double* out, double* lds) {
  double v[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) v[c] = KIND == 1 && c < 3 ? __builtin_inf() : (KIND == 1 && c < 6 ? -__builtin_inf() : 0.0);
  // U partials per thread are loaded before any is combined, so U loads are
  // in flight instead of one; each thread still combines its partials in
  // increasing b (out-of-range slots are skipped), so the tree is unchanged
  constexpr int U = NV > 8 ? 2 : (NV > 4 ? 4 : 8);
  for (int b0 = threadIdx.x; b0 < nblk; b0 += U * blockDim.x) {
    double t[U][NV];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + u * blockDim.x;
#pragma unroll
      for (int c = 0; c < NV; ++c) t[u][c] = b < nblk ? part[c * nblk + b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (b0 + u * (int)blockDim.x < nblk)
#pragma unroll
        for (int c = 0; c < NV; ++c) v[c] = red_op<KIND>(c, v[c], t[u][c]);
  }
  block_reduce<NV, KIND>(v, lds);
#pragma unroll
  for (int c = 0; c < NV; ++c) out[c] = v[c];
}

// First level of a two-level sum over many block partials: block f of the
// grid adds its contiguous chunk of partials (in increasing order per thread,
// then the fixed block tree), so a single-block reduction only reads gridDim
// values per component.  Deterministic for a given (nblk, gridDim).
template <int NV>
__global__ void __launch_bounds__(256) k_fold(const double* __restrict__ part, int nblk, double* __restrict__ out) {
  __shared__ double lds[4 * NV];
  const int nf = gridDim.x, chunk = (nblk + nf - 1) / nf;
  const int b0 = blockIdx.x * chunk, b1 = b0 + chunk < nblk ? b0 + chunk : nblk;
  double v[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) v[c] = 0;
  for (int b = b0 + threadIdx.x; b < b1; b += 256)
#pragma unroll
    for (int c = 0; c < NV; ++c) v[c] += part[c * nblk + b];
  block_reduce<NV, 0>(v, lds);
  if (threadIdx.x == 0)
    for (int c = 0; c < NV; ++c) out[c * nf + blockIdx.x] = v[c];
}

// ---------------------------------------------------------------------------
// Permutations between the caller's particle order and the cell order.  A
// permutation of the six SoA rows and the charge reads (or writes) seven
// 8-byte values per particle at random places, each costing a whole memory
// segment.  Staged through one 64-byte record per particle (rows, charge,
// pad: one aligned segment), the random side moves one segment per particle
// and every other access is sequential.
struct alignas(64) PRec {
  double v[8];  // x, x', y, y', tau, delta, q, 0
};

__device__ __forceinline__ void store_rec(PRec* __restrict__ dst, const double* v) {
  double2* d = (double2*)dst->v;
#pragma unroll
  for (int k = 0; k < 4; ++k) d[k] = double2{v[2 * k], v[2 * k + 1]};
}

__device__ __forceinline__ void load_rec(const PRec* __restrict__ src, double* v) {
  const double2* s = (const double2*)src->v;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double2 a = s[k];
    v[2 * k] = a.x;
    v[2 * k + 1] = a.y;
  }
}

// pass A: per-block sums of the momenta (sc.py:224 numerator)
__device__ __forceinline__ void passA_body(Rows R, KickScalars* __restrict__ scp, double* __restrict__ part,
                                           const int bx, const int nbx) {
  __shared__ double lds[(kPT / 64) * 3];
  const KickScalars s = *scp;
  double acc[3] = {0, 0, 0};
  const int64_t n = s.n;
  for (int64_t p = (int64_t)bx * kPT + threadIdx.x; p < n; p += (int64_t)nbx * kPT) {
    double m[3];
    to_p(R.r[R.ld + p], R.r[3 * R.ld + p], R.r[5 * R.ld + p], s, m);
    acc[0] += m[0];
    acc[1] += m[1];
    acc[2] += m[2];
  }
  block_reduce<3, 0>(acc, lds);
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) part[c * nbx + bx] = acc[c];
}

__global__ void __launch_bounds__(kPT) k_passA(Rows R, KickScalars* __restrict__ scp, double* __restrict__ part) {
  passA_body(R, scp, part, blockIdx.x, gridDim.x);
}

// frame from the mean momentum (sc.py:224-239)
__device__ __forceinline__ void reduceA_body(KickScalars* __restrict__ scp, const double* __restrict__ part, int nblk,
                                             const int bx, const int nbx) {
  __shared__ double lds[16 * 3];
  double acc[3];
  reduce_partials<3, 0>(part, nblk, acc, lds);
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

__global__ void __launch_bounds__(1024) k_reduceA(KickScalars* __restrict__ scp, const double* __restrict__ part,
                                                  int nblk) {
  reduceA_body(scp, part, nblk, blockIdx.x, gridDim.x);
}

// pass B: rest-frame min/max, total charge, charge-weighted sums.
// rec (or null): each particle's record (rows, charge) is written as well,
// rec[p] for the caller's particle p
template <bool FAST>
__device__ __forceinline__ void passB_body(Rows R, const double* __restrict__ q, KickScalars* __restrict__ scp,
                                           double* __restrict__ part, unsigned long long* __restrict__ zg,
                                           int64_t zn, const int bx, const int nbx, PRec* __restrict__ rec = nullptr) {
  __shared__ double lds[(kPT / 64) * kNPartB];
  // the charge grid the deposit after this pass adds to is cleared here (its
  // last reader, the previous kick's z pass, ran earlier on the stream): no
  // launch and no cross-stream wait of its own
  for (int64_t i = (int64_t)bx * kPT + threadIdx.x; i < zn; i += (int64_t)nbx * kPT) zg[i] = 0ull;
  const KickScalars s = *scp;
  const double inf = __builtin_inf();
  double mn[3] = {inf, inf, inf}, mx[3] = {-inf, -inf, -inf};
  double sq = 0, saq = 0, sqx[3] = {0, 0, 0};
  int bad = 0;
  const int64_t n = s.n;
  // two particles per iteration, both loaded before either is used (more
  // loads in flight); accumulated in the order of the plain strided loop
  const int64_t stride = (int64_t)nbx * kPT;
  auto body = [&](const double* v, double qq) {
    double xp[6], xyz[3];
    if constexpr (FAST)
      to_xp_fast(v[0], v[1], v[2], v[3], v[4], v[5], s, xp);
    else
      to_xp(v[0], v[1], v[2], v[3], v[4], v[5], s, xp);
    rot(xp, s.T, xyz);
    xyz[2] = xyz[2] * s.gamma0;
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
  };
  for (int64_t p = (int64_t)bx * kPT + threadIdx.x; p < n; p += 2 * stride) {
    const int64_t p2 = p + stride;
    const bool two = p2 < n;
    double a[6], b[6], qa, qb = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) a[k] = ldnt(R.r + k * R.ld + p);
    qa = s.q_const ? s.q_val : ldnt(q + p);
    if (two) {
#pragma unroll
      for (int k = 0; k < 6; ++k) b[k] = ldnt(R.r + k * R.ld + p2);
      qb = s.q_const ? s.q_val : ldnt(q + p2);
    }
    if (rec) {
      const double va[8] = {a[0], a[1], a[2], a[3], a[4], a[5], qa, 0.0};
      store_rec(rec + p, va);
      if (two) {
        const double vb[8] = {b[0], b[1], b[2], b[3], b[4], b[5], qb, 0.0};
        store_rec(rec + p2, vb);
      }
    }
    body(a, qa);
    if (two) body(b, qb);
  }
  if (bad) atomicOr(&scp->err, (int)kErrNonFinite);
  double vals[kNPartB] = {mn[0], mn[1], mn[2], mx[0], mx[1], mx[2], sq, saq, sqx[0], sqx[1], sqx[2]};
  block_reduce<kNPartB, 1>(vals, lds);
  if (threadIdx.x == 0)
    for (int c = 0; c < kNPartB; ++c) part[c * nbx + bx] = vals[c];
}

template <bool FAST>
__global__ void __launch_bounds__(kPT) k_passB(Rows R, const double* __restrict__ q, KickScalars* __restrict__ scp,
                                               double* __restrict__ part, unsigned long long* __restrict__ zg = nullptr,
                                               int64_t zn = 0, PRec* __restrict__ rec = nullptr) {
  passB_body<FAST>(R, q, scp, part, zg, zn, blockIdx.x, gridDim.x, rec);
}

// cell sizes, grid offset, charge scale and kernel-cache decision
// (sc.py:172-186)
__device__ __forceinline__ void reduceB_body(KickScalars* __restrict__ scp, const double* __restrict__ part, int nblk,
                                             int prec, const int bx, const int nbx) {
  __shared__ double lds[16 * kNPartB];
  double out[kNPartB];
  reduce_partials<kNPartB, 1>(part, nblk, out, lds);
  const int tid = threadIdx.x;
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
    s.inv_steps[k] = 1.0 / h;
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

__global__ void __launch_bounds__(1024) k_reduceB(KickScalars* __restrict__ scp, const double* __restrict__ part,
                                                  int nblk, int prec) {
  reduceB_body(scp, part, nblk, prec, blockIdx.x, gridDim.x);
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

// Workgroups are dispatched to the XCDs round-robin (workgroup b runs on XCD
// b mod 8).  xcd_block renumbers them so that each XCD takes one contiguous
// run of blocks: for a bunch kept in cell order that run covers one slab of
// the mesh, whose field and charge cells then stay in that XCD's own L2.
// A bijection of [0, nblk) for any nblk; only the order of work changes.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nblk) {
  const int64_t q = nblk / kNumXcd, rem = nblk - q * kNumXcd;
  const int64_t x = b % kNumXcd, j = b / kNumXcd;
  return x * q + (x < rem ? x : rem) + j;
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

// NGP cell of a particle from its six rows, -1 if outside the mesh; FAST
// (mixed form) takes the reciprocal-multiply transforms, the FP64 form the
// exact sequence of the reference
template <bool FAST>
__device__ __forceinline__ int64_t ngp_cell(const double* R6, const KickScalars& s) {
  double xp[6], xyz[3], X[3];
  if constexpr (FAST)
    to_xp_fast(R6[0], R6[1], R6[2], R6[3], R6[4], R6[5], s, xp);
  else
    to_xp(R6[0], R6[1], R6[2], R6[3], R6[4], R6[5], s, xp);
  rot(xp, s.T, xyz);
  xyz[2] = xyz[2] * s.gamma0;
  if constexpr (FAST)
    grid_coords_fast(xyz, s, X);
  else
    grid_coords(xyz, s, X);
  const int nx = s.nmesh[0], ny = s.nmesh[1], nz = s.nmesh[2];
  const double f0 = __builtin_floor(X[0]) + 1, f1 = __builtin_floor(X[1]) + 1, f2 = __builtin_floor(X[2]) + 1;
  if (!(f0 >= 0 && f0 < nx && f1 >= 0 && f1 < ny && f2 >= 0 && f2 < nz)) return -1;
  return ((int64_t)f0 * ny + (int64_t)f1) * nz + (int64_t)f2;
}

// Deposit, pass C.  A workgroup handles kDepItems consecutive particles per
// thread and accumulates them in an LDS hash table keyed by cell (64-bit
// integer LDS atomics), then adds each occupied slot to the global grid with
// one atomic.  For a bunch kept in cell order (tracking) a workgroup's
// particles occupy a few hundred neighbouring cells, so global atomics drop by
// one to two orders of magnitude; in random order the table simply passes
// every particle through.  Integer sums: the grid is identical either way.
constexpr int kDepPT = 256;                  // threads per workgroup
constexpr int kDepItems = 8;                 // particles per thread
constexpr int kDepSlots = 2048;              // hash slots (power of two), one-word (mixed) form
constexpr int kDepSlots2 = 1024;             // two-word (FP64) form: half the LDS per workgroup, more
                                             // workgroups per CU (FP64 step 0.964 -> 0.950 ms)

// deposit tiles (option "deposit_tile"): 0 the sizes above, 1 half the
// particles per thread and half the slots, 2 twice both, 3 no table (every
// particle one global atomic: a bunch in random order spreads a workgroup's
// particles over far more cells than a table holds, and each insert then
// probes a nearly full table before it falls through), 4 as 0 with the rows
// of all of a thread's particles loaded before the first is deposited (tile
// 0 loads one particle ahead; the deposit waits on memory most of its time)
template <int TILE>
struct DepTile {
  static constexpr bool direct = TILE == 3;
  static constexpr bool all_loads = TILE == 4;  // as 0, every item's loads issued first
  static constexpr int items = TILE == 1 ? kDepItems / 2 : (TILE == 2 ? kDepItems * 2 : kDepItems);
  template <int NWORD>
  static constexpr int slots() {
    constexpr int base = NWORD == 2 ? kDepSlots2 : kDepSlots;
    return TILE == 1 ? base / 2 : (TILE == 2 ? base * 2 : (TILE == 3 ? 1 : base));
  }
};

template <bool XCD>
__device__ __forceinline__ void flush_cell(unsigned long long* g, int64_t ncell, int64_t idx, long long hi,
                                           long long lo, int nword) {
  grid_add<XCD>(g, idx, hi);
  if (nword == 2) grid_add<XCD>(g, ncell + idx, lo);
}

// cells (may be null): receives the flat cell index each particle's charge
// went to (-1 outside the mesh), in the order of the rows
// one (cell, fixed-point value) into the workgroup's LDS hash table (linear
// probing; cell indices fit in 31 bits: mesh <= 1024^3), or straight to the
// grid when the table is crowded
template <int NWORD, bool XCD, int kSlots>
__device__ __forceinline__ void hash_add(int* hkey, long long (*hval)[kSlots], unsigned long long* g, int64_t ncell,
                                         int64_t idx, long long hi, long long lo) {
  unsigned slot = ((unsigned)idx * 2654435761u) >> (32 - ilog2(kSlots));
  for (int probe = 0; probe < 32; ++probe) {
    const int k = atomicCAS(&hkey[slot], -1, (int)idx);
    if (k == -1 || k == (int)idx) {
      atomicAdd((unsigned long long*)&hval[0][slot], (unsigned long long)hi);
      if constexpr (NWORD == 2) atomicAdd((unsigned long long*)&hval[1][slot], (unsigned long long)lo);
      return;
    }
    slot = (slot + 1) & (kSlots - 1);
  }
  flush_cell<XCD>(g, ncell, idx, hi, lo, NWORD);
}

// Wave-level pre-sum (AGG): in cell order, neighbouring lanes usually carry
// the same cell, and their LDS atomics on one slot serialise.  Runs of equal
// cells in lane order are summed first (segmented inclusive scan with
// shuffles, exact in integers); only the last lane of each run goes to the
// table.  Cells that recur in separate runs are simply added twice.
// RECIN: particle p of the cell order is the record rin[ord[p]] (rows and
// charge) instead of column p of R and q
template <int NWORD, bool XCD, bool AGG, int TILE, bool RECIN = false>
__device__ __forceinline__ void deposit_body(Rows R, const double* __restrict__ q, KickScalars* __restrict__ scp,
                                             unsigned long long* __restrict__ grid, int32_t* __restrict__ cells,
                                             const int bx, const int nbx, const PRec* __restrict__ rin = nullptr,
                                             const uint32_t* __restrict__ ord = nullptr) {
  constexpr int kSlots = DepTile<TILE>::template slots<NWORD>();
  constexpr int kItems = DepTile<TILE>::items;
  __shared__ int hkey[kSlots];
  __shared__ long long hval[NWORD][kSlots];
  const KickScalars& s0 = *scp;
  if (s0.err) return;  // uniform across the grid
  const KickScalars s = s0;
  const int64_t ncell = (int64_t)s.nmesh[0] * s.nmesh[1] * s.nmesh[2];
  unsigned long long* g = grid;
  if constexpr (XCD) g += (int64_t)xcc_id() * NWORD * ncell;
  constexpr bool kDirect = DepTile<TILE>::direct;
  if constexpr (!kDirect) {
    for (int i = threadIdx.x; i < kSlots; i += kDepPT) {
      hkey[i] = -1;
      hval[0][i] = 0;
      if constexpr (NWORD == 2) hval[1][i] = 0;
    }
    __syncthreads();
  }
  const int64_t base = xcd_block(bx, nbx) * (kDepPT * kItems);
  const int lane = threadIdx.x & 63;
  int bad = 0;
  auto load = [&](int64_t p, double* v) {
    if constexpr (RECIN) {
      double t[8];
      load_rec(rin + ord[p], t);
#pragma unroll
      for (int k = 0; k < 6; ++k) v[k] = t[k];
      v[6] = s.q_const ? s.q_val : t[6];
    } else {
#pragma unroll
      for (int k = 0; k < 6; ++k) v[k] = ldnt(R.r + k * R.ld + p);
      v[6] = s.q_const ? s.q_val : ldnt(q + p);
    }
  };
  // one particle (valid: p < n) through the table; cur holds its rows
  auto item = [&](int64_t p, bool valid, const double* cur) {
    int64_t idx = -1;
    long long hi = 0, lo = 0;
    if (valid) {
      idx = ngp_cell<NWORD == 1>(cur, s);  // one word: the mixed form
      if (cells) cells[p] = (int32_t)idx;
      if (idx < 0) {
        bad = 1;
      } else {
        const double v = cur[6] * s.qscale;  // exact: qscale is a power of two
        const double h = __builtin_rint(v);
        hi = (long long)h;
        if constexpr (NWORD == 2) lo = (long long)__builtin_rint((v - h) * 0x1p30);
      }
    }
    if constexpr (AGG) {
      const int key = valid && idx >= 0 ? (int)idx : -1;
      const int kprev = __shfl_up(key, 1);
      const uint64_t heads = __ballot(lane == 0 || key != kprev);
      const uint64_t upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1);
      const int start = 63 - __clzll(heads & upto);  // first lane of this lane's run
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const long long h2 = __shfl_up(hi, off);
        long long l2 = 0;
        if constexpr (NWORD == 2) l2 = __shfl_up(lo, off);
        if (lane - off >= start) {
          hi += h2;
          if constexpr (NWORD == 2) lo += l2;
        }
      }
      const int knext = __shfl_down(key, 1);
      if (key >= 0 && (lane == 63 || knext != key)) {
        if constexpr (kDirect)
          flush_cell<XCD>(g, ncell, key, hi, lo, NWORD);
        else
          hash_add<NWORD, XCD, kSlots>(hkey, hval, g, ncell, key, hi, lo);
      }
    } else if (idx >= 0) {
      if constexpr (kDirect)
        flush_cell<XCD>(g, ncell, idx, hi, lo, NWORD);
      else
        hash_add<NWORD, XCD, kSlots>(hkey, hval, g, ncell, idx, hi, lo);
    }
  };
  if constexpr (DepTile<TILE>::all_loads) {
    // every item's rows in flight before the first goes through the table
    double v[kItems][7];
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      const int64_t p = base + threadIdx.x + (int64_t)it * kDepPT;
      if (p < s.n) load(p, v[it]);
    }
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      const int64_t p = base + threadIdx.x + (int64_t)it * kDepPT;
      const bool valid = p < s.n;
      if constexpr (AGG) {
        if (__ballot(valid) == 0) break;  // uniform across the wave: every lane takes part in the scan
      } else {
        if (!valid) break;
      }
      item(p, valid, v[it]);
    }
  } else {
    // the rows of item it + 1 are loaded before item it goes through the table
    double cur[7], nxt[7] = {0, 0, 0, 0, 0, 0, 0};
    int64_t p = base + threadIdx.x;
    if (p < s.n) load(p, cur);
    for (int it = 0; it < kItems; ++it, p += kDepPT) {
      const bool valid = p < s.n;
      if constexpr (AGG) {
        if (__ballot(valid) == 0) break;  // uniform across the wave: every lane takes part in the scan
      } else {
        if (!valid) break;
      }
      if (valid) {
        const int64_t pn = p + kDepPT;
        if (it + 1 < kItems && pn < s.n) load(pn, nxt);
      }
      item(p, valid, cur);
#pragma unroll
      for (int k = 0; k < 7; ++k) cur[k] = nxt[k];
    }
  }
  if (bad) atomicOr(&scp->err, (int)kErrCell);
  if constexpr (!kDirect) {
    __syncthreads();
    for (int i = threadIdx.x; i < kSlots; i += kDepPT) {
      const int k = hkey[i];
      if (k >= 0) flush_cell<XCD>(g, ncell, k, hval[0][i], NWORD == 2 ? hval[NWORD - 1][i] : 0, NWORD);
    }
  }
}

template <int NWORD, bool XCD = false, bool AGG = false, int TILE = 0, bool RECIN = false>
__global__ void __launch_bounds__(kDepPT) k_deposit(Rows R, const double* __restrict__ q, KickScalars* __restrict__ scp,
                                                 unsigned long long* __restrict__ grid, int32_t* __restrict__ cells,
                                                 const PRec* __restrict__ rin = nullptr,
                                                 const uint32_t* __restrict__ ord = nullptr) {
  deposit_body<NWORD, XCD, AGG, TILE, RECIN>(R, q, scp, grid, cells, blockIdx.x, gridDim.x, rin, ord);
}

// ---------------------------------------------------------------------------
// Cell ordering of a resident bunch (multi-kick tracking): sort keys are the
// NGP cells under the scalars of the last kick; particles move little between
// re-sorts, so consecutive particles keep sharing cells.
// Sort keys: bricks of kBrick cells (8 x 8 x 4, z fastest), bricks in C
// order, cells in C order inside a brick, so a run of consecutive particles
// fills a compact 3D box of the mesh.  Keys only steer locality, so the fast
// transforms are used; out-of-mesh particles sort last.
constexpr int kBx = 8, kBy = 8, kBz = 4;

// number of keys (the out-of-mesh key): bricks times cells per brick, or
// bricks only (cells == 0: the order inside a brick is left as it was)
__host__ __device__ inline int64_t brick_key_count(int nx, int ny, int nz, int cells = 1) {
  const int64_t nbx = (nx + kBx - 1) / kBx, nby = (ny + kBy - 1) / kBy, nbz = (nz + kBz - 1) / kBz;
  return nbx * nby * nbz * (cells ? kBx * kBy * kBz : 1);
}

// key bits the radix sort must look at
__host__ __device__ inline int sort_end_bit(int nx, int ny, int nz, int cells) {
  int b = 1;
  while ((int64_t(1) << b) <= brick_key_count(nx, ny, nz, cells)) ++b;
  return b;
}

// rec (may be null): also write each particle's record (rows and charge q),
// the sequential half of the permutation into cell order.  ahead: the keys
// are taken at the position after a drift of that length (the middle of the
// coming re-sort interval, so particles stay as close to their sorted place
// before it as after it); the rows are not changed.
__device__ __forceinline__ void sort_keys_body(Rows R, const double* __restrict__ q,
                                               const KickScalars* __restrict__ scp, uint32_t* __restrict__ keys,
                                               PRec* __restrict__ rec, int cells, DriftArgs ahead, uint32_t tag,
                                               const int bx, const int nbx) {
  const KickScalars s = *scp;
  const int64_t p = (int64_t)bx * kPT + threadIdx.x;
  if (p >= s.n) return;
  double v[8];
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = ldnt(R.r + k * R.ld + p);
  if (rec) {
    v[6] = ldnt(q + p);
    v[7] = 0.0;
    store_rec(rec + p, v);
  }
  double xp[6], xyz[3], X[3];
  const double x = __builtin_fma(ahead.L, v[1], v[0]), y = __builtin_fma(ahead.L, v[3], v[2]),
               t = __builtin_fma(ahead.r56, v[5], v[4]);
  to_xp_fast(x, v[1], y, v[3], t, v[5], s, xp);
  rot(xp, s.T, xyz);
  xyz[2] = xyz[2] * s.gamma0;
  grid_coords_fast(xyz, s, X);
  const int nx = s.nmesh[0], ny = s.nmesh[1], nz = s.nmesh[2];
  const double g0 = __builtin_floor(X[0]) + 1, g1 = __builtin_floor(X[1]) + 1, g2 = __builtin_floor(X[2]) + 1;
  uint32_t key = (uint32_t)brick_key_count(nx, ny, nz, cells);
  if (g0 >= 0 && g0 < nx && g1 >= 0 && g1 < ny && g2 >= 0 && g2 < nz) {
    const int f0 = (int)g0, f1 = (int)g1, f2 = (int)g2;
    const int nby = (ny + kBy - 1) / kBy, nbz = (nz + kBz - 1) / kBz;
    const int64_t brick = ((int64_t)(f0 / kBx) * nby + f1 / kBy) * nbz + f2 / kBz;
    key = cells ? (uint32_t)(brick * (kBx * kBy * kBz) + ((f0 % kBx) * kBy + f1 % kBy) * kBz + f2 % kBz)
                : (uint32_t)brick;
  }
  keys[p] = tag | key;
}

__global__ void __launch_bounds__(kPT) k_sort_keys(Rows R, const double* __restrict__ q,
                                                   const KickScalars* __restrict__ scp, uint32_t* __restrict__ keys,
                                                   PRec* __restrict__ rec, int cells, DriftArgs ahead,
                                                   uint32_t tag = 0) {
  sort_keys_body(R, q, scp, keys, rec, cells, ahead, tag, blockIdx.x, gridDim.x);
}

// dst[:, j] = rec[perm[j]] (six rows and the charge), optionally applying a
// drift to the rows it writes (as k_drift rounds it)
template <bool DRIFT>
__global__ void __launch_bounds__(256) k_gather_rec(const PRec* __restrict__ rec, const uint32_t* __restrict__ perm,
                                                    double* __restrict__ dst, int64_t ldd, double* __restrict__ qd,
                                                    int64_t n, DriftArgs dr) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  double v[8];
  load_rec(rec + perm[j], v);
  if constexpr (DRIFT) {
    v[0] = __builtin_fma(dr.L, v[1], v[0]);
    v[2] = __builtin_fma(dr.L, v[3], v[2]);
    v[4] = __builtin_fma(dr.r56, v[5], v[4]);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) stnt(v[k], dst + k * ldd + j);
  stnt(v[6], qd + j);
}

// rec[perm[j]] = src[:, j] (six rows; the random half of the way back to the
// caller's order: whole-segment writes)
__global__ void __launch_bounds__(256) k_scatter_rec(const double* __restrict__ src, int64_t lds_,
                                                     const uint32_t* __restrict__ perm, PRec* __restrict__ rec,
                                                     int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  double v[8];
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = ldnt(src + k * lds_ + j);
  v[6] = 0.0;
  v[7] = 0.0;
  store_rec(rec + perm[j], v);
}

// dst[:, i] = rec[i] for the six rows (sequential)
__global__ void __launch_bounds__(256) k_rec_to_rows(const PRec* __restrict__ rec, double* __restrict__ dst,
                                                     int64_t ldd, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double v[8];
  load_rec(rec + i, v);
#pragma unroll
  for (int k = 0; k < 6; ++k) stnt(v[k], dst + k * ldd + i);
}

// k_scatter_rec only if the kick flagged the bunch as invalid (that kick
// then wrote no records; the copy goes back as it stands)
__global__ void __launch_bounds__(256) k_scatter_rec_if_err(const double* __restrict__ src, int64_t lds_,
                                                            const uint32_t* __restrict__ perm, PRec* __restrict__ rec,
                                                            int64_t n, const KickScalars* __restrict__ scp) {
  if (!scp->err) return;
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  double v[8];
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = src[k * lds_ + j];
  v[6] = 0.0;
  v[7] = 0.0;
  store_rec(rec + perm[j], v);
}

// k_rec_to_rows unless the kick flagged the bunch as invalid (then the
// caller's rows stay as they were)
__global__ void __launch_bounds__(256) k_rec_to_rows_checked(const PRec* __restrict__ rec, double* __restrict__ dst,
                                                             int64_t ldd, int64_t n,
                                                             const KickScalars* __restrict__ scp) {
  if (scp->err) return;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double v[8];
  load_rec(rec + i, v);
#pragma unroll
  for (int k = 0; k < 6; ++k) stnt(v[k], dst + k * ldd + i);
}

// Re-sort of an already cell-ordered copy: dst[:, j] = src[:, order[j]] for
// the six rows and the charge, and perm_dst[j] = perm_src[order[j]] (the
// caller's index of the particle now at j), in one pass.  The particles moved
// a few cells since the last sort, so order[j] stays near j and the reads of
// a workgroup fall in a small window of each row.
__global__ void __launch_bounds__(256) k_resort_gather(const double* __restrict__ src, const double* __restrict__ qs,
                                                       const uint32_t* __restrict__ order,
                                                       const uint32_t* __restrict__ perm_src, double* __restrict__ dst,
                                                       double* __restrict__ qd, uint32_t* __restrict__ perm_dst,
                                                       int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const int64_t i = order[j];
  double v[7];
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = src[k * n + i];
  v[6] = qs[i];
  const uint32_t pi = perm_src[i];
#pragma unroll
  for (int k = 0; k < 6; ++k) stnt(v[k], dst + k * n + j);
  stnt(v[6], qd + j);
  perm_dst[j] = pi;
}

// dst[:, j] = src[:, perm[j]] for the six rows and the charge
// One row per blockIdx.y (rows 0-5, then q as row 6): workgroups are
// dispatched in x-major order, so a pass's random reads (gather) or writes
// (scatter) stay within one row of n doubles, which the memory-side cache
// holds, instead of spreading over all seven rows at once.
__global__ void __launch_bounds__(256) k_permute_gather(const double* __restrict__ src, int64_t lds_,
                                                        const double* __restrict__ qs, const uint32_t* __restrict__ perm,
                                                        double* __restrict__ dst, int64_t ldd, double* __restrict__ qd,
                                                        int64_t n) {
  // one contiguous run of destinations per XCD: after a re-sort the sources
  // of a run come from the same region of the mesh, so the lines they share
  // are fetched into one L2
  const int64_t j = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (j >= n) return;
  const int row = blockIdx.y;
  const int64_t p = perm[j];
  if (row < 6)
    dst[row * ldd + j] = src[row * lds_ + p];
  else
    qd[j] = qs[p];
}

// dst[:, perm[j]] = src[:, j] for the six rows (one row per blockIdx.y)
__global__ void __launch_bounds__(256) k_permute_scatter(const double* __restrict__ src, int64_t lds_,
                                                         const uint32_t* __restrict__ perm, double* __restrict__ dst,
                                                         int64_t ldd, int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const int row = blockIdx.y;
  dst[row * ldd + perm[j]] = src[row * lds_ + j];
}

// out[j] = a[b[j]]
__global__ void __launch_bounds__(256) k_compose(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                                                 uint32_t* __restrict__ out, int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j < n) out[j] = a[b[j]];
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
// same zero-plane and edge rules as gather_diff.  The two z-neighbours of a
// corner pair are adjacent in memory and are read with one 8-byte load (the
// address is only 4-byte aligned, which global loads allow): 12 loads per
// particle for the three components instead of 24.
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
  const int64_t o0 = i0[0] * sx + i0[1] * sy + i0[2];
  float v[8];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      float2 pr;
      __builtin_memcpy(&pr, E + o0 + a * sx + b * sy, sizeof(pr));
      bool z0 = false, z1 = false;  // corners on the component's last plane read 0
      if (AX == 0) z0 = z1 = i0[0] + a >= nx - 1;
      if (AX == 1) z0 = z1 = i0[1] + b >= ny - 1;
      if (AX == 2) {
        z0 = i0[2] >= nz - 1;
        z1 = i0[2] + 1 >= nz - 1;
      }
      v[a * 4 + b * 2 + 0] = z0 ? 0.f : pr.x;
      v[a * 4 + b * 2 + 1] = z1 ? 0.f : pr.y;
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

// Mixed form, the three components at once.  Component k samples its grid at
// X on axis k and at X + 1/2 on the other two (sc.py:202-204), so the six
// floors and fractions are computed once; offsets are 32-bit (grids hold at
// most 512^3 nodes), so each pair load is one scalar-based global load.
// Interpolation by nested linear blends in FP32; zero-plane and edge rules as
// in gather_grid.
__device__ __forceinline__ void gather_field3(const float* __restrict__ E, const int* n, const double* X,
                                              float* out) {
  int f[2][3];
  float t[2][3];
  bool ok[2][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const double c = h ? X[k] + 0.5 : X[k];
      ok[h][k] = c >= 0 && c <= n[k] - 1;
      int fl = ok[h][k] ? (int)__builtin_floor(c) : 0;
      fl = fl > n[k] - 2 ? n[k] - 2 : fl;
      f[h][k] = fl;
      t[h][k] = (float)(c - (double)fl);
    }
  const unsigned sy = (unsigned)n[2], sx = (unsigned)(n[1] * n[2]), nc = sx * (unsigned)n[0];
  const char* base = (const char*)E;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int h0 = ax == 0 ? 0 : 1, h1 = ax == 1 ? 0 : 1, h2 = ax == 2 ? 0 : 1;
    const int i0 = f[h0][0], i1 = f[h1][1], i2 = f[h2][2];
    const unsigned o = ax * nc + (unsigned)i0 * sx + (unsigned)i1 * sy + (unsigned)i2;
    float v[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        float2 pr;
        __builtin_memcpy(&pr, base + (size_t)((o + a * sx + b * sy) * 4u), sizeof(pr));
        bool z0 = false, z1 = false;  // corners on the component's last plane read 0
        if (ax == 0) z0 = z1 = i0 + a >= n[0] - 1;
        if (ax == 1) z0 = z1 = i1 + b >= n[1] - 1;
        if (ax == 2) {
          z0 = i2 >= n[2] - 1;
          z1 = i2 + 1 >= n[2] - 1;
        }
        v[a][b][0] = z0 ? 0.f : pr.x;
        v[a][b][1] = z1 ? 0.f : pr.y;
      }
    const float tx = t[h0][0], ty = t[h1][1], tz = t[h2][2];
    float r[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const float q0 = __builtin_fmaf(tz, v[a][0][1] - v[a][0][0], v[a][0][0]);
      const float q1 = __builtin_fmaf(tz, v[a][1][1] - v[a][1][0], v[a][1][0]);
      r[a] = __builtin_fmaf(ty, q1 - q0, q0);
    }
    const float res = __builtin_fmaf(tx, r[1] - r[0], r[0]);
    out[ax] = (ok[h0][0] && ok[h1][1] && ok[h2][2]) ? res : 0.f;
  }
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

// FP64 form: the three staggered difference fields E_AX = (phi[i] -
// phi[i+1]) / h_AX (zero on each component's last plane) formed once per kick
// from the potential, the arrays the reference interpolates (sc.py:198-204);
// the kick then reads 8 corner values per component instead of 16 potential
// values and divides nothing.  Stored one after the other, C order.
__device__ __forceinline__ void diff64_body(const double* __restrict__ phi, const KickScalars* __restrict__ scp,
                                            double* __restrict__ E, const int bx, const int nbx) {
  const int nx = scp->nmesh[0], ny = scp->nmesh[1], nz = scp->nmesh[2];
  const int64_t nc = (int64_t)nx * ny * nz;
  const int64_t o = (int64_t)bx * 256 + threadIdx.x;
  if (o >= nc) return;
  const int64_t sx = (int64_t)ny * nz;
  const int i = (int)(o / sx), j = (int)((o / nz) % ny), k = (int)(o % nz);
  const double p0 = phi[o];
  E[o] = i < nx - 1 ? (p0 - phi[o + sx]) / scp->steps[0] : 0.0;
  E[nc + o] = j < ny - 1 ? (p0 - phi[o + nz]) / scp->steps[1] : 0.0;
  E[2 * nc + o] = k < nz - 1 ? (p0 - phi[o + 1]) / scp->steps[2] : 0.0;
}

__global__ void __launch_bounds__(256) k_diff64(const double* __restrict__ phi, const KickScalars* __restrict__ scp,
                                                double* __restrict__ E) {
  diff64_body(phi, scp, E, blockIdx.x, gridDim.x);
}

// trilinear interpolation of a k_diff64 component, in gather_diff's order of
// corners and operations (the same result bit for bit)
template <int AX>
__device__ __forceinline__ double gather_grid64(const double* __restrict__ E, int nx, int ny, int nz, const double* c) {
  const int n[3] = {nx, ny, nz};
  if (!(c[0] >= 0 && c[0] <= nx - 1 && c[1] >= 0 && c[1] <= ny - 1 && c[2] >= 0 && c[2] <= nz - 1)) return 0.0;
  int i0[3];
  double t[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int f = (int)__builtin_floor(c[k]);
    f = f > n[k] - 2 ? n[k] - 2 : f;
    i0[k] = f;
    t[k] = c[k] - (double)f;
  }
  const int64_t sx = (int64_t)ny * nz, sy = nz;
  double acc = 0;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      // the two z-neighbours of a corner pair in one 16-byte load
      double2 pr;
      __builtin_memcpy(&pr, E + (i0[0] + a) * sx + (i0[1] + b) * sy + i0[2], sizeof(pr));
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const double ev = e ? pr.y : pr.x;
        const double w = (a ? t[0] : 1 - t[0]) * (b ? t[1] : 1 - t[1]) * (e ? t[2] : 1 - t[2]);
        acc += w * ev;
      }
    }
  }
  return acc;
}

__device__ __forceinline__ void gather_field64(const double* __restrict__ E3, const KickScalars& s, const double* X,
                                               double* E) {
  const int nx = s.nmesh[0], ny = s.nmesh[1], nz = s.nmesh[2];
  const int64_t nc = (int64_t)nx * ny * nz;
  const double cx[3] = {X[0], X[1] + 0.5, X[2] + 0.5};
  const double cy[3] = {X[0] + 0.5, X[1], X[2] + 0.5};
  const double cz[3] = {X[0] + 0.5, X[1] + 0.5, X[2]};
  E[0] = gather_grid64<0>(E3, nx, ny, nz, cx) * s.gamma0;
  E[1] = gather_grid64<1>(E3 + nc, nx, ny, nz, cy) * s.gamma0;
  E[2] = gather_grid64<2>(E3 + 2 * nc, nx, ny, nz, cz);
}

// pass D: gather, kick in the rotated frame, rotate back, write all six rows
// in place (sc.py:241-251, coord_transform.py:16-54).
// Tracking fusions: DRIFT applies the next step's drift to the written rows
// (x += L x', y += L y', tau += r56 delta, as numpy.dot(R, X) rounds it);
// SUMP stores per-block sums of the momenta the next kick's pass A would
// compute from the written rows (a drift leaves x', y', delta unchanged).
// threads per workgroup of the kick, one particle each (and one momentum
// partial per workgroup); 128 and 512 measured no faster
constexpr int kKickPT = 256;

// RECOUT: the rows are written as records to rec[perm[p]] (the caller's
// position of the cell-ordered particle p) instead of in place: the last kick
// of a tracking call and the way back to the caller's order in one pass.
// zg, zn (tracking): zn words of the charge grid to clear for the next step's
// deposit (the z pass that read them ran before this kick).
// RECIN (with RECOUT): particle p is read from the record rec[perm[p]] that
// it is written back to, instead of column p of r
template <typename F, bool DRIFT, bool SUMP, bool RECOUT, bool RECIN = false>
__device__ __forceinline__ void kick_body(double* __restrict__ r, int64_t ld, const F* __restrict__ phi,
                                          const KickScalars* __restrict__ scp, DriftArgs dr, double* __restrict__ part,
                                          const uint32_t* __restrict__ perm, PRec* __restrict__ rec,
                                          unsigned long long* __restrict__ zg, int64_t zn, const int bx,
                                          const int nbx) {
  __shared__ double lds[SUMP ? (kKickPT / 64) * 3 : 1];
  const KickScalars& s0 = *scp;
  if (s0.err) return;  // uniform across the grid
  for (int64_t i = (int64_t)bx * kKickPT + threadIdx.x; i < zn; i += (int64_t)nbx * kKickPT) zg[i] = 0ull;
  const KickScalars s = s0;
  const int64_t blk = xcd_block(bx, nbx);
  const int64_t p = blk * kKickPT + threadIdx.x;
  const bool active = p < s.n;
  // both forms take the fast transforms (to_xp_fast, to_ocelot_fast: the
  // rows differ from the reference's operation sequence in the last bits;
  // FP64 step 1.035 -> 0.952 ms at 10M / 127^3); the FP64 form keeps the exact
  // grid coordinates, the FP64 field and the exact pass-A momentum sums
  constexpr bool FAST = sizeof(F) == 4;
  double xp[6], xyz[3], X[3], pr[3], E[3], pb[3], o[8];
  double m[3] = {0, 0, 0};
  const int nn[3] = {s.nmesh[0], s.nmesh[1], s.nmesh[2]};
  static_assert(!RECIN || RECOUT, "records in only with records out");
  if (active) {
    if constexpr (RECIN) {
      double t[8];
      load_rec(rec + perm[p], t);
      to_xp_fast(t[0], t[1], t[2], t[3], t[4], t[5], s, xp);
    } else {
      to_xp_fast(ldnt(r + p), ldnt(r + ld + p), ldnt(r + 2 * ld + p), ldnt(r + 3 * ld + p), ldnt(r + 4 * ld + p),
                 ldnt(r + 5 * ld + p), s, xp);
    }
    rot(xp, s.T, xyz);
    xyz[2] = xyz[2] * s.gamma0;
    if constexpr (FAST)
      grid_coords_fast(xyz, s, X);
    else
      grid_coords(xyz, s, X);
    rot(xp + 3, s.T, pr);
    if constexpr (FAST) {
      float g[3];
      gather_field3((const float*)phi, nn, X, g);
      E[0] = (double)g[0] * s.gamma0;
      E[1] = (double)g[1] * s.gamma0;
      E[2] = (double)g[2];
    } else {
      gather_field64((const double*)phi, s, X, E);  // phi: the k_diff64 fields
    }
    pr[0] = pr[0] + s.kperp * E[0];
    pr[1] = pr[1] + s.kperp * E[1];
    pr[2] = pr[2] + s.cdT * E[2];
    rot_back(pr, s.T, pb);
    to_ocelot_fast(xp, pb, s, o);
    if constexpr (DRIFT) {
      o[0] = __builtin_fma(dr.L, o[1], o[0]);
      o[2] = __builtin_fma(dr.L, o[3], o[2]);
      o[4] = __builtin_fma(dr.r56, o[5], o[4]);
    }
    if constexpr (RECOUT) {
      o[6] = 0.0;
      o[7] = 0.0;
      store_rec(rec + perm[p], o);
    } else {
#pragma unroll
      for (int i = 0; i < 6; ++i) stnt(o[i], r + i * ld + p);
    }
    if constexpr (SUMP) {
      // next kick's pass A: the momentum of the written rows (a drift keeps
      // x', y', delta); the mixed form takes the lab momentum it just wrote
      // them from, equal up to rounding
      if constexpr (FAST) {
        m[0] = pb[0];
        m[1] = pb[1];
        m[2] = pb[2];
      } else {
        to_p(o[1], o[3], o[5], s, m);
      }
    }
  }
  if constexpr (SUMP) {
    block_reduce<3, 0>(m, lds);
    if (threadIdx.x == 0)
      for (int c = 0; c < 3; ++c) part[c * nbx + blk] = m[c];
  }
}

template <typename F, bool DRIFT = false, bool SUMP = false, bool RECOUT = false, bool RECIN = false>
__global__ void __launch_bounds__(kKickPT) k_kick(double* __restrict__ r, int64_t ld, const F* __restrict__ phi,
                                                  const KickScalars* __restrict__ scp, DriftArgs dr = {0, 0},
                                                  double* __restrict__ part = nullptr,
                                                  const uint32_t* __restrict__ perm = nullptr,
                                                  PRec* __restrict__ rec = nullptr,
                                                  unsigned long long* __restrict__ zg = nullptr, int64_t zn = 0) {
  kick_body<F, DRIFT, SUMP, RECOUT, RECIN>(r, ld, phi, scp, dr, part, perm, rec, zg, zn, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// el_field entry (sc.py:170-205): rest-frame positions given directly as
// three rows (x, y, z before the gamma scaling of z).

__global__ void __launch_bounds__(kPT) k_passB_xyz(Rows R, const double* __restrict__ q,
                                                   KickScalars* __restrict__ scp, double* __restrict__ part) {
  __shared__ double lds[(kPT / 64) * kNPartB];
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
  block_reduce<kNPartB, 1>(vals, lds);
  if (threadIdx.x == 0)
    for (int c = 0; c < kNPartB; ++c) part[c * gridDim.x + blockIdx.x] = vals[c];
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
