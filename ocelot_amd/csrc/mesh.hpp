// Open-boundary Poisson solve on the rest-frame mesh (Hockney doubled grid).
//
// Reference: ocelot/cpbd/sc.py:109-168 (sym_kernel, potential).  The reference
// convolves on (2N-1)^3 with numpy/pyfftw; here every axis uses the power of
// two M >= 2N-1 (the circular convolution restricted to [0,N)^3 is the same
// linear convolution for any M >= 2N-1, with the even kernel laid out as
// K2[M-m] = K[m]).  The transforms are pruned: the charge has N non-zeros per
// line and only N outputs per line are kept, so no pass touches the zero
// padding in HBM.
//
// Layouts (C order, last index fastest; Hz = Mz/2+1):
//   rho/phi  [nx][ny][nz]      real
//   A        [nx][ny][Hz]      complex   (after the z pass)
//   B        [nx][My][Hz]      complex   (after the y pass; x pass in place)
//   G        [Hx][Hy][Hz]      real      kernel spectrum, even in every axis,
//                                        scaled by 1/(Mx My Mz 4 pi eps0 hx hy hz)
#pragma once
#include "fft.hpp"

namespace ocsc {

constexpr int kNT = 256;  // threads per workgroup for the mesh kernels

// columns per workgroup: LDS footprint M*(T+1)*sizeof(cpx<F>) stays <= ~72 KB
// (plus the twiddle table W_M^k, k < M, staged behind the tile when the
// workgroup's LDS stays within 64 KB: every Stockham pass reads it)
// V selects the plan (option "fft_plan"): 0 the default, 1 radix 8 in every
// pass, 2 half as many columns per workgroup (less LDS, more workgroups)
template <typename F, int M, int V = 0>
struct Tile {
  static constexpr int cap = sizeof(F) == 4 ? 64 : 32;
  static constexpr int want = (sizeof(F) == 4 ? 4096 : 2048) / M;
  static constexpr int T0 = want < 2 ? 2 : (want > cap ? cap : want);
  static constexpr int T = V == 2 ? (T0 / 2 < 2 ? 2 : T0 / 2) : T0;
  static constexpr int RSEL = V == 1 ? 8 : 0;
  static constexpr int S = T + 1;
  static constexpr size_t data_bytes = (size_t)M * S * sizeof(cpx<F>);
  static constexpr bool stw = data_bytes + M * sizeof(cpx<F>) <= 65536;
  static constexpr size_t lds_bytes = data_bytes + (stw ? M * sizeof(cpx<F>) : 0);
};

// Copy the twiddle table into LDS at dst (ON) and return the table the
// transforms should read; the kernel's barrier before its first transform
// publishes the copy.
template <typename F, int M, int NT, bool ON>
__device__ __forceinline__ const cpx<F>* stage_tw(cpx<F>* dst, const cpx<F>* __restrict__ tw) {
  if constexpr (ON) {
    for (int i = threadIdx.x; i < M; i += NT) dst[i] = tw[i];
    return dst;
  } else {
    return tw;
  }
}

struct MeshDims {
  int nx, ny, nz;
  int mx, my, mz;
  int hx, hy, hz;  // M/2 + 1 per axis
};

enum ChargeIn : int { kChargeFix1 = 0, kChargeFix2 = 1, kChargeReal = 2 };

// charge of cell i from the fixed-point accumulators (see deposit kernel);
// ncopy grid copies (one per XCD, or 1) of NWORD words are summed exactly
template <int CIN, typename F>
__device__ __forceinline__ double charge_at(const void* src, int64_t i, int64_t ncell, double inv_qscale,
                                            int ncopy) {
  if constexpr (CIN == kChargeReal) {
    return ((const double*)src)[i];
  } else if constexpr (CIN == kChargeFix1) {
    const long long* a = (const long long*)src;
    unsigned long long hi = 0;
    for (int c = 0; c < ncopy; ++c) hi += (unsigned long long)a[c * ncell + i];
    return (double)(long long)hi * inv_qscale;
  } else {
    const long long* a = (const long long*)src;
    unsigned long long hi = 0, lo = 0;
    for (int c = 0; c < ncopy; ++c) {
      hi += (unsigned long long)a[2 * c * ncell + i];
      lo += (unsigned long long)a[(2 * c + 1) * ncell + i];
    }
    return ((double)(long long)hi + (double)(long long)lo * 0x1p-30) * inv_qscale;
  }
}

// Spectral forward difference along one axis: (phi[i] - phi[i+1]) / h has
// the transform (1 - exp(+2 pi i k / M)) / h times that of phi.  Evaluated
// in FP64 from the FP64 twiddle table (1 - cos is taken there, where it does
// not cancel) and rounded once.
template <typename F>
__device__ __forceinline__ cpx<F> dfac(const cpx<double>* __restrict__ twd, int k, double inv_h) {
  const cpx<double> w = twd[k];  // (cos, -sin) of 2 pi k / M
  return cpx<F>{(F)((1.0 - w.x) * inv_h), (F)(w.y * inv_h)};
}

// The y and x passes transform columns (a, b) of a [a][m][b] array along m,
// T columns per workgroup.  Columns are numbered flat, c = a * nb + b, so
// every workgroup but the last gets T of them: with b running over Hz = M/2+1
// spectrum planes, tiles of b alone would leave one plane (b = M/2) to a tile
// of its own in every row a.  A thread touches one column in every loop of
// these kernels (NT is a multiple of T), so it splits c once.
struct Col {
  int t, a, b;
  bool ok;
};
template <int T, int NT>
__device__ __forceinline__ Col flat_col(int ncol, int nb, int bx) {
  static_assert(NT % T == 0, "a thread must keep its column in every loop");
  Col c;
  c.t = threadIdx.x % T;
  const int col = bx * T + c.t;
  c.ok = col < ncol;
  c.a = c.ok ? col / nb : 0;
  c.b = c.ok ? col - c.a * nb : 0;
  return c;
}

// Items per thread of a loop over `count` items strided by NT threads.  The
// global loads of a tile are issued in unrolled loops of this many items,
// all before the first use, so a thread has its loads in flight together
// instead of waiting for each before the next (the loop counts depend on
// run-time sizes, so the compiler would otherwise keep them rolled).
constexpr int per_thread(int count, int nt) { return (count + nt - 1) / nt; }

// Plan 4 (option "fft_plan"): register-resident transforms (reg_fft) in
// every pass whose length is M = R*Q, Q = R (16, 64, 256) or Q = 2R (32, 128,
// 512), Q threads per column, both forms; other lengths as plan 0.
template <int M>
constexpr int reg_r() {
  return M == 16 || M == 32 ? 4 : (M == 64 || M == 128 ? 8 : (M == 256 || M == 512 ? 16 : 0));
}
template <int M>
constexpr int reg_q() {
  return reg_r<M>() ? M / reg_r<M>() : 1;
}
template <typename F, int M, int V>
constexpr bool reg_plan() {
  return V == 4 && reg_r<M>() > 0;
}
// The plan the FP64 z forward and y inverse passes take under plan V: at
// M >= 256 a register-resident FP64 column (16 complex doubles per thread)
// leaves too few waves per CU for those two passes, and they keep plan 0
// (127^3, one MI355X: z forward 26.8 vs 39.0 us, y inverse 30.1 vs 31.7 us;
// the x pass 58.3 -> 46.9 us and the z inverse 17.7 -> 15.4 us do gain).
template <typename F, int M, int V>
constexpr int plan_zf_yi() {
  return sizeof(F) == 8 && M >= 256 ? 0 : V;
}

// threads and LDS bytes per workgroup of the z and y passes (plan 4: the
// larger exchange buffer of the two lane orders, twiddles, difference factors)
template <typename F, int M, int V>
constexpr int pass_threads() {
  return reg_plan<F, M, V>() ? reg_q<M>() * Tile<F, M>::T : kNT;
}
template <typename F, int M, int V>
constexpr size_t pass_lds() {
  constexpr int R = reg_plan<F, M, V>() ? reg_r<M>() : 1;
  return reg_plan<F, M, V>() ? (size_t)(RegX<R, reg_q<M>(), Tile<F, M>::T>::entries + 2 * M) * sizeof(cpx<F>)
                             : Tile<F, M, V>::lds_bytes;
}

// Plan 4 z forward: T pairs of real rows (2t, 2t+1) as T complex columns,
// lanes j fastest (rows are contiguous in memory): thread (t, j) holds
// z = Q*n + j of its pair; z < nz <= M/2, so only n < R/2 can be non-zero.
// The spectrum then goes to LDS in natural order and is split into the two
// rows' half spectra as in zfwd_body.
template <typename F, int M, int CIN>
__device__ __forceinline__ void zfwd_reg(const void* __restrict__ rho, cpx<F>* __restrict__ A, MeshDims d,
                                         const KickScalars* __restrict__ sc, const cpx<F>* __restrict__ tw,
                                         int ncopy, const int bx) {
  constexpr int R = reg_r<M>(), Q = reg_q<M>(), T = Tile<F, M>::T, NT = Q * T;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  cpx<F>* twl = s + RegX<R, Q, T>::entries;
  if (sc->err) return;
  const int nz = d.nz, Hz = d.hz;
  const int64_t nrow = (int64_t)d.nx * d.ny;
  const int64_t ncell = nrow * nz;
  const int64_t row0 = (int64_t)bx * (2 * T);
  const double iq = sc->inv_qscale;
  const int j = threadIdx.x % Q, t = threadIdx.x / Q;
  const int64_t ra = row0 + 2 * t, rb = ra + 1;
  cpx<F> x[R];
#pragma unroll
  for (int n = 0; n < R; ++n) {
    const int m = Q * n + j;
    x[n] = cpx<F>{0, 0};
    if (n < R / 2 && m < nz) {
      if (ra < nrow) x[n].x = (F)charge_at<CIN, F>(rho, ra * nz + m, ncell, iq, ncopy);
      if (rb < nrow) x[n].y = (F)charge_at<CIN, F>(rho, rb * nz + m, ncell, iq, ncopy);
    }
  }
  for (int i = threadIdx.x; i < M; i += NT) twl[i] = tw[i];
  __syncthreads();
  reg_fft<F, R, Q, T, -1, R / 2, true>(x, s, t, j, twl);
  __syncthreads();  // the exchange buffer is read; now it holds the spectra
#pragma unroll
  for (int q = 0; q < R; ++q) s[t * M + j + Q * q] = x[q];
  __syncthreads();
  constexpr int NS = per_thread(T * (M / 2 + 1), NT);
#pragma unroll
  for (int k2 = 0; k2 < NS; ++k2) {
    const int i = threadIdx.x + k2 * NT;
    if (i >= T * Hz) break;
    const int tt = i / Hz, k = i - tt * Hz;
    const int64_t a = row0 + 2 * tt, b = a + 1;
    if (a >= nrow) continue;
    const cpx<F> z = s[tt * M + k];
    const cpx<F> w = s[tt * M + ((M - k) & (M - 1))];
    const F h = (F)0.5;
    A[a * Hz + k] = cpx<F>{(z.x + w.x) * h, (z.y - w.y) * h};
    if (b < nrow) A[b * Hz + k] = cpx<F>{(z.y + w.y) * h, (w.x - z.x) * h};
  }
}

// Plan 4 y forward: thread j of column t holds rows y = Q*n + j; y < ny <=
// M/2, so only n < R/2 can be non-zero; all M outputs are kept.
template <typename F, int M>
__device__ __forceinline__ void yfwd_reg(const cpx<F>* __restrict__ A, cpx<F>* __restrict__ B, MeshDims d,
                                         const KickScalars* __restrict__ sc, const cpx<F>* __restrict__ tw,
                                         const int bx) {
  constexpr int R = reg_r<M>(), Q = reg_q<M>(), T = Tile<F, M>::T, NT = Q * T;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  cpx<F>* twl = s + RegX<R, Q, T>::entries;
  if (sc->err) return;
  const int Hz = d.hz, ny = d.ny;
  const Col c = flat_col<T, NT>(d.nx * Hz, Hz, bx);
  const int j = threadIdx.x / T;
  const cpx<F>* src = A + (int64_t)c.a * ny * Hz + c.b;
  cpx<F> x[R];
#pragma unroll
  for (int n = 0; n < R; ++n) {
    const int y = Q * n + j;
    x[n] = cpx<F>{0, 0};
    if (n < R / 2 && c.ok && y < ny) x[n] = src[(int64_t)y * Hz];
  }
  for (int i = threadIdx.x; i < M; i += NT) twl[i] = tw[i];
  __syncthreads();
  reg_fft<F, R, Q, T, -1, R / 2>(x, s, c.t, j, twl);
  if (!c.ok) return;
  cpx<F>* dst = B + (int64_t)c.a * M * Hz + c.b;
#pragma unroll
  for (int q = 0; q < R; ++q) dst[(int64_t)(j + Q * q) * Hz] = x[q];
}

// ---------------------------------------------------------------------------
// z pass: pairs of real rows -> one complex FFT of length Mz -> two half spectra
template <typename F, int M, int CIN, int V>
__device__ __forceinline__ void zfwd_body(const void* __restrict__ rho, cpx<F>* __restrict__ A, MeshDims d,
                                          const KickScalars* __restrict__ sc, const cpx<F>* __restrict__ tw, int ncopy,
                                          const int bx, const int nbx) {
  if constexpr (reg_plan<F, M, V>()) {
    zfwd_reg<F, M, CIN>(rho, A, d, sc, tw, ncopy, bx);
    return;
  }
  using TL = Tile<F, M, V>;
  constexpr int T = TL::T, S = TL::S;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  if (sc->err) return;
  const cpx<F>* twl = stage_tw<F, M, kNT, TL::stw>(s + (size_t)M * S, tw);
  const int nz = d.nz, Hz = d.hz;
  const int64_t R = (int64_t)d.nx * d.ny;
  const int64_t ncell = R * nz;
  const int64_t row0 = (int64_t)bx * (2 * T);
  const double iq = sc->inv_qscale;
  // zero the padding rows m in [nz, M)
  for (int i = threadIdx.x; i < (M - nz) * T; i += kNT) {
    const int t = i % T, m = nz + i / T;
    s[m * S + t] = cpx<F>{0, 0};
  }
  // 2T rows of nz <= M/2 charges
  constexpr int NL = per_thread(T * M, kNT);
  F v[NL];
  int off[NL];
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int i = threadIdx.x + k * kNT;
    v[k] = 0;
    off[k] = -1;
    if (i < 2 * T * nz) {
      const int rl = i / nz, m = i - rl * nz;
      const int64_t row = row0 + rl;
      if (row < R) v[k] = (F)charge_at<CIN, F>(rho, row * nz + m, ncell, iq, ncopy);
      off[k] = (m * S + (rl >> 1)) * 2 + (rl & 1);
    }
  }
#pragma unroll
  for (int k = 0; k < NL; ++k)
    if (off[k] >= 0) ((F*)s)[off[k]] = v[k];
  __syncthreads();
  lds_fft<F, M, T, kNT, -1, S, TL::RSEL>(s, twl);
  constexpr int NS = per_thread(T * (M / 2 + 1), kNT);
#pragma unroll
  for (int k2 = 0; k2 < NS; ++k2) {
    const int i = threadIdx.x + k2 * kNT;
    if (i >= T * Hz) break;
    const int t = i / Hz, k = i - t * Hz;
    const int64_t ra = row0 + 2 * t, rb = ra + 1;
    if (ra >= R) continue;
    const cpx<F> z = s[k * S + t];
    const cpx<F> w = s[((M - k) & (M - 1)) * S + t];
    const F h = (F)0.5;
    A[ra * Hz + k] = cpx<F>{(z.x + w.x) * h, (z.y - w.y) * h};
    if (rb < R) A[rb * Hz + k] = cpx<F>{(z.y + w.y) * h, (w.x - z.x) * h};
  }
}

template <typename F, int M, int CIN, int V = 0>
__global__ void __launch_bounds__((pass_threads<F, M, V>())) k_zfwd(const void* __restrict__ rho, cpx<F>* __restrict__ A, MeshDims d,
                                              const KickScalars* __restrict__ sc,
                                              const cpx<F>* __restrict__ tw, int ncopy = 1) {
  zfwd_body<F, M, CIN, V>(rho, A, d, sc, tw, ncopy, blockIdx.x, gridDim.x);
}

// y pass forward: A[x][y<ny][kz] -> B[x][ky<My][kz]
template <typename F, int M, int V>
__device__ __forceinline__ void yfwd_body(const cpx<F>* __restrict__ A, cpx<F>* __restrict__ B, MeshDims d,
                                          const KickScalars* __restrict__ sc, const cpx<F>* __restrict__ tw,
                                          const int bx, const int nbx) {
  if constexpr (reg_plan<F, M, V>()) {
    yfwd_reg<F, M>(A, B, d, sc, tw, bx);
    return;
  }
  using TL = Tile<F, M, V>;
  constexpr int T = TL::T, S = TL::S;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  if (sc->err) return;
  const cpx<F>* twl = stage_tw<F, M, kNT, TL::stw>(s + (size_t)M * S, tw);
  const int Hz = d.hz, ny = d.ny;
  const Col c = flat_col<T, kNT>(d.nx * Hz, Hz, bx);
  const cpx<F>* src = A + (int64_t)c.a * ny * Hz + c.b;
  constexpr int NL = per_thread(M * T, kNT);
  cpx<F> v[NL];
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int y = (threadIdx.x + k * kNT) / T;
    v[k] = cpx<F>{0, 0};
    if (y < ny && c.ok) v[k] = src[(int64_t)y * Hz];
  }
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int y = (threadIdx.x + k * kNT) / T;
    if (y < M) s[y * S + c.t] = v[k];
  }
  __syncthreads();
  lds_fft<F, M, T, kNT, -1, S, TL::RSEL>(s, twl);
  if (!c.ok) return;
  cpx<F>* dst = B + (int64_t)c.a * M * Hz + c.b;
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int ky = (threadIdx.x + k * kNT) / T;
    if (ky < M) dst[(int64_t)ky * Hz] = s[ky * S + c.t];
  }
}

template <typename F, int M, int V = 0>
__global__ void __launch_bounds__((pass_threads<F, M, V>())) k_yfwd(const cpx<F>* __restrict__ A, cpx<F>* __restrict__ B, MeshDims d,
                                              const KickScalars* __restrict__ sc,
                                              const cpx<F>* __restrict__ tw) {
  yfwd_body<F, M, V>(A, B, d, sc, tw, blockIdx.x, gridDim.x);
}

// Plan 4 x pass of the potential: thread j of column t holds rows x = Q*n + j
// (n < R/2 non-zero: nx <= M/2), forward transform, times the kernel
// spectrum, inverse, rows x = j + Q*q < nx kept.  In place on B.
template <typename F, int M>
__device__ __forceinline__ void xconv_reg(cpx<F>* __restrict__ B, const F* __restrict__ G, MeshDims d,
                                          const KickScalars* __restrict__ sc, const cpx<F>* __restrict__ tw,
                                          const int bx) {
  constexpr int R = reg_r<M>(), Q = reg_q<M>(), T = Tile<F, M>::T, NT = Q * T;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  cpx<F>* twl = s + RegX<R, Q, T>::entries;
  if (sc->err) return;
  const int Hz = d.hz, nx = d.nx, My = d.my, Hy = d.hy;
  const int64_t cs = (int64_t)My * Hz;  // stride between x planes
  const Col c = flat_col<T, NT>(My * Hz, Hz, bx);  // a = ky, b = kz
  const int j = threadIdx.x / T;
  cpx<F>* col = B + (int64_t)c.a * Hz + c.b;
  cpx<F> x[R];
#pragma unroll
  for (int n = 0; n < R; ++n) {
    const int row = Q * n + j;
    x[n] = cpx<F>{0, 0};
    if (n < R / 2 && c.ok && row < nx) x[n] = col[row * cs];
  }
  for (int i = threadIdx.x; i < M; i += NT) twl[i] = tw[i];
  __syncthreads();
  reg_fft<F, R, Q, T, -1, R / 2>(x, s, c.t, j, twl);
  const int fy = c.a <= My - c.a ? c.a : My - c.a;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int kx = j + Q * q;
    const int fx = kx <= M - kx ? kx : M - kx;
    const F g = c.ok ? G[((int64_t)fx * Hy + fy) * Hz + c.b] : (F)0;
    x[q] = cscale(x[q], g);
  }
  reg_fft<F, R, Q, T, +1, R>(x, s, c.t, j, twl);
  if (!c.ok) return;
#pragma unroll
  for (int q = 0; q < R / 2; ++q) {
    const int row = j + Q * q;
    if (row < nx) col[row * cs] = x[q];
  }
}

// Plan 4 y inverse of the potential: all M spectrum rows in, rows
// y = j + Q*q < ny out.
template <typename F, int M>
__device__ __forceinline__ void yinv_reg(const cpx<F>* __restrict__ B, cpx<F>* __restrict__ A, MeshDims d,
                                         const KickScalars* __restrict__ sc, const cpx<F>* __restrict__ tw,
                                         const int bx) {
  constexpr int R = reg_r<M>(), Q = reg_q<M>(), T = Tile<F, M>::T, NT = Q * T;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  cpx<F>* twl = s + RegX<R, Q, T>::entries;
  if (sc->err) return;
  const int Hz = d.hz, ny = d.ny;
  const Col c = flat_col<T, NT>(d.nx * Hz, Hz, bx);  // a = x, b = kz
  const int j = threadIdx.x / T;
  const cpx<F>* src = B + (int64_t)c.a * M * Hz + c.b;
  cpx<F> x[R];
#pragma unroll
  for (int n = 0; n < R; ++n) {
    x[n] = cpx<F>{0, 0};
    if (c.ok) x[n] = src[(int64_t)(Q * n + j) * Hz];
  }
  for (int i = threadIdx.x; i < M; i += NT) twl[i] = tw[i];
  __syncthreads();
  reg_fft<F, R, Q, T, +1, R>(x, s, c.t, j, twl);
  if (!c.ok) return;
  cpx<F>* dst = A + (int64_t)c.a * ny * Hz + c.b;
#pragma unroll
  for (int q = 0; q < R / 2; ++q) {
    const int y = j + Q * q;
    if (y < ny) dst[(int64_t)y * Hz] = x[q];
  }
}

// x pass: forward FFT along x, multiply by the kernel spectrum, inverse FFT,
// keep x < nx.  In place on B.
template <typename F, int M, int V = 0>
__device__ __forceinline__ void xconv_body(cpx<F>* __restrict__ B, const F* __restrict__ G, MeshDims d,
                                           const KickScalars* __restrict__ sc, const cpx<F>* __restrict__ tw,
                                           const int bx, const int nbx) {
  if constexpr (reg_plan<F, M, V>()) {
    xconv_reg<F, M>(B, G, d, sc, tw, bx);
    return;
  }
  constexpr int T = Tile<F, M>::T, S = Tile<F, M>::S;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  if (sc->err) return;
  const cpx<F>* twl = stage_tw<F, M, kNT, Tile<F, M>::stw>(s + (size_t)M * S, tw);
  const int Hz = d.hz, nx = d.nx, My = d.my, Hy = d.hy;
  const int64_t cs = (int64_t)My * Hz;  // stride between x planes
  const Col c = flat_col<T, kNT>(My * Hz, Hz, bx);  // a = ky, b = kz
  cpx<F>* col = B + (int64_t)c.a * Hz + c.b;
  constexpr int NL = per_thread(M * T, kNT);
  cpx<F> v[NL];
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int x = (threadIdx.x + k * kNT) / T;
    v[k] = cpx<F>{0, 0};
    if (x < nx && c.ok) v[k] = col[x * cs];
  }
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int x = (threadIdx.x + k * kNT) / T;
    if (x < M) s[x * S + c.t] = v[k];
  }
  // the kernel spectrum for this thread's values, loaded while the FFT runs
  const int fy = c.a <= My - c.a ? c.a : My - c.a;
  F g[NL];
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int kx = (threadIdx.x + k * kNT) / T;
    const int fx = kx <= M - kx ? kx : M - kx;
    g[k] = (c.ok && kx < M) ? G[((int64_t)fx * Hy + fy) * Hz + c.b] : (F)0;
  }
  __syncthreads();
  lds_fft<F, M, T, kNT, -1>(s, twl);
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int kx = (threadIdx.x + k * kNT) / T;
    if (kx < M) s[kx * S + c.t] = cscale(s[kx * S + c.t], g[k]);
  }
  __syncthreads();
  lds_fft<F, M, T, kNT, +1>(s, twl);
  if (!c.ok) return;
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int x = (threadIdx.x + k * kNT) / T;
    if (x < nx) col[x * cs] = s[x * S + c.t];
  }
}

template <typename F, int M, int V = 0>
__global__ void __launch_bounds__((pass_threads<F, M, V>())) k_xconv(cpx<F>* __restrict__ B, const F* __restrict__ G,
                                                                     MeshDims d, const KickScalars* __restrict__ sc,
                                                                     const cpx<F>* __restrict__ tw) {
  xconv_body<F, M, V>(B, G, d, sc, tw, blockIdx.x, gridDim.x);
}

// y pass inverse: B[x][ky][kz] -> A[x][y<ny][kz]
template <typename F, int M, int V = 0>
__device__ __forceinline__ void yinv_body(const cpx<F>* __restrict__ B, cpx<F>* __restrict__ A, MeshDims d,
                                          const KickScalars* __restrict__ sc, const cpx<F>* __restrict__ tw,
                                          const int bx, const int nbx) {
  if constexpr (reg_plan<F, M, V>()) {
    yinv_reg<F, M>(B, A, d, sc, tw, bx);
    return;
  }
  constexpr int T = Tile<F, M>::T, S = Tile<F, M>::S;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  if (sc->err) return;
  const cpx<F>* twl = stage_tw<F, M, kNT, Tile<F, M>::stw>(s + (size_t)M * S, tw);
  const int Hz = d.hz, ny = d.ny;
  const Col c = flat_col<T, kNT>(d.nx * Hz, Hz, bx);  // a = x, b = kz
  const cpx<F>* src = B + (int64_t)c.a * M * Hz + c.b;
  constexpr int NL = per_thread(M * T, kNT);
  cpx<F> v[NL];
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int ky = (threadIdx.x + k * kNT) / T;
    v[k] = cpx<F>{0, 0};
    if (c.ok && ky < M) v[k] = src[(int64_t)ky * Hz];
  }
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int ky = (threadIdx.x + k * kNT) / T;
    if (ky < M) s[ky * S + c.t] = v[k];
  }
  __syncthreads();
  lds_fft<F, M, T, kNT, +1>(s, twl);
  if (!c.ok) return;
  cpx<F>* dst = A + (int64_t)c.a * ny * Hz + c.b;
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int y = (threadIdx.x + k * kNT) / T;
    if (y < ny) dst[(int64_t)y * Hz] = s[y * S + c.t];
  }
}

template <typename F, int M, int V = 0>
__global__ void __launch_bounds__((pass_threads<F, M, V>())) k_yinv(const cpx<F>* __restrict__ B,
                                                                    cpx<F>* __restrict__ A, MeshDims d,
                                                                    const KickScalars* __restrict__ sc,
                                                                    const cpx<F>* __restrict__ tw) {
  yinv_body<F, M, V>(B, A, d, sc, tw, blockIdx.x, gridDim.x);
}

// z pass inverse: two Hermitian half spectra -> one complex inverse FFT ->
// two real rows of the potential, z < nz
template <typename F, int M, bool DERIV, int V = 0>
__device__ __forceinline__ void zinv_rows(const cpx<F>* __restrict__ A, F* __restrict__ phi, const MeshDims& d,
                                          const KickScalars* __restrict__ sc, const cpx<F>* __restrict__ tw,
                                          const cpx<double>* __restrict__ twd, cpx<F>* s, const int bx) {
  using TL = Tile<F, M, V>;
  constexpr int T = TL::T, S = TL::S;
  const cpx<F>* twl = stage_tw<F, M, kNT, TL::stw>(s + (size_t)M * S, tw);
  const int nz = d.nz, Hz = d.hz;
  const int64_t R = (int64_t)d.nx * d.ny;
  const int64_t row0 = (int64_t)bx * (2 * T);
  // T pairs of half spectra (Hz = M/2 + 1 each)
  constexpr int NL = per_thread(T * (M / 2 + 1), kNT);
  cpx<F> a[NL], b[NL];
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    const int i = threadIdx.x + j * kNT;
    a[j] = cpx<F>{0, 0};
    b[j] = cpx<F>{0, 0};
    if (i < T * Hz) {
      const int t = i / Hz, k = i - t * Hz;
      const int64_t ra = row0 + 2 * t, rb = ra + 1;
      if (ra < R) a[j] = A[ra * Hz + k];
      if (rb < R) b[j] = A[rb * Hz + k];
    }
  }
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    const int i = threadIdx.x + j * kNT;
    if (i >= T * Hz) break;
    const int t = i / Hz, k = i - t * Hz;
    cpx<F> aa = a[j], bb = b[j];
    if constexpr (DERIV) {
      const cpx<F> f = dfac<F>(twd, k, 1.0 / sc->steps[2]);
      aa = cmul(aa, f);
      bb = cmul(bb, f);
    }
    s[k * S + t] = cpx<F>{aa.x - bb.y, aa.y + bb.x};
    if (k > 0 && k < M / 2) s[(M - k) * S + t] = cpx<F>{aa.x + bb.y, bb.x - aa.y};
  }
  __syncthreads();
  lds_fft<F, M, T, kNT, +1, S, TL::RSEL>(s, twl);
  constexpr int NS = per_thread(T * M, kNT);  // 2T rows of nz <= M/2
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const int i = threadIdx.x + j * kNT;
    if (i >= 2 * T * nz) break;
    const int rl = i / nz, z = i - rl * nz;
    const int64_t row = row0 + rl;
    if (row >= R) continue;
    const cpx<F> v = s[z * S + (rl >> 1)];
    phi[row * nz + z] = (rl & 1) ? v.y : v.x;
  }
}

// Plan 4 z inverse: thread (t, j) (lanes j fastest) assembles the values
// m = Q*n + j of the complex column a + i b (half spectra of rows 2t, 2t+1)
// straight from global memory (m > M/2 from the conjugates at M - m, as
// zinv_rows mirrors them), the z field's difference factor applied first
// (DERIV); rows z = j + Q*q < nz <= M/2 are written.
template <typename F, int M, bool DERIV>
__device__ __forceinline__ void zinv_rows_reg(const cpx<F>* __restrict__ A, F* __restrict__ phi, const MeshDims& d,
                                              const KickScalars* __restrict__ sc, const cpx<F>* __restrict__ tw,
                                              const cpx<double>* __restrict__ twd, cpx<F>* s, const int bx) {
  constexpr int R = reg_r<M>(), Q = reg_q<M>(), T = Tile<F, M>::T, NT = Q * T;
  cpx<F>* twl = s + RegX<R, Q, T>::entries;
  cpx<F>* dl = twl + M;
  const int nz = d.nz, Hz = d.hz;
  const int64_t nrow = (int64_t)d.nx * d.ny;
  const int64_t row0 = (int64_t)bx * (2 * T);
  const int j = threadIdx.x % Q, t = threadIdx.x / Q;
  const int64_t ra = row0 + 2 * t, rb = ra + 1;
  cpx<F> a[R], b[R];
#pragma unroll
  for (int n = 0; n < R; ++n) {
    const int m = Q * n + j;
    const int k = m <= M / 2 ? m : M - m;
    a[n] = cpx<F>{0, 0};
    b[n] = cpx<F>{0, 0};
    if (ra < nrow) a[n] = A[ra * Hz + k];
    if (rb < nrow) b[n] = A[rb * Hz + k];
  }
  for (int i = threadIdx.x; i < M; i += NT) {
    twl[i] = tw[i];
    if constexpr (DERIV) dl[i] = dfac<F>(twd, i, 1.0 / sc->steps[2]);
  }
  __syncthreads();
  cpx<F> x[R];
#pragma unroll
  for (int n = 0; n < R; ++n) {
    const int m = Q * n + j;
    const int k = m <= M / 2 ? m : M - m;
    cpx<F> aa = a[n], bb = b[n];
    if constexpr (DERIV) {
      const cpx<F> f = dl[k];
      aa = cmul(aa, f);
      bb = cmul(bb, f);
    }
    x[n] = m <= M / 2 ? cpx<F>{aa.x - bb.y, aa.y + bb.x} : cpx<F>{aa.x + bb.y, bb.x - aa.y};
  }
  reg_fft<F, R, Q, T, +1, R, true>(x, s, t, j, twl);
#pragma unroll
  for (int q = 0; q < R / 2; ++q) {
    const int z = j + Q * q;
    if (z < nz) {
      if (ra < nrow) phi[ra * nz + z] = x[q].x;
      if (rb < nrow) phi[rb * nz + z] = x[q].y;
    }
  }
}

template <typename F, int M, bool DERIV, int V = 0>
__device__ __forceinline__ void zinv_body(const cpx<F>* __restrict__ A, F* __restrict__ phi, MeshDims d,
                                          const KickScalars* __restrict__ sc, const cpx<F>* __restrict__ tw,
                                          const cpx<double>* __restrict__ twd, const int bx, const int nbx) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  if (sc->err) return;
  if constexpr (reg_plan<F, M, V>())
    zinv_rows_reg<F, M, DERIV>(A, phi, d, sc, tw, twd, (cpx<F>*)smem_raw, bx);
  else
    zinv_rows<F, M, DERIV>(A, phi, d, sc, tw, twd, (cpx<F>*)smem_raw, bx);
}

template <typename F, int M, bool DERIV = false, int V = 0>
__global__ void __launch_bounds__((pass_threads<F, M, V>())) k_zinv(const cpx<F>* __restrict__ A,
                                                                    F* __restrict__ phi, MeshDims d,
                                                                    const KickScalars* __restrict__ sc,
                                                                    const cpx<F>* __restrict__ tw,
                                                                    const cpx<double>* __restrict__ twd = nullptr) {
  zinv_body<F, M, DERIV, V>(A, phi, d, sc, tw, twd, blockIdx.x, gridDim.x);
}

// Mixed form: the three field components in one launch (blockIdx.y), the
// third (z) with the spectral difference factor
template <int M, int V>
__device__ __forceinline__ void zinv3_body(const cpx<float>* __restrict__ A, int64_t nyz, float* __restrict__ E,
                                           int64_t ncell, const MeshDims& d, const KickScalars* __restrict__ sc,
                                           const cpx<float>* __restrict__ tw, const cpx<double>* __restrict__ twd,
                                           const int comp, const int bx) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  if (sc->err) return;
  cpx<float>* s = (cpx<float>*)smem_raw;
  if constexpr (reg_plan<float, M, V>()) {
    if (comp == 2)
      zinv_rows_reg<float, M, true>(A + 2 * nyz, E + 2 * ncell, d, sc, tw, twd, s, bx);
    else
      zinv_rows_reg<float, M, false>(A + comp * nyz, E + comp * ncell, d, sc, tw, twd, s, bx);
  } else {
    if (comp == 2)
      zinv_rows<float, M, true, V>(A + 2 * nyz, E + 2 * ncell, d, sc, tw, twd, s, bx);
    else
      zinv_rows<float, M, false, V>(A + comp * nyz, E + comp * ncell, d, sc, tw, twd, s, bx);
  }
}

template <int M, int V = 0>
__global__ void __launch_bounds__((pass_threads<float, M, V>())) k_zinv3(const cpx<float>* __restrict__ A, int64_t nyz, float* __restrict__ E,
                                               int64_t ncell, MeshDims d, const KickScalars* __restrict__ sc,
                                               const cpx<float>* __restrict__ tw,
                                               const cpx<double>* __restrict__ twd) {
  zinv3_body<M, V>(A, nyz, E, ncell, d, sc, tw, twd, blockIdx.y, blockIdx.x);
}

// columns per workgroup of the kernels that carry two or three fields
// The two field kernels run 512 threads per workgroup: half the values per
// thread of a 256-thread workgroup, so twice the waves fit a CU (tracking
// step at 127^3: 0.894 -> 0.887 ms).
constexpr int kNTF = 512;
// plan 3: 256 threads, so every thread gets a radix-16 butterfly (two passes
// for 256 points instead of three of radix 8).
// plan 4: see reg_plan.
template <int M, int V>
constexpr bool reg_field() {
  return reg_plan<float, M, V>();
}
template <int M, int V>
constexpr int field_threads() {
  return reg_field<M, V>() ? reg_q<M>() * Tile<float, M>::T : (V == 3 ? 256 : kNTF);
}

template <int M, int NT = kNT, int V = 0>
struct FieldTile {
  static constexpr bool reg = reg_field<M, V>();
  static constexpr int T = Tile<float, M, V>::T;  // columns per workgroup
  static constexpr int S = T + 1;
  static constexpr int PER = (M * T + NT - 1) / NT;  // spectrum values held per thread
  static constexpr bool stw = Tile<float, M, V>::stw;
  static constexpr int RSEL = Tile<float, M, V>::RSEL;
  static constexpr int R = reg ? reg_r<M>() : 1;
  static constexpr int Q = reg ? reg_q<M>() : 1;
  // plan 4: the exchange buffer, the twiddle table, the difference factors
  static constexpr size_t lds =
      reg ? (size_t)(RegX<R, Q, T>::entries + 2 * M) * sizeof(cpx<float>) : Tile<float, M, V>::lds_bytes;
};

// Plan 4 set-up: twiddles W_M^k and difference factors dfac(k) for k < M
// into LDS (one evaluation per workgroup instead of one per element)
template <int M, int NT>
__device__ __forceinline__ void stage_reg_tables(cpx<float>* twl, cpx<float>* dl, const cpx<float>* __restrict__ tw,
                                                 const cpx<double>* __restrict__ twd, double inv_h) {
  for (int i = threadIdx.x; i < M; i += NT) {
    twl[i] = tw[i];
    dl[i] = dfac<float>(twd, i, inv_h);
  }
}

// Plan 4 x pass (see xconv_field_body): thread j of column t holds rows
// x = Q*n + j; x < nx <= M/2 (M >= 2 nx - 1 and even), so only n < R/2 can
// be non-zero and only outputs x = j + Q*q with q < R/2 are kept.
template <int M>
__device__ __forceinline__ void xconv_field_reg(const cpx<float>* __restrict__ B, cpx<float>* __restrict__ Qx,
                                                cpx<float>* __restrict__ P, const float* __restrict__ G, MeshDims d,
                                                const KickScalars* __restrict__ sc, const cpx<float>* __restrict__ tw,
                                                const cpx<double>* __restrict__ twd, const int bx) {
  using F = float;
  using FT = FieldTile<M, field_threads<M, 4>(), 4>;
  constexpr int R = FT::R, Q = FT::Q, T = FT::T, NT = field_threads<M, 4>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  cpx<F>* twl = s + RegX<R, Q, T>::entries;
  cpx<F>* dl = twl + M;
  if (sc->err) return;
  const int Hz = d.hz, nx = d.nx, My = d.my, Hy = d.hy;
  const int64_t cs = (int64_t)My * Hz;
  const Col c = flat_col<T, NT>(My * Hz, Hz, bx);  // a = ky, b = kz
  const int j = threadIdx.x / T;
  const int64_t off = (int64_t)c.a * Hz + c.b;
  cpx<F> x[R];
#pragma unroll
  for (int n = 0; n < R; ++n) {
    const int row = Q * n + j;
    x[n] = cpx<F>{0, 0};
    if (n < R / 2 && c.ok && row < nx) x[n] = B[row * cs + off];
  }
  stage_reg_tables<M, NT>(twl, dl, tw, twd, 1.0 / sc->steps[0]);
  __syncthreads();
  reg_fft<F, R, Q, T, -1, R / 2>(x, s, c.t, j, twl);
  // x[q]: the spectrum at kx = j + Q*q; v = S*G kept for the potential
  const int fy = c.a <= My - c.a ? c.a : My - c.a;
  cpx<F> v[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int kx = j + Q * q;
    const int fx = kx <= M - kx ? kx : M - kx;
    const F g = c.ok ? G[((int64_t)fx * Hy + fy) * Hz + c.b] : (F)0;
    v[q] = cscale(x[q], g);
    x[q] = cmul(v[q], dl[kx]);
  }
  reg_fft<F, R, Q, T, +1, R>(x, s, c.t, j, twl);
  if (c.ok) {
#pragma unroll
    for (int q = 0; q < R / 2; ++q) {
      const int row = j + Q * q;
      if (row < nx) Qx[row * cs + off] = x[q];
    }
  }
  reg_fft<F, R, Q, T, +1, R>(v, s, c.t, j, twl);
  if (c.ok) {
#pragma unroll
    for (int q = 0; q < R / 2; ++q) {
      const int row = j + Q * q;
      if (row < nx) P[row * cs + off] = v[q];
    }
  }
}

// Plan 4 y inverse (see yinv_field_body): all M spectrum rows are inputs,
// rows y = j + Q*q < ny <= M/2 are outputs.
template <int M>
__device__ __forceinline__ void yinv_field_reg(const cpx<float>* __restrict__ Qx, const cpx<float>* __restrict__ P,
                                               cpx<float>* __restrict__ Ax, cpx<float>* __restrict__ Ay,
                                               cpx<float>* __restrict__ Ap, MeshDims d,
                                               const KickScalars* __restrict__ sc, const cpx<float>* __restrict__ tw,
                                               const cpx<double>* __restrict__ twd, const int bx) {
  using F = float;
  using FT = FieldTile<M, field_threads<M, 4>(), 4>;
  constexpr int R = FT::R, Q = FT::Q, T = FT::T, NT = field_threads<M, 4>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  cpx<F>* twl = s + RegX<R, Q, T>::entries;
  cpx<F>* dl = twl + M;
  if (sc->err) return;
  const int Hz = d.hz, ny = d.ny;
  const Col c = flat_col<T, NT>(d.nx * Hz, Hz, bx);  // a = x, b = kz
  const int j = threadIdx.x / T;
  const int64_t src = (int64_t)c.a * M * Hz + c.b;
  cpx<F> q[R], p[R];
#pragma unroll
  for (int n = 0; n < R; ++n) {
    q[n] = cpx<F>{0, 0};
    p[n] = cpx<F>{0, 0};
    if (c.ok) {
      const int64_t o = src + (int64_t)(Q * n + j) * Hz;
      q[n] = Qx[o];
      p[n] = P[o];
    }
  }
  stage_reg_tables<M, NT>(twl, dl, tw, twd, 1.0 / sc->steps[1]);
  __syncthreads();
  const int64_t dst0 = (int64_t)c.a * ny * Hz + c.b;
  auto store = [&](cpx<F>* dst, const cpx<F>(&v)[R]) {
    if (!c.ok) return;
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      const int y = j + Q * k;
      if (y < ny) dst[dst0 + (int64_t)y * Hz] = v[k];
    }
  };
  reg_fft<F, R, Q, T, +1, R>(q, s, c.t, j, twl);
  store(Ax, q);
#pragma unroll
  for (int n = 0; n < R; ++n) q[n] = cmul(p[n], dl[Q * n + j]);
  reg_fft<F, R, Q, T, +1, R>(q, s, c.t, j, twl);
  store(Ay, q);
  reg_fft<F, R, Q, T, +1, R>(p, s, c.t, j, twl);
  store(Ap, p);
}

// Mixed form, x pass: forward FFT along x, times the kernel spectrum, then
// two inverse FFTs: Qx = IFFT(S * Dx) (the x field) and P = IFFT(S) (the
// potential, from which the y and z fields follow by their own factors).
// The product S*G stays in registers between the two inverse transforms, so
// LDS holds one tile of T columns (several workgroups per CU).
template <int M, int V>
__device__ __forceinline__ void xconv_field_lds(const cpx<float>* __restrict__ B, cpx<float>* __restrict__ Qx,
                                                 cpx<float>* __restrict__ P, const float* __restrict__ G, MeshDims d,
                                                 const KickScalars* __restrict__ sc, const cpx<float>* __restrict__ tw,
                                                 const cpx<double>* __restrict__ twd, const int bx, const int nbx) {
  constexpr int NTF = field_threads<M, V>();
  using F = float;
  constexpr int T = FieldTile<M, NTF, V>::T, S = FieldTile<M, NTF, V>::S, PER = FieldTile<M, NTF, V>::PER;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  if (sc->err) return;
  const cpx<F>* twl = stage_tw<F, M, NTF, FieldTile<M, NTF, V>::stw>(s + (size_t)M * S, tw);
  const int Hz = d.hz, nx = d.nx, My = d.my, Hy = d.hy;
  const int64_t cs = (int64_t)My * Hz;
  const Col c = flat_col<T, NTF>(My * Hz, Hz, bx);  // a = ky, b = kz
  const int64_t off = (int64_t)c.a * Hz + c.b;
  {
    cpx<F> u[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int x = (threadIdx.x + k * NTF) / T;
      u[k] = cpx<F>{0, 0};
      if (x < nx && c.ok) u[k] = B[x * cs + off];
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int x = (threadIdx.x + k * NTF) / T;
      if (x < M) s[x * S + c.t] = u[k];
    }
  }
  __syncthreads();
  lds_fft<F, M, T, NTF, -1, S, FieldTile<M, NTF, V>::RSEL>(s, twl);
  const int fy = c.a <= My - c.a ? c.a : My - c.a;
  const double inv_hx = 1.0 / sc->steps[0];
  cpx<F> v[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = threadIdx.x + k * NTF;
    v[k] = cpx<F>{0, 0};
    if (M * T % NTF == 0 || i < M * T) {
      const int kx = i / T;
      const int fx = kx <= M - kx ? kx : M - kx;
      const F g = c.ok ? G[((int64_t)fx * Hy + fy) * Hz + c.b] : (F)0;
      v[k] = cscale(s[kx * S + c.t], g);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = threadIdx.x + k * NTF;
    if (M * T % NTF == 0 || i < M * T) {
      const int kx = i / T;
      s[kx * S + c.t] = cmul(v[k], dfac<F>(twd, kx, inv_hx));
    }
  }
  __syncthreads();
  lds_fft<F, M, T, NTF, +1, S, FieldTile<M, NTF, V>::RSEL>(s, twl);
  if (c.ok) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int x = (threadIdx.x + k * NTF) / T;
      if (x < nx) Qx[x * cs + off] = s[x * S + c.t];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = threadIdx.x + k * NTF;
    if (M * T % NTF == 0 || i < M * T) s[(i / T) * S + c.t] = v[k];
  }
  __syncthreads();
  lds_fft<F, M, T, NTF, +1, S, FieldTile<M, NTF, V>::RSEL>(s, twl);
  if (c.ok) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int x = (threadIdx.x + k * NTF) / T;
      if (x < nx) P[x * cs + off] = s[x * S + c.t];
    }
  }
}

template <int M, int V>
__device__ __forceinline__ void xconv_field_body(const cpx<float>* __restrict__ B, cpx<float>* __restrict__ Qx,
                                                 cpx<float>* __restrict__ P, const float* __restrict__ G, MeshDims d,
                                                 const KickScalars* __restrict__ sc, const cpx<float>* __restrict__ tw,
                                                 const cpx<double>* __restrict__ twd, const int bx, const int nbx) {
  if constexpr (reg_field<M, V>())
    xconv_field_reg<M>(B, Qx, P, G, d, sc, tw, twd, bx);
  else
    xconv_field_lds<M, V>(B, Qx, P, G, d, sc, tw, twd, bx, nbx);
}

template <int M, int V = 0>
__global__ void __launch_bounds__((field_threads<M, V>())) k_xconv_field(const cpx<float>* __restrict__ B, cpx<float>* __restrict__ Qx,
                                                     cpx<float>* __restrict__ P, const float* __restrict__ G,
                                                     MeshDims d, const KickScalars* __restrict__ sc,
                                                     const cpx<float>* __restrict__ tw,
                                                     const cpx<double>* __restrict__ twd) {
  xconv_field_body<M, V>(B, Qx, P, G, d, sc, tw, twd, blockIdx.x, gridDim.x);
}

// Mixed form, y inverse: Ex' = IFFT_y(Qx), Ey' = IFFT_y(P * Dy), Phi' = IFFT_y(P)
// (all [nx][ny][Hz]); the z field is Dz * Phi', applied in the z inverse.
// Three transforms of one LDS tile in turn, P held in registers.
template <int M, int V>
__device__ __forceinline__ void yinv_field_lds(const cpx<float>* __restrict__ Qx, const cpx<float>* __restrict__ P,
                                                cpx<float>* __restrict__ Ax, cpx<float>* __restrict__ Ay,
                                                cpx<float>* __restrict__ Ap, MeshDims d,
                                                const KickScalars* __restrict__ sc, const cpx<float>* __restrict__ tw,
                                                const cpx<double>* __restrict__ twd, const int bx, const int nbx) {
  constexpr int NTF = field_threads<M, V>();
  using F = float;
  constexpr int T = FieldTile<M, NTF, V>::T, S = FieldTile<M, NTF, V>::S, PER = FieldTile<M, NTF, V>::PER;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  if (sc->err) return;
  const cpx<F>* twl = stage_tw<F, M, NTF, FieldTile<M, NTF, V>::stw>(s + (size_t)M * S, tw);
  const int Hz = d.hz, ny = d.ny;
  const Col c = flat_col<T, NTF>(d.nx * Hz, Hz, bx);  // a = x, b = kz
  const int64_t src = (int64_t)c.a * M * Hz + c.b;
  const double inv_hy = 1.0 / sc->steps[1];
  cpx<F> p[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = threadIdx.x + k * NTF;
    p[k] = cpx<F>{0, 0};
    if (M * T % NTF == 0 || i < M * T) {
      const int ky = i / T;
      cpx<F> q{0, 0};
      if (c.ok) {
        const int64_t o = src + (int64_t)ky * Hz;
        q = Qx[o];
        p[k] = P[o];
      }
      s[ky * S + c.t] = q;
    }
  }
  __syncthreads();
  lds_fft<F, M, T, NTF, +1, S, FieldTile<M, NTF, V>::RSEL>(s, twl);
  const int64_t dst0 = (int64_t)c.a * ny * Hz + c.b;
  auto store = [&](cpx<F>* dst) {
    if (!c.ok) return;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int y = (threadIdx.x + k * NTF) / T;
      if (y < ny) dst[dst0 + (int64_t)y * Hz] = s[y * S + c.t];
    }
  };
  store(Ax);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = threadIdx.x + k * NTF;
    if (M * T % NTF == 0 || i < M * T) s[(i / T) * S + c.t] = cmul(p[k], dfac<F>(twd, i / T, inv_hy));
  }
  __syncthreads();
  lds_fft<F, M, T, NTF, +1, S, FieldTile<M, NTF, V>::RSEL>(s, twl);
  store(Ay);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = threadIdx.x + k * NTF;
    if (M * T % NTF == 0 || i < M * T) s[(i / T) * S + c.t] = p[k];
  }
  __syncthreads();
  lds_fft<F, M, T, NTF, +1, S, FieldTile<M, NTF, V>::RSEL>(s, twl);
  store(Ap);
}

template <int M, int V>
__device__ __forceinline__ void yinv_field_body(const cpx<float>* __restrict__ Qx, const cpx<float>* __restrict__ P,
                                                cpx<float>* __restrict__ Ax, cpx<float>* __restrict__ Ay,
                                                cpx<float>* __restrict__ Ap, MeshDims d,
                                                const KickScalars* __restrict__ sc, const cpx<float>* __restrict__ tw,
                                                const cpx<double>* __restrict__ twd, const int bx, const int nbx) {
  if constexpr (reg_field<M, V>())
    yinv_field_reg<M>(Qx, P, Ax, Ay, Ap, d, sc, tw, twd, bx);
  else
    yinv_field_lds<M, V>(Qx, P, Ax, Ay, Ap, d, sc, tw, twd, bx, nbx);
}

template <int M, int V = 0>
__global__ void __launch_bounds__((field_threads<M, V>())) k_yinv_field(const cpx<float>* __restrict__ Qx,
                                                    const cpx<float>* __restrict__ P, cpx<float>* __restrict__ Ax,
                                                    cpx<float>* __restrict__ Ay, cpx<float>* __restrict__ Ap,
                                                    MeshDims d, const KickScalars* __restrict__ sc,
                                                    const cpx<float>* __restrict__ tw,
                                                    const cpx<double>* __restrict__ twd) {
  yinv_field_body<M, V>(Qx, P, Ax, Ay, Ap, d, sc, tw, twd, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Integrated Green's function (ocelot/cpbd/sc.py:109-133), FP64.
// K[i,j,k] = 8-corner alternating difference of F at the cell faces
// x_i = i*hx - hx/2 (i = 0..N).
__device__ __forceinline__ double igf_F(double x, double y, double z) {
  const double r = __builtin_sqrt(x * x + y * y + z * z);
  return -x * x * 0.5 * atan(y * z / (x * r)) + y * z * log(x + r) - y * y * 0.5 * atan(z * x / (y * r)) +
         z * x * log(y + r) - z * z * 0.5 * atan(x * y / (z * r)) + x * y * log(z + r);
}

// Two passes: F once at every node (i, j, k) <= (nx, ny, nz) of the cell
// faces, then the 8-corner difference per cell from those values (the
// neighbours come from L2).  A tiled single pass (4x8x8 cells from a 5x9x9
// F tile in LDS) evaluates 58% of F twice and leaves most threads idle while
// the tile is filled.
__device__ __forceinline__ void igf_nodes_body(double* __restrict__ Fn, MeshDims d, const KickScalars* __restrict__ sc,
                                               int force, const int bx, const int nbx) {
  if (sc->err || !(force || sc->rebuild)) return;
  const double hx = sc->steps[0], hy = sc->steps[1], hz = sc->steps[2];
  // 32-bit index arithmetic: meshes have at most 513^3 nodes
  const unsigned sz = d.nz + 1, sy = d.ny + 1;
  const unsigned nn = (unsigned)(d.nx + 1) * sy * sz;
  for (unsigned f = bx * 256u + threadIdx.x; f < nn; f += nbx * 256u) {
    const unsigned row = f / sz, k = f - row * sz, i = row / sy, j = row - i * sy;
    Fn[f] = igf_F(hx * (int)i - hx / 2, hy * (int)j - hy / 2, hz * (int)k - hz / 2);
  }
}

__global__ void __launch_bounds__(256) k_igf_nodes(double* __restrict__ Fn, MeshDims d,
                                                   const KickScalars* __restrict__ sc, int force) {
  igf_nodes_body(Fn, d, sc, force, blockIdx.x, gridDim.x);
}

// A work item is one (y, z) cell and a chunk of kIgfChunk x planes: walking up
// in x, the four F values of plane i+1 become those of plane i for the next
// cell, so a cell costs four loads instead of eight (lanes step in z: the
// loads are coalesced).  The eight terms are summed in the same order as
// one cell at a time, so K is the same bit for bit.
constexpr int kIgfChunk = 16;
// work items of igf_diff_body (host and device)
__host__ __device__ inline int64_t igf_diff_items(const MeshDims& d) {
  return (int64_t)((d.nx + kIgfChunk - 1) / kIgfChunk) * d.ny * d.nz;
}

__device__ __forceinline__ void igf_diff_body(const double* __restrict__ Fn, double* __restrict__ K, MeshDims d,
                                              const KickScalars* __restrict__ sc, int force, const int bx,
                                              const int nbx) {
  if (sc->err || !(force || sc->rebuild)) return;
  const unsigned nx = d.nx, ny = d.ny, nz = d.nz;
  const unsigned sy = nz + 1, sx = (ny + 1) * (nz + 1);
  const unsigned nchunk = (nx + kIgfChunk - 1) / kIgfChunk;
  const unsigned nwork = nchunk * ny * nz;
  for (unsigned o = bx * 256u + threadIdx.x; o < nwork; o += nbx * 256u) {
    const unsigned row = o / nz, k = o - row * nz, ch = row / ny, j = row - ch * ny;
    const unsigned i0 = ch * kIgfChunk, i1 = i0 + kIgfChunk < nx ? i0 + kIgfChunk : nx;
    const double* f = Fn + (i0 * sx + j * sy + k);
    // plane i: F(i, j+q, k+r) as p[q][r]
    double p00 = f[0], p01 = f[1], p10 = f[sy], p11 = f[sy + 1];
    double* out = K + ((size_t)i0 * ny + j) * nz + k;
    for (unsigned i = i0; i < i1; ++i) {
      f += sx;
      const double n00 = f[0], n01 = f[1], n10 = f[sy], n11 = f[sy + 1];
      // F111 - F011 - F101 + F001 - F110 + F010 + F100 - F000 (F(x, y, z) offsets)
      *out = n11 - p11 - n01 + p01 - n10 + p10 + n00 - p00;
      out += (size_t)ny * nz;
      p00 = n00;
      p01 = n01;
      p10 = n10;
      p11 = n11;
    }
  }
}

__global__ void __launch_bounds__(256) k_igf_diff(const double* __restrict__ Fn, double* __restrict__ K, MeshDims d,
                                                  const KickScalars* __restrict__ sc, int force) {
  igf_diff_body(Fn, K, d, sc, force, blockIdx.x, gridDim.x);
}

// Even transform along z of the real kernel, two rows per complex FFT:
// e[m] = K[m] (m < nz), e[M-m] = K[m] (1 <= m < nz); both spectra are real,
// so row a is the real part and row b the imaginary part of the output.
template <int M>
__device__ __forceinline__ void zcos_body(const double* __restrict__ K, double* __restrict__ out, MeshDims d,
                                          const KickScalars* __restrict__ sc, const cpx<double>* __restrict__ tw,
                                          int force, const int bx, const int nbx) {
  using F = double;
  constexpr int T = Tile<F, M>::T, S = Tile<F, M>::S;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  if (sc->err || !(force || sc->rebuild)) return;
  const cpx<F>* twl = stage_tw<F, M, kNT, Tile<F, M>::stw>(s + (size_t)M * S, tw);
  const int nz = d.nz, Hz = d.hz;
  const int64_t R = (int64_t)d.nx * d.ny;
  const int64_t row0 = (int64_t)bx * (2 * T);
  for (int i = threadIdx.x; i < M * T; i += kNT) s[(i / T) * S + (i % T)] = cpx<F>{0, 0};
  __syncthreads();
  constexpr int NL = per_thread(T * M, kNT);  // 2T rows of nz <= M/2 values
  F v[NL];
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    const int i = threadIdx.x + j * kNT;
    v[j] = 0;
    if (i < 2 * T * nz) {
      const int rl = i / nz, m = i - rl * nz;
      const int64_t row = row0 + rl;
      if (row < R) v[j] = K[row * nz + m];
    }
  }
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    const int i = threadIdx.x + j * kNT;
    if (i >= 2 * T * nz) break;
    const int rl = i / nz, m = i - rl * nz;
    ((F*)&s[m * S + (rl >> 1)])[rl & 1] = v[j];
    if (m > 0) ((F*)&s[(M - m) * S + (rl >> 1)])[rl & 1] = v[j];
  }
  __syncthreads();
  lds_fft<F, M, T, kNT, -1>(s, twl);
  constexpr int NS = per_thread(2 * T * (M / 2 + 1), kNT);
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const int i = threadIdx.x + j * kNT;
    if (i >= 2 * T * Hz) break;
    const int rl = i / Hz, k = i - rl * Hz;
    const int64_t row = row0 + rl;
    if (row >= R) continue;
    const cpx<F> w = s[k * S + (rl >> 1)];
    out[row * Hz + k] = (rl & 1) ? w.y : w.x;
  }
}

template <int M>
__global__ void __launch_bounds__(kNT) k_zcos(const double* __restrict__ K, double* __restrict__ out, MeshDims d,
                                              const KickScalars* __restrict__ sc,
                                              const cpx<double>* __restrict__ tw, int force) {
  zcos_body<M>(K, out, d, sc, tw, force, blockIdx.x, gridDim.x);
}

// Even transform along a strided axis of a real array [outer][n][inner] ->
// [outer][H][inner]; adjacent inner columns are paired into one complex FFT.
// The last (x) pass scales by gscale and writes the FP32 or FP64 spectrum.
template <int M, typename FO>
__device__ __forceinline__ void scos_body(const double* __restrict__ in, FO* __restrict__ out, int n, int inner, int H,
                                          int nouter, const KickScalars* __restrict__ sc,
                                          const cpx<double>* __restrict__ tw, int force, int last, double inv_m3,
                                          const int bx, const int nbx) {
  using F = double;
  constexpr int T = Tile<F, M>::T, S = Tile<F, M>::S;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<F>* s = (cpx<F>*)smem_raw;
  if (sc->err || !(force || sc->rebuild)) return;
  const cpx<F>* twl = stage_tw<F, M, kNT, Tile<F, M>::stw>(s + (size_t)M * S, tw);
  // real columns (outer, col), numbered flat over nouter * inner, 2T per
  // tile; this thread's column is the same in every loop below
  static_assert(kNT % (2 * T) == 0, "a thread must keep its column in every loop");
  const int cl = threadIdx.x % (2 * T);
  const int fc = bx * 2 * T + cl;
  const bool ok = fc < nouter * inner;
  const int o = ok ? fc / inner : 0, col = ok ? fc - o * inner : 0;
  for (int i = threadIdx.x; i < M * T; i += kNT) s[(i / T) * S + (i % T)] = cpx<F>{0, 0};
  __syncthreads();
  constexpr int NL = per_thread(T * M, kNT);  // n <= M/2 values of 2T columns
  F v[NL];
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    const int m = (threadIdx.x + j * kNT) / (2 * T);
    v[j] = (ok && m < n) ? in[((int64_t)o * n + m) * inner + col] : 0;
  }
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    const int m = (threadIdx.x + j * kNT) / (2 * T);
    if (m >= n) break;
    ((F*)&s[m * S + (cl >> 1)])[cl & 1] = v[j];
    if (m > 0) ((F*)&s[(M - m) * S + (cl >> 1)])[cl & 1] = v[j];
  }
  __syncthreads();
  lds_fft<F, M, T, kNT, -1>(s, twl);
  double g = 1.0;
  if (last) {
    // potential scale 1/(4 pi eps0 hx hy hz) (sc.py:167) and the 1/(Mx My Mz)
    // of the unnormalised inverse transform
    const double hx = sc->steps[0], hy = sc->steps[1], hz = sc->steps[2];
    const double denom = 4 * kPi * kEps0 * hx * hy * hz;
    g = inv_m3 / denom;
  }
  if (!ok) return;
  constexpr int NS = per_thread((M / 2 + 1) * 2 * T, kNT);
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const int k = (threadIdx.x + j * kNT) / (2 * T);
    if (k >= H) break;
    const cpx<F> w = s[k * S + (cl >> 1)];
    const double r = (cl & 1) ? w.y : w.x;
    out[((int64_t)o * H + k) * inner + col] = (FO)(r * g);
  }
}

template <int M, typename FO>
__global__ void __launch_bounds__(kNT) k_scos(const double* __restrict__ in, FO* __restrict__ out, int n, int inner,
                                              int H, int nouter, const KickScalars* __restrict__ sc,
                                              const cpx<double>* __restrict__ tw, int force, int last,
                                              double inv_m3) {
  scos_body<M, FO>(in, out, n, inner, H, nouter, sc, tw, force, last, inv_m3, blockIdx.x, gridDim.x);
}

}  // namespace ocsc
