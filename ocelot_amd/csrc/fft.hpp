// LDS-resident radix-2/4/8/16 Stockham FFT for power-of-two lengths 8..1024.
//
// A workgroup holds T columns of length M in LDS as s[m*S + t] (S = T + 1:
// the pad keeps column-stride accesses off a single bank).  Each pass reads
// R inputs with stride M/R, twiddles them, does an R-point DFT in registers
// and writes the outputs in Stockham (self-sorting) order, so the result is
// in natural order after the last pass.  Twiddles W_M^k come from a table
// computed on the host in double precision.
#pragma once
#include "common.hpp"

namespace ocsc {

// cos / sin of 2 pi k / 16, k = 0..15
__device__ __forceinline__ constexpr double w16c(int k) {
  constexpr double c[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                            0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                            -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                            0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
  return c[k & 15];
}
__device__ __forceinline__ constexpr double w16s(int k) { return w16c(k - 4); }

// y = x * exp(DIR * 2 pi i * k / R) for k/R a multiple of 1/16
template <int R, int K, int DIR, typename F>
__device__ __forceinline__ cpx<F> twc16(cpx<F> v) {
  constexpr int k16 = ((K * 16 / R) % 16 + 16) % 16;
  if constexpr (k16 == 0) {
    return v;
  } else if constexpr (k16 == 8) {
    return {-v.x, -v.y};
  } else if constexpr ((k16 == 4 && DIR > 0) || (k16 == 12 && DIR < 0)) {
    return {-v.y, v.x};  // * i
  } else if constexpr ((k16 == 4 && DIR < 0) || (k16 == 12 && DIR > 0)) {
    return {v.y, -v.x};  // * -i
  } else {
    const F c = (F)w16c(k16);
    const F s = (F)(DIR * w16s(k16));
    return cmul(v, cpx<F>{c, s});
  }
}

// cos of 2 pi k / 32 for odd k (the even k are w16c's)
__device__ __forceinline__ constexpr double w32c_odd(int k) {
  constexpr double c[16] = {0.980785280403230449126,  0.831469612302545237079,  0.555570233019602224743,
                            0.195090322016128267848,  -0.195090322016128267848, -0.555570233019602224743,
                            -0.831469612302545237079, -0.980785280403230449126, -0.980785280403230449126,
                            -0.831469612302545237079, -0.555570233019602224743, -0.195090322016128267848,
                            0.195090322016128267848,  0.555570233019602224743,  0.831469612302545237079,
                            0.980785280403230449126};
  return c[(k & 31) >> 1];
}

// y = x * exp(DIR * 2 pi i * k / R) with constant k/R (R <= 32); DIR = -1 forward
template <int R, int K, int DIR, typename F>
__device__ __forceinline__ cpx<F> twc(cpx<F> v) {
  constexpr int k32 = ((K * 32 / R) % 32 + 32) % 32;
  if constexpr (k32 % 2 == 1) {
    const F c = (F)w32c_odd(k32);
    const F s = (F)(DIR * w32c_odd(k32 + 24));  // sin(a) = cos(a - pi/2)
    return cmul(v, cpx<F>{c, s});
  } else {
    return twc16<R, K, DIR, F>(v);
  }
}

// In-register R-point DFT (radix-2 decimation in time, fully unrolled).
template <int R, int DIR, typename F>
struct Dft {
  template <int K>
  __device__ __forceinline__ static void combine(cpx<F>* v, const cpx<F>* e, const cpx<F>* o) {
    if constexpr (K < R / 2) {
      cpx<F> t = twc<R, K, DIR>(o[K]);
      v[K] = cadd(e[K], t);
      v[K + R / 2] = csub(e[K], t);
      combine<K + 1>(v, e, o);
    }
  }
  __device__ __forceinline__ static void run(cpx<F>* v) {
    if constexpr (R == 2) {
      cpx<F> a = v[0], b = v[1];
      v[0] = cadd(a, b);
      v[1] = csub(a, b);
    } else {
      cpx<F> e[R / 2], o[R / 2];
#pragma unroll
      for (int i = 0; i < R / 2; ++i) {
        e[i] = v[2 * i];
        o[i] = v[2 * i + 1];
      }
      Dft<R / 2, DIR, F>::run(e);
      Dft<R / 2, DIR, F>::run(o);
      combine<0>(v, e, o);
    }
  }
};

// Radix plan of a length: passes of radix RMAX (16 or 8), then one
// remainder pass of radix 2/4/8.
template <int M, int RMAX = 16>
struct Plan {
  static constexpr int L = ilog2(M);
  static constexpr int lr = ilog2(RMAX);
  static constexpr int nbig = L / lr;
  static constexpr int rem = L % lr;
  static constexpr int npass = nbig + (rem ? 1 : 0);
  static constexpr int radix(int p) { return p < nbig ? RMAX : (1 << rem); }
  // smallest radix decides the butterfly count per column
  static constexpr int rmin = rem ? (1 << rem) : RMAX;
};

// largest radix that still gives every thread a butterfly: M/R butterflies
// per column times T columns should cover NT threads
template <int M, int T, int NT>
constexpr int radix_for() {
  return (M / 16) * T >= NT ? 16 : 8;
}

// One Stockham pass of radix R over T columns (LDS row stride S), NT threads.
template <typename F, int M, int R, int T, int NT, int DIR, int S>
__device__ __forceinline__ void stockham_pass(cpx<F>* s, int Ns, const cpx<F>* __restrict__ tw) {
  constexpr int NB = M / R;            // butterflies per column
  constexpr int TOT = NB * T;          // butterflies per workgroup
  constexpr int C = (TOT + NT - 1) / NT;  // per thread
  cpx<F> v[C][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int w = tid + c * NT;
    if (TOT % NT == 0 || w < TOT) {
      const int t = w % T;
      const int j = w / T;
      const int jm = j & (Ns - 1);
#pragma unroll
      for (int r = 0; r < R; ++r) v[c][r] = s[(j + r * NB) * S + t];
      // twiddle exponent jm*r/(Ns*R) of a turn -> table index jm*r*(M/(Ns*R))
      const int step = jm * (M / (Ns * R));
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cpx<F> w_ = tw[(step * r) & (M - 1)];
        if (DIR > 0) w_.y = -w_.y;
        v[c][r] = cmul(v[c][r], w_);
      }
      Dft<R, DIR, F>::run(v[c]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int w = tid + c * NT;
    if (TOT % NT == 0 || w < TOT) {
      const int t = w % T;
      const int j = w / T;
      const int jm = j & (Ns - 1);
      const int base = (j - jm) * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) s[(base + r * Ns) * S + t] = v[c][r];
    }
  }
  __syncthreads();
}

template <typename F, int M, int T, int NT, int DIR, int S, int RSEL, int P = 0>
__device__ __forceinline__ void lds_fft_from(cpx<F>* s, int Ns, const cpx<F>* __restrict__ tw) {
  using PL = Plan<M, RSEL ? RSEL : radix_for<M, T, NT>()>;
  if constexpr (P < PL::npass) {
    constexpr int R = PL::radix(P);
    stockham_pass<F, M, R, T, NT, DIR, S>(s, Ns, tw);
    lds_fft_from<F, M, T, NT, DIR, S, RSEL, P + 1>(s, Ns * R, tw);
  }
}

// In-place FFT of T columns held in LDS as s[m*S + t] (caller synchronises
// before).  DIR = -1: forward exp(-i...), DIR = +1: inverse without 1/M.
// RSEL: largest radix of the plan (0: radix_for's choice, or 8 / 16).
template <typename F, int M, int T, int NT, int DIR, int S = T + 1, int RSEL = 0>
__device__ __forceinline__ void lds_fft(cpx<F>* s, const cpx<F>* __restrict__ tw) {
  lds_fft_from<F, M, T, NT, DIR, S, RSEL, 0>(s, 1, tw);
}

// ---------------------------------------------------------------------------
// Register-resident transforms of length M = R*Q with Q = R (M = 16, 64,
// 256) or Q = 2R (M = 32, 128, 512), Q threads per column, from the four-step
// factorisation X[k + R*q] = sum_j W_Q^{jq} W_M^{jk} sum_n x[Q*n + j] W_R^{nk}
// (n, k < R; j, q < Q).  Thread j of a column holds x[n] = element Q*n + j on
// entry; it takes the R-point DFT over n in registers and multiplies by
// W_M^{jk}; one transpose through LDS then hands thread j' = k + R*h the Q
// values of its k.  For Q = R it takes their R-point DFT; for Q = 2R the half
// of the Q-point DFT with outputs of parity h: DFT_R(v[i] + v[i+R]) (h = 0)
// or DFT_R((v[i] - v[i+R]) W_Q^i) (h = 1).  Either way element j' + Q*m ends
// in x[m]: the output layout is the input layout of the next transform, so
// chained transforms (forward, pointwise product, inverse) touch LDS once
// each instead of once per radix pass.
//
// Exchange buffer for T columns: R blocks of Q*T entries plus padding.
// Lanes take the column index t fastest (JT = false: columns side by side in
// memory) or the in-column index j fastest (JT = true: rows contiguous in
// memory).  JT = false: thread (t, j) writes its value k to [k][j][t] and
// reads [k'][i][t]; consecutive lanes touch consecutive entries in both
// phases, and the block stride Q*T + 16 (16 entries past a multiple of 512
// bytes for FP32 when Q*T is a multiple of 32) puts the blocks one read
// instruction touches into alternate halves of the banks.  JT = true: writes
// go to [k][t][j] (consecutive) and reads to [k'][t][i], whose lanes step by
// the block stride Q*T + 1, odd, so 32 lanes hit 32 different bank pairs.
constexpr int kRegPad = 16;

template <int R, int Q, int T, bool JT = false>
struct RegX {
  static constexpr int blk = Q * T + (JT ? 1 : kRegPad);
  static constexpr int entries = R * blk;
};

// b[K] = a[K] * exp(DIR 2 pi i K / R) for K < R/2 (constant twiddles)
template <int R, int DIR, typename F, int K = 0>
__device__ __forceinline__ void ramp_half(cpx<F>* b, const cpx<F>* a) {
  if constexpr (K < R / 2) {
    b[K] = twc<R, K, DIR>(a[K]);
    ramp_half<R, DIR, F, K + 1>(b, a);
  }
}

// R-point DFT of v in place; NIN = R/2 when v[n] = 0 for n >= R/2: then
// X[2m] = DFT_{R/2}(v)[m] and X[2m+1] = DFT_{R/2}(v[n] W_R^n)[m].
template <int R, int DIR, typename F, int NIN>
__device__ __forceinline__ void dft_pruned(cpx<F>* v) {
  if constexpr (NIN * 2 == R && R >= 4) {
    cpx<F> a[R / 2], b[R / 2];
#pragma unroll
    for (int n = 0; n < R / 2; ++n) a[n] = v[n];
    ramp_half<R, DIR, F>(b, a);
    Dft<R / 2, DIR, F>::run(a);
    Dft<R / 2, DIR, F>::run(b);
#pragma unroll
    for (int m = 0; m < R / 2; ++m) {
      v[2 * m] = a[m];
      v[2 * m + 1] = b[m];
    }
  } else {
    Dft<R, DIR, F>::run(v);
  }
}

// v[I] = odd ? (a[I] - b[I]) W_{2R}^I : a[I] + b[I] (per-lane select)
template <int R, int DIR, typename F, int I = 0>
__device__ __forceinline__ void half_split(cpx<F>* v, const cpx<F>* a, const cpx<F>* b, const bool odd) {
  if constexpr (I < R) {
    const cpx<F> d = twc<2 * R, I, DIR>(csub(a[I], b[I]));
    const cpx<F> e = cadd(a[I], b[I]);
    v[I] = odd ? d : e;
    half_split<R, DIR, F, I + 1>(v, a, b, odd);
  }
}

// One transform of the column (t, j) as described above.  DIR = -1 forward,
// +1 inverse without 1/M; tw[k] = exp(-2 pi i k / M), k < M.  Begins with a
// barrier (the buffer may still be read by the previous transform); the
// caller publishes tw before the first call.
template <typename F, int R, int Q, int T, int DIR, int NIN, bool JT = false>
__device__ __forceinline__ void reg_fft(cpx<F> (&x)[R], cpx<F>* s, const int t, const int j,
                                        const cpx<F>* __restrict__ tw) {
  static_assert(Q == R || Q == 2 * R, "Q threads per column: R or 2R");
  constexpr int BLK = RegX<R, Q, T, JT>::blk;
  dft_pruned<R, DIR, F, NIN>(x);
#pragma unroll
  for (int k = 1; k < R; ++k) {
    cpx<F> w = tw[j * k];
    if (DIR > 0) w.y = -w.y;
    x[k] = cmul(x[k], w);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < R; ++k) s[k * BLK + (JT ? t * Q + j : j * T + t)] = x[k];
  __syncthreads();
  const int k1 = j & (R - 1);
  if constexpr (Q == R) {
#pragma unroll
    for (int i = 0; i < R; ++i) x[i] = s[k1 * BLK + (JT ? t * Q + i : i * T + t)];
  } else {
    cpx<F> a[R], b[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      a[i] = s[k1 * BLK + (JT ? t * Q + i : i * T + t)];
      b[i] = s[k1 * BLK + (JT ? t * Q + i + R : (i + R) * T + t)];
    }
    half_split<R, DIR, F>(x, a, b, j >= R);
  }
  Dft<R, DIR, F>::run(x);
}

}  // namespace ocsc
