// Beam moments of a bunch on the device: the statistics behind Ocelot's
// get_envelope (ocelot/cpbd/beam/analysis.py:27-224).  Every quantity is a
// mean over the particles, taken in passes that each read the six rows once;
// the per-block sums go through the same fixed reduction trees as the kick's
// frame sums, so results are reproducible run to run.
//
//   pass 0 (all particles)   sum tau, sum q
//   pass 1 (all, bounds)     sum (tau - mean)^2            -> std(tau), window
//   pass 2 (selected)        count, sums of p, x, px, y, py, tau
//   pass 3 (sel., auto_disp) sums of (p - <p>)^2 and (p - <p>)(u - <u>)
//   pass 4 (selected)        sums of the dispersion-corrected x, y, px, py
//   pass 5 (selected)        the twelve centred second moments
// With the dispersion known before the passes (given, or zero), pass 4 needs
// nothing from pass 2 and runs in the same read of the rows: pass 7 = 2 + 4;
// without a tau window pass 0's sums are pass 2's too: pass 6 = 0 + 2 + 4.
// Every value is summed in the same order and tree as in its own pass, so the
// results are the same bit for bit.
#pragma once
#include "particles.hpp"

namespace ocsc {

// rows of rparticles (beam/particle.py:17-22): x, px, y, py, tau, p
struct MomState {
  int64_t n;          // particles
  int32_t bounds;     // select lo <= tau <= hi
  int32_t disp_mode;  // 0: D = 0, 1: D given, 2: D estimated (auto_disp)
  double b0, b1;      // window in units of std(tau) around mean(tau)
  double lo, hi;
  double tau_all, std_all, sumq;
  double cnt;
  double m[6];        // means of the selected particles: p, x, px, y, py, tau
  double D[4];        // Dx, Dxp, Dy, Dyp
  double var_p;
  double c[4];        // means of the corrected x, y, px, py
  double mom[12];     // xx, xpx, pxpx, yy, ypy, pypy, tautau, xy, pxpy, xpy, ypx, pp
  double tau_min, tau_max;  // extent of tau (slice = "Imax")
  double z0;                // window centre: mean(tau), or the peak-current position (slice = "Imax")
};

constexpr int kMomNV[8] = {2, 1, 7, 5, 4, 12, 12, 11};

// dispersion-corrected coordinates (analysis.py:112-130): u -= D * p, then
// px, py scaled by 1 - p - p^2/2 + px^2/2 + py^2/2 of the corrected px, py
__device__ __forceinline__ void mom_corrected(const double* R6, const MomState& s, double* x, double* y,
                                              double* px, double* py) {
  const double p = R6[5];
  *x = R6[0] - s.D[0] * p;
  const double px0 = R6[1] - s.D[1] * p;
  *y = R6[2] - s.D[2] * p;
  const double py0 = R6[3] - s.D[3] * p;
  const double f = 1. - p - 0.5 * p * p + 0.5 * px0 * px0 + 0.5 * py0 * py0;
  *px = px0 * f;
  *py = py0 * f;
}

// DRIFT (pass 5 inside tracking, the last read of a step): the next step's
// drift is applied to the rows after their moments are added, to every
// particle (selected or not), with the drift kernel's fused multiply-adds:
// x += L x', y += L y', tau += r56 delta.  rw is R.r, writable.
template <int PASS, bool DRIFT = false>
__global__ void __launch_bounds__(kPT) k_mom(Rows R, const double* __restrict__ q, const MomState* __restrict__ msp,
                                             double* __restrict__ part, double* rw = nullptr, double L = 0,
                                             double r56 = 0) {
  static_assert(!DRIFT || PASS == 5, "the drift rides on the last pass");
  constexpr int NV = kMomNV[PASS];
  __shared__ double lds[(kPT / 64) * NV];
  const MomState s = *msp;
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) acc[c] = 0;
  for (int64_t i = (int64_t)blockIdx.x * kPT + threadIdx.x; i < s.n; i += (int64_t)gridDim.x * kPT) {
    const double tau = R.r[4 * R.ld + i];
    if constexpr (DRIFT) {
      const int64_t ld = R.ld;
      const double x = rw[i], xp = rw[ld + i], y = rw[2 * ld + i], yp = rw[3 * ld + i], d = rw[5 * ld + i];
      if (!s.bounds || (s.lo <= tau && tau <= s.hi)) {
        const double R6[6] = {x, xp, y, yp, tau, d};
        double cx, cy, cpx, cpy;
        mom_corrected(R6, s, &cx, &cy, &cpx, &cpy);
        const double dx = cx - s.c[0], dy = cy - s.c[1], dpx = cpx - s.c[2], dpy = cpy - s.c[3];
        const double dt = tau - s.m[5], dp = d - s.m[0];
        acc[0] += dx * dx;
        acc[1] += dx * dpx;
        acc[2] += dpx * dpx;
        acc[3] += dy * dy;
        acc[4] += dy * dpy;
        acc[5] += dpy * dpy;
        acc[6] += dt * dt;
        acc[7] += dx * dy;
        acc[8] += dpx * dpy;
        acc[9] += dx * dpy;
        acc[10] += dy * dpx;
        acc[11] += dp * dp;
      }
      rw[i] = __builtin_fma(L, xp, x);
      rw[2 * ld + i] = __builtin_fma(L, yp, y);
      rw[4 * ld + i] = __builtin_fma(r56, d, tau);
      continue;
    }
    if constexpr (PASS == 0) {
      acc[0] += tau;
      acc[1] += q[i];
    } else if constexpr (PASS == 1) {
      const double d = tau - s.tau_all;
      acc[0] += d * d;
    } else {
      if (s.bounds && !(s.lo <= tau && tau <= s.hi)) continue;
      double R6[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) R6[k] = k == 4 ? tau : R.r[k * R.ld + i];
      if constexpr (PASS == 2 || PASS == 6 || PASS == 7) {
        acc[0] += 1.0;
        acc[1] += R6[5];
        acc[2] += R6[0];
        acc[3] += R6[1];
        acc[4] += R6[2];
        acc[5] += R6[3];
        acc[6] += tau;
        if constexpr (PASS == 6) acc[7] += q[i];
        if constexpr (PASS != 2) {
          constexpr int o = PASS == 6 ? 8 : 7;
          double x, y, px, py;
          mom_corrected(R6, s, &x, &y, &px, &py);
          acc[o] += x;
          acc[o + 1] += y;
          acc[o + 2] += px;
          acc[o + 3] += py;
        }
      } else if constexpr (PASS == 3) {
        const double dp = R6[5] - s.m[0];
        acc[0] += dp * dp;
        acc[1] += dp * (R6[0] - s.m[1]);
        acc[2] += dp * (R6[1] - s.m[2]);
        acc[3] += dp * (R6[2] - s.m[3]);
        acc[4] += dp * (R6[3] - s.m[4]);
      } else {
        double x, y, px, py;
        mom_corrected(R6, s, &x, &y, &px, &py);
        if constexpr (PASS == 4) {
          acc[0] += x;
          acc[1] += y;
          acc[2] += px;
          acc[3] += py;
        } else {
          const double dx = x - s.c[0], dy = y - s.c[1], dpx = px - s.c[2], dpy = py - s.c[3];
          const double dt = tau - s.m[5], dp = R6[5] - s.m[0];
          acc[0] += dx * dx;
          acc[1] += dx * dpx;
          acc[2] += dpx * dpx;
          acc[3] += dy * dy;
          acc[4] += dy * dpy;
          acc[5] += dpy * dpy;
          acc[6] += dt * dt;
          acc[7] += dx * dy;
          acc[8] += dpx * dpy;
          acc[9] += dx * dpy;
          acc[10] += dy * dpx;
          acc[11] += dp * dp;
        }
      }
    }
  }
  block_reduce<NV, 0>(acc, lds);
  if (threadIdx.x == 0)
    for (int c = 0; c < NV; ++c) part[c * gridDim.x + blockIdx.x] = acc[c];
}

// one block: totals of a pass, then the scalars the next pass needs
template <int PASS>
__global__ void __launch_bounds__(1024) k_mom_reduce(MomState* __restrict__ msp, const double* __restrict__ part,
                                                     int nblk) {
  constexpr int NV = kMomNV[PASS];
  __shared__ double lds[16 * NV];
  double t[NV];
  reduce_partials<NV, 0>(part, nblk, t, lds);
  if (threadIdx.x != 0) return;
  MomState& s = *msp;
  const double n = (double)s.n;
  if constexpr (PASS == 0) {
    s.tau_all = t[0] / n;
    s.sumq = t[1];
    s.z0 = s.tau_all;
  } else if constexpr (PASS == 1) {
    s.std_all = __builtin_sqrt(t[0] / n);
    s.lo = s.tau_all + s.std_all * s.b0;
    s.hi = s.tau_all + s.std_all * s.b1;
  } else if constexpr (PASS == 2 || PASS == 6 || PASS == 7) {
    if constexpr (PASS == 6) {
      s.tau_all = t[6] / n;
      s.sumq = t[7];
      s.z0 = s.tau_all;
    }
    s.cnt = t[0];
    for (int k = 0; k < 6; ++k) s.m[k] = t[k + 1] / t[0];
    if constexpr (PASS != 2) {
      constexpr int o = PASS == 6 ? 8 : 7;
      for (int k = 0; k < 4; ++k) s.c[k] = t[o + k] / s.cnt;
    }
  } else if constexpr (PASS == 3) {
    s.var_p = t[0] / s.cnt;
    for (int k = 0; k < 4; ++k) s.D[k] = t[k + 1] / s.cnt / s.var_p;
  } else if constexpr (PASS == 4) {
    for (int k = 0; k < 4; ++k) s.c[k] = t[k] / s.cnt;
  } else {
    for (int k = 0; k < 12; ++k) s.mom[k] = t[k] / s.cnt;
  }
}

// Per-step moments inside a tracking loop (graph-capturable: no host copies).
// k_mom_init fills the working state from scalars; k_mom_store appends it to
// the step table at a device-side counter.
__global__ void k_mom_init(MomState* __restrict__ s, int64_t n, int bounds, int disp_mode, double b0, double b1,
                           double d0, double d1, double d2, double d3) {
  if (threadIdx.x != 0) return;
  MomState m{};
  m.n = n;
  m.bounds = bounds;
  m.disp_mode = disp_mode;
  m.b0 = b0;
  m.b1 = b1;
  if (disp_mode == 1) {
    m.D[0] = d0;
    m.D[1] = d1;
    m.D[2] = d2;
    m.D[3] = d3;
  }
  *s = m;
}

__global__ void k_mom_store(const MomState* __restrict__ s, MomState* __restrict__ table, int* __restrict__ counter,
                            int cap) {
  if (threadIdx.x != 0) return;
  const int k = *counter;
  if (k < cap) table[k] = *s;
  *counter = k + 1;
}

__global__ void k_counter_reset(int* __restrict__ counter) {
  if (threadIdx.x == 0) *counter = 0;
}

// ---------------------------------------------------------------------------
// slice = "Imax" (analysis.py:61-64): the window is centred on the peak of the
// current profile s_to_cur(tau, 0.01 std(tau), ...) (analysis.py:266-340):
// first-order (cloud-in-cell) weights of every particle on a uniform grid,
// then a Gaussian smoothing (on the host, a few thousand bins).

// extent of tau over all particles
__global__ void __launch_bounds__(kPT) k_tau_range(Rows R, int64_t n, double* __restrict__ part) {
  __shared__ double lds[(kPT / 64) * 6];
  const double inf = __builtin_inf();
  double v[6] = {inf, inf, inf, -inf, -inf, -inf};
  for (int64_t i = (int64_t)blockIdx.x * kPT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPT) {
    const double t = R.r[4 * R.ld + i];
    v[0] = t < v[0] ? t : v[0];
    v[3] = t > v[3] ? t : v[3];
  }
  block_reduce<6, 1>(v, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = v[0];
    part[gridDim.x + blockIdx.x] = v[3];
  }
}

__global__ void __launch_bounds__(1024) k_tau_range_reduce(MomState* __restrict__ msp, const double* __restrict__ part,
                                                           int nblk) {
  __shared__ double lds[16 * 6];
  const double inf = __builtin_inf();
  double v[6] = {inf, inf, inf, -inf, -inf, -inf};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    v[0] = part[b] < v[0] ? part[b] : v[0];
    v[3] = part[nblk + b] > v[3] ? part[nblk + b] : v[3];
  }
  block_reduce<6, 1>(v, lds);
  if (threadIdx.x == 0) {
    msp->tau_min = v[0];
    msp->tau_max = v[3];
  }
}

// Cloud-in-cell weights (analysis.py:254-260, 316-321) as exact fixed-point
// integers: bin I = floor((tau - a)/ds) gets xi = 1 + I - c, bin I + 1 gets
// 1 - xi, each scaled by wscale (a power of two) and rounded.  One LDS
// histogram per workgroup (nbins <= kCicLds), written out whole; k_cic_sum adds
// the workgroups' histograms bin by bin.
constexpr int kCicLds = 8192;

// bins [j0, j0 + nb) of the nbins-bin profile (windows keep the LDS bounded)
__global__ void __launch_bounds__(kPT) k_cic(Rows R, int64_t n, double a, double ds, int nbins, int j0, int nb,
                                             double wscale, long long* __restrict__ hist) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  long long* h = (long long*)smem_raw;
  for (int j = threadIdx.x; j < nb; j += kPT) h[j] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kPT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPT) {
    const double c = (R.r[4 * R.ld + i] - a) / ds;
    const double fl = __builtin_floor(c);
    const double xi = 1 + fl - c;
    int I = (int)(fl < 0 ? 0 : (fl > nbins - 2 ? nbins - 2 : fl));
    if (I >= j0 && I < j0 + nb)
      atomicAdd((unsigned long long*)&h[I - j0], (unsigned long long)(long long)__builtin_rint(xi * wscale));
    if (I + 1 >= j0 && I + 1 < j0 + nb)
      atomicAdd((unsigned long long*)&h[I + 1 - j0], (unsigned long long)(long long)__builtin_rint((1 - xi) * wscale));
  }
  __syncthreads();
  for (int j = threadIdx.x; j < nb; j += kPT) hist[(int64_t)blockIdx.x * nb + j] = h[j];
}

// Totals of the per-workgroup histograms: a workgroup takes 64 bins, its 16
// waves add every 16th histogram, then wave 0 adds the 16 wave sums.  The
// sums are integers, so the result is exact in any order.  nbins_dev, when
// given, overrides nbins (a grid planned on the device); workgroups beyond
// the bins return at once.
constexpr int kCicBlocks = 1024;  // workgroups of the histogram pass

__global__ void __launch_bounds__(1024) k_cic_sum(const long long* __restrict__ hist, int nblk, int nbins,
                                                  const int* __restrict__ nbins_dev, long long* __restrict__ out) {
  __shared__ long long part[16][64];
  const int nb = nbins_dev ? *nbins_dev : nbins;
  if ((int)blockIdx.x * 64 >= nb) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + lane;
  long long acc = 0;
  if (j < nb)
    for (int b = w; b < nblk; b += 16) acc += hist[(int64_t)b * nb + j];
  part[w][lane] = acc;
  __syncthreads();
  if (w == 0 && j < nb) {
    long long t = 0;
    for (int k = 0; k < 16; ++k) t += part[k][lane];
    out[j] = t;
  }
}

// Histogram on a grid read from a plan P (members nbins, a, ds) planned on
// the device: one LDS window of at most the launch's dynamic LDS (bins).
template <class P>
__global__ void __launch_bounds__(kPT) k_cic_grid(Rows R, int64_t n, const P* __restrict__ p, double wscale,
                                                  long long* __restrict__ hist) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  long long* h = (long long*)smem_raw;
  const int nbins = p->nbins;
  if (nbins == 0) return;
  const double a = p->a, ds = p->ds;
  for (int j = threadIdx.x; j < nbins; j += kPT) h[j] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kPT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPT) {
    const double c = (R.r[4 * R.ld + i] - a) / ds;
    const double fl = __builtin_floor(c);
    const double xi = 1 + fl - c;
    const int I = (int)(fl < 0 ? 0 : (fl > nbins - 2 ? nbins - 2 : fl));
    atomicAdd((unsigned long long*)&h[I], (unsigned long long)(long long)__builtin_rint(xi * wscale));
    atomicAdd((unsigned long long*)&h[I + 1], (unsigned long long)(long long)__builtin_rint((1 - xi) * wscale));
  }
  __syncthreads();
  for (int j = threadIdx.x; j < nbins; j += kPT) hist[(int64_t)blockIdx.x * nbins + j] = h[j];
}

// slice = "Imax" inside tracking graphs: the grid of s_to_cur(tau, 0.01
// std(tau), ...) planned on the device from the moments state (the one-shot
// path plans it on the host the same way), then the smoothing and the first
// maximum in one workgroup; the window is centred there.
struct ImaxPlan {
  int nbins, K, err;  // err: 1 no spread / not finite, 2 more bins than the cap
  double a, ds, sig;
};

__global__ void k_imax_plan(const MomState* __restrict__ s, ImaxPlan* __restrict__ p, int cap,
                            int* __restrict__ err) {
  if (threadIdx.x != 0) return;
  ImaxPlan q{};
  const double sigma = 0.01 * s->std_all;
  const double a = s->tau_min - 3 * sigma, b = s->tau_max + 3 * sigma;
  const double ds0 = 0.25 * sigma;
  const double nf = ceil((b - a) / ds0);
  if (!(nf >= 1) || !(nf <= 1e300)) {
    q.err = 1;
  } else if (nf + 1 > (double)cap) {
    q.err = 2;
  } else {
    const int N = (int)nf;
    q.nbins = N + 1;
    q.ds = (b - a) / N;
    q.a = a;
    q.sig = sigma;
    q.K = (int)floor(3 * sigma / q.ds + 0.5);
  }
  *p = q;
  if (q.err && *err == 0) *err = q.err;
}

// smoothed profile (2K+1 Gaussian taps, then divided by their sum) and its
// first maximum; sets the window [z0 + b0 std, z0 + b1 std] in the state
__global__ void __launch_bounds__(1024) k_imax_peak(const long long* __restrict__ C, const ImaxPlan* __restrict__ p,
                                                    MomState* __restrict__ msp) {
  __shared__ double G[128];
  __shared__ double gsum;
  __shared__ double bv[16];
  __shared__ int bj[16];
  const int nbins = p->nbins, K = p->K;
  if (nbins == 0 || 2 * K + 1 > 128) return;
  const double ds = p->ds, sigma = p->sig;
  if (threadIdx.x == 0) {
    double gs = 0;
    for (int m = 0; m <= 2 * K; ++m) {
      const double x = (double)(m - K) * ds / sigma;
      G[m] = exp(-0.5 * (x * x));
      gs += G[m];
    }
    gsum = gs;
  }
  __syncthreads();
  double best = -1;
  int jbest = 0;
  for (int j = threadIdx.x; j < nbins; j += blockDim.x) {
    double acc = 0;
    for (int m = 0; m <= 2 * K; ++m) {
      const int i = j + K - m;
      if (i >= 0 && i < nbins) acc += (double)C[i] * G[m];
    }
    acc = acc / gsum;
    if (acc > best) {
      best = acc;
      jbest = j;
    }
  }
  // largest value, first index among equals
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_down(best, off, 64);
    const int oj = __shfl_down(jbest, off, 64);
    if (ov > best || (ov == best && oj < jbest)) {
      best = ov;
      jbest = oj;
    }
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    bv[w] = best;
    bj[w] = jbest;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k)
      if (bv[k] > best || (bv[k] == best && bj[k] < jbest)) {
        best = bv[k];
        jbest = bj[k];
      }
    MomState& s = *msp;
    const double z0 = (double)jbest * ds + p->a;
    s.z0 = z0;
    s.lo = z0 + s.std_all * s.b0;
    s.hi = z0 + s.std_all * s.b1;
  }
}

}  // namespace ocsc
