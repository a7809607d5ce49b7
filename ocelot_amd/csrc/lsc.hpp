// Longitudinal space charge (Ocelot's LSC process, ocelot/cpbd/sc.py:261-600)
// on the device.  One kick is
//
//   1. particle passes: mean/std/extent of tau, then the transverse size of
//      the central slice mean + b0 std <= tau < mean + b1 std (sc.py:566-577);
//   2. the current profile s_to_cur(tau, smooth std(tau), ...) (cloud-in-cell
//      weights as exact fixed-point sums, moments.hpp k_cic, then a Gaussian
//      smoothing and the normalisation, analysis.py:266-340);
//   3. the wake (sc.py:411-465): the profile's spectrum on the doubled grid
//      times the impedance, inverse real transform.  Grids are a few hundred
//      to a few thousand bins, so both transforms are direct sums, one
//      workgroup per output frequency / position, twiddles from sincospi of an
//      exactly reduced integer phase;
//   4. a particle pass adding W(tau) 1e-9/pc_ref to delta, with numpy.interp's
//      rules (ends held, exact grid hits) (sc.py:589-600).
//
// Every sum goes through fixed reduction trees: results do not depend on
// launch timing.  The kernels after the statistics read one LscPlan, formed
// either on the host (one-shot kicks: any grid size, one synchronisation) or
// by k_lsc_plan on the device (kicks inside captured tracking graphs).
#pragma once
#include "particles.hpp"

namespace ocsc {

struct LscState {
  int64_t n;
  double b0, b1;                      // slice bounds in units of std(tau)
  double tau_min, tau_max, tau_mean, sumq;
  double tau_std, lo, hi;             // slice lo <= tau < hi
  double cnt, xmin, ymin, xmax, ymax, xmean, ymean;
  double xvar, yvar;
};

// pass 0 (all):   min tau (slot 0), max tau (slot 3), sum tau, sum q
// pass 1 (all):   sum (tau - mean)^2
// pass 2 (slice): min x, min y (0, 1), max x, max y (3, 4), count, sum x, sum y
// pass 3 (slice): sums of (x - <x>)^2, (y - <y>)^2 (Gaussian profile only)
constexpr int kLscNV[4] = {8, 1, 9, 2};
constexpr int kLscKind[4] = {1, 0, 1, 0};

template <int PASS>
__global__ void __launch_bounds__(kPT) k_lsc_stat(Rows R, const double* __restrict__ q,
                                                  const LscState* __restrict__ sp, double* __restrict__ part) {
  constexpr int NV = kLscNV[PASS], KIND = kLscKind[PASS];
  __shared__ double lds[(kPT / 64) * NV];
  const LscState s = *sp;
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) acc[c] = (KIND == 1 && c < 3) ? __builtin_inf() : ((KIND == 1 && c < 6) ? -__builtin_inf() : 0.0);
  for (int64_t i = (int64_t)blockIdx.x * kPT + threadIdx.x; i < s.n; i += (int64_t)gridDim.x * kPT) {
    const double tau = R.r[4 * R.ld + i];
    if constexpr (PASS == 0) {
      acc[0] = tau < acc[0] ? tau : acc[0];
      acc[3] = tau > acc[3] ? tau : acc[3];
      acc[6] += tau;
      acc[7] += q[i];
    } else if constexpr (PASS == 1) {
      const double d = tau - s.tau_mean;
      acc[0] += d * d;
    } else {
      if (tau >= s.lo && tau < s.hi) {
        const double x = R.r[i], y = R.r[2 * R.ld + i];
        if constexpr (PASS == 2) {
          acc[0] = x < acc[0] ? x : acc[0];
          acc[1] = y < acc[1] ? y : acc[1];
          acc[3] = x > acc[3] ? x : acc[3];
          acc[4] = y > acc[4] ? y : acc[4];
          acc[6] += 1.0;
          acc[7] += x;
          acc[8] += y;
        } else {
          const double dx = x - s.xmean, dy = y - s.ymean;
          acc[0] += dx * dx;
          acc[1] += dy * dy;
        }
      }
    }
  }
  block_reduce<NV, KIND>(acc, lds);
  if (threadIdx.x == 0)
    for (int c = 0; c < NV; ++c) part[c * gridDim.x + blockIdx.x] = acc[c];
}

template <int PASS>
__global__ void __launch_bounds__(1024) k_lsc_stat_reduce(LscState* __restrict__ sp, const double* __restrict__ part,
                                                          int nblk) {
#pragma clang fp contract(off)
  constexpr int NV = kLscNV[PASS], KIND = kLscKind[PASS];
  __shared__ double lds[16 * NV];
  double t[NV];
  reduce_partials<NV, KIND>(part, nblk, t, lds);
  if (threadIdx.x != 0) return;
  LscState& s = *sp;
  if constexpr (PASS == 0) {
    s.tau_min = t[0];
    s.tau_max = t[3];
    s.tau_mean = t[6] / (double)s.n;
    s.sumq = t[7];
  } else if constexpr (PASS == 1) {
    s.tau_std = __builtin_sqrt(t[0] / (double)s.n);
    s.lo = s.tau_mean + s.tau_std * s.b0;
    s.hi = s.tau_mean + s.tau_std * s.b1;
  } else if constexpr (PASS == 2) {
    s.xmin = t[0];
    s.ymin = t[1];
    s.xmax = t[3];
    s.ymax = t[4];
    s.cnt = t[6];
    s.xmean = t[7] / t[6];
    s.ymean = t[8] / t[6];
  } else {
    s.xvar = t[0] / s.cnt;
    s.yvar = t[1] / s.cnt;
  }
}

// ---------------------------------------------------------------------------
// Smoothed, normalised profile (analysis.py:322-339): B1 = C (*) G with the
// 2K+1 normalised Gaussian taps G_m = exp(-((m - K) ds/sigma)^2/2), the
// 'same'-size centred window of the full convolution; current = q0 v/(ds sum B1)
// B1 (A) and bunch = current/(q0 c) (1/m, sc.py:585).  One workgroup.
constexpr int kLscMaxTaps = 64;

// ---------------------------------------------------------------------------
// Impedances (imaginary parts; both are purely reactive).
constexpr double kEulerGamma = 0.57721566490153286061;

// exp(x) E1(x) for x > 0: the power series of E1 (Abramowitz & Stegun 5.1.11)
// up to x = 1, above it the continued fraction A&S 5.1.22 evaluated by the
// modified Lentz method.  Relative error a few 1e-15 over (0, 40].
__device__ inline double exp_e1(double x) {
  if (x <= 1.0) {
    double term = 1.0, sum = 0.0;
    for (int k = 1; k < 60; ++k) {
      term *= -x / (double)k;
      const double t = term / (double)k;
      sum += t;
      if (__builtin_fabs(t) <= 1e-18 * __builtin_fabs(sum)) break;
    }
    return exp(x) * (-kEulerGamma - log(x) - sum);
  }
  double b = x + 1.0, c = 1e300, d = 1.0 / b, h = d;
  for (int i = 1; i < 1000; ++i) {
    const double a = -(double)i * (double)i;
    b += 2.0;
    d = 1.0 / (a * d + b);
    c = b + a / c;
    const double del = c * d;
    h *= del;
    if (__builtin_fabs(del - 1.0) <= 1e-16) break;
  }
  return h;
}

// 1 - x K1(x) for x > 0.  Up to x = 2 from the ascending series
// (A&S 9.6.11, n = 1) written so nothing cancels:
//   1 - x K1(x) = -x ln(x/2) I1(x) + (x^2/4) sum_k [psi(k+1) + psi(k+2)] (x^2/4)^k / (k! (k+1)!)
// with I1(x) = (x/2) sum_k (x^2/4)^k / (k! (k+1)!).  Above, K1 from its
// integral  int_0^inf exp(-x cosh u) cosh u du  by the trapezoidal rule, which
// converges exponentially for this analytic integrand; the step keeps the
// aliasing terms below 1e-17 relative (h <= pi^2/(40 + x) and <= 0.7/sqrt(x)).
__device__ inline double one_minus_xk1(double x) {
  if (x <= 2.0) {
    const double y = 0.25 * x * x;
    double t = 1.0, i1 = 0.0, s = 0.0, psi1 = -kEulerGamma;
    for (int k = 0; k < 40; ++k) {
      if (k > 0) {
        t *= y / ((double)k * (double)(k + 1));
        psi1 += 1.0 / (double)k;
      }
      const double psi2 = psi1 + 1.0 / (double)(k + 1);
      i1 += t;
      s += (psi1 + psi2) * t;
      if (t <= 1e-18 * i1) break;
    }
    return -x * log(0.5 * x) * (0.5 * x * i1) + y * s;
  }
  if (x > 745.0) return 1.0;
  const double h1 = 9.8696044010893586188 / (40.0 + x), h2 = 0.7 / __builtin_sqrt(x);
  const double h = 0.9 * (h1 < h2 ? h1 : h2);
  double s = 0.5 * exp(-x);
  for (int j = 1; j < 4000; ++j) {
    const double ch = cosh((double)j * h);
    const double v = exp(-x * ch) * ch;
    s += v;
    if (v <= 1e-18 * s) break;
  }
  return 1.0 - x * h * s;
}

struct LscImp {
  int step_profile;
  double coef;    // Gaussian: Z0/(4 pi c gamma^2); step: Z0 c
  double sigma;   // Gaussian rms size or step radius rb
  double gamma_c; // gamma * c (Gaussian) or c * gamma (step)
  double dz;
  double factor;  // 1 + K_max^2 fill_factor / 2
};

// Everything the per-bin and kick kernels need for one kick, formed from the
// slice statistics either on the host (ocsc_lsc_apply) or on the device
// (tracking graphs).  nbins == 0 means "no kick" (err says why).
struct LscPlan {
  int nbins, K, err;
  double a, ds, sig;          // profile grid a + j ds, smoothing width
  double sigma;               // transverse size (Gaussian rms or step radius)
  double q0, v, pc_ref;       // total charge, beam speed, reference momentum (GeV)
  double scale;               // W = scale w: q0 for the kick, -1 for wake_lsc
  double inv_dt0, dt, ndf;    // wake_lsc's frequency and step constants
  LscImp imp;
};

// inputs of a kick that do not come from the particles
struct LscConst {
  int step_profile;
  double smooth, gamma, v, pc_ref, dz, K_max, fill_factor;
  int cap;  // most profile bins allowed
  // tracking through undulators: the impedance factor 1 + K_max^2 fill/2 of
  // kick k is factors[k] (formed on the host as lsc_imp forms it), k read
  // from and advanced in fctr (device), k < nfactors
  const double* factors;
  int* fctr;
  int nfactors;
};

enum { kLscOk = 0, kLscNoSpread = 1, kLscTooManyBins = 2, kLscEmptySlice = 3, kLscWideSmoothing = 4 };

constexpr double kZ0 = 1. / (kC * kEps0);  // globals.py: Z0 = 1/(c eps0)

// impedance constants (sc.py:302-369, 446-449), formed as numpy forms them
__host__ __device__ inline LscImp lsc_imp(int step_profile, double gamma, double sigma, double dz, double K_max,
                                          double fill_factor) {
#pragma clang fp contract(off)
  LscImp p{};
  p.step_profile = step_profile;
  p.sigma = sigma;
  p.dz = dz;
  p.factor = 1 + 0.5 * K_max * K_max * fill_factor;
  if (!step_profile) {
    p.coef = kZ0 / (4 * kPi * kC * (gamma * gamma));
    p.gamma_c = gamma * kC;
  } else {
    p.coef = kZ0 * kC;
    p.gamma_c = kC * gamma;
  }
  return p;
}

// step and frequency constants of wake_lsc for a profile grid starting s0,
// s1 (sc.py:442-465, analysis.py:376-383 and sc.py:403-409)
__host__ __device__ inline void lsc_wake_grid(double s0, double s1, int64_t nb, LscPlan* p) {
#pragma clang fp contract(off)
  const double ds = s1 - s0;
  const double dt0 = ds / kC;
  p->inv_dt0 = 1 / dt0;
  const double n = (double)(2 * nb);
  const double sb0 = s0 + 1.0 * ds, sb1 = s0 + 2.0 * ds;
  p->dt = (sb1 - sb0) / kC;
  const double df = 1 / p->dt * 1.0 / n;
  p->ndf = n * df;
}

// the plan of one kick (sc.py:566-587 with analysis.py:306-315), in the
// reference's operation order
__host__ __device__ inline void lsc_make_plan(const LscState& h, const LscConst& k, LscPlan* p) {
#pragma clang fp contract(off)
  *p = LscPlan{};
  double sigma;
  if (k.step_profile) {
    if (!(h.cnt > 0)) {
      p->err = kLscEmptySlice;
      return;
    }
    const double ex = h.xmax - h.xmin, ey = h.ymax - h.ymin;
    sigma = (ey < ex ? ey : ex) / 2;
  } else {
    // numpy's std of an empty slice is NaN, which zeroes the impedance
    sigma = h.cnt > 0 ? (sqrt(h.xvar) + sqrt(h.yvar)) / 2. : __builtin_nan("");
  }
  const double sig = h.tau_std * k.smooth;
  const double a = h.tau_min - 3 * sig, b = h.tau_max + 3 * sig;
  const double nf = ceil((b - a) / (0.25 * sig));
  if (!(nf >= 1) || !(nf <= 1e300)) {
    p->err = kLscNoSpread;
    return;
  }
  if (nf + 1 > (double)k.cap) {
    p->err = kLscTooManyBins;
    return;
  }
  const int N = (int)nf;
  const double ds = (b - a) / N;
  const int K = (int)floor(3 * sig / ds + 0.5);
  if (2 * K + 1 > kLscMaxTaps) {
    p->err = kLscWideSmoothing;
    return;
  }
  p->nbins = N + 1;
  p->K = K;
  p->a = a;
  p->ds = ds;
  p->sig = sig;
  p->sigma = sigma;
  p->q0 = h.sumq;
  p->v = k.v;
  p->pc_ref = k.pc_ref;
  p->scale = h.sumq;  // W = -wake q (sc.py:587)
  lsc_wake_grid(a, 1.0 * ds + a, p->nbins, p);
  p->imp = lsc_imp(k.step_profile, k.gamma, sigma, k.dz, k.K_max, k.fill_factor);
}

// imaginary part of the impedance at angular frequency w (sc.py:302-369)
__device__ inline double lsc_impedance(double w, const LscImp& p) {
#pragma clang fp contract(off)
  if (!p.step_profile) {
    const double alpha = w * p.sigma / p.gamma_c;
    const double a2 = alpha * alpha;
    double T = 0.0;
    if (a2 > 40.0) {
      // the reference's ten terms of the asymptotic series
      double k = 0.0, fact = 1.0, xp = a2;
      for (int i = 0; i < 10; ++i) {
        if (i > 0) {
          fact *= (double)i;
          xp *= a2;
        }
        k += ((i & 1) ? -fact : fact) / xp;
      }
      T = k;
    } else if (a2 >= 1e-16) {
      T = exp_e1(a2);
    }
    return p.coef * w * T * p.dz * p.factor;
  }
  if (w < 1e-7) return 0.0;
  const double x = w * p.sigma / p.gamma_c;
  return p.coef / (4 * w * p.sigma * p.sigma) * p.dz * one_minus_xk1(x) * p.factor;
}

// e^{i pi 2 m / n} for integer m, reduced exactly to (-n/2, n/2] first
__device__ inline void unit_root(int64_t m, int64_t n, double* c, double* s) {
  int64_t r = m % n;
  if (2 * r > n) r -= n;
  sincospi(2.0 * (double)r / (double)n, s, c);
}

// spectrum on the doubled grid (n = 2 nb): Zb_k = dt sum_j c bunch_j e^{-2 pi i jk/n}
// (the zero padding adds nothing), times the impedance at w_k = 2 pi f_k,
// f_k = (1/dt0) k / n; stored as Z = i Zim_a Zb for k < nb.  One workgroup
// per k; workgroups beyond the plan's bins return at once.
__global__ void __launch_bounds__(256) k_lsc_spectrum(const double* __restrict__ bunch, const LscPlan* __restrict__ p,
                                                      double* __restrict__ Zr, double* __restrict__ Zi) {
  __shared__ double lds[4 * 2];
  const int k = blockIdx.x;
  const int nb = p->nbins;
  if (k >= nb) return;
  const int64_t n = 2 * (int64_t)nb;
  double v[2] = {0.0, 0.0};
  for (int j = threadIdx.x; j < nb; j += 256) {
    const double b = bunch[j] * kC;
    double c, s;
    unit_root((int64_t)j * k, n, &c, &s);
    v[0] += b * c;
    v[1] -= b * s;
  }
  block_reduce<2, 0>(v, lds);
  if (threadIdx.x == 0) {
    const double br = p->dt * v[0], bi = p->dt * v[1];
    const double f = p->inv_dt0 * (double)k / (double)n;
    const double za = lsc_impedance(f * 2 * kPi, p->imp);
    Zr[k] = -za * bi;
    Zi[k] = za * br;
  }
}

// inverse real transform of the Hermitian spectrum whose first nb + 1 terms
// are Z_0..Z_{nb-1}, conj(Z_{nb-1}) (the reference's completion), scaled by
// n df: w_m = df [Re Z_0 + 2 sum_{k=1}^{nb-1} Re(Z_k e^{2 pi i km/n}) + Re Z_{nb-1} (-1)^m];
// out_m = scale w_m.  One workgroup per m.
__global__ void __launch_bounds__(256) k_lsc_wake(const double* __restrict__ Zr, const double* __restrict__ Zi,
                                                  const LscPlan* __restrict__ p, double* __restrict__ out) {
  __shared__ double lds[4];
  const int m = blockIdx.x;
  const int nb = p->nbins;
  if (m >= nb) return;
  const int64_t n = 2 * (int64_t)nb;
  double acc = 0.0;
  for (int k = 1 + threadIdx.x; k < nb; k += 256) {
    double c, s;
    unit_root((int64_t)k * m, n, &c, &s);
    acc += Zr[k] * c - Zi[k] * s;
  }
  block_reduce<1, 0>(&acc, lds);
  if (threadIdx.x == 0) {
    const double tot = Zr[0] + 2.0 * acc + ((m & 1) ? -Zr[nb - 1] : Zr[nb - 1]);
    out[m] = p->ndf * (tot / (double)n) * p->scale;
  }
}

// grid point j of the profile, as numpy forms it: (j ds) + a
__device__ inline double lsc_grid(int j, double a, double ds) {
#pragma clang fp contract(off)
  return (double)j * ds + a;
}

// delta += interp(tau, grid, W) 1e-9 / pc_ref, numpy.interp's rules; with
// DRIFT, then the next step's drift x += L x', y += L y', tau += r56 delta
// (the same fused multiply-adds as the drift kernel).  A refused plan
// (nbins == 0) leaves delta alone but still drifts.  With SUMP, per-block
// sums of the momenta of the written rows for the next space-charge kick,
// exactly as k_passA would form them (the drift keeps x', y', delta).
template <bool DRIFT, bool SUMP = false>
__global__ void __launch_bounds__(kPT) k_lsc_kick(double* __restrict__ r, int64_t ld, int64_t n,
                                                  const double* __restrict__ W, const LscPlan* __restrict__ p,
                                                  double L, double r56, const KickScalars* __restrict__ scp = nullptr,
                                                  double* __restrict__ part = nullptr) {
#pragma clang fp contract(off)
  __shared__ double lds[SUMP ? (kPT / 64) * 3 : 1];
  const int nbins = p->nbins;
  double acc[3] = {0, 0, 0};
  KickScalars ks{};
  if constexpr (SUMP) ks = *scp;
  // SUMP: the next space-charge kick's momentum sums over the written rows,
  // in k_passA's grid, order and partial layout (same result bit for bit)
  auto sump = [&](int64_t i, double d) {
    if constexpr (SUMP) {
      double m[3];
      to_p(r[ld + i], r[3 * ld + i], d, ks, m);
      acc[0] += m[0];
      acc[1] += m[1];
      acc[2] += m[2];
    }
  };
  if (nbins == 0) {
    for (int64_t i = (int64_t)blockIdx.x * kPT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPT) {
      if (DRIFT) {
        r[i] = __builtin_fma(L, r[ld + i], r[i]);
        r[2 * ld + i] = __builtin_fma(L, r[3 * ld + i], r[2 * ld + i]);
        r[4 * ld + i] = __builtin_fma(r56, r[5 * ld + i], r[4 * ld + i]);
      }
      sump(i, r[5 * ld + i]);
    }
  } else {
    const double a = p->a, ds = p->ds, pc_ref = p->pc_ref;
    const double x0 = lsc_grid(0, a, ds), xl = lsc_grid(nbins - 1, a, ds);
    // interpolated wake at tau (numpy.interp)
    auto wake_at = [&](double tau) -> double {
      if (tau != tau) return tau;
      if (tau < x0) return W[0];
      if (tau >= xl) return W[nbins - 1];
      // x_j <= tau < x_{j+1}, 0 <= j <= nbins - 2
      double fj = __builtin_floor((tau - a) / ds);
      int j = fj < 0 ? 0 : (fj > nbins - 2 ? nbins - 2 : (int)fj);
      while (j > 0 && lsc_grid(j, a, ds) > tau) --j;
      while (j < nbins - 2 && lsc_grid(j + 1, a, ds) <= tau) ++j;
      const double xj = lsc_grid(j, a, ds);
      if (xj == tau) return W[j];
      const double xj1 = lsc_grid(j + 1, a, ds);
      const double slope = (W[j + 1] - W[j]) / (xj1 - xj);
      double dE = slope * (tau - xj) + W[j];
      if (dE != dE) {
        dE = slope * (tau - xj1) + W[j + 1];
        if (dE != dE && W[j] == W[j + 1]) dE = W[j];
      }
      return dE;
    };
    struct Part {
      double x, xp, y, yp, tau, d;
    };
    auto load = [&](int64_t i, Part& v) {
      v.tau = r[4 * ld + i];
      v.d = r[5 * ld + i];
      if (DRIFT || SUMP) {
        v.xp = r[ld + i];
        v.yp = r[3 * ld + i];
      }
      if (DRIFT) {
        v.x = r[i];
        v.y = r[2 * ld + i];
      }
    };
    auto kick = [&](int64_t i, const Part& v) {
      const double d = v.d + wake_at(v.tau) * 1e-9 / pc_ref;
      r[5 * ld + i] = d;
      if (DRIFT) {
        r[i] = __builtin_fma(L, v.xp, v.x);
        r[2 * ld + i] = __builtin_fma(L, v.yp, v.y);
        r[4 * ld + i] = __builtin_fma(r56, d, v.tau);
      }
      if constexpr (SUMP) {
        double m[3];
        to_p(v.xp, v.yp, d, ks, m);
        acc[0] += m[0];
        acc[1] += m[1];
        acc[2] += m[2];
      }
    };
    // two particles per iteration, both loaded before either is written (the
    // compiler cannot move loads above stores to the same rows); the thread's
    // particles keep their order, so the sums do too
    const int64_t stride = (int64_t)gridDim.x * kPT;
    int64_t i = (int64_t)blockIdx.x * kPT + threadIdx.x;
    for (; i + stride < n; i += 2 * stride) {
      Part u, w;
      load(i, u);
      load(i + stride, w);
      kick(i, u);
      kick(i + stride, w);
    }
    if (i < n) {
      Part u;
      load(i, u);
      kick(i, u);
    }
  }
  if constexpr (SUMP) {
    block_reduce<3, 0>(acc, lds);
    if (threadIdx.x == 0)
      for (int c = 0; c < 3; ++c) part[c * gridDim.x + blockIdx.x] = acc[c];
  }
}

// smoothed, normalised profile (analysis.py:322-339, sc.py:585), one workgroup
__global__ void __launch_bounds__(1024) k_lsc_profile(const long long* __restrict__ C, const LscPlan* __restrict__ p,
                                                      double* __restrict__ cur, double* __restrict__ bunch) {
  __shared__ double G[kLscMaxTaps];
  __shared__ double lds[16];
  __shared__ double tot;
  const int nbins = p->nbins, K = p->K;
  if (nbins == 0) return;
  const double ds = p->ds, sigma = p->sig;
  if (threadIdx.x == 0) {
    double gs = 0;
    for (int m = 0; m <= 2 * K; ++m) {
      const double x = (double)(m - K) * ds / sigma;
      G[m] = exp(-0.5 * (x * x));
      gs += G[m];
    }
    for (int m = 0; m <= 2 * K; ++m) G[m] = G[m] / gs;
  }
  __syncthreads();
  double part = 0;
  for (int j = threadIdx.x; j < nbins; j += blockDim.x) {
    double acc = 0;
    for (int m = 0; m <= 2 * K; ++m) {
      const int i = j + K - m;
      if (i >= 0 && i < nbins) acc += (double)C[i] * G[m];
    }
    cur[j] = acc;
    part += acc;
  }
  block_reduce<1, 0>(&part, lds);
  if (threadIdx.x == 0) tot = part;
  __syncthreads();
  const double scale = p->q0 * p->v / (ds * tot);
  const double qc = p->q0 * kC;
  for (int j = threadIdx.x; j < nbins; j += blockDim.x) {
    const double b = scale * cur[j];
    cur[j] = b;
    bunch[j] = b / qc;
  }
}

// ---------------------------------------------------------------------------
// Inside tracking graphs: state set-up and the plan on the device (the
// histogram on the plan's grid is moments.hpp k_cic_grid).
__global__ void k_lsc_init(LscState* __restrict__ s, int64_t n, double b0, double b1) {
  if (threadIdx.x != 0) return;
  LscState h{};
  h.n = n;
  h.b0 = b0;
  h.b1 = b1;
  *s = h;
}

__global__ void k_lsc_plan(const LscState* __restrict__ s, LscConst k, LscPlan* __restrict__ p, int* __restrict__ err) {
  if (threadIdx.x != 0) return;
  LscPlan q;
  lsc_make_plan(*s, k, &q);
  if (k.factors) {
    const int i = *k.fctr;
    if (i < k.nfactors) q.imp.factor = k.factors[i];
    *k.fctr = i + 1;
  }
  *p = q;
  if (q.err && *err == 0) *err = q.err;
}


}  // namespace ocsc
