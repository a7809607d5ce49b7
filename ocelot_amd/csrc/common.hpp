// Shared definitions for the space-charge kernels (gfx950).
//
// Physical constants use the exact expressions of the reference
// (ocelot/common/globals.py:13-24): m_e_eV is derived from m_e and q_e and
// epsilon_0 from mu_0 = 4 pi 1e-7, not CODATA-2018 values.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ocsc {

constexpr double kPi = 3.141592653589793;
constexpr double kC = 299792458.0;
constexpr double kQe = 1.6021766208e-19;
constexpr double kMe = 9.10938215e-31;
// evaluated in the same order as the reference, so the doubles are identical
constexpr double kMeEV = kMe * (kC * kC) / kQe;
constexpr double kMeGeV = kMeEV / 1e+9;
constexpr double kMeEV2 = kMeEV * kMeEV;
constexpr double kMu0 = 4 * kPi * 1e-7;
constexpr double kEps0 = 1 / kMu0 / (kC * kC);

template <typename R>
struct cpx {
  R x, y;
};

template <typename R>
__device__ __forceinline__ cpx<R> cadd(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> csub(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
// complex product; fused multiply-adds are used on purpose in the FFT (the
// translation unit is built with contraction off elsewhere)
__device__ __forceinline__ cpx<float> cmul(cpx<float> a, cpx<float> b) {
  return {__builtin_fmaf(a.x, b.x, -a.y * b.y), __builtin_fmaf(a.x, b.y, a.y * b.x)};
}
__device__ __forceinline__ cpx<double> cmul(cpx<double> a, cpx<double> b) {
  return {__builtin_fma(a.x, b.x, -a.y * b.y), __builtin_fma(a.x, b.y, a.y * b.x)};
}
template <typename R>
__device__ __forceinline__ cpx<R> cscale(cpx<R> a, R s) { return {a.x * s, a.y * s}; }

// reciprocal from the hardware estimate plus two Newton steps (error below
// one ulp, not always correctly rounded): the mixed form's particle passes
// use it where the FP32 field dominates the error anyway
__device__ __forceinline__ double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  return __builtin_fma(r, e, r);
}

// 1 / (g*g) rounded once: double-double g^2, then one Newton correction of the
// reciprocal; matches a correctly rounded pow(g, -2).
__device__ __forceinline__ double inv_sq(double g) {
  double hi = g * g;
  double lo = __builtin_fma(g, g, -hi);
  double q = 1.0 / hi;
  double e = __builtin_fma(-hi, q, 1.0);
  double corr = __builtin_fma(-lo, q, e);
  return __builtin_fma(q, corr, q);
}

// Scalars of one kick, produced and consumed on the device so a whole kick
// (and a multi-kick sequence) can be captured in a HIP graph.
struct KickScalars {
  // inputs written by the host before the kick
  double gamref, betaref, pref, gbref;  // gbref = gamref*betaref
  double inv_gbref, inv_gamref, inv_betaref, inv_pref;  // reciprocals for the fast (mixed) transforms
  double cdT;                           // zstep / betaref
  double mesh_scale;                    // 1 or the U(1,1.1) draw
  double mesh_shift;                    // 0 or the U(-0.5,0.5) draw
  int32_t random_mesh;
  int32_t nmesh[3];
  int64_t n;
  // frame (ocelot/cpbd/sc.py:224-239)
  double T[9];  // T[3*j + k]: component j of axis vector t_{k+1}
  double Pav, gamma0, beta0, kperp;  // kperp = cdT*(1-beta0*beta0)
  // grid placement (ocelot/cpbd/sc.py:172-186)
  double steps[3];
  double inv_steps[3];
  double X_off[3];
  double xmin[3], xmax[3];
  double sumq, sumabsq, sumqx[3];
  // fixed-point charge scale: q*qscale summed exactly in int64
  double qscale, inv_qscale;
  // kernel-spectrum cache key (steps of the last build) and rebuild flag
  double key_steps[3];
  int32_t key_nmesh[3];
  int32_t key_prec;
  int32_t rebuild;
  // error flags: 1 = non-finite coordinate, 2 = zero total charge,
  // 4 = cell index out of range, 8 = non-positive/non-finite cell size
  int32_t err;
  int32_t err_sticky;  // errors of earlier kicks not yet reported
  // workgroups of the fused fold (k_fold_frame) that have finished; the last
  // one resets it
  uint32_t fold_count;
  // tracking: every charge of the bunch equals q_val bit for bit (checked at
  // the start of the call); the extent and deposit passes then take q_val
  // instead of loading the charge.  Zero after every other set-up.
  int32_t q_const;
  double q_val;
};

enum ErrBits : int32_t { kErrNonFinite = 1, kErrZeroCharge = 2, kErrCell = 4, kErrSteps = 8 };

__host__ __device__ constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

}  // namespace ocsc
