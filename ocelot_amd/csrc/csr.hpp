// Tail of Ocelot's CSR kick (ocelot/cpbd/csr.py:69-79, 1105-1143) on the
// device: the linear convolution of the equidistant line charge with the
// reversed CSR kernel, and the kick of every particle by the result
// interpolated at z = -tau.  The mesh has tens to a few thousand points, so
// the convolution is a direct sum, one workgroup per output term, reduced in
// a fixed order; the reference's FFT of the zero-padded operands gives the
// same terms up to rounding.
#pragma once
#include "particles.hpp"

namespace ocsc {

// c_m = (sum_j a_j b_{m-j}) / div * mul for m < P + Q - 2 (the reference keeps
// the first P + Q - 2 of the P + Q - 1 terms)
__global__ void __launch_bounds__(256) k_conv(const double* __restrict__ a, int P, const double* __restrict__ b, int Q,
                                              double div, double mul, double* __restrict__ c) {
  __shared__ double lds[4];
  const int m = blockIdx.x;
  const int j0 = m - Q + 1 > 0 ? m - Q + 1 : 0, j1 = m < P - 1 ? m : P - 1;
  double acc = 0.0;
  for (int j = j0 + threadIdx.x; j <= j1; j += 256) acc += a[j] * b[m - j];
  block_reduce<1, 0>(&acc, lds);
  if (threadIdx.x == 0) c[m] = acc / div * mul;
}

// delta += interp(z inv_st + off, 0..L-1, lamK) 1e-9 / pc_ref, z = -tau,
// numpy.interp's rules on the integer grid (ends held, exact hits)
__global__ void __launch_bounds__(kPT) k_csr_kick(double* __restrict__ r, int64_t ld, int64_t n,
                                                  const double* __restrict__ lamK, int L, double inv_st, double off,
                                                  double pc_ref) {
#pragma clang fp contract(off)
  for (int64_t i = (int64_t)blockIdx.x * kPT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPT) {
    const double z = -r[4 * ld + i];
    const double x = z * inv_st + off;
    double dE;
    if (x != x) {
      dE = x;
    } else if (x < 0.0) {
      dE = lamK[0];
    } else if (x >= (double)(L - 1)) {
      dE = lamK[L - 1];
    } else {
      const int j = (int)__builtin_floor(x);
      const double xj = (double)j;
      if (xj == x) {
        dE = lamK[j];
      } else {
        const double slope = (lamK[j + 1] - lamK[j]) / ((double)(j + 1) - xj);
        dE = slope * (x - xj) + lamK[j];
        if (dE != dE) {
          dE = slope * (x - (double)(j + 1)) + lamK[j + 1];
          if (dE != dE && lamK[j] == lamK[j + 1]) dE = lamK[j];
        }
      }
    }
    r[5 * ld + i] = r[5 * ld + i] + dE * 1e-9 / pc_ref;
  }
}

}  // namespace ocsc
