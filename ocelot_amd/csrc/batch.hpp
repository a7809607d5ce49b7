// Kicks of many bunches in one launch per pass (ocsc_apply_batched).
//
// Each bunch is a job: its rows, charges, scalars, partials and mesh
// workspaces.  A pass over the batch is one launch whose workgroups are the
// concatenation of the per-bunch grids of that pass: workgroup b finds its
// job j and its index inside that job's grid from a prefix table (jobs in
// increasing order of first workgroup), then runs the same device body as the
// single-bunch kernel with that index and grid size.  Every bunch therefore
// goes through exactly the arithmetic of a single kick (same reduction
// trees, same transforms), while a launch fills the GPU with the work of all
// bunches instead of one bunch's few workgroups.  The transform passes are
// templated on their FFT length, so those launch once per distinct length
// along their axis.
#pragma once
#include "mesh.hpp"
#include "particles.hpp"

namespace ocsc {

// everything the kernels of one bunch read
struct BJob {
  double* r;  // rows to kick (6, ld)
  int64_t ld;
  const double* q;
  int64_t n;
  KickScalars* sc;
  double* part;  // partials of passes A and B
  int nblk;      // their workgroup count
  MeshDims d;
  int64_t ncell;
  int64_t nzero;             // words of the charge grid (copies x words x cells)
  unsigned long long* grid;  // charge grid
  void* A;                   // transform scratch [nx][ny][Hz] (x3 fields in the mixed form)
  void* B;                   // transform scratch [nx][My][Hz] (x2 in the mixed form)
  double* KF;                // Green's-function primitive at the nodes
  double* KA;                // kernel-spectrum scratch
  double* KB;
  void* G;    // kernel spectrum
  void* phi;  // mixed: three FP32 field grids; FP64: the potential
  double* E64;  // FP64: the three difference fields
  // cell-ordered copy kept across the rounds of a call: rows (ld n),
  // charges, caller index of each copy position; offset of the job's
  // particles in the call-wide sort arrays
  double* crs;
  double* cqs;
  uint32_t* cperm;
  int64_t off;
};

// one launch over some of the jobs: their ids, the prefix of the workgroups
// given to them (pre[0] = 0, pre[nj] = the launch's grid) and the
// workgroups each uses (fewer where a job's share is padded)
struct BSeg {
  const int* ids;
  const int64_t* pre;
  const int64_t* cnt;
  int nj;
};

// job, workgroup index inside the job and the job's grid for this workgroup
__device__ __forceinline__ void seg_find(const BSeg& g, int& job, int& bx, int& nbx) {
  const int64_t b = blockIdx.x;
  int lo = 0, hi = g.nj;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (g.pre[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  job = g.ids[lo];
  bx = (int)(b - g.pre[lo]);
  nbx = (int)g.cnt[lo];
}

#define OCSC_BJOB()          \
  int j_, bx, nbx;           \
  seg_find(g, j_, bx, nbx);  \
  if (bx >= nbx) return;     \
  const BJob& b = J[j_]

__global__ void __launch_bounds__(kPT) kb_passA(const BJob* __restrict__ J, BSeg g) {
  OCSC_BJOB();
  passA_body(Rows{b.r, b.ld}, b.sc, b.part, bx, nbx);
}

template <bool FAST>
__global__ void __launch_bounds__(kPT) kb_passB(const BJob* __restrict__ J, BSeg g) {
  OCSC_BJOB();
  passB_body<FAST>(Rows{b.r, b.ld}, b.q, b.sc, b.part, b.grid, b.nzero, bx, nbx);
}

// one workgroup per job
__global__ void __launch_bounds__(1024) kb_reduceA(const BJob* __restrict__ J) {
  const BJob& b = J[blockIdx.x];
  reduceA_body(b.sc, b.part, b.nblk, 0, 1);
}

__global__ void __launch_bounds__(1024) kb_reduceB(const BJob* __restrict__ J, int prec) {
  const BJob& b = J[blockIdx.x];
  reduceB_body(b.sc, b.part, b.nblk, prec, 0, 1);
}

template <int NWORD, bool XCD, bool AGG, int TILE>
__global__ void __launch_bounds__(kDepPT) kb_deposit(const BJob* __restrict__ J, BSeg g) {
  OCSC_BJOB();
  deposit_body<NWORD, XCD, AGG, TILE>(Rows{b.r, b.ld}, b.q, b.sc, b.grid, nullptr, bx, nbx);
}

template <typename F>
__global__ void __launch_bounds__(kKickPT) kb_kick(const BJob* __restrict__ J, BSeg g) {
  OCSC_BJOB();
  const F* field = sizeof(F) == 8 ? (const F*)b.E64 : (const F*)b.phi;
  kick_body<F, false, false, false>(b.r, b.ld, field, b.sc, DriftArgs{0, 0}, nullptr, nullptr, nullptr, nullptr, 0, bx,
                                    nbx);
}

// Green's function and its cosine transforms (the buffers of the single
// kick's kernel-spectrum build: F nodes -> KB -> KA -> KB -> G)
__global__ void __launch_bounds__(256) kb_igf_nodes(const BJob* __restrict__ J, BSeg g) {
  OCSC_BJOB();
  igf_nodes_body(b.KF, b.d, b.sc, 0, bx, nbx);
}

__global__ void __launch_bounds__(256) kb_igf_diff(const BJob* __restrict__ J, BSeg g) {
  OCSC_BJOB();
  igf_diff_body(b.KF, b.KB, b.d, b.sc, 0, bx, nbx);
}

template <int M>
__global__ void __launch_bounds__(kNT) kb_zcos(const BJob* __restrict__ J, BSeg g, const cpx<double>* __restrict__ tw) {
  OCSC_BJOB();
  zcos_body<M>(b.KB, b.KA, b.d, b.sc, tw, 0, bx, nbx);
}

template <int M>
__global__ void __launch_bounds__(kNT) kb_scos_y(const BJob* __restrict__ J, BSeg g,
                                                 const cpx<double>* __restrict__ tw) {
  OCSC_BJOB();
  scos_body<M, double>(b.KA, b.KB, b.d.ny, b.d.hz, b.d.hy, b.d.nx, b.sc, tw, 0, 0, 1.0, bx, nbx);
}

template <int M, typename F>
__global__ void __launch_bounds__(kNT) kb_scos_x(const BJob* __restrict__ J, BSeg g,
                                                 const cpx<double>* __restrict__ tw) {
  OCSC_BJOB();
  const double inv_m3 = 1.0 / ((double)b.d.mx * (double)b.d.my * (double)b.d.mz);
  scos_body<M, F>(b.KB, (F*)b.G, b.d.nx, b.d.hy * b.d.hz, b.d.hx, 1, b.sc, tw, 0, 1, inv_m3, bx, nbx);
}

// transform passes
template <typename F, int M, int CIN, int V>
__global__ void __launch_bounds__((pass_threads<F, M, V>())) kb_zfwd(const BJob* __restrict__ J, BSeg g,
                                                                     const cpx<F>* __restrict__ tw, int ncopy) {
  OCSC_BJOB();
  zfwd_body<F, M, CIN, V>(b.grid, (cpx<F>*)b.A, b.d, b.sc, tw, ncopy, bx, nbx);
}

template <typename F, int M, int V>
__global__ void __launch_bounds__((pass_threads<F, M, V>())) kb_yfwd(const BJob* __restrict__ J, BSeg g,
                                                                     const cpx<F>* __restrict__ tw) {
  OCSC_BJOB();
  yfwd_body<F, M, V>((const cpx<F>*)b.A, (cpx<F>*)b.B, b.d, b.sc, tw, bx, nbx);
}

template <int M, int V>
__global__ void __launch_bounds__((field_threads<M, V>())) kb_xconv_field(const BJob* __restrict__ J, BSeg g,
                                                       const cpx<float>* __restrict__ tw,
                                                       const cpx<double>* __restrict__ twd) {
  OCSC_BJOB();
  cpx<float>* B = (cpx<float>*)b.B;
  cpx<float>* B2 = B + (int64_t)b.d.nx * b.d.my * b.d.hz;
  xconv_field_body<M, V>(B, B, B2, (const float*)b.G, b.d, b.sc, tw, twd, bx, nbx);
}

template <int M, int V>
__global__ void __launch_bounds__((field_threads<M, V>())) kb_yinv_field(const BJob* __restrict__ J, BSeg g,
                                                      const cpx<float>* __restrict__ tw,
                                                      const cpx<double>* __restrict__ twd) {
  OCSC_BJOB();
  const cpx<float>* B = (const cpx<float>*)b.B;
  const cpx<float>* B2 = B + (int64_t)b.d.nx * b.d.my * b.d.hz;
  cpx<float>* A = (cpx<float>*)b.A;
  const int64_t nyz = (int64_t)b.d.nx * b.d.ny * b.d.hz;
  yinv_field_body<M, V>(B, B2, A, A + nyz, A + 2 * nyz, b.d, b.sc, tw, twd, bx, nbx);
}

// blockIdx.y: the field component
template <int M, int V>
__global__ void __launch_bounds__((pass_threads<float, M, V>())) kb_zinv3(const BJob* __restrict__ J, BSeg g,
                                                                          const cpx<float>* __restrict__ tw,
                                                                          const cpx<double>* __restrict__ twd) {
  OCSC_BJOB();
  (void)nbx;
  const int64_t nyz = (int64_t)b.d.nx * b.d.ny * b.d.hz;
  zinv3_body<M, V>((const cpx<float>*)b.A, nyz, (float*)b.phi, b.ncell, b.d, b.sc, tw, twd, blockIdx.y, bx);
}

// FP64 form: convolution to the potential, then the difference fields
template <int M, int V>
__global__ void __launch_bounds__((pass_threads<double, M, V>())) kb_xconv64(const BJob* __restrict__ J, BSeg g,
                                                                             const cpx<double>* __restrict__ tw) {
  OCSC_BJOB();
  xconv_body<double, M, V>((cpx<double>*)b.B, (const double*)b.G, b.d, b.sc, tw, bx, nbx);
}

template <int M, int V>
__global__ void __launch_bounds__((pass_threads<double, M, V>())) kb_yinv64(const BJob* __restrict__ J, BSeg g,
                                                                            const cpx<double>* __restrict__ tw) {
  OCSC_BJOB();
  yinv_body<double, M, V>((const cpx<double>*)b.B, (cpx<double>*)b.A, b.d, b.sc, tw, bx, nbx);
}

template <int M, int V>
__global__ void __launch_bounds__((pass_threads<double, M, V>())) kb_zinv64(const BJob* __restrict__ J, BSeg g,
                                                                            const cpx<double>* __restrict__ tw) {
  OCSC_BJOB();
  zinv_body<double, M, false, V>((const cpx<double>*)b.A, (double*)b.phi, b.d, b.sc, tw, nullptr, bx, nbx);
}

__global__ void __launch_bounds__(256) kb_diff64(const BJob* __restrict__ J, BSeg g) {
  OCSC_BJOB();
  diff64_body((const double*)b.phi, b.sc, b.E64, bx, nbx);
}

// Cell-ordered copies of all jobs with one sort: every job's keys carry its
// index above the cell key (tag), so one stable radix sort over the
// concatenated keys orders the particles by job, then by cell, and each job's
// run is its own permutation.  Keys and records from the job's rows (its
// scalars give the frame and grid of the kick just done).
__global__ void __launch_bounds__(kPT) kb_sort_keys(const BJob* __restrict__ J, BSeg g, uint32_t* __restrict__ keys,
                                                    PRec* __restrict__ rec, int cells, int jshift) {
  OCSC_BJOB();
  sort_keys_body(Rows{b.r, b.ld}, b.q, b.sc, keys + b.off, rec + b.off, cells, DriftArgs{0, 0},
                 (uint32_t)j_ << jshift, bx, nbx);
}

// copy[:, i] = record perm[off + i] (a call-wide index into the job's own
// records), cperm[i] = its caller index
__global__ void __launch_bounds__(256) kb_gather_copy(const BJob* __restrict__ J, BSeg g,
                                                      const uint32_t* __restrict__ perm,
                                                      const PRec* __restrict__ rec) {
  OCSC_BJOB();
  const int64_t i = (int64_t)bx * 256 + threadIdx.x;
  if (i >= b.n) return;
  const uint32_t src = perm[b.off + i];
  double v[8];
  load_rec(rec + src, v);
#pragma unroll
  for (int k = 0; k < 6; ++k) stnt(v[k], b.crs + k * b.n + i);
  stnt(v[6], b.cqs + i);
  b.cperm[i] = (uint32_t)(src - (uint32_t)b.off);
}

// back to the caller's order: each copy position's record to its caller
// position (whole segments), then the records as rows
__global__ void __launch_bounds__(256) kb_copy_to_rec(const BJob* __restrict__ J, BSeg g, PRec* __restrict__ rec) {
  OCSC_BJOB();
  const int64_t i = (int64_t)bx * 256 + threadIdx.x;
  if (i >= b.n) return;
  double v[8];
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = ldnt(b.crs + k * b.n + i);
  v[6] = 0.0;
  v[7] = 0.0;
  store_rec(rec + b.off + b.cperm[i], v);
}

__global__ void __launch_bounds__(256) kb_rec_to_rows(const BJob* __restrict__ J, BSeg g,
                                                      const PRec* __restrict__ rec) {
  OCSC_BJOB();
  const int64_t i = (int64_t)bx * 256 + threadIdx.x;
  if (i >= b.n) return;
  double v[8];
  load_rec(rec + b.off + i, v);
#pragma unroll
  for (int k = 0; k < 6; ++k) stnt(v[k], b.r + k * b.ld + i);
}

#undef OCSC_BJOB

}  // namespace ocsc
