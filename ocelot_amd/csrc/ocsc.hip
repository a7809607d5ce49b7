// C ABI of the MI355X space-charge kick: contexts, workspaces, launch
// sequences.  Declarations and contracts are in include/ocsc.h.
#include "ocsc.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "fft.hpp"
#include "mesh.hpp"
#include "particles.hpp"
#include "moments.hpp"
#include "lsc.hpp"
#include "csr.hpp"
#include "sort.hpp"
#include "batch.hpp"

using namespace ocsc;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                            \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) return fail(OCSC_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCHCHK(what)                                                                      \
  do {                                                                                       \
    hipError_t e_ = hipGetLastError();                                                       \
    if (e_ != hipSuccess) return fail(OCSC_EHIP, std::string(what) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
};

// counts device (re)allocations of workspaces: a captured launch sequence
// holds the addresses it was captured with and is valid while this is unchanged
std::atomic<uint64_t> g_alloc_gen{0};

int ensure(Buf& b, size_t bytes) {
  if (b.cap >= bytes) return OCSC_OK;
  g_alloc_gen.fetch_add(1);
  if (b.p) {
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
  }
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(OCSC_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  b.cap = bytes;
  return OCSC_OK;
}

int next_pow2(int v) {
  int m = 1;
  while (m < v) m <<= 1;
  return m;
}

// host-side inputs of one kick, passed to the device by kernel argument
struct KickInputs {
  double gamref, betaref, pref, gbref, cdT, mesh_scale, mesh_shift, gamma0;
  double steps[3];
  int32_t random_mesh, mode;  // mode 0: kick; 1: el_field (gamma0 given); 2: steps given
  int32_t nmesh[3];
  int64_t n;
  // tracking with random_mesh: the (scale, shift) pair of kick k is draws[2k],
  // draws[2k + 1], k read from and advanced in draw_ctr (device), k < ndraws
  const double* draws;
  int* draw_ctr;
  int32_t ndraws;
  // keep the equal-charge flag of a tracking call (KickScalars::q_const)
  // instead of clearing it
  int32_t q_keep;
};

__device__ __forceinline__ void setup_one(KickScalars* __restrict__ s, const KickInputs& in) {
  s->gamref = in.gamref;
  s->betaref = in.betaref;
  s->pref = in.pref;
  s->gbref = in.gbref;
  s->inv_gbref = 1.0 / in.gbref;
  s->inv_gamref = 1.0 / in.gamref;
  s->inv_betaref = 1.0 / in.betaref;
  s->inv_pref = 1.0 / in.pref;
  s->cdT = in.cdT;
  s->mesh_scale = in.mesh_scale;
  s->mesh_shift = in.mesh_shift;
  if (in.draws) {
    const int k = *in.draw_ctr;
    if (k < in.ndraws) {
      s->mesh_scale = in.draws[2 * k];
      s->mesh_shift = in.draws[2 * k + 1];
    }
    *in.draw_ctr = k + 1;
  }
  s->random_mesh = in.random_mesh;
  for (int k = 0; k < 3; ++k) s->nmesh[k] = in.nmesh[k];
  s->n = in.n;
  s->err_sticky |= s->err;
  s->err = 0;
  // k_fold_frame's count: zero after every set-up (its last workgroup runs
  // this one; the first step of a call runs it in k_setup), so a count left
  // by an interrupted launch does not reach into later calls
  s->fold_count = 0;
  if (!in.q_keep) s->q_const = 0;
  if (in.mode == 1) {
    s->gamma0 = in.gamma0;
    for (int k = 0; k < 9; ++k) s->T[k] = (k % 4 == 0) ? 1.0 : 0.0;
  }
  if (in.mode == 2) {
    for (int k = 0; k < 3; ++k) {
      s->steps[k] = in.steps[k];
      s->key_steps[k] = -1.0;  // forced builds never satisfy a later cache lookup
    }
    s->key_prec = -1;
    s->inv_qscale = 1.0;
  }
}

__global__ void k_setup(KickScalars* __restrict__ s, KickInputs in) {
  if (threadIdx.x == 0) setup_one(s, in);
}

// Tracking steps: the previous kick's per-workgroup momentum sums folded by
// gridDim.x workgroups (k_fold), and the last workgroup to finish runs this
// kick's set-up and the frame reduction (k_setup, k_reduceA): one launch
// instead of three.  The same sums bit for bit: each workgroup folds the same
// chunk, and reducing the gridDim.x (<= 256) folded values with 256 threads
// takes the tree that 1024 threads would (the extra lanes add only zeros).
__global__ void __launch_bounds__(256) k_fold_frame(const double* __restrict__ part, int nblk,
                                                    double* __restrict__ folded, KickScalars* __restrict__ sc,
                                                    KickInputs in) {
  __shared__ double lds[16 * 3];
  __shared__ int last;
  const int nf = gridDim.x, chunk = (nblk + nf - 1) / nf;
  const int b0 = blockIdx.x * chunk, b1 = b0 + chunk < nblk ? b0 + chunk : nblk;
  double v[3] = {0, 0, 0};
  for (int b = b0 + threadIdx.x; b < b1; b += 256)
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] += part[c * nblk + b];
  block_reduce<3, 0>(v, lds);
  if (threadIdx.x == 0) {
    for (int c = 0; c < 3; ++c) folded[c * nf + blockIdx.x] = v[c];
    __threadfence();  // the folded values before the count
    last = atomicAdd(&sc->fold_count, 1u) == (unsigned)nf - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();  // every workgroup's folded values visible here
  if (threadIdx.x == 0) setup_one(sc, in);  // also resets the count
  __syncthreads();
  reduceA_body(sc, folded, nf, 0, 1);
}

// Tracking calls: whether every charge equals the first one bit for bit
// (KickScalars::q_const and q_val, kept by the steps' set-ups); the check
// follows the init on the same stream.
__global__ void k_q_const_init(const double* __restrict__ q, int64_t n, KickScalars* __restrict__ s) {
  if (threadIdx.x == 0) {
    s->q_val = n > 0 ? q[0] : 0.0;
    s->q_const = n > 0;
  }
}

__global__ void __launch_bounds__(256) k_q_const_check(const double* __restrict__ q, int64_t n,
                                                       KickScalars* __restrict__ s) {
  const long long q0 = __double_as_longlong(s->q_val);
  bool diff = false;
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256)
    diff |= __double_as_longlong(q[p]) != q0;
  if (diff) s->q_const = 0;  // every writer stores the same value
}

// batched calls: the set-up of every job (thread j: job j)
__global__ void kb_setup(const KickInputs* __restrict__ in, const BJob* __restrict__ J, int nb) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nb) setup_one(J[j].sc, in[j]);
}

// error bits of all jobs, or-ed; clear: reset them afterwards
__global__ void kb_errors(const BJob* __restrict__ J, int nb, int* __restrict__ out, int clear) {
  int e = 0;
  for (int j = threadIdx.x; j < nb; j += blockDim.x) {
    KickScalars* s = J[j].sc;
    e |= s->err | s->err_sticky;
    if (clear) {
      s->err = 0;
      s->err_sticky = 0;
    }
  }
  for (int off = 32; off > 0; off >>= 1) e |= __shfl_down(e, off, 64);
  if (threadIdx.x == 0) *out = e;
}

// x += L x', y += L y', tau += r56 delta (numpy.dot(R, X) accumulates with fma)
__global__ void __launch_bounds__(256) k_drift(double* __restrict__ r, int64_t ld, int64_t n, double L, double r56) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  r[p] = __builtin_fma(L, r[ld + p], r[p]);
  r[2 * ld + p] = __builtin_fma(L, r[3 * ld + p], r[2 * ld + p]);
  r[4 * ld + p] = __builtin_fma(r56, r[5 * ld + p], r[4 * ld + p]);
}

}  // namespace

// stages of one kick, in launch order (names in ocsc_stage_name)
enum Stage : int {
  kStFrame = 0,   // setup, pass A, frame
  kStExtent,      // pass B, grid placement
  kStDeposit,     // clear + deposit
  kStKernel,      // IGF + kernel spectrum (device-skipped on a cache hit)
  kStZfwd,
  kStYfwd,
  kStXconv,
  kStYinv,
  kStZinv,
  kStKick,        // gather + kick + write-back
  kNumStages
};
const char* const kStageNames[kNumStages] = {"frame", "extent", "deposit", "green", "fft_z", "fft_y",
                                             "conv_x", "ifft_y", "ifft_z", "kick"};

// Workspace of one kick in flight: device scalars, partials, charge grid,
// FFT scratch, potential / fields, kernel spectrum (with its cache key in
// the scalars).  A context owns lane 0; batched calls add lanes so that
// independent bunches can be kicked concurrently on their own streams.
struct Lane {
  KickScalars* sc = nullptr;       // device
  KickScalars* sc_host = nullptr;  // pinned mirror
  Buf part;                        // per-block partial sums
  Buf grid, A, B, phi, G;
  Buf KA, KB, KF;                  // kernel-spectrum build scratch (its own, so it can overlap the FFTs)
  Buf E64;                         // FP64 form: the three difference fields of phi
  Buf sk[2], sv, sh, srec;         // batched calls: sort keys, values, histograms, records (cell-order copies)
  hipStream_t stream = nullptr;    // extra lanes only
  hipEvent_t done = nullptr;       // extra lanes only
};

// pinned mirror of the LSC state and plan
struct LscHost {
  LscState s;
  LscPlan p;
};

struct ocsc_ctx {
  int device = 0;
  Lane lane0;
  std::vector<Lane*> lanes;        // extra lanes for batched calls
  cpx<float>* twf = nullptr;       // W_M^k for M at offset M (M = 8..1024)
  cpx<double>* twd = nullptr;
  hipStream_t cap_stream = nullptr;  // stream used for graph capture
  hipStream_t side = nullptr;        // second stream: kernel spectrum next to the deposit
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_fork0 = nullptr, ev_zero = nullptr;  // charge-grid clear on the side stream (grid_clear 1)
  int overlap = 1;                   // 1: build the kernel spectrum concurrently with the deposit
  // stage timing (ocsc_set_profiling): events between the stages of a kick
  int profiling = 0;
  // tuning options (ocsc_set_option)
  int deposit_xcd = 0;  // 1: per-XCD grid copies with L2-local atomics
  int deposit_agg = 0;  // 1: runs of equal cells summed across the wave before the LDS table
  int deposit_tile = 0;  // deposit particles per thread and hash slots (DepTile)
  // transform plan (Tile's V): 0 LDS Stockham passes (radix 16 or 8), 1
  // radix 8, 2 half-width tiles, 3 field passes with 256 threads (radix 16)
  // (1-3 mixed form only), 4 (default) register-resident transforms where the
  // length is 16 to 512, else as 0 (10M / 127^3 tracking step, mixed: 0.854
  // -> 0.816 ms, transform chain 172 -> 128 us; FP64: 0.983 -> 0.970 ms)
  int fft_plan = 4;
  int red_blocks = kMaxRedBlocks;  // blocks of the strided reduction passes
  int sort_every = 10;  // tracking: re-sort the resident bunch by cell every k steps (0: never)
  int permute_rec = 1;  // tracking: permutations staged through 64-byte particle records (0: row by row)
  int apply_sort = 2;   // one-shot kicks on a cell-ordered copy: 0 never, 1 always, 2 FP64 from kApplySortMin particles
  int batch_sort = 1;   // batched calls keep a cell-ordered copy of each bunch across the kicks of a call
  int sort_cells = 1;   // sort keys: bricks and the cells inside them (1), or bricks only (0)
  int sort_ahead = 1;   // tracking sort keys at the position half-way through the re-sort interval
  int resort_near = 0;  // re-sorts of the copy: one gather of rows, charges and permutation (1), or records
  // charge grid cleared by a kernel on the side stream (1), by pass B (0), or
  // in tracking by the previous step's kick (2; the first step as 1)
  int grid_clear = 1;
  // workgroups of the Green's-function node and difference kernels (grid-
  // stride loops; 0: one thread per node or cell): fewer leave more of the
  // GPU to the deposit they overlap
  int igf_blocks = 0;
  // tracking: set-up, fold and frame reduction in one launch (1) or three (0)
  int fold_frame = 1;
  // tracking: charges checked at the start of a call; when all are equal, the
  // extent and deposit passes take the charge from the kick scalars (1)
  // instead of loading it per particle (0)
  int q_uniform = 1;
  // one-shot kicks (ocsc_apply) replayed from a captured graph (option
  // "apply_graph"): the launch sequence of the last call, the set-up node
  // whose inputs change from call to call, and what it was captured for
  int apply_graph = 1;
  struct ApplyGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t setup = nullptr;
    KickInputs in{};
    std::vector<int64_t> key;
    hipEvent_t done = nullptr;  // after the last replay: node updates wait for it
  } ag[2];  // launch sequences with a fresh cell sort (0) and with a re-used cell order (1)
  // one-shot kicks on a cell-ordered copy: the cell order of the last sorting
  // call, re-used by the next apply_resort - 1 calls on the same rows (any
  // permutation gives the same kick bit for bit; a recent one keeps the
  // deposit's and the kick's locality), and the rows it was sorted for
  int apply_resort = 10;
  // such kicks read the records through the cell order in the deposit and
  // the kick (1) or gather them into a cell-ordered copy of the rows first (0)
  int rec_direct = 1;
  Buf apply_order;
  const double* ao_r = nullptr;
  int64_t ao_n = 0, ao_ld = 0;
  int ao_age = 0;  // calls since the sort, that one included; 0: no order kept
  // cell-ordered copy of bunch i of a batched call: six rows, charges, permutation
  struct BunchCopy {
    Buf rs, qs, perm;
  };
  std::vector<BunchCopy*> bcopy;
  // batched calls in one launch per pass (option "batch_fused"): a workspace
  // per bunch and the device tables of jobs, inputs and launch segments
  int batch_fused = 1;
  std::vector<Lane*> flanes;
  Buf btab, btab2;
  // step graphs of the last plain tracking call (no moments, LSC kicks or
  // random-mesh draws) and what they were captured for: the next call with
  // the same launch sequence replays them instead of capturing again
  struct TrackGraphs {
    std::vector<int64_t> key;
    hipGraph_t graphs[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipGraphNode_t knode[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  } tg;
  // tracking workspace: two cell-ordered copies of the bunch, sort buffers,
  // particle records for the permutations
  Buf rs[2], qs[2], keys[2], vals, perm[2], sort_tmp, rec;
  Buf draws;  // random-mesh draws of a tracking call and their counter
  // beam moments: device state and its pinned host mirror
  MomState* mom = nullptr;
  MomState* mom_host = nullptr;
  Buf cic;  // current-profile histograms (slice = "Imax")
  Buf mom_part, mom_tab, mom_ctr;  // per-step moments in tracking: partials, table, counter
  // A ring of event sets, one per kick in flight: kicks are timed without a
  // host wait after each one (which would idle the GPU between kicks); a set
  // is read back when its slot comes round again or when the times are asked.
  static constexpr int kEvRing = 64;
  hipEvent_t ev[kEvRing][kNumStages + 1] = {};
  bool ev_pending[kEvRing] = {};
  int64_t ev_seq = 0;
  double stage_ms[kNumStages] = {};
  int64_t stage_calls = 0;
  // kick timing (option "kick_timing"): a pair of events per tracking step,
  // recorded around the kick kernel inside the captured graphs, and the
  // per-step kick durations of the last ocsc_track_drift
  int kick_timing = 0;
  std::vector<hipEvent_t> kick_ev;
  std::vector<double> kick_ms;
  // longitudinal space charge: device state, pinned mirror, per-bin arrays
  // and the profile grid of the last kick
  LscState* lsc = nullptr;
  LscPlan* lsc_plan = nullptr;
  int* lsc_err = nullptr;  // first plan error inside tracking graphs
  LscHost* lsc_host = nullptr;
  Buf lsc_bins, lsc_cic;
  Buf lsc_fac;  // per-step undulator factors of an LSC tracking call and their counter
  Buf csr_bins;  // CSR: line charge, kernel, convolution
  int lsc_nbins = 0;
  ImaxPlan* imax_plan = nullptr;  // slice = "Imax" inside tracking graphs
  int* imax_err = nullptr;
  double lsc_a = 0, lsc_ds = 0;
};

namespace {

struct Geometry {
  MeshDims d;
  int64_t ncell;
};

int make_geometry(const int32_t* nmesh, Geometry& g) {
  for (int k = 0; k < 3; ++k) {
    if (nmesh[k] < 4) return fail(OCSC_EINVAL, "nmesh must be >= 4 on every axis (steps = extent/(n-3))");
    if (nmesh[k] > 512) return fail(OCSC_EINVAL, "nmesh must be <= 512 on every axis");
  }
  MeshDims& d = g.d;
  d.nx = nmesh[0];
  d.ny = nmesh[1];
  d.nz = nmesh[2];
  d.mx = next_pow2(2 * d.nx - 1);
  d.my = next_pow2(2 * d.ny - 1);
  d.mz = next_pow2(2 * d.nz - 1);
  d.hx = d.mx / 2 + 1;
  d.hy = d.my / 2 + 1;
  d.hz = d.mz / 2 + 1;
  g.ncell = (int64_t)d.nx * d.ny * d.nz;
  return OCSC_OK;
}

// bytes of each mesh workspace of one kick (E64: FP64 form only)
struct MeshSizes {
  size_t grid, A, B, KB, KA, KF, phi, G, E64;
};

template <typename F>
MeshSizes mesh_sizes(const Geometry& g, int nword, int deposit_xcd) {
  const MeshDims& d = g.d;
  MeshSizes m;
  m.grid = (size_t)(deposit_xcd ? kNumXcd : 1) * nword * g.ncell * 8;
  m.A = (size_t)d.nx * d.ny * d.hz * sizeof(cpx<F>);
  m.B = (size_t)d.nx * d.my * d.hz * sizeof(cpx<F>);
  if constexpr (sizeof(F) == 4) {
    // mixed form carries three fields through the inverse passes
    m.A = std::max(m.A, 3 * (size_t)d.nx * d.ny * d.hz * sizeof(cpx<float>));
    m.B = std::max(m.B, 2 * (size_t)d.nx * d.my * d.hz * sizeof(cpx<float>));
  }
  // kernel-spectrum scratch: the IGF grid / its y pass, and its z pass
  m.KB = std::max((size_t)g.ncell, (size_t)d.nx * d.hy * d.hz) * 8;
  m.KA = (size_t)d.nx * d.ny * d.hz * 8;
  m.KF = (size_t)(d.nx + 1) * (d.ny + 1) * (d.nz + 1) * 8;  // the Green's-function primitive at the nodes
  m.phi = (size_t)g.ncell * std::max(sizeof(double), 3 * sizeof(float));
  m.G = (size_t)d.hx * d.hy * d.hz * sizeof(double);
  m.E64 = sizeof(F) == 8 ? (size_t)3 * g.ncell * sizeof(double) : 0;
  return m;
}

template <typename F>
int ensure_mesh(ocsc_ctx* c, Lane& L, const Geometry& g, int nword) {
  const MeshSizes m = mesh_sizes<F>(g, nword, c->deposit_xcd);
  int rc;
  if ((rc = ensure(L.grid, m.grid))) return rc;
  if ((rc = ensure(L.A, m.A))) return rc;
  if ((rc = ensure(L.B, m.B))) return rc;
  if ((rc = ensure(L.KB, m.KB))) return rc;
  if ((rc = ensure(L.KA, m.KA))) return rc;
  if ((rc = ensure(L.KF, m.KF))) return rc;
  if ((rc = ensure(L.phi, m.phi))) return rc;
  if ((rc = ensure(L.G, m.G))) return rc;
  if (m.E64 && (rc = ensure(L.E64, m.E64))) return rc;
  return OCSC_OK;
}

template <typename F>
const cpx<F>* tw_for(ocsc_ctx* c, int M) {
  if constexpr (sizeof(F) == 4)
    return c->twf + M;
  else
    return c->twd + M;
}

#define OCSC_CASE_M(V, ...) \
  case V: {                 \
    constexpr int MM = V;   \
    __VA_ARGS__;            \
  } break;
#define OCSC_SWITCH_M(MVAR, ...)                                   \
  switch (MVAR) {                                                  \
    OCSC_CASE_M(8, __VA_ARGS__)                                    \
    OCSC_CASE_M(16, __VA_ARGS__)                                   \
    OCSC_CASE_M(32, __VA_ARGS__)                                   \
    OCSC_CASE_M(64, __VA_ARGS__)                                   \
    OCSC_CASE_M(128, __VA_ARGS__)                                  \
    OCSC_CASE_M(256, __VA_ARGS__)                                  \
    OCSC_CASE_M(512, __VA_ARGS__)                                  \
    OCSC_CASE_M(1024, __VA_ARGS__)                                 \
    default:                                                       \
      return fail(OCSC_EINVAL, "unsupported FFT length");          \
  }

// kernel spectrum G from the current cell sizes (skipped on device when the
// cache key matches, unless force)
template <typename F>
int launch_kernel_spectrum(ocsc_ctx* c, Lane& L, const Geometry& g, hipStream_t st, int force) {
  const MeshDims& d = g.d;
  double* K = (double*)L.KB.p;
  double* Kz = (double*)L.KA.p;
  double* Kzy = (double*)L.KB.p;
  {
    // F once per node, then the 8-corner differences
    const int64_t nn = (int64_t)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
    const int64_t cap = c->igf_blocks > 0 ? c->igf_blocks : 65536;
    k_igf_nodes<<<(int)std::min<int64_t>((nn + 255) / 256, cap), 256, 0, st>>>((double*)L.KF.p, d, L.sc, force);
    k_igf_diff<<<(int)std::min<int64_t>((igf_diff_items(d) + 255) / 256, cap), 256, 0, st>>>(
        (const double*)L.KF.p, K, d, L.sc, force);
  }
  LAUNCHCHK("k_igf");
  const int64_t R = (int64_t)d.nx * d.ny;
  OCSC_SWITCH_M(d.mz, {
    using TL = Tile<double, MM>;
    const int nb = (int)((R + 2 * TL::T - 1) / (2 * TL::T));
    k_zcos<MM><<<nb, kNT, TL::lds_bytes, st>>>(K, Kz, d, L.sc, c->twd + MM, force);
  });
  LAUNCHCHK("k_zcos");
  OCSC_SWITCH_M(d.my, {
    using TL = Tile<double, MM>;
    const int grid = (d.nx * d.hz + 2 * TL::T - 1) / (2 * TL::T);
    k_scos<MM, double><<<grid, kNT, TL::lds_bytes, st>>>(Kz, Kzy, d.ny, d.hz, d.hy, d.nx, L.sc, c->twd + MM, force, 0,
                                                         1.0);
  });
  LAUNCHCHK("k_scos(y)");
  const double inv_m3 = 1.0 / ((double)d.mx * (double)d.my * (double)d.mz);
  OCSC_SWITCH_M(d.mx, {
    using TL = Tile<double, MM>;
    const int inner = d.hy * d.hz;
    const int grid = (inner + 2 * TL::T - 1) / (2 * TL::T);
    k_scos<MM, F><<<grid, kNT, TL::lds_bytes, st>>>(Kzy, (F*)L.G.p, d.nx, inner, d.hx, 1, L.sc, c->twd + MM, force, 1,
                                                     inv_m3);
  });
  LAUNCHCHK("k_scos(x)");
  return OCSC_OK;
}

// phi = G * rho on the doubled grid; rho read through charge_at<CIN>.  V:
// the transform plan (option "fft_plan": 4 or the default passes).
template <typename F, int CIN, int V>
int launch_convolution_v(ocsc_ctx* c, Lane& L, const Geometry& g, const void* rho, F* phi, hipStream_t st,
                         hipEvent_t* ev, hipEvent_t join) {
  // ev[0] is recorded by the caller before the z pass; ev[1..4] after z, y, x, y.
  // join: the kernel spectrum built on another stream, awaited before the x pass
  const MeshDims& d = g.d;
  cpx<F>* A = (cpx<F>*)L.A.p;
  cpx<F>* B = (cpx<F>*)L.B.p;
  const int64_t R = (int64_t)d.nx * d.ny;
  OCSC_SWITCH_M(d.mz, {
    using TL = Tile<F, MM>;
    const int nb = (int)((R + 2 * TL::T - 1) / (2 * TL::T));
    constexpr int VZ = plan_zf_yi<F, MM, V>();
    k_zfwd<F, MM, CIN, VZ><<<nb, pass_threads<F, MM, VZ>(), pass_lds<F, MM, VZ>(), st>>>(
        rho, A, d, L.sc, tw_for<F>(c, MM), c->deposit_xcd ? kNumXcd : 1);
  });
  LAUNCHCHK("k_zfwd");
  if (ev) (void)hipEventRecord(ev[1], st);
  OCSC_SWITCH_M(d.my, {
    using TL = Tile<F, MM>;
    const int grid = (d.nx * d.hz + TL::T - 1) / TL::T;
    k_yfwd<F, MM, V><<<grid, pass_threads<F, MM, V>(), pass_lds<F, MM, V>(), st>>>(A, B, d, L.sc, tw_for<F>(c, MM));
  });
  LAUNCHCHK("k_yfwd");
  if (ev) (void)hipEventRecord(ev[2], st);
  if (join) HIPCHK(hipStreamWaitEvent(st, join, 0));
  OCSC_SWITCH_M(d.mx, {
    using TL = Tile<F, MM>;
    const int grid = (d.my * d.hz + TL::T - 1) / TL::T;
    k_xconv<F, MM, V><<<grid, pass_threads<F, MM, V>(), pass_lds<F, MM, V>(), st>>>(B, (const F*)L.G.p, d, L.sc,
                                                                                   tw_for<F>(c, MM));
  });
  LAUNCHCHK("k_xconv");
  if (ev) (void)hipEventRecord(ev[3], st);
  OCSC_SWITCH_M(d.my, {
    using TL = Tile<F, MM>;
    const int grid = (d.nx * d.hz + TL::T - 1) / TL::T;
    constexpr int VY = plan_zf_yi<F, MM, V>();
    k_yinv<F, MM, VY><<<grid, pass_threads<F, MM, VY>(), pass_lds<F, MM, VY>(), st>>>(B, A, d, L.sc,
                                                                                     tw_for<F>(c, MM));
  });
  LAUNCHCHK("k_yinv");
  if (ev) (void)hipEventRecord(ev[4], st);
  OCSC_SWITCH_M(d.mz, {
    using TL = Tile<F, MM>;
    const int nb = (int)((R + 2 * TL::T - 1) / (2 * TL::T));
    k_zinv<F, MM, false, V><<<nb, pass_threads<F, MM, V>(), pass_lds<F, MM, V>(), st>>>(A, phi, d, L.sc,
                                                                                        tw_for<F>(c, MM));
  });
  LAUNCHCHK("k_zinv");
  return OCSC_OK;
}

template <typename F, int CIN>
int launch_convolution(ocsc_ctx* c, Lane& L, const Geometry& g, const void* rho, F* phi, hipStream_t st,
                       hipEvent_t* ev = nullptr, hipEvent_t join = nullptr) {
  if (c->fft_plan == 4) return launch_convolution_v<F, CIN, 4>(c, L, g, rho, phi, st, ev, join);
  return launch_convolution_v<F, CIN, 0>(c, L, g, rho, phi, st, ev, join);
}

// Mixed form: the three field components on the mesh, E = -grad phi as the
// forward differences of the reference (sc.py:195-200) taken in Fourier
// space, written as three FP32 grids one after the other at E.  V: the FFT
// plan variant (option "fft_plan").
template <int V>
int launch_fields_v(ocsc_ctx* c, Lane& L, const Geometry& g, const void* rho, float* E, hipStream_t st,
                    hipEvent_t* ev, hipEvent_t join) {
  const MeshDims& d = g.d;
  const int64_t nyz = (int64_t)d.nx * d.ny * d.hz;
  cpx<float>* A = (cpx<float>*)L.A.p;
  cpx<float>* B = (cpx<float>*)L.B.p;
  cpx<float>* B2 = B + (int64_t)d.nx * d.my * d.hz;
  const int64_t R = (int64_t)d.nx * d.ny;
  OCSC_SWITCH_M(d.mz, {
    using TL = Tile<float, MM, V>;
    const int nb = (int)((R + 2 * TL::T - 1) / (2 * TL::T));
    k_zfwd<float, MM, kChargeFix1, V><<<nb, pass_threads<float, MM, V>(), pass_lds<float, MM, V>(), st>>>(
        rho, A, d, L.sc, c->twf + MM, c->deposit_xcd ? kNumXcd : 1);
  });
  LAUNCHCHK("k_zfwd");
  if (ev) (void)hipEventRecord(ev[1], st);
  OCSC_SWITCH_M(d.my, {
    using TL = Tile<float, MM, V>;
    const int grid = (d.nx * d.hz + TL::T - 1) / TL::T;
    k_yfwd<float, MM, V><<<grid, pass_threads<float, MM, V>(), pass_lds<float, MM, V>(), st>>>(A, B, d, L.sc,
                                                                                           c->twf + MM);
  });
  LAUNCHCHK("k_yfwd");
  if (ev) (void)hipEventRecord(ev[2], st);
  if (join) HIPCHK(hipStreamWaitEvent(st, join, 0));  // the kernel spectrum, built on another stream
  OCSC_SWITCH_M(d.mx, {
    using FT = FieldTile<MM, field_threads<MM, V>(), V>;
    const int grid = (d.my * d.hz + FT::T - 1) / FT::T;
    k_xconv_field<MM, V><<<grid, field_threads<MM, V>(), FT::lds, st>>>(B, B, B2, (const float*)L.G.p, d, L.sc, c->twf + MM,
                                                      c->twd + MM);
  });
  LAUNCHCHK("k_xconv_field");
  if (ev) (void)hipEventRecord(ev[3], st);
  OCSC_SWITCH_M(d.my, {
    using FT = FieldTile<MM, field_threads<MM, V>(), V>;
    const int grid = (d.nx * d.hz + FT::T - 1) / FT::T;
    k_yinv_field<MM, V><<<grid, field_threads<MM, V>(), FT::lds, st>>>(B, B2, A, A + nyz, A + 2 * nyz, d, L.sc, c->twf + MM,
                                                     c->twd + MM);
  });
  LAUNCHCHK("k_yinv_field");
  if (ev) (void)hipEventRecord(ev[4], st);
  OCSC_SWITCH_M(d.mz, {
    using TL = Tile<float, MM, V>;
    const int nb = (int)((R + 2 * TL::T - 1) / (2 * TL::T));
    k_zinv3<MM, V><<<dim3(nb, 3), pass_threads<float, MM, V>(), pass_lds<float, MM, V>(), st>>>(
        A, nyz, E, g.ncell, d, L.sc, c->twf + MM, c->twd + MM);
  });
  LAUNCHCHK("k_zinv(fields)");
  return OCSC_OK;
}

int launch_fields(ocsc_ctx* c, Lane& L, const Geometry& g, const void* rho, float* E, hipStream_t st,
                  hipEvent_t* ev = nullptr, hipEvent_t join = nullptr) {
  switch (c->fft_plan) {
    case 1:
      return launch_fields_v<1>(c, L, g, rho, E, st, ev, join);
    case 2:
      return launch_fields_v<2>(c, L, g, rho, E, st, ev, join);
    case 3:
      return launch_fields_v<3>(c, L, g, rho, E, st, ev, join);
    case 4:
      return launch_fields_v<4>(c, L, g, rho, E, st, ev, join);
    default:
      return launch_fields_v<0>(c, L, g, rho, E, st, ev, join);
  }
}

KickInputs kick_inputs(const int32_t* nmesh, int64_t n, double E_GeV, double zstep, int random_mesh,
                       double mesh_scale, double mesh_shift) {
  KickInputs in{};
  // sc.py:214-216 and coord_transform.py:19-20, 61 (same libm pow as numpy)
  in.gamref = E_GeV / kMeGeV;
  in.betaref = std::sqrt(1 - std::pow(in.gamref, -2.0));
  in.pref = kMeEV * std::sqrt(std::pow(in.gamref, 2.0) - 1);
  in.gbref = in.gamref * in.betaref;
  in.cdT = zstep / in.betaref;
  in.random_mesh = random_mesh ? 1 : 0;
  in.mesh_scale = random_mesh ? mesh_scale : 1.0;
  in.mesh_shift = random_mesh ? mesh_shift : 0.0;
  for (int k = 0; k < 3; ++k) in.nmesh[k] = nmesh[k];
  in.n = n;
  in.mode = 0;
  return in;
}

// grid of the strided reduction passes (A, B, moments): one full round of
// resident workgroups (no tail of half-empty rounds), at most one per 8 * 256
// particles (every thread sums several particles before the workgroup's
// reduction tree, which costs more than one particle's transform: batched
// passes over bunches of a few 1e5 particles ran at half the bandwidth)
constexpr int64_t kRedParticles = 8 * kPT;
int nred_blocks(const ocsc_ctx* c, int64_t n) {
  int64_t b = (n + kRedParticles - 1) / kRedParticles;
  return (int)(b < c->red_blocks ? (b < 1 ? 1 : b) : c->red_blocks);
}

// read back one timed kick of the profiling ring
int flush_slot(ocsc_ctx* c, int k) {
  if (!c->ev_pending[k]) return OCSC_OK;
  HIPCHK(hipEventSynchronize(c->ev[k][kNumStages]));
  for (int i = 0; i < kNumStages; ++i) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[k][i], c->ev[k][i + 1]));
    c->stage_ms[i] += ms;
  }
  c->stage_calls += 1;
  c->ev_pending[k] = false;
  return OCSC_OK;
}

int flush_all(ocsc_ctx* c) {
  int rc;
  for (int64_t j = c->ev_seq - ocsc_ctx::kEvRing; j < c->ev_seq; ++j)
    if (j >= 0 && (rc = flush_slot(c, (int)(j % ocsc_ctx::kEvRing)))) return rc;
  return OCSC_OK;
}

// the launch sequence of one kick (no host synchronisation, graph-capturable)
struct SortBufs;

// options of one kick inside a tracking loop
struct StepOpts {
  bool passA = true;   // false: the previous kick left its momentum sums in the partials
  int prev_nblk = 0;   // number of those partials
  bool drift = false;  // apply the next step's drift in the write-back
  bool sump = false;   // leave this kick's momentum sums for the next kick
  DriftArgs dr{0, 0};
  hipEvent_t kev0 = nullptr, kev1 = nullptr;  // recorded around the kick kernel (kick timing)
  bool kev_external = false;                  // capturing: the graph gets event-record nodes instead
  int32_t* cells = nullptr;                   // deposit writes each particle's NGP cell here (or null)
  // deposit and kick a cell-ordered copy of the bunch (frame and extent still
  // from the caller's rows), null: in place
  const SortBufs* sort = nullptr;
  // the kick writes records to rec_out[rec_perm[p]] instead of rows (the last
  // kick of a tracking call on a cell-ordered copy)
  const uint32_t* rec_perm = nullptr;
  PRec* rec_out = nullptr;
  // and reads them from there too (rec_out[rec_perm[p]]) instead of rows
  bool rec_in = false;
  // tracking: the charge grid is zero on entry (the previous step's kick
  // cleared it), and this step's kick clears it for the next step (the
  // z pass, its last reader, runs before the kick)
  bool grid_zeroed = false;
  bool zero_next = false;
};

// buffers of a kick on a cell-ordered copy: sort keys and values (ping-pong),
// histograms, particle records, the copy itself; order_out (may be null)
// receives the permutation (copy position -> caller position); scatter_back
// writes the kicked copy back to the caller's rows
struct SortBufs {
  uint32_t* k[2];
  uint32_t* v[2];
  uint32_t* hist;
  PRec* rec;
  double* rs;  // six rows, row stride n
  double* qs;
  uint32_t* order_out;
  bool scatter_back;
  // a cell order to re-use instead of sorting (copy position -> caller
  // position, a permutation of 0..n-1), or null
  const uint32_t* order_in = nullptr;
};

template <int NWORD, bool XCD, bool AGG, int TILE>
void launch_deposit_t(Rows R, const double* q, KickScalars* sc, unsigned long long* grid, int32_t* cells, int64_t n,
                      hipStream_t st) {
  constexpr int per = kDepPT * DepTile<TILE>::items;
  k_deposit<NWORD, XCD, AGG, TILE><<<(int)((n + per - 1) / per), kDepPT, 0, st>>>(R, q, sc, grid, cells);
}

// the deposit variant the context's options select
template <int NWORD>
void launch_deposit(const ocsc_ctx* c, Rows R, const double* q, KickScalars* sc, unsigned long long* grid,
                    int32_t* cells, int64_t n, hipStream_t st) {
  const int v = (c->deposit_xcd ? 4 : 0) + (c->deposit_agg ? 2 : 0);
  const int t = c->deposit_tile;
#define OCSC_DEP(V, T)                                                                                 \
  if (v == V && t == T) return launch_deposit_t<NWORD, (V & 4) != 0, (V & 2) != 0, T>(R, q, sc, grid, cells, n, st);
#define OCSC_DEP3(V) OCSC_DEP(V, 0) OCSC_DEP(V, 1) OCSC_DEP(V, 2) OCSC_DEP(V, 3) OCSC_DEP(V, 4)
  OCSC_DEP3(0) OCSC_DEP3(2) OCSC_DEP3(4) OCSC_DEP3(6)
#undef OCSC_DEP3
#undef OCSC_DEP
}

template <typename F, bool DRIFT, bool SUMP>
void launch_kick(ocsc_ctx* c, Lane& L, int npart, double* r, int64_t ld, const StepOpts& o, int64_t nz64,
                 hipStream_t st) {
  // mixed form: the three FP32 field grids; FP64 form: the k_diff64 fields
  const F* field = sizeof(F) == 8 ? (const F*)L.E64.p : (const F*)L.phi.p;
  unsigned long long* zg = o.zero_next ? (unsigned long long*)L.grid.p : nullptr;
  if (!DRIFT && !SUMP && o.rec_out && o.rec_in)
    k_kick<F, false, false, true, true><<<npart, kKickPT, 0, st>>>(r, ld, field, L.sc, o.dr, (double*)L.part.p,
                                                                   o.rec_perm, o.rec_out, zg, zg ? nz64 : 0);
  else if (!DRIFT && !SUMP && o.rec_out)
    k_kick<F, false, false, true><<<npart, kKickPT, 0, st>>>(r, ld, field, L.sc, o.dr, (double*)L.part.p, o.rec_perm,
                                                             o.rec_out, zg, zg ? nz64 : 0);
  else
    k_kick<F, DRIFT, SUMP><<<npart, kKickPT, 0, st>>>(r, ld, field, L.sc, o.dr, (double*)L.part.p, nullptr, nullptr,
                                                      zg, zg ? nz64 : 0);
}

template <typename F, int NWORD>
int enqueue_kick(ocsc_ctx* c, Lane& L, const Geometry& g, double* r, int64_t ld, const double* q, const KickInputs& in,
                 int prec, hipStream_t st, const StepOpts& o = StepOpts()) {
  const int64_t n = in.n;
  const int nblk = nred_blocks(c, n);
  const int npart = (int)((n + kKickPT - 1) / kKickPT);  // kick workgroups
  Rows R{r, ld};
  const bool prof = c->profiling != 0;
  const int slot = (int)(c->ev_seq % ocsc_ctx::kEvRing);
  hipEvent_t* ev = c->ev[slot];
  if (prof) {
    int rc0;
    if ((rc0 = flush_slot(c, slot))) return rc0;
  }
  auto mark = [&](int i) {
    if (prof) (void)hipEventRecord(ev[i], st);
  };
  // Side-stream work (not when profiling, so the stage times stay separable):
  // the kernel spectrum, which needs only the cell sizes, is built while the
  // deposit and the z and y passes run; the main stream waits for it only
  // before the x pass, its first reader.
  const bool fork = c->overlap && !prof;
  // the charge grid (all copies, all words): cleared by its own kernel on the
  // side stream while the frame passes run, the deposit waiting for it on the
  // second queue (option grid_clear 1), or by pass B (0; also when profiling
  // or without the second stream).  10M / 127^3 mixed, 20-step calls on one
  // GPU: 0.857 vs 0.870 ms per step (pass B's stores cost more than the
  // cross-queue wait).  Tracking with grid_clear 2: the previous step's kick
  // cleared it (o.grid_zeroed), no clear and no wait here.
  const int64_t nz64 = (c->deposit_xcd ? kNumXcd : 1) * NWORD * g.ncell;
  const bool side_clear = c->grid_clear && fork && !o.grid_zeroed;
  unsigned long long* const zg = side_clear || o.grid_zeroed ? nullptr : (unsigned long long*)L.grid.p;
  if (side_clear) {
    HIPCHK(hipEventRecord(c->ev_fork0, st));
    HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork0, 0));
    k_zero64<<<(int)std::min<int64_t>((nz64 + 255) / 256, 8192), 256, 0, c->side>>>(
        (unsigned long long*)L.grid.p, nz64);
    HIPCHK(hipEventRecord(c->ev_zero, c->side));
  }
  mark(0);
  const bool fold_frame = !o.passA && o.prev_nblk > kMaxRedBlocks && c->fold_frame;
  if (!fold_frame) {
    k_setup<<<1, 64, 0, st>>>(L.sc, in);
    LAUNCHCHK("k_setup");
  }
  if (fold_frame) {
    double* folded = (double*)L.part.p + (size_t)3 * o.prev_nblk;
    k_fold_frame<<<kFoldBlocks, 256, 0, st>>>((const double*)L.part.p, o.prev_nblk, folded, L.sc, in);
  } else if (o.passA) {
    k_passA<<<nblk, kPT, 0, st>>>(R, L.sc, (double*)L.part.p);
    k_reduceA<<<1, 1024, 0, st>>>(L.sc, (double*)L.part.p, nblk);
  } else if (o.prev_nblk > kMaxRedBlocks) {
    // one partial per kick block: fold them on many CUs first
    double* folded = (double*)L.part.p + (size_t)3 * o.prev_nblk;
    k_fold<3><<<kFoldBlocks, 256, 0, st>>>((const double*)L.part.p, o.prev_nblk, folded);
    k_reduceA<<<1, 1024, 0, st>>>(L.sc, folded, kFoldBlocks);
  } else {
    k_reduceA<<<1, 1024, 0, st>>>(L.sc, (double*)L.part.p, o.prev_nblk);
  }
  mark(1);
  // a kick that re-uses a cell order takes its particle records from pass B
  PRec* const recB = o.sort && o.sort->order_in ? o.sort->rec : nullptr;
  if (prec == OCSC_MIXED)
    k_passB<true><<<nblk, kPT, 0, st>>>(R, q, L.sc, (double*)L.part.p, zg, zg ? nz64 : 0, recB);
  else
    k_passB<false><<<nblk, kPT, 0, st>>>(R, q, L.sc, (double*)L.part.p, zg, zg ? nz64 : 0, recB);
  k_reduceB<<<1, 1024, 0, st>>>(L.sc, (double*)L.part.p, nblk, prec);
  LAUNCHCHK("frame/extent passes");
  mark(2);
  int rc;
  if (fork) {
    HIPCHK(hipEventRecord(c->ev_fork, st));
    HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    if ((rc = launch_kernel_spectrum<F>(c, L, g, c->side, 0))) return rc;
    HIPCHK(hipEventRecord(c->ev_join, c->side));
    if (side_clear) HIPCHK(hipStreamWaitEvent(st, c->ev_zero, 0));
  }
  // one-shot kick on a cell-ordered copy: the frame and extent above came
  // from the caller's order, the deposit sums integers and the kick is per
  // particle, so the result equals the caller-order kick bit for bit
  const uint32_t* order = nullptr;
  const SortBufs* sb = o.sort;
  // a one-shot kick on the cell order with the default deposit reads the
  // records through the order in the deposit and the kick: no sorted copy
  const bool rec_direct = sb && sb->scatter_back && !o.drift && !o.sump && !c->deposit_xcd && !c->deposit_agg &&
                          c->deposit_tile == 0 && c->rec_direct;
  if (sb) {
    const int nb = (int)((n + 255) / 256);
    if (sb->order_in) {
      // an earlier call's cell order (the records came from pass B): no sort
      order = sb->order_in;
    } else {
      const int end_bit = sort_end_bit(g.d.nx, g.d.ny, g.d.nz, c->sort_cells);
      k_sort_keys<<<(int)((n + kPT - 1) / kPT), kPT, 0, st>>>(R, q, L.sc, sb->k[0], sb->rec, c->sort_cells,
                                                              DriftArgs{0, 0});
      order = sb->v[rs_sort(sb->k, sb->v, n, end_bit, sb->hist, st)];
    }
    if (!rec_direct)
      k_gather_rec<false><<<nb, 256, 0, st>>>(sb->rec, order, sb->rs, n, sb->qs, n, DriftArgs{0, 0});
    if (sb->order_out) HIPCHK(hipMemcpyAsync(sb->order_out, order, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    LAUNCHCHK("cell-order copy");
    if (!rec_direct) {
      R = Rows{sb->rs, n};
      q = sb->qs;
    }
  }
  if (rec_direct) {
    constexpr int per = kDepPT * DepTile<0>::items;
    k_deposit<NWORD, false, false, 0, true><<<(int)((n + per - 1) / per), kDepPT, 0, st>>>(
        R, q, L.sc, (unsigned long long*)L.grid.p, nullptr, sb->rec, order);
  } else {
    launch_deposit<NWORD>(c, R, q, L.sc, (unsigned long long*)L.grid.p, o.cells, n, st);
  }
  LAUNCHCHK("deposit");
  mark(3);
  if (!fork && (rc = launch_kernel_spectrum<F>(c, L, g, st, 0))) return rc;
  mark(4);
  if constexpr (sizeof(F) == 4) {
    if ((rc = launch_fields(c, L, g, L.grid.p, (float*)L.phi.p, st, prof ? ev + 4 : nullptr,
                            fork ? c->ev_join : nullptr)))
      return rc;
  } else {
    if ((rc = launch_convolution<F, NWORD == 2 ? kChargeFix2 : kChargeFix1>(c, L, g, L.grid.p, (F*)L.phi.p, st,
                                                                              prof ? ev + 4 : nullptr,
                                                                              fork ? c->ev_join : nullptr)))
      return rc;
    k_diff64<<<(int)((g.ncell + 255) / 256), 256, 0, st>>>((const double*)L.phi.p, L.sc, (double*)L.E64.p);
    LAUNCHCHK("k_diff64");
  }
  mark(9);
  if (o.kev0 && !o.kev_external) HIPCHK(hipEventRecord(o.kev0, st));
  double* rk = sb ? sb->rs : r;
  const int64_t ldk = sb ? n : ld;
  StepOpts ok = o;
  ok.zero_next = o.zero_next && !(sb && sb->scatter_back && !o.drift && !o.sump);
  if (sb && sb->scatter_back && !o.drift && !o.sump) {
    // the kick writes each record straight to its caller position (and, with
    // rec_direct, reads it from there)
    ok.rec_perm = order;
    ok.rec_out = sb->rec;
    ok.rec_in = rec_direct;
  }
  if (o.drift && o.sump)
    launch_kick<F, true, true>(c, L, npart, rk, ldk, ok, nz64, st);
  else if (o.drift)
    launch_kick<F, true, false>(c, L, npart, rk, ldk, ok, nz64, st);
  else if (o.sump)
    launch_kick<F, false, true>(c, L, npart, rk, ldk, ok, nz64, st);
  else
    launch_kick<F, false, false>(c, L, npart, rk, ldk, ok, nz64, st);
  LAUNCHCHK("k_kick");
  if (sb && sb->scatter_back) {
    // back to the caller's order through whole records; a bunch the kick
    // rejected (error flag) was not written, so nothing is copied back
    const int nb = (int)((n + 255) / 256);
    if (!ok.rec_out) k_scatter_rec<<<nb, 256, 0, st>>>(rk, n, order, sb->rec, n);
    k_rec_to_rows_checked<<<nb, 256, 0, st>>>(sb->rec, r, ld, n, L.sc);
    LAUNCHCHK("caller-order copy");
  }
  if (o.kev1 && !o.kev_external) HIPCHK(hipEventRecord(o.kev1, st));
  mark(10);
  if (prof) {
    c->ev_pending[slot] = true;
    c->ev_seq += 1;
  }
  return OCSC_OK;
}

int enqueue_kick_prec(ocsc_ctx* c, Lane& L, const Geometry& g, double* r, int64_t ld, const double* q, const KickInputs& in,
                      int prec, hipStream_t st, const StepOpts& o = StepOpts()) {
  if (prec == OCSC_FP64) return enqueue_kick<double, 2>(c, L, g, r, ld, q, in, prec, st, o);
  return enqueue_kick<float, 1>(c, L, g, r, ld, q, in, prec, st, o);
}

int ensure_part(ocsc_ctx* c, Lane& L, int64_t n) {
  const size_t need =
      std::max<size_t>((size_t)kMaxRedBlocks * kNPartB, (size_t)3 * ((n + kKickPT - 1) / kKickPT) + 3 * kFoldBlocks);
  return ensure(L.part, need * sizeof(double));
}

int ensure_prec(ocsc_ctx* c, Lane& L, const Geometry& g, int prec) {
  if (prec == OCSC_FP64) return ensure_mesh<double>(c, L, g, 2);
  return ensure_mesh<float>(c, L, g, 1);
}

__global__ void k_clear_err(KickScalars* __restrict__ s) {
  if (threadIdx.x == 0) {
    s->err = 0;
    s->err_sticky = 0;
  }
}

// error bits of the kicks since the last report (current kick and earlier
// deferred ones); reported bits are cleared on the device
int read_errors(ocsc_ctx* c, Lane& L, hipStream_t st, int32_t* err) {
  HIPCHK(hipMemcpyAsync(L.sc_host, L.sc, sizeof(KickScalars), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *err = L.sc_host->err | L.sc_host->err_sticky;
  if (*err) {
    k_clear_err<<<1, 64, 0, st>>>(L.sc);
    LAUNCHCHK("k_clear_err");
    HIPCHK(hipStreamSynchronize(st));
  }
  return OCSC_OK;
}

int value_error(int32_t err) {
  std::string m;
  if (err & kErrNonFinite) m += "non-finite particle coordinates or charges; ";
  if (err & kErrZeroCharge) m += "total charge is zero (charge centroid undefined); ";
  if (err & kErrSteps) m += "zero or non-finite bunch extent on some axis; ";
  if (err & kErrCell) m += "particle outside the mesh; ";
  return fail(OCSC_EVALUE, m.empty() ? std::string("invalid bunch") : m);
}

int check_common(ocsc_ctx* c, const void* r, const double* q, int64_t n, int64_t ld) {
  if (!c) return fail(OCSC_EINVAL, "null context");
  if (n < 0) return fail(OCSC_EINVAL, "negative particle count");
  if (n == 0) return fail(OCSC_EVALUE, "empty bunch (zero-size array to reduction operation)");
  if (!r || !q) return fail(OCSC_EINVAL, "null particle or charge pointer");
  if (ld < n) return fail(OCSC_EINVAL, "row stride ld must be >= n");
  return OCSC_OK;
}

}  // namespace

namespace {

// bitwise equality of two kick inputs (the values baked into a set-up node)
bool same_inputs(const KickInputs& a, const KickInputs& b) {
  auto eq = [](const void* x, const void* y, size_t nbytes) { return std::memcmp(x, y, nbytes) == 0; };
  return eq(&a.gamref, &b.gamref, 8) && eq(&a.betaref, &b.betaref, 8) && eq(&a.pref, &b.pref, 8) &&
         eq(&a.gbref, &b.gbref, 8) && eq(&a.cdT, &b.cdT, 8) && eq(&a.mesh_scale, &b.mesh_scale, 8) &&
         eq(&a.mesh_shift, &b.mesh_shift, 8) && eq(&a.gamma0, &b.gamma0, 8) && eq(a.steps, b.steps, 24) &&
         a.random_mesh == b.random_mesh && a.mode == b.mode && eq(a.nmesh, b.nmesh, 12) && a.n == b.n &&
         a.draws == b.draws && a.draw_ctr == b.draw_ctr && a.ndraws == b.ndraws;
}

void drop_track_graphs(ocsc_ctx* c) {
  for (int k = 0; k < 2; ++k) {
    if (c->tg.exec[k]) (void)hipGraphExecDestroy(c->tg.exec[k]);
    if (c->tg.graphs[k]) (void)hipGraphDestroy(c->tg.graphs[k]);
  }
  c->tg = ocsc_ctx::TrackGraphs{};
}

void drop_apply_slot(ocsc_ctx::ApplyGraph& A) {
  if (A.done) (void)hipEventSynchronize(A.done);
  if (A.exec) (void)hipGraphExecDestroy(A.exec);
  if (A.graph) (void)hipGraphDestroy(A.graph);
  if (A.done) (void)hipEventDestroy(A.done);
  A = ocsc_ctx::ApplyGraph{};
}

void drop_apply_graph(ocsc_ctx* c) {
  for (auto& A : c->ag) drop_apply_slot(A);
}

// One-shot kick from a captured graph.  A kick is some 30 launches, many of
// them a few microseconds long; launched one by one after the host has waited
// for the previous kick, the GPU idles between them.  The sequence depends on
// the caller's pointers and sizes, the mesh, the precision, the options and
// the workspace addresses; while those are unchanged the graph is replayed,
// with only the set-up node's inputs (energy, step, random-mesh draws)
// replaced when they differ.
int apply_via_graph(ocsc_ctx* c, Lane& L, const Geometry& g, double* r, int64_t ld, const double* q,
                    const KickInputs& in, int prec, hipStream_t st, const StepOpts& o) {
  const bool reuse = o.sort && o.sort->order_in;
  auto& A = c->ag[reuse ? 1 : 0];
  const std::vector<int64_t> key = {(int64_t)(intptr_t)r, ld, (int64_t)(intptr_t)q, in.n, g.d.nx, g.d.ny, g.d.nz,
                                    prec, o.sort != nullptr, reuse, o.sort && o.sort->order_out, c->rec_direct,
                                    c->overlap,
                                    c->deposit_xcd, c->deposit_agg,
                                    c->deposit_tile, c->fft_plan, c->sort_cells, c->red_blocks, c->grid_clear,
                                    c->igf_blocks, (int64_t)g_alloc_gen.load()};
  if (!A.exec || A.key != key) {
    drop_apply_slot(A);  // the other slot's sequence stays cached
    HIPCHK(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_kick_prec(c, L, g, r, ld, q, in, prec, c->cap_stream, o);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->cap_stream, &graph);
    if (rc || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc ? rc : fail(OCSC_EHIP, std::string("apply capture: ") + hipGetErrorString(e));
    }
    A.graph = graph;
    size_t nn = 0;
    HIPCHK(hipGraphGetNodes(graph, nullptr, &nn));
    std::vector<hipGraphNode_t> nodes(nn);
    HIPCHK(hipGraphGetNodes(graph, nodes.data(), &nn));
    for (hipGraphNode_t nd : nodes) {
      hipGraphNodeType t;
      if (hipGraphNodeGetType(nd, &t) != hipSuccess || t != hipGraphNodeTypeKernel) continue;
      hipKernelNodeParams p{};
      if (hipGraphKernelNodeGetParams(nd, &p) == hipSuccess && p.func == (void*)k_setup) A.setup = nd;
    }
    if (!A.setup) {
      // the runtime did not expose the node: kick eagerly from now on
      drop_apply_graph(c);
      c->apply_graph = 0;
      return enqueue_kick_prec(c, L, g, r, ld, q, in, prec, st, o);
    }
    const hipError_t ei = hipGraphInstantiate(&A.exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      drop_apply_graph(c);
      return fail(OCSC_EHIP, std::string("apply instantiate: ") + hipGetErrorString(ei));
    }
    HIPCHK(hipEventCreateWithFlags(&A.done, hipEventDisableTiming));
    A.in = in;
    A.key = key;
  } else if (!same_inputs(A.in, in)) {
    // a replay still in flight (calls without sync) keeps its inputs
    HIPCHK(hipEventSynchronize(A.done));
    hipKernelNodeParams p{};
    HIPCHK(hipGraphKernelNodeGetParams(A.setup, &p));
    KickScalars* scp = L.sc;
    KickInputs inv = in;
    void* args[2] = {&scp, &inv};
    p.kernelParams = args;
    p.extra = nullptr;
    const hipError_t e = hipGraphExecKernelNodeSetParams(A.exec, A.setup, &p);
    if (e != hipSuccess) {
      drop_apply_graph(c);
      return fail(OCSC_EHIP, std::string("apply graph update: ") + hipGetErrorString(e));
    }
    A.in = in;
  }
  const hipError_t e = hipGraphLaunch(A.exec, st);
  if (e != hipSuccess) {
    drop_apply_graph(c);
    return fail(OCSC_EHIP, std::string("apply graph launch: ") + hipGetErrorString(e));
  }
  HIPCHK(hipEventRecord(A.done, st));
  return OCSC_OK;
}

constexpr int kMaxLanes = 8;
// one-shot FP64 kicks of at least this many particles go through a
// cell-ordered copy (measured at 10M / 127^3: FP64 2.58 -> 2.42 ms, where the
// two-word deposit's random atomics dominate; mixed 1.77 -> 2.29 ms, so the
// mixed form stays in the caller's order)
constexpr int64_t kApplySortMin = 1000000;

// device address of the no-allocation workspace on `device` (current)
int scratch_for_device(int device, void** out);

int init_lane(Lane& L, bool own_stream) {
  if (hipMalloc(&L.sc, sizeof(KickScalars)) != hipSuccess) return fail(OCSC_ENOMEM, "scalars");
  if (hipHostMalloc(&L.sc_host, sizeof(KickScalars)) != hipSuccess) return fail(OCSC_ENOMEM, "pinned");
  KickScalars init{};
  init.key_prec = -1;
  HIPCHK(hipMemcpy(L.sc, &init, sizeof(init), hipMemcpyHostToDevice));
  if (own_stream) {
    HIPCHK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
  }
  return OCSC_OK;
}

void free_lane(Lane& L) {
  for (Buf* b : {&L.part, &L.grid, &L.A, &L.B, &L.KA, &L.KB, &L.KF, &L.phi, &L.G, &L.E64, &L.sk[0], &L.sk[1], &L.sv,
                 &L.sh, &L.srec})
    if (b->p) (void)hipFree(b->p);
  if (L.sc) (void)hipFree(L.sc);
  if (L.sc_host) (void)hipHostFree(L.sc_host);
  if (L.stream) (void)hipStreamDestroy(L.stream);
  if (L.done) (void)hipEventDestroy(L.done);
}

}  // namespace

extern "C" {

const char* ocsc_last_error(void) { return g_err.c_str(); }

int ocsc_create(ocsc_ctx** out, int device) {
  if (!out) return fail(OCSC_EINVAL, "null output pointer");
  *out = nullptr;
  HIPCHK(hipSetDevice(device));
  ocsc_ctx* c = new ocsc_ctx();
  c->device = device;
  Lane& L = c->lane0;
  auto cleanup = [&](int rc) {
    ocsc_destroy(c);
    return rc;
  };
  {
    // resolve this device's no-allocation workspace once, at set-up
    void* sp = nullptr;
    const int rc = scratch_for_device(device, &sp);
    if (rc) return cleanup(rc);
  }
  if (hipMalloc(&L.sc, sizeof(KickScalars)) != hipSuccess) return cleanup(fail(OCSC_ENOMEM, "scalars"));
  if (hipHostMalloc(&L.sc_host, sizeof(KickScalars)) != hipSuccess) return cleanup(fail(OCSC_ENOMEM, "pinned"));
  if (hipMalloc(&c->twf, sizeof(cpx<float>) * 2048) != hipSuccess) return cleanup(fail(OCSC_ENOMEM, "twiddles"));
  if (hipMalloc(&c->twd, sizeof(cpx<double>) * 2048) != hipSuccess) return cleanup(fail(OCSC_ENOMEM, "twiddles"));
  // W_M^k = exp(-2 pi i k / M), from the first octant by symmetry so every
  // entry is a correctly rounded double of the exact value to within 1 ulp
  cpx<double> hd[2048];
  cpx<float> hf[2048];
  for (int M = 8; M <= 1024; M *= 2) {
    for (int k = 0; k < M; ++k) {
      const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)M;
      const double cr = (double)cosl(a), si = (double)-sinl(a);
      hd[M + k] = cpx<double>{cr, si};
      hf[M + k] = cpx<float>{(float)cosl(a), (float)-sinl(a)};
    }
  }
  hd[0] = hd[1] = hd[2] = hd[3] = hd[4] = hd[5] = hd[6] = hd[7] = cpx<double>{1, 0};
  for (int k = 0; k < 8; ++k) hf[k] = cpx<float>{1, 0};
  if (hipMemcpy(c->twd, hd, sizeof(hd), hipMemcpyHostToDevice) != hipSuccess) return cleanup(fail(OCSC_EHIP, "copy"));
  if (hipMemcpy(c->twf, hf, sizeof(hf), hipMemcpyHostToDevice) != hipSuccess) return cleanup(fail(OCSC_EHIP, "copy"));
  {
    int ncu = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_passB<true>, kPT, 0) == hipSuccess && ncu > 0 &&
        per_cu > 0)
      c->red_blocks = std::min(kMaxRedBlocks, ncu * per_cu);
  }
  KickScalars init{};
  init.key_prec = -1;
  if (hipMemcpy(L.sc, &init, sizeof(init), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(OCSC_EHIP, "copy"));
  if (hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(OCSC_EHIP, "stream"));
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) return cleanup(fail(OCSC_EHIP, "stream"));
  if (hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork0, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_zero, hipEventDisableTiming) != hipSuccess)
    return cleanup(fail(OCSC_EHIP, "event"));
  *out = c;
  return OCSC_OK;
}

int ocsc_destroy(ocsc_ctx* c) {
  if (!c) return OCSC_OK;
  Lane& L = c->lane0;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  for (Buf* b : {&L.part, &L.grid, &L.A, &L.B, &L.KA, &L.KB, &L.KF, &L.phi, &L.G, &c->rs[0], &c->rs[1], &c->qs[0],
                 &c->qs[1],
                 &c->keys[0],
                 &L.E64, &c->draws, &c->keys[1], &c->vals, &c->perm[0], &c->perm[1], &c->sort_tmp, &c->cic,
                 &c->mom_part, &c->mom_tab, &c->mom_ctr, &c->lsc_bins, &c->lsc_cic, &c->lsc_fac, &c->csr_bins,
                 &c->rec, &c->apply_order})
    if (b->p) (void)hipFree(b->p);
  if (c->lsc) (void)hipFree(c->lsc);
  if (c->lsc_plan) (void)hipFree(c->lsc_plan);
  if (c->imax_plan) (void)hipFree(c->imax_plan);
  if (c->imax_err) (void)hipFree(c->imax_err);
  if (c->lsc_err) (void)hipFree(c->lsc_err);
  if (c->lsc_host) (void)hipHostFree(c->lsc_host);
  if (L.sc) (void)hipFree(L.sc);
  if (L.sc_host) (void)hipHostFree(L.sc_host);
  if (c->mom) (void)hipFree(c->mom);
  if (c->mom_host) (void)hipHostFree(c->mom_host);
  if (c->twf) (void)hipFree(c->twf);
  if (c->twd) (void)hipFree(c->twd);
  drop_apply_graph(c);
  drop_track_graphs(c);
  if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_fork0) (void)hipEventDestroy(c->ev_fork0);
  if (c->ev_zero) (void)hipEventDestroy(c->ev_zero);
  for (auto& set : c->ev)
    for (hipEvent_t e : set)
      if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->kick_ev) (void)hipEventDestroy(e);
  for (Lane* l : c->lanes) {
    free_lane(*l);
    delete l;
  }
  for (Lane* l : c->flanes) {
    free_lane(*l);
    delete l;
  }
  if (c->btab.p) (void)hipFree(c->btab.p);
  if (c->btab2.p) (void)hipFree(c->btab2.p);
  for (ocsc_ctx::BunchCopy* bc : c->bcopy) {
    for (Buf* b : {&bc->rs, &bc->qs, &bc->perm})
      if (b->p) (void)hipFree(b->p);
    delete bc;
  }
  delete c;
  return OCSC_OK;
}

int ocsc_apply(ocsc_ctx* c, double* r, int64_t ld, const double* q, int64_t n, const int32_t nmesh[3], double E_GeV,
               double zstep, int random_mesh, double mesh_scale, double mesh_shift, int prec, int sync,
               void* stream) {
  return ocsc_apply_cells(c, r, ld, q, n, nmesh, E_GeV, zstep, random_mesh, mesh_scale, mesh_shift, prec, sync,
                          nullptr, stream);
}

int ocsc_apply_cells(ocsc_ctx* c, double* r, int64_t ld, const double* q, int64_t n, const int32_t nmesh[3],
                     double E_GeV, double zstep, int random_mesh, double mesh_scale, double mesh_shift, int prec,
                     int sync, int32_t* cells, void* stream) {
  if (c && zstep == 0) return OCSC_OK;  // sc.py:210-212
  int rc;
  if ((rc = check_common(c, r, q, n, ld))) return rc;
  if (!nmesh) return fail(OCSC_EINVAL, "null nmesh");
  if (prec != OCSC_FP64 && prec != OCSC_MIXED) return fail(OCSC_EINVAL, "precision must be OCSC_FP64 or OCSC_MIXED");
  if (!(E_GeV > 0) || !std::isfinite(zstep)) return fail(OCSC_EINVAL, "E_GeV must be > 0 and zstep finite");
  HIPCHK(hipSetDevice(c->device));
  Lane& L = c->lane0;
  Geometry g;
  if ((rc = make_geometry(nmesh, g))) return rc;
  if ((rc = ensure_prec(c, L, g, prec))) return rc;
  if ((rc = ensure_part(c, L, n))) return rc;
  hipStream_t st = (hipStream_t)stream;
  KickInputs in = kick_inputs(nmesh, n, E_GeV, zstep, random_mesh, mesh_scale, mesh_shift);
  StepOpts o;
  o.cells = cells;
  // cell-ordered copy for the deposit and the kick (not with cells: those
  // come in the caller's order)
  SortBufs sb{};
  bool reuse = false, keep = false;
  if (!cells && n <= (int64_t)0xFFFFFFFF &&
      (c->apply_sort == 1 || (c->apply_sort == 2 && prec == OCSC_FP64 && n >= kApplySortMin))) {
    for (int k = 0; k < 2; ++k)
      if ((rc = ensure(c->keys[k], (size_t)n * 4))) return rc;
    if ((rc = ensure(c->vals, (size_t)2 * n * 4)) || (rc = ensure(c->sort_tmp, rs_hist_bytes(n))) ||
        (rc = ensure(c->rec, (size_t)n * sizeof(PRec))) || (rc = ensure(c->rs[0], (size_t)6 * n * 8)) ||
        (rc = ensure(c->qs[0], (size_t)n * 8)))
      return rc;
    // the cell order of a recent call on the same rows, or a sort whose order
    // is kept for the next calls
    reuse = c->apply_resort > 1 && c->ao_age > 0 && c->ao_age < c->apply_resort && c->ao_r == r && c->ao_n == n &&
            c->ao_ld == ld;
    keep = !reuse && c->apply_resort > 1;
    if (keep && (rc = ensure(c->apply_order, (size_t)n * 4))) return rc;
    c->ao_age = reuse ? c->ao_age : 0;  // set again once this call is enqueued
    sb = SortBufs{{(uint32_t*)c->keys[0].p, (uint32_t*)c->keys[1].p},
                  {(uint32_t*)c->vals.p, (uint32_t*)c->vals.p + n},
                  (uint32_t*)c->sort_tmp.p, (PRec*)c->rec.p, (double*)c->rs[0].p, (double*)c->qs[0].p,
                  keep ? (uint32_t*)c->apply_order.p : nullptr, true};
    sb.order_in = reuse ? (const uint32_t*)c->apply_order.p : nullptr;
    o.sort = &sb;
  }
  if (c->apply_graph && !cells && !c->profiling)
    rc = apply_via_graph(c, L, g, r, ld, q, in, prec, st, o);
  else
    rc = enqueue_kick_prec(c, L, g, r, ld, q, in, prec, st, o);
  if (rc) {
    c->ao_age = 0;
    return rc;
  }
  if (reuse) {
    c->ao_age += 1;
  } else if (keep) {
    c->ao_r = r;
    c->ao_n = n;
    c->ao_ld = ld;
    c->ao_age = 1;
  }
  if (sync) {
    int32_t err = 0;
    if ((rc = read_errors(c, L, st, &err))) return rc;
    if (err) return value_error(err);
  }
  return OCSC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Kicks without device allocation (ocsc_apply_noalloc): every workspace of
// the kick (scalars, partials, charge grid, FFT scratch, kernel spectrum,
// fields) lives in one statically sized device array.  In host code the
// array's name is only a stand-in; its device address comes from
// hipGetSymbolAddress, taken once per device with that device current (in
// ocsc_create) and kept in a per-device table.
namespace {

constexpr size_t kScratchBytes = OCSC_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device `device`
int scratch_for_device(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return fail(OCSC_EINVAL, "device ordinal out of range for the scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(OCSC_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return OCSC_OK;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// lay the workspaces of one kick out in the device's scratch array
int scratch_lane(ocsc_ctx* c, const Geometry& g, int prec, Lane* S) {
  void* base = nullptr;
  int rc;
  if ((rc = scratch_for_device(c->device, &base))) return rc;
  const MeshSizes m = prec == OCSC_FP64 ? mesh_sizes<double>(g, 2, c->deposit_xcd)
                                        : mesh_sizes<float>(g, 1, c->deposit_xcd);
  size_t off = 0;
  auto take = [&](Buf& b, size_t bytes) {
    b.p = bytes ? (void*)((char*)base + off) : nullptr;
    b.cap = bytes;
    off += align256(bytes);
  };
  S->sc = (KickScalars*)base;
  off = align256(sizeof(KickScalars));
  // partials of the frame and extent passes (the one-shot kick writes none)
  take(S->part, (size_t)kMaxRedBlocks * kNPartB * sizeof(double));
  take(S->grid, m.grid);
  take(S->A, m.A);
  take(S->B, m.B);
  take(S->KB, m.KB);
  take(S->KA, m.KA);
  take(S->KF, m.KF);
  take(S->phi, m.phi);
  take(S->G, m.G);
  take(S->E64, m.E64);
  if (off > kScratchBytes)
    return fail(OCSC_EINVAL, "mesh " + std::to_string(g.d.nx) + "x" + std::to_string(g.d.ny) + "x" +
                                 std::to_string(g.d.nz) + " needs " + std::to_string(off) +
                                 " bytes of workspace; ocsc_apply_noalloc holds at most " +
                                 std::to_string(kScratchBytes) +
                                 " (every mesh up to 64^3 in either precision, with option deposit_xcd = 0)");
  S->sc_host = c->lane0.sc_host;  // pinned host mirror for the error read-back
  return OCSC_OK;
}

}  // namespace

extern "C" {

int ocsc_apply_noalloc(ocsc_ctx* c, double* r, int64_t ld, const double* q, int64_t n, const int32_t nmesh[3],
                       double E_GeV, double zstep, int random_mesh, double mesh_scale, double mesh_shift, int prec,
                       int sync, void* stream) {
  if (c && zstep == 0) return OCSC_OK;  // sc.py:210-212
  int rc;
  if ((rc = check_common(c, r, q, n, ld))) return rc;
  if (!nmesh) return fail(OCSC_EINVAL, "null nmesh");
  if (prec != OCSC_FP64 && prec != OCSC_MIXED) return fail(OCSC_EINVAL, "precision must be OCSC_FP64 or OCSC_MIXED");
  if (!(E_GeV > 0) || !std::isfinite(zstep)) return fail(OCSC_EINVAL, "E_GeV must be > 0 and zstep finite");
  HIPCHK(hipSetDevice(c->device));
  Geometry g;
  if ((rc = make_geometry(nmesh, g))) return rc;
  Lane S;
  if ((rc = scratch_lane(c, g, prec, &S))) return rc;
  hipStream_t st = (hipStream_t)stream;
  const KickInputs in = kick_inputs(nmesh, n, E_GeV, zstep, random_mesh, mesh_scale, mesh_shift);
  if ((rc = enqueue_kick_prec(c, S, g, r, ld, q, in, prec, st))) return rc;
  if (sync) {
    int32_t err = 0;
    if ((rc = read_errors(c, S, st, &err))) return rc;
    if (err) return value_error(err);
  }
  return OCSC_OK;
}

int ocsc_check(ocsc_ctx* c, void* stream) {
  if (!c) return fail(OCSC_EINVAL, "null context");
  HIPCHK(hipSetDevice(c->device));
  Lane& L = c->lane0;
  int32_t err = 0;
  int rc;
  if ((rc = read_errors(c, L, (hipStream_t)stream, &err))) return rc;
  if (err) return value_error(err);
  return OCSC_OK;
}

int ocsc_el_field(ocsc_ctx* c, const double* xyz, int64_t ld, const double* q, int64_t n, const int32_t nmesh[3],
                  double gamma, int random_mesh, double mesh_scale, double mesh_shift, int prec, double* E,
                  int64_t lde, int32_t* cells, void* stream) {
  int rc;
  if ((rc = check_common(c, xyz, q, n, ld))) return rc;
  if (!nmesh || !E || lde < n) return fail(OCSC_EINVAL, "bad nmesh / Exyz / lde");
  if (prec != OCSC_FP64 && prec != OCSC_MIXED) return fail(OCSC_EINVAL, "bad precision");
  HIPCHK(hipSetDevice(c->device));
  Lane& L = c->lane0;
  Geometry g;
  if ((rc = make_geometry(nmesh, g))) return rc;
  if ((rc = ensure_prec(c, L, g, prec))) return rc;
  if ((rc = ensure_part(c, L, n))) return rc;
  hipStream_t st = (hipStream_t)stream;
  KickInputs in = kick_inputs(nmesh, n, 1.0, 1.0, random_mesh, mesh_scale, mesh_shift);
  in.mode = 1;
  struct XcdOff {  // el_field deposits into a single grid copy
    ocsc_ctx* c;
    int saved;
    ~XcdOff() { c->deposit_xcd = saved; }
  } xcd_off{c, c->deposit_xcd};
  c->deposit_xcd = 0;
  in.gamma0 = gamma;
  const int nblk = nred_blocks(c, n);
  const int npart = (int)((n + kPT - 1) / kPT);
  Rows R{xyz, ld};
  k_setup<<<1, 64, 0, st>>>(L.sc, in);
  k_passB_xyz<<<nblk, kPT, 0, st>>>(R, q, L.sc, (double*)L.part.p);
  k_reduceB<<<1, 1024, 0, st>>>(L.sc, (double*)L.part.p, nblk, prec);
  const int nword = prec == OCSC_FP64 ? 2 : 1;
  const int64_t nz64 = nword * g.ncell;
  k_zero64<<<(int)std::min<int64_t>((nz64 + 255) / 256, 4096), 256, 0, st>>>((unsigned long long*)L.grid.p, nz64);
  LAUNCHCHK("el_field setup");
  if (prec == OCSC_FP64) {
    k_deposit_xyz<2><<<npart, kPT, 0, st>>>(R, q, L.sc, (unsigned long long*)L.grid.p);
    if ((rc = launch_kernel_spectrum<double>(c, L, g, st, 0))) return rc;
    if ((rc = launch_convolution<double, kChargeFix2>(c, L, g, L.grid.p, (double*)L.phi.p, st))) return rc;
    k_gather_xyz<double><<<npart, kPT, 0, st>>>(R, (const double*)L.phi.p, L.sc, E, lde, cells);
  } else {
    k_deposit_xyz<1><<<npart, kPT, 0, st>>>(R, q, L.sc, (unsigned long long*)L.grid.p);
    if ((rc = launch_kernel_spectrum<float>(c, L, g, st, 0))) return rc;
    if ((rc = launch_fields(c, L, g, L.grid.p, (float*)L.phi.p, st))) return rc;
    k_gather_xyz<float><<<npart, kPT, 0, st>>>(R, (const float*)L.phi.p, L.sc, E, lde, cells);
  }
  LAUNCHCHK("el_field");
  int32_t err = 0;
  if ((rc = read_errors(c, L, st, &err))) return rc;
  if (err) return value_error(err);
  return OCSC_OK;
}

int ocsc_sym_kernel(ocsc_ctx* c, const int32_t nmesh[3], const double h[3], double* K, void* stream) {
  if (!c || !nmesh || !h || !K) return fail(OCSC_EINVAL, "null argument");
  for (int k = 0; k < 3; ++k)
    if (nmesh[k] < 1 || nmesh[k] > 1024 || !(h[k] > 0)) return fail(OCSC_EINVAL, "bad shape or cell size");
  HIPCHK(hipSetDevice(c->device));
  Lane& L = c->lane0;
  hipStream_t st = (hipStream_t)stream;
  KickInputs in{};
  in.mode = 2;
  for (int k = 0; k < 3; ++k) {
    in.nmesh[k] = nmesh[k];
    in.steps[k] = h[k];
  }
  k_setup<<<1, 64, 0, st>>>(L.sc, in);
  MeshDims d{};
  d.nx = nmesh[0];
  d.ny = nmesh[1];
  d.nz = nmesh[2];
  int rc;
  const int64_t nn = (int64_t)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
  if ((rc = ensure(L.KF, (size_t)nn * 8))) return rc;
  k_igf_nodes<<<(int)std::min<int64_t>((nn + 255) / 256, 65536), 256, 0, st>>>((double*)L.KF.p, d, L.sc, 1);
  k_igf_diff<<<(int)std::min<int64_t>((igf_diff_items(d) + 255) / 256, 65536), 256, 0, st>>>((const double*)L.KF.p,
                                                                                              K, d, L.sc, 1);
  LAUNCHCHK("k_igf");
  HIPCHK(hipStreamSynchronize(st));
  return OCSC_OK;
}

int ocsc_potential(ocsc_ctx* c, const double* rho, const int32_t nmesh[3], const double h[3], void* phi, int prec,
                   void* stream) {
  if (!c || !rho || !nmesh || !h || !phi) return fail(OCSC_EINVAL, "null argument");
  if (prec != OCSC_FP64 && prec != OCSC_MIXED) return fail(OCSC_EINVAL, "bad precision");
  for (int k = 0; k < 3; ++k)
    if (!(h[k] > 0)) return fail(OCSC_EINVAL, "cell sizes must be > 0");
  int rc;
  HIPCHK(hipSetDevice(c->device));
  Lane& L = c->lane0;
  Geometry g;
  if ((rc = make_geometry(nmesh, g))) return rc;
  if ((rc = ensure_prec(c, L, g, prec))) return rc;
  hipStream_t st = (hipStream_t)stream;
  KickInputs in{};
  in.mode = 2;
  for (int k = 0; k < 3; ++k) {
    in.nmesh[k] = nmesh[k];
    in.steps[k] = h[k];
  }
  k_setup<<<1, 64, 0, st>>>(L.sc, in);
  if (prec == OCSC_FP64) {
    if ((rc = launch_kernel_spectrum<double>(c, L, g, st, 1))) return rc;
    if ((rc = launch_convolution<double, kChargeReal>(c, L, g, rho, (double*)phi, st))) return rc;
  } else {
    if ((rc = launch_kernel_spectrum<float>(c, L, g, st, 1))) return rc;
    if ((rc = launch_convolution<float, kChargeReal>(c, L, g, rho, (float*)phi, st))) return rc;
  }
  HIPCHK(hipStreamSynchronize(st));
  return OCSC_OK;
}

int ocsc_drift(ocsc_ctx* c, double* r, int64_t ld, int64_t n, double L, double E_GeV, void* stream) {
  if (!c || (!r && n > 0)) return fail(OCSC_EINVAL, "null argument");
  if (n < 0 || ld < n) return fail(OCSC_EINVAL, "bad n / ld");
  if (n == 0) return OCSC_OK;
  HIPCHK(hipSetDevice(c->device));
  // r_matrix.py:54-85 with k1 = hx = 0
  const double gamma = E_GeV / kMeGeV;
  const double igamma2 = gamma != 0 ? 1. / (gamma * gamma) : 0.;
  const double beta = std::sqrt(1. - igamma2);
  const double r56 = 0.0 - L / (beta * beta) * igamma2;
  k_drift<<<(int)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(r, ld, n, L, r56);
  LAUNCHCHK("k_drift");
  return OCSC_OK;
}

}  // extern "C"

namespace {

// Cloud-in-cell weights of tau on the grid a + j ds, j < nbins, as exact
// fixed-point sums (moments.hpp k_cic); *tot points at the nbins totals in buf.
int cic_profile(Buf& buf, const Rows& R, int64_t n, double a, double ds, int nbins, hipStream_t st, long long** tot) {
  int e = 0;
  (void)std::frexp((double)n, &e);
  const double wscale = std::ldexp(1.0, 62 - e);  // n * wscale < 2^62
  const int ncic = kCicBlocks;
  const int win = std::min(nbins, kCicLds);
  int rc;
  if ((rc = ensure(buf, ((size_t)ncic * win + nbins) * sizeof(long long)))) return rc;
  long long* hist = (long long*)buf.p;
  *tot = hist + (size_t)ncic * win;
  // windows of at most kCicLds bins (one for any ordinary bunch)
  for (int j0 = 0; j0 < nbins; j0 += win) {
    const int nb = std::min(win, nbins - j0);
    k_cic<<<ncic, kPT, (size_t)nb * sizeof(long long), st>>>(R, n, a, ds, nbins, j0, nb, wscale, hist);
    k_cic_sum<<<(nb + 63) / 64, 1024, 0, st>>>(hist, ncic, nb, nullptr, *tot + j0);
  }
  LAUNCHCHK("k_cic");
  return OCSC_OK;
}

// Centre of the peak-current window (analysis.py:61-64): the current profile
// s_to_cur(tau, 0.01 std(tau), ...) (analysis.py:266-340) on its own grid,
// a = min - 3 sigma, b = max + 3 sigma, ds = (b - a)/ceil((b - a)/(sigma/4)),
// cloud-in-cell weights summed on the device as fixed-point integers, then the
// Gaussian smoothing (2K+1 taps, K = floor(3 sigma/ds + 1/2)) and the first
// maximum on the host; returns z0 = i ds + a and stores the window in the
// device state.
int imax_window(ocsc_ctx* c, const Rows& R, int64_t n, hipStream_t st, double* z0_out) {
  const int nblk = nred_blocks(c, n);
  double* part = (double*)c->lane0.part.p;
  k_tau_range<<<nblk, kPT, 0, st>>>(R, n, part);
  k_tau_range_reduce<<<1, 1024, 0, st>>>(c->mom, part, nblk);
  LAUNCHCHK("k_tau_range");
  MomState& h = *c->mom_host;
  HIPCHK(hipMemcpyAsync(&h, c->mom, sizeof(h), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const double sig0 = h.std_all;
  const double sigma = 0.01 * sig0;
  if (!(sigma > 0) || !std::isfinite(sigma)) return fail(OCSC_EVALUE, "slice='Imax' needs a spread in tau");
  const double a = h.tau_min - 3 * sigma, b = h.tau_max + 3 * sigma;
  const double ds0 = 0.25 * sigma;
  const double nf = std::ceil((b - a) / ds0);
  if (!(nf >= 1) || nf > 1e8) return fail(OCSC_EVALUE, "slice='Imax': current profile grid out of range");
  const int N = (int)nf;
  const double ds = (b - a) / N;
  const int nbins = N + 1;
  long long* tot = nullptr;
  int rc;
  if ((rc = cic_profile(c->cic, R, n, a, ds, nbins, st, &tot))) return rc;
  std::vector<long long> C(nbins);
  HIPCHK(hipMemcpyAsync(C.data(), tot, (size_t)nbins * sizeof(long long), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const int K = (int)std::floor(3 * sigma / ds + 0.5);
  std::vector<double> G(2 * K + 1);
  double gs = 0;
  for (int m = 0; m <= 2 * K; ++m) {
    const double x = (double)(m - K) * ds / sigma;
    G[m] = std::exp(-0.5 * (x * x));
    gs += G[m];
  }
  int imax = 0;
  double best = -1;
  for (int j = 0; j < nbins; ++j) {
    double acc = 0;
    for (int m = 0; m <= 2 * K; ++m) {
      const int i = j + K - m;
      if (i >= 0 && i < nbins) acc += (double)C[i] * G[m];
    }
    acc = acc / gs;
    if (acc > best) {
      best = acc;
      imax = j;
    }
  }
  const double z0 = (double)imax * ds + a;
  const double lohi[2] = {z0 + sig0 * h.b0, z0 + sig0 * h.b1};
  HIPCHK(hipMemcpyAsync((char*)c->mom + offsetof(MomState, lo), lohi, sizeof(lohi), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  *z0_out = z0;
  return OCSC_OK;
}

}  // namespace

extern "C" {

int ocsc_moments(ocsc_ctx* c, const double* r, int64_t ld, const double* q, int64_t n, int flags,
                 const double disp[4], double b0, double b1, double* out, int nout, void* stream) {
  if (!c || !out || ((!r || !q) && n > 0)) return fail(OCSC_EINVAL, "null argument");
  if (n < 0 || ld < n) return fail(OCSC_EINVAL, "bad n / ld");
  if ((flags & OCSC_MOM_DISP_GIVEN) && !disp) return fail(OCSC_EINVAL, "OCSC_MOM_DISP_GIVEN needs disp");
  if ((flags & OCSC_MOM_IMAX) && !(flags & OCSC_MOM_BOUNDS))
    return fail(OCSC_EINVAL, "OCSC_MOM_IMAX selects the window centre and needs OCSC_MOM_BOUNDS");
  if (n == 0) return fail(OCSC_EVALUE, "empty bunch");
  HIPCHK(hipSetDevice(c->device));
  Lane& L = c->lane0;
  hipStream_t st = (hipStream_t)stream;
  if (!c->mom) {
    HIPCHK(hipMalloc(&c->mom, sizeof(MomState)));
    HIPCHK(hipHostMalloc(&c->mom_host, sizeof(MomState), hipHostMallocDefault));
  }
  const int nblk = nred_blocks(c, n);
  int rc;
  if ((rc = ensure(L.part, (size_t)12 * nblk * sizeof(double)))) return rc;
  // the pinned mirror may still be read by an earlier copy on another stream
  HIPCHK(hipStreamSynchronize(st));
  MomState& h = *c->mom_host;
  std::memset(&h, 0, sizeof(h));
  h.n = n;
  h.bounds = (flags & OCSC_MOM_BOUNDS) ? 1 : 0;
  h.b0 = b0;
  h.b1 = b1;
  h.disp_mode = (flags & OCSC_MOM_DISP_GIVEN) ? 1 : ((flags & OCSC_MOM_AUTO_DISP) ? 2 : 0);
  if (h.disp_mode == 1)
    for (int k = 0; k < 4; ++k) h.D[k] = disp[k];
  HIPCHK(hipMemcpyAsync(c->mom, &h, sizeof(h), hipMemcpyHostToDevice, st));
  double* part = (double*)L.part.p;
  const Rows R{r, ld};
  double z0 = 0;
  if (!h.bounds && h.disp_mode != 2) {
    k_mom<6><<<nblk, kPT, 0, st>>>(R, q, c->mom, part);  // passes 0, 2, 4 in one read
    k_mom_reduce<6><<<1, 1024, 0, st>>>(c->mom, part, nblk);
  } else {
    k_mom<0><<<nblk, kPT, 0, st>>>(R, q, c->mom, part);
    k_mom_reduce<0><<<1, 1024, 0, st>>>(c->mom, part, nblk);
    if (h.bounds) {
      k_mom<1><<<nblk, kPT, 0, st>>>(R, q, c->mom, part);
      k_mom_reduce<1><<<1, 1024, 0, st>>>(c->mom, part, nblk);
      if (flags & OCSC_MOM_IMAX) {
        if ((rc = imax_window(c, R, n, st, &z0))) return rc;
      } else {
        z0 = -1;  // mean(tau), set below from the device state
      }
    }
    if (h.disp_mode != 2) {
      k_mom<7><<<nblk, kPT, 0, st>>>(R, q, c->mom, part);  // passes 2 and 4 in one read
      k_mom_reduce<7><<<1, 1024, 0, st>>>(c->mom, part, nblk);
    } else {
      k_mom<2><<<nblk, kPT, 0, st>>>(R, q, c->mom, part);
      k_mom_reduce<2><<<1, 1024, 0, st>>>(c->mom, part, nblk);
      k_mom<3><<<nblk, kPT, 0, st>>>(R, q, c->mom, part);
      k_mom_reduce<3><<<1, 1024, 0, st>>>(c->mom, part, nblk);
      k_mom<4><<<nblk, kPT, 0, st>>>(R, q, c->mom, part);
      k_mom_reduce<4><<<1, 1024, 0, st>>>(c->mom, part, nblk);
    }
  }
  k_mom<5><<<nblk, kPT, 0, st>>>(R, q, c->mom, part);
  k_mom_reduce<5><<<1, 1024, 0, st>>>(c->mom, part, nblk);
  LAUNCHCHK("k_mom");
  HIPCHK(hipMemcpyAsync(&h, c->mom, sizeof(h), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double v[OCSC_MOM_NOUT];
  v[0] = h.cnt;
  v[1] = h.sumq;
  v[2] = h.m[0];
  for (int k = 0; k < 4; ++k) v[3 + k] = h.D[k];
  for (int k = 0; k < 4; ++k) v[7 + k] = h.c[k];
  v[11] = h.m[5];
  for (int k = 0; k < 12; ++k) v[12 + k] = h.mom[k];
  v[24] = h.tau_all;
  v[25] = h.std_all;
  v[26] = (flags & OCSC_MOM_IMAX) && h.bounds ? z0 : h.tau_all;
  for (int k = 0; k < nout && k < OCSC_MOM_NOUT; ++k) out[k] = v[k];
  return OCSC_OK;
}

}  // extern "C"

namespace {

// moments requested per tracking step (the tws_track list of track())
struct TwsReq {
  int bounds = 0, disp_mode = 0;  // as in MomState
  int imax = 0;                   // window centred on the peak current (slice = "Imax")
  double b0 = 0, b1 = 0;
  double disp[4] = {0, 0, 0, 0};
  int cap = 0;  // rows of the step table
};

constexpr int kImaxGraphBins = kCicLds;  // slice = "Imax" inside tracking graphs (64 KB of LDS)
double cic_wscale(int64_t n);

// the moments passes of ocsc_moments on the device only, appended to the
// step table (graph-capturable; own partial buffer, so the kick's fused
// momentum sums are left alone)
// dr: apply the next step's drift in the last pass (after the moments)
int enqueue_moments(ocsc_ctx* c, const Rows& R, const double* q, int64_t n, const TwsReq& t, hipStream_t st,
                    const DriftArgs* dr = nullptr) {
  const int nblk = nred_blocks(c, n);
  double* part = (double*)c->mom_part.p;
  MomState* cur = c->mom;
  k_mom_init<<<1, 64, 0, st>>>(cur, n, t.bounds, t.disp_mode, t.b0, t.b1, t.disp[0], t.disp[1], t.disp[2],
                               t.disp[3]);
  const bool one_read = !t.bounds && t.disp_mode != 2;  // passes 0, 2, 4 together
  if (one_read) {
    k_mom<6><<<nblk, kPT, 0, st>>>(R, q, cur, part);
    k_mom_reduce<6><<<1, 1024, 0, st>>>(cur, part, nblk);
  } else {
    k_mom<0><<<nblk, kPT, 0, st>>>(R, q, cur, part);
    k_mom_reduce<0><<<1, 1024, 0, st>>>(cur, part, nblk);
  }
  if (t.bounds) {
    k_mom<1><<<nblk, kPT, 0, st>>>(R, q, cur, part);
    k_mom_reduce<1><<<1, 1024, 0, st>>>(cur, part, nblk);
    if (t.imax) {
      // centre on the peak of the current profile (analysis.py:61-64)
      const int cap = kImaxGraphBins;
      long long* hist = (long long*)c->cic.p;
      long long* tot = hist + (size_t)kCicBlocks * cap;
      k_tau_range<<<nblk, kPT, 0, st>>>(R, n, part);
      k_tau_range_reduce<<<1, 1024, 0, st>>>(cur, part, nblk);
      k_imax_plan<<<1, 64, 0, st>>>(cur, c->imax_plan, cap, c->imax_err);
      k_cic_grid<ImaxPlan><<<kCicBlocks, kPT, (size_t)cap * sizeof(long long), st>>>(R, n, c->imax_plan,
                                                                                     cic_wscale(n), hist);
      k_cic_sum<<<(cap + 63) / 64, 1024, 0, st>>>(hist, kCicBlocks, cap, &c->imax_plan->nbins, tot);
      k_imax_peak<<<1, 1024, 0, st>>>(tot, c->imax_plan, cur);
    }
  }
  if (t.disp_mode == 2) {
    k_mom<2><<<nblk, kPT, 0, st>>>(R, q, cur, part);
    k_mom_reduce<2><<<1, 1024, 0, st>>>(cur, part, nblk);
    k_mom<3><<<nblk, kPT, 0, st>>>(R, q, cur, part);
    k_mom_reduce<3><<<1, 1024, 0, st>>>(cur, part, nblk);
    k_mom<4><<<nblk, kPT, 0, st>>>(R, q, cur, part);
    k_mom_reduce<4><<<1, 1024, 0, st>>>(cur, part, nblk);
  } else if (!one_read) {
    k_mom<7><<<nblk, kPT, 0, st>>>(R, q, cur, part);  // passes 2 and 4 in one read
    k_mom_reduce<7><<<1, 1024, 0, st>>>(cur, part, nblk);
  }
  if (dr)
    k_mom<5, true><<<nblk, kPT, 0, st>>>(R, q, cur, part, const_cast<double*>(R.r), dr->L, dr->r56);
  else
    k_mom<5><<<nblk, kPT, 0, st>>>(R, q, cur, part);
  k_mom_reduce<5><<<1, 1024, 0, st>>>(cur, part, nblk);
  k_mom_store<<<1, 64, 0, st>>>(cur, (MomState*)c->mom_tab.p, (int*)c->mom_ctr.p, t.cap);
  LAUNCHCHK("tracking moments");
  return OCSC_OK;
}

void moments_row(const MomState& h, double* v) {
  v[0] = h.cnt;
  v[1] = h.sumq;
  v[2] = h.m[0];
  for (int k = 0; k < 4; ++k) v[3 + k] = h.D[k];
  for (int k = 0; k < 4; ++k) v[7 + k] = h.c[k];
  v[11] = h.m[5];
  for (int k = 0; k < 12; ++k) v[12 + k] = h.mom[k];
  v[24] = h.tau_all;
  v[25] = h.std_all;
  v[26] = h.z0;
}

// an LSC kick after every space-charge kick of a tracking loop
struct LscReq {
  LscConst k;
  double b0, b1;
};
int lsc_buffers(ocsc_ctx* c, int64_t bins);
LscConst lsc_const(int step_profile, double smooth, double E_GeV, double dz, double K_max, double fill_factor,
                   int cap);
constexpr int kLscGraphBins = 4096;  // kicks inside tracking graphs (device-planned grid; 32 KB of LDS)
// the LSC kick's momentum sums for the next space-charge kick (the pass-A
// partials of the kick scalars sc, nred_blocks of them)
struct LscSump {
  const KickScalars* sc;
  double* part;
};
int enqueue_lsc(ocsc_ctx* c, double* r, int64_t ld, const double* q, int64_t n, const LscConst& k, double b0,
                double b1, hipStream_t st, const DriftArgs* dr = nullptr, const LscSump* sm = nullptr);
const char* lsc_err_text(int e);

int track_impl(ocsc_ctx* c, double* r, int64_t ld, const double* q, int64_t n, const int32_t nmesh[3],
               double E_GeV, double zstep, double Ld, int nsteps, int prec, int use_graph, void* stream,
               const TwsReq* tw, double* tws_out, const LscReq* lq = nullptr, const double* draws = nullptr) {
  int rc;
  if ((rc = check_common(c, r, q, n, ld))) return rc;
  if (!nmesh || nsteps < 0) return fail(OCSC_EINVAL, "bad nmesh / nsteps");
  if (prec != OCSC_FP64 && prec != OCSC_MIXED) return fail(OCSC_EINVAL, "bad precision");
  if (!(E_GeV > 0)) return fail(OCSC_EINVAL, "E_GeV must be > 0");
  if (n > (int64_t)0xFFFFFFFF) return fail(OCSC_EINVAL, "tracking supports up to 2^32-1 particles");
  HIPCHK(hipSetDevice(c->device));
  Lane& L = c->lane0;
  Geometry g;
  if ((rc = make_geometry(nmesh, g))) return rc;
  if ((rc = ensure_prec(c, L, g, prec))) return rc;
  hipStream_t st = (hipStream_t)stream;
  KickInputs in = kick_inputs(nmesh, n, E_GeV, zstep, draws ? 1 : 0, 1.0, 0.0);
  const KickInputs in_keys = in;  // the first sort keys take no draw
  if (draws && nsteps > 0 && zstep != 0) {
    if ((rc = ensure(c->draws, (size_t)2 * nsteps * sizeof(double) + 256))) return rc;
    HIPCHK(hipMemcpyAsync(c->draws.p, draws, (size_t)2 * nsteps * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));  // the caller's (pageable) array may go away after the call
    int* ctr = (int*)((char*)c->draws.p + (size_t)2 * nsteps * sizeof(double));
    k_counter_reset<<<1, 64, 0, st>>>(ctr);
    in.draws = (const double*)c->draws.p;
    in.draw_ctr = ctr;
    in.ndraws = nsteps;
  }
  const double gamma = E_GeV / kMeGeV;
  const double igamma2 = 1. / (gamma * gamma);
  const double beta = std::sqrt(1. - igamma2);
  const double r56 = 0.0 - Ld / (beta * beta) * igamma2;
  const int nb = (int)((n + 255) / 256);
  if (nsteps == 0 && !tw) return OCSC_OK;
  if ((rc = ensure_part(c, L, n))) return rc;
  if (lq) {
    if (c->kick_timing) return fail(OCSC_EINVAL, "kick timing and LSC kicks cannot be combined");
    if ((rc = lsc_buffers(c, lq->k.cap))) return rc;
    if ((rc = ensure(c->lsc_cic, ((size_t)kCicBlocks * lq->k.cap + lq->k.cap) * sizeof(long long)))) return rc;
    if ((rc = ensure(L.part, (size_t)12 * nred_blocks(c, n) * sizeof(double)))) return rc;
    k_counter_reset<<<1, 64, 0, st>>>(c->lsc_err);
    c->lsc_nbins = 0;  // the per-bin arrays now belong to the tracking kicks
  }
  if (tw) {
    if (c->kick_timing) return fail(OCSC_EINVAL, "kick timing and per-step moments cannot be combined");
    if (!c->mom) {
      HIPCHK(hipMalloc(&c->mom, sizeof(MomState)));
      HIPCHK(hipHostMalloc(&c->mom_host, sizeof(MomState), hipHostMallocDefault));
    }
    if ((rc = ensure(c->mom_part, (size_t)12 * c->red_blocks * sizeof(double)))) return rc;
    if ((rc = ensure(c->mom_tab, (size_t)(nsteps + 1) * sizeof(MomState)))) return rc;
    if ((rc = ensure(c->mom_ctr, sizeof(int)))) return rc;
    k_counter_reset<<<1, 64, 0, st>>>((int*)c->mom_ctr.p);
    if (tw->imax) {
      if (!c->imax_plan) {
        HIPCHK(hipMalloc(&c->imax_plan, sizeof(ImaxPlan)));
        HIPCHK(hipMalloc(&c->imax_err, sizeof(int)));
      }
      if ((rc = ensure(c->cic, ((size_t)kCicBlocks * kImaxGraphBins + kImaxGraphBins) * sizeof(long long))))
        return rc;
      k_counter_reset<<<1, 64, 0, st>>>(c->imax_err);
    }
    // the bunch before the first step (tw0 of track())
    if ((rc = enqueue_moments(c, Rows{r, ld}, q, n, *tw, st))) return rc;
  }

  // cell-ordered resident copy (sort_every > 0)
  const bool sorted = c->sort_every > 0 && zstep != 0;
  const int end_bit = sort_end_bit(g.d.nx, g.d.ny, g.d.nz, c->sort_cells);  // keys 0..nkeys (nkeys: outside)
  if (sorted) {
    for (int k = 0; k < 2; ++k) {
      if ((rc = ensure(c->rs[k], (size_t)6 * n * 8))) return rc;
      if ((rc = ensure(c->qs[k], (size_t)n * 8))) return rc;
      if ((rc = ensure(c->keys[k], (size_t)n * 4))) return rc;
      if ((rc = ensure(c->perm[k], (size_t)n * 4))) return rc;
    }
    if ((rc = ensure(c->vals, (size_t)2 * n * 4))) return rc;
    if ((rc = ensure(c->sort_tmp, rs_hist_bytes(n)))) return rc;
    if (c->permute_rec && (rc = ensure(c->rec, (size_t)n * sizeof(PRec)))) return rc;
  }
  const bool via_rec = sorted && c->permute_rec;
  // sort src (row stride lds) by NGP cell under the current kick scalars into
  // the copy `to`; perm[to] = order relative to the caller's particles.
  // drift: apply the first step's drift to the sorted copy (initial sort).
  // keys at the drifted position half-way through the steps this sort serves
  // (option sort_ahead)
  auto ahead_for = [&](int from_step) {
    const int left = std::min(c->sort_every, nsteps - from_step);
    const double h = c->sort_ahead ? 0.5 * std::max(left - 1, 0) + (from_step == 0 ? 1.0 : 0.0) : 0.0;
    return DriftArgs{Ld * h, r56 * h};
  };
  auto resort = [&](const double* src, int64_t lds, const double* qsrc, int to, const uint32_t* perm_prev,
                    bool drift, DriftArgs ahead) -> int {
    uint32_t* kb[2] = {(uint32_t*)c->keys[0].p, (uint32_t*)c->keys[1].p};
    uint32_t* vb[2] = {(uint32_t*)c->vals.p, (uint32_t*)c->vals.p + n};
    // records for the first sort (from the caller's order); a re-sort of the
    // copy permutes nearby rows directly
    const bool resort_near = perm_prev && c->resort_near;
    PRec* rec = via_rec && !resort_near ? (PRec*)c->rec.p : nullptr;
    k_sort_keys<<<(int)((n + kPT - 1) / kPT), kPT, 0, st>>>(Rows{src, lds}, qsrc, L.sc, kb[0], rec, c->sort_cells,
                                                            ahead);
    LAUNCHCHK("k_sort_keys");
    const uint32_t* order = vb[rs_sort(kb, vb, n, end_bit, (uint32_t*)c->sort_tmp.p, st)];  // order of src
    LAUNCHCHK("cell sort");
    double* dst = (double*)c->rs[to].p;
    double* qdst = (double*)c->qs[to].p;
    if (resort_near) {
      k_resort_gather<<<nb, 256, 0, st>>>(src, qsrc, order, perm_prev, dst, qdst, (uint32_t*)c->perm[to].p, n);
      LAUNCHCHK("re-sort gather");
      return OCSC_OK;
    }
    if (perm_prev)
      k_compose<<<nb, 256, 0, st>>>(perm_prev, order, (uint32_t*)c->perm[to].p, n);
    else
      HIPCHK(hipMemcpyAsync(c->perm[to].p, order, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    if (via_rec && drift)
      k_gather_rec<true><<<nb, 256, 0, st>>>(rec, order, dst, n, qdst, n, DriftArgs{Ld, r56});
    else if (via_rec)
      k_gather_rec<false><<<nb, 256, 0, st>>>(rec, order, dst, n, qdst, n, DriftArgs{0, 0});
    else
      k_permute_gather<<<dim3(nb, 7), 256, 0, st>>>(src, lds, qsrc, order, dst, n, qdst, n);
    LAUNCHCHK("cell-order gather");
    if (drift && !via_rec) k_drift<<<nb, 256, 0, st>>>(dst, n, n, Ld, r56);
    return OCSC_OK;
  };

  int cur = 0;
  double* rr = r;
  int64_t lr = ld;
  const double* qq = q;
  if (sorted) {
    // kick scalars of the current bunch for the first sort keys
    const int nblk = nred_blocks(c, n);
    k_setup<<<1, 64, 0, st>>>(L.sc, in_keys);
    k_passA<<<nblk, kPT, 0, st>>>(Rows{r, ld}, L.sc, (double*)L.part.p);
    k_reduceA<<<1, 1024, 0, st>>>(L.sc, (double*)L.part.p, nblk);
    k_passB<false><<<nblk, kPT, 0, st>>>(Rows{r, ld}, q, L.sc, (double*)L.part.p);
    k_reduceB<<<1, 1024, 0, st>>>(L.sc, (double*)L.part.p, nblk, prec);
    LAUNCHCHK("initial frame");
    // the first step's drift rides on the gather (sorted implies zstep != 0)
    if ((rc = resort(r, ld, q, 0, nullptr, nsteps > 0, ahead_for(0)))) return rc;
    rr = (double*)c->rs[0].p;
    lr = n;
    qq = (const double*)c->qs[0].p;
  }
  // equal charges (Ocelot's generated bunches): the steps' extent and deposit
  // passes take the charge from the kick scalars, the same value bit for bit
  if (c->q_uniform && nsteps > 0 && zstep != 0) {
    k_q_const_init<<<1, 64, 0, st>>>(q, n, L.sc);
    k_q_const_check<<<(int)std::min<int64_t>(nb, 4096), 256, 0, st>>>(q, n, L.sc);
    LAUNCHCHK("charge check");
    in.q_keep = 1;
  }
  // Step loop.  The drift of step s+1 and the momentum sums of kick s+1 are
  // fused into kick s, so the sequence is drift_0, then kicks 0..n-1 with
  // kick s < n-1 drifting (first kick: own pass A; later kicks: partials).
  const int npart = (int)((n + kKickPT - 1) / kKickPT);  // kick workgroups: its momentum partials
  const bool ktime = c->kick_timing != 0 && zstep != 0;
  if (ktime) {
    while (c->kick_ev.size() < (size_t)2 * nsteps) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      c->kick_ev.push_back(e);
    }
  }
  c->kick_ms.clear();
  auto opts_for = [&](int step) {
    StepOpts o;
    o.passA = step == 0;
    o.prev_nblk = npart;
    o.drift = o.sump = step < nsteps - 1;
    if (tw) o.drift = false;  // moments are taken between the kick and the next drift
    if (lq) {
      // the LSC kick changes delta after the space-charge kick: no fused
      // drift here; the LSC kick leaves the next kick's momentum sums, in
      // pass A's partial layout
      o.drift = o.sump = false;
      o.prev_nblk = nred_blocks(c, n);
    }
    o.dr = DriftArgs{Ld, r56};
    o.grid_zeroed = c->grid_clear == 2 && step > 0;
    o.zero_next = c->grid_clear == 2 && step < nsteps - 1;
    if (ktime) {
      o.kev0 = c->kick_ev[2 * step];
      o.kev1 = c->kick_ev[2 * step + 1];
    }
    return o;
  };
  // graph event-record nodes of the kick events, per captured graph
  hipGraphNode_t knode[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  if (zstep == 0) {
    for (int step = 0; step < nsteps; ++step) {
      k_drift<<<nb, 256, 0, st>>>(rr, lr, n, Ld, r56);
      if (lq && (rc = enqueue_lsc(c, rr, lr, qq, n, lq->k, lq->b0, lq->b1, st))) return rc;
      if (tw && (rc = enqueue_moments(c, Rows{rr, lr}, qq, n, *tw, st))) return rc;
    }
  } else if (nsteps > 0 && !sorted) {
    k_drift<<<nb, 256, 0, st>>>(rr, lr, n, Ld, r56);  // (sorted: applied by the initial gather)
  }
  // one tracking step after the drift: the kick and, with per-step moments,
  // the moments and the next drift
  auto step_body = [&](int step, const StepOpts& o, hipStream_t s2) -> int {
    int rc3;
    if ((rc3 = enqueue_kick_prec(c, L, g, rr, lr, qq, in, prec, s2, o))) return rc3;
    // the next drift: fused into the LSC kick when there are no moments
    const DriftArgs next{Ld, r56};
    const bool fuse = lq && !tw && step < nsteps - 1;
    const LscSump sm{L.sc, (double*)L.part.p};  // read by the next step's kick
    if (lq && (rc3 = enqueue_lsc(c, rr, lr, qq, n, lq->k, lq->b0, lq->b1, s2, fuse ? &next : nullptr,
                                 step < nsteps - 1 ? &sm : nullptr)))
      return rc3;
    // with moments, the next drift rides on their last pass (after the kick's
    // and the LSC kick's writes); otherwise the kick or the LSC kick carries it
    if (tw && (rc3 = enqueue_moments(c, Rows{rr, lr}, qq, n, *tw, s2, step < nsteps - 1 ? &next : nullptr)))
      return rc3;
    return OCSC_OK;
  };
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipGraph_t graphs[2] = {nullptr, nullptr};
  auto cleanup = [&]() {
    for (int k = 0; k < 2; ++k) {
      if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
      if (graphs[k]) (void)hipGraphDestroy(graphs[k]);
      exec[k] = nullptr;
      graphs[k] = nullptr;
    }
  };
  // The step graphs hold pointers, sizes, step and energy constants and the
  // options in their launches; a plain call with the same ones (and the same
  // workspaces) replays the previous call's graphs.
  const bool cacheable = use_graph && !tw && !lq && !draws && !c->profiling && zstep != 0;
  std::vector<int64_t> tkey;
  if (cacheable) {
    auto bits = [](double v) {
      int64_t b;
      std::memcpy(&b, &v, sizeof(b));
      return b;
    };
    tkey = {(int64_t)(intptr_t)r, ld, (int64_t)(intptr_t)q, n, g.d.nx, g.d.ny, g.d.nz, prec, sorted, via_rec,
            bits(E_GeV), bits(zstep), bits(Ld), ktime, c->overlap, c->deposit_xcd, c->deposit_agg, c->deposit_tile,
            c->fft_plan, c->sort_cells, c->red_blocks, c->grid_clear, c->igf_blocks, c->fold_frame, c->q_uniform,
            (int64_t)g_alloc_gen.load()};
    if (c->tg.key == tkey) {
      for (int k = 0; k < 2; ++k) {
        exec[k] = c->tg.exec[k];
        graphs[k] = c->tg.graphs[k];
        knode[k][0] = c->tg.knode[k][0];
        knode[k][1] = c->tg.knode[k][1];
      }
      c->tg = ocsc_ctx::TrackGraphs{};  // owned by this call now
    }
  }
  drop_track_graphs(c);
  for (int step = 0; step < nsteps && zstep != 0; ++step) {
    if (sorted && step > 0 && step % c->sort_every == 0) {
      if ((rc = resort(rr, n, qq, 1 - cur, (const uint32_t*)c->perm[cur].p, false, ahead_for(step)))) {
        cleanup();
        return rc;
      }
      cur = 1 - cur;
      rr = (double*)c->rs[cur].p;
      qq = (const double*)c->qs[cur].p;
    }
    StepOpts o = opts_for(step);
    // the last kick writes its records straight to the caller's positions
    // (no moments or LSC kick read the rows after it)
    const bool rec_last = via_rec && !tw && !lq && step == nsteps - 1;
    if (rec_last) {
      o.rec_perm = (const uint32_t*)c->perm[cur].p;
      o.rec_out = (PRec*)c->rec.p;
    }
    const bool middle = step > 0 && step < nsteps - 1;
    if (!use_graph || c->profiling || !middle) {
      if ((rc = step_body(step, o, st))) {
        cleanup();
        return rc;
      }
      continue;
    }
    const int gi = sorted ? cur : 0;
    if (!exec[gi]) {
      StepOpts oc = o;
      oc.kev_external = true;
      HIPCHK(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
      int rc2 = step_body(step, oc, c->cap_stream);
      hipError_t e = hipStreamEndCapture(c->cap_stream, &graphs[gi]);
      if (rc2 || e != hipSuccess) {
        cleanup();
        return rc2 ? rc2 : fail(OCSC_EHIP, std::string("capture: ") + hipGetErrorString(e));
      }
      if (ktime) {
        // the kick is the graph's only sink (every other node feeds it);
        // bracket it with event-record nodes: deps -> start -> kick -> end
        size_t nn = 0;
        HIPCHK(hipGraphGetNodes(graphs[gi], nullptr, &nn));
        std::vector<hipGraphNode_t> nodes(nn);
        HIPCHK(hipGraphGetNodes(graphs[gi], nodes.data(), &nn));
        hipGraphNode_t kick = nullptr;
        int nsink = 0;
        for (hipGraphNode_t nd : nodes) {
          size_t nout = 0;
          HIPCHK(hipGraphNodeGetDependentNodes(nd, nullptr, &nout));
          if (nout == 0) {
            kick = nd;
            ++nsink;
          }
        }
        hipGraphNodeType t;
        if (nsink != 1 || hipGraphNodeGetType(kick, &t) != hipSuccess || t != hipGraphNodeTypeKernel) {
          cleanup();
          return fail(OCSC_EHIP, "kick timing: the captured step graph has no single kernel sink");
        }
        size_t nd = 0;
        HIPCHK(hipGraphNodeGetDependencies(kick, nullptr, &nd));
        std::vector<hipGraphNode_t> deps(nd);
        HIPCHK(hipGraphNodeGetDependencies(kick, deps.data(), &nd));
        std::vector<hipGraphNode_t> to(nd, kick);
        if (nd) HIPCHK(hipGraphRemoveDependencies(graphs[gi], deps.data(), to.data(), nd));
        HIPCHK(hipGraphAddEventRecordNode(&knode[gi][0], graphs[gi], deps.data(), nd, o.kev0));
        HIPCHK(hipGraphAddDependencies(graphs[gi], &knode[gi][0], &kick, 1));
        HIPCHK(hipGraphAddEventRecordNode(&knode[gi][1], graphs[gi], &kick, 1, o.kev1));
      }
      e = hipGraphInstantiate(&exec[gi], graphs[gi], nullptr, nullptr, 0);
      if (e != hipSuccess) {
        cleanup();
        return fail(OCSC_EHIP, std::string("instantiate: ") + hipGetErrorString(e));
      }
    }
    if (ktime) {
      HIPCHK(hipGraphExecEventRecordNodeSetEvent(exec[gi], knode[gi][0], o.kev0));
      HIPCHK(hipGraphExecEventRecordNodeSetEvent(exec[gi], knode[gi][1], o.kev1));
    }
    hipError_t e = hipGraphLaunch(exec[gi], st);
    if (e != hipSuccess) {
      cleanup();
      return fail(OCSC_EHIP, std::string("graph launch: ") + hipGetErrorString(e));
    }
  }
  if (via_rec) {
    // back to the caller's order: whole records to their places (done by the
    // last kick unless it was skipped for an invalid bunch or other work read
    // the rows after it), then rows
    const bool rec_last = !tw && !lq && nsteps > 0;
    if (rec_last)
      k_scatter_rec_if_err<<<nb, 256, 0, st>>>(rr, n, (const uint32_t*)c->perm[cur].p, (PRec*)c->rec.p, n, L.sc);
    else
      k_scatter_rec<<<nb, 256, 0, st>>>(rr, n, (const uint32_t*)c->perm[cur].p, (PRec*)c->rec.p, n);
    k_rec_to_rows<<<nb, 256, 0, st>>>((const PRec*)c->rec.p, r, ld, n);
    LAUNCHCHK("caller-order scatter");
  } else if (sorted) {
    k_permute_scatter<<<dim3(nb, 6), 256, 0, st>>>(rr, n, (const uint32_t*)c->perm[cur].p, r, ld, n);
    LAUNCHCHK("k_permute_scatter");
  }
  int32_t err = 0;
  rc = read_errors(c, L, st, &err);
  if (cacheable && !rc && (exec[0] || exec[1])) {
    // keep the step graphs for the next call
    c->tg.key = tkey;
    for (int k = 0; k < 2; ++k) {
      c->tg.exec[k] = exec[k];
      c->tg.graphs[k] = graphs[k];
      c->tg.knode[k][0] = knode[k][0];
      c->tg.knode[k][1] = knode[k][1];
      exec[k] = nullptr;
      graphs[k] = nullptr;
    }
  }
  cleanup();
  if (rc) return rc;
  if (ktime) {
    HIPCHK(hipEventSynchronize(c->kick_ev[2 * nsteps - 1]));
    c->kick_ms.resize(nsteps);
    for (int step = 0; step < nsteps; ++step) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->kick_ev[2 * step], c->kick_ev[2 * step + 1]));
      c->kick_ms[step] = ms;
    }
  }
  if (tw && tws_out) {
    std::vector<MomState> tab(nsteps + 1);
    HIPCHK(hipMemcpyAsync(tab.data(), c->mom_tab.p, (size_t)(nsteps + 1) * sizeof(MomState), hipMemcpyDeviceToHost,
                          st));
    HIPCHK(hipStreamSynchronize(st));
    for (int k = 0; k <= nsteps; ++k) moments_row(tab[k], tws_out + (size_t)k * OCSC_MOM_NOUT);
  }
  if (err) return value_error(err);
  if (tw && tw->imax) {
    int ierr = 0;
    HIPCHK(hipMemcpyAsync(&ierr, c->imax_err, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (ierr == 1) return fail(OCSC_EVALUE, "slice='Imax' needs a spread in tau (a step's window stayed at mean(tau))");
    if (ierr) return fail(OCSC_EVALUE, "slice='Imax' inside tracking: current profile over 8192 bins (a step's window "
                                       "stayed at mean(tau)); take get_envelope per step instead");
  }
  if (lq) {
    int lerr = 0;
    HIPCHK(hipMemcpyAsync(&lerr, c->lsc_err, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (lerr) return fail(OCSC_EVALUE, std::string("LSC kick skipped: ") + lsc_err_text(lerr));
  }
  return OCSC_OK;
}

}  // namespace

namespace {

// Launch tables of a fused batched round: for every pass, the jobs it covers
// and the prefix of their workgroup counts; the transform passes once per
// FFT length along their axis.
struct FusedPlan {
  struct Launch {
    BSeg seg{};
    int64_t grid = 0;
    int M = 0;
  };
  int nb = 0;
  BJob* jobs = nullptr;
  KickInputs* ins = nullptr;
  int* err = nullptr;
  Launch pab, dep, kick, nodes, diff, d64;
  std::vector<Launch> zcos, scos_y, scos_x, zf, yf, xc, yi, zi;
};

// columns per workgroup of the transform kernels (host side of Tile)
int tile_cols(bool dbl, int M) {
  switch (M) {
#define OCSC_TC(V) \
  case V:          \
    return dbl ? Tile<double, V>::T : Tile<float, V>::T;
    OCSC_TC(8) OCSC_TC(16) OCSC_TC(32) OCSC_TC(64) OCSC_TC(128) OCSC_TC(256) OCSC_TC(512) OCSC_TC(1024)
#undef OCSC_TC
    default:
      return 0;
  }
}

int dep_items(int tile) { return tile == 1 ? DepTile<1>::items : (tile == 2 ? DepTile<2>::items : DepTile<0>::items); }

// Build the jobs (one per kicked bunch: rows rr[i], workspace c->flanes[i])
// and the launch tables, and upload them; the tables stay valid while the
// rows, sizes, meshes, options and workspaces are unchanged.
int fused_plan(ocsc_ctx* c, const std::vector<int>& kicked, const std::vector<double*>& rr,
               const std::vector<int64_t>& ldr, const std::vector<const double*>& qq, const std::vector<int64_t>& nn,
               const std::vector<Geometry>& geo, const std::vector<KickInputs>& ins, int prec, hipStream_t st,
               Buf& tab, bool copies, FusedPlan* P) {
  const bool dbl = prec == OCSC_FP64;
  const int nword = dbl ? 2 : 1;
  const int ncopy = c->deposit_xcd ? kNumXcd : 1;
  const int nbk = (int)kicked.size();
  std::vector<BJob> jobs(nbk);
  std::vector<KickInputs> jin(nbk);
  for (int j = 0; j < nbk; ++j) {
    const int i = kicked[j];
    Lane& L = *c->flanes[j];
    const Geometry& g = geo[i];
    BJob& b = jobs[j];
    b.r = rr[i];
    b.ld = ldr[i];
    b.q = qq[i];
    b.n = nn[i];
    b.sc = L.sc;
    b.part = (double*)L.part.p;
    b.nblk = nred_blocks(c, nn[i]);
    b.d = g.d;
    b.ncell = g.ncell;
    b.nzero = (int64_t)ncopy * nword * g.ncell;
    b.grid = (unsigned long long*)L.grid.p;
    b.A = L.A.p;
    b.B = L.B.p;
    b.KF = (double*)L.KF.p;
    b.KA = (double*)L.KA.p;
    b.KB = (double*)L.KB.p;
    b.G = L.G.p;
    b.phi = L.phi.p;
    b.E64 = (double*)L.E64.p;
    b.crs = copies ? (double*)c->bcopy[j]->rs.p : nullptr;
    b.cqs = copies ? (double*)c->bcopy[j]->qs.p : nullptr;
    b.cperm = copies ? (uint32_t*)c->bcopy[j]->perm.p : nullptr;
    b.off = j ? jobs[j - 1].off + jobs[j - 1].n : 0;
    jin[j] = ins[i];
  }
  // segments: host ids and prefixes, packed after the jobs and inputs
  std::vector<int> ids;
  std::vector<int64_t> pre, cnt;
  struct Pend {
    FusedPlan::Launch* l;
    size_t ids_at, pre_at;
  };
  std::vector<Pend> pend;
  // align: every job starts at a multiple of 8 workgroups (the hardware hands
  // workgroup b to XCD b mod 8, and the deposit and the kick give each XCD a
  // contiguous run of a job's particles by that rule); the padding
  // workgroups exit at once
  auto add = [&](FusedPlan::Launch* l, const std::vector<int>& js, const std::vector<int64_t>& blocks, int M,
                 bool align = false) {
    l->M = M;
    l->seg.nj = (int)js.size();
    pend.push_back({l, ids.size(), pre.size()});
    int64_t acc = 0;
    pre.push_back(0);
    for (size_t k = 0; k < js.size(); ++k) {
      ids.push_back(js[k]);
      cnt.push_back(blocks[k]);
      acc += align ? (blocks[k] + kNumXcd - 1) / kNumXcd * kNumXcd : blocks[k];
      pre.push_back(acc);
    }
    l->grid = acc;
  };
  auto all = [&](auto blocks_of, FusedPlan::Launch* l, bool align = false) {
    std::vector<int> js(nbk);
    std::vector<int64_t> bl(nbk);
    for (int j = 0; j < nbk; ++j) {
      js[j] = j;
      bl[j] = blocks_of(jobs[j]);
    }
    add(l, js, bl, 0, align);
  };
  // per FFT length along one axis
  auto by_m = [&](auto m_of, auto blocks_of, std::vector<FusedPlan::Launch>* out) {
    std::vector<int> ms;
    for (int j = 0; j < nbk; ++j)
      if (std::find(ms.begin(), ms.end(), m_of(jobs[j].d)) == ms.end()) ms.push_back(m_of(jobs[j].d));
    out->assign(ms.size(), FusedPlan::Launch{});
    for (size_t k = 0; k < ms.size(); ++k) {
      std::vector<int> js;
      std::vector<int64_t> bl;
      for (int j = 0; j < nbk; ++j)
        if (m_of(jobs[j].d) == ms[k]) {
          js.push_back(j);
          bl.push_back(blocks_of(jobs[j], ms[k]));
        }
      add(&(*out)[k], js, bl, ms[k]);
    }
  };
  const int items = dep_items(c->deposit_tile);
  all([&](const BJob& b) { return (int64_t)b.nblk; }, &P->pab);
  all([&](const BJob& b) { return (b.n + kDepPT * items - 1) / (kDepPT * items); }, &P->dep, true);
  all([&](const BJob& b) { return (b.n + kKickPT - 1) / kKickPT; }, &P->kick, true);
  all([&](const BJob& b) {
    const int64_t nodes = (int64_t)(b.d.nx + 1) * (b.d.ny + 1) * (b.d.nz + 1);
    return std::min<int64_t>((nodes + 255) / 256, 65536);
  }, &P->nodes);
  all([&](const BJob& b) { return std::min<int64_t>((igf_diff_items(b.d) + 255) / 256, 65536); }, &P->diff);
  all([&](const BJob& b) { return (b.ncell + 255) / 256; }, &P->d64);
  auto mz = [](const MeshDims& d) { return d.mz; };
  auto my = [](const MeshDims& d) { return d.my; };
  auto mx = [](const MeshDims& d) { return d.mx; };
  auto rows_pairs = [&](bool d64) {
    return [&, d64](const BJob& b, int M) {
      const int T = tile_cols(d64, M);
      return ((int64_t)b.d.nx * b.d.ny + 2 * T - 1) / (2 * T);
    };
  };
  by_m(mz, rows_pairs(true), &P->zcos);
  by_m(my, [&](const BJob& b, int M) {
    const int T = tile_cols(true, M);
    return ((int64_t)b.d.nx * b.d.hz + 2 * T - 1) / (2 * T);
  }, &P->scos_y);
  by_m(mx, [&](const BJob& b, int M) {
    const int T = tile_cols(true, M);
    return ((int64_t)b.d.hy * b.d.hz + 2 * T - 1) / (2 * T);
  }, &P->scos_x);
  by_m(mz, rows_pairs(dbl), &P->zf);
  by_m(my, [&](const BJob& b, int M) {
    const int T = tile_cols(dbl, M);
    return ((int64_t)b.d.nx * b.d.hz + T - 1) / T;
  }, &P->yf);
  by_m(mx, [&](const BJob& b, int M) {
    const int T = tile_cols(dbl, M);
    return ((int64_t)b.d.my * b.d.hz + T - 1) / T;
  }, &P->xc);
  by_m(my, [&](const BJob& b, int M) {
    const int T = tile_cols(dbl, M);
    return ((int64_t)b.d.nx * b.d.hz + T - 1) / T;
  }, &P->yi);
  by_m(mz, rows_pairs(dbl), &P->zi);
  // one device table: jobs, inputs, error word, ids, prefixes
  auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
  const size_t o_in = al(nbk * sizeof(BJob));
  const size_t o_err = o_in + al(nbk * sizeof(KickInputs));
  const size_t o_ids = o_err + 256;
  const size_t o_pre = o_ids + al(ids.size() * sizeof(int));
  const size_t o_cnt = o_pre + al(pre.size() * sizeof(int64_t));
  const size_t total = o_cnt + cnt.size() * sizeof(int64_t);
  int rc;
  if ((rc = ensure(tab, total))) return rc;
  std::vector<unsigned char> host(total, 0);
  std::memcpy(host.data(), jobs.data(), nbk * sizeof(BJob));
  std::memcpy(host.data() + o_in, jin.data(), nbk * sizeof(KickInputs));
  std::memcpy(host.data() + o_ids, ids.data(), ids.size() * sizeof(int));
  std::memcpy(host.data() + o_pre, pre.data(), pre.size() * sizeof(int64_t));
  std::memcpy(host.data() + o_cnt, cnt.data(), cnt.size() * sizeof(int64_t));
  HIPCHK(hipMemcpyAsync(tab.p, host.data(), total, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  unsigned char* base = (unsigned char*)tab.p;
  P->nb = nbk;
  P->jobs = (BJob*)base;
  P->ins = (KickInputs*)(base + o_in);
  P->err = (int*)(base + o_err);
  for (const Pend& q : pend) {
    q.l->seg.ids = (const int*)(base + o_ids) + q.ids_at;
    q.l->seg.pre = (const int64_t*)(base + o_pre) + q.pre_at;
    q.l->seg.cnt = (const int64_t*)(base + o_cnt) + q.ids_at;
  }
  return OCSC_OK;
}

template <int NWORD>
int launch_deposit_b(const ocsc_ctx* c, const FusedPlan& P, hipStream_t st) {
  const int v = (c->deposit_xcd ? 4 : 0) + (c->deposit_agg ? 2 : 0);
  const int t = c->deposit_tile;
#define OCSC_DEPB(V, T)                                                                                   \
  if (v == V && t == T) {                                                                                 \
    kb_deposit<NWORD, (V & 4) != 0, (V & 2) != 0, T><<<(unsigned)P.dep.grid, kDepPT, 0, st>>>(P.jobs, P.dep.seg); \
    return OCSC_OK;                                                                                       \
  }
#define OCSC_DEPB4(V) OCSC_DEPB(V, 0) OCSC_DEPB(V, 1) OCSC_DEPB(V, 2) OCSC_DEPB(V, 3) OCSC_DEPB(V, 4)
  OCSC_DEPB4(0) OCSC_DEPB4(2) OCSC_DEPB4(4) OCSC_DEPB4(6)
#undef OCSC_DEPB4
#undef OCSC_DEPB
  return fail(OCSC_EINVAL, "deposit options out of range");
}

// One kick of every job: the passes of enqueue_kick, each one launch (per
// FFT length for the transforms) over all jobs; kernel spectrum in line.
// Front: set-up, frame and extent; back: clear, Green's function, deposit,
// transforms, kick (a round that builds the cell-ordered copies runs the
// front on the callers' rows and the back on the copies).
int fused_front(const FusedPlan& P, int prec, hipStream_t st) {
  const BJob* J = P.jobs;
  kb_setup<<<(P.nb + 63) / 64, 64, 0, st>>>(P.ins, J, P.nb);
  kb_passA<<<(unsigned)P.pab.grid, kPT, 0, st>>>(J, P.pab.seg);
  kb_reduceA<<<P.nb, 1024, 0, st>>>(J);
  if (prec == OCSC_MIXED)
    kb_passB<true><<<(unsigned)P.pab.grid, kPT, 0, st>>>(J, P.pab.seg);
  else
    kb_passB<false><<<(unsigned)P.pab.grid, kPT, 0, st>>>(J, P.pab.seg);
  kb_reduceB<<<P.nb, 1024, 0, st>>>(J, prec);
  LAUNCHCHK("batched frame/extent passes");
  return OCSC_OK;
}

template <typename F, int NWORD>
int fused_back(ocsc_ctx* c, const FusedPlan& P, hipStream_t st) {
  const BJob* J = P.jobs;
  kb_igf_nodes<<<(unsigned)P.nodes.grid, 256, 0, st>>>(J, P.nodes.seg);
  kb_igf_diff<<<(unsigned)P.diff.grid, 256, 0, st>>>(J, P.diff.seg);
  for (const auto& l : P.zcos) {
    OCSC_SWITCH_M(l.M, {
      kb_zcos<MM><<<(unsigned)l.grid, kNT, Tile<double, MM>::lds_bytes, st>>>(J, l.seg, c->twd + MM);
    });
  }
  for (const auto& l : P.scos_y) {
    OCSC_SWITCH_M(l.M, {
      kb_scos_y<MM><<<(unsigned)l.grid, kNT, Tile<double, MM>::lds_bytes, st>>>(J, l.seg, c->twd + MM);
    });
  }
  for (const auto& l : P.scos_x) {
    OCSC_SWITCH_M(l.M, {
      kb_scos_x<MM, F><<<(unsigned)l.grid, kNT, Tile<double, MM>::lds_bytes, st>>>(J, l.seg, c->twd + MM);
    });
  }
  LAUNCHCHK("batched kernel spectrum");
  int rc;
  if ((rc = launch_deposit_b<NWORD>(c, P, st))) return rc;
  LAUNCHCHK("batched deposit");
  const int ncopy = c->deposit_xcd ? kNumXcd : 1;
  // forward passes: plan 4 (register-resident transforms) when chosen, else
  // the default plan
  auto fwd_passes = [&](auto plan) -> int {
    constexpr int V = decltype(plan)::value;
    for (const auto& l : P.zf) {
      OCSC_SWITCH_M(l.M, {
        constexpr int VZ = plan_zf_yi<F, MM, V>();
        kb_zfwd<F, MM, NWORD == 2 ? kChargeFix2 : kChargeFix1, VZ>
            <<<(unsigned)l.grid, pass_threads<F, MM, VZ>(), pass_lds<F, MM, VZ>(), st>>>(J, l.seg, tw_for<F>(c, MM),
                                                                                         ncopy);
      });
    }
    for (const auto& l : P.yf) {
      OCSC_SWITCH_M(l.M, {
        kb_yfwd<F, MM, V><<<(unsigned)l.grid, pass_threads<F, MM, V>(), pass_lds<F, MM, V>(), st>>>(
            J, l.seg, tw_for<F>(c, MM));
      });
    }
    return OCSC_OK;
  };
  if ((rc = c->fft_plan == 4 ? fwd_passes(std::integral_constant<int, 4>())
                             : fwd_passes(std::integral_constant<int, 0>())))
    return rc;
  if constexpr (sizeof(F) == 4) {
    // field passes and z inverse: plan 4 (register-resident square
    // transforms) when chosen, else the default plan
    auto field_passes = [&](auto plan) -> int {
      constexpr int V = decltype(plan)::value;
      for (const auto& l : P.xc) {
        OCSC_SWITCH_M(l.M, {
          constexpr int NT = field_threads<MM, V>();
          kb_xconv_field<MM, V><<<(unsigned)l.grid, NT, FieldTile<MM, NT, V>::lds, st>>>(J, l.seg, c->twf + MM,
                                                                                         c->twd + MM);
        });
      }
      for (const auto& l : P.yi) {
        OCSC_SWITCH_M(l.M, {
          constexpr int NT = field_threads<MM, V>();
          kb_yinv_field<MM, V><<<(unsigned)l.grid, NT, FieldTile<MM, NT, V>::lds, st>>>(J, l.seg, c->twf + MM,
                                                                                        c->twd + MM);
        });
      }
      for (const auto& l : P.zi) {
        OCSC_SWITCH_M(l.M, {
          kb_zinv3<MM, V><<<dim3((unsigned)l.grid, 3), pass_threads<float, MM, V>(), pass_lds<float, MM, V>(), st>>>(
              J, l.seg, c->twf + MM, c->twd + MM);
        });
      }
      return OCSC_OK;
    };
    if ((rc = c->fft_plan == 4 ? field_passes(std::integral_constant<int, 4>())
                               : field_passes(std::integral_constant<int, 0>())))
      return rc;
  } else {
    auto conv_passes = [&](auto plan) -> int {
      constexpr int V = decltype(plan)::value;
      for (const auto& l : P.xc) {
        OCSC_SWITCH_M(l.M, {
          kb_xconv64<MM, V><<<(unsigned)l.grid, pass_threads<double, MM, V>(), pass_lds<double, MM, V>(), st>>>(
              J, l.seg, c->twd + MM);
        });
      }
      for (const auto& l : P.yi) {
        OCSC_SWITCH_M(l.M, {
          constexpr int VY = plan_zf_yi<double, MM, V>();
          kb_yinv64<MM, VY><<<(unsigned)l.grid, pass_threads<double, MM, VY>(), pass_lds<double, MM, VY>(), st>>>(
              J, l.seg, c->twd + MM);
        });
      }
      for (const auto& l : P.zi) {
        OCSC_SWITCH_M(l.M, {
          kb_zinv64<MM, V><<<(unsigned)l.grid, pass_threads<double, MM, V>(), pass_lds<double, MM, V>(), st>>>(
              J, l.seg, c->twd + MM);
        });
      }
      return OCSC_OK;
    };
    if ((rc = c->fft_plan == 4 ? conv_passes(std::integral_constant<int, 4>())
                               : conv_passes(std::integral_constant<int, 0>())))
      return rc;
    kb_diff64<<<(unsigned)P.d64.grid, 256, 0, st>>>(J, P.d64.seg);
  }
  LAUNCHCHK("batched transforms");
  kb_kick<F><<<(unsigned)P.kick.grid, kKickPT, 0, st>>>(J, P.kick.seg);
  LAUNCHCHK("batched kick");
  return OCSC_OK;
}

int fused_back_prec(ocsc_ctx* c, const FusedPlan& P, int prec, hipStream_t st) {
  if (prec == OCSC_FP64) return fused_back<double, 2>(c, P, st);
  return fused_back<float, 1>(c, P, st);
}

int fused_round_prec(ocsc_ctx* c, const FusedPlan& P, int prec, hipStream_t st) {
  int rc;
  if ((rc = fused_front(P, prec, st))) return rc;
  return fused_back_prec(c, P, prec, st);
}

// nkicks rounds in one launch per pass.  With cell-ordered copies (option
// batch_sort, more than one round): the first round takes frame and extent
// from the callers' rows, then sorts all bunches into their copies with one
// radix sort (keys tagged by bunch) and kicks the copies; later rounds work on
// the copies; at the end the copies go back to the callers' order.  Same
// arithmetic and the same orders as the per-bunch path, so the same bits.
// *handled = false: the call does not fit (more than 2^32 particles in all
// or keys wider than 32 bits with the bunch tag) and is left to that path.
int apply_batched_fused(ocsc_ctx* c, const ocsc_bunch* b, int nb, const std::vector<Geometry>& geo,
                        const std::vector<KickInputs>& ins, int prec, int nkicks, int use_graph, hipStream_t st,
                        bool* handled) {
  *handled = false;
  std::vector<int> kicked;
  for (int i = 0; i < nb; ++i)
    if (b[i].zstep != 0) kicked.push_back(i);
  const int nk = (int)kicked.size();
  const bool copy = c->batch_sort && nkicks > 1;
  int64_t N = 0;
  int kbits = 0, jbits = 0;
  for (int i : kicked) {
    N += b[i].n;
    kbits = std::max(kbits, sort_end_bit(geo[i].d.nx, geo[i].d.ny, geo[i].d.nz, c->sort_cells));
  }
  while ((1 << jbits) < nk) ++jbits;
  if (copy && (N > (int64_t)0xFFFFFFFF || kbits + jbits > 32)) return OCSC_OK;
  *handled = true;
  if (nk == 0) return OCSC_OK;
  int rc;
  while ((int)c->flanes.size() < nk) {
    Lane* l = new Lane();
    if ((rc = init_lane(*l, false))) {
      free_lane(*l);
      delete l;
      return rc;
    }
    c->flanes.push_back(l);
  }
  for (int j = 0; j < nk; ++j) {
    const int i = kicked[j];
    if ((rc = ensure_prec(c, *c->flanes[j], geo[i], prec)) || (rc = ensure_part(c, *c->flanes[j], b[i].n)))
      return rc;
  }
  if (copy) {
    while ((int)c->bcopy.size() < nk) c->bcopy.push_back(new ocsc_ctx::BunchCopy());
    for (int j = 0; j < nk; ++j) {
      const int64_t n = b[kicked[j]].n;
      ocsc_ctx::BunchCopy& bc = *c->bcopy[j];
      if ((rc = ensure(bc.rs, (size_t)6 * n * 8)) || (rc = ensure(bc.qs, (size_t)n * 8)) ||
          (rc = ensure(bc.perm, (size_t)n * 4)))
        return rc;
    }
    if ((rc = ensure(c->keys[0], (size_t)N * 4)) || (rc = ensure(c->keys[1], (size_t)N * 4)) ||
        (rc = ensure(c->vals, (size_t)2 * N * 4)) || (rc = ensure(c->sort_tmp, rs_hist_bytes(N))) ||
        (rc = ensure(c->rec, (size_t)N * sizeof(PRec))))
      return rc;
  }
  std::vector<double*> rr(nb), rc_rows(nb);
  std::vector<int64_t> ldr(nb), nn(nb), ldc(nb);
  std::vector<const double*> qq(nb), qc(nb);
  for (int i = 0; i < nb; ++i) {
    nn[i] = b[i].n;
    rr[i] = b[i].rparticles;
    ldr[i] = b[i].ld;
    qq[i] = b[i].q_array;
  }
  for (int j = 0; j < nk && copy; ++j) {
    const int i = kicked[j];
    rc_rows[i] = (double*)c->bcopy[j]->rs.p;
    ldc[i] = b[i].n;
    qc[i] = (const double*)c->bcopy[j]->qs.p;
  }
  FusedPlan pc, ps;  // callers' rows, copies
  if ((rc = fused_plan(c, kicked, rr, ldr, qq, nn, geo, ins, prec, st, c->btab, copy, &pc))) return rc;
  if (copy && (rc = fused_plan(c, kicked, rc_rows, ldc, qc, nn, geo, ins, prec, st, c->btab2, true, &ps)))
    return rc;
  const FusedPlan& later = copy ? ps : pc;
  int first = 0;  // rounds done before the replayed ones
  if (copy) {
    // first round: frame and extent from the callers' rows, then the copies
    if ((rc = fused_front(pc, prec, st))) return rc;
    kb_sort_keys<<<(unsigned)pc.kick.grid, kPT, 0, st>>>(pc.jobs, pc.kick.seg, (uint32_t*)c->keys[0].p,
                                                        (PRec*)c->rec.p, c->sort_cells, kbits);
    uint32_t* const kk[2] = {(uint32_t*)c->keys[0].p, (uint32_t*)c->keys[1].p};
    uint32_t* const vv[2] = {(uint32_t*)c->vals.p, (uint32_t*)c->vals.p + N};
    const uint32_t* order = vv[rs_sort(kk, vv, N, kbits + jbits, (uint32_t*)c->sort_tmp.p, st)];
    kb_gather_copy<<<(unsigned)pc.kick.grid, 256, 0, st>>>(pc.jobs, pc.kick.seg, order, (const PRec*)c->rec.p);
    LAUNCHCHK("batched cell-ordered copies");
    if ((rc = fused_back_prec(c, ps, prec, st))) return rc;
    first = 1;
  }
  if (!use_graph || nkicks - first < 2) {
    for (int k = first; k < nkicks; ++k)
      if ((rc = fused_round_prec(c, later, prec, st))) return rc;
  } else {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHK(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
    const int r2 = fused_round_prec(c, later, prec, c->cap_stream);
    hipError_t e = hipStreamEndCapture(c->cap_stream, &graph);
    if (r2 || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return r2 ? r2 : fail(OCSC_EHIP, std::string("batched capture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    for (int k = first; k < nkicks && e == hipSuccess; ++k) e = hipGraphLaunch(exec, st);
    if (exec) (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(OCSC_EHIP, std::string("batched graph: ") + hipGetErrorString(e));
  }
  if (copy) {
    kb_copy_to_rec<<<(unsigned)pc.kick.grid, 256, 0, st>>>(pc.jobs, pc.kick.seg, (PRec*)c->rec.p);
    kb_rec_to_rows<<<(unsigned)pc.kick.grid, 256, 0, st>>>(pc.jobs, pc.kick.seg, (const PRec*)c->rec.p);
    LAUNCHCHK("batched caller-order copy");
  }
  int32_t err = 0;
  kb_errors<<<1, 64, 0, st>>>(pc.jobs, pc.nb, pc.err, 1);
  LAUNCHCHK("batched errors");
  HIPCHK(hipMemcpyAsync(&err, pc.err, sizeof(err), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (err) return value_error(err);
  return OCSC_OK;
}

}  // namespace

extern "C" {

int ocsc_track_drift_lsc(ocsc_ctx* c, double* r, int64_t ld, const double* q, int64_t n, const int32_t nmesh[3],
                         double E_GeV, double zstep, double Ld, int nsteps, int prec, int use_graph, double lsc_dz,
                         int step_profile, double smooth_param, double b0, double b1, const double* und_factors,
                         void* stream) {
  ocsc_track_opts o{};
  o.use_graph = use_graph;
  o.lsc = 1;
  o.lsc_dz = lsc_dz;
  o.lsc_step_profile = step_profile;
  o.lsc_smooth_param = smooth_param;
  o.lsc_b0 = b0;
  o.lsc_b1 = b1;
  o.lsc_und_factors = und_factors;
  return ocsc_track(c, r, ld, q, n, nmesh, E_GeV, zstep, Ld, nsteps, prec, &o, stream);
}

int ocsc_track_drift_random_mesh(ocsc_ctx* c, double* r, int64_t ld, const double* q, int64_t n,
                                 const int32_t nmesh[3], double E_GeV, double zstep, double Ld, int nsteps, int prec,
                                 int use_graph, const double* draws, void* stream) {
  if (!draws && nsteps > 0 && zstep != 0) return fail(OCSC_EINVAL, "null draws");
  ocsc_track_opts o{};
  o.use_graph = use_graph;
  o.mesh_draws = draws;
  return ocsc_track(c, r, ld, q, n, nmesh, E_GeV, zstep, Ld, nsteps, prec, &o, stream);
}

int ocsc_track_drift(ocsc_ctx* c, double* r, int64_t ld, const double* q, int64_t n, const int32_t nmesh[3],
                     double E_GeV, double zstep, double Ld, int nsteps, int prec, int use_graph, void* stream) {
  return track_impl(c, r, ld, q, n, nmesh, E_GeV, zstep, Ld, nsteps, prec, use_graph, stream, nullptr, nullptr);
}

int ocsc_track_drift_tws(ocsc_ctx* c, double* r, int64_t ld, const double* q, int64_t n, const int32_t nmesh[3],
                         double E_GeV, double zstep, double Ld, int nsteps, int prec, int use_graph, int mom_flags,
                         const double disp[4], double b0, double b1, double* tws_out, void* stream) {
  if (!tws_out) return fail(OCSC_EINVAL, "null tws_out");
  if ((mom_flags & OCSC_MOM_IMAX) && !(mom_flags & OCSC_MOM_BOUNDS))
    return fail(OCSC_EINVAL, "OCSC_MOM_IMAX selects the window centre and needs OCSC_MOM_BOUNDS");
  if ((mom_flags & OCSC_MOM_DISP_GIVEN) && !disp) return fail(OCSC_EINVAL, "OCSC_MOM_DISP_GIVEN needs disp");
  TwsReq t;
  t.bounds = (mom_flags & OCSC_MOM_BOUNDS) ? 1 : 0;
  t.imax = (mom_flags & OCSC_MOM_IMAX) ? 1 : 0;
  t.disp_mode = (mom_flags & OCSC_MOM_DISP_GIVEN) ? 1 : ((mom_flags & OCSC_MOM_AUTO_DISP) ? 2 : 0);
  t.b0 = b0;
  t.b1 = b1;
  if (t.disp_mode == 1)
    for (int k = 0; k < 4; ++k) t.disp[k] = disp[k];
  t.cap = nsteps + 1;
  return track_impl(c, r, ld, q, n, nmesh, E_GeV, zstep, Ld, nsteps, prec, use_graph, stream, &t, tws_out);
}

int ocsc_track(ocsc_ctx* c, double* r, int64_t ld, const double* q, int64_t n, const int32_t nmesh[3], double E_GeV,
               double zstep, double Ld, int nsteps, int prec, const ocsc_track_opts* o, void* stream) {
  if (!o) return fail(OCSC_EINVAL, "null opts");
  TwsReq t;
  const TwsReq* tw = nullptr;
  if (o->tws_out) {
    const int f = o->mom_flags;
    if ((f & OCSC_MOM_IMAX) && !(f & OCSC_MOM_BOUNDS))
      return fail(OCSC_EINVAL, "OCSC_MOM_IMAX selects the window centre and needs OCSC_MOM_BOUNDS");
    t.bounds = (f & OCSC_MOM_BOUNDS) ? 1 : 0;
    t.imax = (f & OCSC_MOM_IMAX) ? 1 : 0;
    t.disp_mode = (f & OCSC_MOM_DISP_GIVEN) ? 1 : ((f & OCSC_MOM_AUTO_DISP) ? 2 : 0);
    t.b0 = o->mom_b0;
    t.b1 = o->mom_b1;
    if (t.disp_mode == 1)
      for (int k = 0; k < 4; ++k) t.disp[k] = o->disp[k];
    t.cap = nsteps + 1;
    tw = &t;
  }
  if (o->lsc && o->lsc_dz >= 1e-10) {
    if (!c) return fail(OCSC_EINVAL, "null argument");
    LscReq lq;
    lq.k = lsc_const(o->lsc_step_profile, o->lsc_smooth_param, E_GeV, o->lsc_dz, 0.0, 0.0, kLscGraphBins);
    lq.b0 = o->lsc_b0;
    lq.b1 = o->lsc_b1;
    if (o->lsc_und_factors && nsteps > 0) {
      int rc;
      HIPCHK(hipSetDevice(c->device));
      hipStream_t st = (hipStream_t)stream;
      if ((rc = ensure(c->lsc_fac, (size_t)nsteps * sizeof(double) + 256))) return rc;
      HIPCHK(hipMemcpyAsync(c->lsc_fac.p, o->lsc_und_factors, (size_t)nsteps * sizeof(double), hipMemcpyHostToDevice,
                            st));
      HIPCHK(hipStreamSynchronize(st));  // the caller's (pageable) array may go away after the call
      int* ctr = (int*)((char*)c->lsc_fac.p + (size_t)nsteps * sizeof(double));
      k_counter_reset<<<1, 64, 0, st>>>(ctr);
      lq.k.factors = (const double*)c->lsc_fac.p;
      lq.k.fctr = ctr;
      lq.k.nfactors = nsteps;
    }
    return track_impl(c, r, ld, q, n, nmesh, E_GeV, zstep, Ld, nsteps, prec, o->use_graph, stream, tw, o->tws_out,
                      &lq, o->mesh_draws);
  }
  return track_impl(c, r, ld, q, n, nmesh, E_GeV, zstep, Ld, nsteps, prec, o->use_graph, stream, tw, o->tws_out,
                    nullptr, o->mesh_draws);
}

int ocsc_kick_times(ocsc_ctx* c, double* ms, int nout, int64_t* count) {
  if (!c || (!ms && nout > 0)) return fail(OCSC_EINVAL, "null argument");
  for (int i = 0; i < nout && i < (int)c->kick_ms.size(); ++i) ms[i] = c->kick_ms[i];
  if (count) *count = (int64_t)c->kick_ms.size();
  return OCSC_OK;
}

int ocsc_apply_batched(ocsc_ctx* c, const ocsc_bunch* b, int nb, int prec, int nkicks, int use_graph,
                       void* stream) {
  if (!c || (!b && nb > 0) || nb < 0 || nkicks < 0) return fail(OCSC_EINVAL, "bad arguments");
  if (prec != OCSC_FP64 && prec != OCSC_MIXED) return fail(OCSC_EINVAL, "bad precision");
  if (nb == 0 || nkicks == 0) return OCSC_OK;
  int rc;
  std::vector<Geometry> geo(nb);
  std::vector<KickInputs> ins(nb);
  for (int i = 0; i < nb; ++i) {
    const ocsc_bunch& u = b[i];
    if ((rc = check_common(c, u.rparticles, u.q_array, u.n, u.ld))) return rc;
    if (!(u.E_GeV > 0) || !std::isfinite(u.zstep)) return fail(OCSC_EINVAL, "E_GeV must be > 0 and zstep finite");
    if (u.random_mesh && nkicks > 1) return fail(OCSC_EINVAL, "random_mesh draws differ per kick: use nkicks = 1");
    if ((rc = make_geometry(u.nmesh, geo[i]))) return rc;
    ins[i] = kick_inputs(u.nmesh, u.n, u.E_GeV, u.zstep, u.random_mesh, u.mesh_scale, u.mesh_shift);
  }
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  if (c->batch_fused) {
    bool handled = false;
    rc = apply_batched_fused(c, b, nb, geo, ins, prec, nkicks, use_graph, st, &handled);
    if (handled) return rc;
  }
  const int nl = nb < kMaxLanes ? nb : kMaxLanes;
  while ((int)c->lanes.size() < nl) {
    Lane* l = new Lane();
    if ((rc = init_lane(*l, true))) {
      free_lane(*l);
      delete l;
      return rc;
    }
    c->lanes.push_back(l);
  }
  for (int i = 0; i < nb; ++i) {
    Lane& L = *c->lanes[i % nl];
    if ((rc = ensure_prec(c, L, geo[i], prec))) return rc;
    if ((rc = ensure_part(c, L, b[i].n))) return rc;
  }
  // Cell-ordered copies: the first kick of each bunch takes its frame and
  // extent from the caller's rows (as ocsc_apply), sorts the bunch into its
  // copy and kicks the copy; later kicks work on the copy; the copies go back
  // to the caller's order at the end.  One kick per call equals ocsc_apply
  // bit for bit; later kicks reduce the frame sums in cell order.
  std::vector<char> sorted(nb, 0);
  std::vector<SortBufs> sbs(nb);
  bool any_sorted = false;
  if (c->batch_sort) {
    while ((int)c->bcopy.size() < nb) c->bcopy.push_back(new ocsc_ctx::BunchCopy());
    for (int i = 0; i < nb; ++i) {
      const int64_t n = b[i].n;
      if (b[i].zstep == 0 || n > (int64_t)0xFFFFFFFF) continue;
      ocsc_ctx::BunchCopy& bc = *c->bcopy[i];
      Lane& L = *c->lanes[i % nl];
      if ((rc = ensure(bc.rs, (size_t)6 * n * 8)) || (rc = ensure(bc.qs, (size_t)n * 8)) ||
          (rc = ensure(bc.perm, (size_t)n * 4)) || (rc = ensure(L.sk[0], (size_t)n * 4)) ||
          (rc = ensure(L.sk[1], (size_t)n * 4)) || (rc = ensure(L.sv, (size_t)2 * n * 4)) ||
          (rc = ensure(L.sh, rs_hist_bytes(n))) || (rc = ensure(L.srec, (size_t)n * sizeof(PRec))))
        return rc;
      sorted[i] = 1;
      any_sorted = true;
    }
    for (int i = 0; i < nb; ++i) {
      if (!sorted[i]) continue;
      Lane& L = *c->lanes[i % nl];
      ocsc_ctx::BunchCopy& bc = *c->bcopy[i];
      sbs[i] = SortBufs{{(uint32_t*)L.sk[0].p, (uint32_t*)L.sk[1].p},
                        {(uint32_t*)L.sv.p, (uint32_t*)L.sv.p + b[i].n},
                        (uint32_t*)L.sh.p, (PRec*)L.srec.p, (double*)bc.rs.p, (double*)bc.qs.p,
                        (uint32_t*)bc.perm.p, false};
    }
  }
  // one round: every bunch kicked once, bunch i on lane i % nl; first: the
  // round that makes the cell-ordered copies
  const int prof = c->profiling, ovl = c->overlap;
  c->profiling = 0;  // stage events and the side stream are per context
  c->overlap = 0;
  auto fork = [&](hipStream_t origin) -> int {
    HIPCHK(hipEventRecord(c->ev_fork, origin));
    for (int l = 0; l < nl; ++l) HIPCHK(hipStreamWaitEvent(c->lanes[l]->stream, c->ev_fork, 0));
    return OCSC_OK;
  };
  auto join = [&](hipStream_t origin) -> int {
    for (int l = 0; l < nl; ++l) {
      HIPCHK(hipEventRecord(c->lanes[l]->done, c->lanes[l]->stream));
      HIPCHK(hipStreamWaitEvent(origin, c->lanes[l]->done, 0));
    }
    return OCSC_OK;
  };
  auto round = [&](hipStream_t origin, bool first) -> int {
    int r2;
    if ((r2 = fork(origin))) return r2;
    for (int i = 0; i < nb; ++i) {
      Lane& L = *c->lanes[i % nl];
      if (b[i].zstep == 0) continue;
      StepOpts o;
      double* rr = b[i].rparticles;
      int64_t ldr = b[i].ld;
      const double* qq = b[i].q_array;
      if (sorted[i] && first) {
        o.sort = &sbs[i];
      } else if (sorted[i]) {
        rr = sbs[i].rs;
        ldr = b[i].n;
        qq = sbs[i].qs;
      }
      if ((r2 = enqueue_kick_prec(c, L, geo[i], rr, ldr, qq, ins[i], prec, L.stream, o))) return r2;
    }
    return join(origin);
  };
  // the copies back to the callers' rows (whole records, then rows)
  auto unsort = [&](hipStream_t origin) -> int {
    int r2;
    if ((r2 = fork(origin))) return r2;
    for (int i = 0; i < nb; ++i) {
      if (!sorted[i]) continue;
      Lane& L = *c->lanes[i % nl];
      const int64_t n = b[i].n;
      const int nbk = (int)((n + 255) / 256);
      k_scatter_rec<<<nbk, 256, 0, L.stream>>>(sbs[i].rs, n, sbs[i].order_out, sbs[i].rec, n);
      k_rec_to_rows<<<nbk, 256, 0, L.stream>>>(sbs[i].rec, b[i].rparticles, b[i].ld, n);
      LAUNCHCHK("batched caller-order copy");
    }
    return join(origin);
  };
  auto restore = [&]() {
    c->profiling = prof;
    c->overlap = ovl;
  };
  if (!use_graph || any_sorted) {
    // the copying round eagerly (per bunch); with graphs the later rounds are captured
    if ((rc = round(st, true))) {
      restore();
      return rc;
    }
  }
  const int first_graph_round = (!use_graph || any_sorted) ? 1 : 0;
  if (!use_graph) {
    for (int k = 1; k < nkicks; ++k)
      if ((rc = round(st, false))) {
        restore();
        return rc;
      }
  } else if (nkicks > first_graph_round) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      restore();
      return fail(OCSC_EHIP, std::string("capture: ") + hipGetErrorString(e));
    }
    int r2 = round(c->cap_stream, !any_sorted);
    e = hipStreamEndCapture(c->cap_stream, &graph);
    if (r2 || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      restore();
      return r2 ? r2 : fail(OCSC_EHIP, std::string("capture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    for (int k = first_graph_round; k < nkicks && e == hipSuccess; ++k) e = hipGraphLaunch(exec, st);
    if (exec) (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      restore();
      return fail(OCSC_EHIP, std::string("batched graph: ") + hipGetErrorString(e));
    }
  }
  restore();
  if (any_sorted && (rc = unsort(st))) return rc;
  int32_t err_all = 0;
  for (int l = 0; l < nl; ++l) {
    int32_t err = 0;
    if ((rc = read_errors(c, *c->lanes[l], st, &err))) return rc;
    err_all |= err;
  }

  if (err_all) return value_error(err_all);
  return OCSC_OK;
}

int ocsc_set_option(ocsc_ctx* c, const char* name, int64_t value) {
  if (!c || !name) return fail(OCSC_EINVAL, "null argument");
  const std::string n(name);
  if (n == "deposit_xcd") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "deposit_xcd must be 0 or 1");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipDeviceSynchronize());
    c->deposit_xcd = (int)value;
    return OCSC_OK;
  }
  if (n == "sort_every") {
    if (value < 0 || value > (1 << 30)) return fail(OCSC_EINVAL, "sort_every must be >= 0");
    c->sort_every = (int)value;
    return OCSC_OK;
  }
  if (n == "kick_timing") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "kick_timing must be 0 or 1");
    c->kick_timing = (int)value;
    return OCSC_OK;
  }
  if (n == "overlap") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "overlap must be 0 or 1");
    c->overlap = (int)value;
    return OCSC_OK;
  }
  if (n == "permute_rec") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "permute_rec must be 0 or 1");
    c->permute_rec = (int)value;
    return OCSC_OK;
  }
  if (n == "deposit_agg") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "deposit_agg must be 0 or 1");
    c->deposit_agg = (int)value;
    return OCSC_OK;
  }
  if (n == "apply_sort") {
    if (value < 0 || value > 2) return fail(OCSC_EINVAL, "apply_sort must be 0, 1 or 2");
    c->apply_sort = (int)value;
    return OCSC_OK;
  }
  if (n == "grid_clear") {
    if (value < 0 || value > 2) return fail(OCSC_EINVAL, "grid_clear must be 0, 1 or 2");
    c->grid_clear = (int)value;
    return OCSC_OK;
  }
  if (n == "igf_blocks") {
    if (value < 0 || value > 65536) return fail(OCSC_EINVAL, "igf_blocks must be 0..65536");
    c->igf_blocks = (int)value;
    return OCSC_OK;
  }
  if (n == "fold_frame") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "fold_frame must be 0 or 1");
    c->fold_frame = (int)value;
    return OCSC_OK;
  }
  if (n == "q_uniform") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "q_uniform must be 0 or 1");
    c->q_uniform = (int)value;
    return OCSC_OK;
  }
  if (n == "rec_direct") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "rec_direct must be 0 or 1");
    c->rec_direct = (int)value;
    return OCSC_OK;
  }
  if (n == "apply_resort") {
    if (value < 1 || value > 1000) return fail(OCSC_EINVAL, "apply_resort must be 1..1000");
    c->apply_resort = (int)value;
    c->ao_age = 0;  // the next sorting call starts a new interval
    return OCSC_OK;
  }
  if (n == "batch_fused") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "batch_fused must be 0 or 1");
    c->batch_fused = (int)value;
    return OCSC_OK;
  }
  if (n == "apply_graph") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "apply_graph must be 0 or 1");
    c->apply_graph = (int)value;
    if (!value) drop_apply_graph(c);
    return OCSC_OK;
  }
  if (n == "deposit_tile") {
    if (value < 0 || value > 4) return fail(OCSC_EINVAL, "deposit_tile must be 0, 1, 2, 3 or 4");
    c->deposit_tile = (int)value;
    return OCSC_OK;
  }
  if (n == "batch_sort") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "batch_sort must be 0 or 1");
    c->batch_sort = (int)value;
    return OCSC_OK;
  }
  if (n == "fft_plan") {
    if (value < 0 || value > 4) return fail(OCSC_EINVAL, "fft_plan must be 0, 1, 2, 3 or 4");
    c->fft_plan = (int)value;
    return OCSC_OK;
  }
  if (n == "sort_cells") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "sort_cells must be 0 or 1");
    c->sort_cells = (int)value;
    return OCSC_OK;
  }
  if (n == "sort_ahead") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "sort_ahead must be 0 or 1");
    c->sort_ahead = (int)value;
    return OCSC_OK;
  }
  if (n == "resort_near") {
    if (value != 0 && value != 1) return fail(OCSC_EINVAL, "resort_near must be 0 or 1");
    c->resort_near = (int)value;
    return OCSC_OK;
  }
  return fail(OCSC_EINVAL, "unknown option " + n);
}

int ocsc_get_option(ocsc_ctx* c, const char* name, int64_t* value) {
  if (!c || !name || !value) return fail(OCSC_EINVAL, "null argument");
  const std::string n(name);
  if (n == "deposit_xcd") {
    *value = c->deposit_xcd;
    return OCSC_OK;
  }
  if (n == "sort_every") {
    *value = c->sort_every;
    return OCSC_OK;
  }
  if (n == "overlap") {
    *value = c->overlap;
    return OCSC_OK;
  }
  if (n == "kick_timing") {
    *value = c->kick_timing;
    return OCSC_OK;
  }
  if (n == "permute_rec") {
    *value = c->permute_rec;
    return OCSC_OK;
  }
  if (n == "deposit_agg") {
    *value = c->deposit_agg;
    return OCSC_OK;
  }
  if (n == "apply_sort") {
    *value = c->apply_sort;
    return OCSC_OK;
  }
  if (n == "apply_graph") {
    *value = c->apply_graph;
    return OCSC_OK;
  }
  if (n == "batch_fused") {
    *value = c->batch_fused;
    return OCSC_OK;
  }
  if (n == "grid_clear") {
    *value = c->grid_clear;
    return OCSC_OK;
  }
  if (n == "igf_blocks") {
    *value = c->igf_blocks;
    return OCSC_OK;
  }
  if (n == "fold_frame") {
    *value = c->fold_frame;
    return OCSC_OK;
  }
  if (n == "q_uniform") {
    *value = c->q_uniform;
    return OCSC_OK;
  }
  if (n == "apply_resort") {
    *value = c->apply_resort;
    return OCSC_OK;
  }
  if (n == "rec_direct") {
    *value = c->rec_direct;
    return OCSC_OK;
  }
  if (n == "deposit_tile") {
    *value = c->deposit_tile;
    return OCSC_OK;
  }
  if (n == "batch_sort") {
    *value = c->batch_sort;
    return OCSC_OK;
  }
  if (n == "fft_plan") {
    *value = c->fft_plan;
    return OCSC_OK;
  }
  if (n == "sort_cells") {
    *value = c->sort_cells;
    return OCSC_OK;
  }
  if (n == "sort_ahead") {
    *value = c->sort_ahead;
    return OCSC_OK;
  }
  if (n == "resort_near") {
    *value = c->resort_near;
    return OCSC_OK;
  }
  return fail(OCSC_EINVAL, "unknown option " + n);
}

int ocsc_cell_sort(ocsc_ctx* c, const uint32_t* keys, int64_t n, int end_bit, uint32_t* order, void* stream) {
  if (!c) return fail(OCSC_EINVAL, "null context");
  if (n < 0 || n > (int64_t)0xFFFFFFFF) return fail(OCSC_EINVAL, "n must be in [0, 2^32)");
  if (end_bit < 1 || end_bit > 32) return fail(OCSC_EINVAL, "end_bit must be in [1, 32]");
  if (n == 0) return OCSC_OK;
  if (!keys || !order) return fail(OCSC_EINVAL, "null keys or order");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  int rc;
  for (int k = 0; k < 2; ++k)
    if ((rc = ensure(c->keys[k], (size_t)n * 4))) return rc;
  if ((rc = ensure(c->vals, (size_t)2 * n * 4))) return rc;
  if ((rc = ensure(c->sort_tmp, rs_hist_bytes(n)))) return rc;
  uint32_t* kb[2] = {(uint32_t*)c->keys[0].p, (uint32_t*)c->keys[1].p};
  uint32_t* vb[2] = {(uint32_t*)c->vals.p, (uint32_t*)c->vals.p + n};
  HIPCHK(hipMemcpyAsync(kb[0], keys, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  const int res = rs_sort(kb, vb, n, end_bit, (uint32_t*)c->sort_tmp.p, st);
  LAUNCHCHK("cell sort");
  HIPCHK(hipMemcpyAsync(order, vb[res], (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return OCSC_OK;
}

int ocsc_set_profiling(ocsc_ctx* c, int on) {
  if (!c) return fail(OCSC_EINVAL, "null context");
  HIPCHK(hipSetDevice(c->device));
  if (on && !c->ev[0][0])
    for (auto& set : c->ev)
      for (hipEvent_t& e : set)
        if (hipEventCreate(&e) != hipSuccess) return fail(OCSC_EHIP, "event");
  int rc;
  if ((rc = flush_all(c))) return rc;  // settle kicks still in flight, then start afresh
  c->profiling = on ? 1 : 0;
  for (int i = 0; i < kNumStages; ++i) c->stage_ms[i] = 0;
  c->stage_calls = 0;
  return OCSC_OK;
}

int ocsc_stage_times(ocsc_ctx* c, double* ms, int nout, int64_t* calls) {
  if (!c || (!ms && nout > 0)) return fail(OCSC_EINVAL, "null argument");
  int rc;
  if ((rc = flush_all(c))) return rc;
  for (int i = 0; i < nout && i < kNumStages; ++i)
    ms[i] = c->stage_calls ? c->stage_ms[i] / (double)c->stage_calls : 0.0;
  if (calls) *calls = c->stage_calls;
  return kNumStages;
}

const char* ocsc_stage_name(int i) { return (i >= 0 && i < kNumStages) ? kStageNames[i] : ""; }

int ocsc_last_stats(ocsc_ctx* c, double* out, int nout, void* stream) {
  if (!c || (!out && nout > 0)) return fail(OCSC_EINVAL, "null argument");
  HIPCHK(hipSetDevice(c->device));
  Lane& L = c->lane0;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(L.sc_host, L.sc, sizeof(KickScalars), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const KickScalars& s = *L.sc_host;
  double v[20];
  v[0] = s.gamma0;
  v[1] = s.Pav;
  for (int k = 0; k < 9; ++k) v[2 + k] = s.T[k];
  for (int k = 0; k < 3; ++k) {
    v[11 + k] = s.steps[k];
    v[14 + k] = s.X_off[k];
  }
  v[17] = s.sumq;
  v[18] = (double)(s.err | s.err_sticky);
  v[19] = (double)s.rebuild;
  for (int k = 0; k < nout && k < 20; ++k) out[k] = v[k];
  return OCSC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Longitudinal space charge (LSC, ocelot/cpbd/sc.py:261-600)

namespace {

constexpr int kLscMaxBins = 1 << 22;  // one-shot kicks (host-planned grid)

LscConst lsc_const(int step_profile, double smooth, double E_GeV, double dz, double K_max, double fill_factor,
                   int cap) {
#pragma clang fp contract(off)
  LscConst k{};
  k.step_profile = step_profile;
  k.smooth = smooth;
  k.gamma = E_GeV / kMeGeV;                                               // sc.py:579
  k.v = std::sqrt(1 - 1 / (k.gamma * k.gamma)) * kC;                      // sc.py:580
  k.pc_ref = std::sqrt(E_GeV * E_GeV / (kMeGeV * kMeGeV) - 1) * kMeGeV;  // sc.py:593
  k.dz = dz;
  k.K_max = K_max;
  k.fill_factor = fill_factor;
  k.cap = cap;
  return k;
}

int lsc_buffers(ocsc_ctx* c, int64_t bins) {
  if (!c->lsc) {
    HIPCHK(hipMalloc(&c->lsc, sizeof(LscState)));
    HIPCHK(hipMalloc(&c->lsc_plan, sizeof(LscPlan)));
    HIPCHK(hipMalloc(&c->lsc_err, sizeof(int)));
    HIPCHK(hipHostMalloc(&c->lsc_host, sizeof(LscHost), hipHostMallocDefault));
  }
  return ensure(c->lsc_bins, (size_t)5 * bins * sizeof(double));
}

// slice statistics (sc.py:566-577) into c->lsc; the state is initialised by
// the caller
void lsc_stats(ocsc_ctx* c, const Rows& R, const double* q, int64_t n, int step_profile, hipStream_t st) {
  double* part = (double*)c->lane0.part.p;
  const int nblk = nred_blocks(c, n);
  k_lsc_stat<0><<<nblk, kPT, 0, st>>>(R, q, c->lsc, part);
  k_lsc_stat_reduce<0><<<1, 1024, 0, st>>>(c->lsc, part, nblk);
  k_lsc_stat<1><<<nblk, kPT, 0, st>>>(R, q, c->lsc, part);
  k_lsc_stat_reduce<1><<<1, 1024, 0, st>>>(c->lsc, part, nblk);
  k_lsc_stat<2><<<nblk, kPT, 0, st>>>(R, q, c->lsc, part);
  k_lsc_stat_reduce<2><<<1, 1024, 0, st>>>(c->lsc, part, nblk);
  if (!step_profile) {
    k_lsc_stat<3><<<nblk, kPT, 0, st>>>(R, q, c->lsc, part);
    k_lsc_stat_reduce<3><<<1, 1024, 0, st>>>(c->lsc, part, nblk);
  }
}

// profile, wake and kick from the histogram totals and the plan on the
// device; grids of `grid` workgroups (the plan's bins, or the cap)
int lsc_from_profile(ocsc_ctx* c, double* r, int64_t ld, int64_t n, const long long* tot, int grid, hipStream_t st,
                     const DriftArgs* dr = nullptr, const LscSump* sm = nullptr) {
  double* cur = (double*)c->lsc_bins.p;
  const int64_t cap = (int64_t)(c->lsc_bins.cap / (5 * sizeof(double)));
  double* bunch = cur + cap;
  double* Zr = bunch + cap;
  double* Zi = Zr + cap;
  double* W = Zi + cap;
  k_lsc_profile<<<1, 1024, 0, st>>>(tot, c->lsc_plan, cur, bunch);
  k_lsc_spectrum<<<grid, 256, 0, st>>>(bunch, c->lsc_plan, Zr, Zi);
  k_lsc_wake<<<grid, 256, 0, st>>>(Zr, Zi, c->lsc_plan, W);
  const int nblk = nred_blocks(c, n);
  if (n > 0 && dr && sm)
    k_lsc_kick<true, true><<<nblk, kPT, 0, st>>>(r, ld, n, W, c->lsc_plan, dr->L, dr->r56, sm->sc, sm->part);
  else if (n > 0 && dr)
    k_lsc_kick<true><<<nblk, kPT, 0, st>>>(r, ld, n, W, c->lsc_plan, dr->L, dr->r56);
  else if (n > 0 && sm)
    k_lsc_kick<false, true><<<nblk, kPT, 0, st>>>(r, ld, n, W, c->lsc_plan, 0.0, 0.0, sm->sc, sm->part);
  else if (n > 0)
    k_lsc_kick<false><<<nblk, kPT, 0, st>>>(r, ld, n, W, c->lsc_plan, 0.0, 0.0);
  LAUNCHCHK("lsc kernels");
  return OCSC_OK;
}

double cic_wscale(int64_t n) {
  int e = 0;
  (void)std::frexp((double)n, &e);
  return std::ldexp(1.0, 62 - e);  // n * wscale < 2^62
}

// one LSC kick with no host synchronisation (graph-capturable): state,
// statistics, plan, histogram, profile, wake, kick
int enqueue_lsc(ocsc_ctx* c, double* r, int64_t ld, const double* q, int64_t n, const LscConst& k, double b0,
                double b1, hipStream_t st, const DriftArgs* dr, const LscSump* sm) {
  const int cap = k.cap;
  const int ncic = kCicBlocks;
  k_lsc_init<<<1, 64, 0, st>>>(c->lsc, n, b0, b1);
  lsc_stats(c, Rows{r, ld}, q, n, k.step_profile, st);
  k_lsc_plan<<<1, 64, 0, st>>>(c->lsc, k, c->lsc_plan, c->lsc_err);
  long long* hist = (long long*)c->lsc_cic.p;
  long long* tot = hist + (size_t)ncic * cap;
  k_cic_grid<LscPlan><<<ncic, kPT, (size_t)cap * sizeof(long long), st>>>(Rows{r, ld}, n, c->lsc_plan, cic_wscale(n),
                                                                         hist);
  k_cic_sum<<<(cap + 63) / 64, 1024, 0, st>>>(hist, ncic, cap, &c->lsc_plan->nbins, tot);
  LAUNCHCHK("k_cic_grid");
  return lsc_from_profile(c, r, ld, n, tot, cap, st, dr, sm);
}

const char* lsc_err_text(int e) {
  switch (e) {
    case kLscNoSpread: return "current profile: tau has no spread or is not finite (cannot size the grid)";
    case kLscTooManyBins: return "current profile: more bins than allowed";
    case kLscEmptySlice: return "zero-size array to reduction operation maximum which has no identity";
    case kLscWideSmoothing: return "current profile: smoothing kernel too wide";
    default: return "LSC error";
  }
}

}  // namespace

extern "C" {

int ocsc_lsc_apply(ocsc_ctx* c, double* r, int64_t ld, const double* q, int64_t n, double E_GeV, double dz,
                   int step_profile, double smooth_param, double b0, double b1, double K_max, double fill_factor,
                   double* info, int ninfo, void* stream) {
  if (!c || ((!r || !q) && n > 0) || (ninfo > 0 && !info)) return fail(OCSC_EINVAL, "null argument");
  if (n < 0 || ld < n) return fail(OCSC_EINVAL, "bad n / ld");
  if (dz < 1e-10) return OCSC_OK;  // sc.py:556-558
  if (n == 0) return fail(OCSC_EVALUE, "zero-size array to reduction operation minimum which has no identity");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if ((rc = lsc_buffers(c, kLscGraphBins))) return rc;
  if ((rc = ensure(c->lsc_cic, ((size_t)kCicBlocks * kLscGraphBins + kLscGraphBins) * sizeof(long long)))) return rc;
  if ((rc = ensure(c->lane0.part, (size_t)12 * nred_blocks(c, n) * sizeof(double)))) return rc;
  // the whole kick with its grid planned on the device (no synchronisation
  // in the middle); a profile that needs more bins than that path holds is
  // planned again on the host from the same statistics (the bunch is not
  // kicked by a refused plan)
  k_counter_reset<<<1, 64, 0, st>>>(c->lsc_err);
  if ((rc = enqueue_lsc(c, r, ld, q, n,
                        lsc_const(step_profile, smooth_param, E_GeV, dz, K_max, fill_factor, kLscGraphBins), b0, b1,
                        st)))
    return rc;
  LscHost& H = *c->lsc_host;
  HIPCHK(hipMemcpyAsync(&H.s, c->lsc, sizeof(H.s), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&H.p, c->lsc_plan, sizeof(H.p), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (H.p.err == kLscTooManyBins) {
    const LscConst k = lsc_const(step_profile, smooth_param, E_GeV, dz, K_max, fill_factor, kLscMaxBins);
    lsc_make_plan(H.s, k, &H.p);
    if (H.p.err) return fail(OCSC_EVALUE, lsc_err_text(H.p.err));
    const int nb = H.p.nbins;
    if ((rc = lsc_buffers(c, nb))) return rc;
    HIPCHK(hipMemcpyAsync(c->lsc_plan, &H.p, sizeof(H.p), hipMemcpyHostToDevice, st));
    long long* tot = nullptr;
    if ((rc = cic_profile(c->lsc_cic, Rows{r, ld}, n, H.p.a, H.p.ds, nb, st, &tot))) return rc;
    if ((rc = lsc_from_profile(c, r, ld, n, tot, nb, st))) return rc;
    HIPCHK(hipStreamSynchronize(st));  // the plan is read from pinned memory
  } else if (H.p.err) {
    return fail(OCSC_EVALUE, lsc_err_text(H.p.err));
  }
  const int nbins = H.p.nbins;
  c->lsc_nbins = nbins;
  c->lsc_a = H.p.a;
  c->lsc_ds = H.p.ds;
  const double v_info[OCSC_LSC_NINFO] = {H.p.sigma, H.s.tau_std, (double)nbins, H.p.ds, H.p.a, H.s.sumq, H.s.cnt,
                                         H.s.tau_mean};
  for (int j = 0; j < ninfo && j < OCSC_LSC_NINFO; ++j) info[j] = v_info[j];
  return OCSC_OK;
}

int ocsc_lsc_profile(ocsc_ctx* c, double* s, double* current, double* wake, int64_t cap, int64_t* nbins) {
  if (!c || !nbins) return fail(OCSC_EINVAL, "null argument");
  *nbins = c->lsc_nbins;
  if (c->lsc_nbins == 0) return fail(OCSC_EINVAL, "no LSC kick on this context yet");
  if (cap < c->lsc_nbins) return fail(OCSC_EINVAL, "cap below the number of profile bins");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipDeviceSynchronize());
  const int nb = c->lsc_nbins;
  const int64_t bcap = (int64_t)(c->lsc_bins.cap / (5 * sizeof(double)));
  const double* dcur = (const double*)c->lsc_bins.p;
  if (current) HIPCHK(hipMemcpy(current, dcur, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost));
  if (wake) HIPCHK(hipMemcpy(wake, dcur + 4 * bcap, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost));
  if (s)
    for (int j = 0; j < nb; ++j) {
#pragma clang fp contract(off)
      s[j] = (double)j * c->lsc_ds + c->lsc_a;
    }
  return OCSC_OK;
}

int ocsc_lsc_wake(ocsc_ctx* c, const double* s, const double* bunch, int64_t nb, double gamma, double sigma, double dz,
                  int step_profile, double K_max, double fill_factor, double* wake, void* stream) {
  if (!c || !s || !bunch || !wake) return fail(OCSC_EINVAL, "null argument");
  if (nb < 2 || nb >= kLscMaxBins) return fail(OCSC_EINVAL, "profile length must be in [2, 2^22)");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if ((rc = lsc_buffers(c, nb))) return rc;
  c->lsc_nbins = 0;  // the last kick's profile is overwritten
  HIPCHK(hipStreamSynchronize(st));
  LscPlan& P = c->lsc_host->p;
  P = LscPlan{};
  P.nbins = (int)nb;
  P.scale = -1.0;  // wake_lsc returns minus the transform
  lsc_wake_grid(s[0], s[1], nb, &P);
  P.imp = lsc_imp(step_profile, gamma, sigma, dz, K_max, fill_factor);
  HIPCHK(hipMemcpyAsync(c->lsc_plan, &P, sizeof(P), hipMemcpyHostToDevice, st));
  const int64_t bcap = (int64_t)(c->lsc_bins.cap / (5 * sizeof(double)));
  double* db = (double*)c->lsc_bins.p + bcap;
  double* Zr = db + bcap;
  double* Zi = Zr + bcap;
  double* W = Zi + bcap;
  HIPCHK(hipMemcpyAsync(db, bunch, (size_t)nb * sizeof(double), hipMemcpyHostToDevice, st));
  k_lsc_spectrum<<<(int)nb, 256, 0, st>>>(db, c->lsc_plan, Zr, Zi);
  k_lsc_wake<<<(int)nb, 256, 0, st>>>(Zr, Zi, c->lsc_plan, W);
  LAUNCHCHK("k_lsc_wake");
  HIPCHK(hipMemcpyAsync(wake, W, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return OCSC_OK;
}

}  // extern "C"

namespace {
__global__ void k_lsc_impedance(const double* __restrict__ w, int64_t n, LscImp imp, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = lsc_impedance(w[i], imp);
}
}  // namespace

extern "C" {

int ocsc_lsc_impedance(ocsc_ctx* c, const double* w, int64_t n, double gamma, double sigma, double dz,
                       int step_profile, double* z_imag, void* stream) {
  if (!c || ((!w || !z_imag) && n > 0)) return fail(OCSC_EINVAL, "null argument");
  if (n < 0 || n > ((int64_t)1 << 30)) return fail(OCSC_EINVAL, "bad n");
  if (n == 0) return OCSC_OK;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if ((rc = ensure(c->lsc_cic, (size_t)2 * n * sizeof(double)))) return rc;
  double* dw = (double*)c->lsc_cic.p;
  double* dz_out = dw + n;
  HIPCHK(hipMemcpyAsync(dw, w, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
  const LscImp imp = lsc_imp(step_profile, gamma, sigma, dz, 0.0, 0.0);
  k_lsc_impedance<<<(int)((n + 255) / 256), 256, 0, st>>>(dw, n, imp, dz_out);
  LAUNCHCHK("k_lsc_impedance");
  HIPCHK(hipMemcpyAsync(z_imag, dz_out, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return OCSC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// CSR kick tail (ocelot/cpbd/csr.py:69-79, 1105-1143)

namespace {

// c = conv(a, b)[:P+Q-2] / div * mul on the device; a, b host arrays
int csr_conv_device(ocsc_ctx* c, const double* a, int64_t P, const double* b, int64_t Q, double div, double mul,
                    hipStream_t st, double** out) {
  if (P < 1 || Q < 1 || P + Q > (1 << 24)) return fail(OCSC_EINVAL, "convolution lengths must be >= 1, sum < 2^24");
  const int64_t L = P + Q - 2;
  int rc;
  if ((rc = ensure(c->csr_bins, (size_t)(P + Q + std::max<int64_t>(L, 1)) * sizeof(double)))) return rc;
  double* da = (double*)c->csr_bins.p;
  double* db = da + P;
  double* dc = db + Q;
  HIPCHK(hipMemcpyAsync(da, a, (size_t)P * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(db, b, (size_t)Q * sizeof(double), hipMemcpyHostToDevice, st));
  if (L > 0) k_conv<<<(int)L, 256, 0, st>>>(da, (int)P, db, (int)Q, div, mul, dc);
  LAUNCHCHK("k_conv");
  *out = dc;
  return OCSC_OK;
}

}  // namespace

extern "C" {

int ocsc_csr_convolution(ocsc_ctx* c, const double* a, int64_t na, const double* b, int64_t nb, double* out,
                         void* stream) {
  if (!c || !a || !b || (!out && na + nb > 2)) return fail(OCSC_EINVAL, "null argument");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  double* dc = nullptr;
  int rc;
  if ((rc = csr_conv_device(c, a, na, b, nb, 1.0, 1.0, st, &dc))) return rc;
  if (na + nb > 2) HIPCHK(hipMemcpyAsync(out, dc, (size_t)(na + nb - 2) * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return OCSC_OK;
}

int ocsc_csr_kick(ocsc_ctx* c, double* r, int64_t ld, int64_t n, double E_GeV, const double* lam_ds, int64_t nlam,
                  const double* K1, int64_t nk, double s1, double s2, double Ns, double delta_s, double* lam_K1,
                  void* stream) {
#pragma clang fp contract(off)
  if (!c || !lam_ds || !K1 || (!r && n > 0)) return fail(OCSC_EINVAL, "null argument");
  if (n < 0 || ld < n) return fail(OCSC_EINVAL, "bad n / ld");
  if (nlam + nk - 2 < 1) return fail(OCSC_EVALUE, "array of sample points is empty");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  // csr.py:1105-1106 and 1138-1143, in the reference's operation order
  const double stp = (s2 - s1) / Ns;
  const double sa = s1 + stp / 2.;
  const double inv_st = 1. / stp;
  const double off = 0. - sa / stp;
  const double pc_ref = std::sqrt(E_GeV * E_GeV / (kMeGeV * kMeGeV) - 1) * kMeGeV;
  std::vector<double> krev(nk);
  for (int64_t k = 0; k < nk; ++k) krev[k] = K1[nk - 1 - k];
  double* dlk = nullptr;
  int rc;
  if ((rc = csr_conv_device(c, lam_ds, nlam, krev.data(), nk, stp, delta_s, st, &dlk))) return rc;
  const int L = (int)(nlam + nk - 2);
  if (n > 0) {
    k_csr_kick<<<nred_blocks(c, n), kPT, 0, st>>>(r, ld, n, dlk, L, inv_st, off, pc_ref);
    LAUNCHCHK("k_csr_kick");
  }
  if (lam_K1) HIPCHK(hipMemcpyAsync(lam_K1, dlk, (size_t)L * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));  // the host copy of the reversed kernel is read by the upload
  return OCSC_OK;
}

}  // extern "C"
