/*
 * ocsc -- MI355X (gfx950) 3D space-charge kick, C ABI.
 *
 * Drop-in compute for Ocelot's SpaceCharge physics process
 * (ocelot/cpbd/sc.py:76-258).  The Python class ocelot_amd.SpaceCharge binds
 * these entry points with ctypes; any other host language binds them the
 * same way.  No PyTorch or HIP types appear in the signatures: particle
 * buffers are plain device pointers (allocated by the caller, e.g. as torch
 * tensors), streams are passed as void* (a hipStream_t, NULL = default).
 *
 * Particle layout is Ocelot's ParticleArray.rparticles
 * (ocelot/cpbd/beam/particle.py:79-84): six FP64 rows x, x', y, y', tau,
 * delta, row i starting at rparticles + i*ld (ld >= n; ld = n for a
 * contiguous (6, n) array).  Charges q_array are FP64 [C].
 *
 * Return value of every entry point: 0 on success, otherwise one of the
 * OCSC_E* codes; ocsc_last_error() gives a message for the calling thread.
 * OCSC_EVALUE marks inputs for which the reference raises ValueError (empty
 * bunch, zero total charge, non-finite coordinates); rparticles is left
 * untouched in that case, as in the reference.
 *
 * Contexts are independent and may live on different devices; one context
 * must not be used from two host threads at once.
 */
#ifndef OCSC_H
#define OCSC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OCSC_OK 0
#define OCSC_EINVAL 1 /* bad argument (null pointer, nmesh < 4, n < 0, ...) */
#define OCSC_EVALUE 2 /* input the reference rejects with ValueError */
#define OCSC_EHIP 3   /* HIP runtime error */
#define OCSC_ENOMEM 4 /* device allocation failed */

#define OCSC_FP64 0  /* FP64 everywhere */
#define OCSC_MIXED 1 /* FP32 mesh, FFT and gather; FP64 frame, Green's function, momentum update */

typedef struct ocsc_ctx ocsc_ctx;

/* Per-process state of one SpaceCharge instance on one device: workspaces,
 * twiddle tables, the kernel-spectrum cache.  Replaces the implicit state of
 * SpaceCharge.__init__ (ocelot/cpbd/sc.py:92-102), which holds no buffers. */
int ocsc_create(ocsc_ctx** ctx, int device);
int ocsc_destroy(ocsc_ctx* ctx);

/* Message of the last failing call on this host thread ("" if none). */
const char* ocsc_last_error(void);

/* One space-charge kick, SpaceCharge.apply(p_array, zstep)
 * (ocelot/cpbd/sc.py:208-251).  All six rows of rparticles are rewritten in
 * place.  zstep == 0 is a no-op, as in the reference (sc.py:210-212).
 * random_mesh != 0 applies the two host draws of the reference
 * (sc.py:174-175, 184-185): the extent factor mesh_scale ~ U(1,1.1) and the
 * offset mesh_shift ~ U(-0.5,0.5); the caller draws them in that order.
 * Asynchronous on stream except when an input error must be reported: the
 * call then synchronises to read the device error flags (sync != 0), or
 * defers them to ocsc_check (sync == 0). */
int ocsc_apply(ocsc_ctx* ctx, double* rparticles, int64_t ld, const double* q_array, int64_t n,
               const int32_t nmesh[3], double E_GeV, double zstep, int random_mesh, double mesh_scale,
               double mesh_shift, int precision, int sync, void* stream);

/* ocsc_apply that also writes, for every particle, the flat C-order NGP cell
 * index its charge was deposited to (the inds of ocelot/cpbd/sc.py:187-192:
 * (floor(X0)+1) ny nz + (floor(X1)+1) nz + floor(X2)+1), as computed by the
 * deposit kernel of this kick, into the device array cells[n] (int32).  With
 * cells == NULL it is ocsc_apply. */
int ocsc_apply_cells(ocsc_ctx* ctx, double* rparticles, int64_t ld, const double* q_array, int64_t n,
                     const int32_t nmesh[3], double E_GeV, double zstep, int random_mesh, double mesh_scale,
                     double mesh_shift, int precision, int sync, int32_t* cells, void* stream);

/* ocsc_apply that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * workspace of the kick lives in one static device array of
 * OCSC_NOALLOC_SCRATCH_BYTES (40 MiB) per GPU, held until the process ends;
 * meshes whose workspace does not fit (above 64^3 in either precision, or
 * with option "deposit_xcd" = 1 on large meshes) return OCSC_EINVAL before
 * any work.  The caller passes rparticles and q_array as for ocsc_apply; the
 * result is the same bit for bit.  All contexts on one GPU share its array:
 * calls that overlap in time on one GPU (other streams, other contexts,
 * other host threads) race and must be serialised by the caller.  The
 * kernel-spectrum cache lives in the array too.  With sync == 0, input errors
 * are reported by the next call with sync != 0 on that GPU. */
#define OCSC_NOALLOC_SCRATCH_BYTES (40u << 20)
int ocsc_apply_noalloc(ocsc_ctx* ctx, double* rparticles, int64_t ld, const double* q_array, int64_t n,
                       const int32_t nmesh[3], double E_GeV, double zstep, int random_mesh, double mesh_scale,
                       double mesh_shift, int precision, int sync, void* stream);

/* Error flags left by kicks issued with sync == 0 (synchronises stream);
 * returns OCSC_EVALUE if any kick saw an invalid bunch and clears the flag. */
int ocsc_check(ocsc_ctx* ctx, void* stream);

/* SpaceCharge.el_field (ocelot/cpbd/sc.py:170-205): lab-frame field (V/m) at
 * rest-frame positions xyz (three FP64 rows x, y, z with z not yet scaled by
 * gamma, row stride ld).  Exyz: three FP64 rows, row stride lde.  cells
 * (optional, may be NULL): the flat C-order NGP cell index of each particle
 * (sc.py:191-192), int32.  Synchronous. */
int ocsc_el_field(ocsc_ctx* ctx, const double* xyz, int64_t ld, const double* q_array, int64_t n,
                  const int32_t nmesh[3], double gamma, int random_mesh, double mesh_scale, double mesh_shift,
                  int precision, double* Exyz, int64_t lde, int32_t* cells, void* stream);

/* SpaceCharge.sym_kernel (ocelot/cpbd/sc.py:109-133): integrated Green's
 * function on nmesh[0] x nmesh[1] x nmesh[2] cells of size h, FP64, written
 * to the device array K (C order).  Synchronous. */
int ocsc_sym_kernel(ocsc_ctx* ctx, const int32_t nmesh[3], const double h[3], double* K, void* stream);

/* SpaceCharge.potential (ocelot/cpbd/sc.py:135-168): potential of the device
 * charge grid rho (FP64, C order) with cell sizes h.  phi is FP64 for
 * precision OCSC_FP64 and FP32 for OCSC_MIXED.  Synchronous. */
int ocsc_potential(ocsc_ctx* ctx, const double* rho, const int32_t nmesh[3], const double h[3], void* phi,
                   int precision, void* stream);

/* First-order drift of length L (ocelot/cpbd/r_matrix.py:41-91 with
 * k1 = hx = 0, applied as in ocelot/cpbd/transformations/transfer_map.py:40-50):
 * x += L x', y += L y', tau += r56 delta, r56 = -L / (beta^2 gamma^2). */
int ocsc_drift(ocsc_ctx* ctx, double* rparticles, int64_t ld, int64_t n, double L, double E_GeV, void* stream);

/* nsteps x [drift(L); kick(zstep)] on a device-resident bunch, the tracking
 * loop of ocelot/cpbd/track.py:470-477 restricted to drifts.  use_graph != 0
 * captures one step in a HIP graph once and replays it.  Error flags are
 * checked once at the end (synchronises). */
int ocsc_track_drift(ocsc_ctx* ctx, double* rparticles, int64_t ld, const double* q_array, int64_t n,
                     const int32_t nmesh[3], double E_GeV, double zstep, double L, int nsteps, int precision,
                     int use_graph, void* stream);

/* Beam moments behind get_envelope (ocelot/cpbd/beam/analysis.py:27-224),
 * computed on the device over the six rows and q_array.  flags:
 * OCSC_MOM_BOUNDS selects the particles with
 * z0 + b0 std(tau) <= tau <= z0 + b1 std(tau) (analysis.py:59-67), z0 =
 * mean(tau) (slice=None) or, with OCSC_MOM_IMAX as well, the peak of the
 * smoothed current profile (slice="Imax", analysis.py:61-64, 266-340);
 * OCSC_MOM_AUTO_DISP estimates the dispersion from the selected
 * particles (analysis.py:90-98); OCSC_MOM_DISP_GIVEN takes disp = {Dx, Dxp,
 * Dy, Dyp} (analysis.py:99-103); otherwise D = 0.  out (host) receives
 * min(nout, OCSC_MOM_NOUT) values: [0] selected count, [1] sum of all q,
 * [2] <p>, [3..6] Dx Dxp Dy Dyp, [7..11] <x> <y> <px> <py> <tau> of the
 * corrected coordinates, [12..23] xx xpx pxpx yy ypy pypy tautau xy pxpy xpy
 * ypx pp, [24] mean(tau) and [25] std(tau) over all particles, [26] the
 * window centre z0.  With fewer than 3 selected particles only [0..2] are
 * meaningful.  Synchronous. */
#define OCSC_MOM_BOUNDS 1
#define OCSC_MOM_AUTO_DISP 2
#define OCSC_MOM_DISP_GIVEN 4
#define OCSC_MOM_IMAX 8
#define OCSC_MOM_NOUT 27
int ocsc_moments(ocsc_ctx* ctx, const double* rparticles, int64_t ld, const double* q_array, int64_t n, int flags,
                 const double disp[4], double b0, double b1, double* out, int nout, void* stream);

/* ocsc_track_drift with the beam moments of ocsc_moments taken before the
 * first step and after every step (after the kick, before the next drift),
 * as track() does with get_envelope (ocelot/cpbd/track.py:462-485).  mom_flags,
 * disp, b0, b1 as in ocsc_moments; with OCSC_MOM_IMAX the current profile of
 * each step is planned on the device, up to 8192 bins (a step that would need
 * more keeps the window at mean(tau) and the call returns OCSC_EVALUE).
 * tws_out (host) receives (nsteps + 1) rows of OCSC_MOM_NOUT values.  The
 * particles end up exactly as with ocsc_track_drift. */
int ocsc_track_drift_tws(ocsc_ctx* ctx, double* rparticles, int64_t ld, const double* q_array, int64_t n,
                         const int32_t nmesh[3], double E_GeV, double zstep, double L, int nsteps, int precision,
                         int use_graph, int mom_flags, const double disp[4], double b0, double b1, double* tws_out,
                         void* stream);

/* LSC.apply (ocelot/cpbd/sc.py:538-600): one longitudinal space-charge kick
 * of length dz on the device-resident bunch, delta += W(tau) 1e-9 / pc_ref.
 * The transverse size comes from the slice mean + b0 std <= tau < mean + b1 std
 * of tau (sc.py:566-577): the mean of std(x) and std(y) there, or with
 * step_profile != 0 the smaller half-extent of x and y.  The current profile
 * is s_to_cur(tau, std(tau) smooth_param, sum q, v) (beam/analysis.py:266-340)
 * and the wake that of the round Gaussian (step_profile == 0, sc.py:302-341)
 * or uniform (sc.py:343-369) beam impedance times 1 + K_max^2 fill_factor/2
 * (sc.py:436-441; K_max and fill_factor as LSC.apply derives them from the
 * undulators of the line, 0 without).  dz < 1e-10 does nothing.  info (host,
 * may be NULL) receives min(ninfo, OCSC_LSC_NINFO) values: [0] transverse
 * size, [1] std(tau), [2] profile bins, [3] bin width, [4] first bin
 * position, [5] sum of q, [6] particles in the slice, [7] mean(tau).
 * OCSC_EVALUE where the reference raises (empty bunch, no spread in tau,
 * non-finite tau, empty slice with the step profile); the bunch is then not
 * kicked.  The profile grid is planned on the device (up to 4096 bins;
 * larger grids are planned again on the host from the same statistics);
 * synchronises once, at the end. */
#define OCSC_LSC_NINFO 8
int ocsc_lsc_apply(ocsc_ctx* ctx, double* rparticles, int64_t ld, const double* q_array, int64_t n, double E_GeV,
                   double dz, int step_profile, double smooth_param, double b0, double b1, double K_max,
                   double fill_factor, double* info, int ninfo, void* stream);

/* Options of ocsc_track: everything that can run inside the captured step of
 * a drift-line tracking loop.  Zero-initialise, then set what is wanted. */
typedef struct ocsc_track_opts {
  int use_graph;             /* capture one step in a HIP graph and replay it */
  /* per-step moments (as ocsc_track_drift_tws) when tws_out != NULL */
  double* tws_out;           /* (nsteps + 1) rows of OCSC_MOM_NOUT values */
  int mom_flags;             /* OCSC_MOM_* */
  double mom_b0, mom_b1;
  double disp[4];
  /* an LSC kick after every space-charge kick (as ocsc_track_drift_lsc) when lsc != 0 */
  int lsc;
  double lsc_dz;
  int lsc_step_profile;
  double lsc_smooth_param, lsc_b0, lsc_b1;
  const double* lsc_und_factors; /* per step, or NULL */
  /* random_mesh draws (as ocsc_track_drift_random_mesh), or NULL */
  const double* mesh_draws;
} ocsc_track_opts;

/* nsteps x [drift(L); SpaceCharge kick(zstep) (none if zstep == 0); LSC kick;
 * moments] on a device-resident bunch, the tracking loop of
 * ocelot/cpbd/track.py:462-485 on a drift line with the processes and Twiss
 * that opts asks for.  The other ocsc_track_drift* entry points are this call
 * with one option set. */
int ocsc_track(ocsc_ctx* ctx, double* rparticles, int64_t ld, const double* q_array, int64_t n,
               const int32_t nmesh[3], double E_GeV, double zstep, double L, int nsteps, int precision,
               const ocsc_track_opts* opts, void* stream);

/* ocsc_track_drift with SpaceCharge(random_mesh=True) (ocelot/cpbd/sc.py:
 * 174-175, 184-185): draws (host) holds nsteps pairs (mesh_scale, mesh_shift),
 * the two global numpy draws of each kick in the reference's order; kick k
 * takes pair k through a device-side counter, so the captured step graphs
 * replay with a new pair each time. */
int ocsc_track_drift_random_mesh(ocsc_ctx* ctx, double* rparticles, int64_t ld, const double* q_array, int64_t n,
                                 const int32_t nmesh[3], double E_GeV, double zstep, double L, int nsteps,
                                 int precision, int use_graph, const double* draws, void* stream);

/* ocsc_track_drift with an LSC kick (as ocsc_lsc_apply, step length lsc_dz)
 * after every space-charge kick: per step drift(L),
 * SpaceCharge.apply(zstep), LSC.apply(lsc_dz), the order track() applies two
 * processes added in that order (ocelot/cpbd/track.py:470-477).  zstep == 0
 * leaves the space-charge kick out; lsc_dz < 1e-10 the LSC kick.  Graph-
 * capturable: the profile grid is planned on the device, up to 4096 bins;
 * a step whose profile would need more, or whose bunch has no spread in tau
 * or an empty step-profile slice, gets no LSC kick and the call returns
 * OCSC_EVALUE at the end.  und_factors (host, may be NULL: no undulator)
 * holds the impedance factor 1 + K_max^2 fill_factor/2 of each step, as
 * LSC.apply derives it from the line at that step (sc.py:559-564, 436-441).
 * Cannot be combined with the option "kick_timing". */
int ocsc_track_drift_lsc(ocsc_ctx* ctx, double* rparticles, int64_t ld, const double* q_array, int64_t n,
                         const int32_t nmesh[3], double E_GeV, double zstep, double L, int nsteps, int precision,
                         int use_graph, double lsc_dz, int step_profile, double smooth_param, double b0, double b1,
                         const double* und_factors, void* stream);

/* Profile grid, current (A) and wake W (V, the kick before scaling) of the
 * last ocsc_lsc_apply on this context (OCSC_EINVAL after an LSC tracking call
 * or ocsc_lsc_wake, which reuse the arrays) (the quantities B[:, 0], B[:, 1] and W
 * of sc.py:582-587), copied to host arrays of at least *nbins entries (any
 * may be NULL).  *nbins is set in every case; OCSC_EINVAL if cap < *nbins.
 * Synchronises the device. */
int ocsc_lsc_profile(ocsc_ctx* ctx, double* s, double* current, double* wake, int64_t cap, int64_t* nbins);

/* LSC.wake_lsc (ocelot/cpbd/sc.py:411-465): wake (V/C) of the normalised
 * profile bunch[nb] on the uniform grid s[nb] (host arrays), for the
 * Gaussian (step_profile == 0) or step transverse profile of size sigma.
 * Synchronous. */
int ocsc_lsc_wake(ocsc_ctx* ctx, const double* s, const double* bunch, int64_t nb, double gamma, double sigma,
                  double dz, int step_profile, double K_max, double fill_factor, double* wake, void* stream);

/* LSC.imp_lsc / LSC.imp_step_lsc (ocelot/cpbd/sc.py:302-369): imaginary part
 * of the impedance (ohm; the real part is zero) at the n angular frequencies
 * w (host arrays).  Synchronous. */
int ocsc_lsc_impedance(ocsc_ctx* ctx, const double* w, int64_t n, double gamma, double sigma, double dz,
                       int step_profile, double* z_imag, void* stream);

/* csr_convolution (ocelot/cpbd/csr.py:69-79): the first na + nb - 2 terms of
 * the linear convolution of a[na] and b[nb] (host arrays) into out.
 * Synchronous. */
int ocsc_csr_convolution(ocsc_ctx* ctx, const double* a, int64_t na, const double* b, int64_t nb, double* out,
                         void* stream);

/* The kick at the end of CSR.apply (ocelot/cpbd/csr.py:1105-1143) on the
 * device-resident bunch: lam_K1 = csr_convolution(lam_ds, K1 reversed) / st *
 * delta_s with st = (s2 - s1) / Ns, then delta += interp(-tau / st - sa / st,
 * 0..len-1, lam_K1) 1e-9 / pc_ref with sa = s1 + st / 2 (numpy.interp's
 * rules).  lam_ds[nlam] is the equidistant line charge and (s1, s2, Ns) its
 * mesh, as the reference's binning returns them; K1[nk] the step-averaged CSR
 * kernel (host arrays).  lam_K1 (host, may be NULL) receives the nlam + nk - 2
 * wake values.  Synchronous. */
int ocsc_csr_kick(ocsc_ctx* ctx, double* rparticles, int64_t ld, int64_t n, double E_GeV, const double* lam_ds,
                  int64_t nlam, const double* K1, int64_t nk, double s1, double s2, double Ns, double delta_s,
                  double* lam_K1, void* stream);

/* One bunch of a batched call (fields as in ocsc_apply). */
typedef struct ocsc_bunch {
  double* rparticles;
  int64_t ld;
  const double* q_array;
  int64_t n;
  int32_t nmesh[3];
  double E_GeV;
  double zstep;
  int32_t random_mesh;
  double mesh_scale;
  double mesh_shift;
} ocsc_bunch;

/* nkicks rounds of SpaceCharge.apply over nb independent bunches of uneven
 * size, mesh and energy (each round kicks every bunch once).  With option
 * "batch_fused" = 1 (default) every pass of a round is one launch over all
 * bunches (a device table maps each workgroup to its bunch; the transform
 * passes launch once per FFT length along their axis), each bunch with a
 * workspace of its own; with 0 the bunches are spread over up to 8
 * workspaces with their own streams; both give bit-identical results.  With
 * use_graph != 0 one round is captured in a HIP graph and replayed.  With
 * option "batch_sort" = 1 (default) each bunch is kicked as a
 * cell-ordered copy kept across the rounds and written back to the caller's
 * order at the end: one round equals ocsc_apply bit for bit; in later rounds
 * the frame sums run in cell order (results equal nkicks calls of ocsc_apply
 * to round-off).  With "batch_sort" = 0 every round equals ocsc_apply bit for
 * bit.  random_mesh requires nkicks == 1.  Synchronises at the end to report
 * input errors.  No reference counterpart beyond calling apply per bunch. */
int ocsc_apply_batched(ocsc_ctx* ctx, const ocsc_bunch* bunches, int nb, int precision, int nkicks, int use_graph,
                       void* stream);

/* Tuning options of a context (the autotuner's knobs; no reference
 * counterpart).  "deposit_xcd": 0 = one charge grid, device-scope atomics;
 * 1 = one grid copy per XCD with atomics that complete in that XCD's L2,
 * summed exactly in the first FFT pass.  Both give bit-identical results.
 * "sort_every": ocsc_track_drift keeps the bunch in mesh-cell order and
 * re-sorts it every k steps (0: never; default 10).  "overlap": 1
 * (default) builds the Green's-function spectrum on a second stream
 * concurrently with the charge deposit (the FP64 Green's function is
 * arithmetic-bound, the deposit memory-bound), 0 runs it in line.  Both give
 * bit-identical results.  "kick_timing": 1 times the kick kernel of every
 * tracking step (see ocsc_kick_times), 0 (default) does not.
 * "permute_rec": 1 (default) stages the permutations between the caller's
 * order and the cell order of ocsc_track_drift through 64-byte particle
 * records (one whole memory segment per particle on the random side), 0
 * permutes row by row.  Both give bit-identical results.  "deposit_agg": 1
 * sums runs of particles in the same cell across the wavefront before the
 * deposit's LDS table (fewer same-slot LDS atomics in cell order), 0
 * (default) does not; bit-identical results.  "apply_sort": one-shot kicks
 * (ocsc_apply) deposit and kick a cell-ordered copy of the bunch and write
 * the result back in the caller's order: 0 never, 1 always, 2 (default) in
 * the FP64 form from 1000000 particles on; bit-identical results
 * (ocsc_apply_cells and ocsc_apply_noalloc never sort).  "apply_resort":
 * such a kick sorts on the first of every apply_resort calls (1..1000,
 * default 10) on the same rows (pointer, particle count and row stride) and
 * the calls in between re-use that call's cell order (any order gives the
 * same kick; a recent one keeps the locality): 1 sorts on every call;
 * bit-identical results.  "rec_direct": such a kick with the default
 * deposit layout reads the particle records through the cell order in the
 * deposit and the kick (1, default) or gathers them into a cell-ordered
 * copy of the rows first (0); bit-identical results.  "deposit_tile": particles per thread and
 * LDS hash slots of the deposit, 0 (default) 8 and 2048 (FP64: 1024), 1 half
 * of both, 2 twice both, 3 no table (one global atomic per particle), 4 as 0
 * with all of a thread's particle loads issued first; bit-identical results.  "batch_sort": see
 * ocsc_apply_batched.  "fft_plan": the transform plan, 0 LDS Stockham
 * passes, radix 16 where every thread gets a butterfly (else 8), 1 radix 8
 * in every pass, 2 half as many columns per workgroup, 3 the two field
 * passes with 256 threads per workgroup (radix 16 throughout) (1 to 3: mixed
 * form only, FP64 as 0), 4 (default) register-resident transforms (one LDS
 * transpose per transform) in every pass whose length is 16 to 512 (in the
 * FP64 form not the z forward and y inverse passes of lengths 256 and 512),
 * other lengths as 0; the results differ in round-off only.  "sort_cells": sort keys of the cell-ordered copies
 * are bricks and the cells inside them (1, default) or bricks only (0: fewer
 * radix passes, cells inside a brick stay in their previous order); both give
 * the same particles up to the order of reduction sums.  "sort_ahead": 1
 * (default) takes the tracking sort keys at each particle's drifted position
 * half-way through the steps until the next re-sort (the copy then stays
 * closer to cell order on average), 0 at its present position; same
 * particles up to the order of reduction sums.  "resort_near": 1 re-sorts
 * the copy with one gather of rows, charges and permutation (the copy is
 * nearly in order), 0 (default) through particle records as the first sort
 * (measured faster); bit-identical results.  "apply_graph": 1 (default)
 * replays the launch sequence of ocsc_apply from a HIP graph captured at the
 * first call and kept while the particle and charge pointers, sizes, mesh,
 * precision, options and workspaces stay the same (energy, step and
 * random-mesh draws are updated in place), 0 launches kernel by kernel;
 * bit-identical results.  The graph holds the caller's pointers: a bunch
 * freed and another allocated at the same address with the same sizes is
 * kicked correctly, since only addresses and sizes enter the sequence.
 * "grid_clear": the charge grid is cleared by a kernel of its own on the
 * second stream (1, default; the deposit waits on that stream), by the
 * frame/extent pass B (0), or in tracking by each step's kick for the next
 * step (2; the first step of a call as 1); bit-identical results.
 * "igf_blocks": workgroups of the Green's-function node and difference
 * kernels, 0 (default) as many as their work items, 1..65536 a cap (their
 * loops stride over the rest); bit-identical results.  "fold_frame":
 * tracking steps fold the previous kick's momentum sums, set up and reduce
 * the frame in one launch (1, default) or three (0); bit-identical results.
 * "q_uniform": a tracking call checks its charges once; when every charge
 * equals the first bit for bit, the steps' extent and deposit passes take
 * that value instead of loading a charge per particle (1, default), or they
 * always load them (0); bit-identical results. */
int ocsc_set_option(ocsc_ctx* ctx, const char* name, int64_t value);
int ocsc_get_option(ocsc_ctx* ctx, const char* name, int64_t* value);

/* The cell sort of the tracking path, for diagnostics and tests (no
 * reference counterpart): order[j] = position of the j-th smallest of the n
 * device keys, comparing their low end_bit bits, ties in position order
 * (numpy.argsort(keys, kind="stable")).  Hand-written stable LSD radix sort,
 * 8 bits per pass.  keys and order are device arrays; synchronous. */
int ocsc_cell_sort(ocsc_ctx* ctx, const uint32_t* keys, int64_t n, int end_bit, uint32_t* order, void* stream);

/* Stage timing for performance work: with profiling on, every kick issued
 * outside a graph capture records HIP events between its stages and waits
 * for them (this serialises the host with the device).  ocsc_stage_times
 * writes the mean milliseconds per stage since profiling was switched on and
 * returns the number of stages; ocsc_stage_name names stage i.  No reference
 * counterpart (the reference logs the FFT wall time, ocelot/cpbd/sc.py:150-166). */
int ocsc_set_profiling(ocsc_ctx* ctx, int on);
int ocsc_stage_times(ocsc_ctx* ctx, double* ms, int nout, int64_t* calls);
const char* ocsc_stage_name(int i);

/* Kick-kernel durations of the last ocsc_track_drift run with option
 * "kick_timing" = 1: a pair of HIP events per step is recorded around the
 * kick kernel on the caller's stream, as event-record nodes inside the
 * captured graphs (the timed loop itself, not a separate profiling run).
 * Writes min(nout, nsteps) milliseconds and the step count.  No reference
 * counterpart. */
int ocsc_kick_times(ocsc_ctx* ctx, double* ms, int nout, int64_t* count);

/* Scalars of the last kick, for diagnostics and parity tests (synchronises):
 * out[0] gamma0, [1] Pav, [2..10] T (row-major, T[j][k]), [11..13] steps,
 * [14..16] X_off, [17] total charge, [18] error bits, [19] kernel rebuilt.
 * Writes min(nout, 20) values. */
int ocsc_last_stats(ocsc_ctx* ctx, double* out, int nout, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* OCSC_H */
