"""Space-charge kick throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload flattop_10m_127]

Workload (default): 10M macroparticles, 1 nC, 500 MeV, flat-top 2 mm with
0.2 mm Gaussian edges and chirp -10 1/m, mesh 127^3.  One step = a 0.1 m drift
followed by one space-charge kick with dz = 0.1 m (the kick of
SpaceCharge.apply, ocelot/cpbd/sc.py:208-251), mixed precision, the step
captured once in a HIP graph and replayed K times; the bunch is resident in
HBM before the timed region starts.  The same K steps are then timed in the
plain FP64 form, a few steps are re-run with per-stage HIP events for the
roofline of the dominant kernel, and the FP64 oracle is timed on the host
cores on one kick of a particle sample of the same bunch.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "SC macroparticle-kicks/s at 10M, 127³ mesh (mixed vs FP64); field RMS err"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="flattop_10m_127")
    ap.add_argument("--particles", type=int, default=None, help="override the particle count")
    ap.add_argument("--cpu-sample", type=int, default=10_000_000,
                    help="particles in the CPU-baseline kick (0: skip)")
    ap.add_argument("--profile-steps", type=int, default=None,
                    help="eager steps timed per stage with HIP events (default: one re-sort cycle)")
    ap.add_argument("--sort-every", type=int, default=None,
                    help="re-sort the resident bunch by mesh cell every k tracking steps (0: never); "
                         "default: chosen by the autotuner")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the context measurements of the other configurations")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the per-configuration errors against the FP64 oracle")
    ap.add_argument("--copy-calib", action="store_true",
                    help="first run one elementwise kernel reading and writing 1 GiB each, the reference "
                         "for scaling profiler byte counters")
    ap.add_argument("--deposit-xcd", default="auto", choices=["0", "1", "auto"],
                    help="charge deposit into one grid (0), per-XCD copies (1), or let the autotuner pick "
                         "after checking both give identical kicks (auto)")
    ap.add_argument("--tune-steps", type=int, default=30, help="timed steps per autotuner candidate")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the library's default options (short runs under a profiler)")
    ap.add_argument("--lsc-only", action="store_true",
                    help="only time tracking with an LSC kick after every space-charge kick (--warmup, then "
                         "--steps graph-captured steps) and two one-shot LSC kicks; one JSON line")
    ap.add_argument("--apply-only", choices=["fp64", "mixed"], default=None,
                    help="only time the drop-in SpaceCharge.apply on the resident bunch of --workload in its "
                         "caller's order (--warmup, then --steps kicks, then per-stage events over --steps more); "
                         "one JSON line")
    ap.add_argument("--batch-only", action="store_true",
                    help="only time the 64-bunch batched kicks (mixed, --warmup kicks, then one call of --steps "
                         "kicks); one JSON line")
    return ap.parse_args()


def batch_only(a, torch, dev):
    """Context run for profiling the batched kicks of 64 bunches of uneven
    size and mesh (not the headline)."""
    from ocelot_amd import DeviceParticleArray, ParticleArray, apply_batched
    from ocelot_amd.workloads import batch64
    bunches = batch64(seed=0)
    ntot = sum(b["r"].shape[1] for b in bunches)
    ds = []
    for b in bunches:
        p = ParticleArray()
        p.rparticles, p.q_array, p.E = b["r"], b["q"], b["E"]
        ds.append(DeviceParticleArray.from_host(p, dev))
    items = [(d, b["nmesh"], 0.1) for d, b in zip(ds, bunches)]
    apply_batched(items, precision="mixed", nkicks=max(a.warmup, 1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    apply_batched(items, precision="mixed", nkicks=a.steps)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(a.steps, 1)
    print(json.dumps({"metric": "batched kicks of 64 bunches", "particles": ntot, "bunches": len(bunches),
                      "kicks": a.steps, "ms_per_round": round(dt * 1e3, 4), "value": round(ntot / dt, 1),
                      "unit": "macroparticle-kicks/s", "precision": "mixed"}))


def apply_only(a, torch, host, nmesh, dz, dev):
    """Context run for profiling the per-kick drop-in path that Ocelot's
    track() drives (one SpaceCharge.apply per step, ref track.py:475-477)."""
    import ctypes
    from ocelot_amd import DeviceParticleArray, SpaceCharge, _lib
    lib, ctx = _lib.load(), _lib.context(dev)
    d = DeviceParticleArray.from_host(host, dev)
    sc = SpaceCharge(nmesh_xyz=list(nmesh), precision=a.apply_only)
    for _ in range(a.warmup):
        sc.apply(d, dz)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        sc.apply(d, dz)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(a.steps, 1)
    _lib.check(lib.ocsc_set_profiling(ctx.handle, 1))
    for _ in range(a.steps):
        sc.apply(d, dz)
    ms = (ctypes.c_double * 16)()
    calls = ctypes.c_int64()
    ns = lib.ocsc_stage_times(ctx.handle, ms, 16, ctypes.byref(calls))
    _lib.check(lib.ocsc_set_profiling(ctx.handle, 0))
    n = d.n
    opts = {k: _lib.get_option(dev, k) for k in ("apply_sort", "deposit_xcd", "deposit_tile", "overlap")}
    print(json.dumps({"metric": "SpaceCharge.apply on a resident bunch (caller's order)", "workload": a.workload,
                      "particles": n, "nmesh": list(nmesh), "precision": a.apply_only, "kicks": a.steps,
                      "warmup": a.warmup, "ms_per_kick": round(dt * 1e3, 4), "value": round(n / dt, 1),
                      "unit": "macroparticle-kicks/s", "options": opts,
                      "stages_ms": {lib.ocsc_stage_name(i).decode(): round(ms[i], 4) for i in range(ns)}}))


def lsc_only(a, torch, host, nmesh, dz, dev):
    """Context run for profiling the LSC kernels (not the headline)."""
    from ocelot_amd import LSC, DeviceParticleArray, SpaceCharge, track_drift
    d = DeviceParticleArray.from_host(host, dev)
    sc, lsc = SpaceCharge(nmesh_xyz=list(nmesh), precision="mixed"), LSC()
    track_drift(d, sc, dz, dz, a.warmup, lsc=lsc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    track_drift(d, sc, dz, dz, a.steps, lsc=lsc)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(a.steps, 1)
    t1 = time.perf_counter()
    lsc.apply(d, dz)
    lsc.apply(d, dz)
    torch.cuda.synchronize()
    da = (time.perf_counter() - t1) / 2
    n = d.n
    print(json.dumps({"metric": "SC + LSC tracking step", "workload": a.workload, "particles": n,
                      "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(dt * 1e3, 4),
                      "value": round(n / dt, 1), "unit": "macroparticle-steps/s",
                      "lsc_apply_ms": round(da * 1e3, 4), "profile_bins": lsc.last_info["nbins"]}))


def algorithmic_bytes(n, nmesh, prec):
    """Compulsory HBM bytes of one kick (particle passes A-D + pruned mesh
    passes), and of the kick kernel alone: the kick reads and writes the six
    FP64 rows (96 B per particle) and reads the three FP32 field grids (12 B
    per cell; 24 B of FP64 difference fields in the FP64 form)."""
    nx, ny, nz = nmesh
    mx, my, mz = (1 << int(np.ceil(np.log2(2 * v - 1))) for v in nmesh)
    hz = mz // 2 + 1
    c = 8 if prec == "mixed" else 16            # complex element
    f = 4 if prec == "mixed" else 8             # real element
    w = 8 if prec == "mixed" else 16            # fixed-point charge words
    part = (24 + 56 + 56 + 96) * n
    nf = 3 if prec == "mixed" else 1            # mixed: three field grids, FP64: the potential
    nc = nx * ny * nz
    mesh = (2 * w * nc                                         # zero + read charge grid
            + c * nx * ny * hz + 2 * c * nx * my * hz            # z pass write, y pass read/write
            + c * nx * my * hz + f * (mx // 2 + 1) * (my // 2 + 1) * hz  # x pass read + spectrum
            + (2 if nf == 3 else 1) * c * nx * my * hz           # x pass write (Qx, P)
            + (2 if nf == 3 else 1) * c * nx * my * hz + nf * c * nx * ny * hz  # inverse y
            + nf * c * nx * ny * hz + nf * f * nc)               # inverse z -> fields
    if prec != "mixed":
        mesh += 32 * nc                                        # k_diff64: potential in, three fields out
    kick = 96 * n + (12 if prec == "mixed" else 24) * nc
    return part + mesh, kick


def cpu_baseline(name, n_sample, steps=3, seed=0):
    """FP64 oracle tracking steps (numpy/scipy, the reference's fallback
    library calls): ``steps`` x [drift, kick] on a particle sample of the
    workload (the full bunch by default), BLAS limited to one thread.
    Returns the baseline line and the oracle's result (final rows and the
    cell indices of the first kick) for the accuracy check."""
    from oracle import sc_oracle as orc
    from ocelot_amd.workloads import make_bunch
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:  # pragma: no cover
        threadpool_limits = None
    # same distribution and total charge
    r, q, E, nmesh, dz = make_bunch(name, n=n_sample, seed=seed)
    r0 = r.copy()
    dbg = {}

    def run():
        t0 = time.perf_counter()
        for k in range(steps):
            orc.drift(r, dz, E)
            orc.apply(r, q, E, dz, nmesh, debug=dbg if k == 0 else None)
        return time.perf_counter() - t0

    if threadpool_limits is not None:
        with threadpool_limits(limits=1):
            dt = run()
    else:
        dt = run()
    import importlib.util
    import scipy
    host = dict(cpu_count=os.cpu_count(), blas_threads=1 if threadpool_limits is not None else None,
                fft="numpy.fft (the reference's fallback branch)",
                pyfftw_installed=importlib.util.find_spec("pyfftw") is not None,
                numpy=np.__version__, scipy=scipy.__version__)
    line = dict(value=n_sample * steps / dt, unit="macroparticle-kicks/s", cores=1, kind="port",
                sample=f"{name}: {steps} FP64 oracle steps (drift {dz} m + kick) of {n_sample} particles on "
                       f"{nmesh[0]}^3, {dt:.2f} s", host=host)
    return line, dict(r0=r0, q=q, E=E, nmesh=nmesh, dz=dz, steps=steps, r=r, inds=dbg["inds"])


def kick_errors(r0, got, ref):
    """Errors of a kicked bunch against the oracle's: rows (max relative to
    each row's largest value), the kick = momentum change of rows x', y',
    delta (max relative to its largest value, with the row's FP64 floor, and
    RMS relative); the kick is the lab-frame field at the particles times
    fixed factors, so its RMS error is the field's."""
    eps = np.finfo(np.float64).eps
    row = max(float(np.max(np.abs(got[i] - ref[i])) / np.max(np.abs(ref[i]))) for i in range(6))
    kmax, krms = 0.0, 0.0
    for i in (1, 3, 5):
        dg, dr = got[i] - r0[i], ref[i] - r0[i]
        floor = 16 * eps * ((1.0 if i == 5 else 0.0) + np.max(np.abs(ref[i])))
        kmax = max(kmax, float(np.max(np.abs(dg - dr)) / (np.max(np.abs(dr)) + floor / 1e-9)))
        krms = max(krms, float(np.linalg.norm(dg - dr) / np.linalg.norm(dr)))
    return {"row_max_rel": row, "kick_max_rel": kmax, "kick_rms_rel": krms}


def errors_vs_oracle(torch, dev, c3):
    """Accuracy of both forms against the FP64 oracle, per configuration:
    the 200k and 2M Gaussian bunches on 63^3 (one kick, full size), the
    10M flat-top on 127^3 (the CPU baseline's oracle steps, full size), four
    bunches of the 64-bunch set (the smallest, the largest, the smallest per
    mesh size; one batched kick) and the 50M hollow bunch on 255^3 (at full
    size mixed against FP64; against the oracle on 100k particles on meshes
    255 long on one axis, the oracle's 509^3 transform being minutes long).
    Cell mismatches: the deposit kernel's NGP cells against the oracle's."""
    from oracle import sc_oracle as orc
    from ocelot_amd import DeviceParticleArray, ParticleArray, SpaceCharge, apply_batched, drift, track_drift
    from ocelot_amd.workloads import batch64, make_bunch

    def pa(r, q, E):
        p = ParticleArray()
        p.rparticles, p.q_array, p.E = r.copy(), q, E
        return p

    def one_kick(name, n=None, nmesh=None):
        r, q, E, nm, dz = make_bunch(name, n=n)
        nm = nm if nmesh is None else nmesh
        ro = r.copy()
        dbg = {}
        orc.apply(ro, q, E, dz, nm, debug=dbg)
        out = {"particles": r.shape[1], "nmesh": list(nm)}
        for prec in ("fp64", "mixed"):
            d = DeviceParticleArray.from_host(pa(r, q, E), dev)
            cells = SpaceCharge(nmesh_xyz=list(nm), precision=prec).apply_cells(d, dz).cpu().numpy()
            out[prec] = kick_errors(r, d.rparticles.cpu().numpy(), ro)
            out[prec]["cell_mismatches"] = int(np.sum(cells != dbg["inds"]))
        return out

    res = {"C1_gauss_200k_63": one_kick("gauss_200k_63"), "C2_gauss_2m_63": one_kick("gauss_2m_63")}
    if c3 is not None:
        out = {"particles": c3["r0"].shape[1], "nmesh": list(c3["nmesh"]), "steps": c3["steps"]}
        for prec in ("fp64", "mixed"):
            sc = SpaceCharge(nmesh_xyz=list(c3["nmesh"]), precision=prec)
            d = DeviceParticleArray.from_host(pa(c3["r0"], c3["q"], c3["E"]), dev)
            track_drift(d, sc, c3["dz"], c3["dz"], c3["steps"], use_graph=True)
            out[prec] = kick_errors(c3["r0"], d.rparticles.cpu().numpy(), c3["r"])
            d = DeviceParticleArray.from_host(pa(c3["r0"], c3["q"], c3["E"]), dev)
            drift(d, c3["dz"])
            cells = sc.apply_cells(d, c3["dz"]).cpu().numpy()
            out[prec]["cell_mismatches_first_kick"] = int(np.sum(cells != c3["inds"]))
        res["C3_flattop_10m_127"] = out
    bunches = batch64(seed=0)
    sizes = [b["r"].shape[1] for b in bunches]
    picks = {int(np.argmin(sizes)), int(np.argmax(sizes))}
    for m in (31, 47, 63):
        cand = [i for i, b in enumerate(bunches) if b["nmesh"][0] == m]
        if cand:
            picks.add(min(cand, key=lambda i: sizes[i]))
    picks = sorted(picks)
    worst = {"fp64": {}, "mixed": {}}
    for prec in ("fp64", "mixed"):
        ds = [DeviceParticleArray.from_host(pa(bunches[i]["r"], bunches[i]["q"], bunches[i]["E"]), dev) for i in picks]
        apply_batched([(d, bunches[i]["nmesh"], 0.1) for d, i in zip(ds, picks)], precision=prec, nkicks=1)
        for d, i in zip(ds, picks):
            b = bunches[i]
            if prec == "fp64":
                b["ro"] = b["r"].copy()
                orc.apply(b["ro"], b["q"], b["E"], 0.1, b["nmesh"])
            e = kick_errors(b["r"], d.rparticles.cpu().numpy(), b["ro"])
            for k, v in e.items():
                worst[prec][k] = max(worst[prec].get(k, 0.0), v)
    res["C4_batch64_4_bunches"] = {"bunches": picks, "particles": [sizes[i] for i in picks], **worst}
    # C5: full size, mixed against FP64 (device), then the 255-long axes vs the oracle
    r, q, E, nm, dz = make_bunch("hollow_50m_255")
    outs = {}
    for prec in ("fp64", "mixed"):
        d = DeviceParticleArray.from_host(pa(r, q, E), dev)
        SpaceCharge(nmesh_xyz=list(nm), precision=prec).apply(d, dz)
        outs[prec] = d.rparticles
        del d
    r0 = torch.as_tensor(r, device=outs["fp64"].device)
    rms = max(float(torch.linalg.norm((outs["mixed"][i] - r0[i]) - (outs["fp64"][i] - r0[i]))
                    / torch.linalg.norm(outs["fp64"][i] - r0[i])) for i in (1, 3, 5))
    del outs, r0, r, q
    axes = {}
    for nmesh in ((255, 15, 15), (15, 255, 15), (15, 15, 255)):
        e = one_kick("hollow_50m_255", n=100_000, nmesh=nmesh)
        axes["x".join(map(str, nmesh))] = {p: e[p] for p in ("fp64", "mixed")}
    res["C5_hollow_50m_255"] = {"particles": 50_000_000, "mixed_vs_fp64_kick_rms_rel": rms,
                                "axes_255_vs_oracle_100k": axes}
    return res


def other_configs(torch, dev):
    """Context lines (not the headline): the 64-bunch batched configuration
    (uneven sizes, meshes 31/47/63, 20 kicks, mixed), single drop-in kicks
    of the 2M-particle 63^3 bunch (FP64, device-resident and host numpy), the
    headline tracking loop with per-step Twiss (track()'s tws_track), the
    LSC kick of the same resident bunch, the tracking loop with an LSC kick
    after every space-charge kick, and the CSR kick tail on the same bunch."""
    from ocelot_amd import DeviceParticleArray, ParticleArray, SpaceCharge, apply_batched
    from ocelot_amd.workloads import batch64, make_bunch
    out = {}
    bunches = batch64(seed=0)
    ntot = sum(b["r"].shape[1] for b in bunches)
    ds = []
    for b in bunches:
        p = ParticleArray()
        p.rparticles, p.q_array, p.E = b["r"], b["q"], b["E"]
        ds.append(DeviceParticleArray.from_host(p, dev))
    items = [(d, b["nmesh"], 0.1) for d, b in zip(ds, bunches)]
    apply_batched(items, precision="mixed", nkicks=2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    apply_batched(items, precision="mixed", nkicks=20)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 20
    out["batch64"] = {"particles": ntot, "bunches": len(bunches), "kicks": 20, "ms_per_round": round(dt * 1e3, 3),
                      "value": round(ntot / dt, 1), "unit": "macroparticle-kicks/s", "precision": "mixed",
                      "hip_graph": True}
    del ds, items
    r, q, E, nmesh, dz = make_bunch("gauss_2m_63")
    host = ParticleArray()
    host.rparticles, host.q_array, host.E = r, q, E
    sc = SpaceCharge(nmesh_xyz=list(nmesh))
    d = DeviceParticleArray.from_host(host, dev)
    sc.apply(d, dz)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        sc.apply(d, dz)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 10
    hp = host.copy()
    t1 = time.perf_counter()
    for _ in range(3):
        sc.apply(hp, dz)
    dth = (time.perf_counter() - t1) / 3
    out["gauss_2m_63_fp64_apply"] = {"particles": r.shape[1], "ms_per_kick_device": round(dt * 1e3, 3),
                                     "value_device": round(r.shape[1] / dt, 1),
                                     "ms_per_kick_host_numpy": round(dth * 1e3, 3),
                                     "value_host_numpy_incl_pcie": round(r.shape[1] / dth, 1),
                                     "unit": "macroparticle-kicks/s", "precision": "fp64", "order": "caller's"}
    del d, hp
    # the headline loop with track()'s per-step Twiss (moments on the device)
    from ocelot_amd import track_drift
    r, q, E, nmesh, dz = make_bunch("flattop_10m_127")
    host = ParticleArray()
    host.rparticles, host.q_array, host.E = r, q, E
    d = DeviceParticleArray.from_host(host, dev)
    sc = SpaceCharge(nmesh_xyz=list(nmesh), precision="mixed")
    track_drift(d, sc, dz, dz, 3, calc_tws=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tws = track_drift(d, sc, dz, dz, 20, calc_tws=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 20
    out["flattop_10m_127_track_with_twiss"] = {"particles": r.shape[1], "steps": 20, "ms_per_step": round(dt * 1e3, 4),
                                               "value": round(r.shape[1] / dt, 1), "unit": "macroparticle-kicks/s",
                                               "twiss_rows": len(tws), "precision": "mixed"}
    # longitudinal space charge (LSC.apply) on the same resident bunch
    from ocelot_amd import LSC
    lsc = LSC()
    lsc.apply(d, dz)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(20):
        lsc.apply(d, dz)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 20
    out["flattop_10m_lsc_apply"] = {"particles": r.shape[1], "kicks": 20, "ms_per_kick": round(dt * 1e3, 4),
                                    "value": round(r.shape[1] / dt, 1), "unit": "macroparticle-kicks/s",
                                    "profile_bins": lsc.last_info["nbins"], "precision": "fp64"}
    # the headline loop with an LSC kick after every space-charge kick, one graph per step
    track_drift(d, sc, dz, dz, 3, lsc=lsc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    track_drift(d, sc, dz, dz, 20, lsc=lsc)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 20
    out["flattop_10m_127_track_sc_lsc"] = {"particles": r.shape[1], "steps": 20, "ms_per_step": round(dt * 1e3, 4),
                                           "value": round(r.shape[1] / dt, 1), "unit": "macroparticle-steps/s",
                                           "precision": "mixed SC + fp64 LSC", "hip_graph": True}
    # the CSR kick tail (convolution of a 1001-point line charge with a
    # 1001-point kernel, interpolated energy kick of every particle) on the
    # same resident bunch; per call, including the small host copies
    from ocelot_amd.csr import csr_kick
    ns = 1000
    s1, s2 = -1.6e-3, 1.6e-3
    zg = np.linspace(s1, s2, ns + 1)
    lam = np.exp(-0.5 * (zg / 6e-4) ** 2)
    lam /= lam.sum() * (s2 - s1) / ns
    K1 = -1e3 / (1.0 + np.arange(ns + 1))
    csr_kick(d, lam, K1, s1, s2, ns, dz)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(20):
        csr_kick(d, lam, K1, s1, s2, ns, dz)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 20
    out["flattop_10m_csr_kick"] = {"particles": r.shape[1], "kicks": 20, "mesh": ns + 1, "ms_per_kick": round(dt * 1e3, 4),
                                   "value": round(r.shape[1] / dt, 1), "unit": "macroparticle-kicks/s",
                                   "kick_GBs": round(24 * r.shape[1] / dt / 1e9, 1), "precision": "fp64"}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; running on several "
                                   "(particles sharded by index, one all-reduce of the charge grid per "
                                   "kick) is designed but not built", "n_gpus": a.gpus}))
        sys.exit(2)
    import torch
    from ocelot_amd import DeviceParticleArray, ParticleArray, SpaceCharge, track_drift
    from ocelot_amd import _lib
    from ocelot_amd.workloads import make_bunch

    if a.copy_calib:
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.add(x, 1.0, out=y)
        torch.cuda.synchronize()
        del x, y
    if a.batch_only:
        return batch_only(a, torch, torch.cuda.current_device())
    r, q, E, nmesh, dz = make_bunch(a.workload, n=a.particles)
    n = r.shape[1]
    host = ParticleArray()
    host.rparticles, host.q_array, host.E = r, q, E
    dev = torch.cuda.current_device()
    L = dz
    if a.lsc_only:
        return lsc_only(a, torch, host, nmesh, dz, dev)
    if a.apply_only:
        return apply_only(a, torch, host, nmesh, dz, dev)

    # the library's autotuner picks the speed-only options (re-sort interval,
    # deposit mode, Green's-function overlap) for this bunch and mesh; options
    # given on the command line are held fixed
    from ocelot_amd.autotune import DEFAULT_SPACE, tune_tracking
    space = dict(DEFAULT_SPACE)
    if a.sort_every is not None:
        space.pop("sort_every")
        _lib.set_option(dev, "sort_every", a.sort_every)
    if a.deposit_xcd != "auto":
        space.pop("deposit_xcd")
        _lib.set_option(dev, "deposit_xcd", int(a.deposit_xcd))
    if a.no_tune:
        space = {}
    d_tune = DeviceParticleArray.from_host(host, dev)
    tune = tune_tracking(d_tune, SpaceCharge(nmesh_xyz=list(nmesh), precision="mixed"), L, dz,
                         steps=a.tune_steps, space=space) if space else {"options": {}}
    del d_tune
    opts = {k: _lib.get_option(dev, k) for k in DEFAULT_SPACE}
    sort_every = opts["sort_every"]
    print(f"# options {opts} (autotune: {tune.get('timings_ms')})", file=sys.stderr, flush=True)

    import ctypes
    lib = _lib.load()
    ctx = _lib.context(dev)

    def run(prec, steps, warmup):
        """Timed loop; the kick kernel of every step is also timed with HIP
        events captured into the step graphs (kick_timing), on its stream."""
        sc = SpaceCharge(nmesh_xyz=list(nmesh), precision=prec)
        d = DeviceParticleArray.from_host(host, dev)
        # the warm-up call runs the timed call's exact launch sequence (kick
        # timing included), so the timed call replays its step graphs as any
        # repeated call on the same bunch does
        _lib.set_option(dev, "kick_timing", 1)
        try:
            track_drift(d, sc, L, dz, warmup, use_graph=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            track_drift(d, sc, L, dz, steps, use_graph=True)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
        finally:
            _lib.set_option(dev, "kick_timing", 0)
        ms = (ctypes.c_double * steps)()
        cnt = ctypes.c_int64()
        _lib.check(lib.ocsc_kick_times(ctx.handle, ms, steps, ctypes.byref(cnt)))
        # every step but the last runs the same kick kernel; the last one also
        # writes the particles back to the caller's rows (another variant)
        k = list(ms[:cnt.value])
        kick_ms = float(np.mean(k[:-1] if len(k) > 1 else k)) if k else None
        last_ms = float(k[-1]) if len(k) > 1 else None
        return (t1 - t0) / steps, (kick_ms, last_ms)

    print(f"# workload {a.workload}: {n} particles, mesh {nmesh}", file=sys.stderr, flush=True)
    t_mix, (kick_ms_live, last_kick_ms) = run("mixed", a.steps, a.warmup)
    print(f"# mixed: {t_mix * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
    t_64 = None
    if not a.no_fp64:
        t_64, _k64 = run("fp64", a.steps, a.warmup)
        print(f"# fp64: {t_64 * 1e3:.3f} ms/step", file=sys.stderr, flush=True)

    # field error of one kick, mixed vs FP64, at full size (kick deltas are the
    # lab-frame field times fixed factors)
    outs = {}
    for prec in ("fp64", "mixed"):
        d = DeviceParticleArray.from_host(host, dev)
        SpaceCharge(nmesh_xyz=list(nmesh), precision=prec).apply(d, dz)
        outs[prec] = d.rparticles
    r0 = torch.as_tensor(r, device=outs["fp64"].device)
    errs, errs_max = [], []
    for i in (1, 3, 5):
        ref = outs["fp64"][i] - r0[i]
        got = outs["mixed"][i] - r0[i]
        errs.append(float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref)))
        # max relative error over particles whose kick exceeds 1e-3 of the largest
        big = ref.abs() > 1e-3 * ref.abs().max()
        errs_max.append(float(((got - ref).abs()[big] / ref.abs()[big]).max()))
    field_rms = max(errs)
    field_max = max(errs_max)
    del outs, r0

    # per-stage HIP events over one re-sort cycle of eager tracking steps (same
    # kicks as the timed region, same stream; profiling switches the graph off)
    stages = {}
    prof = {}
    for prec in ("mixed", "fp64"):
        sc = SpaceCharge(nmesh_xyz=list(nmesh), precision=prec)
        d = DeviceParticleArray.from_host(host, dev)
        _lib.check(lib.ocsc_set_profiling(ctx.handle, 1))
        # a whole re-sort cycle, so the per-stage means match the timed loop's mix
        track_drift(d, sc, L, dz, a.profile_steps or max(sort_every, 1) + 1, use_graph=True)
        ms = (ctypes.c_double * 16)()
        calls = ctypes.c_int64()
        ns = lib.ocsc_stage_times(ctx.handle, ms, 16, ctypes.byref(calls))
        _lib.check(lib.ocsc_set_profiling(ctx.handle, 0))
        stages[prec] = {lib.ocsc_stage_name(i).decode(): round(ms[i], 4) for i in range(ns)}
        prof[prec] = [ms[i] for i in range(ns)]
    total_b, kick_b = algorithmic_bytes(n, nmesh, "mixed")
    total_b64, _ = algorithmic_bytes(n, nmesh, "fp64")
    # algorithmic GB/s of the particle stages in the same eager cycle: extent
    # (pass B) and deposit read six rows and q (56 B per particle), the kick
    # reads and writes the rows and reads the field grids
    nc = int(np.prod(nmesh))
    stage_bytes = {"mixed": {"extent": 56 * n, "deposit": 56 * n, "kick": 96 * n + 12 * nc},
                   "fp64": {"extent": 56 * n, "deposit": 56 * n, "kick": 96 * n + 24 * nc}}
    stages_gbs = {p: {k: round(b / (stages[p][k] * 1e-3) / 1e9, 1) for k, b in stage_bytes[p].items()
                      if stages[p].get(k)} for p in stages}
    kick_ms = kick_ms_live
    dom = max(stages["mixed"], key=lambda k: stages["mixed"][k])
    achieved = kick_b / (kick_ms * 1e-3) / 1e9
    traffic, traffic_src = None, None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         f"k_kick_{a.workload}_traffic.json")
    if os.path.exists(tpath) and a.particles is None:
        with open(tpath) as fh:
            tr = json.load(fh)
        traffic = tr["traffic_bytes"]
        traffic_src = (f"profiles/{os.path.basename(tpath)}: rocprofv3 FETCH_SIZE x{tr['calibration']['fetch_scale']}"
                       f" + WRITE_SIZE x{tr['calibration']['write_scale']} per dispatch (copy-kernel calibration)")
    roofline = {"kernel": "k_kick<float> (gather + kick + write-back)", "bound": "hbm",
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                "kernel_ms": round(kick_ms, 4),
                "kernel_ms_source": "HIP events around k_kick captured into the step graphs of the timed loop "
                                    "(mean over its steps but the last, on the kernel's stream)",
                "last_step_kick_ms": None if last_kick_ms is None else round(last_kick_ms, 4),
                "algorithmic_bytes": kick_b, "largest_stage": dom,
                "kick_total_bytes": total_b,
                "kick_total_GBs": round(total_b / (sum(prof["mixed"]) * 1e-3) / 1e9, 1),
                "kick_total_bytes_fp64": total_b64,
                "kick_total_GBs_fp64": round(total_b64 / (sum(prof["fp64"]) * 1e-3) / 1e9, 1)}

    extra = None if a.no_extra else other_configs(torch, dev)

    cpu, c3 = None, None
    if a.cpu_sample > 0:
        cpu, c3 = cpu_baseline(a.workload, min(a.cpu_sample, n))
    acc = None if a.no_accuracy else errors_vs_oracle(torch, dev, c3)

    value = n / t_mix
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "macroparticle-kicks/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(t_mix * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 mesh/FFT/gather + fp64 frame/Green/momentum",
        "data": "synthetic (seeded flat-top bunch, numpy default_rng)",
        "config": {"workload": a.workload, "particles": n, "nmesh": list(nmesh), "kick_dz_m": dz,
                   "drift_m": L, "precision": "mixed", "hip_graph": True, "options": opts},
        "autotune": tune,
        "fp64": None if t_64 is None else {"value": round(n / t_64, 1), "ms_per_step": round(t_64 * 1e3, 4)},
        "mixed_over_fp64": None if t_64 is None else round(t_64 / t_mix, 3),
        "field_rms_rel_err": field_rms,
        "field_max_rel_err": field_max,
        "stages_ms": stages,
        "stages_GBs": stages_gbs,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "errors_vs_oracle": acc,
        "other_configs": extra,
        "gpu_over_cpu": None if cpu is None else round(value / cpu["value"], 1),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
