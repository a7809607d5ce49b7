"""The 127^3 mesh of the 10M-particle configuration (FFT length 256 on every
axis) against the FP64 oracle, and the NGP cell indices of the deposit kernel
(the indices the hot path actually uses) against the oracle's, bit for bit.

Tolerances as in test_gpu_parity.py: FP64 rows within 1e-12 of each row's
largest value and the kick within 1e-9 of its size (plus the row's FP64
floor); mixed kicks within 1e-5 RMS relative; cell indices bit-exact.  The
flat-top bunch at 500 MeV has rest-frame cells about a thousand times longer
in z than across, where the FP64 Green's function cancels: the reference's
own FP64 kick is then 5e-9 to 1.3e-8 away from the same algorithm with the
kernel in extended precision.  There the FP64 result must lie within three
times the reference's own deviation from that extended-precision answer
(as for the 255-long axes in test_gpu_parity.py).
"""
import numpy as np
import pytest

from oracle import envelope_oracle as eo
from oracle import sc_oracle as orc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from test_gpu_parity import (APPLY_CASES, accuracy_vs_extended, case, kick_err, parray, rel_max,  # noqa: E402
                             rel_rms, within_reference_accuracy)


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from ocelot_amd import _lib
    _lib.load()


def oracle_kick(r, q, E, dz, nmesh):
    ro = r.copy()
    dbg = {}
    orc.apply(ro, q, E, dz, nmesh, debug=dbg)
    return ro, dbg["inds"]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("name", APPLY_CASES)
def test_hot_path_cells_golden_cases(golden, name, prec):
    """The cell each particle's charge went to in the deposit kernel (FP64:
    the reference's transform sequence; mixed: the cancelled-factor
    transforms and reciprocal cell sizes) equals the oracle's floor(X)+1."""
    from ocelot_amd import SpaceCharge
    g = case(golden, name)
    r, q, E, dz, nmesh = g["r_in"], g["q"], float(g["E"]), float(g["zstep"]), tuple(g["nmesh"])
    _, inds = oracle_kick(r, q, E, dz, nmesh)
    p = parray(r, q, E)
    cells = SpaceCharge(nmesh_xyz=list(nmesh), precision=prec).apply_cells(p, dz)
    mism = int(np.sum(cells != inds))
    print(f"{name} [{prec}]: {mism} cell mismatches of {len(inds)}")
    assert cells.dtype == np.int32 and mism == 0


@pytest.mark.parametrize("nmesh", [(127, 15, 15), (15, 127, 15), (15, 15, 127)])
def test_m256_axes_vs_oracle(nmesh):
    """FFT length 256 (meshes of 65-128 cells, the 127^3 configuration) on
    each axis against the oracle, with the flat-top bunch: FP64 rows and
    kick, mixed kick RMS, and both forms' cells."""
    from ocelot_amd import SpaceCharge
    from ocelot_amd.workloads import make_bunch
    r, q, E, _, dz = make_bunch("flattop_10m_127", n=200_000, seed=17)
    ro, inds = oracle_kick(r, q, E, dz, nmesh)
    p = parray(r, q, E)
    c64 = SpaceCharge(nmesh_xyz=list(nmesh)).apply_cells(p, dz)
    pm = parray(r, q, E)
    cm = SpaceCharge(nmesh_xyz=list(nmesh), precision="mixed").apply_cells(pm, dz)
    row = max(rel_max(p.rparticles[i], ro[i]) for i in range(6))
    kick = max(kick_err(p.rparticles, r, ro, i) for i in (1, 3, 5))
    kick_m = max(rel_rms(pm.rparticles[i] - r[i], ro[i] - r[i]) for i in (1, 3, 5))
    g, o = accuracy_vs_extended(p.rparticles, r, q, E, dz, nmesh, ro)
    print(f"mesh {nmesh}: fp64 vs oracle row {row:.3e} kick {kick:.3e}; vs extended-kernel oracle: hip row "
          f"{g[0]:.3e} kick {g[1]:.3e}, reference row {o[0]:.3e} kick {o[1]:.3e}; mixed kick rms {kick_m:.3e}; "
          f"cell mismatches {int(np.sum(c64 != inds))} / {int(np.sum(cm != inds))}")
    assert within_reference_accuracy(g, o) and kick_m <= 1e-5
    assert np.array_equal(c64, inds) and np.array_equal(cm, inds)


def test_c3_127_one_kick_vs_oracle():
    """One kick of a 1M-particle flat-top bunch (the 10M configuration's
    distribution, 1 nC, 500 MeV) on 127^3 in both forms, device-resident,
    against the oracle's 253^3 transform."""
    from ocelot_amd import DeviceParticleArray, SpaceCharge
    from ocelot_amd.workloads import make_bunch
    r, q, E, nmesh, dz = make_bunch("flattop_10m_127", n=1_000_000, seed=21)
    ro, inds = oracle_kick(r, q, E, dz, nmesh)
    out, cells = {}, {}
    for prec in ("fp64", "mixed"):
        d = DeviceParticleArray.from_host(parray(r, q, E))
        cells[prec] = SpaceCharge(nmesh_xyz=list(nmesh), precision=prec).apply_cells(d, dz).cpu().numpy()
        out[prec] = d.rparticles.cpu().numpy()
    row = max(rel_max(out["fp64"][i], ro[i]) for i in range(6))
    kick = max(kick_err(out["fp64"], r, ro, i) for i in (1, 3, 5))
    kick_m = max(rel_rms(out["mixed"][i] - r[i], ro[i] - r[i]) for i in (1, 3, 5))
    mism = {p: int(np.sum(cells[p] != inds)) for p in cells}
    g, o = accuracy_vs_extended(out["fp64"], r, q, E, dz, nmesh, ro)
    print(f"1M/127^3: fp64 vs oracle row {row:.3e} kick {kick:.3e}; vs extended-kernel oracle: hip row {g[0]:.3e} "
          f"kick {g[1]:.3e}, reference row {o[0]:.3e} kick {o[1]:.3e}; mixed kick rms {kick_m:.3e}; "
          f"cell mismatches {mism}")
    assert within_reference_accuracy(g, o) and kick_m <= 1e-5
    assert mism == {"fp64": 0, "mixed": 0}


def test_c3_mixed_tracking_vs_oracle():
    """Mixed-precision tracking on 127^3, graph-captured, re-sorting the
    resident copy every 3 steps: 10 steps of drift 0.1 m + kick against the
    oracle's own drift + apply loop.  The accumulated momentum change of
    every row within 1e-5 RMS relative, and the envelope of every step
    (emittances, beam sizes) within 1e-6 of the oracle's."""
    from ocelot_amd import DeviceParticleArray, SpaceCharge, _lib, track_drift
    from ocelot_amd.workloads import make_bunch
    r, q, E, nmesh, dz = make_bunch("flattop_10m_127", n=1_000_000, seed=22)
    nsteps = 10
    ro = r.copy()
    env_o = []
    for _ in range(nsteps):
        orc.drift(ro, dz, E)
        orc.apply(ro, q, E, dz, nmesh)
        env_o.append(eo.get_envelope(ro, q, E))
    dev = torch.cuda.current_device()
    old = _lib.get_option(dev, "sort_every")
    try:
        _lib.set_option(dev, "sort_every", 3)
        d = DeviceParticleArray.from_host(parray(r, q, E))
        tws = track_drift(d, SpaceCharge(nmesh_xyz=list(nmesh), precision="mixed"), dz, dz, nsteps,
                          use_graph=True, calc_tws=True)
    finally:
        _lib.set_option(dev, "sort_every", old)
    got = d.rparticles.cpu().numpy()
    kick = max(rel_rms(got[i] - r[i], ro[i] - r[i]) for i in (1, 3, 5))
    row = max(rel_max(got[i], ro[i]) for i in range(6))
    env_err = max(abs(getattr(tws[k + 1], a) - env_o[k][a]) / abs(env_o[k][a])
                  for k in range(nsteps) for a in ("emit_x", "emit_y", "xx", "yy", "tautau", "pp"))
    print(f"mixed tracking 10 x 127^3 (sort every 3): accumulated kick rms {kick:.3e}, row max {row:.3e}, "
          f"envelope max rel {env_err:.3e}")
    assert kick <= 1e-5 and env_err <= 1e-6
    assert abs(d.s - nsteps * dz) <= 1e-12


def test_c3_full_size_mixed_vs_fp64_tracking():
    """The 10M-particle 127^3 configuration at full size: 20 graph-captured
    steps in both forms; the mixed form's accumulated momentum change within
    1e-5 RMS of the FP64 form's, both bunches finite, the mixed run
    reproducible bit for bit."""
    from ocelot_amd import DeviceParticleArray, SpaceCharge, track_drift
    from ocelot_amd.workloads import make_bunch
    r, q, E, nmesh, dz = make_bunch("flattop_10m_127")
    host = parray(r, q, E)
    outs = []
    for prec in ("fp64", "mixed", "mixed"):
        d = DeviceParticleArray.from_host(host)
        track_drift(d, SpaceCharge(nmesh_xyz=list(nmesh), precision=prec), dz, dz, 20, use_graph=True)
        outs.append(d.rparticles)
        del d
    assert torch.equal(outs[1], outs[2]), "mixed tracking not reproducible"
    r0 = torch.as_tensor(r, device=outs[0].device)
    for i in (1, 3, 5):
        ref = outs[0][i] - r0[i]
        got = outs[1][i] - r0[i]
        e = float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref))
        print(f"10M/127^3 20 steps row {i}: mixed vs fp64 accumulated kick rms {e:.3e}")
        assert e <= 1e-5
    assert bool(torch.isfinite(outs[0]).all()) and bool(torch.isfinite(outs[1]).all())
