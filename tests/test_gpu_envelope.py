"""get_envelope on the GPU (ocsc_moments through the C ABI) against the
reference's own output (golden vectors) and the oracle.  Tolerance: 1e-9
relative to each quantity's natural scale (sums are taken in a different
order than numpy's pairwise mean, so results differ in the last bits)."""
import os

import numpy as np
import pytest

from oracle import envelope_oracle as eo

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


@pytest.fixture(scope="module")
def env_golden():
    with np.load(os.path.join(HERE, "golden", "envelope_golden.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def parray(r, q, E):
    from ocelot_amd import ParticleArray
    p = ParticleArray()
    p.rparticles = np.array(r, dtype=np.float64, copy=True)
    p.q_array = np.array(q, dtype=np.float64, copy=True)
    p.E = float(E)
    return p


SCALE = {"x": "xx", "y": "yy", "px": "pxpx", "py": "pypy", "tau": "tautau", "p": "pp",
         "xy": ("xx", "yy"), "xpx": ("xx", "pxpx"), "ypy": ("yy", "pypy"), "pxpy": ("pxpx", "pypy"),
         "xpy": ("xx", "pypy"), "ypx": ("yy", "pxpx")}


def close(a, got, ref_d):
    """|got - ref| within 1e-9 of the quantity's scale."""
    ref = ref_d[a]
    sc = abs(ref)
    s = SCALE.get(a)
    if isinstance(s, str):
        sc = max(sc, np.sqrt(abs(ref_d[s])))
    elif isinstance(s, tuple):
        sc = max(sc, np.sqrt(abs(ref_d[s[0]] * ref_d[s[1]])))
    return abs(got - ref) <= 1e-9 * sc + 1e-300


def kwargs_for(g, setting):
    from ocelot_amd import Twiss
    if setting == "bounds":
        return {"bounds": list(g["bounds/b"])}
    if setting == "autodisp":
        return {"auto_disp": True}
    if setting == "twsi":
        t = Twiss()
        t.Dx, t.Dxp, t.Dy, t.Dyp = (float(v) for v in g["disp/tws_i"])
        return {"tws_i": t}
    if setting == "imax":
        return {"bounds": list(g["imax/b"]), "slice": "Imax"}
    return {}


@pytest.mark.parametrize("bunch", ["gauss", "flat"])
@pytest.mark.parametrize("setting", ["plain", "bounds", "autodisp", "twsi", "imax"])
def test_envelope_golden(env_golden, bunch, setting):
    from ocelot_amd import get_envelope
    g = env_golden
    attrs = [str(a) for a in g["meta/attrs"]]
    ref = dict(zip(attrs, g[f"{bunch}/{setting}"]))
    p = parray(g[f"{bunch}/r"], g[f"{bunch}/q"], g[f"{bunch}/E"])
    tws = get_envelope(p, **kwargs_for(g, setting))
    bad = [(a, getattr(tws, a), ref[a]) for a in attrs if not close(a, float(getattr(tws, a)), ref)]
    assert not bad, bad


def test_envelope_two_particles(env_golden):
    from ocelot_amd import get_envelope
    g = env_golden
    ref = dict(zip([str(a) for a in g["meta/attrs"]], g["two/plain"]))
    tws = get_envelope(parray(g["two/r"], g["two/q"], g["two/E"]))
    assert tws.q == pytest.approx(ref["q"], rel=1e-15) and tws.p == pytest.approx(ref["p"], rel=1e-14)
    assert tws.xx == 0.0 and tws.emit_x == 0.0


def test_envelope_imax_10m_vs_oracle():
    """slice='Imax' on the 10M flat-top bunch (two current horns): the window
    centre is the oracle's current peak and the selected moments agree."""
    from ocelot_amd import DeviceParticleArray, get_envelope
    from ocelot_amd.envelope import moments
    from ocelot_amd.workloads import make_bunch
    r, q, E, _, _ = make_bunch("hollow_50m_255", n=2_000_000, seed=3)
    ref = eo.get_envelope(r, q, E, bounds=[-0.3, 0.3], slice="Imax")
    d = DeviceParticleArray.from_host(parray(r, q, E), torch.cuda.current_device())
    v = moments(d, bounds=[-0.3, 0.3], imax=True)
    print(f"Imax window centre: device {v[26]:.9e}, oracle {ref['z0']:.9e}; selected {int(v[0])} vs {ref['count']}")
    assert v[26] == pytest.approx(ref["z0"], abs=1e-15) and int(v[0]) == ref["count"]
    tws = get_envelope(d, bounds=[-0.3, 0.3], slice="Imax")
    bad = [a for a in ("p", "x", "y", "px", "py", "tau", "xx", "pxpx", "yy", "pypy", "tautau", "emit_x", "beta_x")
           if not close(a, float(getattr(tws, a)), ref)]
    assert not bad, [(a, getattr(tws, a), ref[a]) for a in bad]


def test_envelope_10m_device_resident_vs_oracle():
    """The 10M-particle flat-top bunch of the tracking benchmark, resident in
    HBM: moments vs the oracle, and the device path equals the host path."""
    from ocelot_amd import DeviceParticleArray, get_envelope
    from ocelot_amd.workloads import make_bunch
    r, q, E, _, _ = make_bunch("flattop_10m_127")
    ref = eo.get_envelope(r, q, E, auto_disp=True)
    host = parray(r, q, E)
    d = DeviceParticleArray.from_host(host, torch.cuda.current_device())
    torch.cuda.synchronize()
    import time
    t0 = time.perf_counter()
    tws = get_envelope(d, auto_disp=True)
    dt = time.perf_counter() - t0
    print(f"get_envelope 10M on device: {dt * 1e3:.2f} ms")
    bad = [a for a in ("p", "Dx", "Dxp", "Dy", "Dyp", "x", "y", "px", "py", "tau", "xx", "xpx", "pxpx", "yy",
                       "ypy", "pypy", "tautau", "xy", "pxpy", "xpy", "ypx", "pp", "emit_x", "emit_y",
                       "eigemit_1", "eigemit_2", "beta_x", "alpha_x")
           if not close(a, float(getattr(tws, a)), ref)]
    assert not bad, [(a, getattr(tws, a), ref[a]) for a in bad]
    th = get_envelope(host, auto_disp=True)
    for a in ("xx", "pxpx", "emit_x", "Dx", "tau"):
        assert getattr(th, a) == getattr(tws, a)


@pytest.mark.parametrize("graph", [False, True])
def test_track_drift_with_twiss_list(graph):
    """track_drift(calc_tws=True) gives the tws_track list of track() on a
    drift lattice: the moments before the first step and after each step
    (kick done, next drift not yet), against the oracle's own loop; the
    particles end exactly as without the moments."""
    from oracle import sc_oracle as orc
    from ocelot_amd import DeviceParticleArray, SpaceCharge, track_drift
    from ocelot_amd.workloads import make_bunch
    r, q, E, _, dz = make_bunch("flattop_10m_127", n=200_000, seed=21)
    nmesh, nsteps = [31, 31, 63], 12
    dev = torch.cuda.current_device()
    sc = SpaceCharge(nmesh_xyz=nmesh, precision="fp64")
    d0 = DeviceParticleArray.from_host(parray(r, q, E), dev)
    track_drift(d0, sc, dz, dz, nsteps, use_graph=graph)
    d = DeviceParticleArray.from_host(parray(r, q, E), dev)
    tws = track_drift(d, sc, dz, dz, nsteps, use_graph=graph, calc_tws=True, auto_disp=True)
    assert torch.equal(d.rparticles, d0.rparticles)
    assert len(tws) == nsteps + 1
    ro = r.copy()
    ref = [eo.get_envelope(ro, q, E, auto_disp=True)]
    for _ in range(nsteps):
        orc.drift(ro, dz, E)
        orc.apply(ro, q, E, dz, nmesh)
        ref.append(eo.get_envelope(ro, q, E, auto_disp=True))
    worst = 0.0
    for k, (t, rf) in enumerate(zip(tws, ref)):
        for a in ("emit_x", "emit_y", "beta_x", "alpha_x", "xx", "pxpx", "tautau", "pp", "Dx", "x", "px"):
            scale = max(abs(rf[a]), np.sqrt(abs(rf[SCALE[a]])) if isinstance(SCALE.get(a), str) else 0)
            worst = max(worst, abs(float(getattr(t, a)) - rf[a]) / scale)
        assert t.s == pytest.approx(k * dz, rel=1e-12)
    print(f"track_drift tws ({'graph' if graph else 'eager'}): worst relative deviation {worst:.2e} over {nsteps} steps")
    assert worst <= 1e-8


@pytest.mark.parametrize("graph", [False, True])
def test_track_drift_twiss_with_imax_window(graph):
    """track()'s Twiss list with bounds and slice='Imax': each step's window is
    centred on the peak of its current profile, planned and found on the
    device inside the step graphs; against the oracle's own loop, and the last
    row against the one-shot get_envelope of the final bunch."""
    from oracle import sc_oracle as orc
    from ocelot_amd import DeviceParticleArray, SpaceCharge, get_envelope, track_drift
    from ocelot_amd.workloads import make_bunch
    print(f"[imax tracking] graph={graph}", flush=True)
    r, q, E, _, dz = make_bunch("flattop_10m_127", n=200_000, seed=23)
    nmesh, nsteps, bounds = [31, 31, 63], 6, [-0.5, 0.5]
    dev = torch.cuda.current_device()
    sc = SpaceCharge(nmesh_xyz=nmesh, precision="fp64")
    d = DeviceParticleArray.from_host(parray(r, q, E), dev)
    tws = track_drift(d, sc, dz, dz, nsteps, use_graph=graph, calc_tws=True, bounds=bounds, slice="Imax")
    ro = r.copy()
    ref = [eo.get_envelope(ro, q, E, bounds=bounds, slice="Imax")]
    for _ in range(nsteps):
        orc.drift(ro, dz, E)
        orc.apply(ro, q, E, dz, nmesh)
        ref.append(eo.get_envelope(ro, q, E, bounds=bounds, slice="Imax"))
    worst = 0.0
    for t, rf in zip(tws, ref):
        for a in ("emit_x", "emit_y", "beta_x", "alpha_x", "xx", "pxpx", "tautau", "pp", "x", "px"):
            scale = max(abs(rf[a]), np.sqrt(abs(rf[SCALE[a]])) if isinstance(SCALE.get(a), str) else 0)
            worst = max(worst, abs(float(getattr(t, a)) - rf[a]) / scale)
    print(f"imax tws ({'graph' if graph else 'eager'}): worst relative deviation {worst:.2e}")
    assert worst <= 1e-8
    last = get_envelope(d, bounds=bounds, slice="Imax")
    for a in ("emit_x", "beta_x", "xx", "tautau", "pp"):
        assert float(getattr(tws[-1], a)) == pytest.approx(float(getattr(last, a)), rel=1e-12)


def test_npz_bunch_loads_into_hbm(tmp_path):
    """A bunch file the reference wrote, loaded straight into HBM, equals the
    host load; saving the device bunch gives the same file contents."""
    from ocelot_amd import io as oio
    ref = os.path.join(HERE, "golden", "ref_bunch_track_wo_lsc.npz")
    h = oio.load_particle_array(ref)
    d = oio.load_particle_array(ref, device=0)
    assert d.rparticles.is_cuda
    np.testing.assert_array_equal(d.rparticles.cpu().numpy(), h.rparticles)
    np.testing.assert_array_equal(d.q_array.cpu().numpy(), h.q_array)
    assert d.E == float(h.E) and d.s == float(h.s)
    f = str(tmp_path / "d.npz")
    oio.save_particle_array(f, d)
    back = oio.load_particle_array(f)
    np.testing.assert_array_equal(back.rparticles, h.rparticles)
