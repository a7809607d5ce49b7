"""LSC on the GPU (ocsc_lsc_* through the C ABI) against the reference's own
output (golden vectors written by tests/golden/make_golden_lsc.py) and the
oracle (oracle/lsc_oracle.py) on larger bunches.

Tolerances (relative to the largest value of each array): transverse size
1e-13; current profile 1e-12 (sums taken in another order than numpy's);
wake and kick 1e-10.  The wake bound covers the reference's own rounding:
its step-profile impedance evaluates 1 - x K1(x) by subtraction, which loses
up to ~1e-5 relative at the lowest frequencies (a change of ~6e-12 in these
wakes), while the device evaluates it from a series without cancellation;
and the device transforms are direct sums instead of numpy's FFT."""
import os

import numpy as np
import pytest

from oracle import lsc_oracle as lo

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = [("gauss", False), ("gauss", True), ("horns", False), ("horns", True)]


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


@pytest.fixture(scope="module")
def lg():
    with np.load(os.path.join(HERE, "golden", "lsc_golden.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def parray(r, q, E):
    from ocelot_amd import ParticleArray
    p = ParticleArray()
    p.rparticles = np.array(r, dtype=np.float64, copy=True)
    p.q_array = np.array(q, dtype=np.float64, copy=True)
    p.E = float(E)
    return p


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


@pytest.mark.parametrize("case,step", CASES)
def test_lsc_matches_reference_golden(lg, case, step):
    from ocelot_amd import LSC
    r, q, E = lg[f"{case}/r"], lg[f"{case}/q"], float(lg[f"{case}/E"])
    t = f"{case}/{'step' if step else 'gauss'}"
    print(f"[lsc golden] {t} n={r.shape[1]}", flush=True)
    p = parray(r, q, E)
    lsc = LSC(step_profile=step)
    lsc.apply(p, 0.5)
    Q = q.sum()
    assert lsc.last_info["sigma"] == pytest.approx(float(lg[t + "/sigma"]), rel=1e-13)
    s, cur, W = lsc.last_profile()
    np.testing.assert_array_equal(s, lg[t + "/s"])
    assert rel(cur / (Q * lo.C), lg[t + "/bunch"]) < 1e-12
    assert rel(-W / Q, lg[t + "/wake"]) < 1e-10
    kick = lg[t + "/delta"] - r[5]
    assert rel(p.rparticles[5] - r[5], kick) < 1e-10
    np.testing.assert_array_equal(p.rparticles[:5], r[:5])


def test_lsc_apply_on_a_foreign_process_object(lg):
    """LSC.apply called on an object with only the reference class's
    attributes (Ocelot's own LSC delegating to the GPU) kicks as the drop-in
    instance does, bit for bit."""
    import types
    from ocelot_amd import LSC
    r, q, E = lg["gauss/r"], lg["gauss/q"], float(lg["gauss/E"])
    a, b = parray(r, q, E), parray(r, q, E)
    LSC(step_profile=True, smooth_param=0.2).apply(a, 0.5)
    ns = types.SimpleNamespace(step_profile=True, smooth_param=0.2, bounds=[-0.4, 0.4], _is_undul_in_beam_line=False)
    LSC.apply(ns, b, 0.5)
    assert np.array_equal(a.rparticles, b.rparticles) and ns.last_info["nbins"] > 0


class _Elem:
    def __init__(self, l):
        self.l = l


class Undulator(_Elem):
    def __init__(self, lperiod, nperiods, Kx=0.0, Ky=0.0):
        super().__init__(lperiod * nperiods)
        self.Kx, self.Ky = Kx, Ky


class _Line:
    def __init__(self, seq):
        self.sequence = seq

    def get_sequence_part(self, start, stop):
        return self.sequence[self.sequence.index(start):self.sequence.index(stop) + 1]


def test_lsc_undulator_step_matches_reference(lg):
    """The line of the golden script (drift 0.7 m, undulator 25 x 0.04 m with
    Kx = 1.8, drift 0.45 m) and a step ending at 0.75 m, half inside the
    undulator: the kick the reference gave there."""
    from ocelot_amd import LSC
    r, q, E = lg["gauss/r"], lg["gauss/q"], float(lg["gauss/E"])
    m1, m2 = _Elem(0.0), _Elem(0.0)
    line = _Line([m1, _Elem(0.7), Undulator(0.04, 25, Kx=1.8), _Elem(0.45), m2])
    lsc = LSC()
    lsc.start_elem, lsc.end_elem, lsc.s_start = m1, m2, 0.0
    lsc.prepare(line)
    lsc.z0 = 0.75
    p = parray(r, q, E)
    lsc.apply(p, 0.1)
    assert (lsc.last_info["K_max"], lsc.last_info["fill_factor"]) == (float(lg["und/Kmax"][0]),
                                                                       float(lg["und/fill"][0]))
    assert rel(p.rparticles[5] - r[5], lg["und/delta"] - r[5]) < 1e-10


def _bunch(n, kind, seed):
    rng = np.random.default_rng(seed)
    r = np.empty((6, n))
    r[0] = rng.normal(0, 50e-6, n)
    r[1] = rng.normal(0, 1e-5, n)
    r[2] = rng.normal(0, 70e-6, n)
    r[3] = rng.normal(0, 1e-5, n)
    if kind == "gauss":
        r[4] = rng.normal(0, 20e-6, n)
    else:  # flat top with a spike
        r[4] = np.where(rng.random(n) < 0.9, rng.uniform(-60e-6, 60e-6, n), rng.normal(20e-6, 3e-6, n))
    r[5] = rng.normal(0, 2e-4, n)
    return r, np.full(n, 0.5e-9 / n)


@pytest.mark.parametrize("kind,step,n", [("gauss", False, 2_000_000), ("flat", True, 2_000_000),
                                         ("flat", False, 4_000_000)])
def test_lsc_device_bunch_against_oracle(kind, step, n):
    """Resident bunch (CUDA tensor rows), kicked in place, against the oracle
    on the same input; the rows other than delta are untouched."""
    from ocelot_amd import LSC, DeviceParticleArray
    print(f"[lsc oracle] kind={kind} step={step} n={n}", flush=True)
    r, q = _bunch(n, kind, 11)
    E = 2.4
    ref, dbg = r.copy(), {}
    lo.apply(ref, q, E, 0.7, step_profile=step, debug=dbg)
    d = DeviceParticleArray(torch.from_numpy(r.copy()).cuda(), torch.from_numpy(q).cuda(), E)
    lsc = LSC(step_profile=step)
    lsc.apply(d, 0.7)
    got = d.rparticles.cpu().numpy()
    np.testing.assert_array_equal(got[:5], r[:5])
    assert lsc.last_info["sigma"] == pytest.approx(dbg["sigma"], rel=1e-13)
    s, cur, W = lsc.last_profile()
    # the grid starts at min(tau) - 3 smooth std(tau): std summed in another
    # order than numpy's moves it by an ulp or so
    assert len(s) == len(dbg["B"]) and rel(s, dbg["B"][:, 0]) < 1e-14
    assert rel(cur, dbg["B"][:, 1]) < 1e-12
    assert rel(W, dbg["W"]) < 1e-10
    assert rel(got[5] - r[5], ref[5] - r[5]) < 1e-10


def test_lsc_is_deterministic():
    from ocelot_amd import LSC, DeviceParticleArray
    r, q = _bunch(1_000_000, "flat", 5)
    outs = []
    for _ in range(2):
        d = DeviceParticleArray(torch.from_numpy(r.copy()).cuda(), torch.from_numpy(q).cuda(), 1.0)
        LSC().apply(d, 0.3)
        outs.append(d.rparticles[5].cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])


def test_impedance_against_oracle_and_series_branch():
    """Gaussian impedance over alpha^2 from below the 1e-16 cut-off through the
    exp(x) E1(x) branch into the asymptotic series (alpha^2 > 40)."""
    from ocelot_amd import LSC
    lsc = LSC()
    gamma, sigma, dz = 5.0, 2e-4, 0.3
    a2 = np.concatenate([[5e-17, 2e-16], np.geomspace(1e-15, 39.9, 300), [40.0, 40.0001], np.geomspace(41, 5e3, 60)])
    w = np.sqrt(a2) * gamma * lo.C / sigma
    got = lsc.imp_lsc(gamma, sigma, w, dz)
    ref = lo.imp_lsc(gamma, sigma, w, dz)
    assert np.all(got.real == 0) and got[0] == 0
    nz = ref.imag != 0
    assert np.array_equal(nz, got.imag != 0)
    assert np.abs(got.imag[nz] / ref.imag[nz] - 1).max() < 2e-14


def test_step_impedance_against_extended_precision():
    """Uniform-beam impedance over x = w rb/(c gamma) from 1e-9 to 800: the
    device's cancellation-free 1 - x K1(x) against 40-digit arithmetic, and
    against the reference form (scipy k1) within that form's own rounding."""
    mpmath = pytest.importorskip("mpmath")
    from ocelot_amd import LSC
    lsc = LSC(step_profile=True)
    gamma, rb, dz = 50.0, 1e-4, 1.0
    x = np.concatenate([np.geomspace(1e-9, 2, 150), np.geomspace(2.0001, 800, 150)])
    w = x * lo.C * gamma / rb
    got = lsc.imp_step_lsc(gamma, rb, w, dz).imag
    mpmath.mp.dps = 40
    xx = w * rb / (lo.C * gamma)
    f = np.array([float(1 - mpmath.mpf(float(v)) * mpmath.besselk(1, float(v))) for v in xx])
    exact = lo.Z0 * lo.C / (4 * w * rb * rb) * dz * f
    assert np.abs(got / exact - 1).max() < 1e-13
    ref = lo.imp_step_lsc(gamma, rb, w, dz).imag
    assert np.all(np.abs(got - ref) <= 2 * np.abs(ref - exact) + 1e-13 * np.abs(exact))
    assert lsc.imp_step_lsc(gamma, rb, np.array([0.0, 5e-8]), dz).imag.tolist() == [0.0, 0.0]


@pytest.mark.parametrize("step", [False, True])
def test_wake_lsc_against_oracle(step):
    """wake_lsc on a hand-made profile (two Gaussians, 3001 bins, undulator
    factor on) against the oracle."""
    from ocelot_amd import LSC
    s = np.arange(3001) * 1.3e-6 - 2e-3
    b = np.exp(-0.5 * ((s + 4e-4) / 2e-4) ** 2) + 0.4 * np.exp(-0.5 * ((s - 5e-4) / 5e-5) ** 2)
    b /= b.sum() * 1.3e-6
    got = LSC(step_profile=step).wake_lsc(s, b, 400.0, 8e-5, 0.25, K_max=2.0, fill_factor=0.3)
    ref = lo.wake_lsc(s, b, 400.0, 8e-5, 0.25, step, 2.0, 0.3)
    assert rel(got, ref) < 1e-10


def test_lsc_dz_below_threshold_is_noop(lg):
    from ocelot_amd import LSC
    p = parray(lg["gauss/r"], lg["gauss/q"], lg["gauss/E"])
    LSC().apply(p, 5e-11)
    np.testing.assert_array_equal(p.rparticles, lg["gauss/r"])


def test_lsc_errors_where_reference_raises(lg):
    from ocelot_amd import LSC
    r, q = lg["gauss/r"], lg["gauss/q"]
    with pytest.raises(ValueError):
        LSC().apply(parray(np.zeros((6, 0)), np.zeros(0), 1.0), 0.1)      # empty bunch
    flat = r.copy()
    flat[4] = 1e-4
    with pytest.raises(ValueError):
        LSC().apply(parray(flat, q, 1.0), 0.1)                            # no spread in tau
    bad = r.copy()
    bad[4, 7] = np.nan
    with pytest.raises(ValueError):
        LSC().apply(parray(bad, q, 1.0), 0.1)                             # non-finite tau
    with pytest.raises(ValueError):
        LSC(step_profile=True, bounds=[50, 60]).apply(parray(r, q, 1.0), 0.1)   # empty slice


def test_lsc_empty_gaussian_slice_gives_no_kick(lg):
    """With no particle in the slice the reference's Gaussian size is NaN,
    which zeroes its impedance: the kick is exactly zero."""
    from ocelot_amd import LSC
    p = parray(lg["gauss/r"], lg["gauss/q"], 1.0)
    LSC(bounds=[50, 60]).apply(p, 0.1)
    np.testing.assert_array_equal(p.rparticles, lg["gauss/r"] + 0.0)


def _resident(r, q, E):
    from ocelot_amd import DeviceParticleArray
    return DeviceParticleArray(torch.from_numpy(r.copy()).cuda(), torch.from_numpy(q.copy()).cuda(), E)


@pytest.mark.parametrize("step", [False, True])
def test_tracking_with_lsc_equals_step_by_step(step):
    """Graph-captured SC + LSC tracking (LSC grid planned on the device) is
    bit-identical to the eager loop drift, SpaceCharge.apply, LSC.apply (grid
    planned on the host), with the resident copy unsorted."""
    from ocelot_amd import LSC, SpaceCharge, drift, track_drift
    from ocelot_amd import _lib
    print(f"[lsc tracking] step_profile={step}", flush=True)
    r, q = _bunch(300_000, "flat", 9)
    E, L, dz, nsteps = 0.13, 0.05, 0.05, 6
    sc = SpaceCharge(nmesh_xyz=[31, 31, 31])
    lsc = LSC(step_profile=step)
    a = _resident(r, q, E)
    b = _resident(r, q, E)
    dev = a.rparticles.device.index
    old = _lib.get_option(dev, "sort_every")
    _lib.set_option(dev, "sort_every", 0)
    try:
        track_drift(a, sc, L, dz, nsteps, lsc=lsc)
        for _ in range(nsteps):
            drift(b, L)
            sc.apply(b, dz)
            lsc.apply(b, dz)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(a.rparticles.cpu().numpy(), b.rparticles.cpu().numpy())
        # sorted resident copy: same physics, sums in another order
        c = _resident(r, q, E)
        _lib.set_option(dev, "sort_every", 3)
        track_drift(c, sc, L, dz, nsteps, lsc=lsc)
        got, ref = c.rparticles.cpu().numpy(), b.rparticles.cpu().numpy()
        assert rel(got[5] - r[5], ref[5] - r[5]) < 1e-9
    finally:
        _lib.set_option(dev, "sort_every", old)


def test_lsc_only_tracking_against_oracle():
    """sc=None: drift + LSC per step, against the oracle's own loop."""
    from ocelot_amd import LSC, track_drift
    from oracle import sc_oracle as so
    print("[lsc only tracking]", flush=True)
    r, q = _bunch(200_000, "gauss", 4)
    E, L, dz, nsteps = 0.13, 0.2, 0.2, 5
    a = _resident(r, q, E)
    track_drift(a, None, L, dz, nsteps, lsc=LSC())
    ref = r.copy()
    for _ in range(nsteps):
        so.drift(ref, L, E)
        lo.apply(ref, q, E, dz)
    got = a.rparticles.cpu().numpy()
    assert rel(got[5] - r[5], ref[5] - r[5]) < 1e-9
    assert rel(got[4], ref[4]) < 1e-12


def test_tracking_lsc_reports_impossible_grid():
    """No spread in tau: the device plan refuses the kick, tracking reports
    it at the end (ValueError), and no LSC kick is applied."""
    from ocelot_amd import LSC, track_drift
    r, q = _bunch(10_000, "gauss", 1)
    r[4] = 0.0
    r[5] = 0.0   # no energy spread: the drifts keep tau at 0
    a = _resident(r, q, 0.13)
    with pytest.raises(ValueError):
        track_drift(a, None, 0.1, 0.1, 2, lsc=LSC())
    np.testing.assert_array_equal(a.rparticles.cpu().numpy()[5], r[5])


@pytest.mark.parametrize("step", [False, True])
def test_lsc_tracking_through_undulator_equals_step_by_step(step):
    """LSC-only tracking through the golden line (drift 0.7 m, undulator
    1.0 m with Kx = 1.8, drift 0.45 m) in steps of 0.1 m from s = 0.45 m: the
    per-step undulator factors taken up front give the same bunch, bit for
    bit, as drift + LSC.apply with the Navigator's z0 at each step end."""
    from ocelot_amd import LSC, drift, track_drift
    print(f"[lsc undulator tracking] step_profile={step}", flush=True)
    r, q = _bunch(200_000, "flat", 12)
    E, L, nsteps, s0 = 0.13, 0.1, 8, 0.45

    def prepared():
        m1, m2 = _Elem(0.0), _Elem(0.0)
        line = _Line([m1, _Elem(0.7), Undulator(0.04, 25, Kx=1.8), _Elem(0.45), m2])
        lsc = LSC(step_profile=step)
        lsc.start_elem, lsc.end_elem, lsc.s_start = m1, m2, 0.0
        lsc.prepare(line)
        return lsc

    a, b = _resident(r, q, E), _resident(r, q, E)
    la, lb = prepared(), prepared()
    track_drift(a, None, L, L, nsteps, lsc=la, z0=s0)
    factors = []
    for k in range(nsteps):
        drift(b, L)
        lb.z0 = s0 + (k + 1) * L
        lb.apply(b, L)
        factors.append(lb.last_info["K_max"])
    assert max(factors) > 0 and min(factors) == 0    # the steps cross into the undulator
    assert la.z0 == pytest.approx(s0 + nsteps * L)
    np.testing.assert_array_equal(a.rparticles.cpu().numpy(), b.rparticles.cpu().numpy())


def test_tracking_sc_lsc_and_twiss_in_one_loop():
    """track()'s usual combination, SpaceCharge + LSC with calc_tws, in one
    captured loop (ocsc_track): the particles equal the same loop without
    moments bit for bit, and equal the eager sequence drift, SpaceCharge.apply,
    LSC.apply (unsorted copy); each Twiss row equals get_envelope of the eager
    bunch after that step."""
    from ocelot_amd import LSC, SpaceCharge, drift, get_envelope, track_drift
    from ocelot_amd import _lib
    print("[sc+lsc+tws tracking]", flush=True)
    r, q = _bunch(300_000, "flat", 13)
    E, L, dz, nsteps = 0.13, 0.05, 0.05, 5
    sc, lsc = SpaceCharge(nmesh_xyz=[31, 31, 31]), LSC()
    dev = torch.cuda.current_device()
    old = _lib.get_option(dev, "sort_every")
    _lib.set_option(dev, "sort_every", 0)
    try:
        a, b, e = _resident(r, q, E), _resident(r, q, E), _resident(r, q, E)
        tws = track_drift(a, sc, L, dz, nsteps, calc_tws=True, lsc=lsc, bounds=[-0.5, 0.5])
        track_drift(b, sc, L, dz, nsteps, lsc=lsc)
        ref = [get_envelope(e, bounds=[-0.5, 0.5])]
        for _ in range(nsteps):
            drift(e, L)
            sc.apply(e, dz)
            lsc.apply(e, dz)
            ref.append(get_envelope(e, bounds=[-0.5, 0.5]))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(a.rparticles.cpu().numpy(), b.rparticles.cpu().numpy())
        np.testing.assert_array_equal(a.rparticles.cpu().numpy(), e.rparticles.cpu().numpy())
        assert len(tws) == nsteps + 1
        for t, rf in zip(tws, ref):
            for k in ("emit_x", "emit_y", "beta_x", "alpha_y", "xx", "pp", "tautau"):
                assert float(getattr(t, k)) == pytest.approx(float(getattr(rf, k)), rel=1e-12, abs=0)
    finally:
        _lib.set_option(dev, "sort_every", old)


def test_lsc_apply_long_tail_takes_host_planned_grid():
    """A far outlier stretches the profile beyond the device-planned capacity
    (4096 bins): the kick is planned again on the host from the same
    statistics and still matches the oracle."""
    from ocelot_amd import LSC, DeviceParticleArray
    r, q = _bunch(500_000, "gauss", 17)
    r[4, 0] = 3e-3          # 150 std away
    ref, dbg = r.copy(), {}
    lo.apply(ref, q, 1.5, 0.4, debug=dbg)
    assert len(dbg["B"]) > 4096
    d = DeviceParticleArray(torch.from_numpy(r.copy()).cuda(), torch.from_numpy(q).cuda(), 1.5)
    lsc = LSC()
    lsc.apply(d, 0.4)
    assert lsc.last_info["nbins"] == len(dbg["B"])
    got = d.rparticles.cpu().numpy()
    assert rel(got[5] - r[5], ref[5] - r[5]) < 1e-10
