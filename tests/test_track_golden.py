"""The oracle's drift map and tracking loop against vectors written by the
reference's own Drift element and track() (tests/golden/make_golden_track.py).

Tolerances: the drift map and the particles after tracking are bit-exact
(the oracle makes the reference's library calls in the reference's order,
including the navigator's section lengths); the Twiss list within 1e-9 of
each quantity's scale (the oracle's envelope takes its means in another
order than the reference's, see tests/test_envelope_oracle.py).
"""
import hashlib
import os

import numpy as np
import pytest

from oracle import sc_oracle as orc
from oracle import track_oracle as to

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["trk_step1", "trk_step5", "trk_step5_bounds", "trk_split"]
LINE = [0.0, 0.1, 0.16, 0.0, 0.04, 0.2, 0.0]   # M0, D0, D1, M1, D2, D3, M2


@pytest.fixture(scope="module")
def tg():
    with np.load(os.path.join(HERE, "golden", "track_golden.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def track_inputs(tg):
    """The 20k-particle bunch of the tracking cases, regenerated from its seed."""
    from ocelot_amd.workloads import gaussian_bunch
    n = int(tg["trk/n"])
    r = gaussian_bunch(n, sigmas=(100e-6, 1e-5, 100e-6, 1e-5, 1e-3, 1e-3), seed=32)
    assert hashlib.sha256(r.tobytes()).hexdigest() == str(tg["trk/in_sha256"])
    return r, np.full(n, float(tg["trk/Q"]) / n)


def schedule(tg, tag):
    """Per navigator step, the kicks the reference made: [(z_step, nmesh)]."""
    nav = tg[f"{tag}/nav_dz"]
    z = np.cumsum(nav)
    kicks = [[] for _ in nav]
    procs = tg[f"{tag}/procs"]
    for k in range(len(procs)):
        for z0, zs, _s in tg[f"{tag}/calls{k}"]:
            j = int(np.argmin(np.abs(z - z0)))
            assert abs(z[j] - z0) < 1e-12
            kicks[j].append((float(zs), tuple(int(v) for v in procs[k][1:])))
    b = tg[f"{tag}/bounds"]
    return nav, kicks, (None if np.isnan(b[0]) else [float(v) for v in b])


def twiss_rel_err(tws, ref_rows, attrs):
    """Largest |got - ref| over the Twiss list, relative to each quantity's
    scale (its own size or the size of the moments it is made of)."""
    scale = {"x": "xx", "y": "yy", "px": "pxpx", "py": "pypy", "tau": "tautau", "p": "pp"}
    worst = 0.0
    for k, t in enumerate(tws):
        ref = dict(zip(attrs, ref_rows[k]))
        for a in attrs:
            if a not in t or not np.isfinite(ref[a]):
                continue
            sc = abs(ref[a])
            if a in scale:
                sc = max(sc, np.sqrt(abs(ref[scale[a]])))
            worst = max(worst, abs(float(np.real(t[a])) - ref[a]) / (sc + 1e-300))
    return worst


@pytest.mark.parametrize("tag", ["hi", "mid", "lo"])
def test_drift_matrix_golden_and_closed_form(tg, tag):
    """R of a drift: the reference's Drift(l).R(E) exactly, and the closed
    form R01 = R23 = L, R45 = -L/(beta^2 gamma^2) with every other entry of
    the identity (a dropped gamma^2 or a sign slip fails at gamma 2.9)."""
    E, L = float(tg[f"drift_{tag}/E"]), float(tg[f"drift_{tag}/L"])
    R = orc.drift_matrix(L, E)
    assert np.array_equal(R, tg[f"drift_{tag}/R"])
    g = E / orc.M_E_GEV
    b2 = 1 - 1 / g ** 2
    ref = np.eye(6)
    ref[0, 1] = ref[2, 3] = L
    ref[4, 5] = -L / (b2 * g * g)
    assert np.allclose(R, ref, rtol=1e-14, atol=0)


@pytest.mark.parametrize("tag", ["hi", "mid", "lo"])
def test_drift_golden(tg, tag):
    """The drift applied to a bunch equals the reference's tracking of a
    Drift element bit for bit; tau moves by r56 * delta."""
    E, L = float(tg[f"drift_{tag}/E"]), float(tg[f"drift_{tag}/L"])
    r0 = tg["drift/r_in"]
    r = orc.drift(r0.copy(), L, E)
    assert np.array_equal(r, tg[f"drift_{tag}/r_out"])
    assert tg[f"drift_{tag}/s_out"] == L
    g = E / orc.M_E_GEV
    # r56 = -L / (beta^2 gamma^2) = -L / (gamma^2 - 1); the difference of tau
    # carries the rounding of tau itself
    eps = np.finfo(np.float64).eps
    assert np.max(np.abs((r[4] - r0[4]) - (-L / (g * g - 1)) * r0[5])) <= 4 * eps * np.max(np.abs(r[4]))


def test_navigator_sections():
    """Section lengths of the navigator on the line M0 D0(0.1) D1(0.16) M1
    D2(0.04) D3(0.2) M2: a step that ends on an element end takes the rest of
    the element (l - start), computed from the element's length."""
    line = to.DriftLine(LINE)
    secs = [line.sections(0.02) for _ in range(25)]
    assert secs[0] == [0.0, 0.02]
    assert secs[4] == [0.1 - (0.08 + 0.1 - 0.1)]      # D0 ends: l - start, not 0.1 - 0.08
    assert abs(line.z0 - 0.5) < 1e-12
    assert sum(sum(s) for s in secs) == pytest.approx(0.5, abs=1e-12)


@pytest.mark.parametrize("tag", CASES)
def test_oracle_track_loop_golden(tg, tag):
    """track() on the drift line with SpaceCharge at step 1, step 5, step 5
    with bounds, and a step-1 process followed by a step-5 process: the
    oracle's loop, fed the reference's schedule, gives its particles bit for
    bit and its Twiss list to rounding."""
    r0, q = track_inputs(tg)
    nav, kicks, bounds = schedule(tg, tag)
    r = r0.copy()
    tws = to.track_line(r, q, float(tg[f"{tag}/E"]), LINE, nav, kicks, bounds=bounds)
    sub = int(tg["trk/sub"])
    assert np.array_equal(r[:, ::sub], tg[f"{tag}/r_out_sub"])
    assert np.array_equal(r.sum(axis=1), tg[f"{tag}/row_sum"])
    attrs = [str(a) for a in tg["meta/attrs"]]
    assert len(tws) == tg[f"{tag}/tws"].shape[0]
    err = twiss_rel_err(tws, tg[f"{tag}/tws"], attrs)
    print(f"{tag}: twiss max rel {err:.3e}")
    assert err <= 1e-9


def test_schedule_of_the_split_processes(tg):
    """The reference's navigator on the split line: the step-1 process kicks
    13 times up to M1 (0.26 m), the step-5 process then takes 0.1, 0.1 and a
    last step cut to 0.04 at the end of the line; dz of every kick equals
    its navigator step."""
    nav = tg["trk_split/nav_dz"]
    c0, c1 = tg["trk_split/calls0"], tg["trk_split/calls1"]
    assert len(c0) == 13 and np.allclose(c0[:, 1], 0.02)
    assert np.allclose(c1[:, 1], [0.1, 0.1, 0.04])
    assert np.allclose(nav, [0.02] * 13 + [0.1, 0.1, 0.04])
    assert np.allclose(c1[:, 0], [0.36, 0.46, 0.5])
