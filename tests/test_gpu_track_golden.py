"""Graph-captured tracking on the GPU (track_drift with calc_tws) against the
reference's own track() on a drift line with SpaceCharge at step 1, step 5,
step 5 with bounds, and a step-1 process followed by a step-5 process
(tests/golden/make_golden_track.py).  The navigator's schedule recorded by
the reference (step lengths, kick z_step) drives the calls: runs of equal
(step, z_step) become one track_drift call each.

Tolerances (FP64 form): particles within 1e-10 of each row's largest value
(the 6.5 MeV bunch is kicked to ~50x its initial divergence over 25 steps,
so per-kick round-off of 1e-12 grows along the line); Twiss within 1e-8 of
each quantity's scale; s and the Twiss positions within 1e-12.  Mixed form:
the accumulated momentum change within 1e-5 RMS relative.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from test_track_golden import CASES, schedule, tg, track_inputs  # noqa: E402,F401
from test_gpu_parity import parray, rel_max, rel_rms  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


def runs_of(nav, kicks):
    """[(L, z_step, nmesh, count)]: consecutive steps with one kick of equal
    step length, z_step and mesh."""
    out = []
    for dz, k in zip(nav, kicks):
        assert len(k) == 1
        key = (float(dz), k[0][0], k[0][1])
        if out and out[-1][:3] == key:
            out[-1] = (*key, out[-1][3] + 1)
        else:
            out.append((*key, 1))
    return out


def track_case(tg, tag, prec):
    from ocelot_amd import DeviceParticleArray, SpaceCharge, track_drift
    r0, q = track_inputs(tg)
    nav, kicks, bounds = schedule(tg, tag)
    E = float(tg[f"{tag}/E"])
    d = DeviceParticleArray.from_host(parray(r0, q, E))
    tws, s_off = [], 0.0
    for L, zs, nmesh, cnt in runs_of(nav, kicks):
        t = track_drift(d, SpaceCharge(nmesh_xyz=list(nmesh), precision=prec), L, zs, cnt, use_graph=True,
                        calc_tws=True, bounds=bounds)
        if tws:
            t = t[1:]          # the first row repeats the previous call's last
        for x in t:
            x.s += s_off
        tws += t
        s_off += cnt * L
    return r0, d, tws


@pytest.mark.parametrize("tag", CASES)
def test_track_drift_vs_reference_track(tg, tag):
    r0, d, tws = track_case(tg, tag, "fp64")
    sub = int(tg["trk/sub"])
    got = d.rparticles.cpu().numpy()
    ref = tg[f"{tag}/r_out_sub"]
    row = max(rel_max(got[i, ::sub], ref[i]) for i in range(6))
    attrs = [str(a) for a in tg["meta/attrs"]]
    rows = tg[f"{tag}/tws"]
    assert len(tws) == rows.shape[0]
    scale = {"x": "xx", "y": "yy", "px": "pxpx", "py": "pypy", "tau": "tautau", "p": "pp"}
    worst = 0.0
    for k, t in enumerate(tws):
        refk = dict(zip(attrs, rows[k]))
        for a in attrs:
            if a == "s" or not np.isfinite(refk[a]):
                continue
            sc = abs(refk[a])
            if a in scale:
                sc = max(sc, np.sqrt(abs(refk[scale[a]])))
            worst = max(worst, abs(float(getattr(t, a)) - refk[a]) / (sc + 1e-300))
        assert abs(t.s - refk["s"]) <= 1e-12
    print(f"{tag}: rows {row:.3e}, twiss {worst:.3e}, s {d.s} (reference {float(tg[f'{tag}/s_out'])})")
    assert row <= 1e-10 and worst <= 1e-8
    assert abs(d.s - float(tg[f"{tag}/s_out"])) <= 1e-12


@pytest.mark.parametrize("tag", ["trk_step1", "trk_split"])
def test_mixed_track_drift_vs_reference_track(tg, tag):
    r0, d, _ = track_case(tg, tag, "mixed")
    sub = int(tg["trk/sub"])
    got = d.rparticles.cpu().numpy()[:, ::sub]
    ref = tg[f"{tag}/r_out_sub"]
    r0s = r0[:, ::sub]
    kick = max(rel_rms(got[i] - r0s[i], ref[i] - r0s[i]) for i in (1, 3, 5))
    print(f"{tag} mixed: accumulated kick rms {kick:.3e}")
    assert kick <= 1e-5
