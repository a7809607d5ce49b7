"""Write golden vectors for get_envelope (ocelot/cpbd/beam/analysis.py:27-224)
by running the reference in the development container (reference checkout at
/root/reference).  The vectors are committed next to this script as
envelope_golden.npz; nothing here runs on a GPU machine and no reference
source is stored.

Usage:  PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden_envelope.py

Bunches (seeded, FP64): a correlated Gaussian with dispersion and x-y
coupling (5000 particles) and a flat-top (3000).  Settings per bunch: plain,
bounds=[-2, 1] (slice=None), auto_disp=True, tws_i with given dispersion,
bounds=[-0.5, 0.5] with slice="Imax" (window on the current peak);
plus a 2-particle bunch (fewer than 3: only E, q, p are set).  numexpr is
absent here, so the reference's numpy branch ran (recorded in meta).
"""
import os
import sys

os.environ.setdefault("PYTHONDONTWRITEBYTECODE", "1")
sys.dont_write_bytecode = True
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, "/root/reference")

import numpy as np  # noqa: E402

from ocelot.cpbd.beam import ParticleArray, Twiss  # noqa: E402
from ocelot.cpbd.beam import analysis as ref_an  # noqa: E402
from ocelot.cpbd.beam.analysis import get_envelope  # noqa: E402

ATTRS = ["E", "q", "p", "Dx", "Dxp", "Dy", "Dyp", "x", "y", "px", "py", "tau", "xx", "xpx", "pxpx", "yy", "ypy",
         "pypy", "tautau", "xy", "pxpy", "xpy", "ypx", "pp", "emit_x", "emit_y", "emit_xn", "emit_yn", "eigemit_1",
         "eigemit_2", "beta_x", "beta_y", "alpha_x", "alpha_y"]


def bunch_gauss(n, seed):
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((6, n))
    r = np.empty((6, n))
    r[5] = 1e-3 * z[5]
    r[4] = 1e-3 * z[4]
    r[0] = 100e-6 * z[0] + 0.25 * r[5]
    r[1] = 1e-5 * z[1] - 0.04 * r[0] + 0.01 * r[5]
    r[2] = 80e-6 * z[2] + 0.3 * r[0] - 0.1 * r[5]
    r[3] = 2e-5 * z[3] + 0.02 * r[2] - 0.05 * r[1]
    q = np.full(n, 250e-12 / n)
    return r, q


def bunch_flat(n, seed):
    rng = np.random.default_rng(seed)
    r = np.empty((6, n))
    r[0] = rng.uniform(-2e-4, 2e-4, n)
    r[1] = rng.uniform(-3e-5, 3e-5, n) + 0.05 * r[0]
    r[2] = rng.uniform(-1e-4, 1e-4, n)
    r[3] = rng.uniform(-2e-5, 2e-5, n)
    r[4] = rng.uniform(-1e-3, 1e-3, n)
    r[5] = -5 * r[4] + 2e-4 * rng.standard_normal(n)
    q = rng.uniform(0.5, 1.5, n) * 1e-15
    return r, q


def run(r, q, E, **kw):
    p = ParticleArray(n=r.shape[1])
    p.rparticles[:] = r
    p.q_array[:] = q
    p.E = E
    tws = get_envelope(p, **kw)
    return np.array([float(np.real(getattr(tws, a, np.nan))) for a in ATTRS])


def main():
    out = {"meta/numpy": np.array(np.__version__), "meta/ne_flag": np.array(bool(ref_an.ne_flag)),
           "meta/attrs": np.array(ATTRS)}
    tws_i = Twiss()
    tws_i.Dx, tws_i.Dxp, tws_i.Dy, tws_i.Dyp = 0.3, 0.01, -0.2, 0.005
    settings = {"plain": {}, "bounds": {"bounds": [-2.0, 1.0]}, "autodisp": {"auto_disp": True},
                "twsi": {"tws_i": tws_i}, "imax": {"bounds": [-0.5, 0.5], "slice": "Imax"}}
    for name, (r, q, E) in {"gauss": (*bunch_gauss(5000, 1), 0.13), "flat": (*bunch_flat(3000, 2), 0.7)}.items():
        out[f"{name}/r"] = r
        out[f"{name}/q"] = q
        out[f"{name}/E"] = np.array(E)
        for s, kw in settings.items():
            out[f"{name}/{s}"] = run(r, q, E, **kw)
    r, q = bunch_gauss(2, 3)
    out["two/r"], out["two/q"], out["two/E"] = r, q, np.array(0.13)
    out["two/plain"] = run(r, q, 0.13)
    out["disp/tws_i"] = np.array([0.3, 0.01, -0.2, 0.005])
    out["bounds/b"] = np.array([-2.0, 1.0])
    out["imax/b"] = np.array([-0.5, 0.5])
    for name in ("gauss", "flat"):
        from ocelot.cpbd.beam.analysis import s_to_cur
        tau = out[f"{name}/r"][4]
        B = s_to_cur(tau, 0.01 * np.std(tau), np.sum(out[f"{name}/q"]))
        out[f"{name}/cur"] = B
    path = os.path.join(HERE, "envelope_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
