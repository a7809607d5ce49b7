"""Write golden vectors for the longitudinal space-charge process LSC
(ocelot/cpbd/sc.py:261-600) by running the reference in the development
container (reference checkout at /root/reference).  The vectors are committed
next to this script as lsc_golden.npz; nothing here runs on a GPU machine and
no reference source is stored.

Usage:  PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden_lsc.py

Cases (FP64): a Gaussian bunch and a flat-top bunch with two current horns,
each with the Gaussian transverse impedance and with the uniform (step)
profile, no undulator in the line (K_max = 0); recorded: the inputs, the
transverse size, the current profile, the wake and the kicked delta row.
Then a line with an undulator (drift 0.7 m, undulator 25 x 0.04 m with
Kx = 1.8, drift 0.45 m): the K(s) profile from LSC.prepare, filling factors
and K_max of a few steps, and the Gaussian bunch kicked at the end of a
step that is half in the undulator.
"""
import os
import sys

os.environ.setdefault("PYTHONDONTWRITEBYTECODE", "1")
sys.dont_write_bytecode = True
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, "/root/reference")

import numpy as np  # noqa: E402

from ocelot.cpbd.sc import LSC  # noqa: E402
from ocelot.cpbd.beam import ParticleArray  # noqa: E402
from ocelot.cpbd.elements import Drift, Marker, Undulator  # noqa: E402
from ocelot.cpbd.magnetic_lattice import MagneticLattice  # noqa: E402


class RecordingLSC(LSC):
    def wake_lsc(self, s, bunch, gamma, sigma, dz, K_max=0, fill_factor=0):
        w = super().wake_lsc(s, bunch, gamma, sigma, dz, K_max, fill_factor)
        self.rec = dict(s=np.array(s), bunch=np.array(bunch), sigma=float(sigma), wake=np.array(w))
        return w


def bunches():
    rng = np.random.default_rng(7)
    n = 20000
    g = rng.standard_normal((6, n)) * np.array([[60e-6], [1e-5], [40e-6], [1e-5], [0.3e-3], [1e-3]])
    yield "gauss", g, np.full(n, 250e-12 / n), 0.13
    n = 30000
    h = np.empty((6, n))
    h[0] = rng.normal(0, 80e-6, n)
    h[1] = rng.normal(0, 1e-5, n)
    h[2] = rng.normal(0, 80e-6, n)
    h[3] = rng.normal(0, 1e-5, n)
    k = rng.integers(0, 4, n)
    tau = rng.uniform(-0.5e-3, 0.5e-3, n)
    tau = np.where(k == 2, -0.5e-3 + rng.normal(0, 20e-6, n), tau)
    tau = np.where(k == 3, 0.5e-3 + rng.normal(0, 20e-6, n), tau)
    h[4] = tau
    h[5] = rng.normal(0, 1e-3, n)
    yield "horns", h, np.full(n, 1e-9 / n), 0.5


def main():
    out = {"meta/numpy": np.array(np.__version__)}
    for name, r, q, E in bunches():
        out[f"{name}/r"], out[f"{name}/q"], out[f"{name}/E"] = r, q, np.array(E)
        for prof in (False, True):
            tag = f"{name}/{'step' if prof else 'gauss'}"
            lsc = RecordingLSC(step_profile=prof)
            p = ParticleArray(n=r.shape[1])
            p.rparticles[:] = r
            p.q_array[:] = q
            p.E = E
            lsc.apply(p, 0.5)
            for k, v in lsc.rec.items():
                out[f"{tag}/{k}"] = np.array(v)
            out[f"{tag}/delta"] = p.rparticles[5].copy()
    m1, m2 = Marker(), Marker()
    lat = MagneticLattice([m1, Drift(l=0.7), Undulator(lperiod=0.04, nperiods=25, Kx=1.8), Drift(l=0.45), m2])
    lsc = RecordingLSC()
    lsc.start_elem, lsc.end_elem, lsc.s_start = m1, m2, 0.0
    lsc.prepare(lat)
    s_pts = np.array([0.0, 0.35, 0.7, 0.71, 1.2, 1.7, 1.71, 2.15, 2.5])
    out["und/s_pts"], out["und/K_pts"] = s_pts, lsc.K_s_func(s_pts)
    steps = np.array([[0.75, 0.1], [0.72, 0.05], [1.72, 0.1], [0.3, 0.1], [2.1, 0.2], [1.75, 0.1]])
    out["und/steps"] = steps
    out["und/fill"] = np.array([lsc.compute_filling_factor(z - d, z, num_points=200) for z, d in steps])
    out["und/Kmax"] = np.array([np.max(lsc.K_s_func(np.linspace(z - d, z, num=200))) for z, d in steps])
    r, q, E = out["gauss/r"], out["gauss/q"], float(out["gauss/E"])
    p = ParticleArray(n=r.shape[1])
    p.rparticles[:] = r
    p.q_array[:] = q
    p.E = E
    lsc.z0 = 0.75
    lsc.apply(p, 0.1)
    out["und/wake"] = lsc.rec["wake"]
    out["und/delta"] = p.rparticles[5].copy()
    path = os.path.join(HERE, "lsc_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
