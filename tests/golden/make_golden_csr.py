"""Write golden vectors for the tail of Ocelot's CSR kick (ocelot/cpbd/csr.py:
csr_convolution at lines 69-79 and CSR.apply at lines 1088-1146) by running
the reference in the development container (reference checkout at
/root/reference).  The vectors are committed next to this script as
csr_golden.npz; no reference source is stored.

Usage:  PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden_csr.py

A line of a marker, a 0.5 m sector bend (angle 0.03, then 0.06 in a second
case, and 0.03 with a fine mesh, sigma_min = 2 um, in a third), a 0.5 m drift
and a marker; the reference's CSR process prepares its
trajectory (arcs) and kicks a 40k-particle Gaussian bunch at z0 = 0.6 m with
delta_s = 0.1 m.  Recorded: the equidistant line charge lam_ds and mesh
(s1, s2, Ns) its binning produced, both operands and the result of each
csr_convolution call, the bunch before and after.  Plus csr_convolution on
random vectors of several lengths.
"""
import os
import sys

os.environ.setdefault("PYTHONDONTWRITEBYTECODE", "1")
sys.dont_write_bytecode = True
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, "/root/reference")

import numpy as np  # noqa: E402

import ocelot.cpbd.csr as csr_mod  # noqa: E402
from ocelot.cpbd.beam import ParticleArray  # noqa: E402
from ocelot.cpbd.elements import Drift, Marker, SBend  # noqa: E402
from ocelot.cpbd.magnetic_lattice import MagneticLattice  # noqa: E402

REC = {}
_conv = csr_mod.csr_convolution


def recording_convolution(a, b):
    c = _conv(a, b)
    REC["conv"] = (np.array(a), np.array(b), np.array(c))
    return c


csr_mod.csr_convolution = recording_convolution


def main():
    out = {"meta/numpy": np.array(np.__version__)}
    rng = np.random.default_rng(3)
    for k, (na, nb) in enumerate([(1, 1), (2, 3), (7, 7), (100, 100), (513, 700), (1000, 999)]):
        a, b = rng.standard_normal(na), rng.standard_normal(nb)
        out[f"conv{k}/a"], out[f"conv{k}/b"], out[f"conv{k}/c"] = a, b, _conv(a, b)
    n = 40000
    r = np.empty((6, n))
    r[0] = rng.normal(0, 50e-6, n)
    r[1] = rng.normal(0, 1e-5, n)
    r[2] = rng.normal(0, 50e-6, n)
    r[3] = rng.normal(0, 1e-5, n)
    r[4] = rng.normal(0, 25e-6, n)
    r[5] = rng.normal(0, 1e-4, n)
    q = np.full(n, 0.5e-9 / n)
    E = 0.13
    out["bunch/r"], out["bunch/q"], out["bunch/E"] = r, q, np.array(E)
    for case, angle, sigma_min in (("bend1", 0.03, 1e-4), ("bend2", 0.06, 1e-4), ("fine", 0.03, 2e-6)):
        m1, m2 = Marker(), Marker()
        lat = MagneticLattice([m1, SBend(l=0.5, angle=angle), Drift(l=0.5), m2])
        csr = csr_mod.CSR(sigma_min=sigma_min)
        csr.indx0, csr.indx1 = 0, len(lat.sequence) - 1
        csr.prepare(lat)
        q2e = csr.bin_smoth.Q2EQUI

        def rec_q2e(*args, _f=q2e):
            res = _f(*args)
            REC["q2e"] = res
            return res

        csr.bin_smoth.Q2EQUI = rec_q2e
        p = ParticleArray(n=n)
        p.rparticles[:] = r
        p.q_array[:] = q
        p.E = E
        csr.z0 = 0.6
        csr.apply(p, 0.1)
        s1, s2, Ns, lam = REC["q2e"]
        a, b, c = REC["conv"]
        out[f"{case}/s1"], out[f"{case}/s2"], out[f"{case}/Ns"] = np.array(s1), np.array(s2), np.array(Ns)
        out[f"{case}/lam_ds"] = np.array(lam)
        out[f"{case}/conv_a"], out[f"{case}/conv_b"], out[f"{case}/conv_c"] = a, b, c
        out[f"{case}/delta_s"] = np.array(0.1)
        out[f"{case}/delta"] = p.rparticles[5].copy()
        print(case, "Ns", Ns, "len(K1)", len(b), "max kick", np.abs(p.rparticles[5] - r[5]).max())
    path = os.path.join(HERE, "csr_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
