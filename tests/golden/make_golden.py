"""Write golden input/output vectors for the space-charge kick by running the
reference implementation (ocelot/cpbd/sc.py, ocelot/cpbd/coord_transform.py)
in the development container, where the reference checkout lives at
/root/reference.  The vectors are committed next to this script as npz data;
nothing here runs on a GPU machine and no reference source is stored.

Usage:  PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py

Cases (all FP64, numpy fallback branches of the reference: pyfftw and numexpr
are absent here, recorded in meta):
  coord_*      xxstg_2_xp_mad / xp_2_xxstg_mad on 1000 particles
  kern_*       sym_kernel on three (shape, h) sets
  pot_*        potential of a random charge grid
  apply_*      SpaceCharge.apply on seeded bunches, with the charge grid,
               cell sizes, potential and field captured at the sub-calls
  rmesh_*      random_mesh=True, two kicks after prepare() (RNG order)
  big_c1       200k particles on 63^3: checksums and a particle subsample
  err_*        the exception classes raised for empty / zero-charge bunches
"""
import os
import sys
import hashlib

os.environ.setdefault("PYTHONDONTWRITEBYTECODE", "1")
sys.dont_write_bytecode = True
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, "/root/reference")

import numpy as np  # noqa: E402
import scipy  # noqa: E402

from ocelot.cpbd.sc import SpaceCharge  # noqa: E402
from ocelot.cpbd import sc as ref_sc  # noqa: E402
from ocelot.cpbd.coord_transform import xxstg_2_xp_mad, xp_2_xxstg_mad  # noqa: E402
from ocelot.cpbd.beam import ParticleArray  # noqa: E402

from ocelot_amd.workloads import gaussian_bunch, make_bunch  # noqa: E402


class RecordingSC(SpaceCharge):
    """Captures the arguments/results of the reference's sub-calls."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.rec = {}

    def potential(self, q, steps):
        self.rec["rho"] = q.copy()
        self.rec["steps"] = np.array(steps, dtype=np.float64)
        phi = super().potential(q, steps)
        self.rec["phi"] = phi.copy()
        return phi

    def el_field(self, X, Q, gamma, nxyz):
        self.rec["gamma0"] = float(gamma)
        E = super().el_field(X, Q, gamma, nxyz)
        self.rec["Exyz"] = E.copy()
        return E


def parray(r, q, E):
    p = ParticleArray(n=r.shape[1])
    p.rparticles[:] = r
    p.q_array = q.copy()
    p.E = E
    return p


def case_apply(out, name, r, q, E, nmesh, zstep, keep_grids=True):
    sc = RecordingSC(nmesh_xyz=list(nmesh))
    p = parray(r, q, E)
    sc.apply(p, zstep)
    out[name + "/r_in"] = r
    out[name + "/q"] = q
    out[name + "/E"] = np.float64(E)
    out[name + "/zstep"] = np.float64(zstep)
    out[name + "/nmesh"] = np.array(nmesh, dtype=np.int64)
    out[name + "/r_out"] = p.rparticles.copy()
    out[name + "/Exyz"] = sc.rec["Exyz"]
    out[name + "/steps"] = sc.rec["steps"]
    out[name + "/gamma0"] = np.float64(sc.rec["gamma0"])
    if keep_grids:
        out[name + "/rho"] = sc.rec["rho"]
        out[name + "/phi"] = sc.rec["phi"]


def main():
    out = {}
    out["meta/numpy"] = np.array(np.__version__)
    out["meta/scipy"] = np.array(scipy.__version__)
    out["meta/pyfftw_flag"] = np.array(bool(ref_sc.pyfftw_flag))
    out["meta/ne_flag"] = np.array(bool(ref_sc.ne_flag))

    # coordinate maps
    rng = np.random.default_rng(101)
    r = gaussian_bunch(1000, sigmas=(1e-3, 2e-3, 1e-3, 2e-3, 1e-3, 2e-2), seed=101)
    r[:, :3] = [[0.0, 0.0, 1e-4], [0.0, 0.05, -0.05], [0.0, 0.0, 2e-3],
                [0.0, -0.04, 0.0], [0.0, 1e-3, -1e-3], [0.0, 0.3, -0.3]]
    for tag, E in (("hi", 0.5), ("lo", 0.0065)):
        gamref = E / ref_sc.m_e_GeV
        xp = xxstg_2_xp_mad(r, np.zeros(r.shape), gamref)
        back = xp_2_xxstg_mad(xp.copy(), np.zeros(r.shape), gamref)
        out[f"coord_{tag}/r"] = r
        out[f"coord_{tag}/E"] = np.float64(E)
        out[f"coord_{tag}/xp"] = xp
        out[f"coord_{tag}/back"] = back

    # integrated Green's function
    sc = SpaceCharge()
    for i, (shape, h) in enumerate((((7, 7, 7), (1e-5, 1e-5, 1e-3)),
                                    ((5, 9, 13), (2e-6, 3e-6, 5e-5)),
                                    ((16, 16, 16), (1.0, 1.0, 1.0)))):
        out[f"kern_{i}/shape"] = np.array(shape, dtype=np.int64)
        out[f"kern_{i}/h"] = np.array(h, dtype=np.float64)
        out[f"kern_{i}/K"] = sc.sym_kernel(shape, np.array(h))

    # potential of a random grid
    rho = rng.uniform(0, 1e-12, (9, 11, 13))
    steps = np.array([1e-5, 2e-5, 3e-5])
    out["pot_0/rho"] = rho
    out["pot_0/steps"] = steps
    out["pot_0/phi"] = sc.potential(rho, steps)

    # apply on seeded bunches
    r = gaussian_bunch(2000, seed=1)
    case_apply(out, "apply_gauss_2k_15", r, np.full(2000, 250e-12 / 2000), 0.130, (15, 15, 15), 0.1)

    r = gaussian_bunch(5000, sigmas=(50e-6, 2e-5, 120e-6, 1e-5, 0.5e-3, 1e-3), seed=2,
                       mean=(10e-6, 2e-3, -5e-6, -1e-3, 0.0, 0.0))
    r[5] += -8.0 * r[4]
    q = rng.uniform(0.5, 1.5, 5000) * 100e-12 / 5000
    case_apply(out, "apply_aniso_5k_9_13_17", r, q, 0.25, (9, 13, 17), 0.05)

    r = gaussian_bunch(10000, seed=3)
    case_apply(out, "apply_gauss_10k_31", r, np.full(10000, 500e-12 / 10000), 0.130, (31, 31, 31), 0.1)

    r = gaussian_bunch(2000, sigmas=(230e-6, 200e-6, 230e-6, 200e-6, 0.8e-3, 0.4e-3), seed=4)
    case_apply(out, "apply_lowE_2k_7", r, np.full(2000, 500e-12 / 2000), 0.0065, (7, 7, 7), -0.02)

    r = make_bunch("flattop_10m_127", n=4000, seed=5)[0]
    case_apply(out, "apply_flattop_4k_23_23_47", r, np.full(4000, 1e-9 / 4000), 0.5, (23, 23, 47), 0.1)

    # random mesh: prepare() seeds the global RNG, then two kicks
    r = gaussian_bunch(3000, seed=6)
    q = np.full(3000, 250e-12 / 3000)
    sc = RecordingSC(nmesh_xyz=[15, 15, 15], random_mesh=True)
    sc.prepare(None)
    p = parray(r, q, 0.130)
    sc.apply(p, 0.1)
    out["rmesh/r_in"] = r
    out["rmesh/q"] = q
    out["rmesh/r_out1"] = p.rparticles.copy()
    out["rmesh/steps1"] = sc.rec["steps"]
    sc.apply(p, 0.1)
    out["rmesh/r_out2"] = p.rparticles.copy()
    out["rmesh/steps2"] = sc.rec["steps"]

    # C1-sized bunch: checksums + subsample (inputs regenerate from the seed)
    r, q, E, nmesh, dz = make_bunch("gauss_200k_63")
    out["big_c1/in_sha256"] = np.array(hashlib.sha256(r.tobytes()).hexdigest())
    sc = RecordingSC(nmesh_xyz=list(nmesh))
    p = parray(r, q, E)
    sc.apply(p, dz)
    ro = p.rparticles
    out["big_c1/sub_idx"] = np.arange(0, ro.shape[1], 97)
    out["big_c1/r_out_sub"] = ro[:, ::97].copy()
    out["big_c1/Exyz_sub"] = sc.rec["Exyz"][::97].copy()
    out["big_c1/steps"] = sc.rec["steps"]
    out["big_c1/row_sum"] = ro.sum(axis=1)
    out["big_c1/row_sumsq"] = (ro * ro).sum(axis=1)
    out["big_c1/rho_sum"] = np.float64(sc.rec["rho"].sum())
    out["big_c1/phi_sub"] = sc.rec["phi"][::4, ::4, ::4].copy()

    # error behaviour
    for tag, (n, qv) in (("empty", (0, 1e-12)), ("zero_charge", (100, 0.0))):
        r = gaussian_bunch(n, seed=7)
        p = parray(r, np.full(n, qv), 0.13)
        try:
            SpaceCharge().apply(p, 0.1)
            out[f"err_{tag}/exc"] = np.array("")
        except Exception as e:  # record the class name only
            out[f"err_{tag}/exc"] = np.array(type(e).__name__)

    path = os.path.join(HERE, "sc_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
