"""Write golden vectors for the drift map and for the reference's own tracking
loop with space charge, by running the reference in the development
container (reference checkout at /root/reference).  The vectors are committed
next to this script as track_golden.npz; nothing here runs on a GPU machine
and no reference source is stored.

Usage:  PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden_track.py

Cases (FP64; numpy fallback branches of the reference, pyfftw and numexpr
absent, recorded in meta):
  drift_<tag>   first-order R matrix of a Drift element (Drift(l).R(E),
                ocelot/cpbd/r_matrix.py:41-91) and the element's map applied
                to a seeded bunch by track() without processes
                (ocelot/cpbd/transformations/transfer_map.py:40-50), at three
                energies (gamma 254, 12.7 and 2.9)
  trk_<tag>     track(lattice, p_array, navi) (ocelot/cpbd/track.py:431-504)
                on a four-drift line with markers, unit_step 0.02, with
                SpaceCharge(step=1), SpaceCharge(step=5), step 5 with
                bounds=[-0.5, 0.5], and the reference test's split of a
                step-1 process on the first part and a step-5 process on the
                rest (unit_tests/ebeam_test/space_charge/space_charge_test.py:
                148-175): the navigator step lengths, every (z0, z_step, s) the
                processes were called with, the Twiss list, every 4th final
                particle, the row sums and the final s.  The 20k-particle input
                bunch regenerates from its seed (checksum stored).
"""
import hashlib
import os
import sys

os.environ.setdefault("PYTHONDONTWRITEBYTECODE", "1")
sys.dont_write_bytecode = True
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, "/root/reference")

import numpy as np  # noqa: E402

from ocelot.cpbd.sc import SpaceCharge  # noqa: E402
from ocelot.cpbd import sc as ref_sc  # noqa: E402
from ocelot.cpbd.beam import ParticleArray  # noqa: E402
from ocelot.cpbd.elements import Drift, Marker  # noqa: E402
from ocelot.cpbd.magnetic_lattice import MagneticLattice  # noqa: E402
from ocelot.cpbd.navi import Navigator  # noqa: E402
from ocelot.cpbd.track import track  # noqa: E402

from ocelot_amd.workloads import gaussian_bunch  # noqa: E402

ATTRS = ["E", "q", "p", "Dx", "Dxp", "Dy", "Dyp", "x", "y", "px", "py", "tau", "xx", "xpx", "pxpx", "yy", "ypy",
         "pypy", "tautau", "xy", "pxpy", "xpy", "ypx", "pp", "emit_x", "emit_y", "emit_xn", "emit_yn", "eigemit_1",
         "eigemit_2", "beta_x", "beta_y", "alpha_x", "alpha_y", "s"]
DRIFTS = (0.1, 0.16, 0.04, 0.2)   # 0.5 m; the marker M1 sits at 0.26 m
SUB = 4                           # final particles kept: every 4th


class RecordingSC(SpaceCharge):
    """Records the (z0, z_step) of every call the tracking loop makes."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.calls = []

    def apply(self, p_array, zstep):
        self.calls.append((float(self.z0), float(zstep), float(p_array.s)))
        super().apply(p_array, zstep)


def parray(r, q, E):
    p = ParticleArray(n=r.shape[1])
    p.rparticles[:] = r
    p.q_array[:] = q
    p.E = E
    return p


def line():
    m0, m1, m2 = Marker(eid="M0"), Marker(eid="M1"), Marker(eid="M2")
    d = [Drift(l=v, eid=f"D{i}") for i, v in enumerate(DRIFTS)]
    seq = (m0, d[0], d[1], m1, d[2], d[3], m2)
    return MagneticLattice(seq), m0, m1, m2


def case_track(out, tag, r, q, E, procs, bounds=None):
    """procs: list of (step, nmesh, start marker name, stop marker name)."""
    lat, m0, m1, m2 = line()
    marks = {"M0": m0, "M1": m1, "M2": m2}
    navi = Navigator(lat)
    scs = []
    for step, nmesh, a, b in procs:
        sc = RecordingSC()
        sc.nmesh_xyz = list(nmesh)
        sc.step = step
        navi.add_physics_proc(sc, marks[a], marks[b])
        scs.append(sc)
    navi.unit_step = 0.02
    p = parray(r, q, E)
    steps = []
    orig = navi.get_next_step

    def get_next_step():
        for item in orig():
            steps.append(float(item[1]))
            yield item
    navi.get_next_step = get_next_step
    tws, p = track(lat, p, navi, print_progress=False, bounds=bounds)
    out[f"{tag}/E"] = np.float64(E)
    out[f"{tag}/procs"] = np.array([[s, *n] for s, n, _, _ in procs], dtype=np.int64)
    out[f"{tag}/bounds"] = np.array(bounds if bounds is not None else [np.nan, np.nan], dtype=np.float64)
    out[f"{tag}/nav_dz"] = np.array(steps)
    for k, sc in enumerate(scs):
        out[f"{tag}/calls{k}"] = np.array(sc.calls, dtype=np.float64).reshape(-1, 3)
    out[f"{tag}/tws"] = np.array([[float(np.real(getattr(t, a, np.nan))) for a in ATTRS] for t in tws])
    out[f"{tag}/r_out_sub"] = p.rparticles[:, ::SUB].copy()
    out[f"{tag}/row_sum"] = p.rparticles.sum(axis=1)
    out[f"{tag}/s_out"] = np.float64(p.s)


def main():
    out = {"meta/numpy": np.array(np.__version__), "meta/pyfftw_flag": np.array(bool(ref_sc.pyfftw_flag)),
           "meta/ne_flag": np.array(bool(ref_sc.ne_flag)), "meta/attrs": np.array(ATTRS),
           "meta/drifts": np.array(DRIFTS)}

    # drift map: R matrix and its application by the tracking loop
    r = gaussian_bunch(1000, sigmas=(1e-3, 2e-3, 1e-3, 2e-3, 1e-3, 2e-2), seed=31)
    out["drift/r_in"] = r
    for tag, E in (("hi", 0.13), ("mid", 0.0065), ("lo", 0.0015)):
        L = 0.37
        d = Drift(l=L)
        R = d.R(E)
        lat = MagneticLattice((Marker(), d, Marker()))
        p = parray(r, np.full(1000, 1e-15), E)
        track(lat, p, Navigator(lat), calc_tws=False, print_progress=False)
        out[f"drift_{tag}/E"] = np.float64(E)
        out[f"drift_{tag}/L"] = np.float64(L)
        out[f"drift_{tag}/R"] = np.array(R[0])
        out[f"drift_{tag}/r_out"] = p.rparticles.copy()
        out[f"drift_{tag}/s_out"] = np.float64(p.s)

    # tracking with space charge through the drift line
    n = 20000
    r = gaussian_bunch(n, sigmas=(100e-6, 1e-5, 100e-6, 1e-5, 1e-3, 1e-3), seed=32)
    q = np.full(n, 250e-12 / n)
    # inputs regenerate from the seed (gaussian_bunch(20000, sigmas, seed=32)); checksum kept
    out["trk/in_sha256"] = np.array(hashlib.sha256(r.tobytes()).hexdigest())
    out["trk/n"] = np.int64(n)
    out["trk/Q"] = np.float64(250e-12)
    out["trk/sub"] = np.int64(SUB)
    case_track(out, "trk_step1", r, q, 0.0065, [(1, (15, 15, 15), "M0", "M2")])
    case_track(out, "trk_step5", r, q, 0.0065, [(5, (31, 31, 31), "M0", "M2")])
    case_track(out, "trk_step5_bounds", r, q, 0.0065, [(5, (31, 31, 31), "M0", "M2")], bounds=[-0.5, 0.5])
    case_track(out, "trk_split", r, q, 0.0065, [(1, (15, 15, 15), "M0", "M1"), (5, (15, 15, 15), "M1", "M2")])

    path = os.path.join(HERE, "track_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
