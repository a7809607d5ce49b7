"""npz bunch files (ocelot_amd.io) against a file the reference wrote:
tests/golden/ref_bunch_track_wo_lsc.npz is the reference test suite's own
data file unit_tests/ebeam_test/long_space_charge/ref_results/
test_track_wo_lsc.npz (a 10k-particle bunch saved by the reference's
save_particle_array)."""
import os
import zipfile

import numpy as np
import pytest

from ocelot_amd import io as oio
from ocelot_amd import ParticleArray

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "golden", "ref_bunch_track_wo_lsc.npz")


def test_loads_reference_written_file():
    p = oio.load_particle_array(REF)
    assert isinstance(p, ParticleArray)
    assert p.rparticles.shape == (6, 10000) and p.rparticles.dtype == np.float64
    assert p.q_array.shape == (10000,)
    assert np.ndim(p.E) == 0 and np.ndim(p.s) == 0
    assert float(p.q_array.sum()) == pytest.approx(250e-12, rel=1e-12)   # the fixture's 250 pC
    assert np.all(np.isfinite(p.rparticles))


def test_round_trip_is_exact_and_in_reference_layout(tmp_path):
    p = oio.load_particle_array(REF)
    f = str(tmp_path / "b.npz")
    oio.save_particle_array(f, p)
    q = oio.load_particle_array(f)
    for k in ("rparticles", "q_array", "E", "s"):
        np.testing.assert_array_equal(getattr(q, k), getattr(p, k))
    # same members, dtypes and shapes as the file the reference wrote
    with zipfile.ZipFile(REF) as a, zipfile.ZipFile(f) as b:
        assert sorted(a.namelist()) == sorted(b.namelist())
    with np.load(REF) as a, np.load(f) as b:
        for k in a.files:
            assert (a[k].dtype, a[k].shape) == (b[k].dtype, b[k].shape)


def test_fresh_bunch_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    p = ParticleArray(n=1234)
    p.rparticles[:] = rng.standard_normal((6, 1234))
    p.q_array[:] = 1e-15
    p.E, p.s = 0.13, 4.25
    f = str(tmp_path / "c.npz")
    oio.save_particle_array2npz(f, p)
    q = oio.load_particle_array_from_npz(f)
    np.testing.assert_array_equal(q.rparticles, p.rparticles)
    assert float(q.E) == 0.13 and float(q.s) == 4.25


def test_unsupported_formats_and_bad_files(tmp_path):
    with pytest.raises(ValueError):
        oio.load_particle_array(str(tmp_path / "beam.ast"))
    with pytest.raises(ValueError):
        oio.save_particle_array(str(tmp_path / "beam.fmt1"), ParticleArray(n=3))
    f = str(tmp_path / "other.npz")
    np.savez(f, s=np.zeros(3))
    with pytest.raises(ValueError):
        oio.load_particle_array(f)
