"""Host contract of the device paths: charges of the wrong length are
rejected before the library is called, and two host threads kicking on one
device give the same bits as the same kicks made one after the other (each
library call holds its context for its whole duration)."""
import threading

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from test_gpu_parity import parray  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


def test_wrong_length_charges_rejected_on_device():
    from ocelot_amd import LSC, DeviceParticleArray, SpaceCharge, apply_batched, get_envelope, track_drift
    from ocelot_amd.workloads import make_bunch
    r, q, E, nmesh, dz = make_bunch("flattop_10m_127", n=10_000, seed=3)
    with pytest.raises(ValueError):
        DeviceParticleArray(torch.from_numpy(r).cuda(), torch.from_numpy(q[:-1]).cuda(), E)
    d = DeviceParticleArray.from_host(parray(r, q, E))
    d.q_array = d.q_array[:-7]          # shortened after construction
    before = d.rparticles.clone()
    sc = SpaceCharge(nmesh_xyz=[15, 15, 15])
    for call in (lambda: sc.apply(d, dz), lambda: track_drift(d, sc, dz, dz, 2), lambda: LSC().apply(d, dz),
                 lambda: apply_batched([(d, (15, 15, 15), dz)]), lambda: get_envelope(d)):
        with pytest.raises(ValueError):
            call()
    assert torch.equal(d.rparticles, before)


def test_two_threads_on_one_device_equal_sequential_kicks():
    """Two threads, each kicking its own resident bunch five times (FP64 and
    mixed, 31^3 and 47^3), started together: every bunch ends bit-identical
    to the same five kicks made alone."""
    from ocelot_amd import DeviceParticleArray, SpaceCharge
    from ocelot_amd.workloads import make_bunch
    jobs = []
    for k, (prec, m) in enumerate((("fp64", 31), ("mixed", 47))):
        r, q, E, _, dz = make_bunch("flattop_10m_127", n=400_000, seed=40 + k)
        jobs.append((parray(r, q, E), SpaceCharge(nmesh_xyz=[m, m, m], precision=prec), dz))
    ref = []
    for p, sc, dz in jobs:
        d = DeviceParticleArray.from_host(p)
        for _ in range(5):
            sc.apply(d, dz)
        ref.append(d.rparticles.cpu())
    ds = [DeviceParticleArray.from_host(p) for p, _, _ in jobs]
    torch.cuda.synchronize()
    start = threading.Barrier(2)
    errors = []

    def work(i):
        try:
            _, sc, dz = jobs[i]
            start.wait()
            for _ in range(5):
                sc.apply(ds[i], dz)
            torch.cuda.synchronize()
        except Exception as e:  # reported in the main thread
            errors.append(e)

    ts = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not errors, errors
    for d, rr in zip(ds, ref):
        assert torch.equal(d.rparticles.cpu(), rr)


@pytest.mark.parametrize("name,good,bad", [
    ("fft_plan", (0, 1, 2, 3, 4), (-1, 5)),
    ("deposit_tile", (0, 1, 2, 3, 4), (-1, 5)),
    ("grid_clear", (0, 1, 2), (-1, 3)),
    ("igf_blocks", (0, 1, 65536), (-1, 65537)),
    ("fold_frame", (0, 1), (-1, 2)),
    ("q_uniform", (0, 1), (-1, 2)),
    ("apply_resort", (1, 2, 1000), (0, 1001)),
    ("rec_direct", (0, 1), (-1, 2)),
])
def test_options_round_trip_and_reject_out_of_range(name, good, bad):
    """Every value an option documents is kept and read back; values outside
    its range raise ValueError and leave the option as it was."""
    from ocelot_amd import _lib
    dev = torch.cuda.current_device()
    old = _lib.get_option(dev, name)
    try:
        for v in good:
            _lib.set_option(dev, name, v)
            assert _lib.get_option(dev, name) == v
        _lib.set_option(dev, name, good[0])
        for v in bad:
            with pytest.raises(ValueError):
                _lib.set_option(dev, name, v)
            assert _lib.get_option(dev, name) == good[0]
    finally:
        _lib.set_option(dev, name, old)
