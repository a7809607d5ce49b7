"""ocsc_apply_noalloc (the kick that allocates no device memory: every
workspace in one static device array per GPU) against ocsc_apply, through
the C ABI: the same rows bit for bit, on two calls in a row with the second
one smaller (fewer particles, smaller mesh), in both precisions; meshes
whose workspace does not fit are rejected before any work; the calls leave
the device's free memory unchanged."""
import ctypes

import numpy as np
import pytest

from ocelot_amd import _lib
from ocelot_amd.workloads import equal_charges, gaussian_bunch

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

E_GEV, DZ = 0.13, 0.1


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    lib = _lib.load()
    h = ctypes.c_void_p()
    _lib.check(lib.ocsc_create(ctypes.byref(h), torch.cuda.current_device()))
    yield lib, h
    torch.cuda.synchronize()
    lib.ocsc_destroy(h)


def bunch(n, seed):
    return gaussian_bunch(n, seed=seed), equal_charges(n, 250e-12)


def call(lib, h, fn, d, dq, nmesh, prec, sync=1):
    n = d.shape[1]
    return getattr(lib, fn)(h, ctypes.c_void_p(d.data_ptr()), n, ctypes.c_void_p(dq.data_ptr()), n,
                            _lib.i32x3(nmesh), E_GEV, DZ, 0, 1.0, 0.0, prec, sync, None)


def kick(lib, h, fn, r, q, nmesh, prec):
    d = torch.tensor(r, device="cuda")
    dq = torch.tensor(q, device="cuda")
    _lib.check(call(lib, h, fn, d, dq, nmesh, prec))
    torch.cuda.synchronize()
    return d.cpu().numpy()


@pytest.mark.parametrize("prec", [_lib.OCSC_FP64, _lib.OCSC_MIXED])
def test_noalloc_matches_apply_on_two_calls(ctx, prec):
    """A 200k-particle bunch on 63^3, then a 20k-particle bunch on 31^3: each
    call's rows equal ocsc_apply's bit for bit (the second call reuses the
    array the first one filled)."""
    lib, h = ctx
    for n, nmesh, seed in ((200_000, (63, 63, 63), 3), (20_000, (31, 31, 31), 4)):
        r, q = bunch(n, seed)
        got = kick(lib, h, "ocsc_apply_noalloc", r, q, nmesh, prec)
        ref = kick(lib, h, "ocsc_apply", r, q, nmesh, prec)
        print(f"prec {prec} n {n} mesh {nmesh}: max |noalloc - apply| = {np.max(np.abs(got - ref)):.3e}")
        assert np.array_equal(got, ref)
        assert not np.array_equal(got, r)


def test_noalloc_repeats_and_alternates_meshes(ctx):
    """Kicks alternating between two meshes and both precisions on one array
    (the kernel-spectrum cache in the array must follow): every result equals
    ocsc_apply's."""
    lib, h = ctx
    r, q = bunch(50_000, 5)
    for nmesh, prec in (((31, 31, 31), 0), ((47, 23, 63), 1), ((31, 31, 31), 0), ((31, 31, 31), 1),
                        ((47, 23, 63), 1)):
        got = kick(lib, h, "ocsc_apply_noalloc", r, q, nmesh, prec)
        ref = kick(lib, h, "ocsc_apply", r, q, nmesh, prec)
        assert np.array_equal(got, ref), (nmesh, prec)


def test_noalloc_rejects_meshes_beyond_the_array(ctx):
    """65^3 needs transforms of 256 points and more workspace than the array
    holds: OCSC_EINVAL with a message naming the limit, bunch untouched."""
    lib, h = ctx
    r, q = bunch(10_000, 6)
    d = torch.tensor(r, device="cuda")
    dq = torch.tensor(q, device="cuda")
    for prec in (_lib.OCSC_FP64, _lib.OCSC_MIXED):
        rc = call(lib, h, "ocsc_apply_noalloc", d, dq, (65, 65, 65), prec)
        assert rc == _lib.OCSC_EINVAL
        assert b"at most" in lib.ocsc_last_error()
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), r)


def test_noalloc_reports_invalid_bunch(ctx):
    """A bunch with a non-finite coordinate: OCSC_EVALUE as ocsc_apply, rows
    untouched; the next valid call succeeds."""
    lib, h = ctx
    r, q = bunch(10_000, 7)
    r[0, 17] = np.nan
    d = torch.tensor(r, device="cuda")
    dq = torch.tensor(q, device="cuda")
    assert call(lib, h, "ocsc_apply_noalloc", d, dq, (31, 31, 31), _lib.OCSC_FP64) == _lib.OCSC_EVALUE
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), r, equal_nan=True)
    r2, q2 = bunch(10_000, 8)
    got = kick(lib, h, "ocsc_apply_noalloc", r2, q2, (31, 31, 31), _lib.OCSC_FP64)
    ref = kick(lib, h, "ocsc_apply", r2, q2, (31, 31, 31), _lib.OCSC_FP64)
    assert np.array_equal(got, ref)


def test_noalloc_leaves_free_memory_unchanged():
    """On a fresh context (no workspaces yet), kicks on the largest meshes the
    array admits do not change the device's free memory, while ocsc_apply on
    the same context allocates its workspaces (about 30 MB at 63^3 FP64)."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    lib = _lib.load()
    h = ctypes.c_void_p()
    _lib.check(lib.ocsc_create(ctypes.byref(h), torch.cuda.current_device()))
    try:
        r, q = bunch(100_000, 9)
        d = torch.tensor(r, device="cuda")
        dq = torch.tensor(q, device="cuda")
        cases = (((64, 64, 64), _lib.OCSC_FP64), ((63, 63, 63), _lib.OCSC_MIXED))
        # a first round launches these kernels once on this device (the
        # runtime loads code objects and sizes its own per-queue resources)
        for nmesh, prec in cases:
            _lib.check(call(lib, h, "ocsc_apply_noalloc", d, dq, nmesh, prec))
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        for _ in range(2):
            for nmesh, prec in cases:
                _lib.check(call(lib, h, "ocsc_apply_noalloc", d, dq, nmesh, prec))
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        _lib.check(call(lib, h, "ocsc_apply", d, dq, (64, 64, 64), _lib.OCSC_FP64))
        torch.cuda.synchronize()
        free2 = torch.cuda.mem_get_info()[0]
        print(f"free memory: {free0} -> {free1} after ocsc_apply_noalloc, {free2} after ocsc_apply")
        assert free1 == free0
        assert free0 - free2 >= 20 << 20  # ocsc_apply's own workspaces
    finally:
        torch.cuda.synchronize()
        lib.ocsc_destroy(h)
