"""The cell sort of the tracking path (hand-written stable LSD radix sort,
8 bits per pass) against numpy.argsort(kind="stable"), through the C ABI
(ocsc_cell_sort): random keys, keys with few distinct values, nearly sorted
keys, all-equal keys; sizes around the 4096-key tile; 1 to 4 passes."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from ocelot_amd import _lib
    return _lib


def gpu_order(_lib, keys, end_bit):
    dev = torch.cuda.current_device()
    k = torch.tensor(keys.astype(np.int64).astype(np.uint32).view(np.int32), device="cuda")
    out = torch.empty(max(len(keys), 1), dtype=torch.int32, device="cuda")
    with _lib.using(dev) as h:
        _lib.check(_lib.load().ocsc_cell_sort(h, ctypes.c_void_p(k.data_ptr()), len(keys), end_bit,
                                              ctypes.c_void_p(out.data_ptr()), None))
        torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32)[:len(keys)]


def check(_lib, keys, end_bit):
    got = gpu_order(_lib, keys, end_bit)
    ref = np.argsort(keys, kind="stable").astype(np.uint32)
    bad = int(np.sum(got != ref))
    print(f"n={len(keys)} end_bit={end_bit}: {bad} positions differ")
    assert bad == 0


@pytest.mark.parametrize("n", [1, 2, 63, 4095, 4096, 4097, 70_001, 1_000_003])
@pytest.mark.parametrize("end_bit", [8, 14, 22, 32])
def test_random_keys(lib, n, end_bit):
    rng = np.random.default_rng(n + end_bit)
    hi = (1 << end_bit) if end_bit < 32 else (1 << 32) - 1
    check(lib, rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32), end_bit)


@pytest.mark.parametrize("n", [5000, 300_000])
def test_few_distinct_keys(lib, n):
    """Ties everywhere: stability decides the whole order."""
    rng = np.random.default_rng(7)
    check(lib, rng.integers(0, 5, n).astype(np.uint32) * 4099, 22)


def test_nearly_sorted_keys(lib):
    """Brick keys of a bunch that drifted a little since its last sort."""
    rng = np.random.default_rng(8)
    keys = np.sort(rng.integers(0, 1 << 21, 2_000_000)).astype(np.int64)
    keys += rng.integers(-40, 41, keys.size) * (rng.random(keys.size) < 0.2)
    check(lib, np.clip(keys, 0, (1 << 21) - 1).astype(np.uint32), 22)


def test_equal_keys_and_sentinel(lib):
    """All keys equal (the identity order), and the largest key value."""
    check(lib, np.full(9000, 123, dtype=np.uint32), 8)
    check(lib, np.full(9000, 0xFFFFFFFF, dtype=np.uint32), 32)


def test_rejects_bad_arguments(lib):
    dev = torch.cuda.current_device()
    with lib.using(dev) as h:
        assert lib.load().ocsc_cell_sort(h, None, 10, 8, None, None) == lib.OCSC_EINVAL
        k = torch.zeros(4, dtype=torch.int32, device="cuda")
        assert lib.load().ocsc_cell_sort(h, ctypes.c_void_p(k.data_ptr()), 4, 33, ctypes.c_void_p(k.data_ptr()),
                                         None) == lib.OCSC_EINVAL
