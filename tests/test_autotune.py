"""The tracking autotuner's coordinate search (CPU, with a synthetic cost
model) and, on a GPU, one tuning run that applies its options."""
import numpy as np
import pytest

from ocelot_amd import autotune


def test_search_finds_coordinate_minimum():
    # cost: sort interval optimum at 7, per-XCD copies slower, overlap faster
    def cost(o):
        return (o["sort_every"] - 7) ** 2 * 0.01 + 1.0 + 0.1 * o["deposit_xcd"] - 0.05 * o["overlap"]
    calls = []

    def measure(o):
        calls.append(dict(o))
        return cost(o), None
    best, timings, _ = autotune.search(autotune.DEFAULT_SPACE, measure, {"sort_every": 10, "deposit_xcd": 1,
                                                                        "overlap": 0})
    assert best == {"sort_every": 7, "deposit_xcd": 0, "overlap": 1}
    assert set(timings) == set(autotune.DEFAULT_SPACE)
    assert len(calls) == 1 + sum(len(v) for v in autotune.DEFAULT_SPACE.values())


def test_search_keeps_start_when_nothing_is_faster():
    best, _, _ = autotune.search({"a": (1, 2)}, lambda o: (1.0 + (o["a"] != 3), None), {"a": 3})
    assert best == {"a": 3}


@pytest.mark.gpu
def test_tune_tracking_on_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from ocelot_amd import DeviceParticleArray, ParticleArray, SpaceCharge, _lib
    from ocelot_amd.workloads import make_bunch
    r, q, E, _, dz = make_bunch("flattop_10m_127", n=300_000, seed=5)
    p = ParticleArray()
    p.rparticles, p.q_array, p.E = r, q, E
    dev = torch.cuda.current_device()
    d = DeviceParticleArray.from_host(p, dev)
    sc = SpaceCharge(nmesh_xyz=[31, 31, 63], precision="mixed")
    old = {k: _lib.get_option(dev, k) for k in autotune.DEFAULT_SPACE}
    try:
        res = autotune.tune_tracking(d, sc, dz, dz, steps=8)
        print("autotune:", res)
        assert not res["cached"] and res["ms_per_step"] > 0
        for k, v in res["options"].items():
            assert _lib.get_option(dev, k) == v
        again = autotune.tune_tracking(d, sc, dz, dz)
        assert again["cached"] and again["options"] == res["options"]
    finally:
        for k, v in old.items():
            _lib.set_option(dev, k, v)


@pytest.mark.gpu
def test_spacecharge_autotune_keyword_in_tracking():
    """SpaceCharge(autotune=True): track_drift tunes first (options set, result
    cached), takes exactly the random-mesh draws of its own kicks from the
    global RNG, and tracks like the untuned process up to reduction order."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import numpy as np
    from ocelot_amd import DeviceParticleArray, ParticleArray, SpaceCharge, _lib, track_drift
    from ocelot_amd.workloads import make_bunch
    r, q, E, _, dz = make_bunch("flattop_10m_127", n=200_000, seed=11)
    p = ParticleArray()
    p.rparticles, p.q_array, p.E = r, q, E
    dev = torch.cuda.current_device()
    old = {k: _lib.get_option(dev, k) for k in autotune.DEFAULT_SPACE}
    try:
        outs, states = [], []
        for tune in (False, True):
            sc = SpaceCharge(nmesh_xyz=[31, 31, 47], precision="mixed", random_mesh=True, autotune=tune)
            np.random.seed(3)
            d = DeviceParticleArray.from_host(p, dev)
            track_drift(d, sc, dz, dz, 6)
            outs.append(d.rparticles.cpu().numpy())
            states.append(np.random.get_state()[1].copy())
        np.testing.assert_array_equal(states[0], states[1])        # same 12 draws consumed
        key = (200_000, (31, 31, 47), _lib.OCSC_MIXED, dev)
        assert key in autotune._cache
        for k, v in autotune._cache[key]["options"].items():
            assert _lib.get_option(dev, k) == v
        a, b = outs
        for i in range(6):
            assert np.abs(a[i] - b[i]).max() <= 1e-8 * np.abs(a[i]).max()
    finally:
        for k, v in old.items():
            _lib.set_option(dev, k, v)
