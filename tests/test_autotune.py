"""The tracking autotuner's coordinate search (CPU, with a synthetic cost
model) and, on a GPU, one tuning run that applies its options."""
import numpy as np
import pytest

from ocelot_amd import autotune


def test_search_finds_coordinate_minimum():
    # cost: sort interval optimum at 7, per-XCD copies slower, overlap faster,
    # the wave pre-sum slower, the small deposit tile faster, records faster
    def cost(o):
        return ((o["sort_every"] - 7) ** 2 * 0.01 + 1.0 + 0.1 * o["deposit_xcd"] - 0.05 * o["overlap"]
                + 0.02 * o["deposit_agg"] - 0.03 * (o["deposit_tile"] == 1) - 0.04 * o["permute_rec"]
                - 0.01 * (o["fft_plan"] == 2) - 0.05 * o["sort_ahead"])
    calls = []

    def measure(o):
        calls.append(dict(o))
        return cost(o), None
    start = {"sort_every": 10, "sort_ahead": 0, "deposit_xcd": 1, "deposit_agg": 1, "deposit_tile": 0,
             "permute_rec": 0, "overlap": 0, "fft_plan": 0}
    best, timings, results = autotune.search(autotune.DEFAULT_SPACE, measure, start)
    assert best == {"sort_every": 7, "sort_ahead": 1, "deposit_xcd": 0, "deposit_agg": 0, "deposit_tile": 1,
                    "permute_rec": 1, "overlap": 1, "fft_plan": 2}
    assert set(timings) == set(autotune.DEFAULT_SPACE) == set(results)
    assert len(calls) == 1 + sum(len(v) for v in autotune.DEFAULT_SPACE.values())


def test_identity_check_compares_within_each_knobs_loop():
    """A bit-identical knob is checked between the results of its own loop
    (the other knobs held fixed), even when a later knob changes the best
    setting; a knob whose values differ is reported."""
    # results depend on sort_every (not bit-identical) and, wrongly, on overlap
    def measure(o):
        return 1.0 + 0.01 * (o["sort_every"] != 5), (o["sort_every"], o["overlap"] if o["overlap"] else 0)
    space = {"overlap": (0, 1), "sort_every": (5, 10)}
    _, _, results = autotune.search(space, measure, {"overlap": 0, "sort_every": 10})
    with pytest.raises(RuntimeError, match="overlap"):
        autotune.check_identical(results, ["overlap"])
    good = lambda o: (1.0 + 0.01 * (o["sort_every"] != 5), (o["sort_every"],))  # noqa: E731
    _, _, results = autotune.search(space, good, {"overlap": 0, "sort_every": 10})
    autotune.check_identical(results, ["overlap"])  # the later sort_every loop does not matter


def test_file_cache_round_trip(tmp_path, monkeypatch):
    """Entries are keyed by architecture, mesh, precision and the power-of-two
    bucket of the bunch size, written atomically, and read back."""
    from ocelot_amd import _lib
    path = str(tmp_path / "sub" / "autotune.json")
    monkeypatch.setenv("OCSC_AUTOTUNE_CACHE", path)
    assert autotune.cache_path() == path
    k1 = autotune.file_key("gfx950", (127, 127, 127), _lib.OCSC_MIXED, 10_000_000)
    k2 = autotune.file_key("gfx950", (127, 127, 127), _lib.OCSC_MIXED, 9_000_000)
    assert k1 == k2 == "gfx950|127x127x127|mixed|n2^23"
    assert autotune.file_key("gfx950", (63, 63, 63), _lib.OCSC_FP64, 2_000_000) == "gfx950|63x63x63|fp64|n2^21"
    autotune.store_file(k1, {"options": {"sort_every": 7}, "ms_per_step": 0.8})
    autotune.store_file("other", {"options": {}})
    data = autotune.load_file()
    assert data[k1]["options"] == {"sort_every": 7} and "other" in data
    monkeypatch.setenv("OCSC_AUTOTUNE_CACHE", "")
    assert autotune.cache_path() is None and autotune.load_file() == {}


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
    import os
    import tempfile
    os.environ["OCSC_AUTOTUNE_CACHE"] = os.path.join(tempfile.mkdtemp(), "autotune.json")
    try:
        res = autotune.tune_tracking(d, sc, dz, dz, steps=8)
        print("autotune:", res)
        assert not res["cached"] and res["ms_per_step"] > 0
        for k, v in res["options"].items():
            assert _lib.get_option(dev, k) == v
        again = autotune.tune_tracking(d, sc, dz, dz)
        assert again["cached"] and again["source"] == "memory" and again["options"] == res["options"]
        autotune._cache.clear()
        from_file = autotune.tune_tracking(d, sc, dz, dz)
        assert from_file["source"] == "file" and from_file["options"] == res["options"]
    finally:
        os.environ.pop("OCSC_AUTOTUNE_CACHE", None)
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
