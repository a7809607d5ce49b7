"""Autotuner for the tracking loop: picks, per bunch size, mesh and precision,
the context options that only change speed, and keeps the choices in a file.

    from ocelot_amd.autotune import tune_tracking
    best = tune_tracking(d_array, sc, L=0.1, dz=0.1)     # sets the options

Knobs (``ocsc_set_option``):

* ``sort_every``: re-sort interval of the cell-ordered copy (a sort costs a
  radix sort and a permutation of the bunch, and buys deposit/gather
  locality that decays as particles drift);
* ``sort_ahead``: sort keys at the drifted position half-way through the
  interval, or at the present one;
* ``deposit_xcd``: one charge grid, or one per XCD with L2-local atomics;
* ``deposit_agg``: runs of equal cells summed across the wavefront before
  the deposit's LDS table;
* ``deposit_tile``: particles per thread and LDS hash slots of the deposit
  (0: 8 x 2048, 1: 4 x 1024, 2: 16 x 4096, 4: as 0 with every particle's
  loads issued first);
* ``permute_rec``: permutations of the tracking copy through 64-byte
  particle records or row by row;
* ``overlap``: Green's-function spectrum on a second stream next to the
  deposit;
* ``fft_plan``: the transform plan (0: LDS Stockham passes, radix 16 where
  every thread gets a butterfly, else 8; 1: radix 8 in every pass; 2: half as
  many columns per workgroup, less LDS and more workgroups; 3: the two field
  passes with 256 threads, radix 16 in every pass (1 to 3 in the mixed form
  only); 4, the default: register-resident transforms, one LDS transpose per
  transform, in every pass whose length is 16 to 512).

The search is coordinate-wise: each knob in turn over its candidates with the
others fixed, timing ``steps`` graph-captured steps on a copy of the bunch.
Every knob but ``sort_every``, ``sort_ahead`` and ``fft_plan`` (which change
reduction orders and FP32 round-off) must leave the particles bit-identical; this is
checked between the values of each knob's own loop (the other knobs held as
they were during that loop).  Choices are kept in memory per (n, nmesh,
precision, device) and in a JSON file keyed by GPU architecture, mesh,
precision and bunch-size bucket (a power of two), so another process on the
same kind of GPU re-applies them without measuring: ``$OCSC_AUTOTUNE_CACHE``
or ``~/.cache/ocelot_amd/autotune.json`` (``OCSC_AUTOTUNE_CACHE=""`` turns
the file off).

``SpaceCharge(autotune=True)`` runs this at the start of ``track_drift``.
"""
import copy
import json
import math
import os
import tempfile
import threading
import time

from . import _lib

DEFAULT_SPACE = {"sort_every": (5, 7, 10, 14, 20), "sort_ahead": (0, 1), "deposit_xcd": (0, 1),
                 "deposit_agg": (0, 1), "deposit_tile": (0, 1, 2, 4), "permute_rec": (0, 1), "overlap": (0, 1),
                 "fft_plan": (0, 1, 2, 3, 4)}
BIT_IDENTICAL = ("deposit_xcd", "deposit_agg", "deposit_tile", "permute_rec", "overlap")

_cache = {}
_file_lock = threading.Lock()


def cache_path():
    """The autotune file (None when switched off with OCSC_AUTOTUNE_CACHE="")."""
    p = os.environ.get("OCSC_AUTOTUNE_CACHE")
    if p is None:
        return os.path.join(os.path.expanduser("~"), ".cache", "ocelot_amd", "autotune.json")
    return p or None


def file_key(arch, nmesh, prec, n):
    """Key of a tuned shape in the file: architecture, mesh, precision and
    the bunch-size bucket round(log2 n)."""
    bucket = int(round(math.log2(max(int(n), 1))))
    return f"{arch}|{'x'.join(str(int(v)) for v in nmesh)}|{'mixed' if prec == _lib.OCSC_MIXED else 'fp64'}|n2^{bucket}"


def load_file(path=None):
    path = path or cache_path()
    if not path or not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def store_file(key, entry, path=None):
    """Add one entry to the file (read, update, write to a temporary file in
    the same directory, rename: concurrent writers never leave a torn file)."""
    path = path or cache_path()
    if not path:
        return
    with _file_lock:
        data = load_file(path)
        data[key] = entry
        d = os.path.dirname(os.path.abspath(path))
        os.makedirs(d, exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=d, prefix=".autotune.", suffix=".json")
        with os.fdopen(fd, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def search(space, measure, start):
    """Coordinate search.  ``measure(options) -> (seconds, result)``; returns
    (best options, {knob: {value: seconds}}, {knob: {value: result}}): the
    results of each knob's own loop, where only that knob varies."""
    best = dict(start)
    t_best, _ = measure(best)
    timings, results = {}, {}
    for knob, values in space.items():
        timings[knob], results[knob] = {}, {}
        for v in values:
            opts = dict(best, **{knob: v})
            t, res = measure(opts)
            timings[knob][v] = t
            results[knob][v] = res
            if t < t_best:
                best, t_best = opts, t
    return best, timings, results


def check_identical(results, knobs=BIT_IDENTICAL):
    """Raise if a bit-identical knob's values gave different particles."""
    for knob in knobs:
        outs = list(results.get(knob, {}).values())
        for o in outs[1:]:
            if o != outs[0]:
                raise RuntimeError(f"option {knob} changed the tracked particles")


def _arch(dev):
    import torch
    return torch.cuda.get_device_properties(dev).gcnArchName.split(":")[0]


def tune_tracking(p_array, sc, L, dz, steps=30, space=None, check_bits=True, force=False, use_file=True):
    """Tune and apply the tracking options for this bunch and mesh; returns
    {"options": ..., "ms_per_step": ..., "timings_ms": ..., "cached": bool,
    "source": "memory" | "file" | "measured"}."""
    import torch
    from .beam import DeviceParticleArray
    from .sc import _nmesh_of, _prec_of, track_drift
    r = p_array.rparticles
    if not (isinstance(r, torch.Tensor) and r.is_cuda):
        raise TypeError("tune_tracking needs a DeviceParticleArray")
    dev = r.device.index
    n, nmesh, prec = int(r.shape[1]), tuple(_nmesh_of(sc)), _prec_of(sc)
    space = dict(DEFAULT_SPACE if space is None else space)
    key = (n, nmesh, prec, dev)
    fkey = file_key(_arch(dev), nmesh, prec, n)

    def apply(out, source):
        for k, v in out["options"].items():
            _lib.set_option(dev, k, v)
        return dict(out, cached=source != "measured", source=source)

    if not force:
        if key in _cache:
            return apply(_cache[key], "memory")
        entry = load_file().get(fkey) if use_file else None
        if entry and set(entry.get("options", {})) >= set(space):
            entry = {**entry, "options": {k: entry["options"][k] for k in space}}
            _cache[key] = entry
            return apply(entry, "file")
    start = {k: _lib.get_option(dev, k) for k in space}
    # the trial runs take no random-mesh draws from the global RNG (the
    # caller's kicks keep their draws) and do not tune again
    sc = copy.copy(sc)
    sc.random_mesh = False
    sc.autotune = False

    def measure(opts):
        for k, v in opts.items():
            _lib.set_option(dev, k, v)
        d = DeviceParticleArray(r.clone(), p_array.q_array, p_array.E, getattr(p_array, "s", 0.0))
        track_drift(d, sc, L, dz, 3, use_graph=True)  # first use of buffers and graphs
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        track_drift(d, sc, L, dz, steps, use_graph=True)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / steps
        bits = d.rparticles.contiguous().view(torch.int64)
        w = torch.arange(1, bits.shape[1] + 1, device=bits.device, dtype=torch.int64)
        # fingerprint of the exact bit patterns (wrapping integer sums)
        return t, (int(bits.sum()), int((bits * w).sum()))

    best, timings, results = search(space, measure, start)
    if check_bits:
        check_identical(results, [k for k in BIT_IDENTICAL if k in space])
    out = {"options": best,
           "ms_per_step": round(min(t for tm in timings.values() for t in tm.values()) * 1e3, 4),
           "timings_ms": {k: {str(v): round(t * 1e3, 4) for v, t in tm.items()} for k, tm in timings.items()},
           "steps": steps, "particles": n}
    _cache[key] = out
    if use_file:
        store_file(fkey, out)
    return apply(out, "measured")
