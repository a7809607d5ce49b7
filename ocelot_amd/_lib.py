"""ctypes binding of libocsc.so (include/ocsc.h).

The library is the product path: there is no CPU fallback.  If it is missing
or fails to load, every entry point raises.
"""
import contextlib
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("OCSC_LIB", os.path.join(HERE, "libocsc.so"))

OCSC_OK = 0
OCSC_EINVAL = 1
OCSC_EVALUE = 2
OCSC_EHIP = 3
OCSC_ENOMEM = 4
OCSC_FP64 = 0
OCSC_MIXED = 1

_c_int = ctypes.c_int
_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)

# name -> (restype, argtypes); the list every header declaration must match
SIGNATURES = {
    "ocsc_create": (_c_int, [ctypes.POINTER(_vp), _c_int]),
    "ocsc_destroy": (_c_int, [_vp]),
    "ocsc_last_error": (ctypes.c_char_p, []),
    "ocsc_apply": (_c_int, [_vp, _vp, _i64, _vp, _i64, _i32p, _dbl, _dbl, _c_int, _dbl, _dbl, _c_int, _c_int, _vp]),
    "ocsc_apply_cells": (_c_int, [_vp, _vp, _i64, _vp, _i64, _i32p, _dbl, _dbl, _c_int, _dbl, _dbl, _c_int, _c_int,
                                  _vp, _vp]),
    "ocsc_apply_noalloc": (_c_int, [_vp, _vp, _i64, _vp, _i64, _i32p, _dbl, _dbl, _c_int, _dbl, _dbl, _c_int, _c_int,
                                    _vp]),
    "ocsc_check": (_c_int, [_vp, _vp]),
    "ocsc_el_field": (_c_int, [_vp, _vp, _i64, _vp, _i64, _i32p, _dbl, _c_int, _dbl, _dbl, _c_int, _vp, _i64, _vp,
                               _vp]),
    "ocsc_sym_kernel": (_c_int, [_vp, _i32p, _dp, _vp, _vp]),
    "ocsc_potential": (_c_int, [_vp, _vp, _i32p, _dp, _vp, _c_int, _vp]),
    "ocsc_drift": (_c_int, [_vp, _vp, _i64, _i64, _dbl, _dbl, _vp]),
    "ocsc_track_drift": (_c_int, [_vp, _vp, _i64, _vp, _i64, _i32p, _dbl, _dbl, _dbl, _c_int, _c_int, _c_int, _vp]),
    "ocsc_track_drift_tws": (_c_int, [_vp, _vp, _i64, _vp, _i64, _i32p, _dbl, _dbl, _dbl, _c_int, _c_int, _c_int,
                                      _c_int, _dp, _dbl, _dbl, _dp, _vp]),
    "ocsc_last_stats": (_c_int, [_vp, _dp, _c_int, _vp]),
    "ocsc_apply_batched": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp]),
    "ocsc_moments": (_c_int, [_vp, _vp, _i64, _vp, _i64, _c_int, _dp, _dbl, _dbl, _dp, _c_int, _vp]),
    "ocsc_set_option": (_c_int, [_vp, ctypes.c_char_p, _i64]),
    "ocsc_get_option": (_c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(_i64)]),
    "ocsc_cell_sort": (_c_int, [_vp, _vp, _i64, _c_int, _vp, _vp]),
    "ocsc_set_profiling": (_c_int, [_vp, _c_int]),
    "ocsc_stage_times": (_c_int, [_vp, _dp, _c_int, ctypes.POINTER(_i64)]),
    "ocsc_stage_name": (ctypes.c_char_p, [_c_int]),
    "ocsc_kick_times": (_c_int, [_vp, _dp, _c_int, ctypes.POINTER(_i64)]),
    "ocsc_lsc_apply": (_c_int, [_vp, _vp, _i64, _vp, _i64, _dbl, _dbl, _c_int, _dbl, _dbl, _dbl, _dbl, _dbl, _dp,
                                _c_int, _vp]),
    "ocsc_lsc_profile": (_c_int, [_vp, _dp, _dp, _dp, _i64, ctypes.POINTER(_i64)]),
    "ocsc_lsc_wake": (_c_int, [_vp, _dp, _dp, _i64, _dbl, _dbl, _dbl, _c_int, _dbl, _dbl, _dp, _vp]),
    "ocsc_lsc_impedance": (_c_int, [_vp, _dp, _i64, _dbl, _dbl, _dbl, _c_int, _dp, _vp]),
    "ocsc_track": (_c_int, [_vp, _vp, _i64, _vp, _i64, _i32p, _dbl, _dbl, _dbl, _c_int, _c_int, _vp, _vp]),
    "ocsc_track_drift_random_mesh": (_c_int, [_vp, _vp, _i64, _vp, _i64, _i32p, _dbl, _dbl, _dbl, _c_int, _c_int,
                                              _c_int, _dp, _vp]),
    "ocsc_track_drift_lsc": (_c_int, [_vp, _vp, _i64, _vp, _i64, _i32p, _dbl, _dbl, _dbl, _c_int, _c_int, _c_int,
                                      _dbl, _c_int, _dbl, _dbl, _dbl, _dp, _vp]),
    "ocsc_csr_convolution": (_c_int, [_vp, _dp, _i64, _dp, _i64, _dp, _vp]),
    "ocsc_csr_kick": (_c_int, [_vp, _vp, _i64, _i64, _dbl, _dp, _i64, _dp, _i64, _dbl, _dbl, _dbl, _dbl, _dp, _vp]),
}

_lib = None
_lock = threading.Lock()


def load(path=None):
    """Load libocsc.so once and declare the C signatures."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise RuntimeError(f"libocsc.so not found at {p}: build it with `python -m ocelot_amd.build` "
                               "(the space-charge kick has no CPU fallback)")
        lib = ctypes.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


class OcscError(RuntimeError):
    pass


def check(rc):
    """Raise for a non-zero status: ValueError where the reference raises
    ValueError (OCSC_EVALUE), RuntimeError otherwise."""
    if rc == OCSC_OK:
        return
    msg = load().ocsc_last_error().decode(errors="replace")
    if rc == OCSC_EVALUE:
        raise ValueError(msg)
    if rc == OCSC_EINVAL:
        raise ValueError("invalid argument: " + msg)
    raise OcscError(f"ocsc error {rc}: {msg}")


class TrackOpts(ctypes.Structure):
    """ocsc_track_opts of include/ocsc.h"""
    _fields_ = [("use_graph", ctypes.c_int), ("tws_out", _dp), ("mom_flags", ctypes.c_int),
                ("mom_b0", ctypes.c_double), ("mom_b1", ctypes.c_double), ("disp", ctypes.c_double * 4),
                ("lsc", ctypes.c_int), ("lsc_dz", ctypes.c_double), ("lsc_step_profile", ctypes.c_int),
                ("lsc_smooth_param", ctypes.c_double), ("lsc_b0", ctypes.c_double), ("lsc_b1", ctypes.c_double),
                ("lsc_und_factors", _dp), ("mesh_draws", _dp)]


class Bunch(ctypes.Structure):
    """ocsc_bunch of include/ocsc.h"""
    _fields_ = [("rparticles", _vp), ("ld", _i64), ("q_array", _vp), ("n", _i64),
                ("nmesh", ctypes.c_int32 * 3), ("E_GeV", _dbl), ("zstep", _dbl),
                ("random_mesh", ctypes.c_int32), ("mesh_scale", _dbl), ("mesh_shift", _dbl)]


def i32x3(v):
    a = (ctypes.c_int32 * 3)(*[int(x) for x in v])
    return a


def d3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


class Context:
    """One library context (device workspaces and caches) per device.  A
    context must not be used by two host threads at once (ctypes releases the
    GIL during a call), so every use holds ``lock`` from the first library
    call to the last host copy that depends on it (see ``using``)."""

    def __init__(self, device):
        lib = load()
        h = _vp()
        check(lib.ocsc_create(ctypes.byref(h), int(device)))
        self.handle = h
        self.device = int(device)
        self.lock = threading.RLock()

    def __del__(self):
        try:
            if self.handle:
                load().ocsc_destroy(self.handle)
        except Exception:
            pass


def set_option(device, name, value):
    with using(device) as h:
        check(load().ocsc_set_option(h, name.encode(), int(value)))


def get_option(device, name):
    v = _i64()
    with using(device) as h:
        check(load().ocsc_get_option(h, name.encode(), ctypes.byref(v)))
    return v.value


_contexts = {}


def context(device):
    """Process-wide context of a device (created lazily, never pickled)."""
    device = int(device)
    with _lock:
        ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        with _lock:
            # two threads may race to create it: the first one stored wins and
            # the other context is released when it goes out of scope
            ctx = _contexts.setdefault(device, ctx)
    return ctx


@contextlib.contextmanager
def using(device):
    """The handle of the device's context, used by this thread alone for the
    duration of the block (calls from other threads on the same device wait)."""
    ctx = context(device)
    with ctx.lock:
        yield ctx.handle


def check_charges(q, n):
    """The charges must be one value per particle: the reference fails with
    ValueError in np.dot(Q, X) otherwise (ocelot/cpbd/sc.py:182); the kernels
    would read past a short array."""
    shape = tuple(q.shape) if hasattr(q, "shape") else np.shape(q)
    if shape != (int(n),):
        raise ValueError(f"q_array must hold one charge per particle: shape {shape}, {int(n)} particles")
