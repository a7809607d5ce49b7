"""Bunch files in the reference's npz format (ocelot/cpbd/io.py:223-241,
262-292): ``rparticles`` (6, n), ``q_array`` (n,), ``E`` and ``s`` (0-d),
written with ``numpy.savez_compressed`` and read with ``numpy.load`` (no
pickles).  Files are interchangeable with the reference's
``save_particle_array`` / ``load_particle_array``.

Beyond the reference: ``device=k`` loads the bunch straight into HBM as a
``DeviceParticleArray`` (rows staged through pinned host memory and copied
asynchronously), and saving accepts a ``DeviceParticleArray``.  Only the npz
format is handled; ASTRA, CSRTrack and openPMD files raise ``ValueError``.
"""
import os

import numpy as np

from .beam import DeviceParticleArray, ParticleArray

_KEYS = ("rparticles", "q_array", "E", "s")


def _host_arrays(p_array):
    if isinstance(p_array, DeviceParticleArray):
        return (p_array.rparticles.cpu().numpy(), p_array.q_array.cpu().numpy(), p_array.E, p_array.s)
    return (p_array.rparticles, p_array.q_array, p_array.E, p_array.s)


def save_particle_array2npz(filename, p_array):
    """np.savez_compressed with the reference's four keys (io.py:223-226)."""
    r, q, E, s = _host_arrays(p_array)
    np.savez_compressed(filename, rparticles=r, q_array=q, E=E, s=s)


def load_particle_array_from_npz(filename, print_params=False, device=None):
    """ParticleArray with every key of the file set as an attribute, as the
    reference loads it (io.py:229-241; E and s stay 0-d arrays); with
    ``device``, a DeviceParticleArray in HBM instead."""
    with np.load(filename, allow_pickle=False) as data:
        fields = {key: data[key] for key in data.keys()}
    missing = [k for k in ("rparticles", "q_array") if k not in fields]
    if missing:
        raise ValueError(f"{filename}: not a particle-array file (missing {', '.join(missing)})")
    if device is None:
        p_array = ParticleArray()
        p_array.__dict__.update(fields)
    else:
        import torch
        dev = torch.device(f"cuda:{int(device)}")
        r = torch.from_numpy(np.ascontiguousarray(fields["rparticles"], dtype=np.float64)).pin_memory()
        q = torch.from_numpy(np.ascontiguousarray(fields["q_array"], dtype=np.float64)).pin_memory()
        with torch.cuda.device(dev):
            p_array = DeviceParticleArray(r.to(dev, non_blocking=True), q.to(dev, non_blocking=True),
                                          float(fields.get("E", 0.0)), float(fields.get("s", 0.0)))
            torch.cuda.current_stream(dev).synchronize()   # the pinned staging buffers are freed after this
    if print_params:
        print(p_array)
    return p_array


def load_particle_array(filename, print_params=False, device=None):
    """Load a bunch file by extension (io.py:244-275); npz only."""
    ext = os.path.splitext(filename)[1]
    if ext == ".npz":
        return load_particle_array_from_npz(filename, print_params=print_params, device=device)
    raise ValueError(f"unsupported beam file format {ext!r}: only *.npz is handled here")


def save_particle_array(filename, p_array):
    """Save a bunch file by extension (io.py:278-300); npz only."""
    ext = os.path.splitext(filename)[1]
    if ext == ".npz":
        return save_particle_array2npz(filename, p_array)
    raise ValueError(f"unsupported beam file format {ext!r}: only *.npz is handled here")
