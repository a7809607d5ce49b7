"""The tail of Ocelot's CSR kick on the MI355X kernels.

``csr_convolution(a, b)`` is the reference's function (ocelot/cpbd/csr.py:
69-79): the first len(a) + len(b) - 2 terms of the linear convolution.
``csr_kick(p_array, lam_ds, K1, s1, s2, Ns, delta_s)`` is the end of
``CSR.apply`` (csr.py:1105-1143): given the equidistant line charge and mesh
that the reference's binning returns and the step-averaged CSR kernel K1, it
convolves them, scales by delta_s/st and adds the interpolated energy change
to every particle's delta, on the device (no sort: the interpolation is per
particle).  The trajectory, the kernel integration along it and the binning
stay with the caller (they depend on the lattice's bends and undulators).
"""
import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


def _dev(device):
    return _torch().cuda.current_device() if device is None else int(device)


def csr_convolution(a, b, device=None):
    lib = _lib.load()
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    b = np.ascontiguousarray(np.asarray(b, dtype=np.float64).reshape(-1))
    out = np.empty(max(a.size + b.size - 2, 0))
    dev = _dev(device)
    torch = _torch()
    with torch.cuda.device(dev), _lib.using(dev) as h:
        _lib.check(lib.ocsc_csr_convolution(h, a.ctypes.data_as(_lib._dp), a.size,
                                            b.ctypes.data_as(_lib._dp), b.size, out.ctypes.data_as(_lib._dp),
                                            torch.cuda.current_stream(dev).cuda_stream))
    return out


def csr_kick(p_array, lam_ds, K1, s1, s2, Ns, delta_s, device=None):
    """Kick p_array in place; returns lam_K1 (the wake on the mesh)."""
    torch = _torch()
    lib = _lib.load()
    lam = np.ascontiguousarray(np.asarray(lam_ds, dtype=np.float64).reshape(-1))
    k1 = np.ascontiguousarray(np.asarray(K1, dtype=np.float64).reshape(-1))
    out = np.empty(max(lam.size + k1.size - 2, 0))
    r = p_array.rparticles
    n = int(r.shape[1])
    args = (lam.ctypes.data_as(_lib._dp), lam.size, k1.ctypes.data_as(_lib._dp), k1.size, float(s1), float(s2),
            float(Ns), float(delta_s), out.ctypes.data_as(_lib._dp))
    if isinstance(r, torch.Tensor) and r.is_cuda:
        if r.dtype != torch.float64 or r.stride(1) != 1 or r.shape[0] != 6:
            raise TypeError("device rparticles must be a (6, n) float64 tensor with contiguous rows")
        dev = r.device.index
        with torch.cuda.device(dev), _lib.using(dev) as h:
            _lib.check(lib.ocsc_csr_kick(h, r.data_ptr(), r.stride(0), n, float(p_array.E),
                                         *args, torch.cuda.current_stream(dev).cuda_stream))
        return out
    dev = _dev(device)
    with torch.cuda.device(dev), _lib.using(dev) as h:
        rt = torch.from_numpy(np.ascontiguousarray(r, dtype=np.float64)).to(f"cuda:{dev}")
        _lib.check(lib.ocsc_csr_kick(h, rt.data_ptr(), n, n, float(p_array.E), *args,
                                     torch.cuda.current_stream(dev).cuda_stream))
        r[5] = rt[5].cpu().numpy()
    return out
