"""Drop-in ``LSC`` (longitudinal space charge) physics process on the MI355X
kernels.

Same surface as the reference class (ocelot/cpbd/sc.py:261-600): constructor
``LSC(step=1, step_profile=False, smooth_param=0.1, bounds=[-0.4, 0.4],
slice=None)``, ``prepare(lat)`` building the undulator strength profile
``K_s_func`` of the line, ``compute_filling_factor``, ``apply(p_array, dz)``
adding the LSC energy kick to ``p_array.rparticles[5]`` in place, and the
helpers ``imp_lsc``, ``imp_step_lsc``, ``wake_lsc``.  One extra keyword
argument, ``device`` (GPU index; default: the device of a resident bunch or
the current torch device).

``apply`` runs on the GPU through ``ocsc_lsc_apply``: slice statistics,
current profile, wake and the interpolated kick; the host only forms the
undulator factor (K_max, filling factor) from ``K_s_func`` as the reference
does.  ``p_array.rparticles`` may be a numpy array (uploaded, kicked, delta
copied back into the same ndarray) or a CUDA float64 tensor (kicked in place).

Errors: ValueError where the reference raises (empty bunch, no spread or
non-finite tau, empty central slice with the step profile).
"""
import ctypes

import numpy as np

from . import _lib
from .physproc import PhysProc


def _torch():
    import torch
    return torch


def _is_cuda_tensor(a):
    try:
        torch = _torch()
    except ImportError:
        return False
    return isinstance(a, torch.Tensor) and a.is_cuda


def _device_of(proc, tensor=None):
    if tensor is not None:
        return tensor.device.index
    dev = getattr(proc, "device", None)
    if dev is not None:
        return int(dev)
    return _torch().cuda.current_device()


def _filling_factor(K_s_func, x0, x1, num_points):
    x = np.linspace(x0, x1, num_points)
    k_vals = K_s_func(x)
    nonzero_intervals = (k_vals[:-1] != 0) | (k_vals[1:] != 0)
    dx = (x1 - x0) / (num_points - 1)
    return np.sum(nonzero_intervals) * dx / (x1 - x0)


def _kick_args(proc, dz):
    if getattr(proc, "_is_undul_in_beam_line", False):
        fill_factor = _filling_factor(proc.K_s_func, proc.z0 - dz, proc.z0, 200)
        K_max = np.max(proc.K_s_func(np.linspace(proc.z0 - dz, proc.z0, num=200)))
        return float(K_max), float(fill_factor)
    return 0.0, 0.0


def _is_undulator(elem):
    return any(k.__name__ == "Undulator" for k in type(elem).__mro__)


class LSC(PhysProc):
    """Longitudinal space-charge kick from the 1D impedance of a round beam
    (Gaussian or uniform transverse profile) acting on the current profile."""

    def __init__(self, step=1, **kwargs):
        PhysProc.__init__(self, step)
        self.K_s_func = None
        self.step_profile = kwargs.get("step_profile", False)
        self.smooth_param = kwargs.get("smooth_param", 0.1)
        self.bounds = kwargs.get("bounds", [-0.4, 0.4])
        self.slice = kwargs.get("slice", None)
        self._is_undul_in_beam_line = False
        self.device = kwargs.get("device", None)

    # ------------------------------------------------------------------ helpers
    def _device(self, tensor=None):
        return _device_of(self, tensor)

    def imp_lsc(self, gamma, sigma, w, dz):
        """Impedance of a round Gaussian beam of rms size sigma over dz
        (sc.py:302-341), complex numpy array (purely imaginary)."""
        return 1j * self._impedance(gamma, sigma, w, dz, 0)

    def imp_step_lsc(self, gamma, rb, w, dz):
        """Impedance of a uniform round beam of radius rb over dz
        (sc.py:343-369), complex numpy array (purely imaginary)."""
        return 1j * self._impedance(gamma, rb, w, dz, 1)

    def _impedance(self, gamma, sigma, w, dz, step):
        lib = _lib.load()
        w = np.ascontiguousarray(np.asarray(w, dtype=np.float64).reshape(-1))
        out = np.empty_like(w)
        dev = _device_of(self)
        torch = _torch()
        with torch.cuda.device(dev), _lib.using(dev) as h:
            _lib.check(lib.ocsc_lsc_impedance(h, w.ctypes.data_as(_lib._dp), w.size,
                                              float(gamma), float(sigma), float(dz), int(step),
                                              out.ctypes.data_as(_lib._dp), torch.cuda.current_stream(dev).cuda_stream))
        return out

    def wake_lsc(self, s, bunch, gamma, sigma, dz, K_max=0, fill_factor=0):
        """Wake (V/C) of the normalised profile bunch on the uniform grid s
        (sc.py:411-465), the impedance chosen by ``step_profile``."""
        lib = _lib.load()
        s = np.ascontiguousarray(np.asarray(s, dtype=np.float64))
        b = np.ascontiguousarray(np.asarray(bunch, dtype=np.float64))
        if s.shape != b.shape or s.ndim != 1:
            raise ValueError("s and bunch must be 1D arrays of one length")
        out = np.empty_like(b)
        dev = _device_of(self)
        torch = _torch()
        with torch.cuda.device(dev), _lib.using(dev) as h:
            _lib.check(lib.ocsc_lsc_wake(h, s.ctypes.data_as(_lib._dp),
                                         b.ctypes.data_as(_lib._dp), b.size, float(gamma), float(sigma), float(dz),
                                         int(bool(self.step_profile)), float(K_max), float(fill_factor),
                                         out.ctypes.data_as(_lib._dp), torch.cuda.current_stream(dev).cuda_stream))
        return out

    # ------------------------------------------------------------------ PhysProc
    def prepare(self, lat):
        """Undulator strength along the line, K(s) = Kx + Ky inside undulators
        that are not both-plane, 0 elsewhere, linear between the element
        edges (sc.py:467-503)."""
        from scipy.interpolate import interp1d
        self.check_step()
        seq = lat.get_sequence_part(self.start_elem, self.end_elem)
        s, k = [], []
        s_current = self.s_start
        for elem in seq:
            s_next = s_current + elem.l
            s.extend([s_current, s_next])
            if _is_undulator(elem) and not (elem.Kx != 0 and elem.Ky != 0):
                k.extend([elem.Kx + elem.Ky, elem.Kx + elem.Ky])
                self._is_undul_in_beam_line = True
            else:
                k.extend([0.0, 0.0])
            s_current = s_next
        self.K_s_func = interp1d(s, k, kind='linear', bounds_error=False, fill_value=(k[0], k[-1]))

    def compute_filling_factor(self, x0, x1, num_points=100):
        """Fraction of [x0, x1] on num_points samples where K is nonzero at
        either end of a sample interval (sc.py:505-536)."""
        return _filling_factor(self.K_s_func, x0, x1, num_points)

    def kick_args(self, dz):
        """(K_max, fill_factor) of the step ending at z0 (sc.py:559-564)."""
        return _kick_args(self, dz)

    def apply(self, p_array, dz):
        """LSC energy kick on rparticles[5] in place (sc.py:538-600).  Reads
        only attributes the reference class defines, so it also runs on
        Ocelot's own LSC instance."""
        if dz < 1e-10:
            return
        K_max, fill_factor = _kick_args(self, dz)
        torch = _torch()
        lib = _lib.load()
        r = p_array.rparticles
        n = int(r.shape[1])
        _lib.check_charges(p_array.q_array, n)
        b0, b1 = (float(v) for v in self.bounds)
        info = (ctypes.c_double * 8)()
        args = (float(p_array.E), float(dz), int(bool(self.step_profile)), float(self.smooth_param), b0, b1,
                K_max, fill_factor, info, 8)
        if _is_cuda_tensor(r):
            dev = _device_of(self, r)
            if r.dtype != torch.float64 or r.stride(1) != 1 or r.shape[0] != 6:
                raise TypeError("device rparticles must be a (6, n) float64 tensor with contiguous rows")
            q = p_array.q_array
            if not _is_cuda_tensor(q):
                q = torch.as_tensor(np.asarray(q, dtype=np.float64), device=r.device)
            q = q.to(dtype=torch.float64).contiguous()
            with torch.cuda.device(dev), _lib.using(dev) as h:
                stream = torch.cuda.current_stream(dev).cuda_stream
                _lib.check(lib.ocsc_lsc_apply(h, r.data_ptr(), r.stride(0), q.data_ptr(), n,
                                              *args, stream))
        else:
            dev = _device_of(self)
            direct = (isinstance(r, np.ndarray) and r.dtype == np.float64 and r.flags.c_contiguous
                      and r.flags.writeable and r.size > 0)
            with torch.cuda.device(dev), _lib.using(dev) as h:
                host = torch.from_numpy(r if direct else np.ascontiguousarray(r, dtype=np.float64))
                rt = host.to(f"cuda:{dev}")
                qt = torch.from_numpy(np.ascontiguousarray(p_array.q_array, dtype=np.float64)).to(f"cuda:{dev}")
                stream = torch.cuda.current_stream(dev).cuda_stream
                _lib.check(lib.ocsc_lsc_apply(h, rt.data_ptr(), n, qt.data_ptr(), n, *args,
                                              stream))
                if direct:
                    host[5].copy_(rt[5])   # only delta changes: one row back, into the caller's ndarray
                else:
                    r[5] = rt[5].cpu().numpy()
        self.last_info = dict(sigma=info[0], sigma_tau=info[1], nbins=int(info[2]), ds=info[3], s0=info[4],
                              q=info[5], count=int(info[6]), mean_tau=info[7], K_max=K_max, fill_factor=fill_factor)

    def last_profile(self, device=None):
        """Grid (m), current (A) and wake W (V) of the last kick on this
        device: the arrays x, B[:, 1] and W of the reference's apply."""
        lib = _lib.load()
        dev = _device_of(self) if device is None else int(device)
        with _lib.using(dev) as h:
            nb = ctypes.c_int64(0)
            rc = lib.ocsc_lsc_profile(h, None, None, None, 0, ctypes.byref(nb))
            if nb.value == 0:
                _lib.check(rc)
            s, cur, W = (np.empty(nb.value) for _ in range(3))
            _lib.check(lib.ocsc_lsc_profile(h, s.ctypes.data_as(_lib._dp), cur.ctypes.data_as(_lib._dp),
                                            W.ctypes.data_as(_lib._dp), nb.value, ctypes.byref(nb)))
        return s, cur, W

    def __repr__(self) -> str:
        cname = type(self).__name__
        step, step_profile, smooth_param, bounds = self.step, self.step_profile, self.smooth_param, self.bounds
        return f"<{cname}: {step=}, {step_profile=}, {smooth_param=}, {bounds=}>"
