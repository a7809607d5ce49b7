"""Beam moments and Twiss parameters of a bunch, computed on the GPU.

``get_envelope(p_array, tws_i=None, bounds=None, slice=None, auto_disp=False)``
has the surface of the reference function (ocelot/cpbd/beam/analysis.py:27-224):
the means and second moments are taken over the particles by ``ocsc_moments``
(passes over the six rows in HBM), and the handful of derived scalars
(emittances, Twiss beta/alpha, eigen-emittances) are formed on the host from
those moments.  ``p_array.rparticles`` may be a numpy array (copied to the GPU
for the reduction) or a CUDA tensor (reduced in place, no copy).

Eigen-emittances: the eigenvalues of Sigma*S for the 4x4 transverse moment
matrix Sigma and the symplectic form S are +-i e1, +-i e2, so e1^2 + e2^2 =
-tr((Sigma S)^2)/2 and e1^2 e2^2 = det(Sigma); e1 >= e2 are the roots of that
quadratic (the reference evaluates the same two roots in closed form).

With ``bounds``, ``slice="Imax"`` centres the window on the peak of the
smoothed current profile (the reference's ``s_to_cur`` with width
0.01 std(tau)): the cloud-in-cell profile is summed on the device, the
smoothing and the peak search run on the host over its few thousand bins.
"""
import numpy as np

from . import _lib

M_E_GEV = 9.10938215e-31 * 299792458.0 ** 2 / 1.6021766208e-19 / 1e+9

_FIELDS_ZERO = ("emit_x", "emit_y", "emit_xn", "emit_yn", "eigemit_1", "eigemit_2", "beta_x", "beta_y",
                "alpha_x", "alpha_y", "Dx", "Dy", "Dxp", "Dyp", "mux", "muy", "s", "q", "x", "y", "p", "tau",
                "xp", "yp", "xx", "xpx", "pxpx", "yy", "ypy", "pypy", "tautau", "xy", "pxpy", "xpy", "ypx", "pp",
                "E")


class Twiss:
    """Plain container with the attributes get_envelope fills, all 0 by
    default as in the reference's Twiss (ocelot/cpbd/beam/core.py:16-75)."""

    def __init__(self, **kwargs):
        for k in _FIELDS_ZERO:
            setattr(self, k, 0.0)
        self.id = ""
        for k, v in kwargs.items():
            setattr(self, k, v)

    def __repr__(self):
        keys = ("emit_x", "emit_y", "beta_x", "beta_y", "alpha_x", "alpha_y", "E", "q")
        return "<Twiss: " + ", ".join(f"{k}={getattr(self, k)!r}" for k in keys) + ">"


def eigen_emittances(Sigma):
    """(e1, e2), e1 >= e2, of a 4x4 moment matrix in (x, y, px, py) order."""
    S = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [-1, 0, 0, 0], [0, -1, 0, 0]], dtype=float)
    M = Sigma @ S
    s = -0.5 * np.trace(M @ M)
    d = np.linalg.det(Sigma)
    disc = np.sqrt(complex(s * s - 4 * d))
    r1 = np.sqrt((s + disc) / 2)
    r2 = np.sqrt((s - disc) / 2)
    return float(abs(r1.real)), float(abs(r2.real))


def moments(p_array, bounds=None, auto_disp=False, disp=None, device=None, imax=False):
    """Raw output of ocsc_moments (OCSC_MOM_NOUT values) for a bunch."""
    import ctypes
    import torch
    lib = _lib.load()
    r = p_array.rparticles
    flags = 0
    b0 = b1 = 0.0
    if bounds is not None:
        flags |= 1
        b0, b1 = float(bounds[0]), float(bounds[1])
        if imax:
            flags |= 8
    if disp is not None:
        flags |= 4
    elif auto_disp:
        flags |= 2
    dbuf = (ctypes.c_double * 4)(*(disp if disp is not None else (0.0, 0.0, 0.0, 0.0)))
    out = (ctypes.c_double * 27)()
    if isinstance(r, torch.Tensor) and r.is_cuda:
        if r.dtype != torch.float64 or r.stride(1) != 1 or r.shape[0] != 6:
            raise TypeError("device rparticles must be a (6, n) float64 tensor with contiguous rows")
        dev = r.device.index
        q = p_array.q_array
        if not (isinstance(q, torch.Tensor) and q.is_cuda):
            q = torch.as_tensor(np.asarray(q, dtype=np.float64), device=r.device)
        q = q.to(dtype=torch.float64).contiguous()
        rt, ld = r, r.stride(0)
    else:
        dev = torch.cuda.current_device() if device is None else int(device)
        rt = torch.from_numpy(np.ascontiguousarray(r, dtype=np.float64)).to(f"cuda:{dev}")
        q = torch.from_numpy(np.ascontiguousarray(p_array.q_array, dtype=np.float64)).to(f"cuda:{dev}")
        ld = rt.shape[1]
    n = int(rt.shape[1])
    _lib.check_charges(q, n)
    with torch.cuda.device(dev), _lib.using(dev) as h:
        stream = torch.cuda.current_stream(dev).cuda_stream
        _lib.check(lib.ocsc_moments(h, rt.data_ptr(), ld, q.data_ptr(), n, flags, dbuf,
                                    b0, b1, out, 27, stream))
    return np.array(out[:])


def get_envelope(p_array, tws_i=None, bounds=None, slice=None, auto_disp=False, device=None):
    """Twiss parameters of a bunch (analysis.py:27-224) from moments taken on
    the GPU.  Returns a ``Twiss``; with fewer than 3 selected particles only
    E, q and p are set, as in the reference."""
    disp = None
    if tws_i is not None:
        disp = (tws_i.Dx, tws_i.Dxp, tws_i.Dy, tws_i.Dyp)
    v = moments(p_array, bounds=bounds, auto_disp=auto_disp and tws_i is None, disp=disp, device=device,
                imax=slice == "Imax")
    return twiss_from_moments(v, float(p_array.E), disp is not None or auto_disp)


def twiss_from_moments(v, E, with_disp):
    """A Twiss from one row of moments (OCSC_MOM_NOUT values, the layout of
    ocsc_moments), with the derived quantities of analysis.py:160-224."""
    tws = Twiss()
    tws.E = float(E)
    tws.q = float(v[1])
    tws.p = float(v[2])
    if v[0] < 3:
        return tws
    if with_disp:
        tws.Dx, tws.Dxp, tws.Dy, tws.Dyp = (float(a) for a in v[3:7])
    # numpy scalars, as in the reference: a zero emittance gives inf/nan with a
    # RuntimeWarning rather than an exception
    tws.x, tws.y, tws.px, tws.py, tws.tau = (np.float64(a) for a in v[7:12])
    (tws.xx, tws.xpx, tws.pxpx, tws.yy, tws.ypy, tws.pypy, tws.tautau, tws.xy, tws.pxpy, tws.xpy, tws.ypx,
     tws.pp) = (np.float64(a) for a in v[12:24])
    tws.emit_x = np.sqrt(tws.xx * tws.pxpx - tws.xpx ** 2)
    tws.emit_y = np.sqrt(tws.yy * tws.pypy - tws.ypy ** 2)
    relgamma = tws.E / M_E_GEV
    relbeta = np.sqrt(1 - relgamma ** -2) if relgamma != 0 else 1.
    tws.emit_xn = tws.emit_x * relgamma * relbeta
    tws.emit_yn = tws.emit_y * relgamma * relbeta
    Sigma = np.array([[tws.xx, tws.xy, tws.xpx, tws.xpy],
                      [tws.xy, tws.yy, tws.ypx, tws.ypy],
                      [tws.xpx, tws.ypx, tws.pxpx, tws.pxpy],
                      [tws.xpy, tws.ypy, tws.pxpy, tws.pypy]])
    tws.eigemit_1, tws.eigemit_2 = eigen_emittances(Sigma)
    tws.beta_x = tws.xx / tws.emit_x
    tws.beta_y = tws.yy / tws.emit_y
    tws.alpha_x = -tws.xpx / tws.emit_x
    tws.alpha_y = -tws.ypy / tws.emit_y
    return tws
