"""Drop-in ``SpaceCharge`` physics process on the MI355X kernels.

Same surface as the reference class (ocelot/cpbd/sc.py:76-258): constructor
``SpaceCharge(step=1, nmesh_xyz=[63,63,63], low_order_kick=True,
random_mesh=False)``, attributes mutable after construction, ``prepare(lat)``
seeding the global numpy RNG with ``random_seed`` (10), ``apply(p_array,
zstep)`` rewriting all six rows of ``p_array.rparticles`` in place, and the
helper methods ``sym_kernel``, ``potential``, ``el_field``.  Three extra
keyword arguments default to the reference behaviour: ``precision`` ("fp64"
or "mixed"), ``autotune`` (False; True: ``track_drift`` tunes the speed-only
tracking options for the bunch shape first) and ``device`` (GPU index;
default: the current torch device or the device of a resident bunch).

``p_array.rparticles`` may be a numpy array (it is copied to the GPU, kicked
and copied back into the same ndarray, so views such as ``p_array.x()`` stay
valid) or a CUDA float64 tensor (kicked in place in HBM, no host copy).

Random mesh: with ``random_mesh`` on, each kick draws from the global numpy
RNG exactly as the reference does, first U(1, 1.1) (extent factor) and then
U(-0.5, 0.5) (one offset for all three axes); draws happen on the host and
are passed to the kernels.

Errors: the reference raises ValueError for an empty bunch, zero total
charge and non-finite coordinates (numpy reductions / bincount); so does this
class, leaving rparticles untouched.  ``nmesh`` below 4 on an axis (where the
reference divides by n-3 <= 0) also raises ValueError here.

Instances hold only plain data: device contexts live in a per-process
registry, so deepcopy and pickle (Navigator, ParameterScanner) work.
"""
import ctypes

import numpy as np

from . import _lib
from .physproc import PhysProc

_PREC = {"fp64": _lib.OCSC_FP64, "mixed": _lib.OCSC_MIXED}


def _torch():
    import torch
    return torch


def _host_direct(a):
    """True for an ndarray the GPU copies can read and write in place."""
    return (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
            and a.flags.writeable and a.size > 0)


def _is_cuda_tensor(a):
    try:
        torch = _torch()
    except ImportError:
        return False
    return isinstance(a, torch.Tensor) and a.is_cuda


# ---------------------------------------------------------------------------
# Host-side kick arguments from the process attributes the reference defines
# (nmesh_xyz, random_mesh) and the optional precision / device.

def _prec_of(proc):
    prec = getattr(proc, "precision", "fp64")
    try:
        return _PREC[prec]
    except KeyError:
        raise ValueError(f"precision must be 'fp64' or 'mixed', not {prec!r}")


def _device_of(proc, tensor=None):
    if tensor is not None:
        return tensor.device.index
    dev = getattr(proc, "device", None)
    if dev is not None:
        return int(dev)
    return _torch().cuda.current_device()


def _nmesh_of(proc):
    nm = np.array(proc.nmesh_xyz).reshape(-1)
    if nm.size != 3:
        raise ValueError("nmesh_xyz must have three entries")
    return _lib.i32x3(nm)


def _draws_of(proc, q=None, n=None):
    """The two random-mesh draws of one kick, sc.py:174-175 then sc.py:184-185,
    global numpy RNG.  With q, the charges are checked between the draws,
    where the reference's np.dot(Q, X) (sc.py:182) raises for a q_array of
    the wrong length."""
    if not proc.random_mesh:
        if q is not None:
            _lib.check_charges(q, n)
        return 0, 1.0, 0.0
    scale = np.random.uniform(low=1, high=1.1)
    if q is not None:
        _lib.check_charges(q, n)
    shift = np.random.uniform(low=-0.5, high=0.5)
    return 1, float(scale), float(shift)


def _kick_args(proc, p_array, zstep):
    n = int(p_array.rparticles.shape[1])
    if n == 0:
        raise ValueError("zero-size array to reduction operation maximum which has no identity")
    nmesh = _nmesh_of(proc)
    prec = _prec_of(proc)
    E = float(p_array.E)
    rm, scale, shift = _draws_of(proc, p_array.q_array, n)
    return dict(n=n, nmesh=nmesh, prec=prec, E=E, zstep=float(zstep), random_mesh=rm, mesh_scale=scale,
                mesh_shift=shift)


def _kick(proc, p_array, zstep, want_cells):
    """One kick of SpaceCharge.apply (sc.py:208-251) on host or device rows;
    with ``want_cells`` the deposit's cell indices are returned."""
    a = _kick_args(proc, p_array, zstep)
    torch = _torch()
    lib = _lib.load()
    r = p_array.rparticles
    n, nmesh, prec, E = a["n"], a["nmesh"], a["prec"], a["E"]
    rm, scale, shift = a["random_mesh"], a["mesh_scale"], a["mesh_shift"]
    if _is_cuda_tensor(r):
        dev = _device_of(proc, r)
        if r.dtype != torch.float64 or r.stride(1) != 1 or r.shape[0] != 6:
            raise TypeError("device rparticles must be a (6, n) float64 tensor with contiguous rows")
        q = p_array.q_array
        if not _is_cuda_tensor(q):
            q = torch.as_tensor(np.asarray(q, dtype=np.float64), device=r.device)
        q = q.to(device=r.device, dtype=torch.float64).contiguous()
        cells = torch.empty(n, dtype=torch.int32, device=r.device) if want_cells else None
        with torch.cuda.device(dev), _lib.using(dev) as h:
            stream = torch.cuda.current_stream(dev).cuda_stream
            _lib.check(lib.ocsc_apply_cells(h, r.data_ptr(), r.stride(0), q.data_ptr(), n, nmesh, E, float(zstep),
                                            rm, scale, shift, prec, 1,
                                            cells.data_ptr() if cells is not None else None, stream))
        return cells
    dev = _device_of(proc)
    # a C-ordered, writable float64 ndarray is copied to and from the GPU
    # as it lies (no host-side temporaries); anything else goes through a
    # contiguous copy and is assigned back element-wise
    direct = _host_direct(r)
    with torch.cuda.device(dev), _lib.using(dev) as h:
        host = torch.from_numpy(r if direct else np.ascontiguousarray(r, dtype=np.float64))
        rt = host.to(f"cuda:{dev}")
        qt = torch.from_numpy(np.ascontiguousarray(p_array.q_array, dtype=np.float64)).to(f"cuda:{dev}")
        cells = torch.empty(n, dtype=torch.int32, device=rt.device) if want_cells else None
        stream = torch.cuda.current_stream(dev).cuda_stream
        _lib.check(lib.ocsc_apply_cells(h, rt.data_ptr(), n, qt.data_ptr(), n, nmesh, E, float(zstep), rm, scale,
                                        shift, prec, 1, cells.data_ptr() if cells is not None else None, stream))
        if direct:
            host.copy_(rt)
        else:
            r[...] = rt.cpu().numpy()
    return cells.cpu().numpy() if cells is not None else None


class SpaceCharge(PhysProc):
    """3D space-charge kick (Poisson solve in the bunch rest frame by Hockney
    FFT convolution with the integrated Green's function)."""

    def __init__(self, step=1, **kwargs):
        PhysProc.__init__(self)
        self.step = step
        self.nmesh_xyz = kwargs.get("nmesh_xyz", [63, 63, 63])
        self.low_order_kick = kwargs.get("low_order_kick", True)
        self.start_elem = None
        self.end_elem = None
        self.debug = False
        self.random_mesh = kwargs.get("random_mesh", False)
        self.random_seed = 10
        self.precision = kwargs.get("precision", "fp64")
        self.autotune = kwargs.get("autotune", False)
        self.device = kwargs.get("device", None)

    # ------------------------------------------------------------------ helpers
    # (module functions below, so the kick also runs on any object with the
    # reference's attributes, e.g. ocelot's own SpaceCharge instance)
    def _prec(self):
        return _prec_of(self)

    def _device(self, tensor=None):
        return _device_of(self, tensor)

    def _nmesh(self):
        return _nmesh_of(self)

    def _draws(self):
        return _draws_of(self)

    # ------------------------------------------------------------------ PhysProc
    def prepare(self, lat):
        self.check_step()
        if self.random_seed is not None:
            np.random.seed(self.random_seed)

    def kick_args(self, p_array, zstep):
        """Host-side arguments of one kick, in the order the reference
        consumes them: mesh, precision, energy, then the random-mesh draws.
        Raises ValueError for an empty bunch before drawing (the reference
        fails in the extent reduction, sc.py:173, before its first draw) and
        for a q_array of the wrong length between the two draws (np.dot(Q, X),
        sc.py:182)."""
        return _kick_args(self, p_array, zstep)

    def apply(self, p_array, zstep):
        if zstep == 0:
            return
        _kick(self, p_array, zstep, want_cells=False)

    def apply_cells(self, p_array, zstep):
        """``apply``, returning for every particle the flat C-order NGP cell
        index (int32; the ``inds`` of sc.py:187-192) its charge was deposited
        to by this kick's deposit kernel: a CUDA tensor for device rows, a
        numpy array for host rows; None for zstep == 0 (no kick)."""
        if zstep == 0:
            return None
        return _kick(self, p_array, zstep, want_cells=True)

    # ------------------------------------------------------------------ sub-steps
    def sym_kernel(self, ijk2, hxyz):
        """Integrated Green's function on shape ijk2 with cell sizes hxyz
        (sc.py:109-133), FP64 numpy array."""
        torch = _torch()
        lib = _lib.load()
        dev = _device_of(self)
        shape = tuple(int(v) for v in ijk2)
        out = torch.empty(shape, dtype=torch.float64, device=f"cuda:{dev}")
        with torch.cuda.device(dev), _lib.using(dev) as ch:
            stream = torch.cuda.current_stream(dev).cuda_stream
            h = (ctypes.c_double * 3)(*[float(v) for v in hxyz])
            _lib.check(lib.ocsc_sym_kernel(ch, _lib.i32x3(shape), h, out.data_ptr(), stream))
        return out.cpu().numpy()

    def potential(self, q, steps):
        """Potential of the charge grid q (sc.py:135-168), FP64 numpy array
        (FP32 values widened in mixed precision)."""
        torch = _torch()
        lib = _lib.load()
        dev = _device_of(self)
        prec = _prec_of(self)
        rho = torch.from_numpy(np.ascontiguousarray(q, dtype=np.float64)).to(f"cuda:{dev}")
        dt = torch.float64 if prec == _lib.OCSC_FP64 else torch.float32
        phi = torch.empty(rho.shape, dtype=dt, device=rho.device)
        with torch.cuda.device(dev), _lib.using(dev) as ch:
            stream = torch.cuda.current_stream(dev).cuda_stream
            h = (ctypes.c_double * 3)(*[float(v) for v in steps])
            _lib.check(lib.ocsc_potential(ch, rho.data_ptr(), _lib.i32x3(rho.shape), h, phi.data_ptr(),
                                          prec, stream))
        return phi.double().cpu().numpy()

    def el_field(self, X, Q, gamma, nxyz, return_cells=False):
        """Lab-frame field at rest-frame positions X (N, 3) (sc.py:170-205).
        Like the reference, scales X[:, 2] by gamma in place and takes the
        random-mesh draws when random_mesh is on."""
        torch = _torch()
        lib = _lib.load()
        dev = _device_of(self)
        prec = _prec_of(self)
        n = X.shape[0]
        if n == 0:
            raise ValueError("zero-size array to reduction operation maximum which has no identity")
        xyz = torch.from_numpy(np.ascontiguousarray(np.asarray(X, dtype=np.float64).T)).to(f"cuda:{dev}")
        qt = torch.from_numpy(np.ascontiguousarray(Q, dtype=np.float64)).to(f"cuda:{dev}")
        E = torch.empty((3, n), dtype=torch.float64, device=xyz.device)
        cells = torch.empty(n, dtype=torch.int32, device=xyz.device) if return_cells else None
        rm, scale, shift = _draws_of(self, np.asarray(Q), n)
        nm = _lib.i32x3(np.array(nxyz).reshape(-1))
        with torch.cuda.device(dev), _lib.using(dev) as ch:
            stream = torch.cuda.current_stream(dev).cuda_stream
            _lib.check(lib.ocsc_el_field(ch, xyz.data_ptr(), n, qt.data_ptr(), n, nm, float(gamma), rm,
                                         scale, shift, prec, E.data_ptr(), n,
                                         cells.data_ptr() if cells is not None else None, stream))
        X[:, 2] = X[:, 2] * gamma
        out = np.ascontiguousarray(E.cpu().numpy().T)
        if return_cells:
            return out, cells.cpu().numpy()
        return out

    def last_stats(self, device=None):
        """Scalars of the last kick on the device: gamma0, Pav, T, steps, X_off,
        total charge, error bits, kernel-rebuild flag."""
        lib = _lib.load()
        dev = _device_of(self) if device is None else int(device)
        out = (ctypes.c_double * 20)()
        with _lib.using(dev) as ch:
            _lib.check(lib.ocsc_last_stats(ch, out, 20, None))
        v = list(out)
        return dict(gamma0=v[0], Pav=v[1], T=np.array(v[2:11]).reshape(3, 3), steps=np.array(v[11:14]),
                    X_off=np.array(v[14:17]), sumq=v[17], err=int(v[18]), rebuilt=int(v[19]))

    def __repr__(self) -> str:
        cname = type(self).__name__
        step = self.step
        nmesh_xyz = self.nmesh_xyz
        random_mesh = self.random_mesh
        return f"<{cname}: {step=}, {nmesh_xyz=}, {random_mesh=}>"


def drift(p_array, L, device=None):
    """First-order drift of length L on a bunch (numpy or CUDA rows), the map
    of a Drift element (ocelot/cpbd/r_matrix.py:41-91, k1 = hx = 0)."""
    torch = _torch()
    lib = _lib.load()
    r = p_array.rparticles
    n = int(r.shape[1])
    if _is_cuda_tensor(r):
        dev = r.device.index
        with torch.cuda.device(dev), _lib.using(dev) as h:
            _lib.check(lib.ocsc_drift(h, r.data_ptr(), r.stride(0), n, float(L),
                                      float(p_array.E), torch.cuda.current_stream(dev).cuda_stream))
        return
    dev = torch.cuda.current_device() if device is None else int(device)
    with torch.cuda.device(dev), _lib.using(dev) as h:
        rt = torch.from_numpy(np.ascontiguousarray(r, dtype=np.float64)).to(f"cuda:{dev}")
        _lib.check(lib.ocsc_drift(h, rt.data_ptr(), n, n, float(L), float(p_array.E),
                                  torch.cuda.current_stream(dev).cuda_stream))
        r[...] = rt.cpu().numpy()


def track_drift(p_array, sc, L, dz, nsteps, use_graph=True, calc_tws=False, bounds=None, auto_disp=False,
                tws_i=None, lsc=None, z0=None, slice=None):
    """nsteps x [drift(L); sc.apply(p_array, dz)] on a device-resident bunch,
    the step loop of ocelot/cpbd/track.py:462-485 for a drift lattice, with
    one step captured in a HIP graph and replayed.

    With ``calc_tws`` the beam moments are taken on the device before the
    first step and after every step, and the list of Twiss (``tws_track`` of
    track(), ``s`` = cumulative length) is returned; ``bounds``, ``slice``
    (None or "Imax"), ``auto_disp`` and ``tws_i`` as in get_envelope.

    With ``lsc`` (an LSC instance) every step also applies ``lsc.apply(p_array,
    dz)`` after the space-charge kick, inside the same graph (the order in
    which track() applies two processes added SpaceCharge first); ``sc`` may
    then be None for LSC-only tracking.  For a prepared LSC with undulators in
    its line, step k ends at ``z0 + (k + 1) L`` (``z0``: the lattice position at
    the start, default ``p_array.s``) and takes the K_max and filling factor of
    that step, as LSC.apply would with the Navigator's z0; ``lsc.z0`` is left at
    the end position.

    With ``sc.random_mesh`` each kick takes the reference's two global numpy
    draws (scale, then shift), all taken before the call in kick order.

    With ``sc.autotune`` the speed-only options of the tracking loop are first
    tuned for this bunch size, mesh and precision (ocelot_amd.autotune, on a
    copy of the bunch, once per shape; no random draws are taken)."""
    torch = _torch()
    lib = _lib.load()
    r = p_array.rparticles
    if not _is_cuda_tensor(r):
        raise TypeError("track_drift needs a DeviceParticleArray")
    if sc is None and lsc is None:
        raise ValueError("track_drift needs a SpaceCharge or an LSC process")
    dev = r.device.index
    n = int(r.shape[1])
    nsteps = int(nsteps)
    _lib.check_charges(p_array.q_array, n)
    if sc is not None and getattr(sc, "autotune", False) and nsteps > 0 and float(dz) != 0:
        from .autotune import tune_tracking
        tune_tracking(p_array, sc, L, dz)
    opts = _lib.TrackOpts()
    opts.use_graph = 1 if use_graph else 0
    zstep = float(dz) if sc is not None else 0.0
    nmesh = _nmesh_of(sc) if sc is not None else _lib.i32x3(np.array([4, 4, 4]))
    prec = _prec_of(sc) if sc is not None else _lib.OCSC_FP64
    keep = []
    if sc is not None and sc.random_mesh and zstep != 0:
        # the reference draws U(1, 1.1) then U(-0.5, 0.5) from the global numpy
        # RNG in every kick (sc.py:174-175, 184-185); no kick, no draw
        draws = np.array([_draws_of(sc)[1:] for _ in range(nsteps)], dtype=np.float64).reshape(-1)
        buf = (ctypes.c_double * max(draws.size, 1))(*draws)
        keep.append(buf)
        opts.mesh_draws = ctypes.cast(buf, _lib._dp)
    if lsc is not None:
        opts.lsc = 1
        opts.lsc_dz = float(dz)
        opts.lsc_step_profile = int(bool(lsc.step_profile))
        opts.lsc_smooth_param = float(lsc.smooth_param)
        opts.lsc_b0, opts.lsc_b1 = (float(v) for v in lsc.bounds)
        if getattr(lsc, "_is_undul_in_beam_line", False) and float(dz) >= 1e-10:
            from .lsc import _kick_args as _lsc_kick_args
            start = float(p_array.s if z0 is None else z0)
            vals = []
            for k in range(nsteps):
                lsc.z0 = start + (k + 1) * float(L)
                K_max, fill_factor = _lsc_kick_args(lsc, float(dz))
                vals.append(1 + 0.5 * K_max * K_max * fill_factor)   # as the impedance factor is formed
            fac = (ctypes.c_double * max(len(vals), 1))(*vals)
            keep.append(fac)
            opts.lsc_und_factors = ctypes.cast(fac, _lib._dp)
    nout = 27
    if calc_tws:
        flags, disp = 0, (0.0, 0.0, 0.0, 0.0)
        if bounds is not None:
            flags |= 1
            opts.mom_b0, opts.mom_b1 = float(bounds[0]), float(bounds[1])
            if slice == "Imax":
                flags |= 8
        if tws_i is not None:
            flags |= 4
            disp = (tws_i.Dx, tws_i.Dxp, tws_i.Dy, tws_i.Dyp)
        elif auto_disp:
            flags |= 2
        opts.mom_flags = flags
        for k in range(4):
            opts.disp[k] = float(disp[k])
        out = (ctypes.c_double * ((nsteps + 1) * nout))()
        opts.tws_out = ctypes.cast(out, _lib._dp)
    with torch.cuda.device(dev), _lib.using(dev) as h:
        stream = torch.cuda.current_stream(dev).cuda_stream
        _lib.check(lib.ocsc_track(h, r.data_ptr(), r.stride(0), p_array.q_array.data_ptr(), n,
                                  nmesh, float(p_array.E), zstep, float(L), nsteps, prec, ctypes.byref(opts),
                                  stream))
    # each step's drift advances the position, as TransferMap.apply does
    # (ocelot/cpbd/transformations/transformation.py:131)
    p_array.s = float(p_array.s) + nsteps * float(L)
    if not calc_tws:
        return None
    from .envelope import twiss_from_moments
    rows = np.frombuffer(out, dtype=np.float64).reshape(nsteps + 1, nout)
    tws, s = [], 0.0
    for k in range(nsteps + 1):
        t = twiss_from_moments(rows[k], p_array.E, tws_i is not None or auto_disp)
        if k:
            s += float(L)  # track.py:481-483 accumulates the step lengths
        t.s = s
        tws.append(t)
    return tws


def apply_batched(items, precision="fp64", nkicks=1, use_graph=True, random_mesh=False):
    """Kick many independent device-resident bunches (uneven sizes, meshes and
    energies) together, ``nkicks`` times each: ``items`` is a list of
    ``(p_array, nmesh_xyz, zstep)`` with ``p_array`` a DeviceParticleArray on
    one device.  Per bunch the result equals ``SpaceCharge(nmesh_xyz).apply``.
    ``random_mesh`` takes the reference's two global draws per bunch (one
    round only)."""
    torch = _torch()
    lib = _lib.load()
    if not items:
        return
    prec = _PREC[precision]
    dev = items[0][0].rparticles.device.index
    arr = (_lib.Bunch * len(items))()
    keep = []
    for b, (p, nmesh, dz) in zip(arr, items):
        r = p.rparticles
        if not _is_cuda_tensor(r) or r.device.index != dev:
            raise TypeError("apply_batched needs DeviceParticleArrays on one device")
        if r.shape[1] == 0:
            raise ValueError("zero-size array to reduction operation maximum which has no identity")
        _lib.check_charges(p.q_array, r.shape[1])
        q = p.q_array.to(dtype=torch.float64).contiguous()
        keep.append(q)
        b.rparticles, b.ld, b.q_array, b.n = r.data_ptr(), r.stride(0), q.data_ptr(), r.shape[1]
        for k, v in enumerate(np.array(nmesh).reshape(-1)):
            b.nmesh[k] = int(v)
        b.E_GeV, b.zstep = float(p.E), float(dz)
        if random_mesh:
            b.random_mesh = 1
            b.mesh_scale = float(np.random.uniform(low=1, high=1.1))
            b.mesh_shift = float(np.random.uniform(low=-0.5, high=0.5))
        else:
            b.random_mesh, b.mesh_scale, b.mesh_shift = 0, 1.0, 0.0
    with torch.cuda.device(dev), _lib.using(dev) as h:
        _lib.check(lib.ocsc_apply_batched(h, ctypes.cast(arr, ctypes.c_void_p), len(items),
                                          prec, int(nkicks), 1 if use_graph else 0,
                                          torch.cuda.current_stream(dev).cuda_stream))
