"""FP64 CPU oracle for Ocelot's 3D space-charge kick.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it restates, in numpy and
scipy, the algorithm of ``SpaceCharge.apply`` from the reference
(ocelot/cpbd/sc.py:76-251) and of the two coordinate maps it uses
(ocelot/cpbd/coord_transform.py:16-96).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``ocelot_amd``) never calls into it.

It makes the same library calls as the reference's fallback branch (the one
taken when pyfftw and numexpr are absent): ``numpy.fft.fftn/ifftn`` on the
(2N-1)^3 doubled grid, ``numpy.bincount`` for the nearest-grid-point deposit,
``scipy.ndimage.map_coordinates(order=1)`` for the trilinear gather and BLAS
``numpy.dot`` for the frame rotation.

Pinning (how each function is checked, see tests/test_oracle_*.py):
  xxstg_2_xp_mad / xp_2_xxstg_mad -- golden vectors written by the reference,
      round-trip identity and closed-form single-particle kinematics.
  sym_kernel   -- golden vectors; brute-force cubature of the cell integral of
      1/r; scaling K(l*h) = l^2 K(h); far field ~ hx*hy*hz/r; the
      extended-precision form against cubature on elongated cells.
  potential    -- golden vectors; direct O(N^6) convolution sum on small grids.
  deposit      -- golden charge grids; brute-force per-particle NGP loop.
  el_field / apply -- golden vectors written by the reference on seeded bunches
      (incl. random_mesh), plus the uniformly charged ellipsoid (linear interior
      field), linearity in charge and in zstep, and the zero-charge identity.
"""
import numpy as np
import scipy.ndimage as ndimage

# Physical constants, with the exact expressions of ocelot/common/globals.py:13-24
# (not CODATA-2018: m_e_eV is derived from m_e and q_e, epsilon_0 from mu_0).
PI = 3.141592653589793
SPEED_OF_LIGHT = 299792458.0
Q_E = 1.6021766208e-19
M_E_KG = 9.10938215e-31
M_E_EV = M_E_KG * SPEED_OF_LIGHT ** 2 / Q_E
M_E_GEV = M_E_EV / 1e+9
MU_0 = 4 * PI * 1e-7
EPSILON_0 = 1 / MU_0 / SPEED_OF_LIGHT ** 2


def xxstg_2_xp_mad(xxstg, gamref):
    """Ocelot coordinates (x, x', y, y', tau, delta) -> common-time positions and
    momenta in eV/c.  Follows ocelot/cpbd/coord_transform.py:57-96 (numpy branch)."""
    n = xxstg.shape[1]
    betaref = np.sqrt(1 - gamref ** -2)
    gamma = (betaref * xxstg[5] + 1) * gamref
    beta = np.sqrt(1 - gamma ** -2)
    pz2pref = np.sqrt(((gamma * beta) / (gamref * betaref)) ** 2 - xxstg[1] ** 2 - xxstg[3] ** 2)
    u = np.c_[xxstg[1] / pz2pref, xxstg[3] / pz2pref, np.ones(n)]
    u = u / np.linalg.norm(u, 2, 1).reshape((n, 1))
    xp = np.empty_like(xxstg)
    xp[0] = xxstg[0] - u[:, 0] * beta * xxstg[4]
    xp[1] = xxstg[2] - u[:, 1] * beta * xxstg[4]
    xp[2] = -u[:, 2] * beta * xxstg[4]
    xp[3] = u[:, 0] * gamma * beta * M_E_EV
    xp[4] = u[:, 1] * gamma * beta * M_E_EV
    xp[5] = u[:, 2] * gamma * beta * M_E_EV
    return xp


def xp_2_xxstg_mad(xp, xxstg, gamref):
    """Common-time positions/momenta -> Ocelot coordinates, written into the
    rows of ``xxstg`` in place.  Follows ocelot/cpbd/coord_transform.py:16-54."""
    n = xp.shape[1]
    pref = M_E_EV * np.sqrt(gamref ** 2 - 1)
    betaref = np.sqrt(1 - gamref ** -2)
    u = np.c_[xp[3], xp[4], xp[5]]
    gamma = np.sqrt(1 + np.sum(u * u, 1) / M_E_EV ** 2)
    beta = np.sqrt(1 - gamma ** -2)
    u = u / np.linalg.norm(u, 2, 1).reshape((n, 1))
    cdt = -xp[2] / (beta * u[:, 2])
    xxstg[0] = xp[0] + beta * u[:, 0] * cdt
    xxstg[2] = xp[1] + beta * u[:, 1] * cdt
    xxstg[5] = (gamma / gamref - 1) / betaref
    xxstg[4] = cdt
    xxstg[1] = xp[3] / pref
    xxstg[3] = xp[4] / pref
    return xxstg


def igf_primitive(x, y, z):
    """Antiderivative F of 1/r used by the integrated Green's function;
    ocelot/cpbd/sc.py:120-126."""
    r = np.sqrt(x * x + y * y + z * z)
    return (-x * x * 0.5 * np.arctan(y * z / (x * r)) + y * z * np.log(x + r)
            - y * y * 0.5 * np.arctan(z * x / (y * r)) + z * x * np.log(y + r)
            - z * z * 0.5 * np.arctan(x * y / (z * r)) + x * y * np.log(z + r))


def sym_kernel(ijk2, hxyz, dtype=np.float64):
    """Integrated Green's function on the first octant, shape ijk2.
    K[i,j,k] = integral of 1/r over the cell centred at (i*hx, j*hy, k*hz),
    evaluated as the 8-corner alternating difference of F at the cell faces
    (i -/+ 1/2) h.  Follows ocelot/cpbd/sc.py:109-133.

    ``dtype=np.longdouble`` evaluates F and the difference in x87 extended
    precision (64-bit mantissa) and rounds K to float64 at the end: the
    accuracy reference for meshes whose cells are far from cubic, where the
    8-corner difference cancels most of F's digits (the reference's float64
    result then carries errors of 1e-9 relative in the kick)."""
    nx, ny, nz = (int(v) for v in ijk2)
    hx, hy, hz = (dtype(float(v)) for v in hxyz)
    xs = hx * np.arange(nx + 1, dtype=dtype) - hx / 2
    ys = hy * np.arange(ny + 1, dtype=dtype) - hy / 2
    zs = hz * np.arange(nz + 1, dtype=dtype) - hz / 2
    F = igf_primitive(*np.ix_(xs, ys, zs))
    hi_x, lo_x = slice(1, nx + 1), slice(0, nx)
    hi_y, lo_y = slice(1, ny + 1), slice(0, ny)
    hi_z, lo_z = slice(1, nz + 1), slice(0, nz)
    K = (F[hi_x, hi_y, hi_z] - F[lo_x, hi_y, hi_z]
         - F[hi_x, lo_y, hi_z] + F[lo_x, lo_y, hi_z]
         - F[hi_x, hi_y, lo_z] + F[lo_x, hi_y, lo_z]
         + F[hi_x, lo_y, lo_z] - F[lo_x, lo_y, lo_z])
    return K.astype(np.float64)


def mirror_kernel(K):
    """Even extension of the octant kernel onto the (2N-1)^3 doubled grid:
    K2[m] = K[m] for m < N and K2[2N-1-m] = K[m] for 1 <= m < N, per axis.
    ocelot/cpbd/sc.py:145-149."""
    nx, ny, nz = K.shape
    K2 = np.zeros((2 * nx - 1, 2 * ny - 1, 2 * nz - 1))
    K2[:nx, :ny, :nz] = K
    K2[:nx, :ny, nz:] = K2[:nx, :ny, nz - 1:0:-1]
    K2[:nx, ny:, :] = K2[:nx, ny - 1:0:-1, :]
    K2[nx:, :, :] = K2[nx - 1:0:-1, :, :]
    return K2


def potential(q, steps, kernel_dtype=np.float64):
    """Open-boundary potential of the charge grid q by Hockney's doubled-grid
    FFT convolution with the integrated Green's function, scaled by
    1/(4 pi eps0 hx hy hz).  ocelot/cpbd/sc.py:135-168 (numpy.fft branch).
    ``kernel_dtype`` as in sym_kernel (float64 = the reference)."""
    hx, hy, hz = (float(v) for v in steps)
    nx, ny, nz = q.shape
    rho2 = np.zeros((2 * nx - 1, 2 * ny - 1, 2 * nz - 1))
    rho2[:nx, :ny, :nz] = q
    K2 = mirror_kernel(sym_kernel(q.shape, steps, kernel_dtype))
    out = np.real(np.fft.ifftn(np.fft.fftn(rho2) * np.fft.fftn(K2)))
    return out[:nx, :ny, :nz] / (4 * PI * EPSILON_0 * hx * hy * hz)


def field_from_potential(p, steps):
    """Forward differences E = -(phi[i+1]-phi[i])/h, zero on the last plane of
    each axis.  ocelot/cpbd/sc.py:195-200."""
    nx, ny, nz = p.shape
    Ex = np.zeros(p.shape)
    Ey = np.zeros(p.shape)
    Ez = np.zeros(p.shape)
    Ex[:nx - 1] = (p[:nx - 1] - p[1:]) / steps[0]
    Ey[:, :ny - 1] = (p[:, :ny - 1] - p[:, 1:]) / steps[1]
    Ez[:, :, :nz - 1] = (p[:, :, :nz - 1] - p[:, :, 1:]) / steps[2]
    return Ex, Ey, Ez


def mesh_setup(xyz, Q, gamma, nxyz, mesh_scale=None, mesh_shift=None):
    """Rest-frame grid placement.  Scales z by gamma in place, sets the cell
    sizes from the bunch extent and the fractional grid coordinates X.
    ``mesh_scale``/``mesh_shift`` are the two host draws U(1,1.1) and
    U(-0.5,0.5) the reference takes when random_mesh is on (None = off).
    ocelot/cpbd/sc.py:172-186."""
    nxyz = np.asarray(nxyz)
    xyz[:, 2] = xyz[:, 2] * gamma
    XX = np.max(xyz, axis=0) - np.min(xyz, axis=0)
    if mesh_scale is not None:
        XX = XX * mesh_scale
    steps = XX / (nxyz - 3)
    X = xyz / steps
    X_min = np.min(X, axis=0)
    X_mid = np.dot(Q, X) / np.sum(Q)
    X_off = np.floor(X_min - X_mid) + X_mid
    if mesh_shift is not None:
        X_off = X_off + mesh_shift
    X = X - X_off
    return X, steps, X_off


def ngp_indices(X, nxyz):
    """Nearest-grid-point cell of each particle, flat C-order index
    (floor(X)+1 per axis, z fastest).  ocelot/cpbd/sc.py:187-192."""
    nx, ny, nz = (int(v) for v in nxyz)
    Xi = np.int_(np.floor(X) + 1)
    return np.int_(Xi[:, 0] * (nz * ny) + Xi[:, 1] * nz + Xi[:, 2])


def deposit(inds, Q, nxyz):
    """Charge per cell; ocelot/cpbd/sc.py:193."""
    nx, ny, nz = (int(v) for v in nxyz)
    return np.bincount(inds, Q, nx * ny * nz).reshape((nx, ny, nz))


def gather(Ex, Ey, Ez, X, gamma):
    """Trilinear interpolation of the three staggered field components at the
    particles (scipy map_coordinates, order 1, mode 'constant', cval 0: zero
    outside [0, n-1]); transverse components times gamma (rest -> lab).
    ocelot/cpbd/sc.py:201-204."""
    out = np.zeros((X.shape[0], 3))
    out[:, 0] = ndimage.map_coordinates(Ex, np.c_[X[:, 0], X[:, 1] + 0.5, X[:, 2] + 0.5].T, order=1) * gamma
    out[:, 1] = ndimage.map_coordinates(Ey, np.c_[X[:, 0] + 0.5, X[:, 1], X[:, 2] + 0.5].T, order=1) * gamma
    out[:, 2] = ndimage.map_coordinates(Ez, np.c_[X[:, 0] + 0.5, X[:, 1] + 0.5, X[:, 2]].T, order=1)
    return out


def el_field(xyz, Q, gamma, nxyz, mesh_scale=None, mesh_shift=None, debug=None,
             kernel_dtype=np.float64):
    """Lab-frame field (V/m) at the particles from rest-frame positions xyz
    (N,3; modified in place like the reference).  ocelot/cpbd/sc.py:170-205.
    If ``debug`` is a dict it receives steps, X_off, X, inds, rho, phi."""
    nxyz = np.asarray(nxyz)
    X, steps, X_off = mesh_setup(xyz, Q, gamma, nxyz, mesh_scale, mesh_shift)
    inds = ngp_indices(X, nxyz)
    rho = deposit(inds, Q, nxyz)
    phi = potential(rho, steps, kernel_dtype)
    Ex, Ey, Ez = field_from_potential(phi, steps)
    E = gather(Ex, Ey, Ez, X, gamma)
    if debug is not None:
        debug.update(steps=steps, X_off=X_off, X=X, inds=inds, rho=rho, phi=phi)
    return E


def frame(xp):
    """Rotation to the mean-velocity frame: t3 = unweighted mean momentum
    direction, t1 = ey x t3 normalised, t2 = t3 x t1.  Returns (T, Pav).
    ocelot/cpbd/sc.py:224-231."""
    t3 = np.mean(xp[3:6], axis=1)
    Pav = np.linalg.norm(t3)
    t3 = t3 / Pav
    t1 = np.cross(np.array([0, 1, 0]), t3)
    t1 = t1 / np.linalg.norm(t1)
    t2 = np.cross(t3, t1)
    return np.c_[t1, t2, t3], Pav


def apply(rparticles, q_array, E, zstep, nmesh_xyz, mesh_scale=None, mesh_shift=None, debug=None,
          kernel_dtype=np.float64):
    """One space-charge kick; rparticles (6,N) is overwritten in place, all
    six rows.  ocelot/cpbd/sc.py:208-251.  ``mesh_scale``/``mesh_shift`` as in
    mesh_setup (None = random_mesh off); ``kernel_dtype`` as in sym_kernel.
    Returns rparticles."""
    if zstep == 0:
        return rparticles
    nmesh_xyz = np.array(nmesh_xyz)
    gamref = E / M_E_GEV
    betref = np.sqrt(1 - gamref ** -2)
    xp = xxstg_2_xp_mad(rparticles, gamref)
    T, Pav = frame(xp)
    xyz = np.dot(xp[0:3].T, T)
    xp[3:6] = np.dot(xp[3:6].T, T).T
    gamma0 = np.sqrt((Pav / M_E_EV) ** 2 + 1)
    beta0 = np.sqrt(1 - gamma0 ** -2)
    Exyz = el_field(xyz, q_array, gamma0, nmesh_xyz, mesh_scale, mesh_shift, debug, kernel_dtype)
    cdT = zstep / betref
    xp[3] = xp[3] + cdT * (1 - beta0 * beta0) * Exyz[:, 0]
    xp[4] = xp[4] + cdT * (1 - beta0 * beta0) * Exyz[:, 1]
    xp[5] = xp[5] + cdT * Exyz[:, 2]
    xp[3:6] = np.dot(xp[3:6].T, T.T).T
    xp_2_xxstg_mad(xp, rparticles, gamref)
    if debug is not None:
        debug.update(T=T, gamma0=gamma0, Exyz=Exyz)
    return rparticles


def drift_matrix(L, E):
    """First-order R matrix of a drift of length L at energy E [GeV]
    (ocelot/cpbd/r_matrix.py:41-91 with k1 = hx = 0): the identity with
    R01 = R23 = L and R45 = -L/(beta^2 gamma^2)."""
    gamma = E / M_E_GEV
    igamma2 = 1. / (gamma * gamma) if gamma != 0 else 0.
    beta = np.sqrt(1. - igamma2)
    R = np.eye(6)
    R[0, 1] = L
    R[2, 3] = L
    R[4, 5] = -L / (beta * beta) * igamma2
    return R


def drift(rparticles, L, E):
    """First-order drift of length L, applied as R.X
    (ocelot/cpbd/transformations/transfer_map.py:40-50); pinned by the
    reference's own Drift element R matrices and tracked particles
    (tests/golden/track_golden.npz)."""
    rparticles[:] = np.dot(drift_matrix(L, E), rparticles)
    return rparticles
