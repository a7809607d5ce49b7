"""Pins of the oracle that do not depend on the reference's own output:
closed forms, brute force on tiny inputs and invariants.  Each is chosen so
that a plausible slip (a dropped term, a sign, a half-cell offset, a swapped
axis, a missing gamma factor) fails at least one of them."""
import numpy as np
import pytest
import scipy.integrate as integ
import scipy.ndimage as ndimage

from oracle import sc_oracle as orc


# ---------------------------------------------------------------- IGF (sc.py:109-133)

def _cell_integral(c, h, nsub=2, npts=10):
    """Gauss-Legendre cubature of 1/r over the box centred at c with sides h,
    split into near-cubic sub-boxes (nsub along the shortest side)."""
    g, w = np.polynomial.legendre.leggauss(npts)
    ns = [max(1, int(round(nsub * h[d] / min(h)))) for d in range(3)]
    tot = 0.0
    for sx in range(ns[0]):
        for sy in range(ns[1]):
            for sz in range(ns[2]):
                lo = [c[d] - h[d] / 2 + h[d] * s / ns[d] for d, s in enumerate((sx, sy, sz))]
                half = [h[d] / ns[d] / 2 for d in range(3)]
                xs = [lo[d] + half[d] * (g + 1) for d in range(3)]
                X, Y, Z = np.meshgrid(*xs, indexing="ij")
                W = w[:, None, None] * w[None, :, None] * w[None, None, :]
                tot += np.sum(W / np.sqrt(X * X + Y * Y + Z * Z)) * half[0] * half[1] * half[2]
    return tot


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (2e-6, 3e-6, 5e-5), (1e-5, 1e-5, 1e-3)])
def test_igf_matches_cubature(h):
    K = orc.sym_kernel((4, 4, 4), h)
    for idx in [(1, 0, 0), (0, 2, 1), (3, 3, 3), (1, 1, 2), (2, 0, 3)]:
        c = [idx[d] * h[d] for d in range(3)]
        ref = _cell_integral(c, h, nsub=4, npts=12)
        assert abs(K[idx] - ref) <= 1e-7 * abs(ref), (idx, K[idx], ref)


@pytest.mark.parametrize("h", [(1.6e-6, 1.6e-6, 1e-4), (2e-5, 2e-5, 4e-6)])
def test_igf_extended_precision_matches_cubature(h):
    """On cells far from cubic the float64 8-corner difference loses digits
    (1e-12 relative here); the extended-precision evaluation, the accuracy
    reference for such meshes, matches the cubature to a few ulps of float64."""
    K64 = orc.sym_kernel((6, 6, 6), h)
    Kx = orc.sym_kernel((6, 6, 6), h, np.longdouble)
    assert Kx.dtype == np.float64
    worst64 = 0.0
    for idx in [(1, 0, 0), (5, 5, 1), (0, 3, 2), (4, 1, 0)]:
        c = [idx[d] * h[d] for d in range(3)]
        ref = _cell_integral(c, h, nsub=2, npts=12)
        assert abs(Kx[idx] - ref) <= 4e-15 * abs(ref), (idx, Kx[idx], ref)
        worst64 = max(worst64, abs(K64[idx] - ref) / abs(ref))
    assert worst64 > 1e-13          # the case is ill-conditioned for float64


def test_igf_origin_cell_closed_form():
    # integral of 1/r over the unit cube centred at 0 = 8 * (1/2)^2 * I1,
    # with I1 = -pi/4 + 3/2 ln(2 + sqrt 3) the corner-cube integral
    I1 = -np.pi / 4 + 1.5 * np.log(2 + np.sqrt(3))
    K = orc.sym_kernel((2, 2, 2), (1.0, 1.0, 1.0))
    assert abs(K[0, 0, 0] - 2 * I1) <= 1e-13
    h = 3.5
    K = orc.sym_kernel((2, 2, 2), (h, h, h))
    assert abs(K[0, 0, 0] - 2 * I1 * h * h) <= 1e-12 * h * h


def test_igf_scaling_and_far_field():
    h = np.array([1e-5, 2e-5, 4e-4])
    K = orc.sym_kernel((12, 12, 12), h)
    K3 = orc.sym_kernel((12, 12, 12), 3.7 * h)
    assert np.max(np.abs(K3 - 3.7 ** 2 * K) / np.abs(K3)) <= 1e-9
    # far field: cell integral ~ volume / distance, error O((h/r)^2)
    i, j, k = np.meshgrid(*[np.arange(12)] * 3, indexing="ij")
    r = np.sqrt((i * h[0]) ** 2 + (j * h[1]) ** 2 + (k * h[2]) ** 2)
    far = r > 8 * h.max()
    approx = h.prod() / r[far]
    assert np.max(np.abs(K[far] - approx) / approx) < 0.01


def test_mirror_kernel_even():
    K = orc.sym_kernel((5, 6, 7), (1.0, 1.3, 0.7))
    K2 = orc.mirror_kernel(K)
    assert K2.shape == (9, 11, 13)
    for ax, n in enumerate(K.shape):
        m = 2 * n - 1
        for s in range(1, n):
            a = np.take(K2, s, axis=ax)
            b = np.take(K2, m - s, axis=ax)
            assert np.array_equal(a, b)


# ---------------------------------------------------------------- potential (sc.py:135-168)

def test_potential_direct_sum():
    rng = np.random.default_rng(3)
    shape = (5, 6, 7)
    h = (1e-5, 2e-5, 7e-5)
    q = rng.uniform(-1, 1, shape) * 1e-12
    K = orc.sym_kernel(shape, h)
    direct = np.zeros(shape)
    I = np.indices(shape).reshape(3, -1).T
    for a in I:
        d = np.abs(I - a)
        direct[tuple(a)] = np.sum(q.reshape(-1) * K[d[:, 0], d[:, 1], d[:, 2]])
    direct /= 4 * orc.PI * orc.EPSILON_0 * h[0] * h[1] * h[2]
    phi = orc.potential(q, h)
    assert np.max(np.abs(phi - direct)) <= 1e-12 * np.max(np.abs(direct))


def test_constants_exact_expressions():
    assert orc.M_E_EV == 9.10938215e-31 * 299792458.0 ** 2 / 1.6021766208e-19
    assert abs(orc.M_E_EV - 510998.8671) < 1e-3
    assert abs(orc.EPSILON_0 - 8.854187817620e-12) < 1e-23


# ---------------------------------------------------------------- deposit and grid placement

def test_mesh_setup_invariants():
    rng = np.random.default_rng(5)
    xyz = rng.standard_normal((4000, 3)) * [1e-4, 2e-4, 1e-3]
    Q = rng.uniform(0.5, 1.5, 4000)
    nxyz = np.array([15, 21, 31])
    X, steps, X_off = orc.mesh_setup(xyz.copy(), Q, 3.0, nxyz)
    ext = np.ptp(xyz * [1, 1, 3.0], axis=0)
    assert np.allclose(steps, ext / (nxyz - 3), rtol=1e-15)
    assert np.all(X.min(axis=0) >= 0) and np.all(X.min(axis=0) < 1)
    assert np.all(X.max(axis=0) < nxyz - 2)
    centroid = Q @ X / Q.sum()
    assert np.allclose(centroid, np.round(centroid), atol=1e-9)


def test_deposit_brute_force():
    rng = np.random.default_rng(6)
    nxyz = (7, 8, 9)
    X = rng.uniform(0, 1, (3000, 3)) * (np.array(nxyz) - 2.0001)
    Q = rng.uniform(0, 1, 3000)
    rho = orc.deposit(orc.ngp_indices(X, nxyz), Q, nxyz)
    ref = np.zeros(nxyz)
    for p in range(X.shape[0]):
        i, j, k = (int(np.floor(v)) + 1 for v in X[p])
        ref[i, j, k] += Q[p]
    assert np.allclose(rho, ref, rtol=1e-13, atol=0)
    assert abs(rho.sum() - Q.sum()) < 1e-12


# ---------------------------------------------------------------- gather (sc.py:201-204)

def _trilinear(F, c):
    n = F.shape
    if any(c[d] < 0 or c[d] > n[d] - 1 for d in range(3)):
        return 0.0
    i0 = [min(int(np.floor(c[d])), n[d] - 2) for d in range(3)]
    t = [c[d] - i0[d] for d in range(3)]
    v = 0.0
    for a in (0, 1):
        for b in (0, 1):
            for e in (0, 1):
                w = (t[0] if a else 1 - t[0]) * (t[1] if b else 1 - t[1]) * (t[2] if e else 1 - t[2])
                v += w * F[i0[0] + a, i0[1] + b, i0[2] + e]
    return v


def test_gather_textbook_trilinear_and_edge_rule():
    rng = np.random.default_rng(8)
    F = rng.standard_normal((5, 6, 7))
    pts = rng.uniform(-0.6, 1.0, (400, 3)) * [4.5, 5.5, 6.5]
    pts[:5] = [[0, 0, 0], [4, 5, 6], [4.0, 2.5, 6.0], [-1e-9, 1, 1], [1, 5 + 1e-9, 1]]
    got = ndimage.map_coordinates(F, pts.T, order=1)
    ref = np.array([_trilinear(F, p) for p in pts])
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-13)
    assert got[3] == 0.0 and got[4] == 0.0


def test_field_differences_sign_and_staggering():
    # phi linear in x: E_x = -dphi/dx everywhere except the last x plane
    n = (6, 5, 4)
    h = (0.5, 1.0, 2.0)
    i = np.arange(n[0])[:, None, None] * np.ones(n)
    phi = 3.0 * i * h[0]
    Ex, Ey, Ez = orc.field_from_potential(phi, h)
    assert np.allclose(Ex[:-1], -3.0) and np.all(Ex[-1] == 0)
    assert np.all(Ey == 0) and np.all(Ez == 0)


# ---------------------------------------------------------------- coordinate maps

def test_coordinate_round_trip_and_kinematics():
    rng = np.random.default_rng(9)
    r = rng.standard_normal((6, 3000)) * np.array([1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 1e-2])[:, None]
    for E in (0.0065, 0.13, 2.0):
        gamref = E / orc.M_E_GEV
        xp = orc.xxstg_2_xp_mad(r, gamref)
        back = orc.xp_2_xxstg_mad(xp, np.zeros_like(r), gamref)
        for row in range(6):
            assert np.max(np.abs(back[row] - r[row])) <= 1e-12 * np.max(np.abs(r[row]))
        # transverse momentum is x' times the reference momentum
        pref = orc.M_E_EV * np.sqrt(gamref ** 2 - 1)
        assert np.allclose(xp[3], r[1] * pref, rtol=1e-13)
        # total momentum from the energy deviation delta = dE/(p0 c)
        betaref = np.sqrt(1 - gamref ** -2)
        gamma = gamref + r[5] * betaref * gamref
        ptot = orc.M_E_EV * np.sqrt(gamma ** 2 - 1)
        assert np.allclose(np.sqrt(xp[3] ** 2 + xp[4] ** 2 + xp[5] ** 2), ptot, rtol=1e-12)


# ---------------------------------------------------------------- end to end

def _ellipsoid_field_factor(a, b, c, axis):
    """int_0^inf ds / ((a_axis^2 + s) sqrt((a^2+s)(b^2+s)(c^2+s))), computed in
    units of a^2 (s = a^2 t) so the quadrature sees O(1) scales."""
    A, B, C = 1.0, (b / a) ** 2, (c / a) ** 2
    s_ax = (A, B, C)[axis]
    f = lambda t: 1.0 / ((s_ax + t) * np.sqrt((A + t) * (B + t) * (C + t)))
    return integ.quad(f, 0, np.inf, limit=200)[0] / a ** 3


def test_uniform_ellipsoid_linear_interior_field():
    """Uniformly charged ellipsoid: the interior rest-frame field is linear,
    E_i = rho/(2 eps0) a b c x_i * int_0^inf ds / ((a_i^2+s) R(s)); the lab
    transverse field carries an extra gamma (sc.py:202-203)."""
    rng = np.random.default_rng(10)
    a, b, c = 1e-3, 1.5e-3, 3e-3
    gamma = 2.0
    n = 400_000
    u = rng.standard_normal((n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    u *= rng.uniform(0, 1, n)[:, None] ** (1 / 3)
    xyz = u * [a, b, c]
    Qtot = 1e-9
    Q = np.full(n, Qtot / n)
    lab = xyz.copy()
    lab[:, 2] /= gamma
    E = orc.el_field(lab.copy(), Q, gamma, np.array([41, 41, 41]))
    rho = Qtot / (4 / 3 * np.pi * a * b * c)
    inner = np.sum((xyz / [a, b, c]) ** 2, axis=1) < 0.5 ** 2
    for ax, fac in ((0, gamma), (1, gamma), (2, 1.0)):
        k = rho / (2 * orc.EPSILON_0) * a * b * c * _ellipsoid_field_factor(a, b, c, ax)
        ref = fac * k * xyz[inner, ax]
        got = E[inner, ax]
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert rel < 0.08, (ax, rel)


def _bunch(n=3000, seed=11):
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((6, n)) * np.array([1e-4, 1e-5, 1e-4, 1e-5, 1e-3, 1e-3])[:, None]
    return r


def test_linear_in_charge_and_step():
    r0 = _bunch()
    q = np.full(r0.shape[1], 1e-13)
    outs = {}
    for tag, qq, dz in (("base", q, 0.1), ("q2", 2 * q, 0.1), ("dz2", q, 0.2)):
        r = r0.copy()
        orc.apply(r, qq, 0.13, dz, (15, 15, 15))
        outs[tag] = r - r0
    for row in (1, 3):
        assert np.allclose(outs["q2"][row], 2 * outs["base"][row], rtol=1e-6, atol=1e-3 * np.abs(outs["base"][row]).max())
        assert np.allclose(outs["dz2"][row], 2 * outs["base"][row], rtol=1e-6, atol=1e-3 * np.abs(outs["base"][row]).max())


def test_negligible_charge_is_identity():
    r0 = _bunch()
    r = r0.copy()
    orc.apply(r, np.full(r0.shape[1], 1e-40), 0.13, 0.1, (15, 15, 15))
    for row in range(6):
        assert np.max(np.abs(r[row] - r0[row])) <= 1e-12 * np.max(np.abs(r0[row]))


def test_translation_invariance():
    r0 = _bunch()
    q = np.full(r0.shape[1], 1e-13)
    a = r0.copy()
    orc.apply(a, q, 0.13, 0.1, (15, 15, 15))
    b = r0.copy()
    b[0] += 3e-4
    b[2] -= 2e-4
    orc.apply(b, q, 0.13, 0.1, (15, 15, 15))
    for row in (1, 3, 5):
        d_a = a[row] - r0[row]
        d_b = b[row] - r0[row]
        assert np.max(np.abs(d_a - d_b)) <= 1e-6 * np.max(np.abs(d_a))
