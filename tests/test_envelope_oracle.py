"""The get_envelope oracle against the reference's own output (golden vectors
written by tests/golden/make_golden_envelope.py) and against closed forms."""
import os

import numpy as np
import pytest

from oracle import envelope_oracle as eo

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def env_golden():
    with np.load(os.path.join(HERE, "golden", "envelope_golden.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def settings(g):
    return {"plain": {}, "bounds": {"bounds": list(g["bounds/b"])}, "autodisp": {"auto_disp": True},
            "twsi": {"disp": tuple(g["disp/tws_i"])}, "imax": {"bounds": list(g["imax/b"]), "slice": "Imax"}}


@pytest.mark.parametrize("bunch", ["gauss", "flat"])
def test_current_profile_matches_reference(env_golden, bunch):
    g = env_golden
    tau, q = g[f"{bunch}/r"][4], g[f"{bunch}/q"]
    B = eo.s_to_cur(tau, 0.01 * np.std(tau), np.sum(q))
    ref = g[f"{bunch}/cur"]
    assert B.shape == ref.shape
    assert np.max(np.abs(B[:, 0] - ref[:, 0])) <= 1e-15 * np.max(np.abs(ref[:, 0]))
    assert np.max(np.abs(B[:, 1] - ref[:, 1])) <= 1e-12 * np.max(np.abs(ref[:, 1]))


def test_current_profile_brute_force():
    """Cloud-in-cell weights by a per-particle loop, the smoothing by a direct
    sum; the integral of the current over s is the total charge times v."""
    rng = np.random.default_rng(3)
    A = rng.normal(0, 1e-4, 300)
    sigma = 2e-6
    B = eo.s_to_cur(A, sigma, 1e-9, v=1.0)
    a, ds, n = B[0, 0], B[1, 0] - B[0, 0], B.shape[0]
    C = np.zeros(n)
    for x in A:
        c = (x - a) / ds
        i = int(np.floor(c))
        C[i] += 1 + i - c
        C[i + 1] += c - i
    K = int(np.floor(3 * sigma / ds + 0.5))
    G = np.exp(-0.5 * (np.arange(-K, K + 1) * ds / sigma) ** 2)
    G /= G.sum()
    S = np.array([sum(C[j + K - m] * G[m] for m in range(2 * K + 1) if 0 <= j + K - m < n) for j in range(n)])
    S *= 1e-9 / (ds * S.sum())
    assert np.max(np.abs(B[:, 1] - S)) <= 1e-9 * np.max(S)
    assert np.sum(B[:, 1]) * ds == pytest.approx(1e-9, rel=1e-12)


@pytest.mark.parametrize("bunch", ["gauss", "flat"])
@pytest.mark.parametrize("setting", ["plain", "bounds", "autodisp", "twsi", "imax"])
def test_oracle_matches_reference(env_golden, bunch, setting):
    g = env_golden
    attrs = [str(a) for a in g["meta/attrs"]]
    r, q, E = g[f"{bunch}/r"], g[f"{bunch}/q"], float(g[f"{bunch}/E"])
    got = eo.get_envelope(r, q, E, **settings(g)[setting])
    got["E"] = E
    ref = g[f"{bunch}/{setting}"]
    for a, v in zip(attrs, ref):
        assert np.isclose(got[a], v, rtol=1e-12, atol=1e-12 * (abs(v) + 1e-30)), (a, got[a], v)


def test_fewer_than_three_particles(env_golden):
    g = env_golden
    got = eo.get_envelope(g["two/r"], g["two/q"], 0.13)
    ref = dict(zip([str(a) for a in g["meta/attrs"]], g["two/plain"]))
    assert got["count"] == 2 and "xx" not in got
    assert got["q"] == ref["q"] and got["p"] == ref["p"]


def test_exact_point_set_closed_form():
    """u = mu + L z over the 8 points +-2 e_i has mean mu and covariance L L^T
    exactly; with p = 0 and px, py ~ 1e-7 the px/py factor is 1 to 1e-14, so
    every moment, emittance and beta/alpha follows in closed form."""
    z = np.concatenate([2 * np.eye(4), -2 * np.eye(4)])       # (8, 4): x, px, y, py
    L = np.array([[3e-4, 0, 0, 0], [-2e-7, 1e-7, 0, 0], [1e-4, 0, 2e-4, 0], [0, 5e-8, -1e-7, 2e-7]])
    mu = np.array([1e-5, 2e-8, -3e-5, 1e-8])
    u = mu + z @ L.T
    r = np.zeros((6, 8))
    r[0], r[1], r[2], r[3] = u[:, 0], u[:, 1], u[:, 2], u[:, 3]
    r[4] = np.array([1, -1, 2, -2, 3, -3, 0.5, -0.5]) * 1e-4
    C = L @ L.T                                                 # x, px, y, py order
    t = eo.get_envelope(r, np.full(8, 1e-12), 0.5)
    tol = 1e-12
    assert abs(t["x"] - mu[0]) <= tol * 1e-4 and abs(t["y"] - mu[2]) <= tol * 1e-4
    for key, (i, j) in {"xx": (0, 0), "xpx": (0, 1), "pxpx": (1, 1), "yy": (2, 2), "ypy": (2, 3),
                        "pypy": (3, 3), "xy": (0, 2), "pxpy": (1, 3), "xpy": (0, 3), "ypx": (2, 1)}.items():
        assert abs(t[key] - C[i, j]) <= 1e-10 * np.sqrt(C[i, i] * C[j, j]), key
    ex = np.sqrt(C[0, 0] * C[1, 1] - C[0, 1] ** 2)
    assert abs(t["emit_x"] - ex) <= 1e-9 * ex
    assert abs(t["beta_x"] - C[0, 0] / ex) <= 1e-9 * C[0, 0] / ex
    assert abs(t["alpha_x"] + C[0, 1] / ex) <= 1e-9 * abs(C[0, 1] / ex)
    assert t["tautau"] == pytest.approx(np.mean(r[4] ** 2), rel=1e-14)
    assert t["q"] == pytest.approx(8e-12, rel=1e-15)


def test_eigen_emittances_against_eigvals():
    """Eigen-emittances = positive imaginary parts of eig(Sigma S), for an
    uncoupled beam the two plane emittances."""
    rng = np.random.default_rng(5)
    r = rng.standard_normal((6, 20000)) * np.array([[1e-4], [1e-5], [2e-4], [3e-6], [1e-3], [1e-3]])
    r[2] += 0.5 * r[0]                                          # couple x into y
    r[3] += 0.1 * r[1]
    t = eo.get_envelope(r, np.ones(20000), 0.2)
    S = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [-1, 0, 0, 0], [0, -1, 0, 0]], dtype=float)
    w = np.sort(np.abs(np.linalg.eigvals(eo.sigma_matrix(t) @ S).imag))[::-1]
    assert t["eigemit_1"] == pytest.approx(w[0], rel=1e-9)
    assert t["eigemit_2"] == pytest.approx(w[2], rel=1e-9)
    ru = rng.standard_normal((6, 20000)) * np.array([[1e-4], [1e-5], [2e-4], [3e-6], [1e-3], [1e-4]])
    tu = eo.get_envelope(ru, np.ones(20000), 0.2)
    assert tu["eigemit_1"] == pytest.approx(max(tu["emit_x"], tu["emit_y"]), rel=1e-3)
    assert tu["eigemit_2"] == pytest.approx(min(tu["emit_x"], tu["emit_y"]), rel=1e-3)


def test_auto_disp_recovers_dispersion():
    """x = x0 + D p with x0 orthogonal to p (exact design): auto_disp gives D."""
    n = 4000
    p = np.repeat([-1.0, 1.0], n // 2) * 1e-3
    x0 = np.tile([2e-4, -2e-4], n // 2)
    r = np.zeros((6, n))
    r[5] = p
    r[0] = x0 + 0.7 * p
    r[1] = -0.02 * p + np.tile([1e-5, 1e-5, -1e-5, -1e-5], n // 4)
    r[2] = np.tile([1e-4, 1e-4, -1e-4, -1e-4], n // 4) - 0.3 * p
    t = eo.get_envelope(r, np.ones(n), 0.1, auto_disp=True)
    assert t["Dx"] == pytest.approx(0.7, rel=1e-12)
    assert t["Dxp"] == pytest.approx(-0.02, rel=1e-12)
    assert t["Dy"] == pytest.approx(-0.3, rel=1e-12)
    assert abs(t["xx"] - 4e-8) <= 1e-20


def test_bounds_selection_brute_force():
    rng = np.random.default_rng(9)
    r = rng.standard_normal((6, 3001)) * 1e-3
    q = rng.uniform(1, 2, 3001)
    tau = r[4]
    m, s = tau.mean(), tau.std()
    cnt = sum(1 for v in tau if m + s * -1.5 <= v <= m + s * 0.5)
    mo = eo.moments(r, q, bounds=[-1.5, 0.5])
    assert mo["count"] == cnt
    assert mo["q"] == pytest.approx(q.sum(), rel=1e-15)
