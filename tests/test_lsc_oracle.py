"""Pins of the LSC oracle (oracle/lsc_oracle.py): golden vectors written by
the reference's LSC.apply (ocelot/cpbd/sc.py:538-600), the impedance against
independent evaluations of exp(x) E1(x), the reactive (lossless) character of
the wake, the wake of a sinusoidal density modulation against the impedance at
that frequency, and the half-open slice window on a hand-made bunch."""
import os

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.special import exp1

from oracle import lsc_oracle as L

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "lsc_golden.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("case", ["gauss", "horns"])
@pytest.mark.parametrize("prof", ["gauss", "step"])
def test_oracle_matches_reference(golden, case, prof):
    g = golden
    r, q, E = g[f"{case}/r"], g[f"{case}/q"], float(g[f"{case}/E"])
    t = f"{case}/{prof}"
    rr, dbg = r.copy(), {}
    L.apply(rr, q, E, 0.5, step_profile=(prof == "step"), debug=dbg)
    Q = q.sum()
    assert dbg["sigma"] == pytest.approx(float(g[t + "/sigma"]), rel=1e-15)
    np.testing.assert_array_equal(dbg["B"][:, 0], g[t + "/s"])
    bunch = dbg["B"][:, 1] / (Q * L.C)
    assert np.abs(bunch - g[t + "/bunch"]).max() <= 1e-14 * np.abs(g[t + "/bunch"]).max()
    wake = -dbg["W"] / Q
    assert np.abs(wake - g[t + "/wake"]).max() <= 1e-13 * np.abs(g[t + "/wake"]).max()
    kick = g[t + "/delta"] - r[5]
    assert np.abs(kick).max() > 1e-7                        # the kick is not trivially zero
    assert np.abs((rr[5] - r[5]) - kick).max() <= 1e-12 * np.abs(kick).max()
    np.testing.assert_array_equal(rr[:5], r[:5])           # only delta moves


def test_oracle_undulator_factor(golden):
    """The undulator factor 1 + K^2 fill/2 (sc.py:436-441) with the K_max and
    filling factor that the reference computed for this step."""
    g = golden
    r, q, E = g["gauss/r"], g["gauss/q"], float(g["gauss/E"])
    K, ff = float(g["und/Kmax"][0]), float(g["und/fill"][0])
    rr = r.copy()
    L.apply(rr, q, E, 0.1, K_max=K, fill_factor=ff)
    kick = g["und/delta"] - r[5]
    assert np.abs((rr[5] - r[5]) - kick).max() <= 1e-12 * np.abs(kick).max()
    rr0 = r.copy()
    L.apply(rr0, q, E, 0.1)
    assert np.abs(rr0[5] - r[5]).max() < 0.8 * np.abs(kick).max()   # the factor matters here


def _T(x):
    """exp(x) E1(x) through imp_lsc: choose w so that alpha^2 = x."""
    gamma, sigma, dz = 3.0, 1e-3, 1.0
    w = np.sqrt(np.asarray(x, dtype=float)) * gamma * L.C / sigma
    Z = L.imp_lsc(gamma, sigma, w, dz)
    return (Z / (1j * L.Z0 / (4 * L.PI * L.C * gamma ** 2) * w * dz)).real


def test_impedance_series_branch_against_exp1():
    # above alpha^2 = 40 the reference switches to ten terms of the asymptotic
    # series; its truncation error there is below 10!/40^10 ~ 3e-10 relative
    x = np.linspace(40.5, 900.0, 60)
    lo = x < 600
    ref = np.empty_like(x)
    ref[lo] = np.exp(x[lo]) * exp1(x[lo])
    ref[~lo] = [_T_quad(v) for v in x[~lo]]
    assert np.abs(_T(x) / ref - 1).max() < 4e-10


def _T_quad(v):
    # exp(x) E1(x) = int_0^inf exp(-x u)/(1 + u) du, an independent form
    return quad(lambda u: np.exp(-v * u) / (1 + u), 0, max(80.0 / v, 1.0), limit=400, epsabs=0,
                epsrel=1e-13)[0] + (quad(lambda u: np.exp(-v * u) / (1 + u), 80.0 / v, np.inf, epsabs=0,
                                         epsrel=1e-13)[0] if v > 80 else 0.0)


def test_impedance_exp1_branch_against_quadrature():
    x = np.array([1e-12, 1e-6, 1e-3, 0.1, 0.7, 1.0, 3.0, 12.0, 39.5])
    ref = np.array([quad(lambda u, v=v: np.exp(-v * u) / (1 + u), 0, np.inf, limit=400, epsabs=0,
                         epsrel=1e-13)[0] for v in x])
    assert np.abs(_T(x) / ref - 1).max() < 1e-10
    assert _T(np.array([5e-17]))[0] == 0.0                     # below 1e-16: zero


@pytest.mark.parametrize("step_profile", [False, True])
def test_wake_is_reactive(step_profile):
    """A purely imaginary, odd impedance does no net work on the bunch: the
    profile-weighted wake sums to zero up to the Nyquist term."""
    s = np.linspace(-1e-3, 1e-3, 801)
    b = np.exp(-0.5 * (s / 2e-4) ** 2) * (1 + 0.3 * np.sin(s / 7e-5))
    b /= np.trapezoid(b, s) if hasattr(np, "trapezoid") else np.trapz(b, s)
    w = L.wake_lsc(s, b, 300.0, 1e-4, 1.0, step_profile)
    assert np.abs(w).max() > 0
    assert abs(np.dot(b, w)) < 1e-9 * np.linalg.norm(b) * np.linalg.norm(w)


@pytest.mark.parametrize("step_profile", [False, True])
def test_wake_of_sinusoidal_modulation(step_profile):
    """Line density 1 + eps cos(k0 s) on a long flat bunch: in the middle the
    wake is eps A Im Z(w0) sin(k0 s) per unit charge, A = the flat density
    in 1/s, Z(w0) the impedance at w0 = k0 c (the FFT scaling, the factor c
    of the time profile and the sign convention of the kernel)."""
    nb, ds = 4001, 2.5e-6
    s = (np.arange(nb) - nb // 2) * ds
    Lb = 0.45 * nb * ds
    k0 = 2 * np.pi * 80 / (2 * nb * ds)            # on the doubled FFT grid
    eps = 0.05
    edge = 0.5 * (1 + np.tanh((Lb - np.abs(s)) / (20 * ds)))
    b = edge * (1 + eps * np.cos(k0 * s))
    norm = b.sum() * ds
    b = b / norm
    gamma, sigma = 200.0, 3e-4
    # the wake is linear in the profile: taking away that of the unmodulated
    # bunch (same edges, same norm) leaves the response to the modulation
    w = (L.wake_lsc(s, b, gamma, sigma, 1.0, step_profile)
         - L.wake_lsc(s, edge / norm, gamma, sigma, 1.0, step_profile))
    w0 = k0 * L.C
    Z = (L.imp_step_lsc(gamma, sigma, np.array([w0]), 1.0) if step_profile
         else L.imp_lsc(gamma, sigma, np.array([w0]), 1.0))[0]
    assert Z.real == 0 and Z.imag > 0
    mid = np.abs(s) < 0.25 * Lb
    # the convolution with the impedance maps eps A cos(w0 t) to
    # -eps A Im Z(w0) sin(w0 t) (Z(-w) = conj Z(w)); wake_lsc returns minus it
    pred = eps * (L.C / norm) * Z.imag * np.sin(k0 * s)
    assert np.abs(w[mid] - pred[mid]).max() < 1e-5 * np.abs(pred[mid]).max()


def test_window_is_half_open():
    # mean 0 and std 1 exactly: the window [-1, 1) takes the -1s, not the +1s
    r = np.zeros((6, 8))
    r[4] = [-1, -1, 1, 1, -1, 1, -1, 1]
    r[0] = [1, 2, 30, 40, 3, 50, 4, 60]
    r[2] = [5, 6, 70, 80, 7, 90, 8, 100]
    sig = L.window_sigma(r, bounds=(-1.0, 1.0))
    assert sig == pytest.approx((np.std([1, 2, 3, 4]) + np.std([5, 6, 7, 8])) / 2, rel=1e-15)
    rb = L.window_sigma(r, bounds=(-1.0, 1.0), step_profile=True)
    assert rb == pytest.approx(min(3.0, 3.0) / 2, rel=1e-15)


def test_dz_below_threshold_is_noop(golden):
    r = golden["gauss/r"].copy()
    L.apply(r, golden["gauss/q"], 0.13, 5e-11)
    np.testing.assert_array_equal(r, golden["gauss/r"])
