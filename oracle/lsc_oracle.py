"""FP64 CPU oracle for Ocelot's longitudinal space-charge process (LSC).

TEST INFRASTRUCTURE ONLY.  A numpy/scipy restatement of ``LSC.apply`` and
the helpers it calls (ocelot/cpbd/sc.py:261-600), for a beam line without
undulators unless ``K_max``/``fill_factor`` are given (they come from the
lattice scan in ``LSC.prepare``, sc.py:467-503).  Only ``tests/`` may import
it; the product path never does.

Pinning (tests/test_lsc_oracle.py): golden vectors written by the reference
(``tests/golden/make_golden_lsc.py``: Gaussian and step transverse profiles,
intermediate quantities and the kicked delta row), the Gaussian impedance's
large-argument asymptotic series against the ``exp(x) E1(x)`` branch where
both hold, and the wake of a long uniform bunch against the low-frequency
limit of the impedance.
"""
import numpy as np
from scipy.special import exp1, factorial, k1

from oracle.envelope_oracle import s_to_cur

PI = 3.141592653589793
C = 299792458.0
Q_E = 1.6021766208e-19
M_E_GEV = 9.10938215e-31 * C ** 2 / Q_E / 1e+9
MU_0 = 4 * PI * 1e-7
EPSILON_0 = 1 / MU_0 / C ** 2
Z0 = 1. / (C * EPSILON_0)


def imp_lsc(gamma, sigma, w, dz):
    """Round Gaussian beam impedance (sc.py:302-341): i Z0 w T dz/(4 pi c
    gamma^2), T = exp(a^2) E1(a^2) with a = w sigma/(gamma c), the asymptotic
    series of ten terms for a^2 > 40, zero below 1e-16."""
    alpha = w * sigma / (gamma * C)
    alpha2 = alpha * alpha
    inda = np.where(alpha2 > 40.0)[0]
    ind = np.where((alpha2 <= 40.0) & (alpha2 >= 1e-16))[0]
    T = np.zeros(w.shape)
    T[ind] = np.exp(alpha2[ind]) * exp1(alpha2[ind])
    x = alpha2[inda]
    k = 0
    for i in range(10):
        k += (-1) ** i * factorial(i) / (x ** (i + 1))
    T[inda] = k
    return 1j * Z0 / (4 * PI * C * gamma ** 2) * w * T * dz


def imp_step_lsc(gamma, rb, w, dz):
    """Uniform round beam of radius rb (sc.py:343-369): i Z0 c dz (1 - x K1(x))
    /(4 w rb^2), x = w rb/(c gamma); frequencies below 1e-7 give zero."""
    w = np.array(w, dtype=float)
    indx = np.where(w < 1e-7)[0]
    w[indx] = 1e-7
    x = w * rb / (C * gamma)
    Z = 1j * Z0 * C / (4 * w * rb * rb) * dz * (1 - x * k1(x))
    Z[indx] = 0
    return Z


def wake_lsc(s, bunch, gamma, sigma, dz, step_profile=False, K_max=0.0, fill_factor=0.0):
    """Wake of the normalised current profile (sc.py:411-465): impedance on
    the first half of a doubled frequency grid, times the profile's spectrum
    (zero-padded to twice its length, dt * fft), Hermitian completion, inverse
    real FFT scaled by n df; minus the first half."""
    ds = s[1] - s[0]
    dt = ds / C
    nb = len(s)
    n = nb * 2
    f = 1 / dt * np.arange(0, n) / n
    if step_profile:
        Za = imp_step_lsc(gamma, sigma, f[0:nb] * 2 * PI, dz) * (1 + 0.5 * K_max * K_max * fill_factor)
    else:
        Za = imp_lsc(gamma, sigma, f[0:nb] * 2 * PI, dz) * (1 + 0.5 * K_max * K_max * fill_factor)
    # the profile's spectrum is taken on the padded grid s0 + (k+1) ds, whose
    # first step (and so dt and the frequency step used below) is recomputed
    # from that grid, as the reference does (sc.py:452-463, analysis.py:376-383)
    sb1 = s[0] + np.cumsum(np.ones(n)) * ds
    dt2 = (sb1[1] - sb1[0]) / C
    f2 = 1 / dt2 * np.arange(0, n) / n
    bunch1 = np.append(bunch, np.zeros(nb)) * C
    Zb = dt2 * np.fft.fft(bunch1, n)
    Z = np.zeros(n, dtype=complex)
    Z[0:nb] = Za * Zb[0:nb]
    Z[nb:n] = np.flipud(np.conj(Z[0:nb]))
    df = f2[1] - f2[0]
    w = n * df * np.fft.irfft(Z, n)
    return -w[0:nb]


def window_sigma(r, bounds=(-0.4, 0.4), step_profile=False):
    """Transverse size of the central slice (sc.py:566-575): particles with
    mean + b0 std <= tau < mean + b1 std; the mean of std(x) and std(y), or
    for the step profile the smaller half-extent of x and y."""
    tau = r[4]
    m, s = np.mean(tau), np.std(tau)
    sel = (tau >= m + s * bounds[0]) & (tau < m + s * bounds[1])
    x, y = r[0][sel], r[2][sel]
    if step_profile:
        return min(np.max(x) - np.min(x), np.max(y) - np.min(y)) / 2
    return (np.std(x) + np.std(y)) / 2.


def apply(r, q_array, E, dz, step_profile=False, smooth_param=0.1, bounds=(-0.4, 0.4), K_max=0.0,
          fill_factor=0.0, debug=None):
    """One LSC kick on rparticles r (6, n), in place (sc.py:538-600):
    delta += W(tau) 1e-9 / pc_ref with W the wake (V) interpolated linearly
    at each particle (numpy.interp: end values outside the grid)."""
    if dz < 1e-10:
        return r
    tau = r[4]
    sigma_tau = np.std(tau)
    sigma = window_sigma(r, bounds, step_profile)
    q = np.sum(q_array)
    gamma = E / M_E_GEV
    v = np.sqrt(1 - 1 / gamma ** 2) * C
    B = s_to_cur(tau, sigma_tau * smooth_param, q, v)
    bunch = B[:, 1] / (q * C)
    x = B[:, 0]
    W = -wake_lsc(x, bunch, gamma, sigma, dz, step_profile, K_max, fill_factor) * q
    dE = np.interp(tau, x, W)
    pc_ref = np.sqrt(E ** 2 / M_E_GEV ** 2 - 1) * M_E_GEV
    r[5] += dE * 1e-9 / pc_ref
    if debug is not None:
        debug.update(sigma=sigma, B=B, W=W)
    return r
