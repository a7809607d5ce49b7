"""Seeded synthetic bunches for the space-charge workloads.

Coordinates are Ocelot's (x [m], x' = px/p0, y, y', tau [m], delta = dE/(p0 c)),
stored as a (6, Np) float64 row-major array like ``ParticleArray.rparticles``
(ocelot/cpbd/beam/particle.py:79-84).  Charges are equal, ``q_array = Q/Np``.

Named workloads (sizes as benchmarked):
  gauss_200k_63   200k particles, 250 pC, 130 MeV, 6D Gaussian, mesh 63^3
  gauss_2m_63     the same bunch with 2M particles
  flattop_10m_127 10M particles, 1 nC, 500 MeV, flat-top 2 mm with 0.2 mm
                  Gaussian edges, chirp h = -10 1/m, mesh 127^3
  batch64         64 Gaussian bunches of lognormal sizes on {31,47,63}^3
  hollow_50m_255  50M particles, 2 nC, 150 MeV, annulus 20-200 um,
                  double-horn current, mesh 255^3
"""
import numpy as np

GAUSS_SIGMAS = (100e-6, 1e-5, 100e-6, 1e-5, 1e-3, 1e-3)


def gaussian_bunch(n, sigmas=GAUSS_SIGMAS, seed=0, mean=None):
    """Uncorrelated 6D Gaussian: row-wise standard normals times sigmas."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((6, n))
    z *= np.asarray(sigmas, dtype=np.float64)[:, None]
    if mean is not None:
        z += np.asarray(mean, dtype=np.float64)[:, None]
    return z


def flattop_bunch(n, seed=0, sigma_t=50e-6, sigma_tp=1e-5, length=2e-3,
                  edge=0.2e-3, chirp=-10.0, sigma_p=5e-4):
    """Transverse Gaussian, longitudinal flat-top with Gaussian edges,
    linear energy chirp delta = chirp * tau plus uncorrelated spread."""
    rng = np.random.default_rng(seed)
    r = np.empty((6, n))
    r[0] = rng.standard_normal(n) * sigma_t
    r[1] = rng.standard_normal(n) * sigma_tp
    r[2] = rng.standard_normal(n) * sigma_t
    r[3] = rng.standard_normal(n) * sigma_tp
    r[4] = rng.uniform(-length / 2, length / 2, n) + rng.standard_normal(n) * edge
    r[5] = chirp * r[4] + rng.standard_normal(n) * sigma_p
    return r


def hollow_bunch(n, seed=0, r_in=20e-6, r_out=200e-6, flat=1e-3, horn=0.5e-3,
                 sigma_horn=20e-6):
    """Uniform transverse annulus; longitudinal mixture 50% flat-top of length
    ``flat`` plus 25% Gaussian spikes at -horn and +horn."""
    rng = np.random.default_rng(seed)
    r = np.empty((6, n))
    rad = np.sqrt(rng.uniform(r_in ** 2, r_out ** 2, n))
    th = rng.uniform(0, 2 * np.pi, n)
    r[0] = rad * np.cos(th)
    r[2] = rad * np.sin(th)
    r[1] = rng.standard_normal(n) * 1e-5
    r[3] = rng.standard_normal(n) * 1e-5
    k = rng.integers(0, 4, n)
    tau = rng.uniform(-flat / 2, flat / 2, n)
    tau = np.where(k == 2, -horn + rng.standard_normal(n) * sigma_horn, tau)
    tau = np.where(k == 3, horn + rng.standard_normal(n) * sigma_horn, tau)
    r[4] = tau
    r[5] = rng.standard_normal(n) * 1e-3
    return r


def equal_charges(n, Q):
    return np.full(n, Q / n, dtype=np.float64)


WORKLOADS = {
    "gauss_200k_63": dict(n=200_000, Q=250e-12, E=0.130, nmesh=(63, 63, 63), kind="gauss", dz=0.1),
    "gauss_2m_63": dict(n=2_000_000, Q=250e-12, E=0.130, nmesh=(63, 63, 63), kind="gauss", dz=0.1),
    "flattop_10m_127": dict(n=10_000_000, Q=1e-9, E=0.5, nmesh=(127, 127, 127), kind="flattop", dz=0.1),
    "hollow_50m_255": dict(n=50_000_000, Q=2e-9, E=0.150, nmesh=(255, 255, 255), kind="hollow", dz=0.05),
}


def make_bunch(name, n=None, seed=0):
    """(rparticles, q_array, E_GeV, nmesh, dz) of a named workload; ``n``
    overrides the particle count (same distribution)."""
    w = WORKLOADS[name]
    n = w["n"] if n is None else int(n)
    if w["kind"] == "gauss":
        r = gaussian_bunch(n, seed=seed)
    elif w["kind"] == "flattop":
        r = flattop_bunch(n, seed=seed)
    else:
        r = hollow_bunch(n, seed=seed)
    return r, equal_charges(n, w["Q"]), w["E"], tuple(w["nmesh"]), w["dz"]


def batch64(seed=0, n_bunches=64):
    """Independent Gaussian bunches of uneven size and mesh: Np lognormal
    clipped to [5e4, 2e6], mesh from {31,47,63}^3, Q ~ U(20,500) pC,
    E ~ U(0.05,1) GeV, sigma_x and sigma_tau scaled by 3^U(-1,1)."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(n_bunches):
        n = int(np.clip(round(np.exp(rng.normal(np.log(3.2e5), 1.0))), 5e4, 2e6))
        m = int(rng.choice([31, 47, 63]))
        Q = rng.uniform(20e-12, 500e-12)
        E = rng.uniform(0.05, 1.0)
        sx = 100e-6 * 3.0 ** rng.uniform(-1, 1)
        st = 1e-3 * 3.0 ** rng.uniform(-1, 1)
        sig = (sx, 1e-5, sx, 1e-5, st, 1e-3)
        out.append(dict(r=gaussian_bunch(n, sigmas=sig, seed=seed * 1000 + b + 1),
                        q=equal_charges(n, Q), E=E, nmesh=(m, m, m)))
    return out
