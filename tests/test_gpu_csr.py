"""CSR kick tail on the GPU (ocsc_csr_* through the C ABI) against the
reference's own output (tests/golden/make_golden_csr.py: csr_convolution on
random vectors and three complete CSR.apply calls) and the oracle.
Tolerances relative to the largest value: convolution 1e-13 (direct sums
against the reference's FFT), kick 1e-12."""
import os

import numpy as np
import pytest

from oracle import csr_oracle as co

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


@pytest.fixture(scope="module")
def cg():
    with np.load(os.path.join(HERE, "golden", "csr_golden.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


@pytest.mark.parametrize("k", range(6))
def test_csr_convolution_matches_reference(cg, k):
    from ocelot_amd.csr import csr_convolution
    a, b, c = cg[f"conv{k}/a"], cg[f"conv{k}/b"], cg[f"conv{k}/c"]
    got = csr_convolution(a, b)
    assert got.shape == c.shape
    if c.size:
        assert rel(got, c) < 1e-13


@pytest.mark.parametrize("case", ["bend1", "bend2", "fine"])
def test_csr_kick_matches_reference(cg, case):
    from ocelot_amd import ParticleArray
    from ocelot_amd.csr import csr_kick
    print(f"[csr golden] {case}", flush=True)
    r0 = cg["bunch/r"]
    p = ParticleArray()
    p.rparticles, p.q_array, p.E = r0.copy(), cg["bunch/q"].copy(), float(cg["bunch/E"])
    lam_K1 = csr_kick(p, cg[f"{case}/lam_ds"], cg[f"{case}/conv_b"][::-1], float(cg[f"{case}/s1"]),
                      float(cg[f"{case}/s2"]), float(cg[f"{case}/Ns"]), float(cg[f"{case}/delta_s"]))
    st = (float(cg[f"{case}/s2"]) - float(cg[f"{case}/s1"])) / float(cg[f"{case}/Ns"])
    assert rel(lam_K1, cg[f"{case}/conv_c"] / st * float(cg[f"{case}/delta_s"])) < 1e-13
    kick = cg[f"{case}/delta"] - r0[5]
    assert rel(p.rparticles[5] - r0[5], kick) < 1e-12
    np.testing.assert_array_equal(p.rparticles[:5], r0[:5])


def test_csr_kick_resident_bunch_against_oracle(cg):
    """2M resident particles (tensor rows) with the fine-mesh line charge and
    kernel, some particles outside the mesh on both sides (held end values)."""
    from ocelot_amd import DeviceParticleArray
    from ocelot_amd.csr import csr_kick
    rng = np.random.default_rng(2)
    n = 2_000_000
    r = rng.normal(size=(6, n)) * np.array([[5e-5], [1e-5], [5e-5], [1e-5], [4e-5], [1e-4]])
    lam, K1 = cg["fine/lam_ds"], cg["fine/conv_b"][::-1].copy()
    s1, s2, Ns = float(cg["fine/s1"]), float(cg["fine/s2"]), float(cg["fine/Ns"])
    ref = r.copy()
    co.csr_kick(ref, 0.13, lam, K1, s1, s2, Ns, 0.1)
    d = DeviceParticleArray(torch.from_numpy(r.copy()).cuda(), torch.full((n,), 1e-15, dtype=torch.float64).cuda(),
                            0.13)
    csr_kick(d, lam, K1, s1, s2, Ns, 0.1)
    got = d.rparticles.cpu().numpy()
    assert rel(got[5] - r[5], ref[5] - r[5]) < 1e-12
    np.testing.assert_array_equal(got[:5], r[:5])


def test_csr_short_inputs():
    from ocelot_amd.csr import csr_convolution
    assert csr_convolution([1.0], [2.0]).shape == (0,)
    np.testing.assert_allclose(csr_convolution([1.0, 2.0], [3.0]), [3.0])
