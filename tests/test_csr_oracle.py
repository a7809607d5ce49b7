"""Pins of the CSR-tail oracle (oracle/csr_oracle.py) against the vectors the
reference wrote (tests/golden/make_golden_csr.py) and numpy.convolve."""
import os

import numpy as np
import pytest

from oracle import csr_oracle as co

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "csr_golden.npz")


@pytest.fixture(scope="module")
def cg():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("k", range(6))
def test_convolution_matches_reference(cg, k):
    a, b, c = cg[f"conv{k}/a"], cg[f"conv{k}/b"], cg[f"conv{k}/c"]
    got = co.csr_convolution(a, b)
    assert got.shape == c.shape == (len(a) + len(b) - 2,)
    np.testing.assert_array_equal(got, c)
    # the first P + Q - 2 terms of the direct linear convolution
    direct = np.convolve(a, b)[:len(a) + len(b) - 2]
    if direct.size:
        assert np.abs(got - direct).max() <= 1e-13 * np.abs(direct).max()


@pytest.mark.parametrize("case", ["bend1", "bend2", "fine"])
def test_kick_matches_reference(cg, case):
    r = cg["bunch/r"].copy()
    E = float(cg["bunch/E"])
    a, b = cg[f"{case}/conv_a"], cg[f"{case}/conv_b"]
    np.testing.assert_array_equal(a, cg[f"{case}/lam_ds"])
    lam_K1 = co.csr_kick(r, E, cg[f"{case}/lam_ds"], b[::-1], float(cg[f"{case}/s1"]), float(cg[f"{case}/s2"]),
                         float(cg[f"{case}/Ns"]), float(cg[f"{case}/delta_s"]))
    assert len(lam_K1) == len(cg[f"{case}/conv_c"])
    kick = cg[f"{case}/delta"] - cg["bunch/r"][5]
    assert np.abs(kick).max() > 1e-6
    assert np.abs((r[5] - cg["bunch/r"][5]) - kick).max() <= 1e-13 * np.abs(kick).max()
    np.testing.assert_array_equal(r[:5], cg["bunch/r"][:5])
