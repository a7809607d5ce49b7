"""FP64 CPU oracle for the tail of Ocelot's CSR kick.

TEST INFRASTRUCTURE ONLY.  A numpy restatement of ``csr_convolution``
(ocelot/cpbd/csr.py:69-79) and of the last steps of ``CSR.apply``
(csr.py:1105-1143): the convolution of the equidistant line charge with the
reversed, step-averaged CSR kernel, scaled by delta_s/st, interpolated at each
particle's z = -tau on the mesh index axis and added to delta.  Only
``tests/`` may import it; the product path never does.

Pinning (tests/test_csr_oracle.py): golden vectors written by the reference
(``tests/golden/make_golden_csr.py``: csr_convolution on random vectors of six
lengths, and three complete CSR.apply calls with their recorded line charge,
kernel and kicked delta), and the convolution against numpy.convolve.
"""
import numpy as np

M_E_GEV = 9.10938215e-31 * 299792458.0 ** 2 / 1.6021766208e-19 / 1e+9


def csr_convolution(a, b):
    """Linear convolution of a (P) and b (Q) by zero-padded FFTs of length
    2^ceil(log2(P + Q - 1)); the first P + Q - 2 terms (csr.py:69-79)."""
    P, Q = len(a), len(b)
    L = P + Q - 1
    K = 2 ** int(np.ceil(np.log2(L)))
    c = np.fft.ifft(np.fft.fft(np.pad(a, (0, K - P))) * np.fft.fft(np.pad(b, (0, K - Q))))
    return c[0:L - 1].real


def csr_kick(r, E, lam_ds, K1, s1, s2, Ns, delta_s):
    """delta += interp(z/st - sa/st, 0..L-1, lam_K1) 1e-9/pc_ref in place, with
    st = (s2 - s1)/Ns, sa = s1 + st/2, lam_K1 = conv(lam_ds, K1 reversed)/st
    delta_s and z = -tau (csr.py:1105-1143)."""
    st = (s2 - s1) / Ns
    sa = s1 + st / 2.
    lam_K1 = csr_convolution(lam_ds, K1[::-1]) / st * delta_s
    z = -r[4]
    dE = np.interp(z * (1. / st) + (0. - sa / st), np.arange(len(lam_K1)), lam_K1)
    pc_ref = np.sqrt(E ** 2 / M_E_GEV ** 2 - 1) * M_E_GEV
    r[5] += dE * 1e-9 / pc_ref
    return lam_K1
