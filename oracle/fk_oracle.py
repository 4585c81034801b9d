"""numpy restatement of pyseistr's FK dip filter (test infrastructure only:
the test suite imports it).

Follows pyseistr/fk.py:2-68 (fkdip) and 138-142 (nextpow2) with numpy's
float64 FFTs.  Pinned (tests/test_fk_oracle.py) by: a cone that keeps every
bin returning the input unchanged; a flat event (zero wavenumber) passing
and a trace-alternating one (Nyquist wavenumber) being removed; linearity;
and an explicit-DFT computation on tiny sizes with the cone written as
|k| < f * nw / nn1 over centred wavenumbers, which fails for a wrong shift,
a transposed mask or a dropped Hermitian half.
"""
import numpy as np


def nextpow2(n):
    p = 1
    while p < n:
        p *= 2
    return p


def cone(nn1, nk, w):
    """The kept bins (fk.py:49-55): row i1 (1-based frequency) and
    fftshift-ed column i2 pass when i1 > (nn1/nw)(i2 - nk/2) and
    i1 > (nn1/nw)(nk/2 - i2)."""
    nw = w * nk
    nk2 = nk // 2
    i1 = np.arange(1, nn1 + 1)[:, None]
    i2 = np.arange(1, nk + 1)[None, :]
    a = nn1 / nw
    return ((i1 > a * (i2 - nk2)) & (i1 > a * (nk2 - i2))).astype(np.float64)


def fkdip(d, w):
    d = np.asarray(d)
    n1, n2 = d.shape
    nf, nk = nextpow2(n1), nextpow2(n2)
    nf2 = nf // 2
    D = np.fft.fft(d, nf, 0)[: nf2 + 1, :]
    D = np.fft.fftshift(np.fft.fft(D, nk, 1), 1)
    D = D * cone(nf2 + 1, nk, w)
    D = np.fft.ifft(np.fft.ifftshift(D, 1), nk, 1)
    full = np.zeros((nf, nk), dtype=np.complex128)
    full[: nf2 + 1, :] = D
    full[nf2 + 1:, :] = np.conj(np.flipud(D[1:-1, :]))
    return np.real(np.fft.ifft(full, nf, 0))[:n1, :n2]
