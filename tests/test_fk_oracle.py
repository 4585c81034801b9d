"""Pins of the FK dip filter oracle (oracle/fk_oracle.py): closed-form
cases, linearity and an explicit-DFT restatement on tiny sizes."""
import numpy as np
import pytest

from oracle import fk_oracle as F


def test_nextpow2():
    assert [F.nextpow2(n) for n in (1, 2, 3, 64, 65, 1000)] == [1, 2, 4, 64, 128, 1024]


@pytest.mark.parametrize("shape", [(100, 37), (64, 64), (5, 3), (1, 9)])
def test_open_cone_is_identity(shape):
    # a cone wide enough to keep every bin gives the input back
    d = np.random.default_rng(1).standard_normal(shape)
    assert np.abs(F.fkdip(d, 1e9) - d).max() <= 1e-12 * np.abs(d).max()


def test_flat_event_passes_and_alternating_event_is_removed():
    # 64 x 64, no padding: a sine of 5 periods (frequency bin 5) constant
    # across traces sits at wavenumber 0 and passes; the same sine with the
    # sign alternating trace to trace sits at the Nyquist wavenumber and is
    # cut.  With w = 0.45 the cone slope is 33 / (0.45 * 64) = 1.15.
    t = np.arange(64)
    s = np.sin(2 * np.pi * 5 * t / 64)
    flat = np.repeat(s[:, None], 64, axis=1)
    alt = flat * (-1.0) ** np.arange(64)[None, :]
    assert np.abs(F.fkdip(flat, 0.45) - flat).max() < 1e-12
    assert np.abs(F.fkdip(alt, 0.45)).max() < 1e-12


def test_linear():
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal((2, 50, 30))
    lhs = F.fkdip(2.5 * x - 0.75 * y, 0.05)
    rhs = 2.5 * F.fkdip(x, 0.05) - 0.75 * F.fkdip(y, 0.05)
    assert np.abs(lhs - rhs).max() <= 1e-12 * np.abs(lhs).max()


def _explicit(d, w):
    """The filter with DFT matrices: bins (f, k), f = 0..nf/2, centred
    wavenumber kc in [-nk/2, nk/2); the reference's cone keeps row f when
    f + 1 > (nn1 / nw) |kc + 1|.  Negative frequencies are the conjugates
    of the positive ones (the input is real)."""
    n1, n2 = d.shape
    nf, nk = F.nextpow2(n1), F.nextpow2(n2)
    nn1 = nf // 2 + 1
    pad = np.zeros((nf, nk))
    pad[:n1, :n2] = d
    Et = np.exp(-2j * np.pi * np.outer(np.arange(nf), np.arange(nf)) / nf)
    Ex = np.exp(-2j * np.pi * np.outer(np.arange(nk), np.arange(nk)) / nk)
    D = Et @ pad @ Ex.T
    kc = np.where(np.arange(nk) < nk // 2, np.arange(nk), np.arange(nk) - nk)
    a = nn1 / (w * nk)
    keep = (np.arange(nn1)[:, None] + 1) > a * np.abs(kc + 1)[None, :]
    G = np.zeros((nf, nk), complex)
    G[:nn1] = D[:nn1] * keep
    H = G @ Ex.conj() / nk                      # inverse along traces
    for f in range(nn1, nf):
        H[f] = np.conj(H[nf - f])
    return np.real(Et.conj() @ H / nf)[:n1, :n2]


@pytest.mark.parametrize("shape,w", [((6, 5), 0.3), ((8, 8), 0.1), ((11, 3), 0.6), ((7, 16), 0.05)])
def test_matches_explicit_dft(shape, w):
    d = np.random.default_rng(sum(shape)).standard_normal(shape)
    got, ref = F.fkdip(d, w), _explicit(d, w)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    # and the cone is not trivial at these sizes
    assert np.abs(got - d).max() > 1e-3
