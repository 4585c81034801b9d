"""GPU parity of the FK dip filter (sh_fkdip, hipFFT in float64) against
the numpy oracle.  The FFTs differ in summation order from numpy's, so the
bar is float64 rounding: max |gpu - oracle| <= 1e-11 * max |oracle| with the
oracle in float64.  numpy 2 transforms float32 input in float32 (the
reference's forward FFTs then round to single precision), so for float32
input the bar against the oracle run on that input is 1e-6."""
import numpy as np
import pytest

from oracle import fk_oracle as F

pytestmark = pytest.mark.gpu
TOL = 1e-11
TOL_F32_INPUT = 1e-6


@pytest.fixture(scope="module")
def fk():
    import torch
    from pyseistr_amd import _lib, fk
    assert torch.cuda.is_available()
    assert not _lib.missing_symbols()
    return fk


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("shape,w", [((1000, 300), 0.02), ((256, 128), 0.005), ((513, 65), 0.3),
                                     ((100, 1), 5.0), ((1, 50), 0.05), ((2, 3), 0.4), ((64, 64), 0.45)])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_fkdip_matches_oracle(fk, shape, w, dt):
    d = np.random.default_rng(shape[0] + 7 * shape[1]).standard_normal(shape).astype(dt)
    got = fk.fkdip(d, w)
    ref = F.fkdip(d.astype(np.float64), w)
    assert got.shape == ref.shape and got.dtype == np.float64
    assert np.abs(ref).max() > 0  # the cone keeps something
    print(shape, w, dt.__name__, "rel err", _rel(got, ref))
    assert _rel(got, ref) <= TOL
    if dt == np.float32:
        assert _rel(got, F.fkdip(d, w)) <= TOL_F32_INPUT


@pytest.mark.parametrize("shape", [(1000, 300), (513, 65), (64, 64)])
def test_one_dimensional_transforms_agree(fk, monkeypatch, shape):
    """SH_FK_1D=1 runs the time and space transforms as separate batched 1-D
    transforms (the path taken anyway when an axis has one sample) instead
    of one 2-D real transform; both meet the float64 bar."""
    d = np.random.default_rng(shape[1]).standard_normal(shape)
    ref = F.fkdip(d, 0.03)
    two = fk.fkdip(d, 0.03)
    monkeypatch.setenv("SH_FK_1D", "1")
    one = fk.fkdip(d, 0.03)
    monkeypatch.delenv("SH_FK_1D")
    assert _rel(one, ref) <= TOL and _rel(two, ref) <= TOL


def test_repeated_and_resized_calls(fk):
    # the plans are cached per size: alternate sizes and repeat
    rng = np.random.default_rng(3)
    for shape in [(300, 200), (64, 48), (300, 200), (64, 48)]:
        d = rng.standard_normal(shape)
        ref = F.fkdip(d, 0.03)
        assert np.abs(fk.fkdip(d, 0.03) - ref).max() <= TOL * np.abs(ref).max()


def test_das_size_properties(fk):
    """A DAS-sized section, 8192 samples x 4096 channels (no padding): an
    open cone returns the input, and the filter is linear."""
    import torch
    n1, n2 = 8192, 4096
    g = torch.Generator(device="cuda").manual_seed(4)
    x = torch.randn(n1 * n2, generator=g, device="cuda", dtype=torch.float64)
    y = torch.randn(n1 * n2, generator=g, device="cuda", dtype=torch.float64)
    same = fk.fkdip_device(x, (n1, n2), 1e12)
    assert float((same - x).abs().max()) <= 1e-12 * float(x.abs().max())
    fx = fk.fkdip_device(x, (n1, n2), 0.02)
    fy = fk.fkdip_device(y, (n1, n2), 0.02)
    fxy = fk.fkdip_device(2.0 * x - 0.5 * y, (n1, n2), 0.02)
    assert float((fxy - (2.0 * fx - 0.5 * fy)).abs().max()) <= 1e-12 * float(fxy.abs().max())
    # the cone removes a real share of the noise's energy
    assert float(fx.norm()) < 0.9 * float(x.norm())


def test_errors(fk):
    with pytest.raises(ValueError):
        fk.fkdip(np.zeros((8, 8)), 0.0)
    with pytest.raises(ValueError):
        fk.fkdip(np.zeros(8), 0.1)
