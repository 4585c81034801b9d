"""GPU parity of patch extraction and reassembly (sh_patch, sh_patch_inv)
against the numpy oracle: bit-exact, dtypes and shapes as the reference's
patch2d / patch3d / patch2d_inv / patch3d_inv (pyseistr/patch.py) give."""
import itertools

import numpy as np
import pytest

from oracle import patch_oracle as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pm():
    import torch
    from pyseistr_amd import _lib, patch
    assert torch.cuda.is_available()
    assert not _lib.missing_symbols()
    return patch


CASES2 = [((100, 60), (8, 8), (4, 4)), ((101, 37), (16, 10), (8, 3)), ((64, 64), (16, 16), (8, 8)),
          ((7, 50), (10, 5), (4, 5)), ((30, 20), (3, 2), (5, 4))]


@pytest.mark.parametrize("n,l,s", CASES2)
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_patch2d_bit_exact(pm, n, l, s, dt):
    rng = np.random.default_rng(n[0] * 1000 + n[1])
    A = rng.standard_normal(n).astype(dt)
    X = pm.patch2d(A, l[0], l[1], s[0], s[1])
    ref = P.cut(A, l, s)
    assert X.shape == ref.shape and X.dtype == ref.dtype
    assert np.array_equal(X, ref)
    back = pm.patch2d_inv(X, n[0], n[1], l[0], l[1], s[0], s[1])
    bref = P.paste(X, n, l, s)
    assert back.shape == n and back.dtype == np.float64
    assert np.array_equal(back, bref, equal_nan=True)
    # patches that are not a cut of anything (each sample's copies differ)
    Y = rng.standard_normal(X.shape).astype(dt)
    assert np.array_equal(pm.patch2d_inv(Y, n[0], n[1], l[0], l[1], s[0], s[1]), P.paste(Y, n, l, s),
                          equal_nan=True)


@pytest.mark.parametrize("pad", list(itertools.product([0, 1], repeat=3)))
def test_patch3d_every_padding_case(pm, pad):
    # l = (6, 4, 4), s = (3, 2, 2): n fits exactly when (n - l) % s == 0
    n = tuple(b + p for b, p in zip((18, 12, 10), pad))
    l, s = (6, 4, 4), (3, 2, 2)
    rng = np.random.default_rng(sum(pad) + 7 * pad[0] + 3 * pad[1])
    A = rng.standard_normal(n).astype(np.float32)
    X = pm.patch3d(A, *l, *s)
    ref = P.cut(A, l, s)[:, :, None]
    assert X.shape == ref.shape and X.dtype == ref.dtype
    assert np.array_equal(X, ref)
    Y = rng.standard_normal(X.shape)
    got = pm.patch3d_inv(Y, *n, *l, *s)
    assert got.shape == n and got.dtype == np.float64
    assert np.array_equal(got, P.paste(Y[:, :, 0], n, l, s))


def test_gaps_and_short_axes(pm):
    # shifts beyond the patch size leave uncovered samples (NaN); an axis
    # shorter than its patch is padded up to it
    rng = np.random.default_rng(11)
    n, l, s = (20, 9, 3), (3, 12, 2), (5, 4, 3)
    A = rng.standard_normal(n)
    X = pm.patch3d(A, *l, *s)
    assert np.array_equal(X[:, :, 0], P.cut(A, l, s))
    back = pm.patch3d_inv(X, *n, *l, *s)
    ref = P.paste(X[:, :, 0], n, l, s)
    assert np.isnan(ref).any()
    assert np.array_equal(back, ref, equal_nan=True)


def test_round_trip_full_size(pm):
    """512 x 256 x 256 in patches of 64 x 32 x 32 with half overlap (3375
    patches, 1.8 GB of float64 on the device): integer-valued data comes
    back exactly from cut then reassembly."""
    import torch
    shape, l, s = (512, 256, 256), (64, 32, 32), (32, 16, 16)
    g = torch.Generator(device="cuda").manual_seed(5)
    a = torch.randint(-2 ** 20, 2 ** 20, (int(np.prod(shape)),), generator=g, device="cuda").double()
    X = pm.patch_device(a, shape, l, s)
    assert tuple(X.shape) == (15 ** 3, 64 * 32 * 32)
    back = pm.patch_inv_device(X, shape, l, s)
    assert torch.equal(back, a)
    # spot check of one patch against the definition
    a3 = a.view(256, 256, 512).permute(2, 1, 0)  # [i1, i2, i3]
    pid, (p1, p2, p3) = 1234, np.unravel_index(1234, (15, 15, 15))
    blk = a3[p1 * 32:p1 * 32 + 64, p2 * 16:p2 * 16 + 32, p3 * 16:p3 * 16 + 32]
    assert torch.equal(X[pid], blk.permute(2, 1, 0).reshape(-1))
    del X, back, a


def test_errors(pm):
    with pytest.raises(ValueError):
        pm.patch2d(np.zeros((3, 4)), 5, 2, 2, 1)
    with pytest.raises(NotImplementedError):
        pm.patch2d(np.zeros((8, 8)), mode=2)
    import torch
    with pytest.raises(ValueError):
        pm.patch_inv_device(torch.zeros((3, 4), dtype=torch.float64, device="cuda"), (8, 8, 1), (4, 4, 1),
                            (2, 2, 1))
