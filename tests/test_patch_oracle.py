"""Pins of the patch oracle (oracle/patch_oracle.py) by hand-worked
examples and by identities the patch operators must satisfy."""
import numpy as np
import pytest

from oracle import patch_oracle as P


def _grid(n1, n2):
    i, j = np.meshgrid(np.arange(n1), np.arange(n2), indexing="ij")
    return (10 * i + j).astype(np.float64)


def test_cut_by_hand():
    # 5 x 3, patches 3 x 2, shifts 2 x 1: corners (0,0) (0,1) (2,0) (2,1),
    # each patch listed down its columns (first axis fastest)
    X = P.cut(_grid(5, 3), (3, 2), (2, 1))
    assert X.tolist() == [[0, 10, 20, 1, 11, 21], [1, 11, 21, 2, 12, 22],
                          [20, 30, 40, 21, 31, 41], [21, 31, 41, 22, 32, 42]]


def test_cut_pads_with_zeros_and_promotes():
    # 4 x 2 needs one padded row for patches of 3 with shift 2
    A = _grid(4, 2).astype(np.float32)
    X = P.cut(A, (3, 2), (2, 1))
    assert X.dtype == np.float64
    assert X.tolist() == [[0, 10, 20, 1, 11, 21], [20, 30, 0, 21, 31, 0]]
    # no padding keeps the dtype
    assert P.cut(_grid(5, 3).astype(np.float32), (3, 2), (2, 1)).dtype == np.float32


def test_paste_by_hand():
    # patch k filled with k + 1: each sample is the mean over its patches
    X = np.repeat(np.arange(1.0, 5.0)[:, None], 6, axis=1)
    A = P.paste(X, (5, 3), (3, 2), (2, 1))
    assert A.tolist() == [[1, 1.5, 2], [1, 1.5, 2], [2, 2.5, 3], [3, 3.5, 4], [3, 3.5, 4]]


def test_gaps_are_nan():
    # patches of 2 every 3 samples leave samples 2 and 5 uncovered
    A = P.paste(np.ones((3, 2)), (7, 1), (2, 1), (3, 1))
    assert np.isnan(A[:, 0]).tolist() == [False, False, True, False, False, True, False]
    assert P.counts((7, 1), (2, 1), (3, 1)) == (3, 1)


@pytest.mark.parametrize("n,l,s", [((9, 7), (4, 3), (2, 2)), ((10, 6, 5), (4, 3, 2), (3, 2, 2)),
                                   ((6, 5, 7), (6, 2, 3), (1, 2, 2)), ((3, 5, 4), (5, 5, 2), (3, 3, 1))])
def test_round_trip_exact(n, l, s):
    rng = np.random.default_rng(sum(n))
    a = rng.integers(-1000, 1000, size=n).astype(np.float64)
    assert np.array_equal(P.paste(P.cut(a, l, s), n, l, s), a)


@pytest.mark.parametrize("n,l,s", [((9, 7), (4, 3), (2, 2)), ((10, 6, 5), (4, 3, 2), (3, 2, 2)),
                                   ((11, 4, 6), (3, 4, 4), (3, 1, 3))])
def test_adjoint_identity(n, l, s):
    # <cut(a), X> = <a, cut^T X>, with cut^T X = paste(X) * coverage
    rng = np.random.default_rng(3)
    a = rng.standard_normal(n)
    np_ = int(np.prod(P.counts(n, l, s)))
    X = rng.standard_normal((np_, int(np.prod(l))))
    lhs = float(np.sum(P.cut(a, l, s) * X))
    rhs = float(np.sum(a * P.paste(X, n, l, s) * P.coverage(n, l, s)))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_coverage_counts_every_patch_sample_once():
    n, l, s = (10, 6, 5), (4, 3, 2), (3, 2, 2)
    N = tuple(P.padded_size(*t) for t in zip(n, l, s))
    full = P.coverage(N, l, s)  # no padding needed at the padded size
    assert full.sum() == np.prod(P.counts(n, l, s)) * np.prod(l)


def test_no_patch_fits_raises():
    # (3 - 5) % 2 == 0: no padding, and 3 samples hold no patch of 5
    with pytest.raises(ValueError):
        P.cut(np.zeros((3, 4)), (5, 2), (2, 1))
