"""numpy restatement of pyseistr's patch extraction and reassembly (test
infrastructure only: the test suite and the smoke check import it).

Follows pyseistr/patch.py: patch2d (2-77), patch3d (86-201), patch2d_inv
(208-300), patch3d_inv (307-446), mode 1.  Arrays are numpy arrays indexed
[i1, i2, i3]; a patch is stored as one row of l1*l2*l3 values in Fortran
order (i1 fastest), and patches are numbered with the first axis outermost
and the last innermost.

Pinned (tests/test_patch_oracle.py) by: the reassembly of a cut giving back
integer-valued data exactly; a hand-worked cut and reassembly of a 5 x 3
array; the inner-product identity <cut(a), X> = <a, cut^T(X)> with cut^T
computed as reassembly times the coverage count, which fails for a
transposed patch or a wrong patch numbering; coverage counts summing to
npatch * L; and NaN exactly on the samples no patch covers (s > l).
"""
import numpy as np


def padded_size(n, l, s):
    """patch.py:66-72 / 174-184: pad so that (n - l) is a multiple of s
    (Python's modulo, so n < l pads up to the next fitting size)."""
    t = (n - l) % s
    return n + (s - t if t else 0)


def counts(shape, l, s):
    """Patches per axis (the reference's range(0, N - l + 1, s))."""
    return tuple(len(range(0, padded_size(n, ll, ss) - ll + 1, ss)) for n, ll, ss in zip(shape, l, s))


def cut(A, l, s):
    """Patches of A (any number of axes, 1-3): (npatch, prod(l)) float64
    when padding was needed (the reference concatenates float64 zeros),
    else A's dtype (patch.py:74-77, 186-193)."""
    A = np.asarray(A)
    N = tuple(padded_size(n, ll, ss) for n, ll, ss in zip(A.shape, l, s))
    if N != A.shape:
        P = np.zeros(N, dtype=np.result_type(A.dtype, np.float64))
        P[tuple(slice(0, n) for n in A.shape)] = A
    else:
        P = A
    starts = [range(0, Ni - ll + 1, ss) for Ni, ll, ss in zip(N, l, s)]
    rows = []
    for corner in np.ndindex(*(len(r) for r in starts)):  # last axis fastest
        sl = tuple(slice(r[c], r[c] + ll) for r, c, ll in zip(starts, corner, l))
        rows.append(P[sl].ravel(order="F"))
    if not rows:
        raise ValueError("no patch fits")
    return np.array(rows)


def paste(X, n, l, s):
    """Reassembly (patch.py:274-296, 407-441): add the patches in patch
    order into a float64 array of the padded size, count the patches per
    sample, divide, crop to n.  Samples no patch covers come out NaN."""
    X = np.asarray(X)
    N = tuple(padded_size(nn, ll, ss) for nn, ll, ss in zip(n, l, s))
    A = np.zeros(N)
    mask = np.zeros(N)
    starts = [range(0, Ni - ll + 1, ss) for Ni, ll, ss in zip(N, l, s)]
    for pid, corner in enumerate(np.ndindex(*(len(r) for r in starts))):
        sl = tuple(slice(r[c], r[c] + ll) for r, c, ll in zip(starts, corner, l))
        A[sl] = A[sl] + np.reshape(X[pid].ravel(), l, order="F")
        mask[sl] = mask[sl] + 1.0
    with np.errstate(invalid="ignore", divide="ignore"):
        A = A / mask
    return A[tuple(slice(0, nn) for nn in n)]


def coverage(n, l, s):
    """Number of patches covering each sample of the (unpadded) volume."""
    N = tuple(padded_size(nn, ll, ss) for nn, ll, ss in zip(n, l, s))
    mask = np.zeros(N)
    starts = [range(0, Ni - ll + 1, ss) for Ni, ll, ss in zip(N, l, s)]
    for corner in np.ndindex(*(len(r) for r in starts)):
        mask[tuple(slice(r[c], r[c] + ll) for r, c, ll in zip(starts, corner, l))] += 1.0
    return mask[tuple(slice(0, nn) for nn in n)]
