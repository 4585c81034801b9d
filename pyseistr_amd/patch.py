"""Patch extraction and reassembly on the GPU: drop-ins for pyseistr's
patch2d, patch3d, patch2d_inv and patch3d_inv (pyseistr/patch.py:2,86,208,
307), computed by sh_patch / sh_patch_inv in libseistr_hip.so.

Each axis is zero-padded so that (n - l) is a multiple of s; patches are
numbered with the first axis outermost, each one a row of l1*l2(*l3) values
in Fortran order.  The inverse averages the patches covering each sample,
summing in float64 in patch order as the reference does, so results are
identical to the reference's.  Only mode 1 exists (the reference writes no
other)."""
import ctypes

import numpy as np

from . import _lib


def patch_count(shape, l, s):
    """Number of patches of size l, shift s, over a volume of shape (n1, n2, n3)."""
    out = ctypes.c_long(0)
    rc = _lib.load().sh_patch_count(*(int(v) for v in shape), *(int(v) for v in l), *(int(v) for v in s),
                                    ctypes.byref(out))
    if rc != 0:
        raise ValueError(f"patch: no patch of size {tuple(l)} fits shape {tuple(shape)} with shift {tuple(s)}")
    return int(out.value)


def _dev64(a):
    torch = _lib._torch()
    flat = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))
    return torch.from_numpy(flat).cuda()


def patch_device(din, shape, l, s, out=None):
    """Device-level cut: flat F-order float64 CUDA tensor in, (npatch, L)
    float64 CUDA tensor out."""
    torch = _lib._torch()
    np_ = patch_count(shape, l, s)
    L = int(np.prod(l))
    if din.dtype != torch.float64 or din.numel() != int(np.prod(shape)) or not din.is_contiguous():
        raise ValueError(f"patch: expected a contiguous float64 volume of {int(np.prod(shape))} samples")
    if out is None:
        out = torch.empty((np_, L), dtype=torch.float64, device="cuda")
    ctx = _lib.context()
    rc = _lib.load().sh_patch(ctx, _lib.ptr(din), _lib.ptr(out), *(int(v) for v in shape),
                              *(int(v) for v in l), *(int(v) for v in s))
    _lib.check(ctx, rc, "patch")
    return out


def patch_inv_device(X, shape, l, s, out=None):
    """Device-level reassembly: (npatch, L) float64 CUDA tensor in, flat
    F-order float64 CUDA tensor of the volume out."""
    torch = _lib._torch()
    np_ = patch_count(shape, l, s)
    L = int(np.prod(l))
    if tuple(X.shape) != (np_, L) or X.dtype != torch.float64 or not X.is_contiguous():
        raise ValueError(f"patch_inv: expected contiguous float64 patches of shape ({np_}, {L}), "
                         f"got {tuple(X.shape)} {X.dtype}")
    if out is None:
        out = torch.empty(int(np.prod(shape)), dtype=torch.float64, device="cuda")
    ctx = _lib.context()
    rc = _lib.load().sh_patch_inv(ctx, _lib.ptr(X), _lib.ptr(out), *(int(v) for v in shape),
                                  *(int(v) for v in l), *(int(v) for v in s))
    _lib.check(ctx, rc, "patch_inv")
    return out


def _padded(shape, l, s):
    return any((n - ll) % ss for n, ll, ss in zip(shape, l, s))


def _cut(A, l, s, mode):
    if mode != 1:
        raise NotImplementedError("patch: only mode 1 exists")
    A = np.asarray(A)
    if np.iscomplexobj(A):
        raise TypeError("patch: complex data are not supported (the device path is real float64)")
    shape = A.shape + (1,) * (3 - A.ndim)
    # the reference pads with float64 zeros, which promotes the whole array
    dtype = np.result_type(A.dtype, np.float64) if _padded(A.shape, l, s) else A.dtype
    X = patch_device(_dev64(A.reshape(shape)), shape, l + (1,) * (3 - len(l)), s + (1,) * (3 - len(s)))
    return X.cpu().numpy().astype(dtype, copy=False)


def _paste(X, n, l, s, mode):
    if mode != 1:
        raise NotImplementedError("patch_inv: only mode 1 exists")
    X = np.asarray(X)
    if np.iscomplexobj(X):
        raise TypeError("patch_inv: complex patches are not supported (the device path is real float64)")
    torch = _lib._torch()
    shape = tuple(int(v) for v in n) + (1,) * (3 - len(n))
    l3, s3 = l + (1,) * (3 - len(l)), s + (1,) * (3 - len(s))
    Xd = torch.from_numpy(np.ascontiguousarray(X.reshape(X.shape[0], -1), np.float64)).cuda()
    out = patch_inv_device(Xd, shape, l3, s3)
    return out.cpu().numpy().reshape(tuple(int(v) for v in n), order="F")


def patch2d(A, l1=8, l2=8, s1=4, s2=4, mode=1):
    """Cut a 2-D array into l1 x l2 patches with shifts s1, s2.  Returns
    (npatch, l1*l2); float64 when padding was needed, else A's dtype."""
    A = np.asarray(A)
    if A.ndim != 2:
        raise ValueError("patch2d: A must be 2-D")
    return _cut(A, (l1, l2), (s1, s2), mode)


def patch3d(A, l1=4, l2=4, l3=4, s1=2, s2=2, s3=2, mode=1):
    """Cut a 3-D array into l1 x l2 x l3 patches.  Returns (npatch,
    l1*l2*l3, 1), as the reference does."""
    A = np.asarray(A)
    if A.ndim != 3:
        raise ValueError("patch3d: A must be 3-D")
    return _cut(A, (l1, l2, l3), (s1, s2, s3), mode)[:, :, None]


def patch2d_inv(X, n1, n2, l1=8, l2=8, s1=4, s2=4, mode=1):
    """Reassemble patch2d's patches into an n1 x n2 float64 array, each
    sample the mean of the patches covering it."""
    return _paste(X, (n1, n2), (l1, l2), (s1, s2), mode)


def patch3d_inv(X, n1, n2, n3, l1=4, l2=4, l3=4, s1=2, s2=2, s3=2, mode=1):
    """Reassemble patch3d's patches into an n1 x n2 x n3 float64 array.

    The reference allocates its padded sum for six of the eight padding
    cases (patch.py:407-429) and raises UnboundLocalError when axes 1 and 3
    need padding but axis 2 does not, or axes 2 and 3 but not axis 1; this
    function handles every case by the same rule."""
    return _paste(X, (n1, n2, n3), (l1, l2, l3), (s1, s2, s3), mode)
