"""Drop-in dip estimation: dip2dc / dip3dc (and the plain names dip2d / dip3d).

Same signatures, defaults and output shapes as pyseistr.dip2dc
(pyseistr/dip2d.py:115) and pyseistr.dip3dc (pyseistr/dip3d.py:59); the
computation is the C extension's dipc (pyseistr/src/dip_cfuns.c:1694), run by
the gfx950 library.  The pure-Python twins dip2d/dip3d of the reference are
a different algorithm (no line search, interleaved 3-D directions); the plain
names here follow the C semantics.

Extensions beyond the reference: ``init`` (initial dip, the warm start) and
``report`` (return the per-iteration record).
"""
import ctypes

import numpy as np

from . import _lib


def dip_device(din, shape, niter=5, liter=20, order=2, eps_dv=0.01, rect=(1, 1, 1), mask=None,
               init=None, out=None, report=False):
    """Device-level call: flat F-order float32 CUDA tensors in and out.

    Returns the flat dip tensor (N, or 2N = inline || crossline for n3 > 1),
    and the iteration record when report=True."""
    n1, n2, n3 = (int(s) for s in shape)
    n = n1 * n2 * n3
    nd = 1 if n3 == 1 else 2
    _lib.check_sizes("dip", n, data=din, mask=mask)
    _lib.check_sizes("dip", nd * n, init=init)
    _lib.check_out(out, nd * n, "dip")
    ctx = _lib.context()
    lib = _lib.load()
    if out is None:
        out = _lib.empty(nd * n)
    rep = _lib.DipReport() if report else None
    rc = lib.sh_dip(ctx, _lib.ptr(din), _lib.ptr(mask), _lib.ptr(init), n1, n2, n3, int(niter), int(liter),
                    int(order), float(eps_dv), int(rect[0]), int(rect[1]), int(rect[2]), _lib.ptr(out),
                    ctypes.byref(rep) if rep is not None else None)
    _lib.check(ctx, rc, "dip")
    if report:
        m = min(int(niter), 64)  # the record keeps the first 64 iterations
        rec = {"cg_iters": [[rep.cg_iters[d][k] for k in range(m)] for d in range(nd)],
               "ls_step": [[rep.ls_step[d][k] for k in range(m)] for d in range(nd)],
               "elapsed_ms": float(rep.elapsed_ms)}
        return out, rec
    return out


def _run(din, shape, niter, liter, order, eps_dv, rect, mask, init, report):
    n = int(np.prod(shape))
    nd = 1 if shape[2] == 1 else 2
    if mask is not None and _lib.nsamples(mask) != n:
        raise ValueError("Mask and Data should have the same dimension")
    _lib.check_sizes("dip", n, data=din)
    _lib.check_sizes("dip", nd * n, init=init)
    d = _lib.to_device(din)
    m = _lib.to_device(mask) if mask is not None else None
    ini = _lib.to_device(init) if init is not None else None
    res = dip_device(d, shape, niter, liter, order, eps_dv, rect, m, ini, report=report)
    if report:
        return _lib.to_host(res[0]), res[1]
    return _lib.to_host(res)


def dip2dc(din, niter=5, liter=20, order=2, eps_dv=0.01, eps_cg=1, tol_cg=0.000001, rect=[10, 10, 1], verb=1,
           mask=None, init=None, report=False):
    """2-D local slope (samples per trace) of din[n1, n2].

    eps_cg and tol_cg are accepted and ignored, as in the reference (the C
    solver hard-codes 1 and 1e-6); eps_dv has no effect either (see
    sh_dip).  init: optional initial dip [n1, n2]."""
    din = np.asarray(din)
    n1, n2 = din.shape[0], din.shape[1]
    res = _run(din, (n1, n2, 1), niter, liter, order, eps_dv, rect, mask, init, report)
    if report:
        return res[0].reshape(n1, n2, order="F"), res[1]
    return res.reshape(n1, n2, order="F")


def dip3dc(din, niter=5, liter=10, order=2, eps_dv=0.01, eps_cg=1, tol_cg=0.000001, rect=[5, 5, 5], verb=1, runc=1,
           mask=None, init=None, report=False):
    """3-D inline and crossline slopes of din[n1, n2, n3] -> (dip_i, dip_x).

    init: optional (dip_i, dip_x) pair or array [n1, n2, n3, 2] used as the
    starting dips instead of zeros (warm start)."""
    din = np.asarray(din)
    n1, n2, n3 = din.shape
    if init is not None:
        if isinstance(init, (tuple, list)):
            init = np.stack([np.asarray(init[0]), np.asarray(init[1])], axis=-1)
        init = np.asarray(init, dtype=np.float32).reshape(n1, n2, n3, 2, order="F")
    res = _run(din, (n1, n2, n3), niter, liter, order, eps_dv, rect, mask, init, report)
    dip = (res[0] if report else res).reshape(n1, n2, n3, 2, order="F")
    out = (dip[:, :, :, 0], dip[:, :, :, 1])
    return (out, res[1]) if report else out


# The reference's plain names; here they carry the C extension's semantics.
dip2d = dip2dc
dip3d = dip3dc
