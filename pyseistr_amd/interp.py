"""Drop-in structure-oriented interpolation.

soint3dc (pyseistr/soint3d.py:65), soint2dc (pyseistr/soint2d.py:92),
sint3dc / sint2dc (pyseistr/sint.py:97, 61): same signatures, defaults and
output shapes; the arithmetic is csoint3d / csoint2d / csint3d / csint2d of
the C extension (soint3d_cfuns.c, soint2d_cfuns.c), run by the gfx950
library.  Extensions: ``report=True`` also returns the solver iteration count.
soint2dc(prec=1), the reference's preconditioned predictive solver, runs its
trace-to-trace prediction chain in one GPU thread (the chain is sequential)
with the banded factors built in parallel.  drift=1 shifts each PWD stencil
by NINT(p) as the reference does.
"""
import ctypes

import numpy as np

from . import _lib


def _iters():
    return ctypes.c_int(0)


def soint3d_device(d, m, p, q, shape, order=1, niter=100, njs=(1, 1), drift=0, seed=202223, hasmask=1, var=0,
                   out=None, report=False):
    """Device call of soint3dc: flat F-order tensors in and out; with
    report=True also the iterations run (filling it synchronises)."""
    n1, n2, n3 = shape
    _lib.check_sizes("soint3dc", n1 * n2 * n3, data=d, mask=m, dipi=p, dipx=q)
    _lib.check_out(out, n1 * n2 * n3, "soint3dc")
    d, m, p, q = (_lib.to_device(a) for a in (d, m, p, q))
    ctx = _lib.context()
    out = _lib.empty(n1 * n2 * n3) if out is None else out
    it = _iters()
    rc = _lib.load().sh_soint3d(ctx, _lib.ptr(d), _lib.ptr(m), _lib.ptr(p), _lib.ptr(q), n1, n2, n3, int(order),
                                int(njs[0]), int(njs[1]), int(niter), int(drift), int(seed), int(hasmask),
                                float(var), _lib.ptr(out), ctypes.byref(it) if report else None)
    _lib.check(ctx, rc, "soint3dc")
    return (out, it.value) if report else out


def soint3dc(din, mask, dipi, dipx, order=1, niter=100, njs=[1, 1], drift=0, seed=202223, hasmask=1, var=0,
             verb=1, report=False):
    din = np.asarray(din)
    n1, n2, n3 = din.shape
    _lib.check_sizes("soint3dc", n1 * n2 * n3, mask=mask, dipi=dipi, dipx=dipx)
    out, it = soint3d_device(_lib.to_device(din), _lib.to_device(mask), _lib.to_device(dipi), _lib.to_device(dipx),
                             (n1, n2, n3), order, niter, njs, drift, seed, hasmask, var, report=True)
    res = _lib.to_host(out).reshape(n1, n2, n3, order="F")
    return (res, it) if report else res


def soint2dc(din, mask, dip, order=1, niter=100, njs=[1, 1], drift=0, hasmask=1, twoplane=0, prec=0, verb=1,
             report=False):
    din = np.asarray(din)
    n1, n2 = din.shape
    if prec:
        return _soint2d_prec(din, mask, dip, order, niter, hasmask, twoplane, report)
    if twoplane:
        dip = np.asarray(dip)[:, :, 0]
    _lib.check_sizes("soint2dc", n1 * n2, mask=mask, dip=dip)
    ctx = _lib.context()
    d, m, p = _lib.to_device(din), _lib.to_device(mask), _lib.to_device(dip)
    out = _lib.empty(n1 * n2)
    it = _iters()
    rc = _lib.load().sh_soint2d(ctx, _lib.ptr(d), _lib.ptr(m), _lib.ptr(p), n1, n2, int(order), int(njs[0]),
                                int(niter), int(drift), int(hasmask), int(twoplane), _lib.ptr(out), ctypes.byref(it))
    _lib.check(ctx, rc, "soint2dc")
    res = _lib.to_host(out).reshape(n1, n2, order="F")
    return (res, it.value) if report else res


def _soint2d_prec(din, mask, dip, order, niter, hasmask, twoplane, report):
    """soint2dc(prec=1): the preconditioned solve (sh_soint2d_prec); with
    twoplane, dip[:, :, 0] and dip[:, :, 1] are the two slope planes."""
    n1, n2 = din.shape
    dip = np.asarray(dip)
    _lib.check_sizes("soint2dc", n1 * n2, mask=mask)
    _lib.check_sizes("soint2dc", n1 * n2 * (2 if twoplane else 1), dip=dip)
    ctx = _lib.context()
    d, m = _lib.to_device(din), _lib.to_device(mask)
    p = _lib.to_device(dip[:, :, 0] if twoplane else dip)
    q = _lib.to_device(dip[:, :, 1]) if twoplane else None
    out = _lib.empty(n1 * n2)
    it = _iters()
    rc = _lib.load().sh_soint2d_prec(ctx, _lib.ptr(d), _lib.ptr(m), _lib.ptr(p), _lib.ptr(q), n1, n2, int(order),
                                     int(niter), int(hasmask), _lib.ptr(out), ctypes.byref(it))
    _lib.check(ctx, rc, "soint2dc")
    res = _lib.to_host(out).reshape(n1, n2, order="F")
    return (res, it.value) if report else res


def sint3d_device(d, m, p, q, shape, niter=100, eps=0.01, ns1=1, ns2=1, order1=1, order2=1, out=None,
                  report=False):
    """Device call of sint3dc (flat F-order tensors; mask m)."""
    n1, n2, n3 = shape
    _lib.check_sizes("sint3dc", n1 * n2 * n3, data=d, mask=m, dipi=p, dipx=q)
    _lib.check_out(out, n1 * n2 * n3, "sint3dc")
    d, m, p, q = (_lib.to_device(a) for a in (d, m, p, q))
    ctx = _lib.context()
    out = _lib.empty(n1 * n2 * n3) if out is None else out
    it = _iters()
    rc = _lib.load().sh_sint3d(ctx, _lib.ptr(d), _lib.ptr(p), _lib.ptr(q), _lib.ptr(m), n1, n2, n3, int(niter),
                               int(ns1), int(ns2), int(order1), int(order2), float(eps), _lib.ptr(out),
                               ctypes.byref(it) if report else None)
    _lib.check(ctx, rc, "sint3dc")
    return (out, it.value) if report else out


def sint3dc(din, mask, dipi, dipx, niter=100, eps=0.01, ns1=1, ns2=1, order1=1, order2=1, verb=1, report=False):
    din = np.asarray(din)
    n1, n2, n3 = din.shape
    _lib.check_sizes("sint3dc", n1 * n2 * n3, mask=mask, dipi=dipi, dipx=dipx)
    out, it = sint3d_device(_lib.to_device(din), _lib.to_device(mask), _lib.to_device(dipi), _lib.to_device(dipx),
                            (n1, n2, n3), niter, eps, ns1, ns2, order1, order2, report=True)
    res = _lib.to_host(out).reshape(n1, n2, n3, order="F")
    return (res, it) if report else res


def sint2dc(din, mask, dip, niter=100, eps=0.01, ns=1, order=1, verb=1, report=False):
    din = np.asarray(din)
    n1, n2 = din.shape
    _lib.check_sizes("sint2dc", n1 * n2, mask=mask, dip=dip)
    ctx = _lib.context()
    d, m, p = _lib.to_device(din), _lib.to_device(mask), _lib.to_device(dip)
    out = _lib.empty(n1 * n2)
    it = _iters()
    rc = _lib.load().sh_sint2d(ctx, _lib.ptr(d), _lib.ptr(p), _lib.ptr(m), n1, n2, int(niter), int(ns),
                               int(order), float(eps), _lib.ptr(out), ctypes.byref(it))
    _lib.check(ctx, rc, "sint2dc")
    res = _lib.to_host(out).reshape(n1, n2, order="F")
    return (res, it.value) if report else res


soint3d = soint3dc
soint2d = soint2dc
