"""Drop-in plane-wave spray and structure-oriented filters.

Signatures, defaults and output shapes follow pyseistr:
  pwspray2d  (pyseistr/pwspray2d.py:130)   somean2dc (pyseistr/somean2d.py:36)
  pwspray3d  (pyseistr/pwspray3d.py:2)     somean3dc (pyseistr/somean3d.py:36)
  somf2dc    (pyseistr/somf2d.py:60)       somf3dc   (pyseistr/somf3d.py:54)
The arithmetic is the C extension's (sof_cfuns.c, sof3d_cfuns.c), run by the
gfx950 library; outputs are float32.
"""
import numpy as np

from . import _lib


def _shape3(dn):
    dn = np.asarray(dn)
    if dn.ndim == 2:
        return dn.shape[0], dn.shape[1], 1
    return dn.shape


def pwspray2d(u1, dip, nr, order, eps):
    """2-D plane-wave spray (forward).  Returns the flat spray volume
    u[(i2*(2nr+1) + slot)*n1 + i1] of length n1*n2*(2nr+1)."""
    u1 = np.asarray(u1)
    n1, n2 = u1.shape
    ns2 = 2 * nr + 1
    _lib.check_sizes("pwspray2d", n1 * n2, dip=dip)
    ctx = _lib.context()
    x = _lib.to_device(u1)
    d = _lib.to_device(dip)
    u = _lib.empty(n1 * n2 * ns2)
    _lib.check(ctx, _lib.load().sh_pwspray2d(ctx, 0, _lib.ptr(d), n1, n2, int(nr), int(order), float(eps),
                                             _lib.ptr(x), _lib.ptr(u)), "pwspray2d")
    return _lib.to_host(u)


def pwspray2d_adjoint(u, dip, nr, order, eps, shape):
    """Adjoint of pwspray2d: flat spray volume -> [n1, n2] (pwspray_lop adj,
    pyseistr/src/sof_cfuns.c:958-996)."""
    n1, n2 = shape
    _lib.check_sizes("pwspray2d_adjoint", n1 * n2, dip=dip)
    _lib.check_sizes("pwspray2d_adjoint", n1 * n2 * (2 * int(nr) + 1), u=u)
    ctx = _lib.context()
    du = _lib.to_device(np.asarray(u, np.float32).ravel())
    d = _lib.to_device(dip)
    x = _lib.empty(n1 * n2)
    _lib.check(ctx, _lib.load().sh_pwspray2d(ctx, 1, _lib.ptr(d), n1, n2, int(nr), int(order), float(eps),
                                             _lib.ptr(x), _lib.ptr(du)), "pwspray2d")
    return _lib.to_host(x).reshape(n1, n2, order="F")


def pwspray3d(din, dipi, dipx, ns2, ns3, order, eps):
    """3-D plane-wave spray; returns u of shape (n1, np, n2, n3),
    np = (2ns2+1)(2ns3+1), slot = k2 + (2ns2+1) k3."""
    n1, n2, n3 = _shape3(din)
    np_ = (2 * ns2 + 1) * (2 * ns3 + 1)
    _lib.check_sizes("pwspray3d", n1 * n2 * n3, dipi=dipi, dipx=dipx)
    ctx = _lib.context()
    d, p, q = _lib.to_device(din), _lib.to_device(dipi), _lib.to_device(dipx)
    u = _lib.empty(n1 * n2 * n3 * np_)
    _lib.check(ctx, _lib.load().sh_pwspray3d(ctx, _lib.ptr(d), _lib.ptr(p), _lib.ptr(q), n1, n2, n3, int(ns2),
                                             int(ns3), int(order), float(eps), _lib.ptr(u)), "pwspray3d")
    return _lib.to_host(u).reshape(n1, np_, n2, n3, order="F")


def somean2dc(dn, dip, ns, order, eps, adj=0, verb=1):
    """Plane-wave smoothing (normalised triangle-weighted spray average)."""
    dn = np.asarray(dn)
    n1, n2, n3 = _shape3(dn)
    _lib.check_sizes("somean2dc", n1 * n2 * n3, dip=dip)
    out = somean2d_device(_lib.to_device(dn), _lib.to_device(dip), (n1, n2, n3), ns, order, eps, adj)
    ds = _lib.to_host(out).reshape(n1, n2, n3, order="F")
    return np.squeeze(ds) if n3 == 1 else ds


def somean2d_device(d, p, shape, ns, order=2, eps=0.01, adj=0, out=None):
    """Device call of somean2dc: flat F-order tensors in and out."""
    n1, n2, n3 = shape
    _lib.check_sizes("somean2dc", n1 * n2 * n3, data=d, dip=p)
    _lib.check_out(out, n1 * n2 * n3, "somean2dc")
    d, p = _lib.to_device(d), _lib.to_device(p)
    ctx = _lib.context()
    out = _lib.empty(n1 * n2 * n3) if out is None else out
    _lib.check(ctx, _lib.load().sh_somean2d(ctx, _lib.ptr(d), _lib.ptr(p), n1, n2, n3, int(ns), int(order),
                                            int(adj), float(eps), _lib.ptr(out)), "somean2dc")
    return out


def somf2d_device(d, p, shape, ns, order=2, eps=0.01, option=1, out=None):
    """Device call: flat F-order tensors; nmf = 2ns+1 as somf2dc passes it."""
    n1, n2, n3 = shape
    _lib.check_sizes("somf2dc", n1 * n2 * n3, data=d, dip=p)
    _lib.check_out(out, n1 * n2 * n3, "somf2dc")
    d, p = _lib.to_device(d), _lib.to_device(p)
    ctx = _lib.context()
    out = _lib.empty(n1 * n2 * n3) if out is None else out
    _lib.check(ctx, _lib.load().sh_somf2d(ctx, _lib.ptr(d), _lib.ptr(p), n1, n2, n3, int(ns), 2 * int(ns) + 1,
                                          int(option), int(order), float(eps), _lib.ptr(out)), "somf2dc")
    return out


def somf2dc(dn, dip, ns, order, eps, option=1, verb=1):
    """Structure-oriented median filter over the 2ns+1 sprayed traces."""
    dn = np.asarray(dn)
    n1, n2, n3 = _shape3(dn)
    _lib.check_sizes("somf2dc", n1 * n2 * n3, dip=dip)
    d, p = _lib.to_device(dn), _lib.to_device(dip)
    out = somf2d_device(d, p, (n1, n2, n3), ns, order, eps, option)
    ds = _lib.to_host(out).reshape(n1, n2, n3, order="F")
    return np.squeeze(ds) if n3 == 1 else ds


def somean3d_device(d, p, q, shape, r1, r2, eps=0.01, order=2, out=None):
    n1, n2, n3 = shape
    _lib.check_sizes("somean3dc", n1 * n2 * n3, data=d, dipi=p, dipx=q)
    _lib.check_out(out, n1 * n2 * n3, "somean3dc")
    d, p, q = _lib.to_device(d), _lib.to_device(p), _lib.to_device(q)
    ctx = _lib.context()
    out = _lib.empty(n1 * n2 * n3) if out is None else out
    _lib.check(ctx, _lib.load().sh_somean3d(ctx, _lib.ptr(d), _lib.ptr(p), _lib.ptr(q), n1, n2, n3, int(r1), int(r2),
                                            float(eps), int(order), _lib.ptr(out)), "somean3dc")
    return out


def somf3d_device(d, p, q, shape, r1, r2, eps=0.01, order=2, option=1, nmf=None, out=None):
    n1, n2, n3 = shape
    nmf = 2 * r1 * r2 + 1 if nmf is None else nmf
    _lib.check_sizes("somf3dc", n1 * n2 * n3, data=d, dipi=p, dipx=q)
    _lib.check_out(out, n1 * n2 * n3, "somf3dc")
    d, p, q = _lib.to_device(d), _lib.to_device(p), _lib.to_device(q)
    ctx = _lib.context()
    out = _lib.empty(n1 * n2 * n3) if out is None else out
    _lib.check(ctx, _lib.load().sh_somf3d(ctx, _lib.ptr(d), _lib.ptr(p), _lib.ptr(q), n1, n2, n3, int(r1), int(r2),
                                          int(nmf), int(option), int(order), float(eps), _lib.ptr(out)), "somf3dc")
    return out


def somean3dc(dn, dipi, dipx, r1, r2, eps, order, verb=0):
    """3-D structure-oriented mean (eps fixed at 0.01 as in the reference)."""
    n1, n2, n3 = _shape3(dn)
    out = somean3d_device(dn, dipi, dipx, (n1, n2, n3), r1, r2, eps, order)
    return _lib.to_host(out).reshape(n1, n2, n3, order="F")


def somf3dc(dn, dipi, dipx, r1, r2, eps, order, option=1, verb=1):
    """3-D structure-oriented median over the 2*r1*r2+1 slots around the
    centre of the (2r1+1)(2r2+1) sprayed traces (eps fixed at 0.01)."""
    n1, n2, n3 = _shape3(dn)
    out = somf3d_device(dn, dipi, dipx, (n1, n2, n3), r1, r2, eps, order, option)
    return _lib.to_host(out).reshape(n1, n2, n3, order="F")
