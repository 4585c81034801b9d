"""Calls into the reference's own C extension modules (test infrastructure).

``build_ref.sh`` compiles dip_cfuns.c, sof_cfuns.c, sof3d_cfuns.c,
soint2d_cfuns.c and soint3d_cfuns.c from the reference sources into
``oracle/_ref``.  This module imports them from there and calls them with the
positional argument lists their PyArg_ParseTuple formats expect (cited per
function), flat Fortran-order float32 in, flat float32 out.  It is used to
pin the C restatement and as the CPU baseline; nothing in the product path
imports it.  ``available()`` is False where the modules were not built.
"""
import contextlib
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_REFDIR = os.path.join(_HERE, "_ref")
_mods = {}


def available():
    try:
        _mod("dipcfun")
        return True
    except ImportError:
        return False


def _mod(name):
    if name not in _mods:
        if _REFDIR not in sys.path:
            sys.path.insert(0, _REFDIR)
        _mods[name] = __import__(name)
    return _mods[name]


@contextlib.contextmanager
def _quiet():
    """The reference prints from C (printf); silence fd 1 around calls."""
    import ctypes
    sys.stdout.flush()
    libc = ctypes.CDLL(None)
    libc.fflush(None)
    saved = os.dup(1)
    null = os.open(os.devnull, os.O_WRONLY)
    os.dup2(null, 1)
    try:
        yield
    finally:
        libc.fflush(None)  # C stdio buffers the reference's printf output
        os.dup2(saved, 1)
        os.close(saved)
        os.close(null)


def _f(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32).ravel(order="F"))


def cbp(din, shape, d1, flo=0.0, fhi=0.5, nplo=6, nphi=6, phase=0):
    """bpcfun.cbp, format "Oiiifffiiii" = (din, n1, n2, n3, d1, flo, fhi,
    nplo, nphi, phase, verb) (bp_cfuns.c:540)."""
    n1, n2, n3 = shape
    with _quiet():
        out = _mod("bpcfun").cbp(_f(din), n1, n2, n3, float(d1), float(flo), float(fhi), int(nplo), int(nphi),
                                 int(phase), 0)
    return np.asarray(out, np.float32).ravel()


def smoothcf(din, shape, rect=(1, 1, 1), diff=(0, 0, 0), box=(0, 0, 0), repeat=1, adj=1):
    """dipcfun.smoothcf, format "Oiiiiiiiiiiiiii" = (din, n1, n2, n3, repeat,
    adj, r1, r2, r3, diff1, diff2, diff3, box1, box2, box3)
    (dip_cfuns.c:2021)."""
    n1, n2, n3 = shape
    with _quiet():
        out = _mod("dipcfun").smoothcf(_f(din), n1, n2, n3, int(repeat), int(adj), *(int(r) for r in rect),
                                       *(int(v) for v in diff), *(int(v) for v in box))
    return np.asarray(out, np.float32).ravel()


def dipc(din, shape, niter=5, liter=20, order=2, eps_dv=0.01, rect=(1, 1, 1), mask=None):
    """dipcfun.dipc, format "Oiiiiiifffiiiii" (dip_cfuns.c:1708)."""
    n1, n2, n3 = shape
    d = _f(din)
    hasmask = 0
    if mask is not None:
        d = np.concatenate([d, _f(mask)])
        hasmask = 1
    with _quiet():
        return _mod("dipcfun").dipc(d, n1, n2, n3, niter, liter, order, eps_dv, 1.0, 1e-6,
                                    rect[0], rect[1], rect[2], hasmask, 0)


def somean2d(dn, dip, shape, ns, order=2, eps=0.01, adj=0):
    """sofcfun.csomean2d, format "OOiiiiiifi" (sof_cfuns.c:1448)."""
    n1, n2, n3 = shape
    with _quiet():
        return _mod("sofcfun").csomean2d(_f(dn), _f(dip), n1, n2, n3, ns, order, adj, eps, 0)


def somf2d(dn, dip, shape, ns, order=2, eps=0.01, option=1, nmf=None):
    """sofcfun.csomf2d, format "OOiiiiiiifi" (sof_cfuns.c:1551)."""
    n1, n2, n3 = shape
    nmf = 2 * ns + 1 if nmf is None else nmf
    with _quiet():
        return _mod("sofcfun").csomf2d(_f(dn), _f(dip), n1, n2, n3, ns, nmf, option, order, eps, 0)


def somean3d(dn, dipi, dipx, shape, r1, r2, order=2, eps=0.01):
    """sof3dcfun.csomean3d, format "OOOiiiiiifi" (sof3d_cfuns.c:1373)."""
    n1, n2, n3 = shape
    with _quiet():
        return _mod("sof3dcfun").csomean3d(_f(dn), _f(dipi), _f(dipx), n1, n2, n3, r1, r2, order, eps, 0)


def somf3d(dn, dipi, dipx, shape, r1, r2, order=2, eps=0.01, option=1, nmf=None):
    """sof3dcfun.csomf3d, format "OOOiiiiiiiifi" (sof3d_cfuns.c:1573)."""
    n1, n2, n3 = shape
    nmf = 2 * r1 * r2 + 1 if nmf is None else nmf
    with _quiet():
        return _mod("sof3dcfun").csomf3d(_f(dn), _f(dipi), _f(dipx), n1, n2, n3, r1, r2, nmf, option,
                                         order, eps, 0)


def soint3d(din, mask, dipi, dipx, shape, order=1, niter=100, njs=(1, 1), drift=0, seed=202223,
            hasmask=1, var=0.0):
    """soint3dcfun.csoint3d, format "OOOOiiiiiiiiiifi" (soint3d_cfuns.c:2427)."""
    n1, n2, n3 = shape
    with _quiet():
        return _mod("soint3dcfun").csoint3d(_f(din), _f(mask), _f(dipi), _f(dipx), n1, n2, n3, order,
                                            njs[0], njs[1], niter, drift, seed, hasmask, var, 0)


def sint3d(din, mask, dipi, dipx, shape, niter=100, eps=0.01, ns1=1, ns2=1, order1=1, order2=1):
    """soint3dcfun.csint3d, format "OOOOiiiiiiiiif" (soint3d_cfuns.c:2530)."""
    n1, n2, n3 = shape
    with _quiet():
        return _mod("soint3dcfun").csint3d(_f(din), _f(dipi), _f(dipx), _f(mask), n1, n2, n3, niter, ns1,
                                           ns2, order1, order2, 0, eps)


def soint2d(din, mask, dip, shape, order=1, niter=100, njs=(1, 1), drift=0, hasmask=1, twoplane=0,
            prec=0, dip2=None):
    """soint2dcfun.csoint2d, format "OOOOiiiiiiiiiii" (soint2d_cfuns.c:2277)."""
    n1, n2 = shape
    with _quiet():
        return _mod("soint2dcfun").csoint2d(_f(din), _f(mask), _f(dip), _f(dip if dip2 is None else dip2), n1, n2,
                                            order, njs[0],
                                            njs[1], niter, drift, hasmask, twoplane, prec, 0)


def sint2d(din, mask, dip, shape, niter=100, eps=0.01, ns=1, order=1):
    """soint2dcfun.csint2d, format "OOOiiiiiif" (soint2d_cfuns.c:2436)."""
    n1, n2 = shape
    with _quiet():
        return _mod("soint2dcfun").csint2d(_f(din), _f(dip), _f(mask), n1, n2, niter, ns, order, 0, eps)
