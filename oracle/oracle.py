"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle is the C restatement in ``seistr_oracle.c`` built twice:
``lib/liboracle_f32.so`` follows the reference C extension's FP32 operation
order, ``lib/liboracle_f64.so`` is the same algorithm in float64.  Only the
test suite, the smoke check and the CPU-baseline leg of the benchmark import
this module.  All arrays are flat Fortran-order (n1 fastest), as the
reference's C entry points take them.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def build():
    """Compile both oracle libraries (gcc, seconds)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


# ORACLE_SANITIZED=1 loads the AddressSanitizer / UBSan builds
# (make -C oracle sanitize; the process needs the ASan runtime preloaded)
_SUFFIX = "_san" if os.environ.get("ORACLE_SANITIZED") == "1" else ""


def _lib(dtype):
    key = np.dtype(dtype).itemsize
    if key not in _LIBS:
        name = ("liboracle_f32" if key == 4 else "liboracle_f64") + _SUFFIX + ".so"
        path = os.path.join(_HERE, "lib", name)
        if not os.path.exists(path):
            build()
        _LIBS[key] = ctypes.CDLL(path)
    return _LIBS[key]


def _real(dtype):
    return ctypes.c_float if np.dtype(dtype).itemsize == 4 else ctypes.c_double


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _arr(x, dtype):
    return np.ascontiguousarray(np.asarray(x).ravel(order="F"), dtype=dtype)


def passfilter(p, nw=2, dtype=np.float32):
    lib = _lib(dtype)
    a = np.zeros(2 * nw + 1, dtype)
    lib.or_passfilter(ctypes.c_int(nw), _real(dtype)(p), _ptr(a))
    return a


def aderfilter(p, nw=2, dtype=np.float32):
    lib = _lib(dtype)
    a = np.zeros(2 * nw + 1, dtype)
    lib.or_aderfilter(ctypes.c_int(nw), _real(dtype)(p), _ptr(a))
    return a


def pwd(x, p, shape, axis, nw=2, nj=1, drift=0, der=0, dtype=np.float32):
    """allpass1 (axis 2) / allpass2 (axis 3) on a flat F-order volume."""
    n1, n2, n3 = (int(v) for v in shape)
    x = _arr(x, dtype); p = _arr(p, dtype)
    y = np.zeros_like(x)
    _lib(dtype).or_pwd(n1, n2, n3, axis, nw, nj, drift, der, _ptr(x), _ptr(p), _ptr(y))
    return y


def allpass3(x_or_y, pp, qq, shape, adj=0, nw=2, nj1=1, nj2=1, drift=0, dtype=np.float32):
    n1, n2, n3 = (int(v) for v in shape)
    n = n1 * n2 * n3
    pp = _arr(pp, dtype); qq = _arr(qq, dtype)
    if adj:
        y = _arr(x_or_y, dtype); x = np.zeros(n, dtype)
    else:
        x = _arr(x_or_y, dtype); y = np.zeros(2 * n, dtype)
    _lib(dtype).or_allpass3_lop(adj, 0, n1, n2, n3, nw, nj1, nj2, drift, _ptr(pp), _ptr(qq), _ptr(x), _ptr(y))
    return x if adj else y


def allpass21(x_or_y, pp, shape, adj=0, nw=2, nj=1, drift=0, dtype=np.float32):
    n1, n2 = (int(v) for v in shape)
    pp = _arr(pp, dtype)
    a = _arr(x_or_y, dtype); b = np.zeros(n1 * n2, dtype)
    x, y = (b, a) if adj else (a, b)
    _lib(dtype).or_allpass21_lop(adj, 0, n1, n2, nw, nj, drift, _ptr(pp), _ptr(x), _ptr(y))
    return x if adj else y


def trianglen(x, shape, rect, dtype=np.float32):
    x = _arr(x, dtype).copy()
    _lib(dtype).or_trianglen(*(int(v) for v in shape), *(int(r) for r in rect), _ptr(x))
    return x


def bandpass(x, shape, d1, flo=0.0, fhi=0.5, nplo=6, nphi=6, phase=0, dtype=np.float32):
    """cbp (bp_cfuns.c:524): flat F-order in and out (traces of n1)."""
    x = _arr(x, dtype).copy()
    n1 = int(shape[0])
    ntr = int(np.prod(shape)) // n1
    r = _real(dtype)
    lib = _lib(dtype)
    lib.or_bandpass.argtypes = [ctypes.c_int, ctypes.c_int, r, r, r, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_void_p]
    if lib.or_bandpass(n1, ntr, d1, flo, fhi, int(nplo), int(nphi), int(phase), _ptr(x)) != 0:
        raise ValueError("too many poles")
    return x


def smoothc(x, shape, rect=(1, 1, 1), diff=(0, 0, 0), box=(0, 0, 0), repeat=1, adj=1, dtype=np.float32):
    """smoothcf (dip_cfuns.c:2006): flat F-order in and out."""
    x = _arr(x, dtype).copy()
    iv = lambda v: (ctypes.c_int * 3)(*(int(a) for a in v))  # noqa: E731
    _lib(dtype).or_smoothc(*(int(v) for v in shape), int(repeat), int(adj), iv(rect), iv(diff), iv(box), _ptr(x))
    return x


def dip(din, shape, niter=5, liter=20, order=2, eps_dv=0.01, rect=(1, 1, 1), mask=None,
        p_init=None, dtype=np.float32, report=False):
    """dipc: returns flat dips (N for n3==1, else 2N = inline || crossline)."""
    n1, n2, n3 = (int(v) for v in shape)
    n = n1 * n2 * n3
    nd = 1 if n3 == 1 else 2
    din = _arr(din, dtype)
    m = _arr(mask, dtype) if mask is not None else None
    pi = _arr(p_init, dtype) if p_init is not None else None
    out = np.zeros(nd * n, dtype)
    rep = np.zeros(2 * niter * nd, np.int32)
    _lib(dtype).or_dip(_ptr(din), _ptr(m), n1, n2, n3, int(niter), int(liter), int(order), _real(dtype)(eps_dv),
                       *(int(r) for r in rect), _ptr(pi), _ptr(out), _ptr(rep))
    return (out, rep.reshape(nd, niter, 2)) if report else out


def predict_step(trace, pp, adj=0, forw=0, nw=2, e=1e-4, dtype=np.float32):
    t = _arr(trace, dtype).copy(); p = _arr(pp, dtype)
    _lib(dtype).or_predict_step(adj, forw, t.size, nw, _real(dtype)(e), _ptr(t), _ptr(p))
    return t


def predict1(tr1, pp, forw=0, nw=2, e=1e-4, dtype=np.float32):
    t = _arr(tr1, dtype); p = _arr(pp, dtype); o = np.zeros_like(t)
    _lib(dtype).or_predict1(forw, t.size, nw, _real(dtype)(e), _ptr(t), _ptr(p), _ptr(o))
    return o


def predict2(tr1, tr2, pp1, pp2, forw1=0, forw2=0, nw=2, e=1e-4, dtype=np.float32):
    a = _arr(tr1, dtype); b = _arr(tr2, dtype); p = _arr(pp1, dtype); q = _arr(pp2, dtype)
    o = np.zeros_like(a)
    _lib(dtype).or_predict2(forw1, forw2, a.size, nw, _real(dtype)(e), _ptr(a), _ptr(b), _ptr(p), _ptr(q), _ptr(o))
    return o


def pwspray2d(x, dip, shape, ns, order=2, eps=0.01, adj=0, dtype=np.float32):
    """Forward: x (n1*n2) -> u (n1*(2ns+1)*n2); adjoint: u -> x."""
    n1, n2 = (int(v) for v in shape)
    ns2 = 2 * ns + 1
    d = _arr(dip, dtype)
    if adj:
        u = _arr(x, dtype); xx = np.zeros(n1 * n2, dtype)
    else:
        xx = _arr(x, dtype).copy(); u = np.zeros(n1 * n2 * ns2, dtype)
    _lib(dtype).or_pwspray2d(adj, n1, n2, ns, order, _real(dtype)(eps), _ptr(d), _ptr(xx), _ptr(u))
    return xx if adj else u


def pwsmooth2d(x, dip, shape, ns, order=2, eps=0.01, adj=0, dtype=np.float32):
    n1, n2 = (int(v) for v in shape)
    d = _arr(dip, dtype)
    a = _arr(x, dtype).copy(); b = np.zeros(n1 * n2, dtype)
    trace, smooth = (b, a) if adj else (a, b)
    _lib(dtype).or_pwsmooth2d(adj, _ptr(d), n1, n2, ns, order, _real(dtype)(eps), _ptr(trace), _ptr(smooth))
    return trace if adj else smooth


def somean2d(dn, dip, shape, ns, order=2, eps=0.01, adj=0, dtype=np.float32):
    n1, n2, n3 = (int(v) for v in shape)
    a = _arr(dn, dtype); d = _arr(dip, dtype); o = np.zeros_like(a)
    _lib(dtype).or_somean2d(_ptr(a), _ptr(d), n1, n2, n3, ns, order, adj, _real(dtype)(eps), _ptr(o))
    return o


def somf2d(dn, dip, shape, ns, order=2, eps=0.01, option=1, nmf=None, dtype=np.float32):
    n1, n2, n3 = (int(v) for v in shape)
    a = _arr(dn, dtype); d = _arr(dip, dtype); o = np.zeros_like(a)
    nmf = 2 * ns + 1 if nmf is None else nmf
    _lib(dtype).or_somf2d(_ptr(a), _ptr(d), n1, n2, n3, ns, nmf, option, order, _real(dtype)(eps), _ptr(o))
    return o


def spray3_plan(ns2, ns3):
    np_ = (2 * ns2 + 1) * (2 * ns3 + 1)
    out = np.zeros(4 * np_, np.int32)
    _lib(np.float32).or_spray3_plan(ns2, ns3, _ptr(out))
    return out.reshape(np_, 4)


def pwspray3d(dn, dipi, dipx, shape, ns2, ns3, order=2, eps=0.01, dtype=np.float32):
    """Returns u flat with layout [(trace*np + slot)*n1 + i1]."""
    n1, n2, n3 = (int(v) for v in shape)
    np_ = (2 * ns2 + 1) * (2 * ns3 + 1)
    a = _arr(dn, dtype); p = _arr(dipi, dtype); q = _arr(dipx, dtype)
    u = np.zeros(n1 * n2 * n3 * np_, dtype)
    _lib(dtype).or_pwspray3d(_ptr(a), _ptr(p), _ptr(q), n1, n2, n3, ns2, ns3, order, _real(dtype)(eps), _ptr(u))
    return u


def somean3d(dn, dipi, dipx, shape, r1, r2, order=2, dtype=np.float32):
    n1, n2, n3 = (int(v) for v in shape)
    a = _arr(dn, dtype); p = _arr(dipi, dtype); q = _arr(dipx, dtype); o = np.zeros_like(a)
    _lib(dtype).or_somean3d(_ptr(a), _ptr(p), _ptr(q), n1, n2, n3, r1, r2, order, _ptr(o))
    return o


def somf3d(dn, dipi, dipx, shape, r1, r2, order=2, option=1, nmf=None, dtype=np.float32):
    n1, n2, n3 = (int(v) for v in shape)
    a = _arr(dn, dtype); p = _arr(dipi, dtype); q = _arr(dipx, dtype); o = np.zeros_like(a)
    nmf = 2 * r1 * r2 + 1 if nmf is None else nmf
    _lib(dtype).or_somf3d(_ptr(a), _ptr(p), _ptr(q), n1, n2, n3, r1, r2, nmf, option, order, _ptr(o))
    return o


def quantile(q, a, dtype=np.float32):
    a = np.ascontiguousarray(a, dtype).copy()
    f = _lib(dtype).or_quantile
    f.restype = _real(dtype)
    return f(ctypes.c_int(q), ctypes.c_int(a.size), _ptr(a))


def randn(seed, n):
    out = np.zeros(n, np.float32)
    _lib(np.float32).or_randn(ctypes.c_uint32(seed), ctypes.c_size_t(n), _ptr(out))
    return out


def soint3d(din, mask, dipi, dipx, shape, order=1, niter=100, njs=(1, 1), drift=0, seed=202223,
            hasmask=1, var=0.0, dtype=np.float32):
    n1, n2, n3 = (int(v) for v in shape)
    a = _arr(din, dtype); m = _arr(mask, dtype); p = _arr(dipi, dtype); q = _arr(dipx, dtype)
    o = np.zeros_like(a)
    _lib(dtype).or_soint3d(_ptr(a), _ptr(m), _ptr(p), _ptr(q), n1, n2, n3, order, njs[0], njs[1], niter,
                           drift, seed, hasmask, _real(dtype)(var), _ptr(o))
    return o


def soint2d(din, mask, dip, shape, order=1, niter=100, njs=(1, 1), drift=0, hasmask=1, twoplane=0,
            prec=0, dip2=None, dtype=np.float32):
    """csoint2d; prec=1 is the preconditioned solve (dip2 = the second plane's
    slopes when twoplane is set)."""
    n1, n2 = (int(v) for v in shape)
    a = _arr(din, dtype); m = _arr(mask, dtype); p = _arr(dip, dtype); o = np.zeros_like(a)
    if prec:
        q = _arr(dip2 if dip2 is not None else dip, dtype) if twoplane else None
        _lib(dtype).or_soint2d_prec(_ptr(a), _ptr(m), _ptr(p), _ptr(q) if q is not None else None, n1, n2, order,
                                    niter, hasmask, _ptr(o))
        return o
    _lib(dtype).or_soint2d(_ptr(a), _ptr(m), _ptr(p), n1, n2, order, njs[0], niter, drift, hasmask,
                           twoplane, _ptr(o))
    return o


def sint2d(din, mask, dip, shape, niter=100, eps=0.01, ns=1, order=1, dtype=np.float32):
    n1, n2 = (int(v) for v in shape)
    a = _arr(din, dtype); m = _arr(mask, dtype); p = _arr(dip, dtype); o = np.zeros_like(a)
    _lib(dtype).or_sint2d(_ptr(a), _ptr(p), _ptr(m), n1, n2, niter, ns, order, _real(dtype)(eps), _ptr(o))
    return o


def sint3d(din, mask, dipi, dipx, shape, niter=100, eps=0.01, ns1=1, ns2=1, order1=1, order2=1,
           dtype=np.float32):
    n1, n2, n3 = (int(v) for v in shape)
    a = _arr(din, dtype); m = _arr(mask, dtype); p = _arr(dipi, dtype); q = _arr(dipx, dtype)
    o = np.zeros_like(a)
    _lib(dtype).or_sint3d(_ptr(a), _ptr(p), _ptr(q), _ptr(m), n1, n2, n3, niter, ns1, ns2, order1, order2,
                          _real(dtype)(eps), _ptr(o))
    return o
