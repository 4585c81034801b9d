"""Stand-alone triangle / box smoothing: drop-in for pyseistr's smoothc
(pyseistr/smooth.py:115), computed by sh_smooth in libseistr_hip.so."""
import numpy as np

from . import _lib


def smooth_device(din, shape, rect=(1, 1, 1), diff=(0, 0, 0), box=(0, 0, 0), repeat=1, adj=1, out=None):
    """Device-level call: flat F-order float32 CUDA tensors in and out."""
    n1, n2, n3 = (int(s) for s in shape)
    _lib.check_sizes("smooth", n1 * n2 * n3, data=din)
    _lib.check_out(out, n1 * n2 * n3, "smooth")
    din = _lib.to_device(din)
    ctx = _lib.context()
    lib = _lib.load()
    if out is None:
        out = _lib.empty(n1 * n2 * n3)
    rc = lib.sh_smooth(ctx, _lib.ptr(din), _lib.ptr(out), n1, n2, n3, int(repeat), int(adj),
                       *(int(r) for r in rect), *(int(v) for v in diff), *(int(v) for v in box))
    _lib.check(ctx, rc, "smooth")
    return out


def smoothc(din, rect=[1, 1, 1], diff=[0, 0, 0], box=[0, 0, 0], repeat=1, adj=1):
    """Smooth din (2-D or 3-D) with a triangle (or box, per axis) of radius
    rect[i] along each axis i whose radius exceeds 1; diff[i] takes the
    derivative instead; repeat applies the filter several times; adj picks
    ps_smooth (1, the reference's default) or ps_smooth2 (0).  Returns
    float32 of din's shape (squeezed, as the reference)."""
    din = np.asarray(din)
    if din.ndim == 2:
        din = din[:, :, None]
    n1, n2, n3 = din.shape
    d = _lib.to_device(din)
    res = _lib.to_host(smooth_device(d, (n1, n2, n3), rect, diff, box, repeat, adj))
    return np.squeeze(res.reshape(n1, n2, n3, order="F"))
