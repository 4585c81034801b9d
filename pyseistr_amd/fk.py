"""FK dip filter on the GPU: drop-in for pyseistr's fkdip (pyseistr/fk.py:2),
computed by sh_fkdip in libseistr_hip.so (float64 FFTs and the cone mask)."""
import numpy as np

from . import _lib


def fkdip_device(din, shape, w, out=None):
    """Device-level call: flat F-order float64 CUDA tensors of shape
    (n1, n2) in and out."""
    torch = _lib._torch()
    n1, n2 = (int(s) for s in shape)
    if din.dtype != torch.float64 or din.numel() != n1 * n2 or not din.is_contiguous():
        raise ValueError(f"fkdip: expected a contiguous float64 array of {n1 * n2} samples")
    if out is None:
        out = torch.empty(n1 * n2, dtype=torch.float64, device="cuda")
    ctx = _lib.context()
    rc = _lib.load().sh_fkdip(ctx, _lib.ptr(din), _lib.ptr(out), n1, n2, float(w))
    _lib.check(ctx, rc, "fkdip")
    return out


def fkdip(d, w):
    """Keep the part of the 2-D section d (time x trace) inside a cone of
    half width w (fraction of the padded trace count) around zero
    wavenumber; returns float64 of d's shape.  pyseistr's DAS workflow
    subtracts it (d - fkdip(d, w)) to remove steep events."""
    d = np.asarray(d)
    if d.ndim != 2:
        raise ValueError("fkdip: d must be 2-D")
    if np.iscomplexobj(d):
        raise TypeError("fkdip: complex data are not supported (the device path takes real float64)")
    n1, n2 = d.shape
    torch = _lib._torch()
    dev = torch.from_numpy(np.ascontiguousarray(d.astype(np.float64, copy=False).ravel(order="F"))).cuda()
    return fkdip_device(dev, (n1, n2), w).cpu().numpy().reshape(n1, n2, order="F")
