"""Butterworth band-pass per trace: drop-in for pyseistr's bandpassc
(pyseistr/bp.py:128), computed by sh_bandpass in libseistr_hip.so."""
import numpy as np

from . import _lib


def bandpass_device(din, shape, dt, flo=0.0, fhi=0.5, nplo=6, nphi=6, phase=0, out=None):
    """Device-level call: flat F-order float32 CUDA tensors in and out."""
    n1, n2, n3 = (int(s) for s in shape)
    _lib.check_sizes("bandpass", n1 * n2 * n3, data=din)
    _lib.check_out(out, n1 * n2 * n3, "bandpass")
    din = _lib.to_device(din)
    ctx = _lib.context()
    lib = _lib.load()
    if out is None:
        out = _lib.empty(n1 * n2 * n3)
    rc = lib.sh_bandpass(ctx, _lib.ptr(din), _lib.ptr(out), n1, n2, n3, float(dt), float(flo), float(fhi),
                         int(nplo), int(nphi), int(phase))
    _lib.check(ctx, rc, "bandpass")
    return out


def bandpassc(din, dt, flo=0, fhi=0.5, nplo=6, nphi=6, phase=0, verb=1):
    """Band-pass every trace (first axis) of din between flo and fhi (in
    1/dt units): nplo / nphi poles, phase=0 zero phase, 1 minimum phase.
    Same shapes as the reference: 1-D and 2-D inputs come back in their
    own shape; 3-D ones keep three axes unless n3 == 1."""
    din = np.asarray(din)
    if din.ndim == 1:
        din = din[:, None]
    if din.ndim == 2:
        din = din[:, :, None]
    n1, n2, n3 = din.shape
    d = _lib.to_device(din)
    res = _lib.to_host(bandpass_device(d, (n1, n2, n3), dt, flo, fhi, nplo, nphi, phase))
    out = res.reshape(n1, n2, n3, order="F")
    return np.squeeze(out) if n3 == 1 else out
