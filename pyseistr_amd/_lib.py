"""ctypes binding of libseistr_hip.so (the gfx950 C ABI in include/seistr_hip.h).

The library is the only compute path: if it is missing or no GPU is visible
the calls raise; there is no CPU fallback.  PyTorch is used for device memory
and the current HIP stream only.
"""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libseistr_hip.so")

_lib = None
_tls = threading.local()

c_int, c_float, c_void_p = ctypes.c_int, ctypes.c_float, ctypes.c_void_p
c_int_p = ctypes.POINTER(ctypes.c_int)


class SeistrError(RuntimeError):
    """Raised for a non-zero status of the HIP library."""


class DipReport(ctypes.Structure):
    _fields_ = [("ndir", c_int), ("niter", c_int),
                ("cg_iters", (c_int * 64) * 2), ("ls_step", (c_int * 64) * 2), ("elapsed_ms", c_float)]


FLAG_CHECK_FINITE = 1
FLAG_HOST_PTRS = 2
# C-order arrays of at least this many samples are reordered on the device
C_TO_F_MIN = 1 << 16


_SIGS = {
    "sh_ctx_create": (c_void_p, [c_void_p]),
    "sh_ctx_destroy": (None, [c_void_p]),
    "sh_ctx_set_stream": (c_int, [c_void_p, c_void_p]),
    "sh_ctx_trim": (c_int, [c_void_p]),
    "sh_ctx_set_flags": (c_int, [c_void_p, c_int]),
    "sh_ctx_get_flags": (c_int, [c_void_p]),
    "sh_count_nonfinite": (c_int, [c_void_p, c_void_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]),
    "sh_last_error": (ctypes.c_char_p, [c_void_p]),
    "sh_version": (c_int, []),
    "sh_sync": (c_int, [c_void_p]),
    "sh_prof_enable": (c_int, [c_void_p, c_int]),
    "sh_c_to_f": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int]),
    "sh_prof_report": (c_int, [c_void_p, ctypes.c_char_p, c_int]),
    "sh_dip": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int,
                       c_float, c_int, c_int, c_int, c_void_p, ctypes.POINTER(DipReport)]),
    "sh_smooth": (c_int, [c_void_p, c_void_p, c_void_p] + [c_int] * 14),
    "sh_bandpass": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_float, c_float, c_float, c_int,
                            c_int, c_int]),
    "sh_patch_count": (c_int, [c_int] * 9 + [ctypes.POINTER(ctypes.c_long)]),
    "sh_patch": (c_int, [c_void_p, c_void_p, c_void_p] + [c_int] * 9),
    "sh_patch_inv": (c_int, [c_void_p, c_void_p, c_void_p] + [c_int] * 9),
    "sh_fkdip": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, ctypes.c_double]),
    "sh_pwspray3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int,
                             c_int, c_float, c_void_p]),
    "sh_somean3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int,
                            c_float, c_int, c_void_p]),
    "sh_somf3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int,
                          c_int, c_int, c_int, c_float, c_void_p]),
    "sh_pwspray2d": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_int, c_float, c_void_p,
                             c_void_p]),
    "sh_somean2d": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int,
                            c_float, c_void_p]),
    "sh_somf2d": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                          c_float, c_void_p]),
    "sh_somf2d_noalloc_bytes": (ctypes.c_size_t, [c_int, c_int, c_int, c_int, c_int]),
    "sh_somf2d_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                  c_float, c_void_p]),
    "sh_soint3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int,
                           c_int, c_int, c_int, c_int, c_int, c_int, c_float, c_void_p, c_int_p]),
    "sh_soint2d_prec": (c_int, [c_void_p] * 5 + [c_int] * 5 + [c_void_p, c_int_p]),
    "sh_soint2d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int,
                           c_int, c_int, c_int, c_void_p, c_int_p]),
    "sh_sint3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int,
                          c_int, c_int, c_int, c_int, c_float, c_void_p, c_int_p]),
    "sh_sint2d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int,
                          c_float, c_void_p, c_int_p]),
    "sh_batch_tile_count": (c_int, []),
    "sh_batch_tile_shape": (c_int, [c_int, c_int_p, c_int_p]),
    "sh_dip2d_batched": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int,
                                 c_int, c_int, c_int, c_void_p, c_void_p, c_int_p]),
    "sh_somf2d_batched": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                  c_int, c_int, c_float, c_void_p]),
    "sh_cnn_nparams": (ctypes.c_long, []),
    "sh_cnn_forward": (c_int, [c_void_p, c_void_p, c_void_p, c_float, c_int_p, c_int, c_int, c_int, c_void_p]),
    "sh_cnn_grad": (c_int, [c_void_p, c_void_p, c_void_p, c_float, c_int_p, c_void_p, c_void_p, c_int, c_int,
                            c_int, c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "sh_cnn_adam": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_float,
                            c_float, c_int, c_float]),
}

EXPORTED = tuple(_SIGS)


def load():
    """Load and declare the library (no GPU needed for loading)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SeistrError(f"{LIB_PATH} is not built; run pyseistr_amd/csrc/build.sh "
                              "(hipcc --offload-arch=gfx950)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            if not hasattr(lib, name):
                continue  # reported by missing_symbols()
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def missing_symbols():
    lib = load()
    return [name for name in _SIGS if not hasattr(lib, name)]


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise SeistrError("pyseistr_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def context():
    """Per-thread library context bound to torch's current stream."""
    torch = _torch()
    lib = load()
    dev = torch.cuda.current_device()
    ctxs = getattr(_tls, "ctxs", None)
    if ctxs is None:
        ctxs = _tls.ctxs = {}
    ctx = ctxs.get(dev)
    if ctx is None:
        ctx = lib.sh_ctx_create(None)
        if not ctx:
            raise SeistrError("sh_ctx_create failed")
        ctxs[dev] = ctx
    stream = torch.cuda.current_stream().cuda_stream
    check(ctx, lib.sh_ctx_set_stream(ctx, c_void_p(stream)), "context")
    return ctx


def check(ctx, rc, what):
    if rc != 0:
        msg = load().sh_last_error(ctx)
        msg = msg.decode() if msg else ""
        if rc in (-1, -6):  # bad argument, or non-finite input (SH_FLAG_CHECK_FINITE)
            raise ValueError(f"{what}: {msg}")
        if rc == -5:
            raise NotImplementedError(f"{what}: {msg}")
        raise SeistrError(f"{what}: {msg} (status {rc})")


def check_finite(on=True):
    """Make the path's calls on this thread's context scan their inputs for
    NaN / Inf first and raise ValueError naming the input (off by default,
    as in the reference, which propagates them)."""
    ctx = context()
    lib = load()
    flags = lib.sh_ctx_get_flags(ctx)
    flags = (flags | FLAG_CHECK_FINITE) if on else (flags & ~FLAG_CHECK_FINITE)
    check(ctx, lib.sh_ctx_set_flags(ctx, flags), "check_finite")


def profile_enable(on=True):
    ctx = context()
    check(ctx, load().sh_prof_enable(ctx, 1 if on else 0), "profile")


def profile_report():
    """{kernel: (launches, total_ms, total_algorithmic_bytes)} since enable."""
    import json
    ctx = context()
    buf = ctypes.create_string_buffer(1 << 16)
    check(ctx, load().sh_prof_report(ctx, buf, len(buf)), "profile")
    return {k: tuple(v) for k, v in json.loads(buf.value.decode()).items()}


def ptr(t):
    return c_void_p(t.data_ptr()) if t is not None else None


def nsamples(a):
    """Number of samples of a numpy array, a sequence or a tensor."""
    return int(a.numel()) if hasattr(a, "numel") else int(np.size(a))


def check_sizes(what, n, **arrays):
    """ValueError unless every given array (numpy or tensor; None skipped)
    holds exactly n samples.  The kernels index each input and output by the
    dimensions alone, so a short buffer would be read or written past its
    end; this runs before the library is called."""
    for name, a in arrays.items():
        if a is None:
            continue
        k = nsamples(a)
        if k != n:
            raise ValueError(f"{what}: {name} has {k} samples, {n} expected")


def check_out(out, n, what):
    """A caller-supplied output must be a flat contiguous float32 device
    tensor of exactly n samples."""
    if out is None:
        return None
    check_sizes(what, n, out=out)
    torch = _torch()
    if not (isinstance(out, torch.Tensor) and out.is_cuda and out.dtype == torch.float32 and out.is_contiguous()):
        raise TypeError(f"{what}: out must be a contiguous float32 device tensor")
    return out


def to_device(a):
    """numpy (any order) -> flat Fortran-order float32 device tensor."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        if a.is_cuda and a.dtype == torch.float32 and a.dim() == 1 and a.is_contiguous():
            return a
        raise TypeError("device tensors must be flat contiguous float32 (Fortran-order data)")
    a = np.asarray(a, dtype=np.float32)
    if a.ndim in (2, 3) and a.flags.c_contiguous and not a.flags.f_contiguous and a.size >= C_TO_F_MIN:
        # a C-order volume goes up as it is laid out and is reordered on the
        # device (sh_c_to_f): numpy's Fortran-order copy of a 512x256x256
        # volume takes a third of a second on one host core
        s = a.shape if a.ndim == 3 else (a.shape[0], 1, a.shape[1])
        if s[1] <= 65535 and s[0] <= 4194240:
            src = torch.from_numpy(a.reshape(-1)).cuda()
            out = torch.empty(a.size, dtype=torch.float32, device="cuda")
            ctx = context()
            check(ctx, load().sh_c_to_f(ctx, ptr(src), ptr(out), s[0], s[1], s[2]), "c_to_f")
            return out
    flat = np.ascontiguousarray(a.ravel(order="F"))
    return torch.from_numpy(flat).cuda()


def empty(n):
    torch = _torch()
    return torch.empty(int(n), dtype=torch.float32, device="cuda")


def to_host(t):
    return t.cpu().numpy()
