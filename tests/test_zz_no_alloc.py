"""sh_somf2d_noalloc (fixed device workspace, no allocation) against
sh_somf2d through the C ABI: bit-identical results on two calls in a row,
the second one smaller, for the median and the space-varying median, and
the size limit refused before any launch."""
import ctypes

import numpy as np
import pytest

from pyseistr_amd import _lib, synth

pytestmark = pytest.mark.gpu


def _gather(n1, n2, seed):
    d = synth.gather2d(n1, n2, seed=seed)
    rng = np.random.default_rng(seed + 100)
    t = np.linspace(0, 3, n1)[:, None]
    x = np.linspace(0, 2, n2)[None, :]
    dip = 1.2 * np.sin(t + 2 * x + rng.uniform(0, 6)) + 0.05 * rng.standard_normal((n1, n2))
    return np.asarray(d, np.float32), dip.astype(np.float32)


def _bits(a):
    a = np.where(a == 0, 0, a).astype(np.float32)  # the sign of a zero median is not fixed
    return a.view(np.uint32)


def _run(fn, ctx, d, p, shape, ns, option, order=2, eps=0.01):
    n1, n2, n3 = shape
    out = _lib.empty(n1 * n2 * n3)
    out.fill_(np.nan)
    rc = fn(ctx, _lib.ptr(d), _lib.ptr(p), n1, n2, n3, ns, 2 * ns + 1, option, order, ctypes.c_float(eps),
            _lib.ptr(out))
    _lib.check(ctx, rc, fn.__name__)
    return out


@pytest.mark.parametrize("option", [1, 2])
def test_noalloc_equals_somf2d_twice(option):
    import torch
    lib = _lib.load()
    ctx = _lib.context()
    cases = [((512, 256, 1), 8, 11), ((200, 96, 2), 5, 12)]  # the second call is smaller
    for shape, ns, seed in cases:
        n1, n2, n3 = shape
        print(f"somf2d_noalloc option={option} shape={shape} ns={ns}", flush=True)
        assert lib.sh_somf2d_noalloc_bytes(n1, n2, n3, ns, 2) <= (64 << 20)
        d, p = zip(*[_gather(n1, n2, seed + i) for i in range(n3)])
        d = np.stack(d, axis=2)
        p = np.stack(p, axis=2)
        dd, pp = _lib.to_device(d), _lib.to_device(p)
        a = _run(lib.sh_somf2d_noalloc, ctx, dd, pp, shape, ns, option)
        b = _run(lib.sh_somf2d, ctx, dd, pp, shape, ns, option)
        torch.cuda.synchronize()
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert np.isfinite(a).all()
        assert np.array_equal(_bits(a), _bits(b)), f"max diff {np.abs(a - b).max()}"


def test_noalloc_refuses_oversize():
    lib = _lib.load()
    ctx = _lib.context()
    n1, n2, n3, ns = 2048, 1024, 1, 8
    assert lib.sh_somf2d_noalloc_bytes(n1, n2, n3, ns, 2) > (64 << 20)
    d = _lib.empty(16)
    out = _lib.empty(16)
    rc = lib.sh_somf2d_noalloc(ctx, _lib.ptr(d), _lib.ptr(d), n1, n2, n3, ns, 2 * ns + 1, 1, 2,
                               ctypes.c_float(0.01), _lib.ptr(out))
    assert rc == -1
    assert b"above the fixed" in lib.sh_last_error(ctx)
