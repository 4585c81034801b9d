"""The C ABI's context flags, called through ctypes:
 * SH_FLAG_HOST_PTRS: host arrays in and out give the same bits as device
   arrays (inputs staged to device workspace, outputs copied back);
 * SH_FLAG_CHECK_FINITE: a NaN or Inf in an input fails the call with
   SH_ENONFINITE naming that input, before anything else runs; finite
   inputs run unchanged;
 * sh_dip_report.elapsed_ms is the call's device time;
 * sh_c_to_f, the device-side Fortran-order flatten of C-order arrays."""
import ctypes

import numpy as np
import pytest

from pyseistr_amd import _lib

pytestmark = pytest.mark.gpu

ENONFINITE = -6


@pytest.fixture()
def fresh():
    """A context of its own (flags do not leak into the drop-ins' one)."""
    import torch
    torch.cuda.init()
    lib = _lib.load()
    ctx = lib.sh_ctx_create(None)
    yield lib, ctx
    torch.cuda.synchronize()
    lib.sh_ctx_destroy(ctx)


def F(a):
    return np.ascontiguousarray(np.asarray(a, np.float32).ravel(order="F"))


def hp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _vol(shape, seed):
    from pyseistr_amd import synth
    v, _ = synth.volume3d(shape, seed=seed, n_events=3, smax=2.0)
    rng = np.random.default_rng(seed)
    p = (0.7 * rng.standard_normal(shape)).astype(np.float32)
    q = (0.7 * rng.standard_normal(shape)).astype(np.float32)
    return F(v), F(p), F(q)


def test_host_pointers_equal_device_pointers(fresh):
    import torch
    lib, ctx = fresh
    s = (48, 14, 12)
    n = int(np.prod(s))
    v, p, q = _vol(s, 3)
    dv = [torch.from_numpy(a).cuda() for a in (v, p, q)]
    # somf3d
    d_out = torch.empty(n, device="cuda")
    assert lib.sh_somf3d(ctx, *[_lib.ptr(t) for t in dv], *s, 2, 2, 9, 1, 2, ctypes.c_float(0.01), _lib.ptr(d_out)) == 0
    # dip with a report
    d_dip = torch.empty(2 * n, device="cuda")
    rep = _lib.DipReport()
    assert lib.sh_dip(ctx, _lib.ptr(dv[0]), None, None, *s, 3, 10, 2, ctypes.c_float(0.01), 3, 3, 3, _lib.ptr(d_dip),
                      ctypes.byref(rep)) == 0
    assert rep.elapsed_ms > 0.0
    d_rec = [rep.cg_iters[0][k] for k in range(3)]
    # soint3d
    mask = F((np.random.default_rng(1).random(s) > 0.5).astype(np.float32))
    dm = torch.from_numpy(mask).cuda()
    d_si = torch.empty(n, device="cuda")
    it_d = ctypes.c_int(0)
    assert lib.sh_soint3d(ctx, _lib.ptr(dv[0]), _lib.ptr(dm), _lib.ptr(dv[1]), _lib.ptr(dv[2]), *s, 2, 1, 1, 6, 0,
                          202223, 1, ctypes.c_float(0.0), _lib.ptr(d_si), ctypes.byref(it_d)) == 0
    # 2-D spray adjoint
    s2 = (48, 14)
    n2 = s2[0] * s2[1]
    u = F(np.random.default_rng(2).standard_normal(n2 * 7))
    du = torch.from_numpy(u.copy()).cuda()
    d_x = torch.empty(n2, device="cuda")
    assert lib.sh_pwspray2d(ctx, 1, _lib.ptr(dv[1]), *s2, 3, 2, ctypes.c_float(0.01), _lib.ptr(d_x), _lib.ptr(du)) == 0
    torch.cuda.synchronize()

    assert lib.sh_ctx_set_flags(ctx, _lib.FLAG_HOST_PTRS) == 0
    h_out = np.full(n, np.nan, np.float32)
    assert lib.sh_somf3d(ctx, hp(v), hp(p), hp(q), *s, 2, 2, 9, 1, 2, ctypes.c_float(0.01), hp(h_out)) == 0
    assert np.array_equal(h_out, d_out.cpu().numpy())
    h_dip = np.full(2 * n, np.nan, np.float32)
    rep2 = _lib.DipReport()
    assert lib.sh_dip(ctx, hp(v), None, None, *s, 3, 10, 2, ctypes.c_float(0.01), 3, 3, 3, hp(h_dip),
                      ctypes.byref(rep2)) == 0
    assert np.array_equal(h_dip, d_dip.cpu().numpy())
    assert [rep2.cg_iters[0][k] for k in range(3)] == d_rec
    h_si = np.full(n, np.nan, np.float32)
    it_h = ctypes.c_int(0)
    assert lib.sh_soint3d(ctx, hp(v), hp(mask), hp(p), hp(q), *s, 2, 1, 1, 6, 0, 202223, 1, ctypes.c_float(0.0),
                          hp(h_si), ctypes.byref(it_h)) == 0
    assert np.array_equal(h_si, d_si.cpu().numpy()) and it_h.value == it_d.value
    h_x = np.full(n2, np.nan, np.float32)
    p2 = np.ascontiguousarray(p[:n2])
    assert lib.sh_pwspray2d(ctx, 1, hp(p2), *s2, 3, 2, ctypes.c_float(0.01), hp(h_x), hp(u)) == 0
    assert np.array_equal(h_x, d_x.cpu().numpy())
    # the allocation-free entry point takes device pointers only
    assert lib.sh_somf2d_noalloc(ctx, hp(v), hp(p), 48, 14, 1, 2, 5, 1, 2, ctypes.c_float(0.01), hp(h_out)) == -1


def test_check_finite(fresh):
    import torch
    lib, ctx = fresh
    s = (40, 10, 8)
    n = int(np.prod(s))
    v, p, q = _vol(s, 4)
    assert lib.sh_ctx_set_flags(ctx, _lib.FLAG_CHECK_FINITE) == 0
    out = torch.empty(2 * n, device="cuda")
    good = [torch.from_numpy(a).cuda() for a in (v, p, q)]
    assert lib.sh_dip(ctx, _lib.ptr(good[0]), None, None, *s, 2, 10, 2, ctypes.c_float(0.01), 3, 3, 3, _lib.ptr(out),
                      None) == 0
    bad = v.copy()
    bad[123] = np.nan
    bad[7] = np.inf
    tb = torch.from_numpy(bad).cuda()
    assert lib.sh_dip(ctx, _lib.ptr(tb), None, None, *s, 2, 10, 2, ctypes.c_float(0.01), 3, 3, 3, _lib.ptr(out),
                      None) == ENONFINITE
    assert b"data holds 2 NaN or Inf" in lib.sh_last_error(ctx)
    qb = q.copy()
    qb[-1] = -np.inf
    tq = torch.from_numpy(qb).cuda()
    assert lib.sh_somf3d(ctx, _lib.ptr(good[0]), _lib.ptr(good[1]), _lib.ptr(tq), *s, 2, 2, 9, 1, 2,
                         ctypes.c_float(0.01), _lib.ptr(out)) == ENONFINITE
    assert b"dipx" in lib.sh_last_error(ctx)
    nbad = ctypes.c_long(-1)
    assert lib.sh_count_nonfinite(ctx, _lib.ptr(tb), n, ctypes.byref(nbad)) == 0 and nbad.value == 2
    # host arrays are scanned too
    assert lib.sh_ctx_set_flags(ctx, _lib.FLAG_CHECK_FINITE | _lib.FLAG_HOST_PTRS) == 0
    h = np.empty(n, np.float32)
    assert lib.sh_somean3d(ctx, hp(v), hp(p), hp(qb), *s, 2, 2, ctypes.c_float(0.01), 2, hp(h)) == ENONFINITE


def test_check_finite_drop_in():
    import pyseistr_amd as ps
    v = np.zeros((32, 8), np.float32)
    v[3, 3] = np.nan
    ps.check_finite(True)
    try:
        with pytest.raises(ValueError, match="NaN or Inf"):
            ps.dip2dc(v)
    finally:
        ps.check_finite(False)


@pytest.mark.parametrize("shape", [(67, 9, 131), (41, 130, 33), (300, 257), (64, 1, 1024), (1025, 65, 3)])
def test_c_order_upload_is_fortran_flatten(shape):
    """sh_c_to_f (used by the drop-ins for C-order arrays of 2^16 samples or
    more) gives exactly numpy's Fortran-order flatten, which is what the
    reference's wrappers pass to C; a drop-in run on the C-order array and
    on its Fortran-order copy gives the same bits."""
    a = np.random.default_rng(len(shape) + shape[0]).standard_normal(shape).astype(np.float32)
    assert a.flags.c_contiguous and a.size >= _lib.C_TO_F_MIN
    got = _lib.to_device(a).cpu().numpy()
    np.testing.assert_array_equal(got, a.ravel(order="F"))
    if len(shape) == 3 and shape == (41, 130, 33):
        import pyseistr_amd as ps
        f = np.asfortranarray(a)
        x = ps.somean3dc(a, 0.3 * a, -0.2 * a, 2, 2, 0.01, 2)
        y = ps.somean3dc(f, np.asfortranarray(0.3 * a), np.asfortranarray(-0.2 * a), 2, 2, 0.01, 2)
        np.testing.assert_array_equal(x, y)


def test_c_to_f_rejects_bad_shapes(fresh):
    lib, ctx = fresh
    import torch
    t = torch.zeros(8, device="cuda")
    assert lib.sh_c_to_f(ctx, _lib.ptr(t), _lib.ptr(t), 0, 1, 8) != 0
    assert lib.sh_c_to_f(ctx, _lib.ptr(t), _lib.ptr(t), 1, 70000, 1) != 0
    assert lib.sh_c_to_f(ctx, None, _lib.ptr(t), 1, 1, 8) != 0
