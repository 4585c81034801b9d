"""Pins of the CPU oracle by what the mathematics fixes: closed forms,
adjoint (dot-product) identities, dense solves of the banded systems, exact
order statistics, brute-force convolutions and plane-wave invariants.
Each check is chosen so that a dropped term, a sign or index error or a
transposed operand in the oracle fails it."""
import numpy as np
import pytest

from oracle import oracle as O

F64 = np.float64


def b5(s):  # closed form of the order-2 all-pass coefficients (pyseistr/dip2d.py:295-307)
    return np.array([(1 - s) * (2 - s) * (3 - s) * (4 - s) / 1680,
                     (4 - s) * (2 - s) * (3 - s) * (4 + s) / 420,
                     (4 - s) * (3 - s) * (3 + s) * (4 + s) / 280,
                     (4 - s) * (2 + s) * (3 + s) * (4 + s) / 420,
                     (1 + s) * (2 + s) * (3 + s) * (4 + s) / 1680])


def b3(s):  # order-1 closed form (pyseistr/dip2d.py:271-281, formulas)
    return np.array([(1 - s) * (2 - s) / 12, (2 + s) * (2 - s) / 6, (1 + s) * (2 + s) / 12])


@pytest.mark.parametrize("p", [-2.3, -1.0, -0.37, 0.0, 0.5, 1.25, 2.9])
def test_passfilter_closed_form(p):
    np.testing.assert_allclose(O.passfilter(p, 2, F64), b5(p), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(O.passfilter(p, 1, F64), b3(p), rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("nw", [1, 2, 3])
@pytest.mark.parametrize("p", [-1.7, -0.2, 0.0, 0.8, 2.1])
def test_aderfilter_is_minus_derivative(nw, p):
    h = 1e-6
    fd = (O.passfilter(p + h, nw, F64) - O.passfilter(p - h, nw, F64)) / (2 * h)
    np.testing.assert_allclose(O.aderfilter(p, nw, F64), -fd, rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("nw", [1, 2, 3, 4])
def test_passfilter_properties(nw):
    a0 = O.passfilter(0.0, nw, F64)
    np.testing.assert_allclose(a0, a0[::-1], rtol=1e-14, atol=0)  # symmetric at zero slope
    for p in (-1.3, 0.4, 2.2):
        np.testing.assert_allclose(O.passfilter(p, nw, F64).sum(), 1.0, rtol=1e-9)  # passes DC


@pytest.mark.parametrize("nw", [1, 2, 3])
@pytest.mark.parametrize("shift", [-2, -1, 1, 2])
def test_pwd_annihilates_integer_slope_plane_wave(nw, shift):
    # x(t, x) = f(t - shift * x): the PWD residual with p = shift vanishes
    n1, n2 = 60, 6
    f = rand(n1 + 20, 40)
    x = np.stack([f[10 - shift * j:10 - shift * j + n1] for j in range(n2)], axis=1)[:, :, None]
    p = np.full(x.shape, float(shift))
    y = O.pwd(x, p, (n1, n2, 1), 2, nw=nw, dtype=F64).reshape(n1, n2, order="F")
    assert np.abs(y[2 * nw + 2:-2 * nw - 2, :-1]).max() < 1e-12


def rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


def test_pwd_matches_brute_force():
    n1, n2, n3 = 20, 6, 5
    x = rand((n1, n2, n3), 1); p = 0.8 * rand((n1, n2, n3), 2)
    for axis, der in ((2, 0), (3, 0), (2, 1), (3, 1)):
        y = O.pwd(x, p, (n1, n2, n3), axis, nw=2, der=der, dtype=F64).reshape(n1, n2, n3, order="F")
        ref = np.zeros_like(x)
        for i3 in range(n3 - (axis == 3)):
            for i2 in range(n2 - (axis == 2)):
                for i1 in range(2, n1 - 2):
                    a = -O.aderfilter(p[i1, i2, i3], 2, F64) if der else b5(p[i1, i2, i3])
                    nxt = x[:, i2 + 1, i3] if axis == 2 else x[:, i2, i3 + 1]
                    ref[i1, i2, i3] = sum((nxt[i1 + iw - 2] - x[i1 - iw + 2, i2, i3]) * a[iw] for iw in range(5))
        if der:
            ref = -ref
        np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_pwd_zero_slope_invariant():
    # p = 0 and x constant along the trace axis -> PWD residual vanishes
    n1, n2, n3 = 30, 7, 4
    tr = rand((n1, 1, 1), 3)
    x = np.broadcast_to(tr, (n1, n2, n3)).copy()
    y = O.pwd(x, np.zeros_like(x), (n1, n2, n3), 2, dtype=F64)
    assert np.abs(y).max() < 1e-13


def dot(a, b):
    return float(np.dot(np.ravel(a).astype(F64), np.ravel(b).astype(F64)))


def test_allpass3_adjoint():
    s = (16, 6, 5); n = int(np.prod(s))
    p, q = 0.7 * rand(n, 4), 0.7 * rand(n, 5)
    x, y = rand(n, 6), rand(2 * n, 7)
    Ax = O.allpass3(x, p, q, s, adj=0, dtype=F64)
    Aty = O.allpass3(y, p, q, s, adj=1, dtype=F64)
    assert abs(dot(Ax, y) - dot(x, Aty)) < 1e-10 * abs(dot(Ax, y))


def test_allpass21_adjoint():
    s = (20, 9); n = 180
    p = 0.7 * rand(n, 8); x, y = rand(n, 9), rand(n, 10)
    lhs = dot(O.allpass21(x, p, s, adj=0, dtype=F64), y)
    rhs = dot(x, O.allpass21(y, p, s, adj=1, dtype=F64))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def tri_brute(x, nb):
    """Triangle (2nb-1 taps, weights (nb-|k|)/nb^2) with even reflection."""
    n = len(x)
    out = np.zeros(n)
    for i in range(n):
        for k in range(-(nb - 1), nb):
            q = i + k
            if q < 0:
                q = -1 - q
            m = q % (2 * n)
            q = m if m < n else 2 * n - 1 - m
            out[i] += (nb - abs(k)) / nb ** 2 * x[q]
    return out


@pytest.mark.parametrize("n,nb", [(30, 5), (12, 10), (7, 20), (40, 1)])
def test_triangle_matches_brute_force(n, nb):
    x = rand(n, 11)
    y = O.trianglen(x, (n, 1, 1), (nb, 1, 1), dtype=F64)
    np.testing.assert_allclose(y, tri_brute(x, nb) if nb > 1 else x, rtol=1e-12, atol=1e-12)


def test_trianglen_separable_and_symmetric():
    s = (14, 9, 6); r = (4, 3, 5)
    x, z = rand(s, 12), rand(s, 13)
    Sx = O.trianglen(x, s, r, dtype=F64).reshape(s, order="F")
    ref = np.apply_along_axis(tri_brute, 0, x, 4)
    ref = np.apply_along_axis(tri_brute, 1, ref, 3)
    ref = np.apply_along_axis(tri_brute, 2, ref, 5)
    np.testing.assert_allclose(Sx, ref, rtol=1e-12, atol=1e-12)
    Sz = O.trianglen(z, s, r, dtype=F64)
    assert abs(dot(Sx.ravel(order="F"), z.ravel(order="F")) - dot(x.ravel(order="F"), Sz)) < 1e-10


def dense_predictor(n1, nw, e, filters, ins):
    """Dense form of the implicit plane-wave prediction (predict1_step /
    predict2_step): with C the PWD filter applied to the source trace,
    (C x)[k] = sum_j a_j(k) x[k+j-nw], and C' the same taps reversed onto the
    target trace, (C' y)[k] = sum_j a_j(k) y[k+nw-j], for k in [nw, n1-nw):
        (e (D^T D + B) + sum C'^T C') out = sum C'^T C in + e B in_mean,
    D the second difference, B the selector of rows 0, 1, n1-2, n1-1."""
    R = np.zeros((n1, n1))
    for i in range(n1 - 2):  # second difference rows, squared
        d = np.zeros(n1); d[i:i + 3] = [1, -2, 1]
        R += np.outer(d, d)
    # soft boundary rows: + I on rows 0, 1, n1-2, n1-1 (pyseistr/src/sof_cfuns.c:561-564)
    for k in (0, 1, n1 - 2, n1 - 1):
        R[k, k] += 1
    M = e * R
    rhs = np.zeros(n1)
    bnd = np.zeros(n1)
    for (p, rev), x in zip(filters, ins):
        C = np.zeros((n1, n1)); Cr = np.zeros((n1, n1))
        for k in range(nw, n1 - nw):
            a = O.passfilter(p[k], nw, F64)
            if rev:
                a = a[::-1]
            for j in range(2 * nw + 1):
                C[k, k + j - nw] = a[j]
                Cr[k, k + nw - j] = a[j]
        M = M + Cr.T @ Cr
        rhs += Cr.T @ C @ x
        bnd += x
    bnd /= max(len(ins), 1)
    for k in (0, 1, n1 - 2, n1 - 1):
        rhs[k] += e * bnd[k]
    return np.linalg.solve(M, rhs), R


def test_regulariser_rows_match_reference():
    # the dense regulariser above reproduces diag 6/-4/1 with the modified
    # end rows (pyseistr/src/sof_cfuns.c:548-565): check the matrix itself
    _, R = dense_predictor(10, 2, 1.0, [], [])
    assert R[0, 0] == 2 and R[1, 1] == 6 and R[2, 2] == 6 and R[0, 1] == -2 and R[1, 2] == -4 and R[0, 2] == 1


@pytest.mark.parametrize("forw", [0, 1])
def test_predict1_dense_solve(forw):
    n1 = 24
    x = rand(n1, 14); p = 0.9 * rand(n1, 15)
    out = O.predict1(x, p, forw=forw, nw=2, e=1e-2, dtype=F64)
    ref, _ = dense_predictor(n1, 2, 1e-2, [(p, forw)], [x])
    np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-10)


def test_predict2_dense_solve():
    n1 = 24
    x1, x2 = rand(n1, 16), rand(n1, 17)
    p1, p2 = 0.9 * rand(n1, 18), 0.9 * rand(n1, 19)
    out = O.predict2(x1, x2, p1, p2, forw1=1, forw2=0, nw=2, e=1e-2, dtype=F64)
    ref, _ = dense_predictor(n1, 2, 1e-2, [(p1, 1), (p2, 0)], [x1, x2])
    np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-10)


def test_predict_step_adjoint():
    n1 = 30
    p = 0.8 * rand(n1, 20)
    for forw in (0, 1):
        x, y = rand(n1, 21), rand(n1, 22)
        Px = O.predict_step(x, p, adj=0, forw=forw, e=1e-2, dtype=F64)
        Pty = O.predict_step(y, p, adj=1, forw=forw, e=1e-2, dtype=F64)
        assert abs(dot(Px, y) - dot(x, Pty)) < 1e-9 * abs(dot(Px, y))


def test_predict_plane_wave():
    # a constant-slope plane wave is predicted to its neighbour
    n1, s = 200, 0.6
    t = np.arange(n1)
    tr0 = np.exp(-((t - 90.0) / 6.0) ** 2)
    tr1 = np.exp(-((t - 90.0 - s) / 6.0) ** 2)
    out = O.predict1(tr0, np.full(n1, s), forw=1, nw=2, e=1e-4, dtype=F64)
    assert np.abs(out - tr1)[20:-20].max() < 2e-3


def test_spray2d_adjoint_and_centre():
    s = (24, 9); n = 216; ns = 3
    d = 0.6 * rand(n, 23); x = rand(n, 24); u = rand(n * 7, 25)
    Sx = O.pwspray2d(x, d, s, ns, adj=0, dtype=F64)
    Stu = O.pwspray2d(u, d, s, ns, adj=1, dtype=F64)
    assert abs(dot(Sx, u) - dot(x, Stu)) < 1e-9 * abs(dot(Sx, u))
    # centre slot is the input trace
    U = Sx.reshape(24, 7, 9, order="F")
    np.testing.assert_array_equal(U[:, ns, :], x.reshape(24, 9, order="F"))
    # slots beyond the edges stay zero
    assert np.all(U[:, ns + 1:, 0] == 0) and np.all(U[:, :ns, 8] == 0)


def test_pwsmooth2d_adjoint_and_unit_gain():
    s = (24, 10); n = 240
    d = 0.5 * rand(n, 26); x = rand(n, 27); y = rand(n, 28)
    Sx = O.pwsmooth2d(x, d, s, 3, adj=0, dtype=F64)
    Sty = O.pwsmooth2d(y, d, s, 3, adj=1, dtype=F64)
    assert abs(dot(Sx, y) - dot(x, Sty)) < 1e-9 * abs(dot(Sx, y))
    ones = np.ones(n)
    np.testing.assert_allclose(O.pwsmooth2d(ones, d, s, 3, dtype=F64), 1.0, rtol=1e-12)  # normalised


def test_quantile_is_exact_order_statistic():
    rng = np.random.default_rng(29)
    for n in (1, 2, 5, 13, 33, 81):
        a = rng.standard_normal(n).astype(np.float32)
        a[rng.random(n) < 0.2] = 0.0  # ties
        for q in {0, (n - 1) // 2, n - 1}:
            assert O.quantile(q, a) == np.partition(a, q)[q]


def test_spray3_plan_structure():
    for ns2, ns3 in ((4, 4), (2, 3), (1, 0)):
        plan = O.spray3_plan(ns2, ns3)
        np2 = 2 * ns2 + 1
        dist = {}
        for k, (slot, upd, up2, up3) in enumerate(plan):
            a, b = slot % np2 - ns2, slot // np2 - ns3
            dist[slot] = k
            assert upd == (1 if a else 0) | (2 if b else 0)
            if a:
                assert up2 == (a > 0)
                assert dist[slot - 1 if a > 0 else slot + 1] < k  # predecessor already visited
            if b:
                assert up3 == (b > 0)
                assert dist[slot - np2 if b > 0 else slot + np2] < k
        assert sorted(dist) == list(range(len(plan)))


def test_somean3d_constant_interior():
    s = (40, 12, 12)
    v = np.ones(s, np.float32)
    z = np.zeros(s, np.float32)
    out = O.somean3d(v, z, z, s, 2, 2, dtype=F64).reshape(s, order="F")
    # interior traces receive all 25 predictions of a flat event
    np.testing.assert_allclose(out[8:-8, 2:-2, 2:-2], 1.0, atol=1e-6)
    # corner trace: 9 of 25 slots are in range
    np.testing.assert_allclose(out[8:-8, 0, 0], 9 / 25, atol=1e-6)


def test_somf3d_is_median_of_spray_window():
    s = (30, 8, 7)
    rng = np.random.default_rng(30)
    v = rng.standard_normal(s).astype(np.float32)
    pi = (0.3 * rng.standard_normal(s)).astype(np.float32)
    px = (0.3 * rng.standard_normal(s)).astype(np.float32)
    u = O.pwspray3d(v, pi, px, s, 2, 2, eps=0.01).reshape(30, 25, 8 * 7, order="F")
    med = np.sort(u[:, 12 - 4:12 + 5, :], axis=1)[:, 4, :]
    out = O.somf3d(v, pi, px, s, 2, 2).reshape(30, 56, order="F")
    np.testing.assert_array_equal(out, med)


def test_dip_recovers_constant_slope():
    n1, n2, n3, sx, sy = 80, 24, 20, 0.7, -0.4
    t = np.arange(n1)[:, None, None]; x = np.arange(n2)[None, :, None]; y = np.arange(n3)[None, None, :]
    d = np.exp(-((t - 35 - sx * (x - 12) - sy * (y - 10)) / 3.0) ** 2) + 0.5 * np.sin((t - sx * x - sy * y) * 0.4)
    dip = O.dip(d, (n1, n2, n3), niter=5, liter=10, order=2, rect=(5, 5, 5), dtype=F64)
    di = dip[:d.size].reshape(d.shape, order="F"); dx = dip[d.size:].reshape(d.shape, order="F")
    assert abs(np.median(di[10:-10, 3:-3, 3:-3]) - sx) < 0.02
    assert abs(np.median(dx[10:-10, 3:-3, 3:-3]) - sy) < 0.02


def test_dip_ignores_eps_dv_like_reference():
    d = rand((30, 10, 1), 31)
    a = O.dip(d, (30, 10, 1), niter=2, liter=5, eps_dv=0.01, rect=(3, 3, 1))
    b = O.dip(d, (30, 10, 1), niter=2, liter=5, eps_dv=5.0, rect=(3, 3, 1))
    np.testing.assert_array_equal(a, b)


def test_soint_keeps_known_samples():
    s = (24, 8, 6)
    rng = np.random.default_rng(32)
    v = rng.standard_normal(s).astype(np.float32)
    m = np.broadcast_to((rng.random(s[1:]) < 0.5)[None], s).astype(np.float32)
    z = np.zeros(s, np.float32)
    out = O.soint3d(v * m, m, z, z, s, order=2, niter=10).reshape(s, order="F")
    np.testing.assert_array_equal(out[m > 0], (v * m)[m > 0])
    out2 = O.sint3d(v * m, m, z, z, s, niter=10, ns1=1, ns2=1).reshape(s, order="F")
    assert np.all(np.isfinite(out2))


# --- stand-alone smoothing (smoothcf, dip_cfuns.c:2006-2123) --------------
SMOOTH_CASES = [((40, 1, 1), (3, 1, 1)), ((37, 23, 1), (1, 4, 1)), ((20, 11, 9), (5, 5, 5)),
                ((5, 4, 3), (12, 2, 7)), ((9, 30, 7), (1, 1, 6))]


@pytest.mark.parametrize("shape,rect", SMOOTH_CASES)
@pytest.mark.parametrize("box", [(0, 0, 0), (1, 1, 1)])
def test_smoothc_preserves_constants(shape, rect, box):
    """Triangle and box weights sum to one and the even reflection keeps a
    constant constant, also when the radius exceeds the axis (multi-wrap)."""
    x = np.full(int(np.prod(shape)), 2.5)
    for adj in (0, 1):
        y = O.smoothc(x, shape, rect, box=box, adj=adj, dtype=np.float64)
        assert np.abs(y - 2.5).max() < 1e-12


@pytest.mark.parametrize("shape,rect", SMOOTH_CASES)
def test_smoothc_derivative_of_constant_is_zero(shape, rect):
    x = np.full(int(np.prod(shape)), -1.75)
    y = O.smoothc(x, shape, rect, diff=(1, 1, 1), adj=1, dtype=np.float64)
    assert np.abs(y).max() < 1e-12


@pytest.mark.parametrize("shape,rect", SMOOTH_CASES)
@pytest.mark.parametrize("diff,box", [((0, 0, 0), (0, 0, 0)), ((1, 0, 1), (0, 0, 0)), ((0, 0, 0), (1, 0, 1))])
def test_smoothc_forms_are_adjoint(shape, rect, diff, box):
    """ps_smooth (adj=1: fold, doubint, triple) and ps_smooth2 (adj=0:
    triple2, doubint2, fold2) are transposes: <S x, y> = <x, S2 y>, per axis
    and hence for one axis at a time (float64)."""
    rng = np.random.default_rng(sum(shape) + sum(rect))
    n = int(np.prod(shape))
    for ax in range(3):
        if rect[ax] <= 1:
            continue
        r1 = tuple(rect[a] if a == ax else 1 for a in range(3))
        x, y = rng.standard_normal(n), rng.standard_normal(n)
        sx = O.smoothc(x, shape, r1, diff, box, adj=1, dtype=np.float64)
        s2y = O.smoothc(y, shape, r1, diff, box, adj=0, dtype=np.float64)
        assert abs(np.dot(sx, y) - np.dot(x, s2y)) < 1e-11 * max(1.0, abs(np.dot(sx, y)))


def test_smoothc_matches_reference():
    """The FP32 restatement against the reference's own smoothcf
    (oracle/_ref, built from dip_cfuns.c where the sources are present),
    bit for bit, over shapes, radii (some beyond the axis), derivative, box,
    repeat and both forms."""
    from oracle import ref as R
    if not R.available():
        pytest.skip("reference modules not built")
    rng = np.random.default_rng(0)
    for shape, rect in SMOOTH_CASES:
        x = rng.standard_normal(int(np.prod(shape))).astype(np.float32)
        for diff in [(0, 0, 0), (1, 0, 1)]:
            for box in [(0, 0, 0), (1, 1, 0)]:
                for repeat in (1, 2):
                    for adj in (0, 1):
                        a = O.smoothc(x, shape, rect, diff, box, repeat, adj)
                        b = R.smoothcf(x, shape, rect, diff, box, repeat, adj)
                        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (shape, rect, diff, box,
                                                                                       repeat, adj)


# --- Butterworth band-pass (cbp, bp_cfuns.c:524-645) ----------------------
BP_CASES = [(0.0, 30.0, 6, 6), (5.0, 40.0, 3, 5), (2.5, 1000.0, 8, 2), (10.0, 60.0, 1, 1)]


def test_bandpass_matches_reference():
    """The FP32 restatement against the reference's own cbp (oracle/_ref),
    bit for bit: cutoffs (one above Nyquist), odd and even pole counts,
    zero and minimum phase."""
    from oracle import ref as R
    if not R.available():
        pytest.skip("reference modules not built")
    rng = np.random.default_rng(1)
    for shape in [(200, 3, 1), (64, 10, 2), (7, 5, 1)]:
        x = rng.standard_normal(int(np.prod(shape))).astype(np.float32)
        for flo, fhi, nplo, nphi in BP_CASES:
            for phase in (0, 1):
                a = O.bandpass(x, shape, 0.004, flo, fhi, nplo, nphi, phase)
                b = R.cbp(x, shape, 0.004, flo, fhi, nplo, nphi, phase)
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (shape, flo, fhi, nplo, nphi, phase)


@pytest.mark.parametrize("nplo", [1, 2, 5, 6])
def test_bandpass_frequency_response(nplo):
    """Butterworth magnitudes (float64): a low cut passes a sinusoid well
    above it and removes a constant; a high cut passes a constant and
    removes a sinusoid well above it; zero phase (phase=0) keeps a centred
    pulse symmetric."""
    n, dt = 4000, 0.004
    t = np.arange(n) * dt
    hi_sig = np.sin(2 * np.pi * 60.0 * t)
    k = slice(1000, 3000)  # away from the start-up transients of both passes
    y = O.bandpass(hi_sig, (n, 1, 1), dt, flo=5.0, fhi=0.5 / dt, nplo=nplo, nphi=6, dtype=np.float64)
    assert abs(np.std(y[k]) / np.std(hi_sig[k]) - 1.0) < 0.02
    y = O.bandpass(np.ones(n), (n, 1, 1), dt, flo=5.0, fhi=0.5 / dt, nplo=nplo, nphi=6, dtype=np.float64)
    assert np.abs(y[k]).max() < 1e-6
    y = O.bandpass(np.ones(n), (n, 1, 1), dt, flo=0.0, fhi=10.0, nplo=6, nphi=nplo, dtype=np.float64)
    assert np.abs(y[k] - 1.0).max() < 1e-6
    y = O.bandpass(hi_sig, (n, 1, 1), dt, flo=0.0, fhi=10.0, nplo=6, nphi=nplo, dtype=np.float64)
    # 60 Hz is six times the cut: even one pole, applied forward and
    # backward, leaves |H|^2 = 1 / (1 + 36) (bilinear warping makes it less)
    assert np.std(y[k]) < 0.05 * np.std(hi_sig[k])
    pulse = np.zeros(n)
    pulse[n // 2] = 1.0
    y = O.bandpass(pulse, (n, 1, 1), dt, flo=5.0, fhi=40.0, nplo=6, nphi=6, phase=0, dtype=np.float64)
    m = 300
    assert np.abs(y[n // 2 - m:n // 2] - y[n // 2 + m:n // 2:-1]).max() < 1e-3 * np.abs(y).max()


SOINT_PREC_CASES = [((40, 24), 1, 10, 1, 0), ((60, 30), 2, 25, 1, 0), ((33, 17), 2, 8, 0, 0),
                    ((40, 24), 2, 12, 1, 1), ((50, 20), 1, 3, 1, 1)]


def _soint_prec_case(shape, hasmask, seed):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal(shape).astype(np.float32)
    mask = (rng.random(shape) > 0.5).astype(np.float32)
    if hasmask == 0:
        d = d * mask
    return d, mask, (0.7 * rng.standard_normal(shape)).astype(np.float32), \
        (0.5 * rng.standard_normal(shape)).astype(np.float32)


def test_soint2d_prec_matches_reference():
    """The FP32 restatement of csoint2d with prec (predict_lop or
    predict2_lop preconditioner, ps_solver_prec) against the reference's
    own csoint2d (oracle/_ref), bit for bit."""
    from oracle import ref as R
    if not R.available():
        pytest.skip("reference modules not built")
    for k, (shape, order, niter, hasmask, twoplane) in enumerate(SOINT_PREC_CASES):
        d, mask, p, q = _soint_prec_case(shape, hasmask, k)
        f = lambda a: a.ravel(order="F")  # noqa: E731
        r = np.asarray(R.soint2d(f(d), f(mask), f(p), shape, order=order, niter=niter, hasmask=hasmask,
                                 twoplane=twoplane, prec=1, dip2=f(q) if twoplane else None), np.float32).ravel()
        o = O.soint2d(f(d), f(mask), f(p), shape, order=order, niter=niter, hasmask=hasmask, twoplane=twoplane,
                      prec=1, dip2=f(q))
        assert np.array_equal(r.view(np.uint32), o.view(np.uint32)), (shape, order, niter, hasmask, twoplane)


@pytest.mark.parametrize("twoplane", [0, 1])
def test_soint2d_prec_residual_decreases(twoplane):
    """In float64 the preconditioned CG fits the known samples ever better:
    |M x_k - d| does not grow with the iteration count k (a wrong adjoint or
    a transposed preconditioner breaks this), and with every sample known
    the one-plane solve approaches the data."""
    shape = (36, 14)
    d, mask, _, _ = _soint_prec_case(shape, 1, 7)
    # gently varying slopes keep the preconditioned system well conditioned
    p = np.repeat((0.5 * np.sin(np.arange(36) / 6.0))[:, None], 14, axis=1).astype(np.float32)
    q = np.full(shape, 0.3, np.float32)
    f = lambda a: a.ravel(order="F")  # noqa: E731
    m = f(mask) != 0
    res = []
    for k in range(1, 41):
        x = O.soint2d(f(d), f(mask), f(p), shape, order=2, niter=k, hasmask=1, twoplane=twoplane, prec=1, dip2=f(q),
                      dtype=np.float64)
        res.append(np.linalg.norm((x - f(d).astype(np.float64))[m]))
    assert all(b <= a * (1 + 1e-9) for a, b in zip(res, res[1:])), res
    # a product of two predictions (twoplane) converges more slowly
    assert res[-1] < (0.6 if twoplane else 0.1) * res[0]
    if not twoplane:
        full = np.ones(shape, np.float32)
        x = O.soint2d(f(d), f(full), f(p), shape, order=2, niter=60, hasmask=1, prec=1, dtype=np.float64)
        assert np.linalg.norm(x - f(d)) < 0.05 * np.linalg.norm(f(d))


def test_soint_drift_matches_reference():
    """soint3d / soint2d with drift=1 (each PWD stencil shifted by NINT(p),
    the filter taken at p - NINT(p)): the FP32 restatement against the
    reference's own csoint3d / csoint2d, bit for bit, with slopes large
    enough that shifts and skipped samples occur."""
    from oracle import ref as R
    if not R.available():
        pytest.skip("reference modules not built")
    rng = np.random.default_rng(0)
    f = lambda a: a.ravel(order="F")  # noqa: E731
    for shape, nj, order, niter in [((40, 12, 9), (1, 1), 2, 8), ((33, 10, 7), (2, 1), 1, 6)]:
        d = rng.standard_normal(shape).astype(np.float32)
        mask = (rng.random(shape) > 0.4).astype(np.float32)
        pi = (2.5 * rng.standard_normal(shape)).astype(np.float32)
        px = (2.0 * rng.standard_normal(shape)).astype(np.float32)
        r = np.asarray(R.soint3d(f(d), f(mask), f(pi), f(px), shape, order=order, niter=niter, njs=nj, drift=1),
                       np.float32).ravel()
        o = O.soint3d(f(d), f(mask), f(pi), f(px), shape, order=order, niter=niter, njs=nj, drift=1)
        assert np.array_equal(r.view(np.uint32), o.view(np.uint32)), shape
    for (n1, n2), nj, order, niter in [((40, 20), 1, 2, 10), ((37, 15), 2, 1, 6)]:
        d = rng.standard_normal((n1, n2)).astype(np.float32)
        mask = (rng.random((n1, n2)) > 0.4).astype(np.float32)
        p = (2.5 * rng.standard_normal((n1, n2))).astype(np.float32)
        r = np.asarray(R.soint2d(f(d), f(mask), f(p), (n1, n2), order=order, niter=niter, njs=(nj, 1), drift=1),
                       np.float32).ravel()
        o = O.soint2d(f(d), f(mask), f(p), (n1, n2), order=order, niter=niter, njs=(nj, 1), drift=1)
        assert np.array_equal(r.view(np.uint32), o.view(np.uint32)), (n1, n2)
