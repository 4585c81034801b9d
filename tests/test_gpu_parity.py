"""GPU parity against the reference C extension's golden vectors and the CPU
oracle.  Bars:
 * spray, somean, somf, soint, sint: bit-exact (sign of zero canonicalised);
 * dips: |gpu - reference| <= 2e-3 samples/trace max and 2e-4 RMS, the same
   CG / line-search decisions, and an error against the FP64 oracle no
   larger than the reference's own FP32 error (+1e-5)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DIP_MAX, DIP_RMS = 2e-3, 2e-4


def load(name):
    return np.load(os.path.join(G, name + ".npz"))


def canon(a):
    a = np.asarray(a, np.float32).ravel(order="F")
    return np.where(a == 0, np.float32(0), a).view(np.uint32)


def same(a, b):
    np.testing.assert_array_equal(canon(a), canon(b))


def dip_close(gpu, ref, ref64=None):
    gpu = np.asarray(gpu, np.float64).ravel(order="F"); ref = np.asarray(ref, np.float64).ravel()
    d = np.abs(gpu - ref)
    assert d.max() <= DIP_MAX, d.max()
    assert np.sqrt(np.mean(d ** 2)) <= DIP_RMS
    if ref64 is not None:
        assert np.abs(gpu - ref64).max() <= np.abs(ref - ref64).max() + 1e-5


@pytest.fixture(scope="module")
def ps():
    import pyseistr_amd
    return pyseistr_amd


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def test_dip3d_golden(ps):
    from oracle import oracle as O
    z = load("dip3d"); s = tuple(int(v) for v in z["shape"])
    (di, dx), rep = ps.dip3dc(z["data"], niter=5, liter=10, rect=[5, 5, 5], report=True)
    g = np.concatenate([di.ravel(order="F"), dx.ravel(order="F")])
    r64 = O.dip(z["data"], s, niter=5, liter=10, order=2, rect=(5, 5, 5), dtype=np.float64)
    dip_close(g, z["dip"], r64)
    _, orep = O.dip(z["data"], s, niter=5, liter=10, order=2, rect=(5, 5, 5), report=True)
    assert rep["cg_iters"] == orep[:, :, 0].tolist() and rep["ls_step"] == orep[:, :, 1].tolist()


def test_dip3d_mask_golden(ps):
    z = load("dip3d_mask")
    di, dx = ps.dip3dc(z["data"], niter=3, liter=10, rect=[3, 3, 3], mask=z["mask"])
    dip_close(np.concatenate([di.ravel(order="F"), dx.ravel(order="F")]), z["dip"])


@pytest.mark.parametrize("name,order", [("dip2d", 2), ("dip2d_o1", 1)])
def test_dip2d_golden(ps, name, order):
    z = load(name)
    g = ps.dip2dc(z["data"], niter=int(z["niter"]), liter=int(z["liter"]), order=order,
                  rect=[int(r) for r in z["rect"]])
    dip_close(g, z["dip"])


def test_dip2d_full_gather(ps):
    """The 512 x 256 gather workload against the FP32 oracle (which equals
    the reference bit for bit) and the FP64 oracle."""
    from oracle import oracle as O
    from pyseistr_amd import synth
    d = synth.gather2d(512, 256, seed=1)
    g, rep = ps.dip2dc(d, niter=5, liter=20, order=2, rect=[10, 20, 1], report=True)
    o32, orep = O.dip(d, (512, 256, 1), niter=5, liter=20, order=2, rect=(10, 20, 1), report=True)
    o64 = O.dip(d, (512, 256, 1), niter=5, liter=20, order=2, rect=(10, 20, 1), dtype=np.float64)
    dip_close(g, o32, o64)
    assert rep["cg_iters"] == orep[:, :, 0].tolist()


def test_dip_warm_start_and_errors(ps):
    from oracle import oracle as O
    from pyseistr_amd import synth
    v, _ = synth.volume3d((48, 16, 12), seed=21, n_events=2, smax=1.0)
    s = v.shape; n = v.size
    init = np.stack([np.full(s, 0.3, np.float32), np.full(s, -0.2, np.float32)], -1)
    di, dx = ps.dip3dc(v, niter=2, liter=10, rect=[3, 3, 3], init=init)
    ref = O.dip(v, s, niter=2, liter=10, order=2, rect=(3, 3, 3),
                p_init=np.concatenate([init[..., 0].ravel(order="F"), init[..., 1].ravel(order="F")]))
    dip_close(np.concatenate([di.ravel(order="F"), dx.ravel(order="F")]), ref)
    with pytest.raises(ValueError):
        ps.dip3dc(v, mask=np.ones((4, 4, 4), np.float32))
    with pytest.raises(NotImplementedError):
        ps.dip2dc(v[:, :, 0], order=7)


def test_somean3d_golden(ps):
    z = load("somean3d")
    same(ps.somean3dc(z["data"], z["dipi"].reshape(z["data"].shape, order="F"),
                      z["dipx"].reshape(z["data"].shape, order="F"), 2, 2, 0.01, 2), z["out"])


def test_somf3d_golden(ps):
    z = load("somf3d"); s = z["data"].shape
    pi = z["dipi"].reshape(s, order="F"); px = z["dipx"].reshape(s, order="F")
    same(ps.somf3dc(z["data"], pi, px, 2, 2, 0.01, 2), z["out"])
    same(ps.somf3dc(z["data"], pi, px, 3, 3, 0.01, 2), z["out_r3"])
    same(ps.somf3dc(z["data"], pi, px, 2, 1, 0.01, 1), z["out_o1"])


def test_pwspray3d_matches_oracle(ps):
    from oracle import oracle as O
    z = load("somf3d"); s = z["data"].shape
    pi = z["dipi"].reshape(s, order="F"); px = z["dipx"].reshape(s, order="F")
    for ns2, ns3, order in ((2, 2, 2), (1, 2, 1), (3, 1, 2)):
        u = ps.pwspray3d(z["data"], pi, px, ns2, ns3, order, 0.05)
        same(u, O.pwspray3d(z["data"], pi, px, s, ns2, ns3, order=order, eps=0.05))


def test_somean2d_somf2d_golden(ps):
    z = load("somean2d"); s = z["data"].shape
    same(ps.somean2dc(z["data"], z["dip"], 3, 2, 0.01), z["out"])
    same(ps.somean2dc(z["data"], z["dip"], 3, 2, 0.01, adj=1), z["out_adj"])
    z = load("somf2d")
    same(ps.somf2dc(z["data"], z["dip"], 3, 2, 0.01), z["out"])


def test_pwspray2d_fwd_adj(ps):
    from oracle import oracle as O
    z = load("somf2d")
    d = z["data"][:, :, 0]; p = z["dip"][:, :, 0]
    same(ps.pwspray2d(d, p, 3, 2, 0.01), O.pwspray2d(d, p, d.shape, 3, 2, 0.01))
    u = np.random.default_rng(3).standard_normal(d.size * 7).astype(np.float32)
    same(ps.filters.pwspray2d_adjoint(u, p, 3, 2, 0.01, d.shape), O.pwspray2d(u, p, d.shape, 3, 2, 0.01, adj=1))


def test_soint_sint_golden(ps):
    z = load("soint3d"); s = z["data"].shape
    args = (z["data"], z["mask"], z["dipi"].reshape(s, order="F"), z["dipx"].reshape(s, order="F"))
    same(ps.soint3dc(*args, order=2, niter=15), z["out"])
    same(ps.soint3dc(*args, order=2, niter=4, var=0.05), z["out_var"])
    same(ps.soint3dc(*args, order=1, niter=10, hasmask=0), z["out_nomask"])
    z = load("sint3d")
    same(ps.sint3dc(z["data"], z["mask"], z["dipi"].reshape(s, order="F"), z["dipx"].reshape(s, order="F"),
                    niter=8, eps=0.01, ns1=2, ns2=2, order1=2, order2=2), z["out"])
    z = load("soint2d")
    same(ps.soint2dc(z["data"], z["mask"], z["dip"], order=2, niter=15), z["out"])
    same(ps.soint2dc(z["data"], z["mask"], z["dip"], order=1, niter=10, hasmask=0), z["out_nomask"])
    same(ps.soint2dc(z["data"], z["mask"], np.stack([z["dip"], z["dip"]], -1), twoplane=1), z["data"])
    z = load("sint2d")
    same(ps.sint2dc(z["data"], z["mask"], z["dip"], niter=15, eps=0.01, ns=3, order=2), z["out"])


def test_somf3d_median_exact_on_own_spray(ps):
    """At a larger size: the GPU median equals numpy's order statistic of the
    GPU's own sprayed volume (the 33-slot window of a radius-4 spray)."""
    from pyseistr_amd import synth
    s = (96, 48, 40)
    v, _ = synth.volume3d(s, seed=22, n_events=4, smax=2.0)
    rng = np.random.default_rng(23)
    pi = (0.8 * rng.standard_normal(s)).astype(np.float32)
    px = (0.8 * rng.standard_normal(s)).astype(np.float32)
    u = ps.pwspray3d(v, pi, px, 4, 4, 2, 0.01)  # (n1, 81, n2, n3)
    med = np.partition(u[:, 40 - 16:40 + 17], 16, axis=1)[:, 16]
    out = ps.somf3dc(v, pi, px, 4, 4, 0.01, 2)
    same(out, med)
    mean = ps.somean3dc(v, pi, px, 4, 4, 0.01, 2)
    acc = np.zeros(s, np.float32)
    for k in range(81):
        acc = acc + u[:, k]
    same(mean, acc / np.float32(81))


def test_dip3d_plane_wave_full_size(ps):
    """256 x 128 x 128 (the 3-D dip workload size): a constant-slope plane
    wave is recovered (size-independent property)."""
    n1, n2, n3, sx, sy = 256, 128, 128, 0.7, -0.4
    t = np.arange(n1, dtype=np.float32)[:, None, None]
    x = np.arange(n2, dtype=np.float32)[None, :, None]; y = np.arange(n3, dtype=np.float32)[None, None, :]
    ph = t - sx * x - sy * y
    d = (np.sin(0.35 * ph) + 0.5 * np.sin(0.21 * ph + 1.0)).astype(np.float32)
    di, dx = ps.dip3dc(d, niter=5, liter=10, rect=[5, 5, 5])
    assert abs(np.median(di[20:-20, 5:-5, 5:-5]) - sx) < 0.01
    assert abs(np.median(dx[20:-20, 5:-5, 5:-5]) - sy) < 0.01


def test_bad_arguments_raise(ps):
    v = np.zeros((4, 4, 4), np.float32)
    with pytest.raises(ValueError):
        ps.somf3dc(v, v, v, 2, 2, 0.01, 2)  # n1 < 8
    w = np.zeros((32, 8, 8), np.float32)
    with pytest.raises(ValueError):
        ps.somf3dc(w, w, w, 2, 2, 0.01, 2, option=3)
    with pytest.raises(NotImplementedError):
        ps.soint3dc(w, w, w, w, order=7)  # PWD orders 1..6
    with pytest.raises(NotImplementedError):
        ps.somf3dc(w, w, w, 2, 2, 0.01, 5)  # predictor orders 1..4


def test_wave_axis1_pass_identical(ps, monkeypatch):
    """The plain axis-1 smoothing pass with one wave per trace chunk (the
    default) and with one block per chunk (SH_SMOOTH_WAVE=0), and the plain
    strided pass from LDS tiles (the default) and from registers per i1
    column (SH_SMOOTH_COL=1), run the same stencils and carry no sums, so
    the dips are bit-identical; n1 = 700 splits each trace into two
    chunks."""
    from pyseistr_amd import synth
    for shape, rect in (((96, 40, 36), [5, 5, 5]), ((700, 12, 5), [9, 3, 2])):
        v, _ = synth.volume3d(shape, seed=4, n_events=3, smax=2.0)
        a = np.stack(ps.dip3dc(v, niter=2, liter=6, rect=rect))
        monkeypatch.setenv("SH_SMOOTH_WAVE", "0")
        b = np.stack(ps.dip3dc(v, niter=2, liter=6, rect=rect))
        monkeypatch.delenv("SH_SMOOTH_WAVE")
        monkeypatch.setenv("SH_SMOOTH_COL", "1")
        c = np.stack(ps.dip3dc(v, niter=2, liter=6, rect=rect))
        monkeypatch.delenv("SH_SMOOTH_COL")
        assert np.array_equal(a, b), shape
        assert np.array_equal(a, c), shape


@pytest.mark.parametrize("mode", ["1", "3"])
def test_fused_smoothing_matches_separable(ps, monkeypatch, mode):
    """The opt-in fused smoothing passes (SH_SMOOTH_FUSED=1: all axes in the
    2.5-D kernel; 3: axis 1 separable, axes 2 and 3 in smooth_pair23)
    compute the same per-axis
    stencils in the same order as the separable passes; only the block
    partition of the double-precision CG sums differs, so the dips agree to
    1e-5 samples/trace."""
    from pyseistr_amd import synth
    v, _ = synth.volume3d((96, 40, 36), seed=9, n_events=3, smax=2.0)
    a = np.stack(ps.dip3dc(v, niter=3, liter=10, rect=[5, 5, 5]))
    monkeypatch.setenv("SH_SMOOTH_FUSED", mode)
    b = np.stack(ps.dip3dc(v, niter=3, liter=10, rect=[5, 5, 5]))
    monkeypatch.delenv("SH_SMOOTH_FUSED")
    assert np.abs(a - b).max() <= 1e-5


@pytest.mark.parametrize("shape,rect", [((40, 1, 1), (3, 1, 1)), ((37, 23, 1), (1, 4, 1)),
                                        ((20, 11, 9), (5, 5, 5)), ((5, 4, 3), (12, 2, 7)),
                                        ((128, 64, 48), (4, 6, 3))])
def test_smoothc_bit_exact(ps, shape, rect, monkeypatch):
    """smoothc (smoothcf) on the device against the FP32 oracle, which is
    pinned bit for bit against the reference's own smoothcf: the reference's
    operation order gives identical bits, for both forms, derivative, box,
    repeat and radii beyond the axis length; adj=1 runs the two-sweep kernel
    and, under SH_SMOOTHC_LINES=1, the one-thread-per-line kernel that adj=0
    always uses."""
    from oracle import oracle as O
    from pyseistr_amd import smooth
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(int(np.prod(shape))).astype(np.float32)
    d = _dev(x)
    for lines in (False, True):
        if lines:
            monkeypatch.setenv("SH_SMOOTHC_LINES", "1")
        for diff, box in [((0, 0, 0), (0, 0, 0)), ((1, 0, 1), (0, 0, 0)), ((0, 0, 0), (1, 1, 0))]:
            for repeat in (1, 2):
                for adj in (0, 1):
                    got = smooth.smooth_device(d, shape, rect, diff, box, repeat, adj).cpu().numpy()
                    exp = O.smoothc(x, shape, rect, diff, box, repeat, adj)
                    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), (diff, box, repeat, adj, lines)
    monkeypatch.delenv("SH_SMOOTHC_LINES")
    # the drop-in: 2-D input, squeezed output, reference defaults (adj=1)
    g = rng.standard_normal((50, 30)).astype(np.float32)
    out = ps.smoothc(g, rect=[1, 5, 1])
    assert out.shape == g.shape and out.dtype == np.float32
    exp = O.smoothc(g.ravel(order="F"), (50, 30, 1), (1, 5, 1)).reshape(50, 30, order="F")
    assert np.array_equal(out, exp)


@pytest.mark.parametrize("shape", [(200, 3, 1), (64, 10, 2), (512, 256, 8), (1, 70, 1), (130, 129, 1)])
def test_bandpassc_bit_exact(ps, shape, monkeypatch):
    """bandpassc (cbp) on the device against the FP32 oracle, which is pinned
    bit for bit against the reference's own cbp: coefficients from the same
    host expressions, every section in the reference's order of float
    operations, the zero-phase pass run backwards.  Covers the fused LDS
    kernel (sections cascaded per sample, 64-trace tiles with ragged sample
    and trace tails), the fused minimum-phase pass of both filters, and the
    one-thread-per-trace kernel that takes passes of more than 8 sections
    (and runs everything under SH_BANDPASS_PLAIN=1)."""
    from oracle import oracle as O
    from pyseistr_amd import bandpass
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(int(np.prod(shape))).astype(np.float32)
    d = _dev(x)
    cases = [(0.0, 30.0, 6, 6), (5.0, 40.0, 3, 5), (2.5, 1000.0, 8, 2), (10.0, 60.0, 1, 1), (3.0, 70.0, 9, 9),
             (20.0, 45.0, 20, 18)]
    for plain in (False, True):
        if plain:
            monkeypatch.setenv("SH_BANDPASS_PLAIN", "1")
        for flo, fhi, nplo, nphi in cases:
            for phase in (0, 1):
                got = bandpass.bandpass_device(d, shape, 0.004, flo, fhi, nplo, nphi, phase).cpu().numpy()
                exp = O.bandpass(x, shape, 0.004, flo, fhi, nplo, nphi, phase)
                assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), (flo, fhi, nplo, nphi, phase, plain)
    monkeypatch.delenv("SH_BANDPASS_PLAIN")
    # the drop-in: 1-D and 2-D inputs keep their shapes
    w = rng.standard_normal(300).astype(np.float32)
    out = ps.bandpassc(w, 0.004, flo=0, fhi=20)
    assert out.shape == w.shape
    assert np.array_equal(out, O.bandpass(w, (300, 1, 1), 0.004, 0.0, 20.0))
    g = rng.standard_normal((100, 7)).astype(np.float32)
    out = ps.bandpassc(g, 0.002, flo=3, fhi=80, nplo=4, nphi=4, phase=1)
    assert out.shape == g.shape
    with pytest.raises(NotImplementedError):
        bandpass.bandpass_device(d, shape, 0.004, 5.0, 40.0, 300, 6, 1)


def test_smoothc_bandpassc_full_size_bit_exact(ps):
    """At the benchmark's 512 x 256 x 256: smoothc (rect 5,5,5, and a
    derivative / box mix) and bandpassc (zero and minimum phase) bit for bit
    against the FP32 oracle."""
    from oracle import oracle as O
    from pyseistr_amd import bandpass, smooth
    shape = (512, 256, 256)
    x = np.random.default_rng(31).standard_normal(int(np.prod(shape))).astype(np.float32)
    d = _dev(x)
    for rect, diff, box in [((5, 5, 5), (0, 0, 0), (0, 0, 0)), ((3, 9, 2), (0, 1, 0), (1, 0, 0))]:
        got = smooth.smooth_device(d, shape, rect, diff, box).cpu().numpy()
        assert np.array_equal(got.view(np.uint32), O.smoothc(x, shape, rect, diff, box).view(np.uint32)), rect
    for phase in (0, 1):
        got = bandpass.bandpass_device(d, shape, 0.004, 5.0, 60.0, 6, 6, phase).cpu().numpy()
        exp = O.bandpass(x, shape, 0.004, 5.0, 60.0, 6, 6, phase)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), phase


def test_dip_directions_on_two_streams_identical(ps, monkeypatch):
    """dip3d runs its two directions side by side (the crossline one on the
    context's second stream, with its own workspace); they are independent,
    so dips and iteration records are bit-identical to running them one
    after the other on one stream (SH_DIP_STREAMS=1), with a mask and a
    warm start as well; a following call on the same context sees the
    joined results."""
    from pyseistr_amd import synth
    v, _ = synth.volume3d((80, 30, 28), seed=12, n_events=3, smax=2.0, spike_frac=0.01)
    mask = (np.random.default_rng(4).random(v.shape) > 0.1).astype(np.float32)
    init = (0.3 * np.ones(v.shape, np.float32), -0.2 * np.ones(v.shape, np.float32))
    kw = dict(niter=3, liter=10, rect=[5, 5, 5], report=True)
    a, ra = ps.dip3dc(v, mask=mask, init=init, **kw)
    a2, _ = ps.dip3dc(v, **kw)
    monkeypatch.setenv("SH_DIP_STREAMS", "1")
    b, rb = ps.dip3dc(v, mask=mask, init=init, **kw)
    b2, _ = ps.dip3dc(v, **kw)
    monkeypatch.delenv("SH_DIP_STREAMS")
    for x, y in ((a, b), (a2, b2)):
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
    assert ra["cg_iters"] == rb["cg_iters"] and ra["ls_step"] == rb["ls_step"]


@pytest.mark.parametrize("tile", ["32x128", "64x128", "64x32"])
def test_strided_smoothing_tiles_agree(ps, monkeypatch, tile):
    """The strided smoothing passes give the same stencils for every tile
    shape (SH_SMOOTH_TILE; 64x128 needs more rows per thread than the
    prefetch ring holds, so it also covers the unpipelined path); only the
    block partition of the double-precision CG sums differs, so the dips
    agree with the default tiles to 1e-5 samples/trace."""
    from pyseistr_amd import synth
    v, _ = synth.volume3d((96, 40, 36), seed=9, n_events=3, smax=2.0)
    a = np.stack(ps.dip3dc(v, niter=3, liter=10, rect=[5, 5, 5]))
    monkeypatch.setenv("SH_SMOOTH_TILE", tile)
    b = np.stack(ps.dip3dc(v, niter=3, liter=10, rect=[5, 5, 5]))
    monkeypatch.delenv("SH_SMOOTH_TILE")
    assert np.abs(a - b).max() <= 1e-5


def test_dip3d_long_run_matches_oracle(ps):
    """Thirty Gauss-Newton iterations (the slow drift of the reference's
    heavily damped update, not a device artefact): the dips stay within the
    dip bars of the FP32 oracle and no farther from the FP64 oracle than the
    FP32 oracle is (+1e-4)."""
    from oracle import oracle as O
    from pyseistr_amd import synth
    s = (96, 40, 36)
    v, _ = synth.volume3d(s, seed=5, n_events=3, smax=3.0, sigma=0.3, spike_frac=0.01)
    di, dx = ps.dip3dc(v, niter=30, liter=10, rect=[5, 5, 5])
    g = np.concatenate([di.ravel(order="F"), dx.ravel(order="F")]).astype(np.float64)
    o32 = O.dip(v, s, niter=30, liter=10, order=2, rect=(5, 5, 5))
    o64 = O.dip(v, s, niter=30, liter=10, order=2, rect=(5, 5, 5), dtype=np.float64)
    d = np.abs(g - o32)
    assert d.max() <= 2e-3 and np.sqrt(np.mean(d ** 2)) <= 2e-4
    assert np.abs(g - o64).max() <= np.abs(o32 - o64).max() + 1e-4


@pytest.mark.parametrize("n1,n2,order,niter,hasmask,twoplane", [(40, 24, 1, 10, 1, 0), (60, 30, 2, 25, 1, 0),
                                                                  (33, 17, 2, 8, 0, 0), (40, 24, 2, 12, 1, 1),
                                                                  (50, 20, 1, 3, 1, 1), (64, 9, 3, 6, 1, 0)])
def test_soint2d_prec_bit_exact(ps, n1, n2, order, niter, hasmask, twoplane):
    """soint2dc(prec=1): the preconditioned solve (predict_lop / predict2_lop
    preconditioner, ps_solver_prec) against the FP32 oracle, which equals the
    reference's csoint2d bit for bit; iteration counts agree too."""
    from oracle import oracle as O
    rng = np.random.default_rng(n1 * 100 + n2 + niter)
    d = rng.standard_normal((n1, n2)).astype(np.float32)
    mask = (rng.random((n1, n2)) > 0.5).astype(np.float32)
    if hasmask == 0:
        d = d * mask
    dip = (0.7 * rng.standard_normal((n1, n2, 2))).astype(np.float32)
    dd = dip if twoplane else dip[:, :, 0]
    got, it = ps.soint2dc(d, mask, dd, order=order, niter=niter, hasmask=hasmask, twoplane=twoplane, prec=1,
                          report=True)
    exp = O.soint2d(d.ravel(order="F"), mask.ravel(order="F"), dip[:, :, 0].ravel(order="F"), (n1, n2), order=order,
                    niter=niter, hasmask=hasmask, twoplane=twoplane, prec=1, dip2=dip[:, :, 1].ravel(order="F"))
    same(got, exp)
    assert it == niter
    # niter = 0 returns the data
    same(ps.soint2dc(d, mask, dd, niter=0, hasmask=hasmask, twoplane=twoplane, prec=1), d)


def test_soint2d_prec_chain_kernels_identical(ps, monkeypatch):
    """The three chain kernels give the same bits: the default one (per-sample
    work spread over the block) and the ones taken when its LDS staging would
    exceed 64 KB (SH_SOINT_PREC_CHAIN=lds: one thread from LDS;
    SH_SOINT_PREC_CHAIN=global: one thread from HBM)."""
    rng = np.random.default_rng(5)
    d = rng.standard_normal((48, 20)).astype(np.float32)
    mask = (rng.random((48, 20)) > 0.4).astype(np.float32)
    dip = (0.6 * rng.standard_normal((48, 20, 2))).astype(np.float32)
    for twoplane in (0, 1):
        dd = dip if twoplane else dip[:, :, 0]
        a = ps.soint2dc(d, mask, dd, order=2, niter=7, twoplane=twoplane, prec=1)
        for mode in ("lds", "global"):
            monkeypatch.setenv("SH_SOINT_PREC_CHAIN", mode)
            b = ps.soint2dc(d, mask, dd, order=2, niter=7, twoplane=twoplane, prec=1)
            monkeypatch.delenv("SH_SOINT_PREC_CHAIN")
            same(a, b)


@pytest.mark.parametrize("shape,nj,order,niter", [((40, 12, 9), (1, 1), 2, 8), ((33, 10, 7), (2, 1), 1, 6),
                                                   ((50, 9, 8), (1, 2), 2, 5)])
def test_soint3d_drift_bit_exact(ps, shape, nj, order, niter):
    """soint3dc(drift=1): stencils shifted by NINT(p) (slopes up to ~8
    samples/trace, so shifts and skipped samples occur) against the FP32
    oracle, which equals the reference bit for bit."""
    from oracle import oracle as O
    rng = np.random.default_rng(sum(shape) + niter)
    d = rng.standard_normal(shape).astype(np.float32)
    mask = (rng.random(shape) > 0.4).astype(np.float32)
    pi = (2.5 * rng.standard_normal(shape)).astype(np.float32)
    px = (2.0 * rng.standard_normal(shape)).astype(np.float32)
    got = ps.soint3dc(d, mask, pi, px, order=order, niter=niter, njs=list(nj), drift=1)
    f = lambda a: a.ravel(order="F")  # noqa: E731
    same(got, O.soint3d(f(d), f(mask), f(pi), f(px), shape, order=order, niter=niter, njs=nj, drift=1))


@pytest.mark.parametrize("n1,n2,nj,order,niter", [(40, 20, 1, 2, 10), (37, 15, 2, 1, 6), (64, 30, 1, 3, 7)])
def test_soint2d_drift_bit_exact(ps, n1, n2, nj, order, niter):
    from oracle import oracle as O
    rng = np.random.default_rng(n1 + n2 + niter)
    d = rng.standard_normal((n1, n2)).astype(np.float32)
    mask = (rng.random((n1, n2)) > 0.4).astype(np.float32)
    p = (2.5 * rng.standard_normal((n1, n2))).astype(np.float32)
    got = ps.soint2dc(d, mask, p, order=order, niter=niter, njs=[nj, 1], drift=1)
    f = lambda a: a.ravel(order="F")  # noqa: E731
    same(got, O.soint2d(f(d), f(mask), f(p), (n1, n2), order=order, niter=niter, njs=(nj, 1), drift=1))


@pytest.mark.parametrize("shape,r", [((40, 12, 10), 2), ((36, 9, 8), 3)])
def test_somf3d_space_varying_median(ps, shape, r):
    """somf3dc(option=2): the space-varying median over the full spray
    against the FP32 oracle (identical sprays, the float mean summed in the
    same order, exact selections; zeros compared without their sign).  The
    oracle takes the edge column where the reference reads one row past its
    panel, so agreement with the reference itself is not pinned."""
    from oracle import oracle as O
    from pyseistr_amd import synth
    v, _ = synth.volume3d(shape, seed=31 + r, n_events=3, smax=2.0, sigma=0.3, spike_frac=0.02)
    rng = np.random.default_rng(r)
    pi = (0.8 * rng.standard_normal(shape)).astype(np.float32)
    px = (0.8 * rng.standard_normal(shape)).astype(np.float32)
    got = ps.somf3dc(v, pi, px, r, r, 0.01, 2, option=2)
    same(got, O.somf3d(v, pi, px, shape, r, r, order=2, option=2))


@pytest.mark.parametrize("shape,ns", [((60, 30, 2), 3), ((48, 20, 1), 5), ((512, 256, 1), 8), ((256, 64, 3), 6)])
def test_somf2d_space_varying_median(ps, shape, ns):
    from oracle import oracle as O
    rng = np.random.default_rng(ns)
    d = rng.standard_normal(shape).astype(np.float32)
    d[:, ::7] *= 8.0  # loud traces so every window length occurs
    p = (0.7 * rng.standard_normal(shape)).astype(np.float32)
    got = ps.somf2dc(d if shape[2] > 1 else d[:, :, 0], p if shape[2] > 1 else p[:, :, 0], ns, 2, 0.01, option=2)
    same(got, O.somf2d(d, p, shape, ns, order=2, option=2))


@pytest.mark.parametrize("order", [5, 6])
def test_high_pwd_orders(ps, order):
    """PWD orders above 4 (apfilt_init takes any nw, dip_cfuns.c:833-856) in
    the stencil paths: dip2dc / dip3dc within the dip bars of the FP32
    oracle with the same decisions, soint3dc / soint2dc bit-exact."""
    from oracle import oracle as O
    from pyseistr_amd import synth
    g = synth.gather2d(96, 40, seed=3)
    got, rep = ps.dip2dc(g, niter=3, liter=10, order=order, rect=[5, 5, 1], report=True)
    exp, orep = O.dip(g, (96, 40, 1), niter=3, liter=10, order=order, rect=(5, 5, 1), report=True)
    dip_close(got, exp)
    assert rep["cg_iters"] == orep[:, :, 0].tolist() and rep["ls_step"] == orep[:, :, 1].tolist()
    s = (48, 16, 12)
    v, _ = synth.volume3d(s, seed=order, n_events=2, smax=1.0)
    (di, dx), rep = ps.dip3dc(v, niter=2, liter=10, order=order, rect=[3, 3, 3], report=True)
    exp, orep = O.dip(v, s, niter=2, liter=10, order=order, rect=(3, 3, 3), report=True)
    dip_close(np.concatenate([di.ravel(order="F"), dx.ravel(order="F")]), exp)
    assert rep["cg_iters"] == orep[:, :, 0].tolist()
    rng = np.random.default_rng(order)
    m = (rng.random(s) > 0.5).astype(np.float32)
    f = lambda a: a.ravel(order="F")  # noqa: E731
    same(ps.soint3dc(v * m, m, di, dx, order=order, niter=6),
         O.soint3d(f(v * m), f(m), f(di), f(dx), s, order=order, niter=6))
    same(ps.soint2dc(v[:, :, 0] * m[:, :, 0], m[:, :, 0], di[:, :, 0], order=order, niter=6),
         O.soint2d(f(v[:, :, 0] * m[:, :, 0]), f(m[:, :, 0]), f(di[:, :, 0]), s[:2], order=order, niter=6))
