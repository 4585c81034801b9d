"""GPU parity of the warm-start network kernels against the CPU oracle.

The oracle (oracle/cnn_oracle.py) rounds activations, weights and
back-propagated gradients to bfloat16 at the same points as the kernels and
sums in float64; the kernels sum in float32 on the matrix cores.  Bars
(stated per test): forward relative L2 error <= 5e-3, per-layer gradient
relative L2 error <= 2e-2, loss relative error <= 1e-3; Adam (FP32)
relative error <= 1e-5 against the float64 formula.
"""
import os

import numpy as np
import pytest

from oracle import cnn_oracle as K

pytestmark = pytest.mark.gpu

SHAPES = [(40, 12, 9), (70, 6, 5)]


@pytest.fixture(scope="module")
def ws():
    import torch
    from pyseistr_amd import _lib, warmstart
    assert torch.cuda.is_available()
    assert not _lib.missing_symbols()
    return warmstart


def _case(seed, shape):
    rng = np.random.default_rng(seed)
    p = K.init_params(seed)
    p = (p + 0.02 * rng.standard_normal(p.size)).astype(np.float32)
    n = int(np.prod(shape))
    x = (rng.standard_normal(n) * 3.0).astype(np.float32)
    t = (rng.standard_normal(2 * n) * 0.5).astype(np.float32)
    wt = (rng.random(n) > 0.2).astype(np.float32)
    return p, x, t, wt


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def _blocks():
    out = {"W1": slice(K.OFF_W1, K.OFF_B1), "b1": slice(K.OFF_B1, K.OFF_B1 + K.C)}
    for l in (2, 3, 4, 5):
        o = K.OFF_L[l]
        out[f"W{l}"] = slice(o, o + K.NT * K.C * K.C)
        out[f"b{l}"] = slice(o + K.NT * K.C * K.C, o + K._L)
    out["W6"] = slice(K.OFF_W6, K.OFF_B6)
    out["b6"] = slice(K.OFF_B6, K.OFF_B6 + 2)
    return out


def test_nparams(ws):
    from pyseistr_amd import _lib
    assert _lib.load().sh_cnn_nparams() == K.NPARAM == ws.NPARAM


@pytest.mark.parametrize("shape", SHAPES)
def test_forward_matches_oracle(ws, shape):
    p, x, _, _ = _case(1, shape)
    xs = 0.37
    got = ws.cnn_forward(_dev(p), _dev(x), shape, xs).cpu().numpy()
    ref = K.predict(p, x, shape, xscale=xs, quant=True)
    print("forward rel err", _rel(got, ref))
    assert _rel(got, ref) <= 5e-3
    exact = K.predict(p, x, shape, xscale=xs, quant=False)
    print("bf16 network vs exact float64 network", _rel(got, exact))


@pytest.mark.parametrize("shape", SHAPES)
def test_gradient_matches_oracle(ws, shape):
    p, x, t, wt = _case(2, shape)
    xs = 0.5
    loss, g = ws.cnn_grad(_dev(p), _dev(x), shape, xs, _dev(t), _dev(wt))
    g = g.cpu().numpy()
    lref, gref = K.loss_and_grad(p, x, t, shape, wt=wt, xscale=xs, quant=True)
    assert abs(loss - lref) <= 1e-3 * abs(lref)
    for name, sl in _blocks().items():
        e = _rel(g[sl], gref[sl])
        print(f"grad {name}: rel err {e:.2e}")
        assert e <= 2e-2, name
    # without weights (wt = NULL) the loss is the plain mean square
    loss2, _ = ws.cnn_grad(_dev(p), _dev(x), shape, xs, _dev(t), None)
    l2, _ = K.loss_and_grad(p, x, t, shape, wt=None, xscale=xs, quant=True)
    assert abs(loss2 - l2) <= 1e-3 * abs(l2)


@pytest.mark.parametrize("dil", [(1, 2, 4, 8, 2, 1), (3, 1, 2, 1, 5, 2)])
def test_dilated_forward_and_gradient(ws, dil):
    shape = (70, 21, 18)
    p, x, t, wt = _case(4, shape)
    xs = 0.6
    got = ws.cnn_forward(_dev(p), _dev(x), shape, xs, dil=dil).cpu().numpy()
    ref = K.predict(p, x, shape, xscale=xs, quant=True, dil=dil)
    print("dilated forward rel err", _rel(got, ref))
    assert _rel(got, ref) <= 5e-3
    loss, g = ws.cnn_grad(_dev(p), _dev(x), shape, xs, _dev(t), _dev(wt), dil=dil)
    g = g.cpu().numpy()
    lref, gref = K.loss_and_grad(p, x, t, shape, wt=wt, xscale=xs, quant=True, dil=dil)
    assert abs(loss - lref) <= 1e-3 * abs(lref)
    for name, sl in _blocks().items():
        e = _rel(g[sl], gref[sl])
        print(f"grad {name}: rel err {e:.2e}")
        assert e <= 2e-2, name


@pytest.mark.parametrize("shape", [(300, 10, 6), (200, 50, 45)])
def test_conv_kernels_identical(ws, monkeypatch, shape):
    """The three dilation-1 convolution kernels (plane-scatter, the
    default; wide tiles; plain) sum every output's taps in the same order,
    so forward and gradient must be bit-identical; so do the two first-layer
    kernels (input staged in LDS, the default; 27 global loads per voxel,
    SH_CNN_C1=g).  (300, 10, 6) has an i1 tail tile and rows past n2;
    (200, 50, 45) has more tiles than blocks, so a block's z-range crosses
    from one column into the next."""
    p, x, t, wt = _case(5, shape)
    xs = 0.4
    got = ws.cnn_forward(_dev(p), _dev(x), shape, xs).cpu().numpy()
    loss, g = ws.cnn_grad(_dev(p), _dev(x), shape, xs, _dev(t), _dev(wt))
    g = g.cpu().numpy()
    for var, kind in (("SH_CNN_CONV", "n"), ("SH_CNN_CONV", "w"), ("SH_CNN_C1", "g")):
        monkeypatch.setenv(var, kind)
        got_k = ws.cnn_forward(_dev(p), _dev(x), shape, xs).cpu().numpy()
        loss_k, g_k = ws.cnn_grad(_dev(p), _dev(x), shape, xs, _dev(t), _dev(wt))
        monkeypatch.delenv(var)
        assert np.array_equal(got, got_k), kind
        assert loss == loss_k and np.array_equal(g, g_k.cpu().numpy()), kind
    if np.prod(shape) <= 20000:
        ref = K.predict(p, x, shape, xscale=xs, quant=True)
        assert _rel(got, ref) <= 5e-3
        lref, gref = K.loss_and_grad(p, x, t, shape, wt=wt, xscale=xs, quant=True)
        for name, sl in _blocks().items():
            assert _rel(g[sl], gref[sl]) <= 2e-2, name


def test_dilation_out_of_range_rejected(ws):
    with pytest.raises(ValueError):
        ws.cnn_forward(_dev(K.init_params(0)), _dev(np.zeros(64, np.float32)), (4, 4, 4), 1.0,
                       dil=(1, 1, 9, 1, 1, 1))


def test_adam_matches_formula(ws):
    rng = np.random.default_rng(3)
    p = rng.standard_normal(K.NPARAM).astype(np.float32)
    pd, md, vd = _dev(p), _dev(np.zeros_like(p)), _dev(np.zeros_like(p))
    ph, mh, vh = p.astype(np.float64), np.zeros(K.NPARAM), np.zeros(K.NPARAM)
    for step in range(1, 4):
        g = rng.standard_normal(K.NPARAM).astype(np.float32)
        ws.cnn_adam(pd, _dev(g), md, vd, 1e-3, step, wd=1e-4)
        ph, mh, vh = K.adam(ph, g, mh, vh, 1e-3, 0.9, 0.999, 1e-8, step, wd=1e-4)
    got = pd.cpu().numpy()
    assert np.abs(got - ph).max() <= 1e-5 * max(1.0, np.abs(ph).max())


def test_training_reduces_loss(ws):
    """A few hundred Adam steps on one window fit its dip3d dips."""
    import torch
    from pyseistr_amd import dip, synth
    shape = (64, 32, 32)
    v, _ = synth.volume3d(shape, seed=21, n_events=3, smax=3.0, sigma=0.1, spike_frac=0.0)
    d = torch.from_numpy(v.ravel(order="F")).cuda()
    tgt = dip.dip_device(d, shape, niter=10, liter=10, order=2, rect=(5, 5, 5))
    xs = ws.input_scale(d)
    p = _dev(ws.init_params(0))
    m, vv = torch.zeros_like(p), torch.zeros_like(p)
    l0, g = ws.cnn_grad(p, d, shape, xs, tgt)
    for s in range(1, 301):
        ws.cnn_grad(p, d, shape, xs, tgt, grad=g, want_loss=False)
        ws.cnn_adam(p, g, m, vv, 1e-3, s)
    l1, _ = ws.cnn_grad(p, d, shape, xs, tgt)
    print("loss", l0, "->", l1)
    assert np.isfinite(l1) and l1 < 0.3 * l0


def test_device_input_scale(ws):
    """xscale <= 0 makes the library scale the input by 1 / RMS on the
    device: the same prediction as the host-computed scale (to float
    rounding of the scale), and scale 1 for an all-zero volume."""
    import torch
    shape = (40, 18, 12)
    rng = np.random.default_rng(8)
    p = _dev(K.init_params(2) + 0.02 * rng.standard_normal(K.NPARAM).astype(np.float32))
    d = torch.from_numpy((3.0 * rng.standard_normal(int(np.prod(shape)))).astype(np.float32)).cuda()
    host = ws.cnn_forward(p, d, shape, ws.input_scale(d)).cpu().numpy()
    dev = ws.cnn_forward(p, d, shape, 0.0).cpu().numpy()
    assert np.allclose(dev, host, rtol=1e-4, atol=1e-5), np.abs(dev - host).max()
    z = torch.zeros_like(d)
    np.testing.assert_array_equal(ws.cnn_forward(p, z, shape, 0.0).cpu().numpy(),
                                  ws.cnn_forward(p, z, shape, 1.0).cpu().numpy())


def test_errors(ws):
    import torch
    from pyseistr_amd import _lib
    ctx = _lib.context()
    lib = _lib.load()
    p = torch.zeros(K.NPARAM, device="cuda")
    assert lib.sh_cnn_forward(ctx, None, _lib.ptr(p), 1.0, None, 4, 4, 4, _lib.ptr(p)) == -2
    assert lib.sh_cnn_forward(ctx, _lib.ptr(p), _lib.ptr(p), 1.0, None, 0, 4, 4, _lib.ptr(p)) == -1
    assert lib.sh_cnn_adam(ctx, _lib.ptr(p), _lib.ptr(p), _lib.ptr(p), _lib.ptr(p), 1e-3, 0.9, 0.999, 1e-8, 0,
                           0.0) == -1


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "pyseistr_amd", "weights",
                                                    "warmstart.npz")), reason="no trained weights committed")
def test_warm_start_gain(ws):
    """On a held-out volume (the seed-5 generator; training used seeds
    10000+), the trained network + 2 Gauss-Newton iterations is closer to the
    converged dip3d than the reference's default 5-iteration cold start, and
    the cold start needs at least 8 iterations to match it."""
    import torch
    from pyseistr_amd import dip, synth
    shape = (256, 128, 128)
    v, _ = synth.volume3d(shape, seed=5, n_events=4, smax=3.0, sigma=0.3, spike_frac=0.01)
    d = torch.from_numpy(v.ravel(order="F")).cuda()
    kw = dict(liter=10, order=2, rect=(5, 5, 5))
    conv = dip.dip_device(d, shape, niter=400, **kw)
    net = ws.WarmStart.load()

    def rmse(a):
        return float(torch.sqrt(torch.mean((a - conv) ** 2)).item())

    warm2 = net.dip_device(d, shape, niter=2, **kw)
    cold, errs = None, []
    for _ in range(8):
        cold = dip.dip_device(d, shape, niter=1, init=cold, **kw)
        errs.append(rmse(cold))
    print({"warm2": rmse(warm2), "cold1..8": [round(e, 4) for e in errs]})
    assert rmse(warm2) < errs[4]
    assert all(e > rmse(warm2) for e in errs[:7])
