"""Pins of the warm-start network's CPU oracle (no GPU).

The oracle's exact (float64, unquantised) network is checked against an
independent implementation, torch's conv3d and autograd in float64, and
against central finite differences; its bfloat16 rounding against torch's.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import cnn_oracle as K

SHAPE = (9, 7, 5)


def _torch_w(W):
    ci, co = W.shape[1], W.shape[2]
    return torch.from_numpy(W.reshape(3, 3, 3, ci, co).transpose(4, 3, 0, 1, 2).copy())


def _torch_net(p, x, shape):
    W, b = K.unpack(p)
    n1, n2, n3 = shape
    a = torch.from_numpy(np.asarray(x, np.float64).reshape(1, 1, n3, n2, n1))
    for l in range(1, 7):
        a = F.conv3d(a, _torch_w(W[l]), torch.from_numpy(b[l]), padding=1)
        if l < 6:
            a = torch.relu(a)
    return a  # (1, 2, n3, n2, n1)


def _case(seed=0, shape=SHAPE):
    rng = np.random.default_rng(seed)
    p = K.init_params(seed).astype(np.float64)
    p += 0.05 * rng.standard_normal(p.size)  # non-zero biases too
    n = int(np.prod(shape))
    x = rng.standard_normal(n)
    t = rng.standard_normal(2 * n) * 0.5
    return p, x, t


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    v = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 10000),
                        # exact ties: low 16 bits 0x8000, both parities of the kept bit
                        (np.arange(64, dtype=np.uint32) << 16 | 0x3F800000 | 0x8000).view(np.float32),
                        np.array([0.0, -0.0, 1.0, -1.0, 3.0e38], np.float32)])
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(K.bf16(v), ref)


def test_forward_matches_torch_conv3d():
    p, x, _ = _case(0)
    _, y = K.forward(p, x, SHAPE, quant=False)
    ref = _torch_net(p, x, SHAPE)[0].numpy().transpose(1, 2, 3, 0)
    assert np.abs(y - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_gradient_matches_torch_autograd():
    p, x, t = _case(1)
    n1, n2, n3 = SHAPE
    n = n1 * n2 * n3
    wt = np.random.default_rng(5).random(n)
    loss, g = K.loss_and_grad(p, x, t, SHAPE, wt=wt, quant=False)
    pt = torch.from_numpy(p.copy()).requires_grad_(True)
    W, b = {}, {}
    W[1] = pt[K.OFF_W1:K.OFF_B1].reshape(K.NT, 1, K.C)
    b[1] = pt[K.OFF_B1:K.OFF_B1 + K.C]
    for l in (2, 3, 4, 5):
        o = K.OFF_L[l]
        W[l] = pt[o:o + K.NT * K.C * K.C].reshape(K.NT, K.C, K.C)
        b[l] = pt[o + K.NT * K.C * K.C:o + K._L]
    W[6] = pt[K.OFF_W6:K.OFF_B6].reshape(K.NT, K.C, 2)
    b[6] = pt[K.OFF_B6:K.OFF_B6 + 2]
    a = torch.from_numpy(x.reshape(1, 1, n3, n2, n1))
    for l in range(1, 7):
        ci, co = W[l].shape[1], W[l].shape[2]
        wt_ = W[l].reshape(3, 3, 3, ci, co).permute(4, 3, 0, 1, 2)
        a = F.conv3d(a, wt_, b[l], padding=1)
        if l < 6:
            a = torch.relu(a)
    tt = torch.from_numpy(t.reshape(2, n3, n2, n1))
    m = torch.from_numpy(wt.reshape(n3, n2, n1))
    L = torch.sum(m * (a[0] - tt) ** 2) / (2 * n)
    L.backward()
    assert abs(L.item() - loss) < 1e-12 * max(1.0, abs(loss))
    ref = pt.grad.numpy()
    assert np.abs(g - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


def test_gradient_finite_differences():
    p, x, t = _case(2, (6, 5, 4))
    shape = (6, 5, 4)
    _, g = K.loss_and_grad(p, x, t, shape, quant=False)
    rng = np.random.default_rng(3)
    idx = list(rng.choice(K.NPARAM, 12, replace=False)) + [K.OFF_B1 + 3, K.OFF_B6 + 1, K.OFF_W6 + 5]
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = K.loss_and_grad(pp, x, t, shape, quant=False)
        lm, _ = K.loss_and_grad(pm, x, t, shape, quant=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1e-3, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_quantised_network_close_to_exact():
    p, x, t = _case(4)
    _, y = K.forward(p, x, SHAPE, quant=False)
    _, yq = K.forward(p, x, SHAPE, quant=True)
    assert np.linalg.norm(yq - y) / np.linalg.norm(y) < 3e-2
    _, g = K.loss_and_grad(p, x, t, SHAPE, quant=False)
    _, gq = K.loss_and_grad(p, x, t, SHAPE, quant=True)
    assert np.linalg.norm(gq - g) / np.linalg.norm(g) < 5e-2


def test_adam_matches_torch():
    rng = np.random.default_rng(6)
    p = rng.standard_normal(50)
    m = np.zeros(50)
    v = np.zeros(50)
    pt = torch.from_numpy(p.copy()).requires_grad_(True)
    opt = torch.optim.Adam([pt], lr=1e-2, betas=(0.9, 0.99), eps=1e-8, weight_decay=1e-3)
    for step in range(1, 4):
        g = rng.standard_normal(50)
        p, m, v = K.adam(p, g, m, v, 1e-2, 0.9, 0.99, 1e-8, step, wd=1e-3)
        pt.grad = torch.from_numpy(g.copy())
        opt.step()
    assert np.abs(p - pt.detach().numpy()).max() < 1e-12


@pytest.mark.parametrize("l", [2, 5])
def test_input_grad_is_adjoint_of_conv(l):
    """<conv(a), g> == <a, conv_input_grad(g)> (the transpose identity)."""
    rng = np.random.default_rng(l)
    a = rng.standard_normal((4, 5, 6, K.C))
    g = rng.standard_normal((4, 5, 6, K.C))
    W = rng.standard_normal((K.NT, K.C, K.C))
    lhs = np.sum(K.conv(a, W) * g)
    rhs = np.sum(a * K.conv_input_grad(g, W))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_dilated_network_matches_torch():
    """Dilated layers (offsets scaled by the layer's dilation) against torch
    conv3d with the same dilation, forward and gradient, float64."""
    dil = (1, 2, 3, 2, 1, 1)
    shape = (11, 9, 8)
    p, x, t = _case(7, shape)
    n1, n2, n3 = shape
    n = n1 * n2 * n3
    loss, g = K.loss_and_grad(p, x, t, shape, quant=False, dil=dil)
    _, y = K.forward(p, x, shape, quant=False, dil=dil)
    pt = torch.from_numpy(p.copy()).requires_grad_(True)
    W, b = {}, {}
    W[1] = pt[K.OFF_W1:K.OFF_B1].reshape(K.NT, 1, K.C)
    b[1] = pt[K.OFF_B1:K.OFF_B1 + K.C]
    for l in (2, 3, 4, 5):
        o = K.OFF_L[l]
        W[l] = pt[o:o + K.NT * K.C * K.C].reshape(K.NT, K.C, K.C)
        b[l] = pt[o + K.NT * K.C * K.C:o + K._L]
    W[6] = pt[K.OFF_W6:K.OFF_B6].reshape(K.NT, K.C, 2)
    b[6] = pt[K.OFF_B6:K.OFF_B6 + 2]
    a = torch.from_numpy(x.reshape(1, 1, n3, n2, n1))
    for l in range(1, 7):
        ci, co = W[l].shape[1], W[l].shape[2]
        a = F.conv3d(a, W[l].reshape(3, 3, 3, ci, co).permute(4, 3, 0, 1, 2), b[l], padding=dil[l - 1],
                     dilation=dil[l - 1])
        if l < 6:
            a = torch.relu(a)
    assert np.abs(a[0].detach().numpy().transpose(1, 2, 3, 0) - y).max() < 1e-11 * max(1.0, np.abs(y).max())
    L = torch.sum((a[0] - torch.from_numpy(t.reshape(2, n3, n2, n1))) ** 2) / (2 * n)
    L.backward()
    assert abs(L.item() - loss) < 1e-12 * max(1.0, abs(loss))
    ref = pt.grad.numpy()
    assert np.abs(g - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())
