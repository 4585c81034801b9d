"""CPU oracle of the learned warm-start network (test infrastructure only).

The network has no counterpart in the reference: it only produces the
initial dip that dip3d's Gauss-Newton loop starts from (the reference keeps
that hook, idip/xdip, commented out at pyseistr/src/dip_cfuns.c:1861-1875).
Only the test suite and the smoke check import this module.

Six 3x3x3 convolutions (1 -> 32 -> 32 -> 32 -> 32 -> 32 -> 2 channels),
ReLU after the first five, zero padding at the faces.  Voxels are in
Fortran order (i1 fastest); arrays here are (n3, n2, n1, channels).

Flat parameter vector (float32):
  W1[27][32] b1[32] | {W[27][32 ci][32 co] b[32]} x 4 (layers 2..5) |
  W6[27][32 ci][2 co] b6[2]
Tap t covers the offset dil * (t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1) along
(i1, i2, i3), dil the layer's dilation: out[v] = sum_t in[v + off(t)] W[t]
(a correlation).

``quant=True`` mirrors the device arithmetic's rounding points: the
activations of layers 1..5, the weights of layers 2..6 and the
back-propagated gradients are rounded to bfloat16 (round to nearest even);
everything else is float64 (the device accumulates in float32).
``quant=False`` is the exact float64 network.
"""
import numpy as np

C = 32
NT = 27
_L = NT * C * C + C
OFF_W1 = 0
OFF_B1 = NT * C
OFF_L = {l: OFF_B1 + C + (l - 2) * _L for l in (2, 3, 4, 5)}
OFF_W6 = OFF_B1 + C + 4 * _L
OFF_B6 = OFF_W6 + NT * C * 2
NPARAM = OFF_B6 + 2


def offsets(dil=1):
    return [(dil * (t % 3 - 1), dil * ((t // 3) % 3 - 1), dil * (t // 9 - 1)) for t in range(NT)]


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    W = {1: p[OFF_W1:OFF_B1].reshape(NT, 1, C)}
    b = {1: p[OFF_B1:OFF_B1 + C]}
    for l in (2, 3, 4, 5):
        o = OFF_L[l]
        W[l] = p[o:o + NT * C * C].reshape(NT, C, C)
        b[l] = p[o + NT * C * C:o + _L]
    W[6] = p[OFF_W6:OFF_B6].reshape(NT, C, 2)
    b[6] = p[OFF_B6:OFF_B6 + 2]
    return W, b


def pack(W, b):
    p = np.zeros(NPARAM, np.float64)
    p[OFF_W1:OFF_B1] = W[1].ravel()
    p[OFF_B1:OFF_B1 + C] = b[1]
    for l in (2, 3, 4, 5):
        o = OFF_L[l]
        p[o:o + NT * C * C] = W[l].ravel()
        p[o + NT * C * C:o + _L] = b[l]
    p[OFF_W6:OFF_B6] = W[6].ravel()
    p[OFF_B6:OFF_B6 + 2] = b[6]
    return p


def init_params(seed=0):
    """He-normal weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    W, b = {}, {}
    W[1] = rng.standard_normal((NT, 1, C)) * np.sqrt(2.0 / NT)
    b[1] = np.zeros(C)
    for l in (2, 3, 4, 5):
        W[l] = rng.standard_normal((NT, C, C)) * np.sqrt(2.0 / (NT * C))
        b[l] = np.zeros(C)
    W[6] = rng.standard_normal((NT, C, 2)) * np.sqrt(1.0 / (NT * C))
    b[6] = np.zeros(2)
    return pack(W, b).astype(np.float32)


def _shift(a, d1, d2, d3):
    """a[v + (d1, d2, d3)] with zeros outside; a is (n3, n2, n1, c)."""
    n3, n2, n1 = a.shape[:3]
    h = max(abs(d1), abs(d2), abs(d3), 1)
    p = np.pad(a, ((h, h), (h, h), (h, h), (0, 0)))
    return p[h + d3:h + d3 + n3, h + d2:h + d2 + n2, h + d1:h + d1 + n1]


def conv(a, W, b=None, dil=1):
    """out[v, co] = sum_t sum_ci a[v + off(t), ci] W[t, ci, co] (+ b)."""
    out = np.zeros(a.shape[:3] + (W.shape[2],), np.float64)
    for t, (d1, d2, d3) in enumerate(offsets(dil)):
        out += _shift(a, d1, d2, d3) @ W[t]
    if b is not None:
        out += b
    return out


def conv_wgrad(a, g, dil=1):
    """dW[t, ci, co] = sum_v a[v + off(t), ci] g[v, co]."""
    ci, co = a.shape[3], g.shape[3]
    gf = g.reshape(-1, co)
    return np.stack([_shift(a, d1, d2, d3).reshape(-1, ci).T @ gf for (d1, d2, d3) in offsets(dil)])


def conv_input_grad(g, W, dil=1):
    """dL/da[u, ci] = sum_t sum_co g[u - off(t), co] W[t, ci, co]."""
    out = np.zeros(g.shape[:3] + (W.shape[1],), np.float64)
    for t, (d1, d2, d3) in enumerate(offsets(dil)):
        out += _shift(g, -d1, -d2, -d3) @ W[t].T
    return out


def _vol(x, shape):
    n1, n2, n3 = shape
    return np.asarray(x, np.float64).reshape(n3, n2, n1, 1)


def forward(params, x, shape, xscale=1.0, quant=True, dil=None):
    """Returns (activations [a1..a5], y) with y (n3, n2, n1, 2); dil:
    per-layer dilations (None = all 1)."""
    q = bf16 if quant else (lambda v: v)
    D = [1] * 6 if dil is None else [int(v) for v in dil]
    W, b = unpack(params)
    xs = _vol(x, shape) * np.float64(np.float32(xscale))
    if quant:
        xs = np.asarray(xs, np.float32).astype(np.float64)
    acts = [None]
    a = q(np.maximum(conv(xs, W[1], b[1], D[0]), 0.0))
    acts.append(a)
    for l in (2, 3, 4, 5):
        a = q(np.maximum(conv(a, q(W[l]), b[l], D[l - 1]), 0.0))
        acts.append(a)
    y = conv(a, q(W[6]), b[6], D[5])
    return acts, y


def predict(params, x, shape, xscale=1.0, quant=True, dil=None):
    """Dips as the device returns them: flat [inline N || crossline N]."""
    _, y = forward(params, x, shape, xscale, quant, dil)
    return np.concatenate([y[..., 0].ravel(), y[..., 1].ravel()])


def loss_and_grad(params, x, target, shape, wt=None, xscale=1.0, quant=True, dil=None):
    """loss = sum wt (y - target)^2 / (2N) over both components; returns
    (loss, grad[NPARAM])."""
    q = bf16 if quant else (lambda v: v)
    D = [1] * 6 if dil is None else [int(v) for v in dil]
    n1, n2, n3 = shape
    n = n1 * n2 * n3
    W, b = unpack(params)
    acts, y = forward(params, x, shape, xscale, quant, D)
    t = np.asarray(target, np.float64).reshape(2, n3, n2, n1).transpose(1, 2, 3, 0)
    m = np.ones((n3, n2, n1, 1)) if wt is None else np.asarray(wt, np.float64).reshape(n3, n2, n1, 1)
    if quant:  # the device forms y - t in float32
        d = (np.asarray(y, np.float32) - np.asarray(t, np.float32)).astype(np.float64)
    else:
        d = y - t
    loss = float(np.sum(d * d * m) / (2 * n))
    g = q(2.0 * d * m / (2 * n))
    dW, db = {}, {}
    dW[6] = conv_wgrad(acts[5], g, D[5])
    db[6] = g.reshape(-1, 2).sum(0)
    g = q(conv_input_grad(g, q(W[6]), D[5]) * (acts[5] > 0))
    for l in (5, 4, 3, 2):
        dW[l] = conv_wgrad(acts[l - 1], g, D[l - 1])
        db[l] = g.reshape(-1, C).sum(0)
        g = q(conv_input_grad(g, q(W[l]), D[l - 1]) * (acts[l - 1] > 0))
    xs = _vol(x, shape) * np.float64(np.float32(xscale))
    if quant:
        xs = np.asarray(xs, np.float32).astype(np.float64)
    dW[1] = conv_wgrad(xs, g, D[0])
    db[1] = g.reshape(-1, C).sum(0)
    return loss, pack(dW, db)


def adam(params, grad, m, v, lr, b1, b2, eps, step, wd=0.0):
    """One Adam step (decoupled-free L2 form: g += wd * p), float64."""
    p = np.asarray(params, np.float64)
    g = np.asarray(grad, np.float64) + wd * p
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
