"""Learned warm start for dip3d: a small 3-D CNN predicts the initial slope
field so that the Gauss-Newton plane-wave-destruction loop (the reference's
dipc, pyseistr/src/dip_cfuns.c:1694) needs fewer iterations.

The network runs entirely in libseistr_hip.so (sh_cnn_forward / sh_cnn_grad /
sh_cnn_adam: BF16 3x3x3 convolutions on the matrix cores, FP32 accumulation);
torch only holds the device buffers.  The plain cold-start dip3d stays the
fallback and the measure of the result: the prediction is only an ``init``
for sh_dip, whose iterations still decide the answer.

Training (``python -m pyseistr_amd.warmstart train``): seeded synthetic
volumes with 2-5 curved Ricker events, slopes within +-3 samples/trace and
noise sigma in [0, 0.5]; targets are the converged cold-start dips that the
library's own dip3d produces on each volume; 64^3 windows are cut from them
and the network is fit by Adam on the weighted mean-square dip error.
"""
import argparse
import ctypes
import json
import os
import time

import numpy as np

from . import _lib
from .dip import dip_device

NT, C = 27, 32
NPARAM = NT * C + C + 4 * (NT * C * C + C) + NT * C * 2 + 2
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warmstart.npz")


def init_params(seed=0):
    """He-normal weights, zero biases, in the library's flat layout."""
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal(NT * C) * np.sqrt(2.0 / NT), np.zeros(C)]
    for _ in range(4):
        parts += [rng.standard_normal(NT * C * C) * np.sqrt(2.0 / (NT * C)), np.zeros(C)]
    parts += [rng.standard_normal(NT * C * 2) * np.sqrt(1.0 / (NT * C)), np.zeros(2)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == NPARAM
    return p


def input_scale(d):
    """1 / RMS of the volume (the network sees unit-RMS data)."""
    rms = float(d.double().pow(2).mean().sqrt().item())
    return 1.0 / rms if rms > 0 else 1.0


def _dil(dil):
    if dil is None:
        return None
    d = (ctypes.c_int * 6)(*[int(v) for v in dil])
    return ctypes.cast(d, _lib.c_int_p)


def cnn_forward(params, d, shape, xscale, out=None, dil=None):
    """Device call: dips [2N] (inline || crossline) for the flat volume d;
    dil: per-layer dilations (None = all 1)."""
    n1, n2, n3 = (int(s) for s in shape)
    n = n1 * n2 * n3
    _lib.check_sizes("cnn_forward", n, data=d)
    _lib.check_sizes("cnn_forward", NPARAM, params=params)
    _lib.check_out(out, 2 * n, "cnn_forward")
    ctx = _lib.context()
    if out is None:
        out = _lib.empty(2 * n)
    rc = _lib.load().sh_cnn_forward(ctx, _lib.ptr(params), _lib.ptr(d), float(xscale), _dil(dil), n1, n2, n3,
                                    _lib.ptr(out))
    _lib.check(ctx, rc, "cnn_forward")
    return out


def cnn_grad(params, d, shape, xscale, target, wt=None, grad=None, want_loss=True, dil=None):
    """Device call: (loss or None, grad [NPARAM])."""
    n1, n2, n3 = (int(s) for s in shape)
    n = n1 * n2 * n3
    _lib.check_sizes("cnn_grad", n, data=d, wt=wt)
    _lib.check_sizes("cnn_grad", 2 * n, target=target)
    _lib.check_sizes("cnn_grad", NPARAM, params=params)
    _lib.check_out(grad, NPARAM, "cnn_grad")
    ctx = _lib.context()
    if grad is None:
        grad = _lib.empty(NPARAM)
    loss = ctypes.c_double(0.0)
    rc = _lib.load().sh_cnn_grad(ctx, _lib.ptr(params), _lib.ptr(d), float(xscale), _dil(dil), _lib.ptr(target),
                                 _lib.ptr(wt), n1, n2, n3, _lib.ptr(grad), ctypes.byref(loss) if want_loss else None)
    _lib.check(ctx, rc, "cnn_grad")
    return (loss.value if want_loss else None), grad


def cnn_adam(params, grad, m, v, lr, step, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    ctx = _lib.context()
    rc = _lib.load().sh_cnn_adam(ctx, _lib.ptr(params), _lib.ptr(grad), _lib.ptr(m), _lib.ptr(v), float(lr),
                                 float(b1), float(b2), float(eps), int(step), float(wd))
    _lib.check(ctx, rc, "cnn_adam")


class WarmStart:
    """Trained network + the warm-started dip3d."""

    def __init__(self, params=None, meta=None):
        if params is None:
            params = init_params(0)
        params = np.asarray(params, np.float32).ravel()
        if params.size != NPARAM:
            raise ValueError(f"expected {NPARAM} parameters, got {params.size}")
        self.params_host = params
        self.meta = dict(meta or {})
        self.dil = [int(v) for v in self.meta.get("dilation", [1] * 6)]
        self._dev = None

    @classmethod
    def load(cls, path=None):
        """Weights from `path`, else from $SH_WARMSTART_WEIGHTS, else the
        committed pyseistr_amd/weights/warmstart.npz."""
        path = path or os.environ.get("SH_WARMSTART_WEIGHTS") or WEIGHTS
        with np.load(path, allow_pickle=False) as z:
            meta = json.loads(str(z["meta"])) if "meta" in z else {}
            return cls(z["params"], meta)

    def save(self, path=WEIGHTS):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.params_host, meta=np.array(json.dumps(self.meta)))

    @property
    def params(self):
        if self._dev is None:
            import torch
            self._dev = torch.from_numpy(self.params_host).cuda()
        return self._dev

    def predict_device(self, d, shape, out=None, xscale=None):
        """Initial dips [2N] for the flat Fortran-order device volume d
        (xscale None: the library scales d by 1 / RMS on the device)."""
        return cnn_forward(self.params, d, shape, 0.0 if xscale is None else xscale, out, dil=self.dil)

    def dip_device(self, d, shape, niter=2, liter=10, order=2, rect=(5, 5, 5), out=None, init=None, report=False):
        """Warm-started dip3d: CNN prediction, then niter GN iterations."""
        if init is None:
            init = self.predict_device(d, shape)
        return dip_device(d, shape, niter=niter, liter=liter, order=order, rect=rect, init=init, out=out,
                          report=report)

    def dip3dc(self, din, niter=2, liter=10, order=2, rect=(5, 5, 5)):
        """Host-array convenience with dip3dc's output convention."""
        din = np.asarray(din)
        shape = din.shape
        d = _lib.to_device(din)
        res = _lib.to_host(self.dip_device(d, shape, niter, liter, order, rect))
        n = int(np.prod(shape))
        return res[:n].reshape(shape, order="F"), res[n:].reshape(shape, order="F")


# --------------------------------------------------------------------------
# Training
# --------------------------------------------------------------------------
def training_volume(seed, shape=(256, 128, 128)):
    """Seeded volume: 2-5 curved events (25-35 Hz), slopes within +-3,
    noise sigma in [0, 0.5], spikes (1%, amplitude 5) in half of them."""
    from . import synth
    rng = np.random.default_rng(seed)
    ne = int(rng.integers(2, 6))
    sigma = float(rng.uniform(0.0, 0.5))
    spikes = 0.01 if rng.random() < 0.5 else 0.0
    v, _ = synth.volume3d(shape, seed=int(rng.integers(1 << 31)), n_events=ne, smax=3.0, sigma=sigma,
                          spike_frac=spikes)
    return v


def _gen(args):
    seed, shape = args
    return training_volume(seed, shape)


def make_dataset(seeds, vol_shape, win, per_vol, target_niter, liter=10, rect=(5, 5, 5), pool=None, log=print,
                 threads=1):
    """Cut 64^3 windows (x, target dips) from seeded volumes; everything
    stays on the device.  Returns list of (x[W], target[2W], xscale), in
    seed order.  Window positions follow from each volume's seed alone.
    pool: optional multiprocessing pool for the volume generator, created
    before the process first touches the GPU.  threads > 1 computes the
    targets of that many volumes at once, each thread on a stream and
    library context of its own (small volumes leave the GPU idle between
    one volume's launches otherwise)."""
    import queue
    import threading

    import torch
    w1, w2, w3 = win
    wn = w1 * w2 * w3
    n = int(np.prod(vol_shape))
    t0 = t_log = time.time()
    jobs = [(s, vol_shape) for s in seeds]
    gen = pool.imap(_gen, jobs, chunksize=2) if pool is not None else map(_gen, jobs)
    per_seed = [None] * len(seeds)
    lock = threading.Lock()
    done = [0]
    errors = []
    todo = queue.Queue(maxsize=4 * max(threads, 1))

    def windows(k, v):
        d = torch.from_numpy(v.ravel(order="F")).cuda(non_blocking=False)
        dips = dip_device(d, vol_shape, niter=target_niter, liter=liter, order=2, rect=rect)
        xs = input_scale(d)
        vol = d.view(vol_shape[2], vol_shape[1], vol_shape[0])
        di = dips[:n].view(vol_shape[2], vol_shape[1], vol_shape[0])
        dx = dips[n:].view(vol_shape[2], vol_shape[1], vol_shape[0])
        rng = np.random.default_rng((12345, int(seeds[k])))
        out = []
        for _ in range(per_vol):
            o1 = int(rng.integers(0, vol_shape[0] - w1 + 1))
            o2 = int(rng.integers(0, vol_shape[1] - w2 + 1))
            o3 = int(rng.integers(0, vol_shape[2] - w3 + 1))
            sl = (slice(o3, o3 + w3), slice(o2, o2 + w2), slice(o1, o1 + w1))
            x = vol[sl].contiguous().view(-1)
            t = torch.cat([di[sl].contiguous().view(-1), dx[sl].contiguous().view(-1)])
            assert x.numel() == wn
            out.append((x, t, xs))
        return out

    def worker():
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            while True:
                item = todo.get()
                if item is None:
                    break
                k, v = item
                try:
                    res = windows(k, v)
                except Exception as e:  # reported by the caller
                    with lock:
                        errors.append(e)
                    continue
                with lock:
                    per_seed[k] = res
                    done[0] += 1
        stream.synchronize()

    pool_t = [threading.Thread(target=worker, daemon=True) for _ in range(max(threads, 1))]
    for th in pool_t:
        th.start()
    for k, v in enumerate(gen):
        if errors:
            break
        todo.put((k, v))
        if time.time() - t_log > 30 or k + 1 == len(seeds):
            t_log = time.time()
            with lock:
                nd = done[0]
            log(f"dataset: {nd}/{len(seeds)} volumes, {nd * per_vol} windows, {time.time() - t0:.0f} s")
    for _ in pool_t:
        todo.put(None)
    for th in pool_t:
        th.join()
    torch.cuda.synchronize()
    if errors:
        raise errors[0]
    log(f"dataset: {done[0]}/{len(seeds)} volumes, {time.time() - t0:.0f} s")
    return [w for res in per_seed for w in res]


def window_weights(win, margin):
    """1 inside, 0 within `margin` voxels of a window face (those voxels see
    zero padding that the full-volume inference does not)."""
    import torch
    w1, w2, w3 = win
    m = np.zeros((w3, w2, w1), np.float32)
    m[margin:w3 - margin, margin:w2 - margin, margin:w1 - margin] = 1.0
    return torch.from_numpy(m.ravel()).cuda()


def train(data, steps, lr=1e-3, lr_end=1e-5, seed=0, win=(64, 64, 64), margin=4, log=print, log_every=100,
          params=None, dil=None):
    """Adam over random windows; cosine learning-rate decay.  Returns the
    FP32 parameters (host) and the loss history."""
    import torch
    p = torch.from_numpy(init_params(seed) if params is None else np.asarray(params, np.float32)).cuda()
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    g = torch.empty_like(p)
    wt = window_weights(win, margin)
    rng = np.random.default_rng(seed + 1)
    hist = []
    t0 = time.time()
    run = 0.0
    for s in range(1, steps + 1):
        x, t, xs = data[int(rng.integers(len(data)))]
        want = s % 10 == 0 or s == 1
        loss, _ = cnn_grad(p, x, win, xs, t, wt, grad=g, want_loss=want, dil=dil)
        a = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * (s - 1) / max(steps - 1, 1)))
        cnn_adam(p, g, m, v, a, s)
        if want:
            run = loss if s == 1 else 0.9 * run + 0.1 * loss
            hist.append((s, loss))
        if s % log_every == 0:
            log(f"step {s}/{steps} loss {run:.5f} lr {a:.2e} {time.time() - t0:.0f} s")
    torch.cuda.synchronize()
    return p.cpu().numpy(), hist


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the dip3d warm-start network on the GPU")
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--volumes", type=int, default=500)
    ap.add_argument("--vol-shape", type=str, default="256,128,128")
    ap.add_argument("--per-vol", type=int, default=8)
    ap.add_argument("--target-niter", type=int, default=20)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed0", type=int, default=10000)
    ap.add_argument("--out", type=str, default=WEIGHTS)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--threads", type=int, default=1, help="volumes whose targets are computed at once")
    ap.add_argument("--minutes", type=float, default=None,
                    help="wall-clock budget of the whole run: fewer steps if the dataset took long "
                         "(the learning-rate schedule spans the steps that fit)")
    ap.add_argument("--step-ms", type=float, default=3.2, help="expected time per step, for --minutes")
    ap.add_argument("--dilation", type=str, default="1,1,1,1,1,1", help="per-layer dilations")
    ap.add_argument("--window", type=int, default=64, help="cube edge of the training windows")
    ap.add_argument("--margin", type=int, default=4, help="voxels at each window face left out of the loss")
    ap.add_argument("--init", type=str, default=None,
                    help="continue from these weights (npz) instead of a fresh initialisation")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    from multiprocessing import Pool
    pool = Pool(a.workers)  # forked before the GPU is touched
    vol_shape = tuple(int(s) for s in a.vol_shape.split(","))
    seeds = list(range(a.seed0, a.seed0 + a.volumes))
    t0 = time.time()
    win = (a.window,) * 3
    data = make_dataset(seeds, vol_shape, win, a.per_vol, a.target_niter, pool=pool, log=log, threads=a.threads)
    pool.close()
    log(f"dataset ready: {len(data)} windows in {time.time() - t0:.0f} s")
    dil = [int(v) for v in a.dilation.split(",")]
    steps = a.steps
    if a.minutes is not None:
        left = a.minutes * 60 - (time.time() - t0) - 30
        steps = max(100, min(steps, int(left / (a.step_ms / 1e3))))
        log(f"{steps} steps fit the remaining {left:.0f} s")
    prior = WarmStart.load(a.init) if a.init else None
    if prior is not None and prior.dil != dil:
        raise SystemExit("--init weights have dilations %s, not %s" % (prior.dil, dil))
    params, hist = train(data, steps, lr=a.lr, log=log, dil=dil, win=win, margin=a.margin,
                         params=prior.params_host if prior is not None else None)
    meta = {"volumes": a.volumes, "vol_shape": vol_shape, "seeds": [a.seed0, a.seed0 + a.volumes - 1],
            "windows": len(data), "per_vol": a.per_vol, "target_niter": a.target_niter, "steps": steps,
            "lr": a.lr, "dilation": dil, "window": a.window, "margin": a.margin,
            "final_loss": hist[-1][1] if hist else None,
            "loss_history": hist[::max(1, len(hist) // 50)]}
    if prior is not None:  # the training this run continued
        meta["continued_from"] = {k: v for k, v in prior.meta.items() if k != "loss_history"}
    ws = WarmStart(params, meta)
    ws.save(a.out)
    log(f"saved {a.out}; total {time.time() - t0:.0f} s")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
