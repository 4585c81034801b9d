"""Seeded synthetic seismic data for the benchmark workloads and tests.

Events are Ricker wavelets (1 - 2b) exp(-b), b = (pi f (t - t_k))^2, the same
pulse as pyseistr/ricker.py:21-36, placed along curved traveltime surfaces
t_k(x, y) = t0 + a x + b y + c (x - xk)^2 + d (y - yk)^2 (samples), with the
slope bounds of each workload.  Arrays are float32, shape (n1, n2[, n3]).
"""
import numpy as np


def _ricker(tau, f, dt):
    b = (np.pi * f * tau * dt) ** 2
    return (1.0 - 2.0 * b) * np.exp(-b)


def _events(shape, n_events, freqs, smax, rng, amp=(0.5, 1.0), dt=0.004, curved=True, rec=None):
    """Sum of curved Ricker events; rec (a list) receives each event's
    (t0, a, b, c, d, xk, yk, f, amplitude)."""
    n1 = shape[0]
    n2 = shape[1]
    n3 = shape[2] if len(shape) > 2 else 1
    t = np.arange(n1, dtype=np.float64)[:, None, None]
    x = np.arange(n2, dtype=np.float64)[None, :, None]
    y = np.arange(n3, dtype=np.float64)[None, None, :]
    out = np.zeros((n1, n2, n3), np.float64)
    for k in range(n_events):
        f = freqs[k % len(freqs)] if isinstance(freqs, (list, tuple)) else freqs
        a = rng.uniform(-0.5, 0.5) * smax
        b = rng.uniform(-0.5, 0.5) * smax if n3 > 1 else 0.0
        xk, yk = rng.uniform(0, n2), rng.uniform(0, n3)
        c = d = 0.0
        if curved:
            c = rng.uniform(-1, 1) * smax / (4.0 * max(n2, 2))
            d = rng.uniform(-1, 1) * smax / (4.0 * max(n3, 2)) if n3 > 1 else 0.0
        t0 = rng.uniform(0.15, 0.85) * n1
        tt = t0 + a * (x - n2 / 2) + b * (y - n3 / 2) + c * (x - xk) ** 2 + d * (y - yk) ** 2
        am = rng.uniform(*amp)
        out += am * _ricker(t - tt, f, dt)
        if rec is not None:
            rec.append((t0, a, b, c, d, xk, yk, f, am))
    return out


def gather2d(n1=512, n2=256, seed=1, sigma=0.2, dt=0.004, dx=10.0):
    """Gather with one linear and two hyperbolic events, 30 Hz Ricker, noise
    sigma * max|clean| (the 512 x 256 2-D dip + somean workload)."""
    rng = np.random.default_rng(seed)
    t = np.arange(n1)[:, None] * dt
    x = np.arange(n2)[None, :] * dx
    d = np.zeros((n1, n2))
    d += 1.0 * _ricker((t - (0.5 + 0.8 * dt * np.arange(n2)[None, :])) / dt, 30.0, dt)
    for t0, v, xa, amp in ((0.9, 2000.0, 128 * dx, 0.7), (1.4, 1500.0, 64 * dx, 0.5)):
        th = np.sqrt(t0 ** 2 + ((x - xa) / v) ** 2)
        d += amp * _ricker((t - th) / dt, 30.0, dt)
    d /= np.abs(d).max()
    d += sigma * rng.standard_normal(d.shape)
    return d.astype(np.float32)


def volume3d(shape=(256, 128, 128), seed=2, n_events=4, freqs=(25.0, 28.0, 32.0, 35.0), smax=2.0, sigma=0.3,
             spike_frac=0.01, spike_amp=5.0, dt=0.004):
    """Curved dipping events, Gaussian noise and impulsive spikes (the 3-D
    dip3d + somf3d workloads)."""
    rng = np.random.default_rng(seed)
    clean = _events(shape, n_events, list(freqs), smax, rng, dt=dt)
    clean /= max(np.abs(clean).max(), 1e-12)
    noisy = clean + sigma * rng.standard_normal(clean.shape)
    if spike_frac > 0:
        m = rng.random(clean.shape) < spike_frac
        noisy[m] += spike_amp * rng.choice([-1.0, 1.0], size=int(m.sum()))
    return noisy.astype(np.float32), clean.astype(np.float32)


def true_slopes3d(shape=(256, 128, 128), seed=2, n_events=4, freqs=(25.0, 28.0, 32.0, 35.0), smax=2.0, lobe=1.5,
                  dt=0.004):
    """The generator's own slopes of a volume3d call with the same
    arguments: (inline, crossline, support).  Slopes are dt_k/dx and
    dt_k/dy of the event whose main lobe covers the sample (samples per
    trace, the sign convention of dip3d); support marks samples within
    `lobe` samples of exactly one event's traveltime surface (crossings and
    event-free samples, where the slope is not defined by the data, are
    left out)."""
    rng = np.random.default_rng(seed)
    rec = []
    _events(shape, n_events, list(freqs), smax, rng, dt=dt, rec=rec)
    n1, n2, n3 = shape
    t = np.arange(n1, dtype=np.float64)[:, None, None]
    x = np.arange(n2, dtype=np.float64)[None, :, None]
    y = np.arange(n3, dtype=np.float64)[None, None, :]
    si = np.zeros(shape, np.float32)
    sx = np.zeros(shape, np.float32)
    hits = np.zeros(shape, np.int8)
    for (t0, a, b, c, d, xk, yk, f, am) in rec:
        tt = t0 + a * (x - n2 / 2) + b * (y - n3 / 2) + c * (x - xk) ** 2 + d * (y - yk) ** 2
        near = np.abs(t - tt) <= lobe
        gi = np.broadcast_to(a + 2 * c * (x - xk), shape)
        gx = np.broadcast_to(b + 2 * d * (y - yk), shape)
        si = np.where(near, gi, si).astype(np.float32)
        sx = np.where(near, gx, sx).astype(np.float32)
        hits += near.astype(np.int8)
    return si, sx, hits == 1


def interp3d(shape=(300, 100, 100), seed=3, mask_seed=33, keep=0.3, n_events=3, smax=1.5):
    """Clean curved events and a Bernoulli trace mask (the soint/sint
    workload); returns (observed, mask, complete)."""
    rng = np.random.default_rng(seed)
    clean = _events(shape, n_events, [30.0], smax, rng)
    clean = (clean / max(np.abs(clean).max(), 1e-12)).astype(np.float32)
    mrng = np.random.default_rng(mask_seed)
    tr = (mrng.random(shape[1:]) < keep).astype(np.float32)
    mask = np.broadcast_to(tr[None], shape).astype(np.float32)
    return (clean * mask).astype(np.float32), mask, clean


def patches2d(count=2048, seed=4, n1s=(128, 192, 256, 384, 512), n2s=(64, 96, 128, 192, 256), sigma=0.5,
              spike_frac=0.02, smax=2.0):
    """Uneven DAS-style patches: list of float32 arrays (n1, n2)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n1 = int(rng.choice(n1s)); n2 = int(rng.choice(n2s))
        d = _events((n1, n2), int(rng.integers(1, 4)), [30.0], smax, rng)[:, :, 0]
        d /= max(np.abs(d).max(), 1e-12)
        d += sigma * rng.standard_normal(d.shape)
        bad = rng.random(n2) < spike_frac
        d[:, bad] += rng.uniform(-1, 1, size=(n1, int(bad.sum()))) * 10
        out.append(d.astype(np.float32))
    return out


def snr(g, f):
    """20 log10(|g| / |g - f|) (pyseistr/snr.py:60)."""
    g = np.asarray(g, np.float64); f = np.asarray(f, np.float64)
    return 20.0 * np.log10(np.linalg.norm(g) / np.linalg.norm(g - f))
