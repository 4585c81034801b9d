"""Batched 2-D path: dip2d + somf2d over thousands of uneven patches.

Every patch goes through the same algorithm as the per-patch drop-ins
(dip2dc = dipc with n3 = 1, pyseistr/src/dip_cfuns.c:1694; somf2dc =
csomf2d, pyseistr/src/sof_cfuns.c:1534), but each pipeline stage is ONE
launch over the whole batch (sh_dip2d_batched / sh_somf2d_batched): a work
list of blocks, per-patch CG and line-search state on the device, and
per-patch reductions in a fixed order.

The smoothing tile shape of the batched dip is picked by ``autotune``: an
offline sweep of the tile shapes the library offers, timed per (n1, n2)
bucket, stored as JSON next to this module; ``pick_tile`` chooses for a
batch the shape with the least predicted time over its bucket mix.
"""
import ctypes
import json
import os
import time

import numpy as np

from . import _lib

TUNING = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "batch2d.json")


def tile_shapes():
    lib = _lib.load()
    out = []
    for c in range(lib.sh_batch_tile_count()):
        a, b = ctypes.c_int(), ctypes.c_int()
        lib.sh_batch_tile_shape(c, ctypes.byref(a), ctypes.byref(b))
        out.append((a.value, b.value))
    return out


class Batch:
    """Patches packed back to back on the device (each F-order)."""

    def __init__(self, patches=None, shapes=None, flat=None):
        import torch
        if patches is not None:
            shapes = [tuple(int(s) for s in np.shape(p)) for p in patches]
        self.shapes = [(int(a), int(b)) for a, b in shapes]
        sizes = np.array([a * b for a, b in self.shapes], np.int64)
        self.offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.n1s = np.array([a for a, _ in self.shapes], np.int32)
        self.n2s = np.array([b for _, b in self.shapes], np.int32)
        self.ntot = int(sizes.sum())
        if patches is not None:
            host = np.concatenate([np.asarray(p, np.float32).ravel(order="F") for p in patches])
            self.data = torch.from_numpy(host).cuda()
        else:
            self.data = flat
            if flat is not None and flat.numel() != self.ntot:
                raise ValueError("flat data does not match the patch shapes")

    def __len__(self):
        return len(self.shapes)

    def _tables(self):
        return (self.offs.ctypes.data_as(ctypes.c_void_p), self.n1s.ctypes.data_as(ctypes.c_void_p),
                self.n2s.ctypes.data_as(ctypes.c_void_p))

    def split(self, flat):
        """Host list of (n1, n2) arrays from a flat device result."""
        h = flat.cpu().numpy()
        return [h[o:o + a * b].reshape(a, b, order="F") for o, (a, b) in zip(self.offs, self.shapes)]


def dip2d_batched(batch, niter=5, liter=20, order=2, rect=(10, 10, 1), tile=None, init=None, out=None,
                  report=False):
    """Device call: dips of every patch, flat like batch.data."""
    _lib.check_sizes("dip2d_batched", batch.ntot, data=batch.data, init=init)
    _lib.check_out(out, batch.ntot, "dip2d_batched")
    ctx = _lib.context()
    if out is None:
        out = _lib.empty(batch.ntot)
    if tile is None:
        tile = pick_tile(batch.shapes)
    rec = None
    if report and niter > 0:
        rec = np.zeros(len(batch) * 2 * niter, np.int32)
    o, a, b = batch._tables()
    rc = _lib.load().sh_dip2d_batched(ctx, _lib.ptr(batch.data), o, a, b, len(batch), int(niter), int(liter),
                                      int(order), int(rect[0]), int(rect[1]), int(tile), _lib.ptr(init),
                                      _lib.ptr(out), rec.ctypes.data_as(_lib.c_int_p) if rec is not None else None)
    _lib.check(ctx, rc, "dip2d_batched")
    if report:
        r = rec.reshape(len(batch), niter, 2) if rec is not None else None
        return out, r
    return out


def somf2d_batched(batch, dip, ns=6, order=2, eps=0.01, option=1, nmf=None, out=None):
    """Device call: structure-oriented median of every patch (nmf = 2ns+1,
    as somf2dc passes it, pyseistr/somf2d.py:96)."""
    _lib.check_sizes("somf2d_batched", batch.ntot, data=batch.data, dip=dip)
    _lib.check_out(out, batch.ntot, "somf2d_batched")
    ctx = _lib.context()
    if out is None:
        out = _lib.empty(batch.ntot)
    if nmf is None:
        nmf = 2 * ns + 1
    o, a, b = batch._tables()
    rc = _lib.load().sh_somf2d_batched(ctx, _lib.ptr(batch.data), _lib.ptr(dip), o, a, b, len(batch), int(ns),
                                       int(nmf), int(option), int(order), float(eps), _lib.ptr(out))
    _lib.check(ctx, rc, "somf2d_batched")
    return out


# --------------------------------------------------------------------------
# Autotuner
# --------------------------------------------------------------------------
def bucket(n1, n2):
    return f"{n1}x{n2}"


def load_tuning(path=TUNING):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def default_tile():
    """Index of the library's default tile shape."""
    a, b = ctypes.c_int(), ctypes.c_int()
    _lib.load().sh_batch_tile_shape(-1, ctypes.byref(a), ctypes.byref(b))
    return tile_shapes().index((a.value, b.value))


def pick_tile(shapes, tuning=None, min_gain=0.10):
    """Tile index with the least predicted time for this bucket mix; the
    library default unless the table predicts at least min_gain less time
    than the default for the mix, and when no tuning table is present.
    The table times the dip alone (niter=2) per bucket; in the whole
    batched pipeline (dip niter=5 + somf2d) its predictions do not carry
    over below ~10 %: for 1024 patches of 128 x 64 it predicts 9 % less
    time for 64 x 32 tiles, and the pipeline ran 3.5 % longer with them
    than with the default 64 x 16."""
    if tuning is None:
        tuning = load_tuning()
    if not tuning:
        return -1
    table = tuning["buckets"]
    ntile = len(tuning["tiles"])
    cost = np.zeros(ntile)
    for n1, n2 in shapes:
        key = bucket(n1, n2)
        if key in table:
            cost += np.asarray(table[key]["ns_per_sample"]) * n1 * n2
        else:  # nearest tuned bucket by size
            k = min(table, key=lambda s: abs(np.log(int(s.split("x")[0]) * int(s.split("x")[1]) / (n1 * n2))))
            cost += np.asarray(table[k]["ns_per_sample"]) * n1 * n2
    best = int(np.argmin(cost))
    dflt = tuning["tiles"].index(list(tile_shapes()[default_tile()]))
    return best if cost[best] <= (1.0 - min_gain) * cost[dflt] else dflt


def _time_tiles(b, tiles, niter, liter, repeats):
    """Best-of-repeats ns/sample of the batched dip for every tile shape."""
    import torch
    out = _lib.empty(b.ntot)
    times = []
    for t in range(len(tiles)):
        dip2d_batched(b, niter=1, liter=2, tile=t, out=out)  # warm up
        torch.cuda.synchronize()
        best = None
        for _ in range(repeats):
            t0 = time.perf_counter()
            dip2d_batched(b, niter=niter, liter=liter, tile=t, out=out)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        times.append(best / b.ntot * 1e9)
    return times


def autotune(n1s=(128, 192, 256, 384, 512), n2s=(64, 96, 128, 192, 256), per_bucket=256, niter=2, liter=20,
             repeats=2, path=TUNING, seed=7, mix=2048, log=print):
    """Time every tile shape on a synthetic batch of each (n1, n2) bucket
    (per_bucket patches, enough to fill the GPU so the times measure the
    tiles rather than launch overhead) and store ns/sample per tile; with
    mix > 0 also on the uneven mix of that many patches (synth.patches2d,
    seed 4) at dip2d's default niter=5, as a check of the bucket model.
    Returns the table."""
    import torch

    from . import synth
    tiles = tile_shapes()
    table = {}
    rng = np.random.default_rng(seed)
    for n1 in n1s:
        for n2 in n2s:
            pats = []
            for _ in range(per_bucket):
                d = synth._events((n1, n2), int(rng.integers(1, 4)), [30.0], 2.0, rng)[:, :, 0]
                d = d / max(np.abs(d).max(), 1e-12) + 0.5 * rng.standard_normal((n1, n2))
                pats.append(d.astype(np.float32))
            times = _time_tiles(Batch(pats), tiles, niter, liter, repeats)
            table[bucket(n1, n2)] = {"ns_per_sample": [round(x, 4) for x in times], "best": int(np.argmin(times))}
            log(f"autotune {n1}x{n2}: best tile {tiles[int(np.argmin(times))]} "
                f"{min(times):.3f} ns/sample (worst {max(times):.3f})")
    res = {"tiles": tiles, "niter": niter, "liter": liter, "per_bucket": per_bucket, "buckets": table,
           "device": torch.cuda.get_device_name(0)}
    if mix:
        pats = synth.patches2d(count=mix, seed=4)
        b = Batch(pats)
        times = _time_tiles(b, tiles, 5, 20, repeats)
        res["mix"] = {"patches": mix, "niter": 5, "ns_per_sample": [round(x, 4) for x in times],
                      "best": int(np.argmin(times)), "bucket_model_pick": pick_tile(b.shapes, res)}
        log(f"autotune mix of {mix}: best tile {tiles[int(np.argmin(times))]}, bucket model picks "
            f"{tiles[res['mix']['bucket_model_pick']]}")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    return res


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="autotune the batched 2-D dip tiles on this GPU")
    ap.add_argument("cmd", choices=["autotune"])
    ap.add_argument("--per-bucket", type=int, default=256)
    ap.add_argument("--mix", type=int, default=2048, help="patches of the uneven mix timed as a check (0: none)")
    ap.add_argument("--out", default=TUNING)
    a = ap.parse_args(argv)
    autotune(per_bucket=a.per_bucket, mix=a.mix, path=a.out, log=lambda m: print(m, flush=True))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
