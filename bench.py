#!/usr/bin/env python
"""Benchmark of the structural hot path: dip3d + somf3d on one MI355X.

One step = cold-start dip3d (niter=5, liter=10, order=2, rect=5x5x5) followed
by somf3d (r1=r2=4, order=2, median of 2*r1*r2+1 sprayed slots) over a
device-resident FP32 512 x 256 x 256 synthetic volume (four curved events,
slopes within +-3 samples/trace, Gaussian noise 0.3 and 1% spikes of
amplitude 5).  Inputs are resident in HBM before the timed region; the timed
region covers exactly --steps steps between two synchronisations.

Prints one JSON line (metric, value in samples/s, roofline of the dominant
kernel from live HIP-event timing, CPU baseline = the reference's own C
extension on a bounded sample of the same workload, one host core).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "samples/s dip3d+somf3d at fixed dip RMSE, warm vs cold start; % HBM peak"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shape", type=str, default="512,256,256")
    ap.add_argument("--cpu-shape", type=str, default="128,48,48")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-rmse", action="store_true", help="skip the converged-dip RMSE check")
    return ap.parse_args()


def cpu_baseline(shape):
    """The reference C extension (oracle/_ref) on a bounded sample, one core;
    falls back to the FP32 restatement (kind 'port') where it is not built."""
    import numpy as np

    from pyseistr_amd import synth
    v, _ = synth.volume3d(shape, seed=5, n_events=4, smax=3.0, sigma=0.3, spike_frac=0.01)
    n = int(np.prod(shape))
    try:
        from oracle import ref
        if not ref.available():
            raise ImportError
        kind = "reference"
        t = time.perf_counter()
        d = ref.dipc(v, shape, niter=5, liter=10, order=2, rect=(5, 5, 5))
        ref.somf3d(v, d[:n], d[n:], shape, 4, 4, order=2)
        dt = time.perf_counter() - t
    except ImportError:
        from oracle import oracle as O
        kind = "port"
        t = time.perf_counter()
        d = O.dip(v, shape, niter=5, liter=10, order=2, rect=(5, 5, 5))
        O.somf3d(v, d[:n], d[n:], shape, 4, 4, order=2)
        dt = time.perf_counter() - t
    return {"value": n / dt, "unit": "samples/s", "cores": 1, "kind": kind,
            "sample": f"dip3d(niter=5,liter=10,rect=5)+somf3d(4,4) on a {shape[0]}x{shape[1]}x{shape[2]} "
                      f"volume of the same generator, {dt:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only, see DESIGN.md)", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    from pyseistr_amd import _lib, dip, filters, synth

    shape = tuple(int(s) for s in args.shape.split(","))
    n = int(np.prod(shape))
    vol, _ = synth.volume3d(shape, seed=5, n_events=4, smax=3.0, sigma=0.3, spike_frac=0.01)
    d = torch.from_numpy(vol.ravel(order="F")).cuda()
    dips = torch.empty(2 * n, dtype=torch.float32, device="cuda")
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    dip_kw = dict(niter=5, liter=10, order=2, rect=(5, 5, 5))

    def step():
        dip.dip_device(d, shape, out=dips, **dip_kw)
        filters.somf3d_device(d, dips[:n], dips[n:], shape, 4, 4, order=2, out=out)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = _lib.profile_report()
    _lib.profile_enable(False)
    ms_step = ev0.elapsed_time(ev1) / args.steps
    value = n / (ms_step / 1e3)

    # dominant kernel by total time; achieved = algorithmic bytes / time
    name, (cnt, ms, by) = max(prof.items(), key=lambda kv: kv[1][1])
    achieved = by / (ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "launches": int(cnt), "avg_us": round(ms / cnt * 1e3, 2),
                "bytes_per_launch": by / cnt}
    top = sorted(prof.items(), key=lambda kv: -kv[1][1])[:6]
    breakdown = {k: {"ms_per_step": round(v[1] / args.steps, 3),
                     "GB/s": round(v[2] / (v[1] / 1e3) / 1e9, 1) if v[1] > 0 else None} for k, v in top}

    rmse = None
    if not args.no_rmse:
        ref = dip.dip_device(d, shape, niter=10, liter=10, order=2, rect=(5, 5, 5))
        rmse = float(torch.sqrt(torch.mean((dips - ref) ** 2)).item())

    cpu = None if args.no_cpu else cpu_baseline(tuple(int(s) for s in args.cpu_shape.split(",")))
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded curved Ricker events + noise + spikes)",
        "config": {"workload": "dip3d(cold start, niter=5, liter=10, order=2, rect=5,5,5) + somf3d(r1=r2=4, "
                               "order=2, median of 33 slots)",
                   "shape": list(shape), "samples_per_step": n, "start": "cold",
                   "dip_rmse_vs_niter10": rmse, "wall_s": round(wall, 3)},
        "roofline": roofline, "cpu_baseline": cpu, "breakdown": breakdown,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
