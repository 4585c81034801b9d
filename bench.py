#!/usr/bin/env python
"""Benchmark of the structural hot path: dip3d + somf3d on one MI355X at a
fixed dip error, learned warm start against cold start.

Workload: a device-resident FP32 512 x 256 x 256 synthetic volume (seed 5:
four curved Ricker events, slopes within +-3 samples/trace, Gaussian noise
0.3, 1% spikes of amplitude 5).  dip3d runs with liter=10, order=2,
rect=5x5x5; somf3d with r1=r2=4, order=2 (median of 2*r1*r2+1 = 33 slots).

1. p_conv = the plain cold-start dip3d run for --conv-niter Gauss-Newton
   iterations (its change per iteration is reported).
2. Cold start: the smallest niter whose dips are within --target-rmse
   samples/trace (RMS over both components) of p_conv.
3. Warm start: the trained network's prediction, then the smallest niter
   that reaches the same error.
4. Timed: --steps steps of [warm start: network + dip3d(niter_warm)] +
   somf3d, then the cold path [dip3d(niter_cold) + somf3d] once more, both
   between synchronisations with inputs resident in HBM.

Prints one JSON line: value = samples/s of the warm path at the target
error (the cold path when no trained network is present), the cold path's
rate and the speed-up in config, the roofline of the dominant kernel of the
warm path from live HIP-event timing, and the CPU baseline = the
reference's own C extension on a bounded sample, one host core.
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
BF16_PEAK_TFLOPS = 2500.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shape", type=str, default="512,256,256")
    ap.add_argument("--cpu-shape", type=str, default="128,48,48")
    ap.add_argument("--target-rmse", type=float, default=0.02)
    ap.add_argument("--conv-niter", type=int, default=300)
    ap.add_argument("--max-niter", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--fixed-niter", type=int, default=0,
                    help="skip the error search: time cold dip3d(niter) + somf3d only (profiling runs)")
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
                      f"volume of the same generator (5 GN iterations, not the error-matched count), "
                      f"{dt:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only, see DESIGN.md)", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    from pyseistr_amd import _lib, dip, filters, synth, warmstart

    shape = tuple(int(s) for s in args.shape.split(","))
    n = int(np.prod(shape))
    vol, _ = synth.volume3d(shape, seed=5, n_events=4, smax=3.0, sigma=0.3, spike_frac=0.01)
    d = torch.from_numpy(vol.ravel(order="F")).cuda()
    dips = torch.empty(2 * n, dtype=torch.float32, device="cuda")
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    kw = dict(liter=10, order=2, rect=(5, 5, 5))

    def rmse(a, b):
        return float(torch.sqrt(torch.mean((a - b) ** 2)).item())

    info = {}
    net = None
    if os.path.exists(warmstart.WEIGHTS):
        net = warmstart.WarmStart.load()

    if args.fixed_niter:
        niter_cold, niter_warm = args.fixed_niter, None
    else:
        # 1. converged reference (continuation is exact: the state between
        #    Gauss-Newton iterations is the dip field alone)
        t = time.perf_counter()
        p = dip.dip_device(d, shape, niter=args.conv_niter - 10, **kw)
        prev = p.clone()
        p = dip.dip_device(d, shape, niter=10, init=prev, **kw)
        conv = p.clone()
        info["conv_niter"] = args.conv_niter
        info["conv_change_per_iter_last10"] = round(rmse(conv, prev) / 10, 6)
        info["conv_seconds"] = round(time.perf_counter() - t, 1)

        def search(init):
            """Smallest niter reaching the target from init (None = zeros)."""
            cur = init
            for k in range(0, args.max_niter + 1):
                if k > 0:
                    cur = dip.dip_device(d, shape, niter=1, init=cur.clone() if cur is not None else None, **kw)
                if cur is not None and rmse(cur, conv) <= args.target_rmse:
                    return k, rmse(cur, conv)
            return None, rmse(cur, conv)

        niter_cold, r_cold = search(None)
        info["cold_niter"] = niter_cold
        info["cold_rmse"] = round(r_cold, 5)
        niter_warm = None
        if net is not None:
            init = net.predict_device(d, shape)
            info["warm_init_rmse"] = round(rmse(init, conv), 5)
            niter_warm, r_warm = search(init)
            info["warm_niter"] = niter_warm
            info["warm_rmse"] = round(r_warm, 5)
        if niter_cold is None:
            niter_cold = args.max_niter

    def cold_step():
        dip.dip_device(d, shape, niter=niter_cold, out=dips, **kw)
        filters.somf3d_device(d, dips[:n], dips[n:], shape, 4, 4, order=2, out=out)

    init_buf = torch.empty(2 * n, dtype=torch.float32, device="cuda")

    def warm_step():
        net.predict_device(d, shape, out=init_buf)
        dip.dip_device(d, shape, niter=niter_warm, init=init_buf, out=dips, **kw)
        filters.somf3d_device(d, dips[:n], dips[n:], shape, 4, 4, order=2, out=out)

    use_warm = net is not None and niter_warm is not None
    step = warm_step if use_warm else cold_step

    def timed(fn, steps, profile):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        if profile:
            _lib.profile_enable(True)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record()
        for _ in range(steps):
            fn()
        ev1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        prof = _lib.profile_report() if profile else None
        if profile:
            _lib.profile_enable(False)
        return ev0.elapsed_time(ev1) / steps, wall, prof

    ms_step, wall, prof = timed(step, args.steps, True)
    value = n / (ms_step / 1e3)
    cold_ms = None
    if use_warm:
        cold_ms, _, _ = timed(cold_step, 1, False)
        info["cold_ms_per_step"] = round(cold_ms, 3)
        info["cold_samples_per_s"] = round(n / (cold_ms / 1e3), 1)
        info["speedup_warm_vs_cold"] = round(cold_ms / ms_step, 3)

    # dominant plain (HBM-bound) kernel by total time; achieved =
    # algorithmic bytes / time.  The network's convolutions are MFMA work and
    # are reported separately.
    plain = {k: v for k, v in prof.items() if not k.startswith("conv")}
    name, (cnt, ms, by) = max(plain.items(), key=lambda kv: kv[1][1])
    achieved = by / (ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "launches": int(cnt), "avg_us": round(ms / cnt * 1e3, 2), "bytes_per_launch": by / cnt}
    top = sorted(prof.items(), key=lambda kv: -kv[1][1])[:8]
    breakdown = {k: {"ms_per_step": round(v[1] / args.steps, 3),
                     "GB/s": round(v[2] / (v[1] / 1e3) / 1e9, 1) if v[1] > 0 else None} for k, v in top}
    conv_k = {k: v for k, v in prof.items() if k.startswith("conv")}
    if conv_k:
        cms = sum(v[1] for v in conv_k.values()) / args.steps
        info["cnn_ms_per_step"] = round(cms, 3)
        info["cnn_tflops"] = round(n * 226.0e3 / (cms / 1e3) / 1e12, 1)
        info["cnn_frac_bf16_peak"] = round(info["cnn_tflops"] / BF16_PEAK_TFLOPS, 4)

    cpu = None if args.no_cpu else cpu_baseline(tuple(int(s) for s in args.cpu_shape.split(",")))
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded curved Ricker events + noise + spikes)",
        "config": dict({"workload": ("warm start: CNN + dip3d" if use_warm else "cold start: dip3d") +
                        "(liter=10, order=2, rect=5,5,5) + somf3d(r1=r2=4, order=2, median of 33 slots) "
                        f"at dip RMSE <= {args.target_rmse} vs converged dip3d",
                        "shape": list(shape), "samples_per_step": n, "start": "warm" if use_warm else "cold",
                        "niter": niter_warm if use_warm else niter_cold, "wall_s": round(wall, 3)}, **info),
        "roofline": roofline, "cpu_baseline": cpu, "breakdown": breakdown,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
