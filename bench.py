#!/usr/bin/env python
"""Benchmark of the structural hot path: dip3d + somf3d on one MI355X at
equal dip error, learned warm start against cold start.

Workload: a device-resident FP32 512 x 256 x 256 synthetic volume (seed 5:
four curved Ricker events, slopes within +-3 samples/trace, Gaussian noise
0.3, 1% spikes of amplitude 5).  dip3d runs with liter=10, order=2,
rect=5x5x5; somf3d with r1=r2=4, order=2 (median of 2*r1*r2+1 = 33 slots).

1. p_conv = the plain cold-start dip3d run until the RMS change per
   Gauss-Newton iteration drops below 1e-3 (or --conv-niter iterations).
2. Cold start: dip3d(niter=5) (the reference's default) + somf3d; its dip
   RMSE against p_conv is the error to match.
3. Warm start: the trained network's prediction, then --warm-niter
   Gauss-Newton iterations; it must reach an RMSE no larger than the cold
   start's (the equal-error check is reported).
4. Timed: --steps steps of the warm path, then --steps of the cold path,
   each between synchronisations with inputs resident in HBM.
Also reported: how many cold iterations reach --target-rmse of p_conv and
the lowest error the warm start reaches.

Prints one JSON line: value = samples/s of the warm path (the cold path when
no trained network is present or it misses the equal-error check), the cold
rate and the speed-up in config, the roofline of the dominant plain kernel
from live HIP-event timing, and the CPU baseline = the reference's own C
extension on a bounded sample, one host core.
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
    ap.add_argument("--cold-niter", type=int, default=5)
    ap.add_argument("--warm-niter", type=int, default=2)
    ap.add_argument("--target-rmse", type=float, default=0.02)
    ap.add_argument("--conv-niter", type=int, default=600, help="cap of the convergence run")
    ap.add_argument("--conv-tol", type=float, default=1e-3, help="RMS change per iteration of p_conv")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-conv", action="store_true",
                    help="skip p_conv and the error checks (profiling runs: timing only)")
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
    init_buf = torch.empty(2 * n, dtype=torch.float32, device="cuda")
    nc, nw = args.cold_niter, args.warm_niter

    if not args.no_conv:
        # 1. converged reference; continuation is exact (the state between
        #    Gauss-Newton iterations is the dip field alone)
        t = time.perf_counter()
        p = dip.dip_device(d, shape, niter=10, **kw)
        k, ch = 10, None
        while k < args.conv_niter:
            q = dip.dip_device(d, shape, niter=10, init=p, **kw)
            ch = rmse(q, p) / 10
            p, k = q, k + 10
            if ch < args.conv_tol:
                break
        conv = p
        info["conv_niter"] = k
        info["conv_change_per_iter"] = round(ch, 6)
        info["conv_seconds"] = round(time.perf_counter() - t, 1)
        cold = dip.dip_device(d, shape, niter=nc, **kw)
        r_cold = rmse(cold, conv)
        info["cold_rmse_vs_conv"] = round(r_cold, 5)
        # cold iterations needed for the absolute target
        cur, k2 = cold, nc
        while rmse(cur, conv) > args.target_rmse and k2 < k:
            cur = dip.dip_device(d, shape, niter=1, init=cur, **kw)
            k2 += 1
        info[f"cold_niter_to_rmse_{args.target_rmse}"] = k2
        if net is not None:
            init = net.predict_device(d, shape, out=init_buf, xscale=warmstart.input_scale(d))
            info["warm_init_rmse_vs_conv"] = round(rmse(init, conv), 5)
            best, cur = rmse(init, conv), init.clone()
            for j in range(1, max(nw, 5) + 1):
                cur = dip.dip_device(d, shape, niter=1, init=cur, **kw)
                if j == nw:
                    info["warm_rmse_vs_conv"] = round(rmse(cur, conv), 5)
                best = min(best, rmse(cur, conv))
            if nw == 0:
                info["warm_rmse_vs_conv"] = info["warm_init_rmse_vs_conv"]
            info["warm_min_rmse_vs_conv"] = round(best, 5)
            info["warm_reaches_target_rmse"] = best <= args.target_rmse
            info["equal_error"] = info["warm_rmse_vs_conv"] <= r_cold

    def cold_step():
        dip.dip_device(d, shape, niter=nc, out=dips, **kw)
        filters.somf3d_device(d, dips[:n], dips[n:], shape, 4, 4, order=2, out=out)

    def warm_step():
        net.predict_device(d, shape, out=init_buf)  # includes the input RMS (one host sync)
        dip.dip_device(d, shape, niter=nw, init=init_buf, out=dips, **kw)
        filters.somf3d_device(d, dips[:n], dips[n:], shape, 4, 4, order=2, out=out)

    use_warm = net is not None and info.get("equal_error", args.no_conv)
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
    if use_warm:
        cold_ms, _, _ = timed(cold_step, args.steps, False)
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
        "config": dict({"workload": (f"warm start: CNN + dip3d(niter={nw}" if use_warm else
                                     f"cold start: dip3d(niter={nc}") +
                        ", liter=10, order=2, rect=5,5,5) + somf3d(r1=r2=4, order=2, median of 33 slots); "
                        f"equal dip RMSE vs converged dip3d against the cold start dip3d(niter={nc})",
                        "shape": list(shape), "samples_per_step": n, "start": "warm" if use_warm else "cold",
                        "wall_s": round(wall, 3)}, **info),
        "roofline": roofline, "cpu_baseline": cpu, "breakdown": breakdown,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
