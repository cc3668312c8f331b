"""Benchmark of the BlackTracer hot path on one MI355X.

Workload (configs[1], "full frame"): camera at r = 5M, theta = pi/2 on the
prograde circular orbit of an a/M = 0.999 hole; adaptive Grid(10, 1)
(1025 x 2048 camera sky, refined where neighbouring celestial angles differ);
1920 x 1080 frame with 100k seeded stars and a seeded 4096 x 2048 sky.

One step = one frame from scratch: the adaptive grid build followed by the
render of one 1080p frame.  `value` is the grid build's rays/s (rays traced /
grid-build seconds, device-resident, HIP-event timed).  Frame latency p50/p99
is measured separately over frames rendered from the resident grid.

Extra objects: `roofline` for the dominant kernel (the Cash-Karp integrator,
VALU-bound: algorithmic FLOPs of Metric.h:198-270 per attempt x attempts /
kernel time) and `cpu_baseline` (the C oracle's grid build of the same pose
on the host cores, a bounded sample).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--precision fp64|mixed]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic work per Cash-Karp attempt (odeint1 loop body, Metric.h:292-297):
# 6 x derivs (136 +,-,*,/ each, Metric.h:198-226) + rkck combinations (270,
# Metric.h:228-251) + error scale (15, Metric.h:293-294) + rkqs control (14,
# Metric.h:258-269).  sin/cos/pow are tallied separately (6 sin, 9 cos as
# written, 1 pow per attempt) and not counted as FLOPs.
FLOPS_PER_ATTEMPT = 6 * 136 + 270 + 15 + 14
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 vector
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector
PEAK_HBM_GBS = 8000.0

POSE = dict(a=0.999, theta=math.pi / 2, phi=0.0, r=5.0)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--frames", type=int, default=50, help="frames for p50/p99")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    return ap.parse_args()


def cpu_baseline(seconds):
    """The C oracle's Grid(10, 1) build of the same pose on the host cores."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    ps = O.pose(POSE["a"], POSE["theta"], POSE["phi"], POSE["r"])
    rays, t0, builds = 0, time.perf_counter(), 0
    while True:
        g = O.OracleGrid(ps, O.params(), 10, 1, True, nthreads=threads)
        rays += g.rays_traced
        builds += 1
        el = time.perf_counter() - t0
        if el >= seconds or builds >= 50:
            break
    return {"value": rays / el, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"{builds} oracle builds of Grid(10,1) at a=0.999 r=5M "
                      f"({rays // builds} rays each), OpenMP over rays"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import blacktracer_amd as bt

    ctx = bt.default_context()
    bh = bt.BlackHole(POSE["a"])
    cam = bt.Camera(POSE["theta"], POSE["phi"], POSE["r"], black_hole=bh)
    prec = bt.BT_PREC_MIXED if args.precision == "mixed" else bt.BT_PREC_FP64
    prm = bt.IntegParams(prec)
    rk_name = "rk_mixed" if prec == bt.BT_PREC_MIXED else "rk_f64"

    for _ in range(args.warmup):
        g = bt.Grid(10, 1, True, cam, bh, params=prm, ctx=ctx)
        g.close()
    ctx.sync()
    ctx.timing(True)
    ctx.reset_timers()
    grid_ms, rays, attempts = [], 0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        g = bt.Grid(10, 1, True, cam, bh, params=prm, ctx=ctx)
        ctx.sync()
        grid_ms.append(1e3 * (time.perf_counter() - s0))
        rays += g.rays_traced
        attempts += g.stats()["attempts"]
        g.close()
    ctx.sync()
    wall = time.perf_counter() - t0
    n_rk, rk_ms = ctx.timer(rk_name)
    ctx.timing(False)

    achieved = FLOPS_PER_ATTEMPT * attempts / (rk_ms * 1e-3) / 1e12 if rk_ms > 0 else 0.0
    peak = PEAK_FP64_TFLOPS if prec == bt.BT_PREC_FP64 else PEAK_FP32_TFLOPS
    grid_s = sum(grid_ms) / 1e3
    out = {
        "metric": "rays/s (mixed vs FP64 at <=1 arcsec err); p50/p99 frame ms at 1080p",
        "value": rays / grid_s,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == bt.BT_PREC_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "configs[1] full frame: a=0.999, r=5M, theta=pi/2 circular orbit, "
                               "adaptive Grid(10,1) 1025x2048",
                   "precision": args.precision, "rays_per_grid": rays // args.steps,
                   "parallelism": "single GPU"},
        "grid_build_ms_p50": sorted(grid_ms)[len(grid_ms) // 2],
        "roofline": {"bound": "valu", "kernel": rk_name, "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": None,
                     "flops_per_attempt": FLOPS_PER_ATTEMPT,
                     "attempts_per_step": attempts // args.steps,
                     "avg_launch_ms": rk_ms / max(1, n_rk)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
