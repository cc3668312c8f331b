"""Benchmark of the BlackTracer hot path on one MI355X.

Workload (configs[1], "full frame"): camera at r = 5M, theta = pi/2 on the
prograde circular orbit of an a/M = 0.999 hole; adaptive Grid(10, 1)
(1025 x 2048 camera sky, refined where neighbouring celestial angles
differ); 1920 x 1080 frame with 100k seeded stars and a seeded 4096 x 2048
sky map.

One step = one frame from scratch: the adaptive grid build followed by the
render of one 1080p frame into HBM.  `value` is the grid build's rays/s
(rays traced / grid-build seconds).  Frame latency p50/p99 is measured over
frames rendered from the resident grid while the camera advances along its
orbit (the per-frame sky offset of MakeImageKernel.cu:1412-1413).

`roofline`: the dominant kernel of the grid build, the Cash-Karp integrator,
which is VALU-bound: algorithmic FLOPs (FLOPS_PER_ATTEMPT x attempts) per
second of that kernel, timed with HIP events on the library's stream.
`render_roofline`: the frame's algorithmic HBM bytes per second of render
kernel time against 8 TB/s.  `cpu_baseline`: the C oracle's grid build of
the same pose on the host cores, a bounded sample.

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

# Algorithmic work per Cash-Karp attempt (odeint1's loop body, Metric.h:292-297):
# 6 x derivs (136 +,-,*,/ each as written, Metric.h:198-226) + rkck stage and
# solution combinations (270, Metric.h:228-251) + error scale (15,
# Metric.h:293-294) + rkqs control (14, Metric.h:258-269).  sin/cos/pow are
# tallied separately (per attempt, as written: 12 sin, 18 cos, 1 pow) and
# not counted as FLOPs.
FLOPS_PER_ATTEMPT = 6 * 136 + 270 + 15 + 14
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 vector
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector
PEAK_HBM_GBS = 8000.0

POSE = dict(a=0.999, theta=math.pi / 2, phi=0.0, r=5.0)
W, H = 1920, 1080
RENDER_KERNELS = ("pix_interp", "bh_area", "smooth", "star_light", "star_heavy", "star_sum", "env_map", "composite")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--frames", type=int, default=60, help="frames for p50/p99")
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


def pct(xs, p):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, max(0, int(math.ceil(p * len(xs))) - 1))]


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import blacktracer_amd as bt
    from blacktracer_amd import scene as S

    ctx = bt.default_context()
    bh = bt.BlackHole(POSE["a"])
    cam = bt.Camera(POSE["theta"], POSE["phi"], POSE["r"], black_hole=bh)
    prec = bt.BT_PREC_MIXED if args.precision == "mixed" else bt.BT_PREC_FP64
    prm = bt.IntegParams(prec)
    rk_name = "rk_mixed" if prec == bt.BT_PREC_MIXED else "rk_f64"

    stars = S.synthetic_stars(100_000, seed=1234)
    sky = S.synthetic_sky(4096, seed=99)
    sc = bt.Scene(stars, sky, tree_level=8, ctx=ctx)
    view = bt.View(W, H, ctx=ctx)
    frame_buf = ctx.alloc(4 * W * H)

    def one_step():
        g0 = time.perf_counter()
        g = bt.Grid(10, 1, True, cam, bh, params=prm, ctx=ctx)
        ctx.sync()
        g1 = time.perf_counter()
        bt.render_frame_device(g, sc, view, frame_buf, flags=bt.RenderFlags())
        ctx.sync()
        return g, g1 - g0

    for _ in range(args.warmup):
        g, _ = one_step()
        g.close()
    ctx.sync()
    ctx.timing(True)
    ctx.reset_timers()
    grid_s, rays, attempts = [], 0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g, gs = one_step()
        grid_s.append(gs)
        rays += g.rays_traced
        attempts += g.stats()["attempts"]
        g.close()
    ctx.sync()
    wall = time.perf_counter() - t0
    n_rk, rk_ms = ctx.timer(rk_name)
    ctx.timing(False)
    achieved = FLOPS_PER_ATTEMPT * attempts / (rk_ms * 1e-3) / 1e12 if rk_ms > 0 else 0.0
    peak = PEAK_FP64_TFLOPS if prec == bt.BT_PREC_FP64 else PEAK_FP32_TFLOPS

    # ---- frames from the resident grid, the camera advancing along its orbit
    g = bt.Grid(10, 1, True, cam, bh, params=prm, ctx=ctx)
    speed = 1.0 / float(cam.getParamArray()[0])
    frame_ms = []
    for q in range(args.frames + 3):
        off = 2 * math.pi * q / (0.25 * speed * W)
        if q == 3:
            ctx.sync()
            ctx.timing(True)
            ctx.reset_timers()
        s0 = time.perf_counter()
        bt.render_frame_device(g, sc, view, frame_buf, flags=bt.RenderFlags(offset=off))
        ctx.sync()
        if q >= 3:
            frame_ms.append(1e3 * (time.perf_counter() - s0))
    kern = {n: ctx.timer(n) for n in RENDER_KERNELS}
    ctx.timing(False)
    nf = len(frame_ms)
    kern_ms = sum(v[1] for v in kern.values()) / nf
    # algorithmic HBM bytes of one frame: the grid (float2), the stars
    # (2 x float2), >= 2 float4 summed-table reads per pixel, the BGRA write
    b_frame = 8 * g.N * g.M + 16 * stars.shape[0] + 32 * W * H + 4 * W * H
    ach_gbs = b_frame / (kern_ms * 1e-3) / 1e9

    out = {
        "metric": "rays/s (mixed vs FP64 at <=1 arcsec err); p50/p99 frame ms at 1080p",
        "value": rays / sum(grid_s),
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
                               "adaptive Grid(10,1) 1025x2048, 1920x1080, 100k stars, 4096x2048 sky",
                   "precision": args.precision, "rays_per_grid": rays // args.steps,
                   "parallelism": "single GPU"},
        "grid_build_ms_p50": 1e3 * pct(grid_s, 0.5),
        "frame_ms_p50": pct(frame_ms, 0.5),
        "frame_ms_p99": pct(frame_ms, 0.99),
        "frames_per_s": 1e3 * nf / sum(frame_ms),
        "roofline": {"bound": "valu", "kernel": rk_name, "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                     "flops_per_attempt": FLOPS_PER_ATTEMPT,
                     "attempts_per_step": attempts // args.steps,
                     "avg_launch_ms": rk_ms / max(1, n_rk)},
        "render_roofline": {"bound": "hbm", "achieved": ach_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                            "frac": ach_gbs / PEAK_HBM_GBS, "bytes_per_frame": b_frame,
                            "kernel_ms_per_frame": kern_ms,
                            "per_kernel_ms": {n: v[1] / max(1, v[0]) for n, v in kern.items()}},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
