"""Benchmark of the BlackTracer hot path on one MI355X.

Headline workload (the precision config): camera at r = 5M, theta = pi/2 on
the prograde circular orbit of an a/M = 0.999 hole; the uniform
Grid(11, 11) camera sky of 2049 x 4096 directions (8,384,514 rays traced,
the poles once), reference integrator settings (eps 1e-5, z_end -1e7,
MAXSTP 1000, a-priori capture).  One step = one grid build of that sky with
the mixed-precision integrator (FP32 Cash-Karp bulk, FP64 set-up, warm-up,
fallback and accumulation), directions generated and results written in
HBM.  `value` = rays per second of that build when it meets the metric's
gate (every ray within 1 arcsec of the FP64 build, capture flags equal),
else of the FP64 build (`gate`).  The plain FP64 build of the same sky is
timed beside it, and the two grids are compared ray by ray: great-circle
arcsec error of mixed vs FP64 (mean, p99, p99.9, max, share above 1 arcsec)
and capture-flag agreement.  For scale, the FP64 path's own
truncation error is measured against a converged solution (FP64, eps 1e-10)
on the 513 x 1024 sub-lattice.

Also reported:
  frames     the full-frame config: adaptive Grid(10, 1) of the same pose
             (FP64), 1920 x 1080, 100k seeded stars, seeded 4096 x 2048 sky;
             p50/p99 latency of frames rendered from the resident grid while
             the camera advances along its orbit (MakeImageKernel.cu:1412-1413).
  batch      the batching config: 64 seeded poses (a in [0, 0.999], r in
             [3, 50], theta in [0.1, pi - 0.1], max level 6..11), built in
             one bt_grid_build_batch call, mixed and FP64.
  serving    the serving config: a Poisson stream (200 requests/s) of
             1280x720 / 1920x1080 / 3840x2160 frames (60/30/10%) of four
             orbiting cameras through the continuous-batching server; frames/s
             and p50/p99 latency (arrival -> frame in host memory).
`roofline`: the dominant kernel of a step, rk_mixed (FP32 VALU): algorithmic
FLOPs (FLOPS_PER_ATTEMPT x its attempts) per second of its time, HIP events
on the library's stream, against the FP32 vector peak.  `render_roofline`:
algorithmic bytes of a 1080p frame (grid, stars, view, two summed-table
reads per sky column of each pixel's footprint, the BGRA frame) per second
of frame wall time against 8 TB/s.  `cpu_baseline`: the reference's own
integrator (the C oracle where it is not built) tracing a bounded sample of
the same sky (every 64th ray) on the host's cores and on one thread.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-batch] [--no-serve] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

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
# HBM bytes of one rk_mixed launch over the 8.4M-ray sky: rocprofv3 FETCH_SIZE
# and WRITE_SIZE of that dispatch, each scaled by the same counter on a 1 GiB
# device copy (profiles/integrator_precision_config_counters.md).
TRAFFIC_RK_MIXED_BYTES = 984.7e6
ARCSEC = math.pi / 180 / 3600

POSE = dict(a=0.999, theta=math.pi / 2, phi=0.0, r=5.0)
LEVEL = 11
W, H = 1920, 1080
# timers of disjoint launches (star_heavy covers star_tree, star_chunk and star_reduce)
RENDER_KERNELS = ("pix_interp", "bh_area", "smooth", "star_light", "star_odd", "star_heavy", "star_sum", "env_map",
                  "env_heavy", "composite")
MIXED_KERNELS = ("ic_split", "rk_f64_warmup", "rk_mixed", "rk_f64_fallback")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=60, help="frames for p50/p99")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-pose batch")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving stream")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--profile-run", action="store_true",
                    help="a short fixed program for rocprofv3 (copy calibration, one build each, one frame)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--precision-sweep", action="store_true",
                    help="mixed vs FP64 error and rays/s per pose of the 9-pose precision workload")
    return ap.parse_args()


def pct(xs, p):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, max(0, int(math.ceil(p * len(xs))) - 1))]


def gc_err(t1, p1, t2, p2):
    h = np.sin((t2 - t1) / 2) ** 2 + np.sin(t1) * np.sin(t2) * np.sin((p2 - p1) / 2) ** 2
    return 2 * np.arcsin(np.sqrt(np.clip(h, 0, 1)))


_T0 = time.perf_counter()


def progress(msg):
    """A progress line on stderr (stdout carries only the result line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def g_rays(level):
    """Rays of the uniform Grid(level, level) sky: the autotuner calibrates at
    the size it tunes for."""
    n = (1 << level) + 1
    return n * 2 * (n - 1)


def err_stats(cel_a, cel_b):
    """Great-circle errors of grid b against grid a over the rays both trace
    to the sky."""
    ta, pa = cel_a[..., 0].ravel(), cel_a[..., 1].ravel()
    tb, pb = cel_b[..., 0].ravel(), cel_b[..., 1].ravel()
    traced = ta != -2
    cap_a, cap_b = ta == -1, tb == -1
    ok = traced & ~cap_a & ~cap_b
    e = gc_err(ta[ok], pa[ok], tb[ok], pb[ok]) / ARCSEC
    return {"rays_compared": int(ok.sum()),
            "capture_flags_equal": bool(np.array_equal(cap_a[traced], cap_b[traced])),
            "mean_arcsec": float(e.mean()), "p99_arcsec": float(np.quantile(e, 0.99)),
            "p999_arcsec": float(np.quantile(e, 0.999)), "max_arcsec": float(e.max()),
            "above_1arcsec": int(np.sum(e > 1.0)), "frac_above_1arcsec": float(np.mean(e > 1.0))}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(seconds, keep=None):
    """The reference's own CPU integrator (Metric.h's checkCelest and
    rkckIntegrate1, compiled from its sources into oracle/_ref) tracing every
    64th ray of the 2049 x 4096 sky on the host cores, once on one thread and
    once on the host share of cores (OMP_NUM_THREADS where set, else nproc,
    at most 16: the share of the host one GPU gets); the C oracle where that
    build is absent."""
    from oracle import oracle as O
    nproc = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc
    threads = max(1, min(16, share, nproc))
    ps = O.pose(POSE["a"], POSE["theta"], POSE["phi"], POSE["r"])
    N = 2 ** LEVEL + 1
    M = 2 * (N - 1)
    i, j = np.meshgrid(np.arange(N), np.arange(M), indexing="ij")
    th = (i / (N - 1) * math.pi).ravel()[::64]
    ph = (j / M * 2 * math.pi).ravel()[::64]
    chunk = 16384
    host = {"cpu_model": _cpu_model(), "nproc": nproc}
    if O.load_ref() is not None:
        # initial conditions (Grid.h:524-547, restated) outside the timed region
        ic = O.initial_conditions_batch(ps, th, ph)

        def run(nthreads, budget, res=None):
            rays, el, k = 0, 0.0, 0
            while el < budget:
                part = ic[k:k + chunk]
                t0 = time.perf_counter()
                r_th, r_ph, r_st = O.ref_trace_batch(POSE["a"], POSE["r"], POSE["theta"], POSE["phi"], part,
                                                      nthreads=nthreads)
                el += time.perf_counter() - t0
                if res is not None and rays < th.size:   # the first pass: kept for the parity check
                    res[0][k:k + part.shape[0]], res[1][k:k + part.shape[0]], res[2][k:k + part.shape[0]] = \
                        r_th, r_ph, r_st
                rays += part.shape[0]
                k = (k + chunk) % th.size
            return rays, el

        res = [np.full(th.size, np.nan), np.full(th.size, np.nan), np.zeros(th.size, np.int32)]
        rays, el = run(threads, seconds, res)
        rays1, el1 = run(1, max(2.0, seconds / 4))
        if keep is not None:
            keep.extend(res)
        return {"value": rays / el, "unit": "rays/s", "cores": threads, "kind": "reference",
                "single_thread": {"value": rays1 / el1, "unit": "rays/s", "cores": 1,
                                  "sample": f"{rays1} rays, {el1:.1f} s"},
                "host": host,
                "sample": f"{rays} rays of every 64th direction of the 2049x4096 sky (reference settings): "
                          f"the reference's checkCelest + rkckIntegrate1 built from its sources, OpenMP over "
                          f"rays on {threads} threads, {el:.1f} s (initial conditions computed beforehand, "
                          f"untimed); then on 1 thread"}

    def run_port(nthreads, budget):
        rays, k = 0, 0
        t0 = time.perf_counter()
        while True:
            O.trace_rays(ps, O.params(), th[k:k + chunk], ph[k:k + chunk], nthreads=nthreads)
            rays += min(chunk, th.size - k)
            k = (k + chunk) % th.size
            el = time.perf_counter() - t0
            if el >= budget:
                return rays, el

    rays, el = run_port(threads, seconds)
    rays1, el1 = run_port(1, max(2.0, seconds / 4))
    return {"value": rays / el, "unit": "rays/s", "cores": threads, "kind": "port",
            "single_thread": {"value": rays1 / el1, "unit": "rays/s", "cores": 1, "sample": f"{rays1} rays, {el1:.1f} s"},
            "host": host,
            "sample": f"{rays} rays of every 64th direction of the 2049x4096 sky (reference settings), "
                      f"C oracle, OpenMP over rays on {threads} threads, {el:.1f} s; then on 1 thread"}


def fp64_vs_reference(gpu, ref):
    """The FP64 grid against the reference's own integrator (the CPU
    baseline's first pass) on every 64th direction of the sky, the pole rows
    left out (the grid copies its pole rays along them, Grid.h:306-312)."""
    N = 2 ** LEVEL + 1
    M = 2 * (N - 1)
    idx = np.arange(N * M)[::64]
    rows = idx // M
    m = (rows > 0) & (rows < N - 1) & np.isfinite(ref[0])
    gt, gp, gs = gpu[0][m], gpu[1][m], gpu[2][m]
    rt, rp, rs = ref[0][m], ref[1][m], ref[2][m]
    cap_g, cap_r = gt == -1, rt == -1
    esc = ~cap_g & ~cap_r
    dt = np.abs(gt[esc] - rt[esc])
    dp = np.abs(gp[esc] - rp[esc])
    dp = np.minimum(dp, 2 * math.pi - dp)
    return {"rays": int(m.sum()), "capture_flags_equal": bool(np.array_equal(cap_g, cap_r)),
            "steps_equal_frac": float(np.mean(gs == rs)), "max_abs_rad": float(max(dt.max(initial=0), dp.max(initial=0))),
            "frac_beyond_1e-9_rad": float(np.mean(np.maximum(dt, dp) > 1e-9)),
            "note": "GPU FP64 path vs the reference's checkCelest + rkckIntegrate1 (built from its sources), "
                    "every 64th direction of the 2049x4096 sky, pole rows excluded"}


def env_columns(thphi, offset, rows):
    """Sky-map columns each pixel's footprint spans (the pixNum of
    distortEnvironmentMap, MakeImageKernel.cu:640-660) from a frame's pixel
    corners: the quad's phi range after the frame offset and the 2 pi seam
    fix (factor 0.2), in float as the kernel computes it; 0 for shadow pixels
    and for pixel 0, which the reference skips."""
    t = thphi[..., 0].astype(np.float32)
    p = thphi[..., 1].astype(np.float32)
    T = np.stack([t[1:, :-1], t[:-1, :-1], t[:-1, 1:], t[1:, 1:]], -1)
    P = np.stack([p[1:, :-1], p[:-1, :-1], p[:-1, 1:], p[1:, 1:]], -1)
    live = (T >= 0).all(-1) & (P >= 0).all(-1)
    two_pi = np.float32(2 * math.pi)
    P = np.fmod(P + np.float32(offset), two_pi).astype(np.float32)
    hi = (P > two_pi * np.float32(0.8)).any(-1, keepdims=True)
    P = np.where(hi & (P < two_pi * np.float32(0.2)), P + two_pi, P).astype(np.float32)
    pix = np.float32(math.pi) / np.float32(rows)
    phmax, phmin = P.max(-1), P.min(-1)
    pmax = np.trunc(phmax / pix).astype(np.int64)
    pmin = np.trunc(phmin / pix).astype(np.int64)
    n = pmax - pmin + 1
    n = np.where((n > 1) & (pmax.astype(np.float32) * pix > phmax), n - 1, n)
    live[0, 0] = False
    return np.where(live, n, 0)


def frames(bt, S, ctx, nframes):
    bh = bt.BlackHole(POSE["a"])
    cam = bt.Camera(POSE["theta"], POSE["phi"], POSE["r"], black_hole=bh)
    stars = S.synthetic_stars(100_000, seed=1234)
    sky = S.synthetic_sky(4096, seed=99)
    sc = bt.Scene(stars, sky, tree_level=8, ctx=ctx)
    view = bt.View(W, H, ctx=ctx)
    buf = ctx.alloc(4 * W * H)
    t0 = time.perf_counter()
    g = bt.Grid(10, 1, True, cam, bh, ctx=ctx)
    ctx.sync()
    grid_ms = 1e3 * (time.perf_counter() - t0)
    speed = 1.0 / float(cam.getParamArray()[0])
    ms = []
    for q in range(nframes + 3):
        off = 2 * math.pi * q / (0.25 * speed * W)
        if q == 3:
            ctx.sync()
            ctx.timing(True)
            ctx.reset_timers()
        s0 = time.perf_counter()
        bt.render_frame_device(g, sc, view, buf, flags=bt.RenderFlags(offset=off))
        ctx.sync()
        if q >= 3:
            ms.append(1e3 * (time.perf_counter() - s0))
    kern = {n: ctx.timer(n) for n in RENDER_KERNELS}
    ctx.timing(False)
    nf = len(ms)
    # one frame to the host, for the CPU leg's check against the oracle
    off_chk = 2 * math.pi * 7 / (0.25 * speed * W)
    check = {"img": bt.render_frame(g, sc, view, flags=bt.RenderFlags(offset=off_chk)),
             "grid": g.cel.astype(np.float32), "gridlvl": g.MAXLEVEL, "cam7": cam.getParamArray(),
             "offset": off_chk, "stars": stars, "sky": sky}
    extra = display_effects(bt, ctx, g, sc, view, buf, cam, bh, speed)
    kern_ms = sum(v[1] for v in kern.values()) / nf
    # algorithmic bytes of a frame: the FP32 grid, the stars, the view
    # (corner directions and pixel solid angles), two float4 summed-table
    # reads per sky column of each pixel's footprint, the BGRA frame
    _, stg = bt.render_frame(g, sc, view, flags=bt.RenderFlags(offset=off_chk), stages=True)
    sky_cols = int(env_columns(stg["thphi"], off_chk, sky.shape[0]).sum())
    b_frame = (8 * g.N * g.M + 16 * stars.shape[0] + 8 * (W + 1) * (H + 1) + 4 * W * H + 32 * sky_cols +
               4 * W * H)
    # per frame of wall time: the star and sky passes overlap on two streams,
    # so the kernels' times add up to more than the frame
    ach = b_frame / (pct(ms, 0.5) * 1e-3) / 1e9
    out = {"workload": "adaptive Grid(10,1) 1025x2048 (FP64) + 1920x1080 frame, 100k stars, 4096x2048 sky",
           "grid_build_ms": grid_ms, "grid_rays": g.rays_traced,
           "frame_ms_p50": pct(ms, 0.5), "frame_ms_p99": pct(ms, 0.99), "frames_per_s": 1e3 * nf / sum(ms),
           "display_effects": extra}
    roof = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
            "bytes_per_frame": b_frame, "sky_columns_per_frame": sky_cols,
            "bytes_note": "8 B per FP32 grid point + 16 B per star + view corners (8 B) and pixel solid angles "
                          "(4 B) + 32 B per sky column of each pixel's footprint (two float4 summed-table reads) "
                          "+ 4 B BGRA per pixel",
            "frame_ms_p50": pct(ms, 0.5), "kernel_ms_per_frame": kern_ms,
            "per_kernel_ms": {n: v[1] / max(1, v[0]) for n, v in kern.items()}}
    g.close()
    return out, roof, check


def cpu_render(check, threads):
    """The C oracle's 1080p frame of the full-frame config on the host cores
    (a port: the reference renders only on CUDA), and the GPU frame checked
    against it."""
    from oracle import oracle as O
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    tree, pos, mag = O.make_tree(check["stars"], 8)
    table = O.sum_table(check["sky"])
    view = O.camera_view(math.pi / 2, 0.0, 0.25 * math.pi / 4, W, H)
    t0 = time.perf_counter()
    img = O.frame(check["grid"], check["gridlvl"], view, check["cam7"], tree, 8, pos, mag, table,
                  offset=check["offset"])
    el = time.perf_counter() - t0
    d = np.abs(img.astype(np.int16) - check["img"].astype(np.int16))
    return {"frame_ms": 1e3 * el, "cores": threads, "kind": "port",
            "sample": "one 1920x1080 frame (stars, sky, lensing) by the C oracle; the star tree and summed table "
                      "are built beforehand, untimed",
            "gpu_frame_vs_oracle": {"max_channel_diff": int(d.max()),
                                    "frac_pixels_differing": float((d.max(axis=-1) > 0).mean())}}


def display_effects(bt, ctx, g, sc, view, buf, cam, bh, speed, nframes=20):
    """Frame times of the display loop's extras at the frame config: star
    trails, trails with a diffraction sprite, and a frame blended between
    this grid and a second pose's (r + 0.5M)."""
    def timed(render):
        ms = []
        for q in range(nframes + 3):
            s0 = time.perf_counter()
            render(q)
            ctx.sync()
            if q >= 3:
                ms.append(1e3 * (time.perf_counter() - s0))
        return {"frame_ms_p50": pct(ms, 0.5), "frame_ms_p99": pct(ms, 0.99)}

    def off(q):
        return bt.RenderFlags(offset=2 * math.pi * q / (0.25 * speed * W))

    out = {}
    T = bt.Trails(sc, W, H, ctx=ctx)
    out["trails"] = timed(lambda q: T.render(g, sc, view, flags=off(q), out_buffer=buf))
    T.close()
    fs = W // 16
    y, x = np.mgrid[0:fs, 0:fs] - fs // 2
    sprite = np.zeros((fs, fs, 3), np.uint8)
    sprite[(x == 0) | (y == 0)] = (255, 220, 180)
    T = bt.Trails(sc, W, H, diffraction=sprite, ctx=ctx)
    out["trails_diffraction"] = timed(lambda q: T.render(g, sc, view, flags=off(q), out_buffer=buf))
    T.close()
    cam2 = bt.Camera(POSE["theta"], POSE["phi"], POSE["r"] + 0.5, black_hole=bh)
    g2 = bt.Grid(10, 1, True, cam2, bh, ctx=ctx)
    out["blend"] = timed(lambda q: bt.render_frame_blend(g, g2, (q % 20) / 20.0, sc, view, flags=off(q),
                                                         out_buffer=buf))
    g2.close()
    out["note"] = ("1920x1080 frames as above: trails (per-star image caches over consecutive frames), "
                   "trails + a 120x120 diffraction sprite, and two-grid blending (r=5M and 5.5M)")
    return out


def batch(bt, ctx):
    from blacktracer_amd import scene as S
    cams, bhs, mls = [], [], []
    for a, r, th, ml in S.batch_poses(64, seed=2024):
        bh = bt.BlackHole(a)
        bhs.append(bh)
        cams.append(bt.Camera(th, 0.0, r, black_hole=bh))
        mls.append(ml)
    out = {"workload": "64 seeded poses, adaptive grids max level 6..11 (start 1), one bt_grid_build_batch; "
                       "median of 3 builds after a warm-up"}
    for name, prec in (("fp64", bt.BT_PREC_FP64), ("mixed", bt.BT_PREC_MIXED)):
        prm = bt.IntegParams(prec)
        grids, _ = bt.Grid.build_batch(cams, bhs, mls, [1] * 64, params=prm, ctx=ctx)   # warm-up
        for g in grids:
            g.close()
        ctx.sync()
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            grids, st = bt.Grid.build_batch(cams, bhs, mls, [1] * 64, params=prm, ctx=ctx)
            ctx.sync()
            times.append(time.perf_counter() - t0)
            rays = sum(g.rays_traced for g in grids)
            for g in grids:
                g.close()
        el = float(np.median(times))
        out[name] = {"rays": rays, "seconds": el, "seconds_runs": times, "rays_per_s": rays / el,
                     "fp64_attempt_share": st["attempts_fp64"] / max(1, st["attempts"])}
    out["mixed_over_fp64"] = out["mixed"]["rays_per_s"] / out["fp64"]["rays_per_s"]
    return out


def serve(bt, S, ctx, seconds, rate=200.0):
    """Poisson stream of render requests (mean `rate`/s) over four orbiting
    cameras, resolutions 1280x720 / 1920x1080 / 3840x2160 at 60/30/10%,
    stars and sky of the full-frame config, through the serving loop.  The
    four grids are built (and so cached) by a first request each, and a
    warm-up makes the view of every resolution and sizes every slot's
    buffers; the timed stream then runs against a warm server."""
    stars = S.synthetic_stars(100_000, seed=1234)
    sky = S.synthetic_sky(4096, seed=99)
    sc = bt.Scene(stars, sky, tree_level=8, ctx=ctx)
    srv = bt.Server(sc, hbm_budget=64 << 30, slots=4, max_level=10, start_level=1, ctx=ctx)
    rng = np.random.default_rng(7)
    traj = [(0.999, 5.0, math.pi / 2), (0.9, 8.0, 1.3), (0.5, 12.0, 1.8), (0.99, 6.0, 1.0)]
    speed = [1.0 / float(bt.Camera(th, 0.0, r, black_hole=bt.BlackHole(a)).getParamArray()[0])
             for a, r, th in traj]
    res = [(1280, 720), (1920, 1080), (3840, 2160)]
    bufs = {wh: [np.zeros((wh[1], wh[0], 4), dtype=np.uint8) for _ in range(24)] for wh in res}
    t0 = time.perf_counter()
    for k, pose in enumerate(traj):                       # cold: one request per pose
        srv.submit(*pose, 0.0, bufs[res[0]][k], 10_000 + k)
    srv.drain(300)
    cold_s = time.perf_counter() - t0
    # warm-up: the view of every resolution, and every slot's frame buffers
    # at the largest resolution (four 4K frames in flight at once)
    for k in range(4):
        srv.submit(*traj[k], 0.0, bufs[res[2]][k], 20_000 + k)
    srv.submit(*traj[0], 0.0, bufs[res[1]][0], 20_010)
    srv.drain(300)
    srv.poll(64)
    warm = srv.stats()
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * seconds * 1.2)))
    arrivals = arrivals[arrivals < seconds]
    kinds = rng.choice(3, size=arrivals.size, p=[0.6, 0.3, 0.1])
    who = rng.integers(0, len(traj), size=arrivals.size)
    frame_no = [0] * len(traj)
    lat, qms, fms = [], [], []
    by_res = {wh: [] for wh in res}
    tag_res = {}

    def take(comps):
        for c in comps:
            lat.append(c["latency_ms"])
            qms.append(c["queue_ms"])
            fms.append(c["frame_ms"])
            if c["tag"] in tag_res:
                by_res[tag_res.pop(c["tag"])].append((c["latency_ms"], c["queue_ms"], c["frame_ms"]))

    start = time.perf_counter()
    for i, t_arr in enumerate(arrivals):
        while True:
            now = time.perf_counter() - start
            if now >= t_arr:
                break
            time.sleep(min(0.0005, t_arr - now))
        j = int(who[i])
        wh = res[int(kinds[i])]
        q = frame_no[j]
        frame_no[j] += 1
        off = 2 * math.pi * q / (0.25 * speed[j] * wh[0])   # MakeImageKernel.cu:1412-1413
        tag_res[i] = wh
        srv.submit(*traj[j], off, bufs[wh][i % 24], i)
        take(srv.poll(256))
    srv.drain(600)
    take(srv.poll(4096))
    wall = time.perf_counter() - start
    st = srv.stats()
    srv.close()
    return {"workload": f"Poisson {rate:g} req/s for {seconds:g} s, 4 orbiting cameras (Grid(10,1) each), "
                        "1280x720/1920x1080/3840x2160 at 60/30/10%, 100k stars, 4096x2048 sky, 4 slots",
            "requests": len(lat), "offered_per_s": arrivals.size / seconds, "frames_per_s": len(lat) / wall,
            "latency_ms_p50": pct(lat, 0.5), "latency_ms_p99": pct(lat, 0.99), "latency_ms_mean": float(np.mean(lat)),
            "queue_ms_p50": pct(qms, 0.5), "frame_ms_p50": pct(fms, 0.5), "frame_ms_p99": pct(fms, 0.99),
            "by_resolution": {f"{w}x{h}": {"requests": len(v),
                                            "latency_ms_p50": pct([x[0] for x in v], 0.5),
                                            "latency_ms_p99": pct([x[0] for x in v], 0.99),
                                            "queue_ms_p99": pct([x[1] for x in v], 0.99),
                                            "frame_ms_p50": pct([x[2] for x in v], 0.5),
                                            "frame_ms_p99": pct([x[2] for x in v], 0.99)}
                              for (w, h), v in by_res.items() if v},
            "cold_start_s": cold_s, "grid_builds": st["grid_builds"], "grid_hits": st["grid_hits"] - warm["grid_hits"],
            "peak_inflight": st["peak_inflight"], "peak_hbm_gib": st["peak_bytes"] / 2 ** 30}


def profile_run(bt, S, ctx):
    """A short fixed program for rocprofv3 counter passes: a 1 GiB device copy
    (the runtime's copy kernel; the bytes calibration of FETCH_SIZE /
    WRITE_SIZE), two mixed and two FP64 builds of the precision-config sky
    (read the second of each), 1080p frames."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    nbytes = 1 << 30
    src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    if hip.hipMemcpy(dst.ptr, src.ptr, nbytes, 3) != 0:   # hipMemcpyDeviceToDevice
        raise RuntimeError("calibration copy failed")
    ctx.sync()
    src.close()
    dst.close()
    bh = bt.BlackHole(POSE["a"])
    cam = bt.Camera(POSE["theta"], POSE["phi"], POSE["r"], black_hole=bh)
    out = {}
    for rep in range(2):   # the second pass is the one to read (warm clocks and caches)
        for name, prec in (("mixed", bt.BT_PREC_MIXED), ("fp64", bt.BT_PREC_FP64)):
            g = bt.Grid(LEVEL, LEVEL, True, cam, bh, params=bt.IntegParams(prec), ctx=ctx)
            out[name] = g.stats()
            g.close()
    fr, _, _ = frames(bt, S, ctx, 2)
    out["frame_ms"] = fr["frame_ms_p50"]
    print(json.dumps(out))


def precision_poses(bt, ctx):
    """Mixed against FP64 on the 2049 x 4096 sky of each of the nine poses of
    the precision config, at the default critical-curve width (0.03) and at
    1e-4: one build of each (at the current launch shape), every ray
    compared."""
    rows = []
    for a in (0.5, 0.9, 0.999):
        for r in (3.0, 5.0, 20.0):
            bh = bt.BlackHole(a)
            cam = bt.Camera(math.pi / 2, 0.0, r, black_hole=bh)
            ctx.sync()
            t0 = time.perf_counter()
            g64 = bt.Grid(LEVEL, LEVEL, True, cam, bh, params=bt.IntegParams(bt.BT_PREC_FP64), ctx=ctx)
            ctx.sync()
            t64 = time.perf_counter() - t0
            row = {"a": a, "r": r, "fp64_ms": 1e3 * t64}
            for name, ce in (("default", None), ("crit_eps_1e-4", 1e-4)):
                prm = bt.IntegParams(bt.BT_PREC_MIXED) if ce is None else bt.IntegParams(bt.BT_PREC_MIXED,
                                                                                         crit_eps=ce)
                ctx.sync()
                t0 = time.perf_counter()
                gm = bt.Grid(LEVEL, LEVEL, True, cam, bh, params=prm, ctx=ctx)
                ctx.sync()
                tm = time.perf_counter() - t0
                e = err_stats(g64.cel, gm.cel)
                st = gm.stats()
                row[name] = {"max_arcsec": e["max_arcsec"], "p99_arcsec": e["p99_arcsec"],
                             "mean_arcsec": e["mean_arcsec"], "above_1arcsec": e["above_1arcsec"],
                             "frac_above_1arcsec": e["frac_above_1arcsec"],
                             "capture_flags_equal": e["capture_flags_equal"],
                             "fp64_attempt_share": st["attempts_fp64"] / max(1, st["attempts"]),
                             "fp64_finished_rays": st["fallback_rays"] / max(1, st["integrated"]),
                             "mixed_ms": 1e3 * tm, "mixed_over_fp64": t64 / tm}
                gm.close()
            g64.close()
            rows.append(row)
    return rows


def precision_sweep(bt, ctx):
    """Mixed against FP64, pose by pose: the 2049 x 4096 sky (uniform
    Grid(11,11)) at a in {0.5, 0.9, 0.999} x r in {3, 5, 20} M, theta = pi/2,
    reference integrator settings; then the radius-stop variant (rel tol 1e-8,
    escape r > 1000M, capture r < 1.01 r+) on the 257 x 513 sky.  Each
    precision runs at the launch shape the autotuner picks for it."""
    rows = []

    def timed(level, cam, bh, prm):
        bt.Grid(level, level, True, cam, bh, params=prm, ctx=ctx).close()
        ts = []
        for _ in range(5):
            ctx.sync()
            t0 = time.perf_counter()
            g = bt.Grid(level, level, True, cam, bh, params=prm, ctx=ctx)
            ctx.sync()
            ts.append(time.perf_counter() - t0)
            if len(ts) < 5:
                g.close()
        return g, sorted(ts)[2]

    cases = [(a, r, LEVEL, "reference") for a in (0.5, 0.9, 0.999) for r in (3.0, 5.0, 20.0)]
    cases.append((0.999, 5.0, 8, "radius-stop"))
    for a, r, level, kind in cases:
        bh = bt.BlackHole(a)
        cam = bt.Camera(math.pi / 2, 0.0, r, black_hole=bh)
        mk = (lambda prec: bt.radius_stop_params(prec)) if kind == "radius-stop" else bt.IntegParams
        res, shapes = {}, {}
        for name, prec in (("fp64", bt.BT_PREC_FP64), ("mixed", bt.BT_PREC_MIXED)):
            shapes[name] = ctx.autotune(cam, params=mk(prec), n_rays=g_rays(level))
            g, t = timed(level, cam, bh, mk(prec))
            res[name] = (g, t)
        g64, t64 = res["fp64"]
        gm, tm = res["mixed"]
        row = {"a": a, "r": r, "grid": f"Grid({level},{level})", "params": kind, "rays": gm.rays_traced}
        row.update(err_stats(g64.cel, gm.cel))
        st = gm.stats()
        row["fp64_attempt_share"] = st["attempts_fp64"] / max(1, st["attempts"])
        row["fallback_rays"] = st["fallback_rays"]
        row["fp64_rays_per_s"] = g64.rays_traced / t64
        row["mixed_rays_per_s"] = gm.rays_traced / tm
        row["mixed_over_fp64"] = t64 / tm
        row["launch_shape"] = shapes
        g64.close()
        gm.close()
        print(json.dumps(row), flush=True)
        rows.append(row)
    return rows


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import blacktracer_amd as bt
    from blacktracer_amd import scene as S

    ctx = bt.default_context()
    if args.profile_run:
        profile_run(bt, S, ctx)
        return
    if args.precision_sweep:
        precision_sweep(bt, ctx)
        return
    bh = bt.BlackHole(POSE["a"])
    cam = bt.Camera(POSE["theta"], POSE["phi"], POSE["r"], black_hole=bh)
    p_mixed = bt.IntegParams(bt.BT_PREC_MIXED)
    p_fp64 = bt.IntegParams(bt.BT_PREC_FP64)

    def build(prm):
        return bt.Grid(LEVEL, LEVEL, True, cam, bh, params=prm, ctx=ctx)

    # ---- headline: mixed-precision build of the 2049 x 4096 sky, at the
    # launch shape the autotuner picks for this pose and precision
    shapes = {}
    progress("autotune (mixed)")
    if not args.no_tune:
        shapes["mixed"] = ctx.autotune(cam, params=p_mixed, n_rays=g_rays(LEVEL))
    progress(f"timed mixed builds: {args.steps} steps after {args.warmup} warm-up")
    for _ in range(args.warmup):
        build(p_mixed).close()
    ctx.sync()
    ctx.timing(True)
    ctx.reset_timers()
    times, rays, attempts, fp64_att = [], 0, 0, 0
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        g = build(p_mixed)
        ctx.sync()
        times.append(time.perf_counter() - t0)
        st = g.stats()
        rays += g.rays_traced
        attempts += st["attempts"] - st["attempts_fp64"]
        fp64_att += st["attempts_fp64"]
        g.close()
    ctx.sync()
    wall = time.perf_counter() - t_all
    kt = {n: ctx.timer(n) for n in MIXED_KERNELS}
    ctx.timing(False)
    value = rays / sum(times)
    n_rk, rk_ms = kt["rk_mixed"]
    achieved = FLOPS_PER_ATTEMPT * attempts / (rk_ms * 1e-3) / 1e12 if rk_ms > 0 else 0.0

    # ---- the plain FP64 build of the same sky (tuned for FP64 in turn)
    progress("FP64 builds")
    if not args.no_tune:
        shapes["fp64"] = ctx.autotune(cam, params=p_fp64, n_rays=g_rays(LEVEL))
    build(p_fp64).close()
    ctx.sync()
    ctx.timing(True)
    ctx.reset_timers()
    f_times = []
    for _ in range(max(2, args.steps // 3)):
        t0 = time.perf_counter()
        g = build(p_fp64)
        ctx.sync()
        f_times.append(time.perf_counter() - t0)
        f_rays = g.rays_traced
        f_att = g.stats()["attempts"]
        g.close()
    n64, rk64_ms = ctx.timer("rk_f64")
    ctx.timing(False)
    fp64_value = f_rays * len(f_times) / sum(f_times)

    if "mixed" in shapes:
        ctx.set_tune(*(shapes["mixed"][k] for k in ("chunk", "block", "waves")))
    # ---- precision: mixed vs FP64 ray by ray; FP64 vs converged on the sub-lattice
    progress("precision comparison")
    gm, g64 = build(p_mixed), build(p_fp64)
    prec = err_stats(g64.cel, gm.cel)
    stm = gm.stats()
    prec["fp64_attempt_share"] = stm["attempts_fp64"] / max(1, stm["attempts"])
    prec["fallback_rays"] = stm["fallback_rays"]
    sub64 = g64.cel[::4, ::4]
    # the FP64 grid's values at the CPU baseline's sample (every 64th direction)
    ref_sample = (g64.cel[..., 0].ravel()[::64].copy(), g64.cel[..., 1].ravel()[::64].copy(),
                  g64.steps.ravel()[::64].copy())
    gm.close()
    g64.close()
    conv = bt.Grid(LEVEL, LEVEL, True, cam, bh, ctx=ctx,
                   params=bt.IntegParams(bt.BT_PREC_FP64, eps=1e-10, max_steps=20000))
    prec["fp64_vs_converged"] = err_stats(conv.cel[::4, ::4], sub64)
    prec["fp64_vs_converged"]["note"] = "FP64 path (eps 1e-5) against FP64 at eps 1e-10, 513x1024 sub-lattice"
    conv.close()

    progress("1080p frames")
    fr, render_roof, frame_check = frames(bt, S, ctx, args.frames)
    # the metric's gate: the mixed build counts only when every ray of the
    # sky is within 1 arcsec of the FP64 build; otherwise the headline is the
    # FP64 path's throughput (which meets the gate trivially)
    mixed_value = value
    gate_ok = prec["capture_flags_equal"] and prec["max_arcsec"] <= 1.0
    value = mixed_value if gate_ok else fp64_value
    out = {
        "metric": "rays/s (mixed vs FP64 at <=1 arcsec err); p50/p99 frame ms at 1080p",
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if gate_ok else "f64",
        "data": "synthetic",
        "config": {"workload": "precision config: a=0.999, r=5M, theta=pi/2 circular orbit; uniform Grid(11,11) "
                               "2049x4096 camera sky (8,384,514 rays); eps 1e-5, z_end -1e7, MAXSTP 1000; "
                               "mixed FP32/FP64 integrator",
                   "rays_per_step": rays // args.steps, "parallelism": "single GPU",
                   "launch_shape": shapes or "default (chunk 128, block 256, waves 0)"},
        "fp64_rays_per_s": fp64_value,
        "mixed_over_fp64": mixed_value / fp64_value,
        "mixed_rays_per_s": mixed_value,
        "gate": {"max_arcsec": prec["max_arcsec"], "capture_flags_equal": prec["capture_flags_equal"],
                 "passed": gate_ok, "value_is": "mixed" if gate_ok else "fp64",
                 "note": "value is the mixed build's rays/s when its max error against the FP64 build over "
                         "every ray of the sky is <= 1 arcsec with equal capture flags, else the FP64 build's"},
        "kernel_ms_per_step": {n: v[1] / max(1, v[0]) for n, v in kt.items()},
        "fp64_rk_ms_per_step": rk64_ms / max(1, n64),
        "precision": prec,
        "frames": fr,
        "roofline": {"bound": "valu", "kernel": "rk_mixed", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": TRAFFIC_RK_MIXED_BYTES,
                     "traffic_note": "HBM bytes of the 8.4M-ray rk_mixed launch from rocprofv3 FETCH_SIZE/WRITE_SIZE",
                     "flops_per_attempt": FLOPS_PER_ATTEMPT, "attempts_per_step": attempts // args.steps,
                     "fp64_attempts_per_step": fp64_att // args.steps,
                     "avg_launch_ms": rk_ms / max(1, n_rk),
                     "fp64_path": {"kernel": "rk_f64", "achieved": FLOPS_PER_ATTEMPT * f_att /
                                   (rk64_ms / max(1, n64) * 1e-3) / 1e12, "peak": PEAK_FP64_TFLOPS}},
        "render_roofline": render_roof,
    }
    out["roofline"]["fp64_path"]["frac"] = out["roofline"]["fp64_path"]["achieved"] / PEAK_FP64_TFLOPS
    progress("nine precision poses")
    out["precision_poses"] = precision_poses(bt, ctx)
    dflt = [p["default"] for p in out["precision_poses"]]
    out["gate"]["all_poses"] = {
        "max_arcsec": max(d["max_arcsec"] for d in dflt),
        "above_1arcsec": sum(d["above_1arcsec"] for d in dflt),
        "capture_flags_equal": all(d["capture_flags_equal"] for d in dflt),
        "passed": all(d["max_arcsec"] <= 1.0 and d["capture_flags_equal"] for d in dflt),
        "note": "the same gate over every ray of each of the nine precision poses (a 0.5/0.9/0.999 x r 3/5/20M, "
                "2049x4096 sky, default critical-curve width)"}
    if not args.no_batch:
        progress("64-pose batch")
        out["batch"] = batch(bt, ctx)
    if not args.no_serve:
        progress("serving stream")
        out["serving"] = serve(bt, S, ctx, args.serve_seconds)
    if not args.no_cpu:
        progress("CPU baseline")
        kept = []
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds, kept)
        if kept:
            out["fp64_vs_reference"] = fp64_vs_reference(ref_sample, kept)
        out["cpu_render_baseline"] = cpu_render(frame_check, out["cpu_baseline"]["cores"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
