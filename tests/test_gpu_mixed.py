"""Mixed-precision ray path (BT_PREC_MIXED) against the plain FP64 path.

What is exact: capture flags (the a-priori test runs in FP64 in both
paths); every ray sent to FP64 -- near the critical curve, equatorial,
step collapse, attempt cap, flagged by the rounding-noise accounting --
runs the exact FP64 path from its start and ends exactly as there.  What is not: rays
traced in FP32 take a step sequence of their own, so their celestial angle
differs from the FP64 path's by up to that path's own truncation error; the
bound tested here is statistical (mean, p99) with the maximum reported.
"""
import math

import numpy as np
import pytest

import blacktracer_amd as bt

pytestmark = pytest.mark.gpu

ARCSEC = math.pi / 180 / 3600
# Bounds on mixed-vs-FP64 great-circle error (arcsec) over integrated rays.
# Measured on MI355X over nine precision-config poses (a in {0.5, 0.9,
# 0.999}, r in {3, 5, 20}): mean 0.014-0.070, p99 0.055-0.33; the bounds
# leave a factor of about two.
MEAN_BOUND = 0.15
P99_BOUND = 0.6


def _lattice(L):
    N = 2 ** L + 1
    M = 2 * (N - 1)
    i, j = np.meshgrid(np.arange(N), np.arange(M), indexing="ij")
    return (i / (N - 1) * math.pi).ravel(), (j / M * 2 * math.pi).ravel()


def _gc(t1, p1, t2, p2):
    h = np.sin((t2 - t1) / 2) ** 2 + np.sin(t1) * np.sin(t2) * np.sin((p2 - p1) / 2) ** 2
    return 2 * np.arcsin(np.sqrt(np.clip(h, 0, 1)))


POSES = [(0.5, 3.0), (0.999, 5.0), (0.9, 20.0)]


@pytest.mark.parametrize("a,r", POSES)
def test_mixed_vs_fp64_error_and_flags(ctx, a, r):
    print(f"mixed vs fp64 a={a} r={r}", flush=True)
    cam = bt.Camera(math.pi / 2, 0.0, r, black_hole=bt.BlackHole(a))
    t, p = _lattice(8)
    t64, p64, s64 = bt.trace_rays(cam, t, p, ctx=ctx)
    tm, pm, sm, stm = bt.trace_rays(cam, t, p, params=bt.IntegParams(bt.BT_PREC_MIXED), ctx=ctx,
                                    return_stats=True)
    cap64, capm = t64 == -1, tm == -1
    assert np.array_equal(cap64, capm)
    e = _gc(t64[~cap64], p64[~cap64], tm[~cap64], pm[~cap64]) / ARCSEC
    print(f"  arcsec: mean {e.mean():.3f} p99 {np.quantile(e, 0.99):.3f} max {e.max():.1f}; "
          f"fp64 share {stm['attempts_fp64'] / stm['attempts']:.3f}", flush=True)
    assert e.mean() < MEAN_BOUND
    assert np.quantile(e, 0.99) < P99_BOUND
    assert stm["attempts_fp64"] < stm["attempts"]


def test_mixed_all_critical_is_fp64(ctx):
    """crit_eps larger than any distance: every ray is near-critical and
    runs the exact FP64 path, so the mixed mode returns the FP64 path's
    results bit for bit."""
    cam = bt.Camera(math.pi / 2, 0.0, 5.0, black_hole=bt.BlackHole(0.999))
    t, p = _lattice(6)
    t64, p64, s64 = bt.trace_rays(cam, t, p, ctx=ctx)
    tm, pm, sm, st = bt.trace_rays(cam, t, p, params=bt.IntegParams(bt.BT_PREC_MIXED, crit_eps=1e30),
                                   ctx=ctx, return_stats=True)
    assert np.array_equal(t64, tm) and np.array_equal(p64, pm) and np.array_equal(s64, sm)
    assert st["attempts_fp64"] == st["attempts"]
    assert st["fallback_rays"] == st["integrated"]


def test_mixed_flagged_rays_run_the_fp64_path(ctx):
    """A noise budget below any attempt's noise flags every ray of the FP32
    bulk at its first FP32 attempt, and each then runs the exact FP64 path
    from its start: step counts and flags equal the FP64 path's, and so do
    the angles, bit for bit on all but the rays that end during the FP64
    warm-up (far from the hole, in the regrouped FP64 form), which agree to
    1e-9 rad."""
    cam = bt.Camera(math.pi / 2, 0.0, 5.0, black_hole=bt.BlackHole(0.999))
    t, p = _lattice(6)
    t64, p64, s64 = bt.trace_rays(cam, t, p, ctx=ctx)
    tm, pm, sm, st = bt.trace_rays(cam, t, p, params=bt.IntegParams(bt.BT_PREC_MIXED, noise_budget=1e-30),
                                   ctx=ctx, return_stats=True)
    assert np.array_equal(s64, sm) and np.array_equal(t64 == -1, tm == -1)
    d = np.maximum(np.abs(t64 - tm), np.abs(p64 - pm))
    print(f"  bit-equal rays {np.mean(d == 0):.4f}, max |d angle| {d.max():.3g} rad", flush=True)
    assert d.max() < 1e-9
    assert st["fallback_rays"] > 0


def _polar_directions(ps, targets):
    """Camera-sky directions whose b is each of targets (and its negative):
    the zero of b along phi at several polar angles (bisection on the
    oracle's initial conditions), then a first-order step to each target."""
    from oracle import oracle as O

    def b_of(th, ph):
        return O.initial_conditions_batch(ps, np.array([th]), np.array([ph]))[0, 2]

    th_out, ph_out = [], []
    for thd in np.linspace(0.6, 2.6, 8):
        grid = np.linspace(0.05, math.pi - 0.05, 121)
        bs = O.initial_conditions_batch(ps, np.full(grid.size, thd), grid)[:, 2]
        k = int(np.nonzero(np.diff(np.sign(bs)) != 0)[0][0])
        lo, hi = grid[k], grid[k + 1]
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            if np.sign(b_of(thd, mid)) == np.sign(b_of(thd, lo)):
                lo = mid
            else:
                hi = mid
        ph0 = 0.5 * (lo + hi)
        slope = (b_of(thd, ph0 + 1e-6) - b_of(thd, ph0 - 1e-6)) / 2e-6
        for t in targets:
            for sg in (1, -1):
                th_out.append(thd)
                ph_out.append(ph0 + sg * t / slope)
    return np.array(th_out), np.array(ph_out)


def test_mixed_polar_rays_run_the_fp64_path(ctx):
    """Rays with |b| below polar_b pass within about |b| / sqrt(q) of a pole,
    where the FP32 error estimate is hypersensitive to the step size; they
    restart in the exact FP64 path and end as there (bit for bit, or within
    1e-9 rad for rays that end in the FP64 warm-up)."""
    from oracle import oracle as O
    cam = bt.Camera(math.pi / 2, 0.0, 5.0, black_hole=bt.BlackHole(0.999))
    ps = O.pose(0.999, math.pi / 2, 0.0, 5.0)
    t, p = _polar_directions(ps, (1e-7, 1e-5, 3e-4))
    b = O.initial_conditions_batch(ps, t, p)[:, 2]
    assert np.all(np.abs(b) < 1e-3) and np.all(np.abs(b) > 0)
    t64, p64, s64 = bt.trace_rays(cam, t, p, ctx=ctx)
    tm, pm, sm, st = bt.trace_rays(cam, t, p, params=bt.IntegParams(bt.BT_PREC_MIXED), ctx=ctx,
                                   return_stats=True)
    assert np.array_equal(s64, sm) and np.array_equal(t64 == -1, tm == -1)
    d = np.maximum(np.abs(t64 - tm), np.abs(p64 - pm))
    print(f"  {t.size} polar rays (|b| 1e-7 .. 3e-4): bit-equal {np.mean(d == 0):.3f}, max |d angle| "
          f"{d.max():.3g} rad, fallback {st['fallback_rays']}", flush=True)
    assert d.max() < 1e-9


def test_mixed_attempt_cap_restarts_in_fp64(ctx):
    """Rays of the FP32 bulk that near the attempt cap restart in FP64, so a
    capped ray returns the FP64 path's sentinel (the camera's own angles).
    crit_eps = 0 sends every ray through the FP32 bulk."""
    cam = bt.Camera(math.pi / 2, 0.0, 5.0, black_hole=bt.BlackHole(0.999))
    t, p = _lattice(6)
    prm64 = bt.IntegParams(bt.BT_PREC_FP64, max_steps=40)
    prmm = bt.IntegParams(bt.BT_PREC_MIXED, max_steps=40, crit_eps=0.0)
    t64, p64, s64 = bt.trace_rays(cam, t, p, params=prm64, ctx=ctx)
    tm, pm, sm, st = bt.trace_rays(cam, t, p, params=prmm, ctx=ctx, return_stats=True)
    capped = s64 == 39
    assert capped.sum() > 10
    assert np.array_equal(t64[capped], tm[capped]) and np.array_equal(p64[capped], pm[capped])
    assert st["fallback_rays"] >= capped.sum()


def test_mixed_below_fp32_tolerance_runs_fp64(ctx):
    """eps below what FP32 step control can carry (1e-6): the mixed mode
    traces every ray in FP64, identical to the FP64 path, and counts them
    all as fallback rays."""
    cam = bt.Camera(math.pi / 2, 0.0, 5.0, black_hole=bt.BlackHole(0.999))
    t, p = _lattice(5)
    for eps in (1e-8, 1e-12):
        print(f"eps={eps}", flush=True)
        t64, p64, s64, st64 = bt.trace_rays(cam, t, p, params=bt.IntegParams(bt.BT_PREC_FP64, eps=eps),
                                            ctx=ctx, return_stats=True)
        tm, pm, sm, stm = bt.trace_rays(cam, t, p, params=bt.IntegParams(bt.BT_PREC_MIXED, eps=eps),
                                        ctx=ctx, return_stats=True)
        assert np.array_equal(t64, tm) and np.array_equal(p64, pm) and np.array_equal(s64, sm)
        assert stm["attempts_fp64"] == stm["attempts"] == st64["attempts"]
        assert stm["fallback_rays"] == stm["integrated"] > 0


def test_mixed_grid_build(ctx):
    """The adaptive grid built with the mixed path: same layout, capture
    pattern equal wherever both grids hold a point."""
    bh = bt.BlackHole(0.999)
    cam = bt.Camera(math.pi / 2, 0.0, 5.0, black_hole=bh)
    g64 = bt.Grid(8, 1, True, cam, bh, ctx=ctx)
    gm = bt.Grid(8, 1, True, cam, bh, params=bt.IntegParams(bt.BT_PREC_MIXED), ctx=ctx)
    assert (g64.N, g64.M) == (gm.N, gm.M)
    both = (g64.cel[..., 0] != -2) & (gm.cel[..., 0] != -2)
    c64, cm = g64.cel[both], gm.cel[both]
    assert np.array_equal(c64[:, 0] == -1, cm[:, 0] == -1)
    assert abs(g64.rays_traced - gm.rays_traced) <= 0.01 * g64.rays_traced
