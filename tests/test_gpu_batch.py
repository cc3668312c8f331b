"""Batched grid build (bt_grid_build_batch): every grid of a multi-pose
batch equals the same grid built alone, bit for bit (FP64: same kernels,
same arithmetic per ray; only the scheduling of rays differs)."""
import math

import numpy as np
import pytest

import blacktracer_amd as bt

pytestmark = pytest.mark.gpu

POSES = [  # (a, r, theta, max_level, start_level)
    (0.999, 5.0, math.pi / 2, 8, 1),
    (0.3, 12.0, 1.1, 6, 1),
    (0.9, 3.5, 2.0, 7, 2),
    (0.6, 40.0, 0.3, 6, 6),       # uniform grid (start == max)
    (0.05, 8.0, math.pi / 2, 9, 1),
]


def _grids(poses, params=None, ctx=None):
    cams, bhs = [], []
    for a, r, th, _, _ in poses:
        bh = bt.BlackHole(a)
        bhs.append(bh)
        cams.append(bt.Camera(th, 0.0, r, black_hole=bh))
    return cams, bhs


def test_batch_equals_single_builds(ctx):
    cams, bhs = _grids(POSES)
    grids, st = bt.Grid.build_batch(cams, bhs, [p[3] for p in POSES], [p[4] for p in POSES], ctx=ctx)
    total = 0
    for k, (cam, bh, p) in enumerate(zip(cams, bhs, POSES)):
        print(f"batch grid {k} pose={p}", flush=True)
        single = bt.Grid(p[3], p[4], True, cam, bh, ctx=ctx)
        g = grids[k]
        assert (g.N, g.M) == (single.N, single.M)
        assert g.rays_traced == single.rays_traced
        assert np.array_equal(g.cel, single.cel)
        assert np.array_equal(g.levels, single.levels)
        assert np.array_equal(g.steps, single.steps)
        total += g.rays_traced
    assert st["rays"] == total


def test_batch_mixed_precision(ctx):
    cams, bhs = _grids(POSES[:3])
    prm = bt.IntegParams(bt.BT_PREC_MIXED)
    grids, st = bt.Grid.build_batch(cams, bhs, [p[3] for p in POSES[:3]], [p[4] for p in POSES[:3]],
                                    params=prm, ctx=ctx)
    assert st["attempts_fp64"] < st["attempts"]
    for k, (cam, bh, p) in enumerate(zip(cams, bhs, POSES[:3])):
        single = bt.Grid(p[3], p[4], True, cam, bh, params=prm, ctx=ctx)
        # same rays, same per-ray arithmetic: the mixed batch equals the mixed single build
        assert np.array_equal(grids[k].cel, single.cel)


HALF_SKY = [  # cameras in the equatorial plane, moving within it: (a, r, motion, max_level, start_level)
    (0.999, 5.0, dict(), 9, 1),
    (0.5, 2.5, dict(br=0.6, btheta=0.0, bphi=0.8), 8, 1),
    (0.3, 15.0, dict(speed=0.2), 7, 7),
]
MOVING = [  # (a, r, theta, motion, max_level, start_level)
    (0.9, 4.0, 1.2, dict(br=-0.3, btheta=0.5, bphi=math.sqrt(0.66)), 8, 2),
    (0.6, 9.0, 0.7, dict(br=0.0, btheta=-0.6, bphi=0.8), 7, 1),
]


@pytest.mark.parametrize("angle", [False, True])
@pytest.mark.parametrize("prec", [bt.BT_PREC_FP64, bt.BT_PREC_MIXED])
def test_batch_half_sky_and_moving_cameras(ctx, angle, prec):
    """Half-sky grids (angle = false) and cameras moving along another
    direction than the orbit, or at a caller's speed: every grid of the
    batch equals its single build bit for bit."""
    if angle:
        specs = [(a, r, th, m, ml, sl) for a, r, th, m, ml, sl in MOVING]
    else:
        specs = [(a, r, math.pi / 2, m, ml, sl) for a, r, m, ml, sl in HALF_SKY]
    bhs = [bt.BlackHole(s[0]) for s in specs]
    cams = [bt.Camera(th, 0.0, r, black_hole=bh, **m) for (a, r, th, m, _, _), bh in zip(specs, bhs)]
    prm = bt.IntegParams(prec)
    grids, st = bt.Grid.build_batch(cams, bhs, [s[4] for s in specs], [s[5] for s in specs], angle=angle,
                                    params=prm, ctx=ctx)
    for k, (cam, bh, s) in enumerate(zip(cams, bhs, specs)):
        print(f"batch angle={angle} prec={prec} grid {k} spec={s}", flush=True)
        single = bt.Grid(s[4], s[5], angle, cam, bh, params=prm, ctx=ctx)
        g = grids[k]
        assert (g.N, g.M, g.equafactor) == (single.N, single.M, 1 if angle else 0)
        assert g.rays_traced == single.rays_traced
        assert np.array_equal(g.cel, single.cel)
        assert np.array_equal(g.levels, single.levels)
        single.close()
        g.close()


@pytest.mark.parametrize("angle", [True, False])
def test_batch_tiny_grids(ctx, angle):
    """Grids smaller than one workgroup's share of the set-up launches
    (levels 1-3, a few to a few hundred points) beside a larger one: each
    equals its single build."""
    specs = [(1, 1), (2, 1), (2, 2), (3, 1), (3, 3), (7, 1), (1, 1)]
    bhs = [bt.BlackHole(0.9) for _ in specs]
    cams = [bt.Camera(math.pi / 2 if not angle else 1.0 + 0.1 * k, 0.0, 4.0 + k, black_hole=bh)
            for k, bh in enumerate(bhs)]
    grids, st = bt.Grid.build_batch(cams, bhs, [s[0] for s in specs], [s[1] for s in specs], angle=angle, ctx=ctx)
    for k, (cam, bh, s) in enumerate(zip(cams, bhs, specs)):
        print(f"tiny batch angle={angle} grid {k} Grid{s}", flush=True)
        single = bt.Grid(s[0], s[1], angle, cam, bh, ctx=ctx)
        g = grids[k]
        assert (g.N, g.M) == (single.N, single.M)
        assert g.rays_traced == single.rays_traced
        assert np.array_equal(g.cel, single.cel)
        assert np.array_equal(g.levels, single.levels)
        single.close()
        g.close()


def test_batch_rejects_bad_sizes(ctx):
    cams, bhs = _grids(POSES[:1])
    with pytest.raises(bt.BtError):
        bt.Grid.build_batch(cams, bhs, [13], [1], ctx=ctx)
