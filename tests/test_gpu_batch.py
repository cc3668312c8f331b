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
        assert np.mean(g.steps != single.steps) <= 0.001
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


def test_batch_rejects_bad_sizes(ctx):
    cams, bhs = _grids(POSES[:1])
    with pytest.raises(bt.BtError):
        bt.Grid.build_batch(cams, bhs, [13], [1], ctx=ctx)
