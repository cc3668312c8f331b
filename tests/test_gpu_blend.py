"""Parity of the blended (G > 1) frame with the CPU oracle.

Bar: shadow border and per-corner celestial angles bit-exact; frames within
1/255 per channel with at most 0.1% of pixels differing.
"""
import math

import numpy as np
import pytest

import blacktracer_amd as bt
from blacktracer_amd import scene
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world(ctx):
    bh = bt.BlackHole(0.999)
    cams = [bt.Camera(math.pi / 2, 0.0, 5.0, black_hole=bh), bt.Camera(math.pi / 2, 0.0, 6.5, black_hole=bh)]
    grids = [bt.Grid(8, 1, True, c, bh, ctx=ctx) for c in cams]
    stars = scene.synthetic_stars(20000, 21)
    sky = scene.synthetic_sky(512, seed=5)
    sc = bt.Scene(stars, sky, tree_level=8, ctx=ctx)
    return dict(bh=bh, cams=cams, grids=grids, stars=stars, sky=sky, scene=sc)


def compare_frames(a, b, max_frac=1e-3):
    d = np.abs(a.astype(np.int16) - b.astype(np.int16))
    assert d.max() <= 1, f"max channel diff {d.max()} at {np.unravel_index(d.argmax(), d.shape)}"
    assert (d.max(axis=-1) > 0).mean() <= max_frac


@pytest.mark.parametrize("alpha,sphere", [(0.0, False), (0.35, False), (1.0, False), (0.6, True)])
def test_blend_matches_oracle(ctx, world, alpha, sphere):
    W, H = 320, 180
    print(f"blend {W}x{H} alpha={alpha} sphere={sphere}", flush=True)
    f = bt.RenderFlags()
    view = bt.View(W, H, sphere=sphere, ctx=ctx)
    g0, g1 = world["grids"]
    img, st = bt.render_frame_blend(g0, g1, alpha, world["scene"], view, flags=f, stages=True)
    if sphere:
        oview = O.sphere_view(W, H)
    else:
        oview = O.camera_view(math.pi / 2, 0.0, 0.25 * math.pi / 4, W, H)
    tree, pos, mag = O.make_tree(world["stars"], 8)
    table = O.sum_table(world["sky"])
    cam14 = list(world["cams"][0].getParamArray()) + list(world["cams"][1].getParamArray())
    oimg, ost = O.frame_blend(g0.cel.astype(np.float32), g1.cel.astype(np.float32), alpha, g0.MAXLEVEL,
                              oview, cam14, tree, 8, pos, mag, table, star=f.star, map=f.map, lr=f.lr,
                              hor=f.hor, ver=f.ver, offset=f.offset, stages=True)
    assert np.array_equal(st["border"], ost["border"])
    assert np.array_equal(st["thphi"], ost["thphi"])
    compare_frames(img, oimg)
    assert img[..., 3].min() == 255
    if sphere:
        # the whole sky is in view: the warp inside the stretch circle is exercised
        b = ost["border"].astype(np.float64)
        c = np.array([0.5 * (b[0, 0] + b[0, 1]), 0.5 * (b[1, 0] + b[1, 1])])
        rad = max(b[0, 1] - b[0, 0], b[1, 0] - b[1, 1]) * 0.75
        v = oview.astype(np.float64)
        inside = (v[..., 0] - c[0]) ** 2 + (v[..., 1] - c[1]) ** 2 < rad * rad
        assert (inside & (ost["thphi"][..., 0] != -1)).sum() > 100


def test_blend_endpoint_equals_one_grid_frame_outside_shadow(ctx, world):
    W, H = 160, 90
    view = bt.View(W, H, sphere=True, ctx=ctx)
    g0, g1 = world["grids"]
    f = bt.RenderFlags(star=False, map=True, lr=False)
    _, st = bt.render_frame_blend(g0, g1, 0.0, world["scene"], view, flags=f, stages=True)
    _, one = bt.render_frame(g0, world["scene"], view, flags=f, stages=True)
    vis = st["thphi"][..., 0] != -1
    assert vis.mean() > 0.5
    assert np.array_equal(st["thphi"][vis][:, 0], one["thphi"][vis][:, 0])


def test_blend_rejects_bad_input(ctx, world):
    view = bt.View(32, 18, ctx=ctx)
    g0, g1 = world["grids"]
    with pytest.raises(bt.BtError):
        bt.render_frame_blend(g0, g1, 1.5, world["scene"], view)
    bh = world["bh"]
    other = bt.Grid(7, 1, True, world["cams"][1], bh, ctx=ctx)
    with pytest.raises(bt.BtError):
        bt.render_frame_blend(g0, other, 0.5, world["scene"], view)
    other.close()
