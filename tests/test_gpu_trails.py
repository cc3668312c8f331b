"""Parity of frame sequences with star trails and diffraction with the CPU
oracle.

Bar, per frame: trail buffer and summed star light within 1e-4 relative
(the device evaluates the redshift magnitude with the FP32 library log10 and
reassociates the tap sums); frames within 1/255 per channel with at most
0.1% of pixels differing.
"""
import math

import numpy as np
import pytest

import blacktracer_amd as bt
from blacktracer_amd import scene
from oracle import oracle as O

pytestmark = pytest.mark.gpu

W, H = 320, 180


@pytest.fixture(scope="module")
def world(ctx):
    bh = bt.BlackHole(0.999)
    cam = bt.Camera(math.pi / 2, 0.0, 5.0, black_hole=bh)
    grid = bt.Grid(8, 1, True, cam, bh, ctx=ctx)
    stars = scene.synthetic_stars(20000, 21)
    sky = scene.synthetic_sky(512, seed=5)
    sc = bt.Scene(stars, sky, tree_level=8, ctx=ctx)
    tree, pos, mag = O.make_tree(stars, 8)
    return dict(cam=cam, grid=grid, stars=stars, sky=sky, scene=sc, tree=tree, pos=pos, mag=mag,
                table=O.sum_table(sky))


def sprite():
    fs = W // 16
    y, x = np.mgrid[0:fs, 0:fs] - fs // 2
    s = np.zeros((fs, fs, 3), np.uint8)
    spike = (np.abs(x) <= 0) | (np.abs(y) <= 0)
    s[spike] = (255, 220, 180)
    s[(x * x + y * y) <= 4] = (120, 120, 120)
    return s


def compare_frames(a, b, max_frac=1e-3):
    d = np.abs(a.astype(np.int16) - b.astype(np.int16))
    assert d.max() <= 1, f"max channel diff {d.max()} at {np.unravel_index(d.argmax(), d.shape)}"
    assert (d.max(axis=-1) > 0).mean() <= max_frac


@pytest.mark.parametrize("diffraction", [False, True])
def test_trail_sequence_matches_oracle(ctx, world, diffraction):
    view = bt.View(W, H, ctx=ctx)
    oview = O.camera_view(math.pi / 2, 0.0, 0.25 * math.pi / 4, W, H)
    sp = sprite() if diffraction else None
    T = bt.Trails(world["scene"], W, H, diffraction=sp, ctx=ctx)
    OT = O.Trails(world["pos"].shape[0], W, H, sp)
    gridf = world["grid"].cel.astype(np.float32)
    cam7 = world["cam"].getParamArray()
    drew = 0
    for k in range(3):
        off = k * 3 * 2 * math.pi / W
        print(f"trails frame {k} offset {off:.4f} diffraction={diffraction}", flush=True)
        f = bt.RenderFlags(offset=off)
        img, st = T.render(world["grid"], world["scene"], view, flags=f, stages=True)
        oimg, olight, otrail = OT.frame(gridf, world["grid"].MAXLEVEL, oview, cam7, world["tree"], 8,
                                        world["pos"], world["mag"], world["table"], offset=off)
        assert np.allclose(st["trail"], otrail, rtol=1e-4, atol=1e-3)
        assert np.allclose(st["star"], olight, rtol=1e-4, atol=1e-3, equal_nan=True)
        compare_frames(img, oimg)
        if k == 0:
            assert st["trail"].max() == 0
        drew += int((st["trail"][..., 0] > 0).sum())
    assert drew > 50
    assert T.stats() == dict(frames=3, dropped=0)
    T.close()
    OT.close()


def test_trails_reset_and_bad_state(ctx, world):
    view = bt.View(W, H, ctx=ctx)
    T = bt.Trails(world["scene"], W, H, ctx=ctx)
    f0 = T.render(world["grid"], world["scene"], view)
    T.render(world["grid"], world["scene"], view, flags=bt.RenderFlags(offset=0.1))
    T.reset()
    again = T.render(world["grid"], world["scene"], view)
    assert np.array_equal(f0, again)
    # without trails the first frame is the plain frame
    plain = bt.render_frame(world["grid"], world["scene"], view)
    assert np.array_equal(plain, f0)
    small = bt.View(W // 2, H // 2, ctx=ctx)
    with pytest.raises(bt.BtError):
        T.render(world["grid"], world["scene"], small)
    T.close()
    # the sprite is (width / 16)^2 pixels: no sprite fits a frame narrower than 16
    with pytest.raises((bt.BtError, ValueError)):
        bt.Trails(world["scene"], 8, 8, diffraction=np.zeros((1, 1, 3), np.uint8), ctx=ctx)
