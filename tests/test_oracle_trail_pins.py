"""Pins of the oracle's star trails and diffraction (MakeImageKernel.cu:
98-174,517-537,618-629,775-800) on an unlensed sky: an identity grid
(celestial direction = camera direction) seen through the whole-sky view, so
that a star's image moves by a known number of pixels when the frame's
celestial offset or view offset changes.

The reference's trails and diffraction exist only as CUDA code with no tests
or fixtures; these checks pin them by geometry and by the conditions the
reference states:
  * a star moving along the celestial motion direction (the gradient field
    rotated by 90 degrees) by 2..M/25 pixels leaves a trail on the pixels
    between its two images, fading as (d_max - d)^2 / d_max^2 from the new
    image (brightness ratios between trail pixels are the squared ratios);
  * no trail on the first frame, for a move of one pixel, for a move beyond
    M/25, or for a move across the motion direction;
  * a zero diffraction sprite changes nothing; a uniform sprite adds light
    exactly on the sprite's window around each pixel brighter than 255^2.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

W, H = 200, 100
GN, GM, LVL = 129, 256, 7


def identity_grid():
    i = np.arange(GN, dtype=np.float64)[:, None] * np.ones((1, GM))
    j = np.ones((GN, 1)) * np.arange(GM, dtype=np.float64)[None, :]
    return np.stack([i * math.pi / (GN - 1), j * 2 * math.pi / GM], axis=-1).astype(np.float32)


def scene(bright_mag=3.0):
    rng = np.random.default_rng(7)
    n = 3000
    s = np.zeros((n, 4), np.float32)
    # faint filler stars away from the test star, so the tree's leaves are small
    s[:, 0] = rng.uniform(0.2, 1.0, n)
    s[:, 1] = rng.uniform(0.0, 2 * math.pi, n)
    s[:, 2] = 30.0
    s[:, 3] = 0.6
    # the test star: row 70 of 100, column about 100
    s[0] = [70.5 * math.pi / H, 100.3 * 2 * math.pi / W, bright_mag, 0.6]
    tree, pos, mag = O.make_tree(s, 8)
    return tree, pos, mag


def run(frames, bright_mag=3.0, sprite=None):
    """frames: list of dict(offset=, ver=); returns [(img, light, trail)]."""
    g = identity_grid()
    view = O.sphere_view(W, H)
    tree, pos, mag = scene(bright_mag)
    table = O.sum_table(np.zeros((8, 16, 3), np.uint8))
    cam7 = [0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 1.0]
    T = O.Trails(pos.shape[0], W, H, sprite)
    out = [T.frame(g, LVL, view, cam7, tree, 8, pos, mag, table, map=False, lr=False, **f) for f in frames]
    T.close()
    return out


def star_pixel(light):
    i, j = np.unravel_index(np.argmax(light[..., 0]), light.shape[:2])
    return int(i), int(j)


def trail_pixels(trail, floor=1e-6):
    return {tuple(int(v) for v in p) for p in np.argwhere(trail[..., 0] > floor)}


def test_no_trail_on_the_first_frame():
    (_, light, trail), = run([dict(offset=0.0)])
    assert light[..., 0].max() > 0
    assert trail.max() == 0


@pytest.mark.parametrize("move", [3, 5, 8])
def test_trail_along_the_motion(move):
    d = move * 2 * math.pi / W
    (_, l0, _), (_, l1, tr) = run([dict(offset=0.0), dict(offset=d)])
    i0, j0 = star_pixel(l0)
    i1, j1 = star_pixel(l1)
    assert (i1, j1) == (i0, j0 - move)
    # the pixels strictly between the two images, on the star's row
    want = {(i0, j) for j in range(j1 + 1, j0)}
    assert trail_pixels(tr) == want
    # fading: pixel at distance k from the new image has weight (move - k)^2 / move^2
    r2 = {k: float(tr[i0, j1 + k, 0]) for k in range(1, move)}
    for k in range(2, move):
        f1, fk = (move - 1) ** 2, (move - k) ** 2
        assert r2[k] / r2[1] == pytest.approx((fk / f1) ** 2, rel=1e-4)


@pytest.mark.parametrize("move", [1, 9])
def test_no_trail_for_tiny_or_long_moves(move):
    d = move * 2 * math.pi / W
    (_, l0, _), (_, l1, tr) = run([dict(offset=0.0), dict(offset=d)])
    i0, j0 = star_pixel(l0)
    assert star_pixel(l1) == (i0, j0 - move)
    assert trail_pixels(tr, 1e-12) == set()


def test_no_trail_across_the_motion():
    # a view offset in theta moves the image along the grid's theta gradient
    dv = 4 * math.pi / H
    (_, l0, _), (_, l1, tr) = run([dict(offset=0.0), dict(offset=0.0, ver=dv)])
    i0, j0 = star_pixel(l0)
    i1, j1 = star_pixel(l1)
    assert j1 == j0 and abs(i1 - i0) == 4
    assert trail_pixels(tr, 1e-12) == set()


def test_zero_sprite_changes_nothing():
    fs = W // 16
    frames = [dict(offset=0.0), dict(offset=5 * 2 * math.pi / W)]
    plain = run(frames, bright_mag=-2.0)
    zero = run(frames, bright_mag=-2.0, sprite=np.zeros((fs, fs, 3), np.uint8))
    for a, b in zip(plain, zero):
        assert np.array_equal(a[0], b[0])
        assert np.array_equal(a[1], b[1])


def test_uniform_sprite_covers_its_window():
    fs = W // 16
    fh = fs // 2
    (_, light0, _), = run([dict(offset=0.0)], bright_mag=-2.0)
    (_, light, _), = run([dict(offset=0.0)], bright_mag=-2.0, sprite=np.ones((fs, fs, 3), np.uint8))
    hot = np.argwhere(light0.max(axis=-1) > 65025.0)
    assert len(hot) >= 1
    cover = np.zeros((H, W), bool)
    for i, j in hot:
        for q in range(fs):
            r = i - fh + q
            if 0 <= r < H:
                cover[r, (j - fh + np.arange(fs)) % W] = True
    # outside every window nothing changes; inside, every dark pixel lights up
    # (on bright pixels the addition can be below their float spacing)
    assert np.array_equal(light[~cover], light0[~cover])
    dark = cover & (light0[..., 0] == 0)
    assert dark.sum() > fs * fs // 2
    assert (light[..., 0][dark] > 0).all()
