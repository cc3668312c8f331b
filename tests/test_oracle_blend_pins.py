"""Pins of the oracle's multi-grid blending (the reference's G > 1 frame,
MakeImageKernel.cu:386-510), on synthetic whole-sky grids with a captured
disk of known radius.

The reference's blending exists only as CUDA code and has no tests or
fixtures; these checks pin it by geometry and by the limits the blend must
reduce to:
  * the shadow-border walk lands on the disk's edge in every direction
    (a sign, axis or grid swap moves it by several cells);
  * blending a grid with itself gives the one-grid frame;
  * alpha = 0 gives grid A's frame and alpha = 1 grid B's, wherever the
    blend is not in the shadow.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

GN, GM, LVL = 129, 256, 7   # Grid(7, ., angle): N = 2^7 + 1, M = 2^8
CI, CJ = 64, 128            # disk centre (grid rows, columns)


def disk_grid(radius, tilt):
    """Dense grid: smooth celestial angles, (-1, -1) inside the disk."""
    i = np.arange(GN, dtype=np.float64)[:, None]
    j = np.arange(GM, dtype=np.float64)[None, :]
    th = 0.4 + 2.2 * i / (GN - 1) + tilt * np.sin(2 * np.pi * j / GM)
    ph = np.mod(0.3 + 2 * np.pi * j / GM + tilt * np.cos(np.pi * i / (GN - 1)), 2 * np.pi)
    g = np.stack([th, ph], axis=-1).astype(np.float32)
    inside = (i - CI) ** 2 + (j - CJ) ** 2 < radius ** 2
    g[inside] = -1.0
    return g


def thphi_of(grid_a, grid_b=None, alpha=0.0, w=128, h=64):
    view = O.sphere_view(w, h)
    tree, pos, mag = O.make_tree(np.array([[1.0, 1.0, 5.0, 0.5]], np.float32), 2)
    table = O.sum_table(np.zeros((8, 16, 3), np.uint8))
    cam7 = [1.0, 0.9, 0.01, 4.0, 0.0, 0.0, 1.0]
    if grid_b is None:
        _, st = O.frame(grid_a, LVL, view, cam7, tree, 2, pos, mag, table, star=False, map=False,
                        lr=False, stages=True)
        return st["thphi"]
    _, st = O.frame_blend(grid_a, grid_b, alpha, LVL, view, cam7 + cam7, tree, 2, pos, mag, table,
                          star=False, map=False, lr=False, stages=True)
    return st["thphi"]


def outside_stretch(grid_a, grid_b, w=128, h=64):
    """Corners outside the circle in which the blend warps the grids."""
    b = O.blend_border(grid_a, grid_b).astype(np.float64)
    c = np.array([0.5 * (b[0, 0] + b[0, 1]), 0.5 * (b[1, 0] + b[1, 1])])
    rad = max(b[0, 1] - b[0, 0], b[1, 0] - b[1, 1]) * 0.75
    v = O.sphere_view(w, h).astype(np.float64)
    return (v[..., 0] - c[0]) ** 2 + (v[..., 1] - c[1]) ** 2 > rad * rad * 1.01


def test_border_walk_lands_on_disk_edge():
    ra, rb = 20.0, 14.0
    border = O.blend_border(disk_grid(ra, 0.05), disk_grid(rb, 0.02))
    # findBhCenter: extremes of pi/GN * (row, column) over captured points
    cap = np.zeros((GN, GM), bool)
    for r in (ra, rb):
        i, j = np.nonzero((np.arange(GN)[:, None] - CI) ** 2 + (np.arange(GM)[None, :] - CJ) ** 2 < r ** 2)
        cap[i, j] = True
    i, j = np.nonzero(cap)
    gs = np.float32(math.pi) / np.float32(GN)
    assert border[0, 0] == gs * np.float32(i.min()) and border[0, 1] == gs * np.float32(i.max())
    assert border[1, 0] == gs * np.float32(j.min()) and border[1, 1] == gs * np.float32(j.max())
    # walk start, in grid cells (the reference's centre, converted with GM)
    c0 = 0.5 * (border[0, 0] + border[0, 1]) / (2 * math.pi) * GM
    c1 = 0.5 * (border[1, 0] + border[1, 1]) / (2 * math.pi) * GM
    cell = 2 * math.pi / GM
    for s in range(0, 1000, 7):
        a = 2 * math.pi * s / 1000
        d = np.array([-math.sin(a), math.cos(a)])
        for slot, r in ((0, ra), (1, rb)):
            # exit of the ray start + t d from the disk
            p0 = np.array([c0 - CI, c1 - CJ])
            b = p0 @ d
            t = -b + math.sqrt(b * b - (p0 @ p0 - r * r))
            want = (np.array([c0, c1]) + t * d) * cell
            got = border[2 + 2 * s + slot].astype(np.float64)
            assert np.abs(got - want).max() < 2.0 * cell, (s, slot, got / cell, want / cell)


def test_blend_of_a_grid_with_itself_is_the_one_grid_frame():
    g = disk_grid(18.0, 0.05)
    one = thphi_of(g)
    both = thphi_of(g, g, alpha=0.4)
    vis = (both[..., 0] != -1)
    assert vis.mean() > 0.8
    assert np.array_equal(one[vis] == -1, np.zeros_like(one[vis], bool))
    assert np.abs(both[vis] - one[vis]).max() < 2e-5
    # the blended shadow is the shadow of the grid, up to the border band
    assert ((both[..., 0] == -1) != (one[..., 0] == -1)).mean() < 0.01


@pytest.mark.parametrize("alpha", [0.0, 1.0])
def test_blend_endpoints_are_the_two_grids(alpha):
    ga, gb = disk_grid(20.0, 0.05), disk_grid(14.0, -0.04)
    blend = thphi_of(ga, gb, alpha=alpha)
    ref = thphi_of(ga if alpha == 0.0 else gb)
    vis = blend[..., 0] != -1
    assert vis.mean() > 0.8
    d = np.abs(blend[vis] - ref[vis])
    # theta exactly; phi up to the 2 pi unwrapping round trip
    assert np.array_equal(blend[vis][:, 0], ref[vis][:, 0])
    assert d[:, 1].max() < 1e-6
    # the other grid only adds shadow: every visible blended corner is visible there too
    assert (ref[vis][:, 0] != -1).all()


def test_blend_moves_between_the_grids():
    ga, gb = disk_grid(20.0, 0.05), disk_grid(14.0, -0.04)
    ta, tb = thphi_of(ga), thphi_of(gb)
    prev = None
    for alpha in (0.25, 0.5, 0.75):
        t = thphi_of(ga, gb, alpha=alpha)
        vis = (t[..., 0] != -1) & (ta[..., 0] != -1) & (tb[..., 0] != -1)
        # far from the shadow (no warp): theta is the linear blend of the two grids' values
        far = vis & outside_stretch(ga, gb)
        assert far.mean() > 0.5
        want = (1 - alpha) * ta[far][:, 0].astype(np.float64) + alpha * tb[far][:, 0]
        assert np.abs(t[far][:, 0] - want).max() < 1e-5
        # the shadow shrinks as the blend moves towards the smaller disk
        sh = (t[..., 0] == -1).sum()
        if prev is not None:
            assert sh <= prev
        prev = sh
