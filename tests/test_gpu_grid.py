"""Parity of the device-resident adaptive grid build (bt_grid_build) with
the CPU oracle: block levels / refinement set bit-exact, traced-point set
bit-exact, step counts equal, celestial angles within 1e-9 rad, and the
device fixTvertices bit-exact against the oracle's pinned order on the same
pre-fix values."""
import math

import numpy as np
import pytest

import blacktracer_amd as bt
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cmp_cel(a, b, present):
    m = present & (a[..., 0] != -1)
    dt = np.abs(a[..., 0] - b[..., 0])[m]
    dp = np.abs(a[..., 1] - b[..., 1])[m]
    dp = np.minimum(dp, np.abs(dp - 2 * math.pi))
    return max(dt.max(initial=0), dp.max(initial=0))


@pytest.mark.parametrize("pose,lv,start,angle", [
    ((0.999, math.pi / 2, 0.0, 5.0), 8, 8, True),     # C0: uniform 257x513 sky
    ((0.999, math.pi / 2, 0.0, 5.0), 9, 1, True),     # C1 pose, adaptive
    ((0.6, 1.0, 0.3, 4.0), 8, 1, True),
    ((0.9, 0.3, 0.0, 20.0), 7, 2, True),
    ((0.999, math.pi / 2, 0.0, 5.0), 9, 1, False),    # half sky (camera in the
    ((0.5, math.pi / 2, 0.0, 12.0), 8, 8, False),     # equatorial plane)
    ((0.9, 1.0, 0.0, 4.0), 1, 1, True),               # the smallest grids
    ((0.9, math.pi / 2, 0.0, 4.0), 2, 1, False),
    ((0.9, 1.2, 0.0, 4.0), 3, 3, True),
])
def test_grid_matches_oracle(ctx, pose, lv, start, angle):
    print(f"grid pose={pose} Grid({lv},{start},angle={angle})", flush=True)
    a, th, ph, r = pose
    bh = bt.BlackHole(a)
    cam = bt.Camera(th, ph, r, black_hole=bh)
    g = bt.Grid(lv, start, angle, cam, bh, ctx=ctx)
    og = O.OracleGrid(O.pose(a, th, ph, r), O.params(), lv, start, angle)
    assert g.equafactor == (1 if angle else 0)
    assert (g.N, g.M) == (og.N, og.M)
    assert g.rays_traced == og.rays_traced
    assert np.array_equal(g.levels, og.level)
    present = g.cel[..., 0] != -2
    assert np.array_equal(present, og.present == 1)
    assert np.array_equal(g.steps, og.steps), f"{np.sum(g.steps != og.steps)} step counts differ"
    err = _cmp_cel(g.cel, og.cel, present)
    print(f"  rays {g.rays_traced}, max |d angle| {err:.3g}", flush=True)
    assert err < 1e-9


def _pole_tvertices(levels, present, max_level, start_level):
    """Meridian T-vertices in the first row of blocks (rows 0..gap): the
    repairs that interleave with the row-0 top edges in the pinned order."""
    n = 0
    M = levels.shape[1]
    for L in range(start_level, max_level):
        ug = 1 << (max_level - L)
        for j in range(0, M, ug):
            if levels[0, j] == L:
                n += int(present[ug // 2, j]) + int(present[ug // 2, (j + ug) % M])
    return n


@pytest.mark.parametrize("pose,lv,pole,angle", [
    ((0.999, math.pi / 2, 0.0, 5.0), 9, False, True),
    ((0.999, 0.3, 0.0, 3.0), 8, True, True),     # close cameras: T-vertices on the
    ((0.5, math.pi / 2, 0.0, 2.5), 8, True, True),   # meridians next to the north pole
    ((0.5, 0.3, 0.0, 3.0), 8, True, True),
    ((0.999, math.pi / 2, 0.0, 5.0), 10, False, False),   # half sky
    ((0.5, math.pi / 2, 0.0, 2.5), 9, True, False),
])
def test_fix_tvertices_bit_exact(ctx, pose, lv, pole, angle):
    """Device line walkers == oracle's sequential pinned order, on identical
    pre-fix values (the device grid built without the fix)."""
    print(f"fix_tvertices pose={pose} Grid({lv},1,angle={angle})", flush=True)
    a, th, ph, r = pose
    bh = bt.BlackHole(a)
    cam = bt.Camera(th, ph, r, black_hole=bh)
    g_nofix = bt.Grid(lv, 1, angle, cam, bh, fix_tvertices=False, ctx=ctx)
    g_fix = bt.Grid(lv, 1, angle, cam, bh, fix_tvertices=True, ctx=ctx)
    og = O.OracleGrid(O.pose(a, th, ph, r), O.params(), lv, 1, angle, fix_tvertices=False)
    if pole:
        assert _pole_tvertices(g_nofix.levels, g_nofix.cel[..., 0] != -2, lv, 1) > 0
    N, M = og.N, og.M
    cel = np.ctypeslib.as_array(og.g.cel, shape=(N, M, 2))
    cel[...] = g_nofix.cel
    np.ctypeslib.as_array(og.g.level, shape=(N, M))[...] = g_nofix.levels
    og.fix_tvertices()
    changed = np.any(g_nofix.cel != g_fix.cel, axis=-1).sum()
    assert changed > 100
    assert np.array_equal(og.cel, g_fix.cel)


@pytest.mark.parametrize("pose,lv,start", [((0.999, math.pi / 2, 0.0, 5.0), 10, 1),
                                           ((0.3, math.pi / 2, 0.0, 30.0), 11, 11)])
def test_half_sky_grid_is_upper_half_of_full_grid(ctx, pose, lv, start):
    """At sizes beyond the oracle: the half-sky grid (angle = false) equals
    the full grid's rows above the equator bit for bit before fixTvertices
    (same ray directions, and a block's refinement reads only its corners;
    Grid.h:291-333,605-615)."""
    print(f"half vs full pose={pose} Grid({lv},{start})", flush=True)
    a, th, ph, r = pose
    bh = bt.BlackHole(a)
    cam = bt.Camera(th, ph, r, black_hole=bh)
    h = bt.Grid(lv, start, False, cam, bh, fix_tvertices=False, ctx=ctx)
    f = bt.Grid(lv, start, True, cam, bh, fix_tvertices=False, ctx=ctx)
    n = h.N
    assert (h.N, h.M, f.N) == (2 ** (lv - 1) + 1, f.M, 2 * n - 1)
    assert np.array_equal(h.levels[:n - 1], f.levels[:n - 1])
    assert np.array_equal(h.cel[:n - 1], f.cel[:n - 1])
    assert np.array_equal(h.steps[:n - 1], f.steps[:n - 1])
    eq = h.cel[n - 1, :, 0] != -2
    assert np.array_equal(h.cel[n - 1][eq], f.cel[n - 1][eq])
    h.close()
    f.close()


def test_grid_psh_roundtrip(ctx):
    """saveAsGpuHash: every traced key is found with its float2 value; absent
    keys read (-2,-2) (kernel.cuh:705-733)."""
    bh = bt.BlackHole(0.999)
    cam = bt.Camera(math.pi / 2, 0.0, 5.0, black_hole=bh)
    g = bt.Grid(7, 1, True, cam, bh, ctx=ctx)
    ht, ot, tag, hw, ow = g.psh()
    cel = g.cel
    N, M = g.N, g.M
    ii, jj = np.meshgrid(np.arange(N), np.arange(M), indexing="ij")
    ix, iy = (ii + ow) % ow, (jj + ow) % ow
    off = ot.reshape(ow, ow, 2)[ix, iy]
    hx = ((ii + hw) % hw + off[..., 0] + hw) % hw
    hy = ((jj + hw) % hw + off[..., 1] + hw) % hw
    t = tag.reshape(hw, hw, 2)[hx, hy]
    found = (t[..., 0] == ii) & (t[..., 1] == jj)
    present = cel[..., 0] != -2
    assert np.array_equal(found, present)
    vals = ht.reshape(hw, hw, 2)[hx, hy]
    assert np.array_equal(vals[present], cel[present].astype(np.float32))
