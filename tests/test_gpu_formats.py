"""The reference's caches through the device path: a grid written to a
.grid file and read back renders through makeImage exactly as the resident
grid; a scene made from a .star file renders exactly as the scene built from
the same stars; a .grid file has the bytes cereal itself writes for the
same tables (oracle/_ref/libbtcereal.so, built from the reference's vendored
cereal headers where the checkout exists).  The element layout of CUDA's
float2/int2 in those tables (x then y) is the one assumption: the
reference's tree has no serializer for them."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import blacktracer_amd as bt
from blacktracer_amd import scene

pytestmark = pytest.mark.gpu


def test_grid_file_renders_as_resident_grid(ctx, tmp_path):
    bh = bt.BlackHole(0.9)
    cam = bt.Camera(1.3, 0.0, 8.0, black_hole=bh)
    g = bt.Grid(7, 1, True, cam, bh, ctx=ctx)
    p = tmp_path / "pose.grid"
    g.write(p)
    d = bt.read_grid_file(p)
    assert (d["max_level"], d["N"], d["M"]) == (g.MAXLEVEL, g.N, g.M)
    ht, ot, tag, hw, ow = g.psh()
    assert (d["hw"], d["ow"]) == (hw, ow)
    assert np.array_equal(d["hash_table"], ht) and np.array_equal(d["offset_table"], ot)
    stars = scene.synthetic_stars(5000, 2)
    sky = scene.synthetic_sky(256, seed=3)
    W, H = 200, 120
    view = bt.viewer_matrix(math.pi / 2, 0.0, 0.25 * math.pi / 4, W, H)
    tree, pos, mag = bt.star_tree(stars, 8)
    img = bt.make_image(pos, tree, cam.getParamArray(), mag, 8, W, H, 1, sky, view, d["M"], d["N"],
                        d["max_level"], d["offset_table"], d["hash_table"], d["hash_pos_tag"], d["hw"], d["ow"],
                        ctx=ctx)
    sc = bt.Scene(stars, sky, tree_level=8, ctx=ctx)
    ref = bt.render_frame(g, sc, bt.View(W, H, ctx=ctx))
    assert np.array_equal(img, ref)
    # a uniform grid builds no table (Grid.h:623): empty tables in the file
    u = bt.Grid(5, 5, True, cam, bh, ctx=ctx)
    u.write(tmp_path / "u.grid")
    du = bt.read_grid_file(tmp_path / "u.grid")
    assert du["hw"] == 0 and du["hash_table"].size == 0


def test_star_file_scene_renders_as_built_scene(ctx, tmp_path):
    stars = scene.synthetic_stars(8000, 5)
    sky = scene.synthetic_sky(256, seed=6)
    tree, pos, mag = bt.star_tree(stars, 7)
    bt.write_star_file(tmp_path / "s.star", tree, 7, pos, mag)
    t2, l2, p2, m2 = bt.read_star_file(tmp_path / "s.star")
    a = bt.Scene(stars, sky, tree_level=7, ctx=ctx)
    b = bt.Scene.from_tree(t2, l2, p2, m2, sky, ctx=ctx)
    bh = bt.BlackHole(0.999)
    cam = bt.Camera(math.pi / 2, 0.0, 5.0, black_hole=bh)
    g = bt.Grid(7, 1, True, cam, bh, ctx=ctx)
    v = bt.View(160, 90, ctx=ctx)
    assert np.array_equal(bt.render_frame(g, a, v), bt.render_frame(g, b, v))
    with pytest.raises(bt.BtError):   # a tree that does not index the stars
        bt.Scene.from_tree(t2, l2, p2[:-1], m2[:-1], sky, ctx=ctx)


CEREAL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                      "libbtcereal.so")


@pytest.mark.skipif(not os.path.exists(CEREAL), reason="oracle/_ref/libbtcereal.so not built (no reference checkout)")
def test_grid_file_equals_cereal_archive(ctx, tmp_path):
    """Grid's cache (Grid.h:34-39 with PSHOffsetTable.h:44-48) written by
    cereal's BinaryOutputArchive from the grid's exported tables and by
    bt_grid_file_write: the same bytes, adaptive and uniform grids."""
    lib = C.CDLL(CEREAL)
    bh = bt.BlackHole(0.7)
    cam = bt.Camera(1.1, 0.0, 12.0, black_hole=bh)
    for lv, start in ((8, 1), (5, 5)):
        g = bt.Grid(lv, start, True, cam, bh, ctx=ctx)
        ours = tmp_path / f"ours{lv}.grid"
        g.write(ours)
        ht, ot, tag, hw, ow = g.psh() if start != lv else (np.zeros(0, np.float32), np.zeros(0, np.int32),
                                                           np.zeros(0, np.int32), 0, 0)
        ht = np.ascontiguousarray(ht, np.float32).reshape(-1)
        ot = np.ascontiguousarray(ot, np.int32).reshape(-1)
        tag = np.ascontiguousarray(tag, np.int32).reshape(-1)
        theirs = tmp_path / f"cereal{lv}.grid"
        ip = C.POINTER(C.c_int)
        rc = lib.btcereal_grid_write(str(theirs).encode(), C.c_int(g.MAXLEVEL), C.c_int(g.N), C.c_int(g.M),
                                     ht.ctypes.data_as(C.POINTER(C.c_float)), ot.ctypes.data_as(ip),
                                     tag.ctypes.data_as(ip), C.c_int(hw), C.c_int(ow))
        assert rc == 0
        assert ours.read_bytes() == theirs.read_bytes()
