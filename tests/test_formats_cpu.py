"""The reference's star catalogue reader and its cereal caches (host only).

Pins: the catalogue conversion keeps each star's direction (equatorial
(ra, dec) and the wrapped (theta, phi) give the same unit vector), the
magnitude cut and file order; the .star layout is checked byte by byte
against cereal's binary archive rules (a 32-bit int raw, a vector as a
64-bit count then its raw elements, little endian) and against a file that
cereal itself writes (the reference's vendored cereal 1.2.2, built into
oracle/_ref/libbtcereal.so by oracle/Makefile where the checkout exists)."""
import ctypes as C
import math
import os
import struct

import numpy as np
import pytest

import blacktracer_amd as bt


def _unit(theta, phi):
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)


def test_catalogue_keeps_directions_cut_and_order():
    rng = np.random.default_rng(11)
    n = 500
    mag = rng.uniform(-1.5, 9.0, n)
    col = rng.uniform(-0.3, 2.0, n)
    ra = rng.uniform(-7.0, 13.0, n)          # any radians: the reader wraps
    dec = rng.uniform(-4.0, 4.0, n)
    text = "\n".join(f"{m:.6f} {c:.6f} {r:.6f} {d:.6f} 0.5" for m, c, r, d in zip(mag, col, ra, dec)) + "\n"
    s4 = bt.read_star_catalog(text, 6.0)
    m32 = np.array([float(np.float32(f"{m:.6f}")) for m in mag])
    keep = m32 <= 6.0
    assert s4.shape == (keep.sum(), 4)
    assert np.array_equal(s4[:, 2], m32[keep].astype(np.float32))                  # file order, cut
    assert np.all((s4[:, 0] >= 0) & (s4[:, 0] <= math.pi + 1e-6))
    assert np.all((s4[:, 1] >= 0) & (s4[:, 1] < 2 * math.pi + 1e-6))
    r32 = np.array([float(np.float32(f"{r:.6f}")) for r in ra])[keep]
    d32 = np.array([float(np.float32(f"{d:.6f}")) for d in dec])[keep]
    # equatorial unit vector: z = sin(dec), theta = pi/2 - dec
    want = _unit(math.pi / 2 - d32, r32)
    got = _unit(s4[:, 0].astype(np.float64), s4[:, 1].astype(np.float64))
    assert np.max(np.abs(got - want)) < 2e-6


def test_catalogue_edge_cases():
    assert bt.read_star_catalog("", 6.0).shape == (0, 4)
    # a trailing incomplete record is dropped (the reference appended a record
    # made of a failed read)
    s4 = bt.read_star_catalog("1 0 0.5 0.25 0\n2 0 0.5", 6.0)
    assert s4.shape == (1, 4)
    with pytest.raises(bt.BtError):
        bt.read_star_catalog("1 0 x 0.25 0\n", 6.0)


def test_star_file_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    S, L = 300, 5
    stars = np.stack([np.arccos(1 - 2 * rng.random(S)), 2 * math.pi * rng.random(S),
                      rng.uniform(0, 8, S), rng.uniform(0, 1.5, S)], -1).astype(np.float32)
    tree, pos, mag = bt.star_tree(stars, L)
    p = tmp_path / "a.star"
    bt.write_star_file(p, tree, L, pos, mag)
    raw = p.read_bytes()
    T = (1 << (L + 1)) - 1
    off = 0

    def vec(fmt, n_expected):
        nonlocal off
        (n,) = struct.unpack_from("<Q", raw, off)
        off += 8
        assert n == n_expected
        a = np.frombuffer(raw, dtype=fmt, count=n, offset=off)
        off += a.nbytes
        return a

    assert np.array_equal(vec("<i4", T), tree)
    assert np.array_equal(vec("<f4", 2 * S), pos.ravel())
    assert np.array_equal(vec("<f4", 2 * S), mag.ravel())
    assert struct.unpack_from("<iii", raw, off) == (S, T, L)
    assert off + 12 == len(raw)
    t2, l2, p2, m2 = bt.read_star_file(p)
    assert l2 == L and np.array_equal(t2, tree) and np.array_equal(p2, pos) and np.array_equal(m2, mag)


def test_star_file_written_by_cereal_rules_reads_back(tmp_path):
    """A file laid out by hand from cereal's rules reads back as written."""
    L, S = 1, 2
    tree = np.array([1, 1, 2], dtype="<i4")
    pos = np.array([0.5, 1.0, 2.0, 3.0], dtype="<f4")
    mag = np.array([4.0, 0.5, 6.0, 1.0], dtype="<f4")
    raw = (struct.pack("<Q", 3) + tree.tobytes() + struct.pack("<Q", 4) + pos.tobytes() +
           struct.pack("<Q", 4) + mag.tobytes() + struct.pack("<iii", S, 3, L))
    p = tmp_path / "b.star"
    p.write_bytes(raw)
    t2, l2, p2, m2 = bt.read_star_file(p)
    assert l2 == L and np.array_equal(t2, tree) and np.array_equal(p2.ravel(), pos)
    assert np.array_equal(m2.ravel(), mag)
    p.write_bytes(raw[:-4])                 # truncated: refused
    with pytest.raises(bt.BtError):
        bt.read_star_file(p)
    with pytest.raises(bt.BtError):
        bt.read_star_file(tmp_path / "missing.star")


def test_grid_file_header_rules(tmp_path):
    """A .grid laid out by hand (MAXLEVEL, N, M, three vectors of pairs, two
    widths) reads back; inconsistent widths are refused."""
    hw, ow = 2, 1
    ht = np.arange(2 * hw * hw, dtype="<f4")
    ot = np.array([1, 0], dtype="<i4")
    tag = np.arange(2 * hw * hw, dtype="<i4")
    raw = (struct.pack("<iii", 9, 513, 1024) + struct.pack("<Q", hw * hw) + ht.tobytes() +
           struct.pack("<Q", ow * ow) + ot.tobytes() + struct.pack("<Q", hw * hw) + tag.tobytes() +
           struct.pack("<ii", hw, ow))
    p = tmp_path / "g.grid"
    p.write_bytes(raw)
    g = bt.read_grid_file(p)
    assert (g["max_level"], g["N"], g["M"], g["hw"], g["ow"]) == (9, 513, 1024, hw, ow)
    assert np.array_equal(g["hash_table"], ht) and np.array_equal(g["offset_table"], ot)
    assert np.array_equal(g["hash_pos_tag"], tag)
    p.write_bytes(raw[:-8] + struct.pack("<ii", hw + 1, ow))
    with pytest.raises(bt.BtError):
        bt.read_grid_file(p)


CEREAL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                      "libbtcereal.so")


@pytest.mark.skipif(not os.path.exists(CEREAL), reason="oracle/_ref/libbtcereal.so not built (no reference checkout)")
@pytest.mark.parametrize("S,L", [(0, 1), (1, 2), (2500, 8)])
def test_star_file_equals_cereal_archive(tmp_path, S, L):
    """StarProcessor's cache (StarProcessor.h:19-23) written by cereal's own
    BinaryOutputArchive and by bt_star_file_write: the same bytes."""
    rng = np.random.default_rng(S + L)
    stars = np.stack([np.arccos(1 - 2 * rng.random(S)), 2 * math.pi * rng.random(S),
                      rng.uniform(0, 8, S), rng.uniform(0, 1.5, S)], -1).astype(np.float32).reshape(S, 4)
    tree, pos, mag = bt.star_tree(stars, L)
    ours = tmp_path / "ours.star"
    bt.write_star_file(ours, tree, L, pos, mag)
    lib = C.CDLL(CEREAL)
    theirs = tmp_path / "cereal.star"
    tree = np.ascontiguousarray(tree, np.int32)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
    mag = np.ascontiguousarray(mag, np.float32).reshape(-1)
    fp = C.POINTER(C.c_float)
    rc = lib.btcereal_star_write(str(theirs).encode(), tree.ctypes.data_as(C.POINTER(C.c_int)), C.c_int(tree.size),
                                 pos.ctypes.data_as(fp), mag.ctypes.data_as(fp), C.c_int(S), C.c_int(L))
    assert rc == 0
    assert ours.read_bytes() == theirs.read_bytes()
