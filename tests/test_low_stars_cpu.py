"""Stars below the catalogue cut painted into the sky image
(StarProcessor::addLowLightStarsToImage, StarProcessor.h:108-158): the
library's host implementation against the oracle, and both against closed
forms of the splat (host only, no device)."""
import math
import os

import numpy as np

import blacktracer_amd as bt
from blacktracer_amd import scene as S
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def blackbody_table():
    vals = []
    with open(os.path.join(HERE, "..", "oracle", "blackbody_rgb.inc")) as f:
        for line in f:
            vals += [int(v) for v in line.split("/*")[0].replace(",", " ").split()]
    return np.array(vals, dtype=np.int64).reshape(-1, 3)


def star_at(row, col, rows, cols, mag, T, fy=0.5, fx=0.5):
    """A star whose image position is (row + fy, col + fx) in sky pixels."""
    theta = (row + fy) / rows * math.pi
    phi = (col + fx) / cols * 2 * math.pi
    return [theta, phi, mag, float(S.temperature_to_bv(T))]


def test_no_faint_stars_leaves_the_sky_unchanged():
    sky = S.synthetic_sky(256, seed=4)
    stars = S.synthetic_stars(500, seed=9)
    out = bt.sky_add_low_stars(sky, stars, magnitude_cut=100.0)   # every star is bright: none painted
    assert np.array_equal(out, sky)


def test_single_star_splat_closed_form():
    """One star of magnitude m at a pixel centre on a black sky: tap (du, dv)
    holds 400 g(d) 10^(-0.4 m) times the squared blackbody colour, with
    g(d) = exp(-2 d) / pi (0.318310) of d = du^2 + dv^2, for d <= 6.25;
    the pixel is floor(sqrt(.)) of it, capped at 255; nothing else lights."""
    rows, cols = 64, 128
    sky = np.zeros((rows, cols, 3), np.uint8)
    T, m = 6500.0, 2.0
    st = np.array([star_at(30, 60, rows, cols, m, T)], np.float32)
    out = bt.sky_add_low_stars(sky, st, magnitude_cut=1.0)
    rgb = blackbody_table()[int(T / 100) - 10]            # Ballesteros at z = 1: index T/100 - 10
    want = np.zeros((rows, cols, 3), np.int64)
    for du in range(-2, 3):
        for dv in range(-2, 3):
            d = du * du + dv * dv
            if d > 6.25:
                continue
            b = 400.0 * 0.318310 * math.exp(-2 * d) * 10 ** (-0.4 * m)
            light = np.floor(b * rgb[::-1].astype(np.float64) ** 2)     # BGR
            want[30 + du, 60 + dv] = np.minimum(255, np.floor(np.sqrt(light)))
    assert np.abs(out.astype(np.int64) - want).max() <= 1
    lit = out.max(axis=-1) > 0
    assert lit.sum() == 21            # the 5 x 5 taps within 2.5 pixels: 25 - 4 corners
    assert np.array_equal(out, O.sky_low_stars(sky, st))


def test_later_star_overwrites_taps_and_bright_background_adds_in_quadrature():
    rows, cols = 32, 64
    sky = np.full((rows, cols, 3), 100, np.uint8)
    a = star_at(10, 20, rows, cols, 3.0, 4000.0)
    b = star_at(10, 21, rows, cols, 4.0, 9000.0)
    both = bt.sky_add_low_stars(sky, np.array([a, b], np.float32), 0.0)
    only_b = bt.sky_add_low_stars(sky, np.array([b], np.float32), 0.0)
    # every tap of the later star shows only the later star
    taps_b = bt.sky_add_low_stars(np.zeros_like(sky), np.array([b], np.float32), 0.0).max(axis=-1) > 0
    assert np.array_equal(both[taps_b], only_b[taps_b])
    # a tap adds to the background in quadrature: out^2 <= light + bg^2 < (out + 1)^2
    light_sq = only_b.astype(np.int64) ** 2 - 100 ** 2
    assert (light_sq[taps_b] >= 0).all()
    assert np.array_equal(both, O.sky_low_stars(sky, np.array([a, b], np.float32)))


def test_seam_and_poles():
    """Splats wrap across phi = 0 / 2 pi and are cut at the first and last rows."""
    rows, cols = 32, 64
    sky = np.zeros((rows, cols, 3), np.uint8)
    st = np.array([star_at(0, 0, rows, cols, 1.0, 5000.0, fy=0.2, fx=0.1),
                   star_at(rows - 1, cols - 1, rows, cols, 1.0, 7000.0, fy=0.9, fx=0.8)], np.float32)
    out = bt.sky_add_low_stars(sky, st, 0.5)
    assert out[0, cols - 1].max() > 0 and out[0, 1].max() > 0          # across the seam
    assert out[rows - 1, 0].max() > 0                                  # the second star wraps too
    assert np.array_equal(out, O.sky_low_stars(sky, st))


def test_library_equals_oracle_on_a_catalogue():
    sky = S.synthetic_sky(512, seed=8)
    stars = S.synthetic_stars(20000, seed=12)
    stars[:, 2] = np.float32(4.0) + np.float32(8.0) * stars[:, 2] / np.float32(8.0)   # magnitudes 4..12
    cut = 7.5
    out = bt.sky_add_low_stars(sky, stars, cut)
    ref = O.sky_low_stars(sky, stars[stars[:, 2] > cut])
    assert np.array_equal(out, ref)
    assert not np.array_equal(out, sky)
