"""Pins the render-pass oracle (Path B) with closed forms that do not restate
its code: an unlensed (identity) grid, one star at a known sub-pixel
position, skies whose colours are known per texel, and the physics of the
redshift.  Each test names the reference code it pins.

Identity grid: every grid point holds its own camera-sky direction, so pixel
corners map to themselves (pixInterpolation reproduces affine fields
exactly), the lensing factor is 1 and the pixel's celestial footprint is the
pixel itself.
"""
import math
import os

import numpy as np
import pytest

from blacktracer_amd import scene
from oracle import oracle as O

L = O.lib()
FP, IP, U8P = O.FP, O.IP, O.U8P
HERE = os.path.dirname(os.path.abspath(__file__))
SIGMA = 0.4            # kernel.cuh:335-338: sigma = 1 / 2.5


def f32(x):
    return np.ascontiguousarray(x, dtype=np.float32)


def blackbody_table():
    """(391, 3) RGB8 of 1000 K + 100 K * k, read as data from the table file."""
    vals = []
    with open(os.path.join(HERE, "..", "oracle", "blackbody_rgb.inc")) as f:
        for line in f:
            line = line.split("/*")[0]
            vals += [int(v) for v in line.replace(",", " ").split()]
    t = np.array(vals, dtype=np.int64).reshape(-1, 3)
    assert t.shape == (391, 3)
    return t


def grid_field(lv, fn):
    """A full-sky dense grid of level lv whose point (i, j) holds fn(theta, phi)
    of its own camera direction (theta = 2 pi i / GM, phi = 2 pi j / GM)."""
    GN, GM = (1 << lv) + 1, 1 << (lv + 1)
    i, j = np.meshgrid(np.arange(GN), np.arange(GM), indexing="ij")
    th, ph = 2 * np.pi * i / GM, 2 * np.pi * j / GM
    t2, p2 = fn(th, ph)
    return np.stack([t2, p2], -1).astype(np.float32)


def identity_grid(lv):
    return grid_field(lv, lambda t, p: (t, p))


def square_view(th0, ph0, d, W, H):
    """Pixel corners on an axis-aligned lattice: corner (i, j) = (th0 + i d, ph0 + j d)."""
    i, j = np.meshgrid(np.arange(H + 1), np.arange(W + 1), indexing="ij")
    return np.stack([th0 + i * d, ph0 + j * d], -1).astype(np.float32)


def static_cam(alpha=1.0):
    # getParamArray order {speed, alpha, w, wbar, br, btheta, bphi}: a camera at rest, no frame dragging
    return np.array([0.0, alpha, 0.0, 1.0, 0.0, 0.0, 1.0], np.float32)


def run_frame(grid, lv, view, cam, stars=None, sky=None, star=True, map=False, lr=True, step=1, tree_level=4):
    if stars is None:
        stars = np.array([[0.1, 0.1, 20.0, 0.5]], np.float32)
    tree, pos, mag = O.make_tree(stars, tree_level)
    if sky is None:
        sky = np.zeros((32, 64, 3), np.uint8)
    table = O.sum_table(sky)
    return O.frame(grid, lv, view, cam, tree, tree_level, pos, mag, table, star=star, map=map, lr=lr,
                   step=step, stages=True)


# ----------------------------------------------------------------------------- small functions


def test_gaussian_is_the_normalised_profile():
    """gaussian (kernel.cuh:335-338) takes a squared distance: g(d^2) =
    exp(-d^2 / (2 sigma^2)) / (sigma sqrt(2 pi)) with sigma = 0.4, i.e. a unit-area
    1-D normal profile."""
    xs = np.linspace(-4, 4, 8001)
    g = np.array([L.btor_gaussian(float(x * x)) for x in xs])
    assert np.trapezoid(g, xs) == pytest.approx(1.0, rel=1e-5)
    for d2 in (0.0, 0.08, 0.5, 1.28, 2.25):
        want = math.exp(-d2 / (2 * SIGMA ** 2)) / (SIGMA * math.sqrt(2 * math.pi))
        assert L.btor_gaussian(d2) == pytest.approx(want, rel=1e-6)


@pytest.mark.parametrize("mirror", [False, True])
def test_sub_pixel_position_is_the_affine_coordinate(mirror):
    """interpolateStar (kernel.cuh:112-200) bisects the quad through its edge
    midpoints 10 times; on a parallelogram that is the affine coordinate of the
    star from corner 1 (pixel corner (i, j)): Y along corner 1 -> 0 (rows),
    X along corner 1 -> 2 (columns), to 2^-11 or the 1e-5 rad early stop."""
    rng = np.random.default_rng(5)
    for _ in range(300):
        c1 = rng.uniform([1.0, 2.0], [1.5, 3.0])
        ey = rng.uniform([0.02, -0.01], [0.06, 0.01])
        ex = rng.uniform([-0.01, 0.02], [0.01, 0.06])
        if mirror:
            ex = -ex
        c0, c2 = c1 + ey, c1 + ex
        c3 = c1 + ex + ey
        t = f32([c0[0], c1[0], c2[0], c3[0]])
        p = f32([c0[1], c1[1], c2[1], c3[1]])
        # orientation of the corner cycle 0-1-2-3 by the shoelace formula
        area2 = sum(t[q] * p[(q + 1) % 4] - t[(q + 1) % 4] * p[q] for q in range(4))
        sgn = 1 if area2 < 0 else -1
        u, v = rng.uniform(0.05, 0.95, 2)
        star = c1 + u * ey + v * ex
        yx = np.zeros(2, np.float32)
        L.btor_interpolate_star(O._ptr(t, FP), O._ptr(p, FP), float(star[0]), float(star[1]), sgn,
                                O._ptr(yx, FP))
        tol = 2.0 ** -10 + 1e-5 / min(np.linalg.norm(ey), np.linalg.norm(ex)) * 2
        assert abs(yx[0] - u) <= tol and abs(yx[1] - v) <= tol, (u, v, yx)


def test_pi_check_unwraps_only_across_the_seam():
    """piCheck (kernel.cuh:274-295): when some phi exceeds 2 pi (1 - f), every
    phi below 2 pi f gains 2 pi; otherwise nothing changes."""
    f = 0.2
    cases = [([6.2, 0.05, 0.1, 6.1], [6.2, 0.05 + 2 * math.pi, 0.1 + 2 * math.pi, 6.1], 1),
             ([3.0, 3.1, 3.2, 3.05], [3.0, 3.1, 3.2, 3.05], 0),
             ([6.0, 1.5, 6.1, 1.4], [6.0, 1.5, 6.1, 1.4], 0),
             ([5.1, 1.2, 0.3, 5.2], [5.1, 1.2 + 2 * math.pi, 0.3 + 2 * math.pi, 5.2], 1)]
    for p, want, flag in cases:
        pp = f32(p)
        got = L.btor_pi_check(O._ptr(pp, FP), f)
        assert got == flag
        assert np.allclose(pp, want, atol=1e-6)


def test_catmull_rom_is_exact_on_quadratics():
    """hermite with tension = bias = 0 (kernel.cuh:790-815) is Catmull-Rom:
    central-difference tangents are exact for quadratics, so the cubic Hermite
    reproduces x0..x3 sampled from any quadratic at every v in [0, 1]."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        a, b, c = rng.normal(size=3)
        d, e, f = rng.normal(size=3)
        q1 = lambda s: a * s * s + b * s + c  # noqa: E731
        q2 = lambda s: d * s * s + e * s + f  # noqa: E731
        x = f32([q1(-1), q2(-1), q1(0), q2(0), q1(1), q2(1), q1(2), q2(2)])
        out = np.zeros(2, np.float32)
        for v in (0.0, 0.25, 0.5, 0.8, 1.0):
            L.btor_hermite(v, O._ptr(x, FP), O._ptr(out, FP))
            assert out[0] == pytest.approx(q1(v), abs=2e-5)
            assert out[1] == pytest.approx(q2(v), abs=2e-5)


def test_line_intersection():
    """intersection (kernel.cuh:754-772): the meeting point of lines AB and CD;
    parallel lines give (-1, -1)."""
    rng = np.random.default_rng(4)
    for _ in range(100):
        P = rng.uniform(-2, 2, 2)
        u, w = rng.normal(size=2), rng.normal(size=2)
        s1, s2, s3, s4 = rng.uniform(-1, 1, 4)
        abcd = f32([*(P + s1 * u), *(P + s2 * u), *(P + s3 * w), *(P + s4 * w)])
        out = np.zeros(2, np.float32)
        L.btor_intersection(O._ptr(abcd, FP), O._ptr(out, FP))
        if abs(u[0] * w[1] - u[1] * w[0]) > 0.05 and abs(s1 - s2) > 0.05 and abs(s3 - s4) > 0.05:
            assert np.allclose(out, P, atol=5e-4)
    par = f32([0, 0, 1, 1, 0, 1, 1, 2])
    out = np.zeros(2, np.float32)
    L.btor_intersection(O._ptr(par, FP), O._ptr(out, FP))
    assert tuple(out) == (-1.0, -1.0)


def test_wrap_to_pi_float_brute_force():
    """wrapToPi (kernel.cuh:694-703): the same direction on the sphere, theta in
    [0, pi], phi in [0, 2 pi)."""
    rng = np.random.default_rng(6)
    for t0, p0 in rng.uniform([-3, -7], [9, 13], size=(300, 2)):
        t, p = f32([t0]), f32([p0])
        L.btor_wrap_to_pi(O._ptr(t, FP), O._ptr(p, FP))
        assert -1e-6 <= t[0] <= math.pi + 1e-6 and -1e-6 <= p[0] < 2 * math.pi + 1e-5
        v0 = (math.sin(t0) * math.cos(p0), math.sin(t0) * math.sin(p0), math.cos(t0))
        v1 = (math.sin(t[0]) * math.cos(p[0]), math.sin(t[0]) * math.sin(p[0]), math.cos(t[0]))
        assert np.allclose(v0, v1, atol=2e-5)


# ----------------------------------------------------------------------------- viewer


@pytest.mark.parametrize("fov,ver,hor,w,h", [(math.pi / 2, 0.0, math.pi / 16, 64, 36),
                                             (1.2, 0.1, 0.3, 40, 30)])
def test_camera_view_closed_form(fov, ver, hor, w, h):
    """makeCameraView (Viewer.h:109-133): with s = tan(fov/2), step = 2 s / W,
    row i looks at theta = pi/2 - atan(s H / W - i step) - offVer and
    column W - j at phi = pi - atan(s - j step) - offHor."""
    v = O.camera_view(fov, ver, hor, w, h)
    s = math.tan(fov / 2)
    step = 2 * s / w
    for i in range(h + 1):
        th = math.pi / 2 - math.atan(s * h / w - i * step) - ver
        assert np.allclose(v[i, :, 0], th, atol=2e-6)
    for j in range(w + 1):
        ph = math.pi - math.atan(s - j * step) - hor
        assert np.allclose(v[:, w - j, 1], ph, atol=2e-6)


def test_sphere_view_closed_form():
    """makeSphereView (Viewer.h:82-88): corner (i, j) = (i pi / H, 2 pi j / W)."""
    v = O.sphere_view(48, 24)
    i, j = np.meshgrid(np.arange(25), np.arange(49), indexing="ij")
    assert np.allclose(v[..., 0], i * math.pi / 24, atol=1e-6)
    assert np.allclose(v[..., 1], 2 * math.pi * j / 48, atol=1e-6)


# ----------------------------------------------------------------------------- grid lookups


def test_perfect_hash_lookup_round_trip():
    """hashLookup (kernel.cuh:705-733) inside makeGrid (MakeImageKernel.cu:369-384):
    a perfect hash built here by random offsets (displacement per offset cell)
    returns every stored point and -2 for every absent one."""
    rng = np.random.default_rng(8)
    GN, GM = 33, 64
    present = rng.random((GN, GM)) < 0.3
    vals = rng.uniform(0, 3, (GN, GM, 2)).astype(np.float32)
    pts = np.argwhere(present)
    n = len(pts)
    hw = int(math.sqrt(2.5 * n)) + 1
    ow = 11
    for _ in range(200):  # randomised displacement search
        off = np.zeros((ow * ow, 2), np.int32)
        used = np.zeros((hw, hw), bool)
        ok = True
        groups = {}
        for i, j in pts:
            groups.setdefault((i % ow, j % ow), []).append((i, j))
        for cell, members in sorted(groups.items(), key=lambda kv: -len(kv[1])):
            for _try in range(400):
                d = rng.integers(0, hw, 2)
                slots = [((i + d[0]) % hw, (j + d[1]) % hw) for i, j in members]
                if len(set(slots)) == len(slots) and not any(used[s] for s in slots):
                    for s in slots:
                        used[s] = True
                    off[cell[0] * ow + cell[1]] = d
                    break
            else:
                ok = False
                break
        if ok:
            break
    assert ok
    ht = np.zeros((hw * hw, 2), np.float32)
    tag = np.full((hw * hw, 2), -1, np.int32)
    for i, j in pts:
        d = off[(i % ow) * ow + (j % ow)]
        s = ((i + d[0]) % hw) * hw + (j + d[1]) % hw
        ht[s] = vals[i, j]
        tag[s] = (i, j)
    grid = O.make_grid(ht, off, tag, hw, ow, GN, GM)
    assert grid.shape == (GN, GM, 2)
    assert np.array_equal(grid[present], vals[present])
    assert (grid[~present] == -2).all()


def test_pix_interpolation_reproduces_quadratic_fields():
    """pixInterpolation on a uniform grid (MakeImageKernel.cu:458-515,
    kernel.cuh:736-968): bicubic Catmull-Rom reproduces fields quadratic in
    theta and phi, so every pixel corner gets the field's own value there."""
    lv = 8
    c_t, c_p = 1.5, 2.5
    fn = lambda t, p: (c_t + 0.9 * (t - c_t) + 0.2 * (t - c_t) ** 2 + 0.05 * (t - c_t) * (p - c_p),  # noqa: E731
                       c_p + 1.1 * (p - c_p) - 0.15 * (p - c_p) ** 2 + 0.1 * (t - c_t))
    grid = grid_field(lv, fn)
    view = square_view(1.3, 2.2, 0.0137, 40, 30)
    _, st = run_frame(grid, lv, view, static_cam(), star=False, map=False, lr=False)
    want_t, want_p = fn(view[..., 0].astype(np.float64), view[..., 1].astype(np.float64))
    assert np.abs(st["thphi"][..., 0] - want_t).max() < 2e-5
    assert np.abs(st["thphi"][..., 1] - want_p).max() < 2e-5
    assert (st["bh"] == 0).all()


# ----------------------------------------------------------------------------- redshift


@pytest.mark.parametrize("a,r,th,motion", [(0.999, 5.0, math.pi / 2, (0.0, 0.0, 1.0)),
                                           (0.9, 8.0, 1.1, (0.3, 0.2, math.sqrt(1 - 0.13))),
                                           (0.5, 3.0, 0.7, (0.0, 0.6, 0.8))])
def test_redshift_is_energy_at_infinity_over_energy_at_camera(a, r, th, motion):
    """redshift (kernel.cuh:348-365) for a camera moving at v along B relative
    to the ZAMO: 1 + z = E_inf / E_cam.  With E = 1 and L = b at infinity the
    ZAMO measures E_F = (1 - omega b) / alpha; the camera, moving at v along
    its own y axis, measures E_F / (gamma (1 - v y)) for a photon arriving from
    view direction y = sin(theta) sin(phi).  b is the impact parameter the
    grid build's initial conditions give the same camera direction (the Path A
    oracle, itself pinned bit for bit to the reference's Metric.h/Grid.h)."""
    ps = O.pose(a, th, 0.0, r, *motion)
    cam = f32([ps.speed, ps.alpha, ps.w, ps.wbar, ps.br, ps.btheta, ps.bphi])
    v = ps.speed
    gamma = 1 / math.sqrt(1 - v * v)
    rng = np.random.default_rng(11)
    for tc, pc in zip(rng.uniform(0.1, 3.0, 40), rng.uniform(0, 2 * math.pi, 40)):
        b = O.initial_conditions(ps, float(tc), float(pc))[2]
        y = math.sin(tc) * math.sin(pc)
        want = ps.alpha * gamma * (1 - v * y) / (1 - ps.w * b)
        got = L.btor_redshift(float(tc), float(pc), O._ptr(cam, FP))
        assert got == pytest.approx(want, rel=2e-6)


@pytest.mark.parametrize("v", [0.1, 0.5, 0.9])
def test_redshift_flat_space_is_special_relativistic_doppler(v):
    """alpha = 1, omega = 0: 1 + z = gamma (1 - v cos psi), psi the angle between
    the view direction and the motion (the camera's y axis)."""
    cam = f32([v, 1.0, 0.0, 1.0, 0.0, 0.0, 1.0])
    g = 1 / math.sqrt(1 - v * v)
    for tc, pc in [(math.pi / 2, math.pi / 2), (math.pi / 2, 3 * math.pi / 2), (0.4, 1.0), (2.5, 4.0)]:
        cospsi = math.sin(tc) * math.sin(pc)
        assert L.btor_redshift(tc, pc, O._ptr(cam, FP)) == pytest.approx(g * (1 - v * cospsi), rel=1e-6)
    # looking along the motion: blueshift sqrt((1 - v) / (1 + v))
    assert L.btor_redshift(math.pi / 2, math.pi / 2, O._ptr(cam, FP)) == pytest.approx(
        math.sqrt((1 - v) / (1 + v)), rel=1e-6)


# ----------------------------------------------------------------------------- star light


def star_frame(T, mag, fy, fx, alpha=1.0, lr=True, step=1):
    """One star at sub-pixel (fy, fx) of pixel (8, 10) on the identity grid."""
    lv = 9
    W, H, d = 24, 16, 0.05
    th0, ph0 = 1.2, 2.0
    view = square_view(th0, ph0, d, W, H)
    bv = scene.temperature_to_bv(T)
    stars = np.array([[th0 + (8 + fy) * d, ph0 + (10 + fx) * d, mag, bv]], np.float32)
    _, st = run_frame(identity_grid(lv), lv, view, static_cam(alpha), stars=stars, lr=lr, step=step)
    return st


@pytest.mark.parametrize("T,mag,fy,fx,alpha,lr,step", [
    (5750.0, 3.0, 0.3, 0.7, 1.0, True, 1),
    (3050.0, 1.5, 0.62, 0.21, 1.0, True, 1),
    (9950.0, 5.0, 0.5, 0.5, 1.0, False, 1),
    (12040.0, 2.0, 0.45, 0.8, 0.8, True, 1),
    (25030.0, 0.5, 0.15, 0.35, 1.0, True, 2),
])
def test_star_taps_closed_form(T, mag, fy, fx, alpha, lr, step):
    """distortStarMap (MakeImageKernel.cu:206-326) + sumStarLight (:753-773):
    a star at sub-pixel (Y, X) of pixel (i, j) lights pixel (i + di, j + dj) with
    20 (100 frac g(d2) 10^(-0.4 (m + 4 log10 z)) c)^2 per channel, where
    d2 = (di + 0.5 - Y)^2 + (dj + 0.5 - X)^2 <= (step + 0.5)^2, frac = 1 on the
    identity grid, z = alpha for a camera at rest (1 with lr off), and c is the
    blackbody colour of row floor(T / (100 z)) - 10 (temperatures 1000 K +
    100 K k; B-V from inverting Ballesteros' formula, solved here by
    numpy.roots)."""
    # B-V by an independent root find of Ballesteros: T = 4600 (1/(0.92 BV + 1.7) + 1/(0.92 BV + 0.62))
    k = T / 4600.0
    roots = np.roots([k * 0.92 ** 2, 0.92 * (2.32 * k - 2.0), 1.054 * k - 2.32])
    bv = float(max(r.real for r in roots if abs(r.imag) < 1e-12 and 0.92 * r.real > -0.62))
    assert bv == pytest.approx(scene.temperature_to_bv(T), abs=1e-9)
    st = star_frame(T, mag, fy, fx, alpha, lr, step)["star"]
    z = alpha if lr else 1.0
    table = blackbody_table()
    row = int(math.floor(T / (100.0 * z))) - 10
    c = table[row].astype(np.float64)
    expect = np.zeros_like(st, dtype=np.float64)
    for di in range(-step, step + 1):
        for dj in range(-step, step + 1):
            d2 = (di + 0.5 - fy) ** 2 + (dj + 0.5 - fx) ** 2
            if d2 > (step + 0.5) ** 2:
                continue
            g = math.exp(-d2 / (2 * SIGMA ** 2)) / (SIGMA * math.sqrt(2 * math.pi))
            bright = 100.0 * g * 10 ** (-0.4 * (mag + 4 * math.log10(z)))
            expect[8 + di, 10 + dj] = 20.0 * (bright * c) ** 2
    lit = expect.sum(-1) > 0
    assert ((st.sum(-1) > 0) == lit).all(), "star light on the wrong pixels"
    assert np.allclose(st[lit], expect[lit], rtol=1e-2)


def test_star_light_adds_over_stars():
    """Two stars in different pixels: the light is the sum of each alone."""
    lv = 9
    W, H, d = 24, 16, 0.05
    view = square_view(1.2, 2.0, d, W, H)
    mk = lambda fy, fx, i, j, m, T: [1.2 + (i + fy) * d, 2.0 + (j + fx) * d, m, scene.temperature_to_bv(T)]  # noqa
    s1 = mk(0.3, 0.6, 5, 5, 2.0, 4500.0)
    s2 = mk(0.7, 0.4, 6, 6, 3.0, 8000.0)
    g = identity_grid(lv)
    a = run_frame(g, lv, view, static_cam(), stars=np.array([s1], np.float32))[1]["star"]
    b = run_frame(g, lv, view, static_cam(), stars=np.array([s2], np.float32))[1]["star"]
    ab = run_frame(g, lv, view, static_cam(), stars=np.array([s1, s2], np.float32))[1]["star"]
    assert np.allclose(ab, a + b, rtol=1e-5)


# ----------------------------------------------------------------------------- makePix, composite


def test_make_pix_below_and_above_saturation():
    """makePix (MakeImageKernel.cu:802-835): light L per channel gives
    floor(sqrt(L)) while every channel is <= 255; brighter colours keep their
    hue and saturation (a common scale factor) and are scaled to perceived
    brightness 255 in the HSP model, P = sqrt(.299 R^2 + .587 G^2 + .114 B^2)."""
    rng = np.random.default_rng(9)
    out = np.zeros(4, np.uint8)
    for _ in range(300):
        rgb = rng.uniform(0, 255, 3)
        L.btor_make_pix(*(float(x * x) for x in rgb), O._ptr(out, U8P))
        assert list(out[:3][::-1]) == [int(math.floor(np.float32(x))) for x in np.sqrt((rgb ** 2).astype(np.float32))]
        assert out[3] == 255
    for _ in range(300):
        rgb = rng.uniform(0, 255, 3) * rng.uniform(1.2, 30)
        if rgb.max() <= 256:
            continue
        L.btor_make_pix(*(float(x * x) for x in rgb), O._ptr(out, U8P))
        P = math.sqrt(0.299 * rgb[0] ** 2 + 0.587 * rgb[1] ** 2 + 0.114 * rgb[2] ** 2)
        want = np.minimum(255.0, rgb * 255.0 / P)
        got = out[:3][::-1].astype(np.float64)
        assert np.all(np.abs(got - want) <= 1.01), (rgb, got, want)


def test_composite_is_root_sum_of_squares():
    """addStarsAndBackground (MakeImageKernel.cu:837-853): each channel is
    min(255, floor(sqrt(star^2 + sky^2))) of the two 8-bit images."""
    lv = 9
    W, H, d = 24, 16, 0.05
    view = square_view(1.2, 2.0, d, W, H)
    stars = np.array([[1.2 + 8.4 * d, 2.0 + 10.6 * d, 1.0, scene.temperature_to_bv(6000.0)],
                      [1.2 + 3.5 * d, 2.0 + 4.5 * d, 4.0, scene.temperature_to_bv(4000.0)]], np.float32)
    sky = np.zeros((128, 256, 3), np.uint8)
    sky[...] = (90, 60, 30)
    g = identity_grid(lv)
    s, _ = run_frame(g, lv, view, static_cam(), stars=stars, sky=sky, star=True, map=False)
    m, _ = run_frame(g, lv, view, static_cam(), stars=stars, sky=sky, star=False, map=True)
    both, _ = run_frame(g, lv, view, static_cam(), stars=stars, sky=sky, star=True, map=True)
    want = np.minimum(255, np.floor(np.sqrt(s[..., :3].astype(np.float64) ** 2 + m[..., :3].astype(np.float64) ** 2)))
    assert np.abs(both[..., :3] - want).max() <= 1
    assert (s[..., :3].max() > 100) and (m[..., :3].max() > 20)


# ----------------------------------------------------------------------------- area smoothing


def test_smoothing_keeps_constant_fields_and_spreads_a_delta():
    """smoothAreaH / smoothAreaV (MakeImageKernel.cu:570-616): a box mean of
    radius r = floor(gap / 2) over the non-shadow pixels, wrapping in j and
    clamped in i.  A constant field stays constant; a delta of 1 at the centre
    becomes 1 / (2r + 1)^2 over the (2r + 1)^2 window and 0 outside."""
    W, H = 40, 30
    for gap in (1, 2, 4, 6):
        r = gap // 2
        gapm = np.full((H + 1, W + 1), gap, np.int32)
        bh = np.zeros((H, W), np.uint8)
        a = np.full((H, W), 0.37, np.float32)
        L.btor_smooth_area(O._ptr(a, FP), O._ptr(bh, U8P), O._ptr(gapm, IP), W, H)
        assert np.allclose(a, 0.37, rtol=1e-6)
        a = np.zeros((H, W), np.float32)
        a[15, 20] = 1.0
        L.btor_smooth_area(O._ptr(a, FP), O._ptr(bh, U8P), O._ptr(gapm, IP), W, H)
        want = np.zeros((H, W))
        want[15 - r:15 + r + 1, 20 - r:20 + r + 1] = 1.0 / (2 * r + 1) ** 2
        assert np.allclose(a, want, atol=1e-7)
    # wrap in j: a delta at column 0 spreads to column W - 1
    gapm = np.full((H + 1, W + 1), 2, np.int32)
    a = np.zeros((H, W), np.float32)
    a[10, 0] = 1.0
    L.btor_smooth_area(O._ptr(a, FP), O._ptr(bh, U8P), O._ptr(gapm, IP), W, H)
    assert a[10, W - 1] == pytest.approx(1 / 9) and a[10, 1] == pytest.approx(1 / 9)
    # clamp in i: at row 0 the vertical window holds 2 rows
    a = np.zeros((H, W), np.float32)
    a[0, 5] = 1.0
    L.btor_smooth_area(O._ptr(a, FP), O._ptr(bh, U8P), O._ptr(gapm, IP), W, H)
    assert a[0, 5] == pytest.approx(1 / 3 / 2) and a[1, 5] == pytest.approx(1 / 3 / 3)
    # shadow pixels neither contribute nor count, and are left as they were
    bh = np.zeros((H, W), np.uint8)
    bh[10, 11] = 1
    a = np.zeros((H, W), np.float32)
    a[10, 10] = 0.9
    a[10, 11] = 5.0
    gapm = np.full((H + 1, W + 1), 2, np.int32)
    L.btor_smooth_area(O._ptr(a, FP), O._ptr(bh, U8P), O._ptr(gapm, IP), W, H)
    assert a[10, 11] == 5.0
    # row 10 horizontal pass at j = 10: (a[10,9] + a[10,10]) / 2 (j = 11 is shadow)
    # then the vertical pass over rows 9..11 (all lit at column 10): mean of (0, 0.45, 0)
    assert a[10, 10] == pytest.approx(0.45 / 3, rel=1e-6)


# ----------------------------------------------------------------------------- environment map


def env_frame(sky, view, grid=None, lv=9, alpha=1.0, lr=True):
    if grid is None:
        grid = identity_grid(lv)
    out, st = run_frame(grid, lv, view, static_cam(alpha), sky=sky, star=False, map=True, lr=lr)
    return out, st


def test_env_map_constant_sky():
    """distortEnvironmentMap (MakeImageKernel.cu:631-751): the summed-table
    differences, the gamma 2.2 round trip and HSP give a constant sky back
    (within the final float -> 8-bit truncation), lensing on or off.  Pixel
    index 0 is skipped by the reference (`ind > 0`) and stays black."""
    sky = np.zeros((256, 512, 3), np.uint8)
    sky[...] = (40, 120, 200)
    view = square_view(1.2 + 0.0031, 2.0 + 0.0017, 0.02, 24, 16)
    for lr in (True, False):
        out, _ = env_frame(sky, view, lr=lr)
        body = out.reshape(-1, 4)[1:, :3].astype(int)
        assert (np.abs(body - [40, 120, 200]) <= 1).all()
        assert (out[0, 0, :3] == 0).all()


def texel_weights(rows, cols):
    """Solid angle of each texel row of an equirect map: (cos t0 - cos t1) dphi."""
    px = math.pi / rows
    t = np.arange(rows + 1) * px
    return (np.cos(t[:-1]) - np.cos(t[1:])) * (2 * math.pi / cols)


@pytest.mark.parametrize("d", [0.0051, 0.021, 0.047])
def test_env_map_is_the_footprint_mean(d):
    """For an axis-aligned footprint [ta, tb] x [pa, pb] on an identity grid,
    the colour is (sum w c^2.2 / sum w)^(1/2.2) over the texels of rows
    floor(ta / px) .. floor(tb / px), weights the texel solid angles, and of
    the columns the reference reads (MakeImageKernel.cu:650-708): column
    floor(pa / px) when the footprint lies in one column, otherwise columns
    floor(pa / px) + 1 .. floor(pa / px) + n, n the number of columns the
    footprint spans (its table index is one column right of the span it
    measures).  Computed here directly from the texels, not from the summed
    table."""
    rows, cols = 256, 512
    px = math.pi / rows
    rng = np.random.default_rng(12)
    sky = rng.integers(0, 256, (rows, cols, 3)).astype(np.uint8)
    W, H = 12, 10
    view = square_view(1.1 + 0.37 * px, 1.9 + 0.29 * px, d, W, H)
    out, st = env_frame(sky, view, lr=False)
    thphi = st["thphi"].astype(np.float64)
    wrow = texel_weights(rows, cols)
    lin = sky.astype(np.float64) ** 2.2
    worst = 0
    for i in range(H):
        for j in range(W):
            if i == 0 and j == 0:
                continue
            tq = [thphi[i + 1, j, 0], thphi[i, j, 0], thphi[i, j + 1, 0], thphi[i + 1, j + 1, 0]]
            pq = [thphi[i + 1, j, 1], thphi[i, j, 1], thphi[i, j + 1, 1], thphi[i + 1, j + 1, 1]]
            ta, tb, pa, pb = min(tq), max(tq), min(pq), max(pq)
            r0, r1 = int(ta // px), int(tb // px)
            c0, c1 = int(pa // px), int(pb // px)
            n = c1 - c0 + 1
            colsel = [c0] if n == 1 else [c0 + s for s in range(1, n + 1)]
            w = wrow[r0:r1 + 1][:, None] * np.ones(len(colsel))[None, :]
            num = np.einsum("rc,rcx->x", w, lin[r0:r1 + 1][:, colsel])
            mean = (num / w.sum()) ** (1 / 2.2)       # B, G, R
            worst = max(worst, np.abs(out[i, j, :3] - np.floor(mean)).max())
    assert worst <= 1


@pytest.mark.parametrize("alpha", [1.0, 0.8, 1.25])
def test_env_map_lensing_and_redshift_in_hsp(alpha):
    """With lensing on, the sky colour's perceived brightness P (HSP) becomes
    min(1, (P frac)^(1+z)) at fixed hue and saturation (MakeImageKernel.cu:
    737-745).  A grid that doubles angles about the view centre makes each
    footprint 4x the pixel: frac = Omega_pixel / Omega_footprint from the exact
    areas of the two theta-phi rectangles; a camera at rest has 1 + z = alpha."""
    c = np.array([70.0, 140.0, 210.0])      # B, G, R
    sky = np.zeros((512, 1024, 3), np.uint8)
    sky[...] = c.astype(np.uint8)
    lv = 9
    ct, cp = 1.5, 2.5
    grid = grid_field(lv, lambda t, p: (ct + 2 * (t - ct), cp + 2 * (p - cp)))
    d = 0.01
    W, H = 10, 8
    view = square_view(ct - H / 2 * d + 0.0013, cp - W / 2 * d + 0.0021, d, W, H)
    out, _ = env_frame(sky, view, grid=grid, lv=lv, alpha=alpha, lr=True)
    P = math.sqrt(0.299 * c[2] ** 2 + 0.587 * c[1] ** 2 + 0.114 * c[0] ** 2) / 255.0
    for i in range(H):
        for j in range(W):
            if i == 0 and j == 0:
                continue      # pixel index 0 stays black
            ta, tb = float(view[i, j, 0]), float(view[i + 1, j, 0])
            pa, pb = float(view[i, j, 1]), float(view[i, j + 1, 1])
            om_pix = (math.cos(ta) - math.cos(tb)) * (pb - pa)
            ta2, tb2 = ct + 2 * (ta - ct), ct + 2 * (tb - ct)
            om_cel = (math.cos(ta2) - math.cos(tb2)) * 2 * (pb - pa)
            frac = om_pix / om_cel
            P2 = min(1.0, (P * frac) ** alpha)
            want = np.minimum(255, c * P2 / P)
            assert np.abs(out[i, j, :3] - want).max() <= 1.01, (i, j, out[i, j], want)


def test_pix_interpolation_through_coarse_blocks_and_captured_points():
    """findBlock, findPoint and interpolateLinear (kernel.cuh:736-938) on an
    adaptive grid holding an affine field: in a region refined only to gap 4
    (the finer points absent, -2) the quadtree descent stops at gap 4 and
    the stencil points missing at that gap are rebuilt from their
    neighbours; next to a captured block (-1) the linear fallback is taken.
    Every corner that is not captured gets the affine field's value."""
    lv = 8
    ct, cp = 1.5, 2.5
    fn = lambda t, p: (ct + 0.8 * (t - ct) + 0.1 * (p - cp), cp + 1.2 * (p - cp) - 0.05 * (t - ct))  # noqa: E731
    grid = grid_field(lv, fn)
    GM = grid.shape[1]
    h = 2 * math.pi / GM
    i0, i1, j0, j1 = 48, 80, 96, 160          # coarse region, aligned to 8
    ii, jj = np.meshgrid(np.arange(i0 + 1, i1), np.arange(j0 + 1, j1), indexing="ij")
    coarse = (ii % 4 != 0) | (jj % 4 != 0)
    grid[ii[coarse], jj[coarse]] = -2.0
    grid[60:62, 176:178] = -1.0               # a captured 2 x 2 block in the fine part
    view = square_view(44.3 * h, 90.7 * h, 0.61 * h, 150, 60)
    _, st = run_frame(grid, lv, view, static_cam(), star=False, map=False, lr=False)
    got = st["thphi"].astype(np.float64)
    want_t, want_p = fn(view[..., 0].astype(np.float64), view[..., 1].astype(np.float64))
    live = got[..., 0] != -1
    assert (~live).sum() > 0 and live.sum() > 0.9 * live.size
    assert np.abs(got[..., 0][live] - want_t[live]).max() < 2e-5
    assert np.abs(got[..., 1][live] - want_p[live]).max() < 2e-5
    # the descent reports the block size it stopped at
    ci = np.floor(view[..., 0] / h).astype(int)
    cj = np.floor(view[..., 1] / h).astype(int)
    inside = (ci > i0 + 1) & (ci < i1 - 2) & (cj > j0 + 1) & (cj < j1 - 2)
    assert (st["gap"][inside] == 4).all() and (st["gap"][~inside & live] <= 4).all()
    # captured corners are exactly those whose cell touches the captured block
    cap_cell = (ci >= 59) & (ci <= 61) & (cj >= 175) & (cj <= 177)
    assert (~live == cap_cell).all()
