"""Seeded synthetic inputs: the render pass's stars and sky map, and the
camera poses of the 64-pose batched grid build.

Stars of the 1080p full-frame workload: uniform on the sphere, magnitude uniform in [0, 8],
temperature log-uniform in [3000, 30000] K.  The reference stores a B-V
colour per star and derives temperature with Ballesteros' formula
(MakeImageKernel.cu:295-297, index = 46/z (1/(0.92 BV + 1.7) + 1/(0.92 BV +
0.62)) - 10 = T/100 - 10 at z = 1), so temperatures are written into the
B-V slot through the inverse: with x = 0.92 BV and k = T / 4600,
k x^2 + (2.32 k - 2) x + (1.054 k - 2.32) = 0, root with x > -0.62.

Sky: an equirectangular RGB8 image (rows = cols / 2) from an integer hash
(value noise on three octaves plus a galactic band), stored BGR like the
reference's cv::Mat.  Everything is a pure function of the seed.
"""
import numpy as np


def temperature_to_bv(T):
    k = np.asarray(T, dtype=np.float64) / 4600.0
    b = 2.32 * k - 2.0
    c = 1.054 * k - 2.32
    x = (-b + np.sqrt(b * b - 4 * k * c)) / (2 * k)
    return x / 0.92


def synthetic_stars(n=100_000, seed=1234):
    """n x 4 float32: (theta, phi, magnitude, B-V)."""
    rng = np.random.default_rng(seed)
    u = rng.random((n, 4))
    theta = np.arccos(1.0 - 2.0 * u[:, 0])
    phi = 2.0 * np.pi * u[:, 1]
    mag = 8.0 * u[:, 2]
    T = 3000.0 * 10.0 ** u[:, 3]
    out = np.stack([theta, phi, mag, temperature_to_bv(T)], axis=1).astype(np.float32)
    # keep phi strictly inside [0, 2 pi) after the float32 cast
    out[:, 1] = np.where(out[:, 1] >= np.float32(2 * np.pi), np.float32(0.0), out[:, 1])
    return out


def _hash(x, y, s):
    h = (x.astype(np.uint32) * np.uint32(0x8DA6B343)) ^ (y.astype(np.uint32) * np.uint32(0xD8163841))
    h ^= np.uint32((int(s) * 0xCB1AB31F) & 0xFFFFFFFF)
    h ^= h >> np.uint32(13)
    h *= np.uint32(0x5BD1E995)
    h ^= h >> np.uint32(15)
    return h


def _value_noise(rows, cols, cell, seed):
    gy, gx = rows // cell + 2, cols // cell + 1
    yy, xx = np.meshgrid(np.arange(gy, dtype=np.uint32), np.arange(gx, dtype=np.uint32), indexing="ij")
    lattice = (_hash(xx % np.uint32(gx - 1 if gx > 1 else 1), yy, seed) & np.uint32(0xFFFF)) / 65535.0
    r = np.arange(rows) / cell
    c = np.arange(cols) / cell
    r0 = np.floor(r).astype(np.int64)
    c0 = np.floor(c).astype(np.int64)
    fr = (r - r0)[:, None]
    fc = (c - c0)[None, :]
    fr = fr * fr * (3 - 2 * fr)
    fc = fc * fc * (3 - 2 * fc)
    v00 = lattice[r0][:, c0]
    v01 = lattice[r0][:, c0 + 1]
    v10 = lattice[r0 + 1][:, c0]
    v11 = lattice[r0 + 1][:, c0 + 1]
    return (v00 * (1 - fc) + v01 * fc) * (1 - fr) + (v10 * (1 - fc) + v11 * fc) * fr


def synthetic_sky(cols=4096, seed=99):
    """(cols/2, cols, 3) uint8, BGR channel order."""
    rows = cols // 2
    n = (0.55 * _value_noise(rows, cols, max(cols // 16, 2), seed)
         + 0.3 * _value_noise(rows, cols, max(cols // 64, 2), seed + 1)
         + 0.15 * _value_noise(rows, cols, max(cols // 256, 2), seed + 2))
    theta = (np.arange(rows) + 0.5) / rows * np.pi
    phi = (np.arange(cols) + 0.5) / cols * 2 * np.pi
    band = np.exp(-((theta[:, None] - np.pi / 2 - 0.35 * np.sin(phi[None, :])) / 0.18) ** 2)
    base = 0.25 * n + 0.55 * band * n
    b = np.clip(40 + 160 * base + 30 * n, 0, 255)
    g = np.clip(25 + 150 * base, 0, 255)
    r = np.clip(20 + 190 * base * (0.6 + 0.4 * n), 0, 255)
    return np.stack([b, g, r], axis=-1).astype(np.uint8)


def batch_poses(n=64, seed=2024):
    """n camera poses of the batched grid build: (a, r, theta, max_level) with
    a ~ U[0, 0.999] (at least 1e-3: Metric.h:133-143 divides by a), r ~ U[3, 50] M,
    theta ~ U[0.1, pi - 0.1], max_level ~ U{6..11} (start level 1)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = max(1e-3, float(rng.uniform(0, 0.999)))
        r = float(rng.uniform(3, 50))
        th = float(rng.uniform(0.1, np.pi - 0.1))
        out.append((a, r, th, int(rng.integers(6, 12))))
    return out
