"""Seeded synthetic inputs: ellipsoid tables and a portable random generator.

Ellipsoid records are 16 floats: centre[3], semi_axes[3], R[9] (row-major,
body coordinates r = R (p - centre)), value.  The generator is splitmix64
seeding xoshiro256**, so a seed gives the same numbers on every platform.
"""
from __future__ import annotations

import math

import numpy as np

_MASK = (1 << 64) - 1


def _rotl(x: int, k: int) -> int:
    return ((x << k) | (x >> (64 - k))) & _MASK


class Xoshiro256:
    """xoshiro256** seeded through splitmix64 (Blackman & Vigna, 2018)."""

    def __init__(self, seed: int):
        s = seed & _MASK
        self.s = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & _MASK
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
            self.s.append(z ^ (z >> 31))

    def next(self) -> int:
        s = self.s
        result = (_rotl((s[1] * 5) & _MASK, 7) * 9) & _MASK
        t = (s[1] << 17) & _MASK
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def normal(self) -> float:
        # Box-Muller
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def _rotation_z(deg: float) -> np.ndarray:
    t = math.radians(deg)
    c, s = math.cos(t), math.sin(t)
    return np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])


# Modified 3-D Shepp-Logan table (ASTRA generateSheppLogan3D as used by the
# reference's tests: Yktest/YkTestPhantoms.hpp:51-69, dlltest/test_shepp_logan.cpp:59-65):
# x, y, z, a, b, c (normalized), rotation in degrees about z, modified value
SHEPP_LOGAN = [
    (0.0, 0.0, 0.0, 0.69, 0.92, 0.81, 0.0, 1.0),
    (0.0, -0.0184, 0.0, 0.6624, 0.874, 0.78, 0.0, -0.8),
    (0.22, 0.0, 0.0, 0.11, 0.31, 0.22, -18.0, -0.2),
    (-0.22, 0.0, 0.0, 0.16, 0.41, 0.28, 18.0, -0.2),
    (0.0, 0.35, 0.0, 0.21, 0.25, 0.41, 0.0, 0.1),
    (0.0, 0.1, 0.0, 0.046, 0.046, 0.05, 0.0, 0.1),
    (0.0, -0.1, 0.0, 0.046, 0.046, 0.05, 0.0, 0.1),
    (-0.08, -0.605, 0.0, 0.046, 0.023, 0.05, 0.0, 0.1),
    (0.0, -0.605, 0.0, 0.023, 0.023, 0.02, 0.0, 0.1),
    (0.06, -0.605, 0.0, 0.023, 0.046, 0.02, 0.0, 0.1),
]


def shepp_logan(half_extent_mm: float, scale: float = 0.02) -> np.ndarray:
    """Shepp-Logan ellipsoids in world mm.  The table's normalized frame has
    y and z pointing against the world axes (px = x/h, py = -y/h, pz = -z/h)."""
    out = []
    h = half_extent_mm
    for (x, y, z, a, b, c, deg, val) in SHEPP_LOGAN:
        t = math.radians(deg)
        co, si = math.cos(t), math.sin(t)
        # body = [[c, -s], [s, c]] applied to normalized offsets (dx, dy) with
        # dx = dwx / h, dy = -dwy / h, dz = -dwz / h
        R = np.array([[co, si, 0.0], [si, -co, 0.0], [0.0, 0.0, -1.0]])
        rec = [h * x, -h * y, -h * z, h * a, h * b, h * c, *R.ravel().tolist(), val * scale]
        out.append(rec)
    return np.array(out, dtype=np.float64)


def random_rotation(rng: Xoshiro256) -> np.ndarray:
    """Uniform random rotation from a unit quaternion (Shoemake)."""
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    q = (math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
         math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3))
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def random_ellipsoids(seed: int, count: int, semi_axes=(5.0, 40.0), density=(0.0, 0.05),
                      center_radius=50.0) -> np.ndarray:
    """`count` ellipsoids: semi-axes uniform in `semi_axes` mm, values uniform in
    `density` mm^-1 (additive), centres uniform in a ball of `center_radius` mm,
    uniformly random orientations."""
    rng = Xoshiro256(seed)
    out = []
    for _ in range(count):
        while True:
            c = [rng.uniform(-1.0, 1.0) for _ in range(3)]
            if c[0] ** 2 + c[1] ** 2 + c[2] ** 2 <= 1.0:
                break
        centre = [center_radius * v for v in c]
        axes = [rng.uniform(*semi_axes) for _ in range(3)]
        R = random_rotation(rng)
        val = rng.uniform(*density)
        out.append([*centre, *axes, *R.ravel().tolist(), val])
    return np.array(out, dtype=np.float64)


def sphere(radius_mm: float, value: float, centre=(0.0, 0.0, 0.0)) -> np.ndarray:
    return np.array([[*centre, radius_mm, radius_mm, radius_mm, *np.eye(3).ravel().tolist(), value]])
