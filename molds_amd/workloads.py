"""Deterministic synthetic molecules for the benchmark and tests.

One RNG everywhere: SplitMix64-seeded xoshiro256** (Blackman & Vigna) with
Box-Muller normals, written out here so every platform draws the same numbers.

* alkane(n, seed, jitter)      all-trans C_nH_{2n+2}: C-C 1.54 A, C-H 1.09 A,
                               tetrahedral angles, zig-zag carbons in the xy plane
* water_cluster(m, seed, ...)  (H2O)_m: O uniform in a cube of 29.9 A^3 per
                               molecule with O-O >= 2.6 A (seeded rejection),
                               O-H 0.96 A, HOH 104.5 deg, uniform random orientation
                               (Shoemake quaternions)
* batch_mix(count, seed)       even index -> alkane n ~ U{1..40}, odd -> water m ~ U{2..64}
* poisson_requests(...)        arrivals and log-uniform sizes on [5, 600] atoms
Coordinates are returned in Angstrom; to_bohr() converts with MolDS's constant.
"""
from __future__ import annotations

import math

import numpy as np

M64 = (1 << 64) - 1
ANG2AU = 1.0 / 0.5291772
TETRA = 109.4712206344907   # degrees


def _splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return x, z ^ (z >> 31)


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & M64


class Rng:
    """xoshiro256** seeded by SplitMix64."""

    def __init__(self, seed):
        x = seed & M64
        s = []
        for _ in range(4):
            x, z = _splitmix64(x)
            s.append(z)
        self.s = s
        self._spare = None

    def next_u64(self):
        s = self.s
        result = (_rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def uniform(self):
        """[0, 1) with 53 random bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))

    def normal(self):
        """Box-Muller; returns both variates in turn."""
        if self._spare is not None:
            v, self._spare = self._spare, None
            return v
        u1 = 1.0 - self.uniform()   # (0, 1]
        u2 = self.uniform()
        r = math.sqrt(-2.0 * math.log(u1))
        self._spare = r * math.sin(2.0 * math.pi * u2)
        return r * math.cos(2.0 * math.pi * u2)


def mix_seed(*parts):
    """Seed for an item of a config: SplitMix64 over the integer parts."""
    x = 0x2545F4914F6CDD1D
    for p in parts:
        x, z = _splitmix64((x ^ (p & M64)) & M64)
        x = z
    return x


def _unit(v):
    return v / np.linalg.norm(v)


def _rodrigues(v, axis, ang):
    k = _unit(axis)
    return v * math.cos(ang) + np.cross(k, v) * math.sin(ang) + k * np.dot(k, v) * (1 - math.cos(ang))


def alkane(n, seed=0, jitter=0.02):
    """All-trans C_nH_{2n+2}; returns (Z, xyz[A]) with carbons first."""
    cc, ch = 1.54, 1.09
    half = math.radians(TETRA) / 2.0
    if n == 1:
        C = np.zeros((1, 3))
        d = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float) / math.sqrt(3.0)
        H = C[0] + ch * d
    else:
        C = np.array([[i * cc * math.sin(half), (i % 2) * cc * math.cos(half), 0.0] for i in range(n)])
        H = []
        for i in range(n):
            if 0 < i < n - 1:
                b = -_unit((C[i - 1] - C[i]) + (C[i + 1] - C[i]))
                for sgn in (1.0, -1.0):
                    d = math.cos(half) * b + sgn * math.sin(half) * np.array([0, 0, 1.0])
                    H.append(C[i] + ch * _unit(d))
            else:
                nb = C[1] if i == 0 else C[n - 2]
                # trans hydrogen: where the next zig-zag carbon would sit
                step = 1 if i == n - 1 else -1
                ghost = np.array([(i + step) * cc * math.sin(half), ((i + step) % 2) * cc * math.cos(half), 0.0])
                da = _unit(ghost - C[i])
                axis = C[i] - nb
                for ang in (0.0, 2 * math.pi / 3, -2 * math.pi / 3):
                    H.append(C[i] + ch * _unit(_rodrigues(da, axis, ang)))
        H = np.array(H)
    xyz = np.vstack([C, H])
    Z = np.array([6] * len(C) + [1] * len(H), dtype=np.int32)
    if jitter:
        rng = Rng(seed)
        xyz = xyz + np.array([[rng.normal() * jitter for _ in range(3)] for _ in range(len(xyz))])
    return Z, xyz


def _quat_rotation(rng):
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    q = (a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
         b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(m, seed=1, jitter=0.0, vol_per_molecule=29.9, min_oo=2.6):
    """(H2O)_m; returns (Z, xyz[A]) ordered O, H, H per molecule."""
    rng = Rng(seed)
    L = (m * vol_per_molecule) ** (1.0 / 3.0)
    O = np.zeros((0, 3))
    tries = 0
    while len(O) < m:
        p = np.array([rng.uniform() * L, rng.uniform() * L, rng.uniform() * L])
        tries += 1
        if len(O) == 0 or np.min(np.sum((O - p) ** 2, axis=1)) >= min_oo * min_oo:
            O = np.vstack([O, p])
        if tries > 2000 * m + 100000:
            raise RuntimeError("water cluster packing failed")
    half = math.radians(104.5) / 2.0
    body = np.array([[math.sin(half), 0.0, math.cos(half)], [-math.sin(half), 0.0, math.cos(half)]]) * 0.96
    xyz = []
    for i in range(m):
        R = _quat_rotation(rng)
        xyz.append(O[i])
        for h in body:
            xyz.append(O[i] + R @ h)
    xyz = np.array(xyz)
    Z = np.array([8, 1, 1] * m, dtype=np.int32)
    if jitter:
        xyz = xyz + np.array([[rng.normal() * jitter for _ in range(3)] for _ in range(len(xyz))])
    return Z, xyz


def to_bohr(xyz_ang):
    return np.asarray(xyz_ang) * ANG2AU


def decane(seed=0):
    """n-decane C10H22 with N(0, 0.02 A) jitter."""
    return alkane(10, seed=seed, jitter=0.02)


def batch_mix(count=4096, base_seed=2, jitter=0.02):
    """Uneven batch: even i -> alkane n ~ U{1..40}; odd i -> (H2O)_m, m ~ U{2..64}.
    Yields (kind, size, Z, xyz[A])."""
    for i in range(count):
        rng = Rng(mix_seed(base_seed, i))
        if i % 2 == 0:
            n = rng.randint(1, 40)
            Z, xyz = alkane(n, seed=mix_seed(base_seed, i, 1), jitter=jitter)
            yield ("alkane", n, Z, xyz)
        else:
            m = rng.randint(2, 64)
            Z, xyz = water_cluster(m, seed=mix_seed(base_seed, i, 1), jitter=jitter)
            yield ("water", m, Z, xyz)


def poisson_requests(rate, count, seed=3, lo=5, hi=600):
    """Poisson arrivals at `rate` per second; sizes log-uniform on [lo, hi] atoms,
    mapped to the nearest alkane (3n+2 atoms) or water cluster (3m atoms) by a
    fair coin.  Yields (t_arrival_s, kind, size, Z, xyz[A])."""
    rng = Rng(mix_seed(seed, int(rate)))
    t = 0.0
    for i in range(count):
        t += -math.log(1.0 - rng.uniform()) / rate
        atoms = math.exp(math.log(lo) + rng.uniform() * (math.log(hi) - math.log(lo)))
        if rng.uniform() < 0.5:
            n = max(1, int(round((atoms - 2) / 3.0)))
            Z, xyz = alkane(n, seed=mix_seed(seed, i, 1), jitter=0.02)
            yield (t, "alkane", n, Z, xyz)
        else:
            m = max(2, int(round(atoms / 3.0)))
            Z, xyz = water_cluster(m, seed=mix_seed(seed, i, 1), jitter=0.02)
            yield (t, "water", m, Z, xyz)
