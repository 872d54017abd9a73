"""Seeded molecular geometries of the benchmark workloads (Angstrom in, bohr out).

* ``water()``               -- rigid H2O: O at the origin, O-H 0.9572 A, H-O-H 104.52 deg
* ``water_lattice(seed)``   -- 8 rigid waters on a 2x2x2 cubic lattice, 3.0 A
                               spacing, +-0.3 A uniform jitter, uniformly random
                               orientations (Shoemake quaternions)
* ``water_box(n, seed)``    -- n waters placed by rejection sampling in a 6 A box,
                               minimum O-O distance 2.6 A, random orientations
* ``alkane(n)``             -- all-trans CnH2n+2, C-C 1.54 A, C-H 1.09 A,
                               tetrahedral angles

Coordinates are converted to bohr with the reference's A_to_a0 = 1.889726126
(src/chem/Constants.hpp:15).
"""
from __future__ import annotations

import math

import numpy as np

A_TO_A0 = 1.889726126
R_OH = 0.9572
THETA_HOH = math.radians(104.52)


def water_frame() -> np.ndarray:
    s, c = math.sin(THETA_HOH / 2), math.cos(THETA_HOH / 2)
    return np.array([[0.0, 0.0, 0.0], [R_OH * s, R_OH * c, 0.0], [-R_OH * s, R_OH * c, 0.0]])


def water():
    return np.array([8, 1, 1], dtype=np.int32), water_frame() * A_TO_A0


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform random rotation from a Shoemake quaternion."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _place(origins, rng):
    Z, xyz = [], []
    frame = water_frame()
    for o in origins:
        R = random_rotation(rng)
        for k in range(3):
            Z.append(8 if k == 0 else 1)
            xyz.append(o + R @ frame[k])
    return np.array(Z, dtype=np.int32), np.array(xyz) * A_TO_A0


def water_lattice(seed: int = 1, spacing: float = 3.0, jitter: float = 0.3):
    rng = np.random.default_rng(seed)
    origins = []
    for i in range(2):
        for j in range(2):
            for k in range(2):
                origins.append(np.array([i, j, k], float) * spacing + rng.uniform(-jitter, jitter, 3))
    return _place(origins, rng)


def water_box(n: int, seed: int, box: float = 6.0, min_oo: float = 2.6, max_tries: int = 100000):
    rng = np.random.default_rng(seed)
    origins = []
    tries = 0
    while len(origins) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place waters")
        p = rng.uniform(0.0, box, 3)
        if all(np.linalg.norm(p - q) >= min_oo for q in origins):
            origins.append(p)
    return _place(origins, rng)


def alkane(n: int):
    """All-trans CnH2n+2: zigzag carbon backbone in the xy plane (C-C 1.54 A,
    C-C-C 109.47 deg); each carbon's hydrogens (C-H 1.09 A) sit on the
    tetrahedral directions left by its two chain neighbours, with virtual
    neighbours continuing the zigzag past the chain ends (which gives the
    terminal in-plane hydrogen)."""
    cc, ch = 1.54, 1.09
    tet = math.radians(109.4712206)
    dx, dy = cc * math.sin(tet / 2), cc * math.cos(tet / 2)

    def carbon(i):
        return np.array([i * dx, (i % 2) * dy, 0.0])

    if n == 1:
        t = ch / math.sqrt(3.0)
        xyz = [np.zeros(3)] + [np.array(v, float) * t for v in [(1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)]]
        return np.array([6, 1, 1, 1, 1], dtype=np.int32), np.array(xyz) * A_TO_A0
    Z, xyz = [], []
    for i in range(n):
        Z.append(6)
        xyz.append(carbon(i))
    for i in range(n):
        c = carbon(i)
        b1 = carbon(i - 1) - c
        b2 = carbon(i + 1) - c
        u = -(b1 + b2)
        u /= np.linalg.norm(u)
        # the two out-of-plane H: H-C-H angle tetrahedral, bisected by u
        for sz in (1.0, -1.0):
            d = math.cos(tet / 2) * u + math.sin(tet / 2) * np.array([0.0, 0.0, sz])
            Z.append(1)
            xyz.append(c + ch * d)
        if i == 0 or i == n - 1:  # terminal: H where the next chain carbon would be
            v = b1 if i == 0 else b2
            Z.append(1)
            xyz.append(c + ch * v / np.linalg.norm(v))
    return np.array(Z, dtype=np.int32), np.array(xyz) * A_TO_A0
