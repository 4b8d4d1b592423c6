"""Seeded synthetic particle systems for the GSE workloads.

All generators return float64 arrays of shape (n, 4): x, y, z, q (the
reference's xyzq_t record, TypesGSE.h:20-22).  Seeds use numpy's PCG64
Generator, so a (generator, seed) pair names one input exactly.
"""
import math

import numpy as np

# TIP3P-like rigid triplet
Q_O = -0.834
Q_H = 0.417
R_OH = 0.9572          # Angstrom
THETA_HOH = 104.52     # degrees
TRIPLET_DENSITY = 0.0334  # triplets per A^3


def pair(r=1.0, L=20.0, seed=1, orientation="random"):
    """Two charges -1 at the origin and +1 at r*u.

    orientation="random": u uniform on the sphere from `seed`.
    orientation="diagonal": u = (1,1,1)/sqrt(3), the placement of the
    reference's testPair (cpu_gse.cpp:689-698).
    """
    if orientation == "diagonal":
        u = np.ones(3) / math.sqrt(3.0)
    else:
        rng = np.random.default_rng(seed)
        v = rng.normal(size=3)
        u = v / np.linalg.norm(v)
    a = np.zeros((2, 4))
    a[0, 3] = -1.0
    a[1, :3] = r * u
    a[1, 3] = 1.0
    return a


def random_rotations(rng, n):
    """n uniform random rotation matrices from unit quaternions (Shoemake)."""
    u1, u2, u3 = rng.random((3, n))
    q0 = np.sqrt(1 - u1) * np.sin(2 * np.pi * u2)
    q1 = np.sqrt(1 - u1) * np.cos(2 * np.pi * u2)
    q2 = np.sqrt(u1) * np.sin(2 * np.pi * u3)
    q3 = np.sqrt(u1) * np.cos(2 * np.pi * u3)
    w, x, y, z = q3, q0, q1, q2
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_box(n_triplets, L, seed=0, wrap=True):
    """Neutral box of rigid TIP3P-like triplets, centres uniform in [0,L)^3,
    uniformly random orientations; every atom wrapped into [0, L)."""
    rng = np.random.default_rng(seed)
    centres = rng.random((n_triplets, 3)) * L
    R = random_rotations(rng, n_triplets)
    half = math.radians(THETA_HOH) / 2.0
    body = np.array([[0.0, 0.0, 0.0],
                     [R_OH * math.sin(half), 0.0, R_OH * math.cos(half)],
                     [-R_OH * math.sin(half), 0.0, R_OH * math.cos(half)]])
    pos = centres[:, None, :] + np.einsum("nij,aj->nai", R, body)
    if wrap:
        pos = np.mod(pos, L)
        pos[pos >= L] -= L  # np.mod can return L for tiny negatives
    out = np.empty((n_triplets, 3, 4))
    out[:, :, :3] = pos
    out[:, 0, 3] = Q_O
    out[:, 1:, 3] = Q_H
    return out.reshape(-1, 4)


def random_alternating(n, L, seed=1234567890):
    """N charges uniform in the box with q = (i%2)*2-1 (the charge pattern
    of the reference's testRandom, cpu_gse.cpp:433-438)."""
    rng = np.random.default_rng(seed)
    a = np.empty((n, 4))
    a[:, :3] = rng.random((n, 3)) * L
    a[:, 3] = (np.arange(n) % 2) * 2 - 1
    return a


GRID_CHOICES = (24, 32, 40, 48, 64)


def nearest_grid(L, choices=GRID_CHOICES):
    """Nearest of `choices` to L; ties go to the smaller grid."""
    best = None
    for m in sorted(choices):
        if best is None or abs(m - L) < abs(best - L):
            best = m
    return best


def batched_systems(n_systems=256, seed=4, nmin=600, nmax=8000):
    """Independent neutral water boxes with uneven sizes: atom counts uniform
    in [nmin, nmax] (multiples of 3), L set for TRIPLET_DENSITY, M the
    nearest grid of GRID_CHOICES to L.  Returns a list of (xyzq, L, M)."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(nmin // 3, nmax // 3 + 1, size=n_systems) * 3
    out = []
    for k, n in enumerate(counts):
        nt = int(n) // 3
        L = (nt / TRIPLET_DENSITY) ** (1.0 / 3.0)
        M = nearest_grid(L)
        out.append((water_box(nt, L, seed=seed * 100003 + k), L, M))
    return out


# Named workloads (sigma and grid as the configurations state them)
WORKLOADS = {
    "pair_L20_M20": dict(kind="pair", L=20.0, M=20, sigma=2.121320344, r=1.0, seed=1),
    "water_48k_M80": dict(kind="water", n_triplets=16000, L=78.3, M=80, sigma=2.12, seed=2),
    "water_3k_M32": dict(kind="water", n_triplets=1000, L=31.0, M=32, sigma=2.12, seed=3),
    "water_1M_M216": dict(kind="water", n_triplets=333334, L=215.3, M=216, sigma=2.12, seed=5),
}


def make(name):
    w = WORKLOADS[name]
    if w["kind"] == "pair":
        xyzq = pair(w["r"], w["L"], seed=w["seed"])
    else:
        xyzq = water_box(w["n_triplets"], w["L"], seed=w["seed"])
    return xyzq, w["L"], w["M"], w["sigma"]
