"""Seeded inputs of the golden fixtures (tests/golden/ref_vectors.npz)."""
import math

import numpy as np

from gse_amd import systems

SIGMA_PAIR = 2.121320344


def pair_case():
    sigma1 = SIGMA_PAIR * math.sqrt(0.4)
    return dict(dims=(20, 20, 20), box=(20.0, 20.0, 20.0), sigma=sigma1,
                rcut=3.0 * math.sqrt(2.0) * sigma1,
                xyzq=systems.pair(1.0, 20.0, orientation="diagonal"),
                phi=np.random.default_rng(101).normal(size=(20, 20, 20)))


def rand_case():
    rng = np.random.default_rng(102)
    a = np.empty((25, 4))
    a[:, :3] = rng.random((25, 3)) * np.array([12.5, 10.0, 14.5])
    a[:, 3] = rng.normal(size=25)
    return dict(dims=(12, 10, 14), box=(12.5, 10.0, 14.5), sigma=1.2,
                rcut=3.0 * math.sqrt(2.0) * 1.2, xyzq=a)


def greens_case():
    rng = np.random.default_rng(103)
    shape = (7, 6, 9 // 2 + 1)
    spec = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    return dict(dims=(9, 6, 7), box=(9.5, 6.25, 7.0), sigma=0.9487, spec=spec)


def conv_case():
    R = 5
    return dict(dims=(16, 8, 6), R=R,
                kx=np.exp(-np.arange(-R, R + 1) ** 2 / (2 * 2.12 ** 2)),
                grid=np.random.default_rng(104).normal(size=(6, 8, 16)))


def ongrid_case():
    sigma = 0.9487
    return dict(dims=(12, 10, 14), box=(12.5, 10.0, 13.0), sigma=sigma,
                rcut=3.0 * math.sqrt(2.0) * sigma,
                grid=np.random.default_rng(105).normal(size=(14, 10, 12)))
