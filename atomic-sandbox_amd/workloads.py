"""Synthetic inputs of the measured workloads, shared by bench.py and the tests.

neutral_blobs: the Poisson microbenchmark's density, 32 seeded normalised Gaussians (centres uniform in the inner
60 % of the box, sigma uniform in [0.3, 1.0] Angstrom, charges +1 x 16 and -1 x 16, so the net charge is 0), on the
n^3 grid at spacing h (Bohr) in the qm layout.
"""
from __future__ import annotations

import math

import numpy as np

BOHR_PER_ANGSTROM = 1.8897261  # scenes/aspirin.scene:3


def neutral_blobs(n: int, h: float, seed: int = 3, count: int = 32) -> np.ndarray:
    rng = np.random.default_rng(seed)
    L = (n - 1) * h
    cen = rng.uniform(-0.3 * L, 0.3 * L, (count, 3))
    sig = rng.uniform(0.3, 1.0, count) * BOHR_PER_ANGSTROM
    q = np.array([1.0] * (count // 2) + [-1.0] * (count - count // 2))
    ax = -(n - 1) * h / 2.0 + h * np.arange(n)
    rho = np.zeros((n, n, n))
    for c, s_, qq in zip(cen, sig, q):
        gx = np.exp(-(ax - c[0]) ** 2 / (2 * s_ * s_))
        gy = np.exp(-(ax - c[1]) ** 2 / (2 * s_ * s_))
        gz = np.exp(-(ax - c[2]) ** 2 / (2 * s_ * s_))
        rho += qq / (2 * math.pi * s_ * s_) ** 1.5 * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return rho.reshape(-1)
