"""Orientation grids for the powder context.

Each grid is a dict of float64 arrays alphas, betas, gammas, weights (the
fields powder() loads from kernel/grids/*.mat, kernel/contexts/powder.m:173-179).
"""
from __future__ import annotations

import math

import numpy as np


def zcwn(n):
    """Craciun's two-angle spiral grid with equal weights
    (kernel/grids/grid_fibon.m:70-78): beta_k = acos(2k/n - 1),
    gamma_k = 2*pi*k/phi^2, alpha_k = 0."""
    k = np.arange(n, dtype=np.float64)
    phi = (1 + math.sqrt(5)) / 2
    return dict(alphas=np.zeros(n), betas=np.arccos(2 * k / n - 1), gammas=2 * math.pi * k / phi ** 2,
                weights=np.full(n, 1.0 / n))


def haar(n, rng):
    """n rotations drawn uniformly on SO(3) (Haar measure) with equal weights:
    alpha, gamma ~ U[0, 2pi), beta = arccos(U[-1, 1])."""
    a = rng.uniform(0, 2 * math.pi, n)
    b = np.arccos(rng.uniform(-1, 1, n))
    g = rng.uniform(0, 2 * math.pi, n)
    return dict(alphas=a, betas=b, gammas=g, weights=np.full(n, 1.0 / n))


def single_crystal(alpha=0.0, beta=0.0, gamma=0.0):
    """One orientation with unit weight (kernel/grids/single_crystal.mat)."""
    return dict(alphas=np.array([alpha], float), betas=np.array([beta], float),
                gammas=np.array([gamma], float), weights=np.array([1.0]))


def subset(grid, idx):
    idx = np.asarray(idx)
    return {k: np.asarray(v)[idx] for k, v in grid.items()}
