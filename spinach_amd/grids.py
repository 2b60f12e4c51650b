"""Orientation grids for the powder context.

Each grid is a dict of float64 arrays alphas, betas, gammas, weights (the
fields powder() loads from kernel/grids/*.mat, kernel/contexts/powder.m:173-179).
"""
from __future__ import annotations

import math
import os

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


_FILE_CACHE = {}


def load_grid(name, root=None):
    """A Spinach grid file read as powder.m:173-179 does
    (load([root_dir '/kernel/grids/' parameters.grid], 'alphas', 'betas',
    'gammas', 'weights')): name is a grid file name (with or without .mat)
    under <root>/kernel/grids, root defaulting to $SPINACH_ROOT, or a path to
    a .mat file. Read with scipy.io.loadmat, which loads data only; the
    result is cached per file."""
    import scipy.io as sio
    path = name if os.path.isabs(name) or os.path.sep in name else None
    if path is None:
        root = root or os.environ.get('SPINACH_ROOT')
        if not root:
            raise ValueError(f'grid {name!r}: set SPINACH_ROOT to the Spinach root directory or pass a path')
        path = os.path.join(root, 'kernel', 'grids', name)
    if not path.endswith('.mat'):
        path += '.mat'
    path = os.path.abspath(path)
    if path not in _FILE_CACHE:
        d = sio.loadmat(path, variable_names=['alphas', 'betas', 'gammas', 'weights'])
        g = {}
        for k in ('alphas', 'betas', 'gammas', 'weights'):
            if k not in d:
                raise ValueError(f'{path}: no {k} array')
            g[k] = np.ascontiguousarray(np.asarray(d[k], dtype=np.float64).ravel())
        n = len(g['weights'])
        if n == 0 or any(len(v) != n for v in g.values()):
            raise ValueError(f'{path}: alphas, betas, gammas and weights must have one common positive length')
        if not all(np.all(np.isfinite(v)) for v in g.values()):
            raise ValueError(f'{path}: non-finite grid entry')
        _FILE_CACHE[path] = g
    return {k: v.copy() for k, v in _FILE_CACHE[path].items()}
