"""Seeded synthetic powder workloads (static solids under 'nmr').

static_2spin  two-spin 1H-13C, complete basis (dim 16), zcwn 256 grid,
              2048 points at 10 us
static_8spin  5x1H, 2x13C, 1x15N, IK-0 level 3 (dim 1789), 4000 Haar
              orientations, 4096 steps at 5 us
batch_uneven  512 systems of 2-10 spins (1H/13C/15N), IK-0 level <=3,
              100-1000 Haar orientations each, 1024 steps at 5 us
mas_6spin     4x13C + 2x1H in Hilbert space (dim 64), magic-angle spinning at
              12.5 kHz, 128 rotor phases per period, 64 periods, 3000 Haar
              orientations

Every system draws CSA eigenvalues uniformly in [-100,100] ppm with Euler
angles uniform on SO(3); dipolar couplings come from coordinates.
"""
from __future__ import annotations

import math

import numpy as np

from . import grids
from . import spinsys as S


def _euler(rng):
    return [rng.uniform(0, 2 * math.pi), math.acos(rng.uniform(-1, 1)), rng.uniform(0, 2 * math.pi)]


def _coords(rng, n, box=4.0, min_sep=1.0):
    pts = []
    tries = 0
    while len(pts) < n:
        p = rng.uniform(0, box, 3)
        tries += 1
        if all(np.linalg.norm(p - q) >= min_sep for q in pts):
            pts.append(p)
        if tries > 100000:
            raise RuntimeError('could not place spins')
    return pts


def build(sys, inter, bas, channel='1H'):
    """create -> assume('nmr') -> basis -> hamiltonian; rho0 = Lx and coil =
    L+ on every spin of `channel` (state(ss,'Lx','1H') etc)."""
    ss = S.create(sys, inter)
    ss = S.basis(ss, bas)
    ssa = S.assume(ss, 'nmr')
    I, Q = S.hamiltonian(ssa)
    rho0 = S.state(ss, 'Lx', channel)
    coil = S.state(ss, 'L+', channel)
    return dict(ss=ss, I=I, Q=S.q_list(Q), Qnested=Q, rho0=rho0, coil=coil)


def static_2spin(seed=0, npoints=2048, n_orient=256, relax=None):
    """relax: optional relaxation fields of inter (relaxation, r1_rates,
    r2_rates, damp_rate, rlx_keep, equilibrium), as in kernel/create.m."""
    rng = np.random.default_rng(seed)
    sys = {'magnet': 14.1, 'isotopes': ['1H', '13C']}
    d = rng.normal(size=3)
    d = 1.1 * d / np.linalg.norm(d)
    inter = {'zeeman': {'eigs': [list(rng.uniform(-100, 100, 3)) for _ in range(2)],
                        'euler': [_euler(rng) for _ in range(2)]},
             'coupling': {'scalar': {(0, 1): 140.0}},
             'coordinates': [np.zeros(3), d]}
    inter.update(relax or {})
    w = build(sys, inter, {'formalism': 'sphten-liouv', 'approximation': 'none'})
    w.update(grid=grids.zcwn(n_orient), dt=1e-5, nsteps=npoints - 1, name='static_2spin')
    return w


def static_8spin(seed=1, nsteps=4096, n_orient=4000, relax=None):
    rng = np.random.default_rng(seed)
    iso = ['1H'] * 5 + ['13C'] * 2 + ['15N']
    sys = {'magnet': 14.1, 'isotopes': iso}
    inter = {'zeeman': {'eigs': [list(rng.uniform(-100, 100, 3)) for _ in iso],
                        'euler': [_euler(rng) for _ in iso]},
             'coordinates': _coords(rng, len(iso))}
    inter.update(relax or {})
    w = build(sys, inter, {'formalism': 'sphten-liouv', 'approximation': 'IK-0', 'level': 3})
    w.update(grid=grids.haar(n_orient, rng), dt=5e-6, nsteps=nsteps, name='static_8spin')
    return w


def random_system(rng, nspins, level=3):
    iso = list(rng.choice(['1H', '13C', '15N'], size=nspins))
    if '1H' not in iso:
        iso[0] = '1H'  # the coil channel must exist
    sys = {'magnet': 14.1, 'isotopes': iso}
    scalar = {}
    for a in range(nspins):
        for b in range(a + 1, nspins):
            if rng.uniform() < 0.3:
                scalar[(a, b)] = float(rng.uniform(-20, 200))
    inter = {'zeeman': {'eigs': [list(rng.uniform(-100, 100, 3)) for _ in iso],
                        'euler': [_euler(rng) for _ in iso]},
             'coupling': {'scalar': scalar},
             'coordinates': _coords(rng, nspins, box=2.0 + nspins ** (1 / 3) * 1.5)}
    appr = {'formalism': 'sphten-liouv', 'approximation': 'IK-0', 'level': min(level, nspins)}
    return build(sys, inter, appr)


def _batch_member(args):
    seed, s, min_spins, max_spins, min_or, max_or = args
    rng = np.random.default_rng([seed, s])
    n = int(rng.integers(min_spins, max_spins + 1))
    w = random_system(rng, n)
    n_or = int(rng.integers(min_or, max_or + 1))
    w['grid'] = grids.haar(n_or, rng)
    w['nspins'] = n
    del w['ss'], w['Qnested']
    return w


def batch_member(seed, s, min_or=100, max_or=1000, min_spins=2, max_spins=10):
    """Member s of batch_uneven(seed, ...) on its own (same draws, so member
    363 of the 512-system batch can be rebuilt without the other 511)."""
    return _batch_member((seed, s, min_spins, max_spins, min_or, max_or))


def batch_uneven(seed=2, n_systems=512, nsteps=1024, min_or=100, max_or=1000, min_spins=2, max_spins=10,
                 workers=None):
    """Every member system s is drawn from its own generator seeded with
    (seed, s), so the batch does not depend on how it is built."""
    jobs = [(seed, s, min_spins, max_spins, min_or, max_or) for s in range(n_systems)]
    if workers is None:
        import os
        workers = min(16, len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else 4)
    if workers > 1 and n_systems > 8:
        import multiprocessing as mp
        with mp.get_context('fork').Pool(workers) as pool:
            out = pool.map(_batch_member, jobs, chunksize=4)
    else:
        out = [_batch_member(j) for j in jobs]
    return dict(systems=out, dt=5e-6, nsteps=nsteps, name='batch_uneven')


def mas_6spin(seed=3, n_orient=3000, periods=64, ticks=128, rate_hz=12500.0):
    """Hilbert-space MAS workload: rho0 = Lx(13C), coil = L+(13C) (carbon
    detection), rotor axis [1 1 1] (magic angle), 'magnet' frame."""
    rng = np.random.default_rng(seed)
    iso = ['13C'] * 4 + ['1H'] * 2
    sys = {'magnet': 14.1, 'isotopes': iso}
    inter = {'zeeman': {'eigs': [list(rng.uniform(-100, 100, 3)) for _ in iso],
                        'euler': [_euler(rng) for _ in iso]},
             'coordinates': _coords(rng, len(iso), box=3.5, min_sep=1.0)}
    ss = S.create(sys, inter)
    ss = S.basis(ss, {'formalism': 'zeeman-hilb'})
    I, Q = S.hamiltonian(S.assume(ss, 'nmr'))
    rho0 = S.state(ss, 'Lx', '13C')
    coil = S.state(ss, 'L+', '13C')
    return dict(ss=ss, I=I, Q=S.q_list(Q), Qnested=Q, rho0=rho0, coil=coil, grid=grids.haar(n_orient, rng),
                axis=(1.0, 1.0, 1.0), rate_hz=rate_hz, ticks=ticks, nsteps=ticks * periods,
                dt=1.0 / (rate_hz * ticks), name='mas_6spin')
