"""The committed fixtures of tests/golden/ (make_golden.py) against the GPU
path, through the C-ABI: the seeded two-spin workload's oracle FIDs and the
test_ctx_powder_average.m setup on Spinach's own leb_2ang_rank_5 grid."""
import os

import numpy as np
import pytest

import pyoracle as O
from spinach_amd import _native, grids, spinsys as S, workloads

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


@pytest.mark.parametrize('mode', ['fp64', 'fp64+propagator', 'fp64+krylov'])
def test_static_2spin_fixture_on_the_gpu(mode):
    d = np.load(os.path.join(GOLD, 'static_2spin_fid.npz'))
    w = workloads.static_2spin(npoints=int(d['npoints']))
    sub = grids.subset(w['grid'], d['idx'])
    fid, per, st = _native.powder_fid(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], w['nsteps'], mode=mode,
                                      per_orientation=True)
    e = O.rel_err(fid, d['fid'])
    ep = max(O.rel_err(per[k], d['per'][k]) for k in range(len(d['per'])))
    print(f'[{mode}] 2-spin fixture (64 orientations x 511 steps): sum {e:.2e}, per-orientation {ep:.2e}', flush=True)
    assert e < 1e-10 and ep < 1e-10


def test_ctx_powder_average_fixture_on_the_gpu():
    d = np.load(os.path.join(GOLD, 'ctx_powder_average.npz'))
    gz = np.load(os.path.join(GOLD, 'grids.npz'))
    grid = {k: gz[f'leb_2ang_rank_5__{k}'] for k in ('alphas', 'betas', 'gammas', 'weights')}
    ss = S.create({'magnet': 14.1, 'isotopes': ['1H']},
                  {'zeeman': {'eigs': [[-2.0, -2.0, 4.0]], 'euler': [[0.0, 0.0, 0.0]]}})
    ss = S.basis(ss, {'approximation': 'none', 'projections': [1]})
    I, Q = S.hamiltonian(S.assume(ss, 'nmr'))
    rho = S.state(ss, 'L+', '1H')
    fid, per, _ = _native.powder_fid(I, S.q_list(Q), grid, rho, rho, 1 / 2000, 2, per_orientation=True)
    tol = lambda ref: 1e-12 + 1e-12 * max(1.0, np.linalg.norm(ref))
    assert np.linalg.norm(fid - d['fid']) <= tol(d['fid'])
    assert np.linalg.norm(per - d['per']) <= tol(d['per'])
