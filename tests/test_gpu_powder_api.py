"""The host mirror of the reference call, powder(spin_system, @acquire,
parameters, 'nmr') (spinach_amd/powder.py), on the GPU: the weighted sum,
sum_up = false, dead_time, a zcwn grid by name, and sys.enable = {'krylov'},
against the two-spin closed form and the oracle."""
import numpy as np
import pytest

import pyoracle as O
from spinach_amd import grids, workloads
from spinach_amd import powder as P
from spinach_amd import spinsys as S

pytestmark = pytest.mark.gpu


def _params(w, **kw):
    p = {'rho0': w['rho0'], 'coil': w['coil'], 'sweep': 1.0 / w['dt'], 'npoints': w['nsteps'] + 1,
         'grid': w['grid']}
    p.update(kw)
    return p


def test_powder_call_sum_and_per_orientation():
    w = workloads.static_2spin(npoints=256, n_orient=32)
    fid, grid = P.powder(w['ss'], P.acquire, _params(w), 'nmr')
    cf = O.closed_form_2spin(w['ss'], grid, w['dt'], w['nsteps'])
    e = O.rel_err(fid, cf)
    per, _ = P.powder(w['ss'], P.acquire, _params(w, sum_up=False), 'nmr')
    print(f'powder(): vs closed form {e:.2e}; {len(per)} per-orientation FIDs', flush=True)
    assert e < 5e-9
    assert len(per) == 32 and O.rel_err(np.sum([wt * f for wt, f in zip(grid['weights'], per)], axis=0), fid) < 1e-14


def test_powder_call_dead_time_named_grid_and_krylov():
    w = workloads.static_2spin(npoints=128, n_orient=16)
    fid, grid = P.powder(w['ss'], P.acquire, _params(w, grid='zcwn_144', dead_time=12e-6), 'nmr')
    assert len(grid['weights']) == 144
    cf = O.closed_form_2spin(w['ss'], grid, w['dt'], w['nsteps'], t0=12e-6)
    ss_k = dict(w['ss'])
    ss_k['enable'] = ['krylov']
    fk, _ = P.powder(ss_k, P.acquire, _params(w, grid='zcwn_144', dead_time=12e-6), 'nmr')
    e, ek = O.rel_err(fid, cf), O.rel_err(fk, fid)
    print(f'powder() zcwn_144 + dead time: vs shifted closed form {e:.2e}; krylov vs default path {ek:.2e}',
          flush=True)
    assert e < 5e-9 and ek < 1e-12
