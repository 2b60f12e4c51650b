"""GPU parity: libspinach_powder (through the C-ABI) against the oracle."""
import numpy as np
import pytest

import pyoracle as O
from spinach_amd import _native, grids, workloads

pytestmark = pytest.mark.gpu


def _run(w, grid=None, mode='fp64', per=False):
    g = grid if grid is not None else w['grid']
    return _native.powder_fid(w['I'], w['Q'], g, w['rho0'], w['coil'], w['dt'], w['nsteps'], mode=mode,
                              per_orientation=per)


def test_static_2spin_small_vs_oracle():
    w = workloads.static_2spin(npoints=128, n_orient=16)
    fid, _, st = _run(w)
    print('stats', st, flush=True)
    ref = O.refcpu_powder(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'])
    assert O.rel_err(fid, ref) < 1e-10
    assert st['reduced_dim'] == 2


def test_static_2spin_full_vs_oracle_and_closed_form():
    w = workloads.static_2spin()
    fid, _, st = _run(w)
    ref = O.refcpu_powder(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'])
    cf = O.closed_form_2spin(w['ss'], w['grid'], w['dt'], w['nsteps'])
    e_ref, e_cf = O.rel_err(fid, ref), O.rel_err(fid, cf)
    print(f'2-spin 256x2047: err vs oracle {e_ref:.2e}, vs closed form {e_cf:.2e}', flush=True)
    assert e_ref < 1e-10
    assert e_cf < 5e-8


def test_static_8spin_subset_vs_oracle():
    w = workloads.static_8spin(n_orient=4000)
    sub = grids.subset(w['grid'], np.arange(6))
    fid, per, st = _run(w, grid=sub, per=True)
    print('stats', st, flush=True)
    ref, rper = O.refcpu_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], w['nsteps'], variant=0,
                                per_orientation=True)
    e = O.rel_err(fid, ref)
    print(f'8-spin 6 orientations x 4096 steps: rel err vs faithful oracle {e:.2e}', flush=True)
    assert e < 1e-10
    for k in range(6):
        assert O.rel_err(per[k], rper[k]) < 1e-10


def test_static_8spin_fp32_error_reported():
    w = workloads.static_8spin(n_orient=4000)
    sub = grids.subset(w['grid'], np.arange(6))
    f64, _, _ = _run(w, grid=sub)
    f32, _, _ = _run(w, grid=sub, mode='fp32')
    e = O.rel_err(f32, f64)
    print(f'8-spin FP32 vs FP64 rel err {e:.3e}', flush=True)
    assert np.isfinite(e) and e < 1e-2
