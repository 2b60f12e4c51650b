"""The MEX gateway (spinach_amd/mex/powder_hip.cpp) on the GPU: compiled
against the mex.h stand-in and linked to the real libspinach_powder.so, run as
powder.m would call it, against the ctypes binding and the oracle."""
import os

import numpy as np
import pytest

import boundary_io as B
import pyoracle as O
from spinach_amd import _native, grids, workloads

pytestmark = pytest.mark.gpu

EXE = os.path.join(B.BUILD, 'mex_driver_gpu')


def _call(rhs, nlhs, tmp):
    assert os.path.exists(EXE), 'tests/native/_build/mex_driver_gpu is built by __graft_entry__.build()'
    r, outs = B.run_driver(EXE, rhs, nlhs, tmp)
    assert r.returncode == 0 and r.stdout.startswith('OK'), (r.stdout, r.stderr[-2000:])
    return outs


def test_gateway_eight_spin_matches_ctypes_and_oracle(tmp_path):
    w = workloads.static_8spin(n_orient=4000)
    sub = grids.subset(w['grid'], np.arange(8))
    rhs = B.powder_hip_rhs(w['I'], w['Qnested'], sub, w['rho0'], w['coil'], w['dt'], w['nsteps'],
                           tols={'krylov_tol': 10000}, opts={})
    fid, per = _call(rhs, 2, str(tmp_path))
    ref, _, _ = _native.powder_fid(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], w['nsteps'])
    orc = O.refcpu_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], w['nsteps'], variant=0)
    print(f'MEX gateway 8-spin 8 x 4096: vs ctypes max |diff| {np.max(np.abs(fid[:, 0] - ref)):.2e}, '
          f'vs oracle {O.rel_err(fid[:, 0], orc):.2e}', flush=True)
    assert np.array_equal(fid[:, 0], ref)
    assert O.rel_err(fid[:, 0], orc) < 1e-10
    assert O.rel_err(per @ sub['weights'], fid[:, 0]) < 1e-14


def test_gateway_dead_time_krylov_sum_up_false(tmp_path):
    w = workloads.static_2spin(npoints=128, n_orient=6)
    opts = {'dead_time': B.mx_dense(np.array([[9.1e-6]])), 'krylov': B.mx_logical(True),
            'sum_up': B.mx_logical(False)}
    rhs = B.powder_hip_rhs(w['I'], w['Qnested'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'], opts=opts)
    per, = _call(rhs, 1, str(tmp_path))
    assert per.shape == (128, 6)
    fid = per @ w['grid']['weights']
    cf = O.closed_form_2spin(w['ss'], w['grid'], w['dt'], w['nsteps'], t0=9.1e-6)
    orc = O.refcpu_powder(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'], dead_time=9.1e-6)
    print(f'MEX gateway 2-spin stepping + dead time: vs oracle {O.rel_err(fid, orc):.2e}, '
          f'vs shifted closed form {O.rel_err(fid, cf):.2e}', flush=True)
    assert O.rel_err(fid, orc) < 1e-10
    assert O.rel_err(fid, cf) < 5e-9
