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


def test_gateway_relaxation_matches_ctypes_and_closed_form(tmp_path):
    """opts.rlx_diag and opts.R0 through the gateway, with the scalar rates
    moved from g0 into R0 = diag(g0): the same FID as the ctypes plan given
    the same sp_relax (bit for bit), the ctypes plan with every rate in the
    planes, and the two-spin closed form with relaxation."""
    import scipy.sparse as sps
    from spinach_amd import spinsys as S
    t1 = np.array([[1.0, 0.3, -0.2], [0.3, 3.0, 0.5], [-0.2, 0.5, 5.0]]) * 40.0
    relax = {'relaxation': ['t1_t2', 'damp'], 'r1_rates': [4.0, t1], 'r2_rates': [t1 * 1.5, 9.0],
             'damp_rate': t1 * 0.5, 'rlx_keep': 'secular'}
    w = workloads.static_2spin(npoints=512, n_orient=24, relax=relax)
    g = S.relaxation_terms(w['ss'])['diag']
    r0 = sps.csc_matrix(sps.diags(g[0]))
    r0.eliminate_zeros()
    split = {'R0': r0, 'diag': np.vstack([np.zeros_like(g[:1]), g[1:]])}
    assert np.abs(g[0]).max() > 0
    opts = {'rlx_diag': B.mx_dense(split['diag'].T), 'R0': B.mx_sparse(split['R0'])}
    rhs = B.powder_hip_rhs(w['I'], w['Qnested'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'], opts=opts)
    fid, = _call(rhs, 1, str(tmp_path))
    refs = []
    for rl in (split, {'R0': None, 'diag': g}):
        plan = _native.plan_powder(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'], relax=rl)
        try:
            plan.execute()
            refs.append(plan.result())
        finally:
            plan.close()
    cf = O.closed_form_2spin(w['ss'], w['grid'], w['dt'], w['nsteps'], relax=w['ss']['rlx'])
    print(f'MEX gateway 2-spin relaxation 24 x 511: vs ctypes max |diff| {np.max(np.abs(fid[:, 0] - refs[0])):.2e}, '
          f'vs planes only {O.rel_err(fid[:, 0], refs[1]):.2e}, vs closed form {O.rel_err(fid[:, 0], cf):.2e}',
          flush=True)
    assert np.array_equal(fid[:, 0], refs[0])
    assert O.rel_err(fid[:, 0], refs[1]) < 1e-12
    assert O.rel_err(fid[:, 0], cf) < 5e-9
