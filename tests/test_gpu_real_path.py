"""GPU parity of the paths a real generator can take: the spectral path
(default in FP64), the real-product propagator path (SP_MODE_PROPAGATOR) and
the general complex path (SP_MODE_COMPLEX), against the oracle and against
each other on the same inputs."""
import numpy as np
import pytest
import scipy.sparse as sps

import pyoracle as O
from spinach_amd import _native, grids, workloads

pytestmark = pytest.mark.gpu


def _run(w, grid, mode):
    return _native.powder_fid(w['I'], w['Q'], grid, w['rho0'], w['coil'], w['dt'], w['nsteps'], mode=mode,
                              per_orientation=True)


def test_8spin_three_paths_agree_with_oracle():
    w = workloads.static_8spin(n_orient=4000)
    sub = grids.subset(w['grid'], np.arange(4))
    out = {}
    for mode in ('fp64', 'fp64+propagator', 'fp64+complex'):
        f, per, st = _run(w, sub, mode)
        print(mode, 'stats', st, flush=True)
        out[mode] = (f, per, st)
    assert out['fp64'][2]['spectral_systems'] == 1 and out['fp64'][2]['real_systems'] == 1
    assert out['fp64'][2]['eig_max_sweeps'] > 0
    assert out['fp64+propagator'][2]['spectral_systems'] == 0 and out['fp64+propagator'][2]['real_systems'] == 1
    assert out['fp64+complex'][2]['spectral_systems'] == 0 and out['fp64+complex'][2]['real_systems'] == 0
    ref = O.refcpu_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], w['nsteps'], variant=0)
    for mode, (f, per, _) in out.items():
        e = O.rel_err(f, ref)
        print(f'8-spin 4 x 4096 {mode}: vs oracle {e:.2e}', flush=True)
        assert e < 1e-10
    fs, ps, _ = out['fp64']
    for mode in ('fp64+propagator', 'fp64+complex'):
        f, per, _ = out[mode]
        assert O.rel_err(fs, f) < 1e-11
        for k in range(4):
            assert O.rel_err(ps[k], per[k]) < 1e-11


def test_2spin_spectral_and_real_propagator_vs_closed_form():
    w = workloads.static_2spin()
    fr, _, sr = _run(w, w['grid'], 'fp64')
    fp, _, sp = _run(w, w['grid'], 'fp64+propagator')
    assert sr['spectral_systems'] == 1 and sp['spectral_systems'] == 0 and sp['real_systems'] == 1
    assert O.rel_err(fr, fp) < 1e-11
    cf = O.closed_form_2spin(w['ss'], w['grid'], w['dt'], w['nsteps'])
    ref = O.refcpu_powder(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'])
    print(f'2-spin real path: vs oracle {O.rel_err(fr, ref):.2e}, vs closed form {O.rel_err(fr, cf):.2e}',
          flush=True)
    assert O.rel_err(fr, ref) < 1e-10
    assert O.rel_err(fr, cf) < 5e-8


def _complex_hermitian(w, scale=1e3, seed=3):
    d = w['I'].shape[0]
    rng = np.random.default_rng(seed)
    K = np.triu(rng.standard_normal((d, d)) * (rng.uniform(size=(d, d)) < 0.05), 1)
    K = K - K.T
    return dict(w, I=sps.csc_matrix(w['I'].toarray() + 1j * scale * K))


def test_complex_generator_takes_the_complex_path():
    w = _complex_hermitian(workloads.static_2spin(npoints=256, n_orient=8))
    assert not _native.generator_is_real(w['I'], w['Q'])
    f, _, st = _run(w, w['grid'], 'fp64')
    assert st['real_systems'] == 0 and st['spectral_systems'] == 0
    ref = O.refcpu_powder(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'])
    e = O.rel_err(f, ref)
    print(f'complex generator 2-spin: vs oracle {e:.2e}', flush=True)
    assert e < 1e-10


def test_batch_mixes_real_and_complex_systems():
    b = workloads.batch_uneven(seed=7, n_systems=4, nsteps=200, min_or=2, max_or=4, min_spins=2, max_spins=5,
                               workers=1)
    systems = list(b['systems'])
    systems[1] = _complex_hermitian(systems[1], seed=4)
    plan = _native.plan_powder_batch(systems, b['dt'], b['nsteps'])
    st = plan.execute()
    fids = plan.result_batch()
    assert st['real_systems'] == 3 and st['spectral_systems'] == 3
    for s, (y, f) in enumerate(zip(systems, fids)):
        # the synthetic complex member is compared with the exact-reduction
        # oracle: on it the reference's zero-track elimination (variant 0)
        # drops states at the 5e-8 level, which dense numpy confirms
        ref = O.refcpu_powder(y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], b['dt'], b['nsteps'],
                              variant=2 if s == 1 else 0)
        e = O.rel_err(f, ref)
        print(f'mixed batch system {s}: vs oracle {e:.2e}', flush=True)
        assert e < 1e-10


def test_fp32_real_path_error_reported():
    w = workloads.static_8spin(n_orient=4000)
    sub = grids.subset(w['grid'], np.arange(4))
    f64, _, _ = _run(w, sub, 'fp64')
    r32, _, s32 = _run(w, sub, 'fp32')
    c32, _, _ = _run(w, sub, 'fp32+complex')
    e_r, e_c = O.rel_err(r32, f64), O.rel_err(c32, f64)
    print(f'8-spin FP32 vs FP64: real path {e_r:.3e}, complex path {e_c:.3e}', flush=True)
    assert s32['real_systems'] == 1
    assert np.isfinite(e_r) and e_r < 1e-2


def test_dense_real_generator_spectral_path():
    # a dense real symmetric block on the 8-spin system's 185 coil-connected states: the
    # union pattern (~34K entries) is far beyond the entries whose values the
    # tridiagonalisation computes up front, so most are read during its assembly passes
    w = workloads.static_8spin(n_orient=4000)
    sub = grids.subset(w['grid'], np.arange(3))
    idx = O.union_states(w['I'], w['Q'], w['coil'])
    rng = np.random.default_rng(11)
    m = len(idx)
    K = rng.standard_normal((m, m))
    K = (K + K.T) * (1e-3 * abs(w['I']).max())
    I = w['I'].toarray()
    I[np.ix_(idx, idx)] += K
    y = dict(w, I=sps.csc_matrix(I))
    assert _native.generator_is_real(y['I'], y['Q'])
    f, _, st = _run(y, sub, 'fp64')
    print('dense generator stats', st, flush=True)
    assert st['spectral_systems'] == 1 and st['reduced_dim'] == m and st['reduced_nnz'] > 8 * 2048
    ref = O.refcpu_powder(y['I'], y['Q'], sub, y['rho0'], y['coil'], y['dt'], y['nsteps'], variant=2)
    e = O.rel_err(f, ref)
    print(f'dense real generator, 185 states, 3 x 4096: vs oracle {e:.2e}', flush=True)
    assert e < 1e-10


def test_tridiagonal_generator_takes_no_reflection():
    # a generator that is already tridiagonal on the 8-spin system's 185 coil-connected states
    # (kept in index order by the reduction) and orientation-independent (no Q): every
    # Householder step finds nothing below the subdiagonal and publishes the column as it is
    w = workloads.static_8spin(n_orient=4000)
    sub = grids.subset(w['grid'], np.arange(2))
    idx = O.union_states(w['I'], w['Q'], w['coil'])
    rng = np.random.default_rng(12)
    m = len(idx)
    scale = abs(w['I']).max()
    I = sps.lil_matrix(w['I'].shape)
    for k in range(m):
        I[idx[k], idx[k]] = scale * rng.uniform(-1, 1)
        if k + 1 < m:
            e = 0.1 * scale * rng.uniform(0.5, 1.0)
            I[idx[k], idx[k + 1]] = e
            I[idx[k + 1], idx[k]] = e
    y = dict(w, I=sps.csc_matrix(I), Q=[None] * len(w['Q']))
    f, _, st = _run(y, sub, 'fp64')
    print('tridiagonal generator stats', st, flush=True)
    assert st['spectral_systems'] == 1 and st['reduced_dim'] == m
    ref = O.refcpu_powder(y['I'], y['Q'], sub, y['rho0'], y['coil'], y['dt'], y['nsteps'], variant=2)
    e = O.rel_err(f, ref)
    print(f'tridiagonal generator, 185 states, 2 x 4096: vs oracle {e:.2e}', flush=True)
    assert e < 1e-10
