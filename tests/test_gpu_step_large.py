"""GPU parity of the stepping path (k_step) at the sizes it exists for: reduced
dimensions above the dense limit of 1024, which evolution() steps
(kernel/evolution.m:574 via krylov.m:139-163), up to the path's limit of 2304.
At these sizes the sliced-ELL values no longer fit in LDS beside the four state
vectors, so every workgroup reads them from its own slot of a global workspace
(the k_step<REAL, false> variant). Random 11- and 12-spin systems (IK-0 level 4)
against the oracle's stepping restatement (refcpu step_liouv), real and complex
generator arithmetic; the norm bound |coil'*rho(t)| <= |coil| |rho0| over a
longer run; rho and the coil copy in the workgroup's global slot (the default at
these sizes: two workgroups per CU) and in LDS; a dimension above 2304 is refused before anything is launched.
"""
import time

import numpy as np
import pytest

import pyoracle as O
from spinach_amd import _native, grids, workloads

pytestmark = pytest.mark.gpu

KRY = O.make_tols(krylov_tol=2)
DT = 5e-6


def _system(nspins, seed):
    return workloads.random_system(np.random.default_rng([7, nspins, seed]), nspins, level=4)


@pytest.mark.parametrize('nspins,seed,dim', [(11, 1, 1419), (11, 2, 1690), (12, 2, 1806)])
@pytest.mark.parametrize('mode', ['fp64', 'fp64+krylov+complex'])
@pytest.mark.parametrize('vectors', ['global', 'lds'])
def test_stepping_above_the_dense_limit_vs_oracle(nspins, seed, dim, mode, vectors, monkeypatch):
    if vectors == 'lds':
        monkeypatch.setenv('SPINACH_STEP_VEC_LDS', '1')
    w = _system(nspins, seed)
    assert len(O.union_states(w['I'], w['Q'], w['coil'])) == dim
    g = grids.haar(3, np.random.default_rng(5))
    print(f'[{nspins} spins, seed {seed}, {mode}, vectors in {vectors}] {dim} coil-connected states, 3 orientations x 64 steps',
          flush=True)
    plan = _native.plan_powder(w['I'], w['Q'], g, w['rho0'], w['coil'], DT, 64, mode=mode)
    st = plan.execute()
    f, per = plan.result(per_orientation=True)
    plan.close()
    ref, rper = O.refcpu_powder(w['I'], w['Q'], g, w['rho0'], w['coil'], DT, 64, variant=2, tols=KRY,
                                per_orientation=True)
    e = O.rel_err(f, ref)
    print(f'  vs stepping oracle {e:.2e}; {st["step_products"]} products', flush=True)
    assert e < 1e-11
    for k in range(3):
        assert O.rel_err(per[k], rper[k]) < 1e-11
    assert st['krylov_systems'] == 1 and st['reduced_dim'] == dim, st


def test_stepping_norm_bound_and_rate_at_1806_states():
    w = _system(12, 2)
    n_or = 1024
    g = grids.haar(n_or, np.random.default_rng(6))
    nsteps = 128
    plan = _native.plan_powder(w['I'], w['Q'], g, w['rho0'], w['coil'], DT, nsteps)
    plan.execute()  # warm-up
    t0 = time.perf_counter()
    st = plan.execute()
    sec = time.perf_counter() - t0
    f, per = plan.result(per_orientation=True)
    plan.close()
    c0 = np.vdot(w['coil'], w['rho0'])
    assert abs(f[0] - c0 * g['weights'].sum()) <= 1e-12 * max(1.0, abs(c0))
    bound = np.linalg.norm(w['coil']) * np.linalg.norm(w['rho0'])
    assert np.all(np.abs(per) <= bound * (1 + 1e-10))
    assert np.all(np.isfinite(per))
    print(f'1806 states, {n_or} orientations x {nsteps} steps: {sec * 1e3:.1f} ms, '
          f'{n_or * nsteps / sec:.3e} orientation-steps/s, {st["step_products"] / sec:.3e} products/s, '
          f'stage ms {st["stage_ms"]}', flush=True)


def test_dimension_above_the_stepping_limit_is_refused():
    w = _system(12, 0)
    g = grids.haar(2, np.random.default_rng(5))
    with pytest.raises(_native.SpinachPowderError, match='dimension 2560.*stepping path up to 2304'):
        _native.plan_powder(w['I'], w['Q'], g, w['rho0'], w['coil'], DT, 8)
