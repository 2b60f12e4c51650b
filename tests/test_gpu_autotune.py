"""sp_plan_autotune: timed coordinate descent over the launch configuration
(time block, workgroups per CU of each path), the per-process cache by plan
shape, and sp_plan_set_tuning. Workgroup counts leave every orientation's
arithmetic unchanged, so spectral and stepping results must be identical bit
for bit; the time block changes only the propagator path's rounding."""
import numpy as np
import pytest

import pyoracle as O
from spinach_amd import _native, grids, workloads

pytestmark = pytest.mark.gpu


def test_spectral_plan_tuning_is_exact_and_cached():
    w = workloads.static_8spin(n_orient=4000)
    sub = grids.subset(w['grid'], np.arange(1024))
    plan = _native.plan_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], w['nsteps'])
    plan.execute()
    f0 = plan.result()
    t = plan.autotune(passes=1)
    print('spectral tuning', t, flush=True)
    plan.execute()
    assert np.array_equal(plan.result(), f0)
    assert t['candidates'] > 1 and t['best_ms'] <= t['default_ms'] * 1.0001
    plan.close()
    plan2 = _native.plan_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], w['nsteps'])
    got = plan2.tuning()
    assert (got['tri_wg_per_cu'], got['syn_wg_per_cu']) == (t['tri_wg_per_cu'], t['syn_wg_per_cu'])
    again = plan2.autotune(passes=1)
    assert again['candidates'] == 0
    plan2.execute()
    assert np.array_equal(plan2.result(), f0)
    plan2.close()


def test_propagator_plan_tuning_keeps_parity():
    ys = []
    for s in (363, 0, 86):
        y = workloads.batch_member(2, s)
        y['grid'] = grids.subset(y['grid'], np.arange(8))
        ys.append(y)
    plan = _native.plan_powder_batch(ys, 5e-6, 1024)
    plan.execute()
    f0 = plan.result_batch()
    t = plan.autotune(passes=1)
    print('propagator tuning', t, flush=True)
    plan.execute()
    f1 = plan.result_batch()
    for y, a, b in zip(ys, f0, f1):
        assert O.rel_err(b, a) < 1e-12
        ref = O.refcpu_powder(y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], 5e-6, 1024, variant=2)
        assert O.rel_err(b, ref) < 1e-10
    plan.close()


def test_stepping_plan_tuning_and_explicit_settings():
    w = workloads.static_8spin(n_orient=4000)
    sub = grids.subset(w['grid'], np.arange(16))
    plan = _native.plan_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], 128, mode='fp64+krylov+graph')
    plan.execute()
    f0 = plan.result()
    t = plan.autotune(passes=1)
    print('stepping tuning', t, flush=True)
    plan.execute()
    assert np.array_equal(plan.result(), f0)
    plan.set_tuning(dict(t, step_wg_per_cu=99))  # clamped to what the LDS holds
    assert 1 <= plan.tuning()['step_wg_per_cu'] <= 4
    plan.execute()
    assert np.array_equal(plan.result(), f0)
    plan.close()
