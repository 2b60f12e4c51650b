"""GPU parity of the time-stepping path (k_step): step() per time step with
detection after each (kernel/krylov.m:139-163, kernel/step.m:204-375), taken
with SP_MODE_KRYLOV (sys.enable={'krylov'}, kernel/evolution.m:531-533) or at
krylov_tol, and of the dead-time step (experiments/acquire.m:65-67) on every
path. The oracle steps the same way (refcpu step_liouv, pinned against expm in
test_oracle_pins.py); both agree with the exact exponential to ~1e-13.
"""
import numpy as np
import pytest

import pyoracle as O
from spinach_amd import _native, grids, workloads

pytestmark = pytest.mark.gpu

KRY = O.make_tols(krylov_tol=2)  # the oracle steps every subspace of two or more states


def _fid(w, grid, mode, nsteps=None, tols=None, dead_time=0.0, per=False):
    ns = w['nsteps'] if nsteps is None else nsteps
    plan = _native.plan_powder(w['I'], w['Q'], grid, w['rho0'], w['coil'], w['dt'], ns, tols=tols, mode=mode)
    if dead_time:
        plan.set_dead_time(dead_time)
    st = plan.execute()
    out = plan.result(per_orientation=per)
    plan.close()
    return out, st


def test_krylov_two_spin_vs_oracle_and_closed_form():
    w = workloads.static_2spin(npoints=256, n_orient=16)
    f, st = _fid(w, w['grid'], 'fp64+krylov')
    print('stats', st, flush=True)
    assert st['krylov_systems'] == 1 and st['spectral_systems'] == 0
    ref = O.refcpu_powder(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'], variant=1, tols=KRY)
    cf = O.closed_form_2spin(w['ss'], w['grid'], w['dt'], w['nsteps'])
    e, ec = O.rel_err(f, ref), O.rel_err(f, cf)
    print(f'2-spin stepping: vs stepping oracle {e:.2e}, vs closed form {ec:.2e}', flush=True)
    assert e < 1e-12
    assert ec < 5e-9


@pytest.mark.parametrize('mode', ['fp64+krylov', 'fp64+krylov+complex'])
def test_krylov_eight_spin_vs_oracle(mode):
    """185 coil-connected states, real (k_step<true>) and complex (k_step<false>)
    generator arithmetic; per-orientation FIDs (sum_up = false) included."""
    w = workloads.static_8spin(n_orient=4000)
    sub = grids.subset(w['grid'], np.arange(4))
    (f, per), st = _fid(w, sub, mode, nsteps=512, per=True)
    print(f'[{mode}] stats', st, flush=True)
    assert st['krylov_systems'] == 1
    ref, rper = O.refcpu_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], 512, variant=2, tols=KRY,
                                per_orientation=True)
    spec, _ = _fid(w, sub, 'fp64', nsteps=512)
    e, es = O.rel_err(f, ref), O.rel_err(f, spec)
    print(f'[{mode}] 8-spin 4 x 512 stepping: vs stepping oracle {e:.2e}, vs GPU spectral path {es:.2e}', flush=True)
    assert e < 1e-11
    assert es < 1e-10
    for k in range(4):
        assert O.rel_err(per[k], rper[k]) < 1e-11


def test_krylov_tol_selects_the_stepping_path():
    """Without the flag, systems with a coil-connected dimension >= krylov_tol
    are stepped (evolution.m:531-533); the batch mixes both paths."""
    ys = []
    for s in (363, 0, 5, 17):
        y = workloads.batch_member(2, s)
        y['grid'] = grids.subset(y['grid'], np.arange(2))
        ys.append(y)
    dims = [len(O.union_states(y['I'], y['Q'], y['coil'])) for y in ys]
    tol = 200
    plan = _native.plan_powder_batch(ys, 5e-6, 256, tols={'krylov_tol': tol})
    st = plan.execute()
    fids = plan.result_batch()
    print('stats', st, 'dims', dims, flush=True)
    assert st['krylov_systems'] == sum(d >= tol for d in dims) >= 1
    for y, f, d in zip(ys, fids, dims):
        ref = O.refcpu_powder(y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], 5e-6, 256, variant=2)
        e = O.rel_err(f, ref)
        print(f'reduced dim {d} ({"stepped" if d >= tol else "dense"}): vs exact oracle {e:.2e}', flush=True)
        assert e < 1e-10


@pytest.mark.parametrize('mode', ['fp64', 'fp64+propagator', 'fp64+krylov', 'fp64+propagator+complex'])
def test_dead_time_all_paths(mode):
    """parameters.dead_time on the spectral, propagator (real and complex) and
    stepping paths against the oracle, and against the shifted closed form."""
    w = workloads.static_2spin(npoints=128, n_orient=8)
    td = 23.4e-6
    f, st = _fid(w, w['grid'], mode, dead_time=td)
    ref = O.refcpu_powder(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'], dead_time=td)
    cf = O.closed_form_2spin(w['ss'], w['grid'], w['dt'], w['nsteps'], t0=td)
    e, ec = O.rel_err(f, ref), O.rel_err(f, cf)
    print(f'[{mode}] dead time {td:.3e}: vs oracle {e:.2e}, vs shifted closed form {ec:.2e}', flush=True)
    assert e < 1e-10
    assert ec < 5e-9


def test_dead_time_eight_spin_and_removal():
    w = workloads.static_8spin(n_orient=4000)
    sub = grids.subset(w['grid'], np.arange(3))
    td = 17e-6
    plan = _native.plan_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], 256)
    plan.execute()
    f0 = plan.result()
    plan.set_dead_time(td)
    plan.execute()
    f1 = plan.result()
    plan.set_dead_time(0.0)
    plan.execute()
    f2 = plan.result()
    plan.close()
    ref = O.refcpu_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], 256, variant=2, dead_time=td)
    e = O.rel_err(f1, ref)
    print(f'8-spin dead time: vs oracle {e:.2e}', flush=True)
    assert e < 1e-10
    assert np.array_equal(f0, f2)
