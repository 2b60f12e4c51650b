"""Orientation-dependent relaxation on the GPU (sp_plan_powder_relax): the
propagator path with L = H + iR(Omega) against the oracle's restatement
(np_powder_relax), the closed form of the two-spin system, and through the
host mirror powder()."""
import numpy as np
import pytest

import pyoracle as O
from spinach_amd import _native, grids, workloads
from spinach_amd import powder as P
from spinach_amd import spinsys as S

pytestmark = pytest.mark.gpu

T1 = np.array([[1.0, 0.3, -0.2], [0.3, 3.0, 0.5], [-0.2, 0.5, 5.0]]) * 40.0
T2 = np.array([[2.0, -0.4, 0.6], [-0.4, 4.0, 0.1], [0.6, 0.1, 6.0]]) * 60.0
TD = np.array([[1.0, 0.2, 0.0], [0.2, 2.0, -0.3], [0.0, -0.3, 1.5]]) * 25.0
RELAX2 = {'relaxation': ['t1_t2', 'damp'], 'r1_rates': [4.0, T1], 'r2_rates': [T2, 9.0], 'damp_rate': TD,
          'rlx_keep': 'secular'}


def _run(w, grid, mode='fp64', per=False):
    plan = _native.plan_powder(w['I'], w['Q'], grid, w['rho0'], w['coil'], w['dt'], w['nsteps'], mode=mode,
                               relax=S.relaxation_terms(w['ss']))
    try:
        st = plan.execute()
        out = plan.result(per_orientation=per)
    finally:
        plan.close()
    return out, st


@pytest.mark.parametrize('mode', ['fp64', 'fp64+graph'])
def test_two_spin_relaxation_vs_oracle_and_closed_form(mode):
    w = workloads.static_2spin(npoints=1024, n_orient=64, relax=RELAX2)
    g = grids.haar(64, np.random.default_rng(8))
    (fid, per), st = _run(w, g, mode=mode, per=True)
    print(f'[{mode}] stats', {k: st[k] for k in ('spectral_systems', 'real_systems', 'reduced_dim')}, flush=True)
    assert st['spectral_systems'] == 0 and st['real_systems'] == 0
    ref, rper = O.np_powder_relax(w['I'], w['Q'], g, w['rho0'], w['coil'], w['dt'], w['nsteps'], w['ss']['rlx'],
                                  w['ss']['basis'], per_orientation=True)
    cf = O.closed_form_2spin(w['ss'], g, w['dt'], w['nsteps'], relax=w['ss']['rlx'])
    e, ec = O.rel_err(fid, ref), O.rel_err(fid, cf)
    ep = max(O.rel_err(per[k], rper[k]) for k in range(0, 64, 7))
    print(f'[{mode}] 2-spin relaxation 64 x 1023: vs oracle {e:.2e} (per orientation {ep:.2e}), '
          f'vs closed form {ec:.2e}', flush=True)
    assert e < 1e-10 and ep < 1e-10 and ec < 5e-9


def test_eight_spin_relaxation_vs_oracle():
    """185 coil-connected states, t1_t2 rates on every spin (tensors on the
    protons), 4 orientations x 4096 steps; the oracle on the reduced matrices."""
    iso = ['1H'] * 5 + ['13C'] * 2 + ['15N']
    r1 = [T1 * 0.2 if i == '1H' else 0.7 for i in iso]
    r2 = [T2 * 0.3 if i == '1H' else 15.0 for i in iso]
    w = workloads.static_8spin(n_orient=4000, relax={'relaxation': ['t1_t2'], 'r1_rates': r1, 'r2_rates': r2,
                                                     'rlx_keep': 'secular'})
    sub = grids.subset(w['grid'], np.arange(4))
    (fid, per), st = _run(w, sub, per=True)
    assert st['reduced_dim'] == 185
    ref, rper = O.np_powder_relax(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], w['nsteps'], w['ss']['rlx'],
                                  w['ss']['basis'], per_orientation=True)
    f0 = O.np_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], w['nsteps'])
    e = O.rel_err(fid, ref)
    print(f'8-spin relaxation 4 x 4096: vs oracle {e:.2e}; relaxation changes the FID by {O.rel_err(f0, ref):.2f}',
          flush=True)
    assert e < 1e-10
    assert O.rel_err(f0, ref) > 0.05
    for k in range(4):
        assert O.rel_err(per[k], rper[k]) < 1e-10


def test_powder_host_mirror_applies_relaxation():
    w = workloads.static_2spin(npoints=256, n_orient=16, relax=RELAX2)
    params = {'rho0': w['rho0'], 'coil': w['coil'], 'sweep': 1 / w['dt'], 'npoints': 256, 'grid': w['grid']}
    fid, grid = P.powder(w['ss'], P.acquire, params, 'nmr')
    cf = O.closed_form_2spin(w['ss'], grid, w['dt'], 255, relax=w['ss']['rlx'])
    e = O.rel_err(fid, cf)
    print(f'powder() with relaxation: vs closed form {e:.2e}', flush=True)
    assert e < 5e-9


@pytest.mark.parametrize('keep', ['secular', 'kite'])
def test_redfield_powder_vs_oracle(keep):
    """Redfield theory (isotropic tumbling, tau_c = 0.5 ns) on the two-spin system through
    powder(): the host builds the orientation-independent superoperator, the library takes it
    as R0 (sp_plan_powder_relax); against the oracle with the same R0 per orientation. The
    rates themselves are pinned on the CPU (CSA and Solomon dipolar rates)."""
    relax = {'relaxation': ['redfield'], 'tau_c': [5e-10], 'rlx_keep': keep, 'equilibrium': 'zero'}
    w = workloads.static_2spin(npoints=256, n_orient=16, relax=relax)
    params = {'rho0': w['rho0'], 'coil': w['coil'], 'sweep': 1 / w['dt'], 'npoints': 256, 'grid': w['grid'],
              'sum_up': False}
    per, grid = P.powder(w['ss'], P.acquire, params, 'nmr')
    fid = (np.asarray(grid['weights'])[:, None] * np.array(per)).sum(axis=0)
    R0 = S.relaxation_terms(S.assume(w['ss'], 'nmr'))['R0'].toarray()
    ref, rper = O.np_powder_relax(w['I'], w['Q'], grid, w['rho0'], w['coil'], w['dt'], 255, {}, w['ss']['basis'],
                                  reduce=False, per_orientation=True, R0=R0)
    free = O.np_powder(w['I'], w['Q'], grid, w['rho0'], w['coil'], w['dt'], 255)
    e = O.rel_err(fid, ref)
    ep = max(O.rel_err(per[k], rper[k]) for k in range(len(per)))
    print(f'Redfield ({keep}) 2-spin powder: vs oracle {e:.2e} (per orientation {ep:.2e}); relaxation changes '
          f'the FID by {O.rel_err(ref, free):.2f}', flush=True)
    assert e < 1e-10 and ep < 1e-10 and O.rel_err(ref, free) > 1e-3


def test_stepping_path_with_relaxation_vs_oracle():
    """The time-stepping path (k_step, SP_MODE_KRYLOV) with L = H + iR(Omega): the rate tensors
    and damping of the two-spin system on the diagonal, against the oracle."""
    w = workloads.static_2spin(npoints=256, n_orient=16, relax=RELAX2)
    w['nsteps'] = 255
    (fid, per), st = _run(w, w['grid'], mode='fp64+krylov', per=True)
    ref, rper = O.np_powder_relax(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], 255, w['ss']['rlx'],
                                  w['ss']['basis'], per_orientation=True)
    e = O.rel_err(fid, ref)
    ep = max(O.rel_err(per[k], rper[k]) for k in range(16))
    print(f"stepping path with relaxation: {st['krylov_systems']} stepping system(s), vs oracle {e:.2e} "
          f"(per orientation {ep:.2e})", flush=True)
    assert st['krylov_systems'] == 1 and e < 1e-10 and ep < 1e-10


def test_redfield_on_the_stepping_path_vs_oracle():
    relax = {'relaxation': ['redfield'], 'tau_c': [5e-10], 'rlx_keep': 'kite', 'equilibrium': 'zero'}
    w = workloads.static_2spin(npoints=128, n_orient=8, relax=relax)
    w['nsteps'] = 127
    (fid, _), st = _run(w, w['grid'], mode='fp64+krylov', per=True)
    R0 = S.relaxation_terms(S.assume(w['ss'], 'nmr'))['R0'].toarray()
    ref = O.np_powder_relax(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], 127, {}, w['ss']['basis'],
                            reduce=False, R0=R0)
    e = O.rel_err(fid, ref)
    print(f'Redfield on the stepping path: vs oracle {e:.2e}', flush=True)
    assert st['krylov_systems'] == 1 and e < 1e-10


@pytest.mark.parametrize('mode', ['fp64', 'fp64+krylov'])
def test_dead_time_with_relaxation_closed_form(mode):
    """acquire.m:65-67 with relaxation: the dead-time step (k_step) under L = H + iR, then the
    acquisition on the propagator or the stepping path; the two-spin closed form from t0."""
    w = workloads.static_2spin(npoints=128, n_orient=16, relax=RELAX2)
    t0 = 3.3e-5
    plan = _native.plan_powder(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], 127, mode=mode,
                               relax=S.relaxation_terms(w['ss']))
    try:
        plan.set_dead_time(t0)
        plan.execute()
        fid = plan.result()
    finally:
        plan.close()
    cf = O.closed_form_2spin(w['ss'], w['grid'], w['dt'], 127, t0=t0, relax=w['ss']['rlx'])
    e = O.rel_err(fid, cf)
    print(f'dead time {t0:g} s with relaxation ({mode}): vs closed form {e:.2e}', flush=True)
    assert e < 5e-9
