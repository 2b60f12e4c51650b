"""cpmg() and sp_acquire() through the host mirror powder() on the GPU
(sp_plan_set_initial_states, sp_plan_propagate) against the oracle's
restatements and the closed forms."""
import math

import numpy as np
import pytest

import pyoracle as O
from spinach_amd import grids, workloads
from spinach_amd import powder as P
from spinach_amd import spinsys as S
from test_sequences import cpmg_closed_form, one_spin_csa

pytestmark = pytest.mark.gpu


def test_cpmg_one_spin_closed_form():
    ss, I, q = one_spin_csa()
    g = grids.haar(10, np.random.default_rng(4))
    params = {'rho0': S.state(ss, 'Lz', '1H'), 'coil': S.state(ss, 'L+', '1H'), 'pulse_op': S.operator(ss, 'Ly', '1H'),
              'timestep': 2e-5, 'npoints': 40, 'nloops': 3, 'grid': g}
    fid, _ = P.powder(ss, P.cpmg, params, 'nmr')
    f0 = np.vdot(params['coil'], S.state(ss, 'Lx', '1H'))
    cf = cpmg_closed_form(ss, g, 2e-5, 40, 3, f0)
    e = O.rel_err(fid, cf)
    print(f'CPMG one spin on the GPU: vs closed form {e:.2e}', flush=True)
    assert e < 1e-9


def test_cpmg_two_spin_powder_vs_oracle():
    w = workloads.static_2spin(npoints=64, n_orient=16)
    ss = w['ss']
    g = grids.haar(16, np.random.default_rng(6))
    pop = S.operator(ss, 'Ly', '1H')
    rho0 = S.state(ss, 'Lz', '1H')
    params = {'rho0': rho0, 'coil': w['coil'], 'pulse_op': pop, 'timestep': w['dt'], 'npoints': 64, 'nloops': 3,
              'grid': g, 'sum_up': False}
    per, _ = P.powder(ss, P.cpmg, params, 'nmr')
    fid = (np.asarray(g['weights'])[:, None] * np.array(per)).sum(axis=0)
    ref, rper = O.np_cpmg(w['I'], w['Q'], g, rho0, w['coil'], pop, w['dt'], 64, 3, per_orientation=True)
    e = O.rel_err(fid, ref)
    ep = max(O.rel_err(per[k], rper[k]) for k in range(16))
    print(f'CPMG 2-spin, 16 orientations, 3 loops: vs oracle {e:.2e} (per orientation {ep:.2e})', flush=True)
    assert len(fid) == 64 + 3 * 128 and e < 1e-10 and ep < 1e-10


def test_sp_acquire_two_spin_vs_oracle():
    w = workloads.static_2spin(npoints=256, n_orient=8)
    ss = w['ss']
    g = grids.haar(8, np.random.default_rng(7))
    ep = S.operator(ss, 'L+', '1H')
    ex, ey = (ep + ep.conj().T) / 2, (ep - ep.conj().T) / 2j
    rho0 = S.state(ss, 'Lz', '1H')
    pulse = dict(spins=['1H'], pulse_frq=3.0e3, pulse_pwr=2 * math.pi * 2.0e4, pulse_dur=1.25e-5, pulse_phi=0.4,
                 pulse_rnk=2)
    params = dict(pulse, rho0=rho0, coil=w['coil'], sweep=1 / w['dt'], npoints=256, grid=g)
    fid, _ = P.powder(ss, P.sp_acquire, params, 'nmr')
    ref = O.np_sp_acquire(w['I'], w['Q'], g, rho0, w['coil'], ex, ey, 3.0e3, pulse['pulse_pwr'], 1.25e-5, 0.4, 2,
                          w['dt'], 255)
    e = O.rel_err(fid, ref)
    print(f'sp_acquire 2-spin (Fokker-Planck rank 2): vs oracle {e:.2e}', flush=True)
    assert e < 1e-10


def test_sp_acquire_reference_known_answer():
    sys = {'magnet': 14.1, 'isotopes': ['1H']}
    ss = S.basis(S.create(sys, {'zeeman': {'scalar': [0.0]}}), {'formalism': 'sphten-liouv', 'approximation': 'none'})
    ep = S.operator(ss, 'L+', '1H')
    ex, ey = (ep + ep.conj().T) / 2, (ep - ep.conj().T) / 2j
    rho0, coil = S.state(ss, 'Lz', '1H'), S.state(ss, 'Lx', '1H')
    amp, phi, dur = 2 * math.pi * 9e3, math.pi / 7, 1.8e-5
    params = dict(spins=['1H'], pulse_frq=0.0, pulse_pwr=amp, pulse_dur=dur, pulse_phi=phi, pulse_rnk=1, rho0=rho0,
                  coil=coil, sweep=1e5, npoints=4, grid='single_crystal')
    fid, _ = P.powder(ss, P.sp_acquire, params, 'nmr')
    import scipy.linalg as sla
    ref = np.vdot(coil, sla.expm(-1j * dur * amp * (math.cos(phi) * ex + math.sin(phi) * ey).toarray()) @ rho0)
    assert np.allclose(fid, ref, rtol=1e-10, atol=1e-13)


def test_hp_acquire_echo_refocuses_one_spin():
    """hp_acquire with an echo on the GPU: pi/2 and pi pulses about y around two echo_time
    evolutions refocus a one-spin chemical shift at every orientation, so the FID is that of
    -Lx (oracle without the echo)."""
    from test_sequences import one_spin_csa
    ss, I, q = one_spin_csa()
    g = grids.haar(10, np.random.default_rng(5))
    ly = S.operator(ss, 'Ly', '1H')
    params = {'rho0': S.state(ss, 'Lz', '1H'), 'coil': S.state(ss, 'L+', '1H'), 'sweep': 5e4, 'npoints': 41,
              'grid': g, 'pulse_op': ly, 'pulse_angle': math.pi / 2, 'echo_time': 3.7e-5, 'echo_oper': ly,
              'echo_angle': math.pi}
    fid, _ = P.powder(ss, P.hp_acquire, params, 'nmr')
    ref = O.np_powder(I, q, g, -S.state(ss, 'Lx', '1H'), params['coil'], 2e-5, 40, reduce=False)
    e = O.rel_err(fid, ref)
    print(f'hp_acquire echo, one spin: vs refocused oracle {e:.2e}', flush=True)
    assert e < 1e-10


def test_hp_acquire_echo_two_spin_vs_oracle():
    w = workloads.static_2spin(npoints=128, n_orient=12)
    ss = w['ss']
    g = grids.haar(12, np.random.default_rng(8))
    ly, lx = S.operator(ss, 'Ly', '1H'), S.operator(ss, 'Lx', '1H')
    rho0 = S.state(ss, 'Lz', '1H')
    params = {'rho0': rho0, 'coil': w['coil'], 'sweep': 1 / w['dt'], 'npoints': 128, 'grid': g, 'pulse_op': ly,
              'pulse_angle': math.pi / 2, 'echo_time': 2.9e-5, 'echo_oper': lx, 'echo_angle': 2 * math.pi / 3,
              'sum_up': False}
    per, _ = P.powder(ss, P.hp_acquire, params, 'nmr')
    fid = (np.asarray(g['weights'])[:, None] * np.array(per)).sum(axis=0)
    ref, rper = O.np_hp_echo(w['I'], w['Q'], g, rho0, w['coil'], ly, math.pi / 2, lx, 2 * math.pi / 3, 2.9e-5,
                             w['dt'], 127, per_orientation=True)
    e = O.rel_err(fid, ref)
    ep = max(O.rel_err(per[k], rper[k]) for k in range(12))
    print(f'hp_acquire echo, 2-spin: vs oracle {e:.2e} (per orientation {ep:.2e})', flush=True)
    assert e < 1e-10 and ep < 1e-10



def test_cpmg_and_echo_with_relaxation_vs_oracle():
    """cpmg (rate tensors and damping on the diagonal) and hp_acquire's echo (Redfield R0) with
    L = H + iR between the pulses and during the acquisitions, against the oracle."""
    T1 = np.array([[1.0, 0.3, -0.2], [0.3, 3.0, 0.5], [-0.2, 0.5, 5.0]]) * 40.0
    relax = {'relaxation': ['t1_t2', 'damp'], 'r1_rates': [4.0, T1], 'r2_rates': [T1 * 2, 9.0],
             'damp_rate': 25.0, 'rlx_keep': 'secular'}
    w = workloads.static_2spin(npoints=32, n_orient=8, relax=relax)
    ss = w['ss']
    g = grids.haar(8, np.random.default_rng(9))
    pop, rho0 = S.operator(ss, 'Ly', '1H'), S.state(ss, 'Lz', '1H')
    params = {'rho0': rho0, 'coil': w['coil'], 'pulse_op': pop, 'timestep': w['dt'], 'npoints': 32, 'nloops': 2,
              'grid': g}
    fid, _ = P.powder(ss, P.cpmg, params, 'nmr')
    ref = O.np_cpmg(w['I'], w['Q'], g, rho0, w['coil'], pop, w['dt'], 32, 2, rlx=ss['rlx'], states=ss['basis'])
    free = O.np_cpmg(w['I'], w['Q'], g, rho0, w['coil'], pop, w['dt'], 32, 2)
    e = O.rel_err(fid, ref)
    print(f'CPMG with relaxation: vs oracle {e:.2e} (relaxation changes it by {O.rel_err(ref, free):.2f})', flush=True)
    assert e < 1e-10 and O.rel_err(ref, free) > 1e-3
    rf = {'relaxation': ['redfield'], 'tau_c': [5e-10], 'rlx_keep': 'kite', 'equilibrium': 'zero'}
    w = workloads.static_2spin(npoints=64, n_orient=8, relax=rf)
    ss = w['ss']
    lx = S.operator(ss, 'Lx', '1H')
    params = {'rho0': rho0, 'coil': w['coil'], 'sweep': 1 / w['dt'], 'npoints': 64, 'grid': g, 'pulse_op': pop,
              'pulse_angle': math.pi / 2, 'echo_time': 4.1e-5, 'echo_oper': lx, 'echo_angle': math.pi}
    fid, _ = P.powder(ss, P.hp_acquire, params, 'nmr')
    R0 = S.relaxation_terms(S.assume(ss, 'nmr'))['R0'].toarray()
    ref = O.np_hp_echo(w['I'], w['Q'], g, rho0, w['coil'], pop, math.pi / 2, lx, math.pi, 4.1e-5, w['dt'], 63,
                       R0=R0)
    e = O.rel_err(fid, ref)
    print(f'hp_acquire echo with Redfield relaxation: vs oracle {e:.2e}', flush=True)
    assert e < 1e-10



def test_sp_acquire_with_relaxation_vs_oracle():
    """sp_acquire under L = H + iR(Omega) (rate tensors and damping): the relaxation planes
    replicated over the Fokker-Planck phase blocks during the pulse, and the acquisition with R."""
    T1 = np.array([[1.0, 0.3, -0.2], [0.3, 3.0, 0.5], [-0.2, 0.5, 5.0]]) * 400.0
    relax = {'relaxation': ['t1_t2', 'damp'], 'r1_rates': [400.0, T1], 'r2_rates': [T1 * 2, 900.0],
             'damp_rate': 2500.0, 'rlx_keep': 'secular'}
    w = workloads.static_2spin(npoints=128, n_orient=6, relax=relax)
    ss = w['ss']
    g = grids.haar(6, np.random.default_rng(12))
    ep = S.operator(ss, 'L+', '1H')
    ex, ey = (ep + ep.conj().T) / 2, (ep - ep.conj().T) / 2j
    rho0 = S.state(ss, 'Lz', '1H')
    pulse = dict(spins=['1H'], pulse_frq=3.0e3, pulse_pwr=2 * math.pi * 2.0e4, pulse_dur=1.25e-5, pulse_phi=0.4,
                 pulse_rnk=2)
    params = dict(pulse, rho0=rho0, coil=w['coil'], sweep=1 / w['dt'], npoints=128, grid=g)
    fid, _ = P.powder(ss, P.sp_acquire, params, 'nmr')
    ref = O.np_sp_acquire(w['I'], w['Q'], g, rho0, w['coil'], ex, ey, 3.0e3, pulse['pulse_pwr'], 1.25e-5, 0.4, 2,
                          w['dt'], 127, rlx=ss['rlx'], states=ss['basis'])
    free = O.np_sp_acquire(w['I'], w['Q'], g, rho0, w['coil'], ex, ey, 3.0e3, pulse['pulse_pwr'], 1.25e-5, 0.4, 2,
                           w['dt'], 127)
    e = O.rel_err(fid, ref)
    print(f'sp_acquire with relaxation: vs oracle {e:.2e} (relaxation changes it by {O.rel_err(ref, free):.2f})',
          flush=True)
    assert e < 1e-10 and O.rel_err(ref, free) > 1e-3
