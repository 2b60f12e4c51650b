"""experiments/cpmg.m and experiments/sp_acquire.m (kernel/pulses/shaped_pulse_af.m)
in the oracle: the closed form of a CPMG echo train on one spin, the
reference's own known answer for the Fokker-Planck soft pulse
(tests/kernel/test_dynamic_shaped_pulses_deep.m:87-117), and the free-evolution
limit of the soft pulse, each with a planted mistake that must fail; the host
mirror's input checks."""
import math

import numpy as np
import pytest
import scipy.linalg as sla

import pyoracle as O
from spinach_amd import contexts as Cx
from spinach_amd import grids, workloads
from spinach_amd import powder as P
from spinach_amd import spinsys as S


def one_spin_csa(approximation='none'):
    sys = {'magnet': 14.1, 'isotopes': ['1H']}
    inter = {'zeeman': {'eigs': [[-20.0, -5.0, 25.0]], 'euler': [[0.3, 0.9, -0.4]]}}
    ss = S.basis(S.create(sys, inter), {'formalism': 'sphten-liouv', 'approximation': approximation})
    I, Q = S.hamiltonian(S.assume(ss, 'nmr'))
    return ss, I, S.q_list(Q)


def cpmg_closed_form(ss, grid, dt, npoints, nloops, f0):
    """One spin, rho0 = Lz, pulse_op = Ly, coil = L+: the pi/2 pulse leaves
    a = 1/2 on L+; free evolution multiplies a by exp(-i w t); the pi pulse
    about y maps L+ to -L- and L- to -L+, so a -> -conj(a); f = 2 f(0) a."""
    from spinach_amd.spinsys import euler2dcm
    z = ss['zeeman'][0]
    ziso = np.trace(z) / 3 - ss['basefrqs'][0]
    n_or = len(grid['weights'])
    per = []
    for o in range(n_or):
        R = euler2dcm(grid['alphas'][o], grid['betas'][o], grid['gammas'][o])
        w = ziso + (R @ (z - np.eye(3) * np.trace(z) / 3) @ R.T)[2, 2]
        a, row = 0.5, []
        for seg in range(nloops + 1):
            n = npoints if seg == 0 else 2 * npoints
            if seg > 0:
                a = -np.conj(a)
            for t in range(n):
                row.append(a)
                if t < n - 1:
                    a = a * np.exp(-1j * w * dt)
        per.append(2 * f0 * np.array(row))
    per = np.array(per)
    return (np.asarray(grid['weights'])[:, None] * per).sum(axis=0)


def test_oracle_fourier_collocation_equals_the_host_one():
    for N in (1, 2, 3, 5, 8, 11):
        x0, d0 = Cx.fourdif(N)
        x1, d1 = O.fourdif1(N)
        assert np.allclose(x0, x1, atol=0) and np.allclose(d0, d1, rtol=1e-15, atol=1e-15)
        # exact on trigonometric polynomials of order < N/2
        if N >= 3:
            assert np.allclose(d1 @ np.sin(x1), np.cos(x1), atol=1e-12)


def test_oracle_cpmg_echo_train_closed_form():
    ss, I, q = one_spin_csa()
    g = grids.haar(10, np.random.default_rng(4))
    rho0, coil = S.state(ss, 'Lz', '1H'), S.state(ss, 'L+', '1H')
    pop = S.operator(ss, 'Ly', '1H')
    dt, npts, nl = 2e-5, 40, 3
    f = O.np_cpmg(I, q, g, rho0, coil, pop, dt, npts, nl)
    f0 = np.vdot(coil, S.state(ss, 'Lx', '1H'))
    cf = cpmg_closed_form(ss, g, dt, npts, nl, f0)
    e = O.rel_err(f, cf)
    print(f'CPMG one spin, {nl} loops: oracle vs closed form {e:.2e}')
    assert len(f) == npts + nl * 2 * npts and e < 1e-10
    assert O.rel_err(O.np_cpmg(I, q, g, rho0, coil, pop, dt, npts, nl, angle=math.pi / 4), cf) > 0.1


def test_oracle_hp_echo_refocuses_a_one_spin_csa():
    """hp_acquire with an echo (hp_acquire.m:51-58): pi/2 pulse about y on Lz, echo_time,
    a pi pulse about y, echo_time. For one spin the chemical shift is refocused at every
    orientation, so the echo equals the pulses alone: acquire() from exp(-i pi Ly) Lx = -Lx.
    A pi/2 echo pulse (planted) does not refocus."""
    ss, I, q = one_spin_csa()
    g = grids.haar(10, np.random.default_rng(5))
    rho0, coil, ly = S.state(ss, 'Lz', '1H'), S.state(ss, 'L+', '1H'), S.operator(ss, 'Ly', '1H')
    dt, nsteps, tau = 2e-5, 40, 3.7e-5
    f = O.np_hp_echo(I, q, g, rho0, coil, ly, math.pi / 2, ly, math.pi, tau, dt, nsteps)
    ref = O.np_powder(I, q, g, -S.state(ss, 'Lx', '1H'), coil, dt, nsteps, reduce=False)
    assert O.rel_err(f, ref) < 1e-10
    bad = O.np_hp_echo(I, q, g, rho0, coil, ly, math.pi / 2, ly, math.pi / 2, tau, dt, nsteps)
    assert O.rel_err(bad, ref) > 0.1


def test_oracle_hp_echo_without_a_pulse_is_a_time_shift():
    """With echo_angle = 0 the echo is free evolution for 2 echo_time: the two-spin FID of Lx
    from t0 = 2 echo_time (closed form). An evolution of the wrong sign (planted through a
    negative echo time) gives the FID from -2 echo_time instead."""
    w = workloads.static_2spin(npoints=64, n_orient=12)
    ss, g = w['ss'], w['grid']
    rho0, ly = S.state(ss, 'Lz', '1H'), S.operator(ss, 'Ly', '1H')
    tau = 1.3e-5
    f = O.np_hp_echo(w['I'], w['Q'], g, rho0, w['coil'], ly, math.pi / 2, ly, 0.0, tau, w['dt'], 63)
    cf = O.closed_form_2spin(ss, g, w['dt'], 63, t0=2 * tau)
    assert O.rel_err(f, cf) < 1e-9
    bad = O.np_hp_echo(w['I'], w['Q'], g, rho0, w['coil'], ly, math.pi / 2, ly, 0.0, -tau, w['dt'], 63)
    assert O.rel_err(bad, cf) > 1e-2


def test_oracle_soft_pulse_reference_known_answer():
    """Zero drift and zero offset: the Fokker-Planck pulse folds back to
    step(amp*(cos(phi) Lx + sin(phi) Ly), rho, dur) (the reference's test)."""
    sys = {'magnet': 14.1, 'isotopes': ['1H']}
    ss = S.basis(S.create(sys, {'zeeman': {'scalar': [0.0]}}), {'formalism': 'sphten-liouv', 'approximation': 'none'})
    I, Q = S.hamiltonian(S.assume(ss, 'nmr'))
    assert abs(I).max() == 0
    q = S.q_list(Q)
    ep = S.operator(ss, 'L+', '1H')
    ex, ey = (ep + ep.conj().T) / 2, (ep - ep.conj().T) / 2j
    rho0 = S.state(ss, 'Lz', '1H')
    amp, phi, dur = 2 * math.pi * 9e3, math.pi / 7, 1.8e-5
    g = grids.single_crystal()
    f = O.np_sp_acquire(I, q, g, rho0, rho0, ex, ey, 0.0, amp, dur, phi, 1, 1e-5, 3)
    gen = amp * (math.cos(phi) * ex + math.sin(phi) * ey).toarray()
    ref_rho = sla.expm(-1j * dur * gen) @ rho0
    assert np.allclose(f, np.vdot(rho0, ref_rho), rtol=1e-12, atol=1e-14)
    for coil in (S.state(ss, 'L+', '1H'), S.state(ss, 'Lx', '1H')):
        assert abs(O.np_sp_acquire(I, q, g, rho0, coil, ex, ey, 0.0, amp, dur, phi, 1, 1e-5, 0)[0]
                   - np.vdot(coil, ref_rho)) < 1e-12


def test_oracle_soft_pulse_without_power_is_free_evolution():
    """pulse_pwr = 0 with an offset: the phase blocks mix, but the columns of
    exp(c*d/dphi) sum to one, so the folded state is exp(-i H dur) rho0 and the
    FID is the two-spin closed form shifted by dur; the first-block plant fails."""
    w = workloads.static_2spin(npoints=128, n_orient=8)
    ss = w['ss']
    ep = S.operator(ss, 'L+', '1H')
    ex, ey = (ep + ep.conj().T) / 2, (ep - ep.conj().T) / 2j
    g, dur = w['grid'], 3e-5
    args = (w['I'], w['Q'], g, w['rho0'], w['coil'], ex, ey, 4.0e3, 0.0, dur, 0.3, 2, w['dt'], w['nsteps'])
    f = O.np_sp_acquire(*args)
    cf = O.closed_form_2spin(ss, g, w['dt'], w['nsteps'], t0=dur)
    assert O.rel_err(f, cf) < 5e-9
    assert O.rel_err(O.np_sp_acquire(*args, plant=O.PLANT_FOLD), cf) > 1e-2


@pytest.mark.parametrize('seq, bad, match', [
    (P.cpmg, {'nloops': 0}, 'nloops'),
    (P.cpmg, {'timestep': -1.0}, 'timestep'),
    (P.sp_acquire, {'pulse_dur': 0.0}, 'pulse_dur'),
    (P.sp_acquire, {'pulse_pwr': -1.0}, 'pulse_pwr'),
    (P.sp_acquire, {'pulse_rnk': 1.5}, 'pulse_rnk'),
])
def test_sequence_inputs_are_validated(seq, bad, match):
    w = workloads.static_2spin(npoints=16, n_orient=2)
    ss = w['ss']
    base = {'rho0': w['rho0'], 'coil': w['coil'], 'grid': w['grid'], 'sweep': 1e5, 'npoints': 16}
    if seq is P.cpmg:
        base.update(pulse_op=S.operator(ss, 'Ly', '1H'), timestep=1e-5, nloops=1)
    else:
        base.update(spins=['1H'], pulse_frq=0.0, pulse_pwr=1e4, pulse_dur=1e-5, pulse_phi=0.0, pulse_rnk=1)
    base.update(bad)
    with pytest.raises(ValueError, match=match):
        P.powder(ss, seq, base, 'nmr')


@pytest.mark.parametrize('bad, match', [
    ({'echo_time': 0.0, 'echo_oper': None, 'echo_angle': 1.0}, 'echo_time'),
    ({'echo_time': 1e-5, 'echo_angle': 1.0}, 'echo_oper'),
    ({'echo_time': 1e-5, 'echo_oper': None}, 'echo_angle'),
    ({'echo_oper': None, 'echo_angle': 1.0}, 'echo time'),
    ({'echo_time': 1e-5, 'echo_oper': None, 'echo_angle': 1j}, 'echo_angle'),
])
def test_hp_acquire_echo_inputs_are_validated(bad, match):
    """hp_acquire.m:112-130; checked on the host before any device work."""
    w = workloads.static_2spin(npoints=16, n_orient=2)
    ss = w['ss']
    ly = S.operator(ss, 'Ly', '1H')
    params = {'rho0': S.state(ss, 'Lz', '1H'), 'coil': w['coil'], 'grid': w['grid'], 'sweep': 1e5, 'npoints': 16,
              'pulse_op': ly, 'pulse_angle': math.pi / 2}
    params.update({k: (ly if v is None else v) for k, v in bad.items()})
    with pytest.raises(ValueError, match=match):
        P.powder(ss, P.hp_acquire, params, 'nmr')
