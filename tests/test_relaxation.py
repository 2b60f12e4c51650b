"""Orientation-dependent relaxation (kernel/relaxation.m 't1_t2' and 'damp',
kernel/utilities/rlx_t1_t2.m) on the host and in the oracle: the reference's
own known answers (tests/kernel/test_dynamic_relaxation_models_suite.m,
test_relaxation_t2_rate.m), the library's quadratic-form expansion against
the oracle's direct restatement, and the oracle against closed forms with
planted mistakes that must fail."""
import math

import numpy as np
import pytest

import pyoracle as O
from spinach_amd import grids, spinsys as S, workloads

T1 = np.array([[1.0, 0.3, -0.2], [0.3, 3.0, 0.5], [-0.2, 0.5, 5.0]]) * 40.0
T2 = np.array([[2.0, -0.4, 0.6], [-0.4, 4.0, 0.1], [0.6, 0.1, 6.0]]) * 60.0
TD = np.array([[1.0, 0.2, 0.0], [0.2, 2.0, -0.3], [0.0, -0.3, 1.5]]) * 25.0


def _one_spin(inter_extra, approximation='none'):
    sys = {'magnet': 14.1, 'isotopes': ['1H']}
    inter = {'zeeman': {'scalar': [0.0]}}
    inter.update(inter_extra)
    ss = S.create(sys, inter)
    return S.basis(ss, {'formalism': 'sphten-liouv', 'approximation': approximation})


def test_reference_known_answers_t1_t2_tensors_at_zero_angles():
    """test_dynamic_relaxation_models_suite.m:25-42: R1 = diag(1,3,5), R2 =
    diag(2,4,6) at zero Euler angles give -5 on Lz and -6 on L+."""
    ss = _one_spin({'relaxation': ['t1_t2'], 'r1_rates': [np.diag([1.0, 3, 5])], 'r2_rates': [np.diag([2.0, 4, 6])],
                    'equilibrium': 'zero', 'rlx_keep': 'secular'})
    R = S.relaxation(ss, (0.0, 0.0, 0.0))
    lz, lp = S.state(ss, 'Lz', '1H'), S.state(ss, 'L+', '1H')
    assert np.allclose(R @ lz, -5 * lz, atol=1e-12) and np.allclose(R @ lp, -6 * lp, atol=1e-12)


def test_reference_known_answers_scalar_rates_and_damping():
    """test_relaxation_t2_rate.m:25-40 (R2 = 7 on L+) and the damping block of
    the suite (:58-70: unit state untouched, -5 on L+)."""
    ss = _one_spin({'relaxation': ['t1_t2'], 'r1_rates': [2.0], 'r2_rates': [7.0], 'equilibrium': 'zero',
                    'rlx_keep': 'secular'})
    lp = S.state(ss, 'L+', '1H')
    assert np.allclose(S.relaxation(ss) @ lp, -7 * lp, atol=1e-12)
    ss = _one_spin({'relaxation': ['damp'], 'damp_rate': 5.0, 'equilibrium': 'zero', 'rlx_keep': 'labframe'})
    R = S.relaxation(ss)
    unit = np.zeros(len(ss['basis']))
    unit[0] = 1.0
    assert np.allclose(R @ unit, 0.0) and np.allclose(R @ S.state(ss, 'L+', '1H'), -5 * S.state(ss, 'L+', '1H'))


@pytest.mark.parametrize('bad, exc', [
    ({'relaxation': ['lindblad'], 'rlx_keep': 'secular'}, NotImplementedError),
    ({'relaxation': ['redfield'], 'rlx_keep': 'secular'}, ValueError),                      # no tau_c
    ({'relaxation': ['redfield'], 'rlx_keep': 'secular', 'tau_c': [1e-9, 2e-9]}, ValueError),  # 2 times, 1 species
    ({'relaxation': ['t1_t2'], 'r1_rates': [1.0], 'r2_rates': [1.0], 'rlx_keep': 'secular',
      'equilibrium': 'IME'}, NotImplementedError),
    ({'relaxation': ['t1_t2'], 'r1_rates': [1.0], 'r2_rates': [1.0]}, ValueError),           # no rlx_keep
    ({'relaxation': ['t1_t2'], 'r1_rates': [1.0], 'rlx_keep': 'secular'}, ValueError),       # no r2_rates
    ({'relaxation': ['t1_t2'], 'r1_rates': [np.ones((2, 2))], 'r2_rates': [1.0], 'rlx_keep': 'secular'}, ValueError),
    ({'relaxation': ['damp'], 'rlx_keep': 'secular'}, ValueError),                          # no damp_rate
])
def test_relaxation_specification_is_validated(bad, exc):
    with pytest.raises(exc):
        _one_spin(bad)


def test_quadratic_form_expansion_equals_the_direct_restatement():
    """The library's R(Omega) = diag(g0 + sum_p q_p g_p) (relaxation_terms and
    the kernel's monomials q_p of ort) against the oracle's ort*T*ort' per
    state, and the host's own relaxation(ss, euler), on the two-spin system
    with tensor and scalar rates and a damping tensor."""
    w = workloads.static_2spin(npoints=8, n_orient=2, relax={
        'relaxation': ['t1_t2', 'damp'], 'r1_rates': [3.0, T1], 'r2_rates': [T2, 11.0], 'damp_rate': TD,
        'rlx_keep': 'secular'})
    ss = w['ss']
    g = S.relaxation_terms(ss)['diag']
    rng = np.random.default_rng(3)
    for _ in range(12):
        eul = tuple(rng.uniform(-math.pi, math.pi, 3))
        b, c = eul[1], eul[2]
        ox, oy, oz = -math.sin(b) * math.cos(c), math.sin(b) * math.sin(c), math.cos(b)
        q = np.array([ox * ox, oy * oy, oz * oz, 2 * ox * oy, 2 * ox * oz, 2 * oy * oz])
        lib = g[0] + q @ g[1:]
        ref = O.relax_diag_at(ss['rlx'], ss['basis'], eul)
        host = S.relaxation(ss, eul).diagonal()
        assert np.allclose(lib, ref, rtol=1e-13, atol=1e-11)
        assert np.allclose(host, ref, rtol=1e-13, atol=1e-11)
        assert not np.allclose(O.relax_diag_at(ss['rlx'], ss['basis'], eul, plant=O.PLANT_RLX_ORT), ref,
                               rtol=1e-6, atol=1e-6)


def _check(f, ref, tol):
    return O.rel_err(f, ref) < tol


def test_oracle_two_spin_closed_form_with_relaxation():
    """np_powder_relax against the closed form of the exchanging I+ / 2I+Sz
    pair with R2 (tensor) on spin 0, R1 (tensor) on spin 1 and a damping
    tensor; the sign and orientation-vector plants must fail."""
    relax = {'relaxation': ['t1_t2', 'damp'], 'r1_rates': [4.0, T1], 'r2_rates': [T2, 9.0], 'damp_rate': TD,
             'rlx_keep': 'secular'}
    w = workloads.static_2spin(npoints=256, n_orient=16, relax=relax)
    ss = w['ss']
    g = grids.haar(16, np.random.default_rng(5))
    args = (w['I'], w['Q'], g, w['rho0'], w['coil'], w['dt'], w['nsteps'], ss['rlx'], ss['basis'])
    f = O.np_powder_relax(*args)
    cf = O.closed_form_2spin(ss, g, w['dt'], w['nsteps'], relax=ss['rlx'])
    f0 = O.closed_form_2spin(ss, g, w['dt'], w['nsteps'])
    e = O.rel_err(f, cf)
    print(f'2-spin with relaxation: oracle vs closed form {e:.2e}; relaxation changes the FID by {O.rel_err(f0, cf):.2f}')
    assert e < 5e-9
    assert O.rel_err(f0, cf) > 0.05
    for plant in (O.PLANT_RLX_SIGN, O.PLANT_RLX_ORT):
        assert O.rel_err(O.np_powder_relax(*args, plant=plant), cf) > 1e-3, plant


def test_oracle_one_spin_anisotropic_t2_decay():
    """One CSA spin, rho0 = coil = L+: each orientation's FID is the static one
    times exp(-R2(Omega) t) with R2(Omega) = ort*T2*ort'."""
    sys = {'magnet': 14.1, 'isotopes': ['1H']}
    inter = {'zeeman': {'eigs': [[-20.0, -5.0, 25.0]], 'euler': [[0.3, 0.9, -0.4]]},
             'relaxation': ['t1_t2'], 'r1_rates': [T1], 'r2_rates': [T2], 'rlx_keep': 'secular'}
    ss = S.basis(S.create(sys, inter), {'formalism': 'sphten-liouv', 'approximation': 'none', 'projections': [1]})
    I, Q = S.hamiltonian(S.assume(ss, 'nmr'))
    q = S.q_list(Q)
    rho = S.state(ss, 'L+', '1H')
    g = grids.haar(12, np.random.default_rng(2))
    dt, nst = 2e-5, 400
    f, per = O.np_powder_relax(I, q, g, rho, rho, dt, nst, ss['rlx'], ss['basis'], per_orientation=True)
    _, per0 = O.np_powder(I, q, g, rho, rho, dt, nst, per_orientation=True)
    t = np.arange(nst + 1) * dt
    for o in range(12):
        R = O._dcm(g['alphas'][o], g['betas'][o], g['gammas'][o])
        r2 = R[2, :] @ T2 @ R[2, :]
        assert O.rel_err(per[o], per0[o] * np.exp(-r2 * t)) < 1e-11


SIX_EULER = np.array([[0, np.pi / 2, np.pi], [0, np.pi / 2, np.pi / 2], [0, 0, 0],
                      [0, np.pi / 2, 3 * np.pi / 4], [0, np.pi / 4, np.pi], [0, np.pi / 4, np.pi / 2]])


def test_six_orientation_sampling_recovers_the_planes():
    """The MATLAB caller's recipe for the gateway's opts.rlx_diag: the
    diagonal of relaxation() at the six orientations whose ort is e_x, e_y,
    e_z, (e_x+e_y)/sqrt2, (e_x+e_z)/sqrt2, (e_y+e_z)/sqrt2 gives g0 = 0,
    g1..g3 = d1..d3, g4 = d4-(d1+d2)/2, g5 = d5-(d1+d3)/2, g6 = d6-(d2+d3)/2;
    the planes so obtained must reproduce the oracle's rates at any
    orientation (g0 folds into g1+g2+g3 because ox^2+oy^2+oz^2 = 1)."""
    w = workloads.static_2spin(npoints=8, n_orient=2, relax={
        'relaxation': ['t1_t2', 'damp'], 'r1_rates': [3.0, T1], 'r2_rates': [T2, 11.0], 'damp_rate': TD,
        'rlx_keep': 'secular'})
    ss = w['ss']
    d = np.stack([S.relaxation(ss, tuple(e)).diagonal() for e in SIX_EULER], axis=1)
    for k, e in enumerate(SIX_EULER):   # the angles give the stated orientation vectors
        ort = O._dcm(*e)[2, :]
        want = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [0, 1, 1]][k]
        assert np.allclose(ort, np.array(want) / np.linalg.norm(want), atol=1e-15)
    g = np.stack([np.zeros(d.shape[0]), d[:, 0], d[:, 1], d[:, 2], d[:, 3] - (d[:, 0] + d[:, 1]) / 2,
                  d[:, 4] - (d[:, 0] + d[:, 2]) / 2, d[:, 5] - (d[:, 1] + d[:, 2]) / 2])
    rng = np.random.default_rng(11)
    for _ in range(8):
        eul = tuple(rng.uniform(-math.pi, math.pi, 3))
        ox, oy, oz = O._dcm(*eul)[2, :]
        q = np.array([ox * ox, oy * oy, oz * oz, 2 * ox * oy, 2 * ox * oz, 2 * oy * oz])
        assert np.allclose(g[0] + q @ g[1:], O.relax_diag_at(ss['rlx'], ss['basis'], eul), rtol=1e-13, atol=1e-11)
