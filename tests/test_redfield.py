"""Redfield relaxation theory on the host (kernel/relaxation.m:96-137,
kernel/includes/redfield_integral_serial.m, kernel/utilities/corrfun.m, expmint.m):
the orientation-independent superoperator the library takes as R0, pinned by the
reference's own known answers (test_dynamic_relaxation_models_suite.m:103-156: the
correlation-function weights and rates, a nonzero Redfield matrix that preserves the
unit state) and by the textbook chemical-shift-anisotropy rates of one spin under
isotropic rotational diffusion,
    R1 = (1/5) w0^2 S J(w0),   R2 = (1/30) w0^2 S (4 J(0) + 3 J(w0)),
with S the sum of squares of the traceless shielding eigenvalues and
J(w) = tau_c / (1 + w^2 tau_c^2) (Kowalewski and Maler, Nuclear Spin Relaxation in
Liquids, the CSA mechanism written for the traceless tensor). Checked without a GPU."""
import math

import numpy as np
import pytest

from spinach_amd import spinsys as S


def _one_spin(eigs, tau_c, keep='secular', iso='1H', dfs='ignore'):
    sys = {'magnet': 14.1, 'isotopes': [iso]}
    inter = {'zeeman': {'eigs': [eigs], 'euler': [[0.4, 1.1, -0.7]]}, 'relaxation': ['redfield'], 'tau_c': [tau_c],
             'rlx_keep': keep, 'equilibrium': 'zero', 'rlx_dfs': dfs}
    return S.basis(S.create(sys, inter), {'formalism': 'sphten-liouv', 'approximation': 'none'})


def test_corrfun_reference_known_answers():
    """test_dynamic_relaxation_models_suite.m:116-146 (1-based indices there, 0-based here)."""
    w, r = S.corrfun(1e-9, 2, 2, 1, 2, 1)
    assert abs(w[0] - 0.2) < 1e-15 and abs(r[0] + 1e9) < 1e-6 * 1e9
    w, r = S.corrfun([2e-9, 5e-9], 2, 2, 1, 2, 1)
    dax, deq = 1 / 12e-9, 1 / 30e-9
    assert abs(w[0] - 0.2) < 1e-15 and abs(r[0] + (6 * deq + (dax - deq))) < 1e-6 * abs(r[0])
    _, r = S.corrfun([1e-9, 2e-9, 3e-9], 2, 2, 2, 2, 2)
    dxx, dyy, dzz = 1 / 6e-9, 1 / 12e-9, 1 / 18e-9
    dl = math.sqrt(dxx ** 2 + dyy ** 2 + dzz ** 2 - dxx * dyy - dxx * dzz - dyy * dzz)
    ref = -np.array([4 * dxx + dyy + dzz, dxx + 4 * dyy + dzz, dxx + dyy + 4 * dzz, 2 * (dxx + dyy + dzz) - 2 * dl,
                     2 * (dxx + dyy + dzz) + 2 * dl])
    assert np.allclose(r, ref, rtol=1e-12)
    # isotropic: only the autocorrelation of the same component
    assert S.corrfun(1e-9, 2, 2, 1, 2, 3)[0] == [0.0]
    with pytest.raises(ValueError, match='different'):
        S.corrfun([1e-9, 1e-9, 3e-9], 2, 2, 2, 2, 2)


def test_redfield_is_nonzero_and_preserves_the_unit_state():
    """test_dynamic_relaxation_models_suite.m:148-156: an anisotropic one-spin system."""
    sys = {'magnet': 14.1, 'isotopes': ['1H']}
    inter = {'zeeman': {'matrix': [np.diag([1.0, 2.0, -3.0])]}, 'relaxation': ['redfield'], 'tau_c': [1e-9],
             'rlx_keep': 'labframe', 'equilibrium': 'zero', 'rlx_dfs': 'ignore'}
    ss = S.basis(S.create(sys, inter), {'formalism': 'sphten-liouv', 'approximation': 'none'})
    R = S.relaxation_terms(ss)['R0'].toarray()
    unit = np.eye(len(R))[0]   # the unit state is the first basis row
    assert np.count_nonzero(R) > 0
    assert np.abs(R @ unit).max() < 1e-12


@pytest.mark.parametrize('iso, tau_c', [('1H', 1e-9), ('1H', 5e-11), ('13C', 2e-9)])
def test_redfield_csa_rates_match_the_textbook(iso, tau_c):
    eigs = [-20.0, -5.0, 25.0]                    # ppm, traceless
    ss = _one_spin(eigs, tau_c, iso=iso)
    R = S.relaxation_terms(ss)['R0'].toarray()
    lz, lp = S.state(ss, 'Lz', iso), S.state(ss, 'L+', iso)
    r1 = -np.vdot(lz, R @ lz).real / np.vdot(lz, lz).real
    r2 = -np.vdot(lp, R @ lp).real / np.vdot(lp, lp).real
    # Lz and L+ are eigenvectors of the secular R
    assert np.abs(R @ lz + r1 * lz).max() < 1e-9 * r1 * np.abs(lz).max()
    assert np.abs(R @ lp + r2 * lp).max() < 1e-9 * r2 * np.abs(lp).max()
    w0 = abs(ss['basefrqs'][0])
    Ssum = sum((e * 1e-6) ** 2 for e in eigs)
    J = lambda w: tau_c / (1 + (w * tau_c) ** 2)
    R1 = w0 ** 2 * Ssum * J(w0) / 5
    R2 = w0 ** 2 * Ssum * (4 * J(0.0) + 3 * J(w0)) / 30
    print(f'{iso} tau_c {tau_c:g}: R1 {r1:.6g} vs {R1:.6g}, R2 {r2:.6g} vs {R2:.6g}')
    assert abs(r1 / R1 - 1) < 1e-4 and abs(r2 / R2 - 1) < 1e-4
    # planted: the correlation time of rank 1 (rate -2D) in place of rank 2's (-6D) moves both
    assert abs((w0 ** 2 * Ssum * (tau_c * 3) / (1 + (w0 * tau_c * 3) ** 2) / 5) / r1 - 1) > 1e-2


def test_redfield_retention_policies():
    """relaxation.m:556-634: 'secular' keeps elements between states of equal Zeeman frequency,
    'kite' the longitudinal block and the diagonal, 'diagonal' the diagonal without the unit
    state; all agree on the diagonal."""
    Rs = {k: S.relaxation_terms(_one_spin([-20.0, -5.0, 25.0], 1e-9, keep=k))['R0'].toarray()
          for k in ('labframe', 'secular', 'kite', 'diagonal')}
    for k in ('secular', 'kite', 'diagonal'):
        assert np.allclose(np.diag(Rs[k]), np.diag(Rs['labframe']), rtol=0, atol=1e-6)
    assert np.count_nonzero(Rs['diagonal'] - np.diag(np.diag(Rs['diagonal']))) == 0
    assert np.count_nonzero(Rs['labframe']) >= np.count_nonzero(Rs['secular'])


def test_redfield_inputs_are_validated():
    sys = {'magnet': 14.1, 'isotopes': ['1H']}
    base = {'zeeman': {'eigs': [[-20.0, -5.0, 25.0]], 'euler': [[0, 0, 0]]}, 'relaxation': ['redfield'],
            'rlx_keep': 'secular', 'equilibrium': 'zero'}
    with pytest.raises(ValueError, match='tau_c'):
        S.create(sys, base)
    with pytest.raises(ValueError, match='zero rotational correlation'):
        S.create(sys, dict(base, tau_c=[0.0]))
    with pytest.raises(ValueError, match='rlx_dfs'):
        S.create(sys, dict(base, tau_c=[1e-9], rlx_dfs='maybe'))


@pytest.mark.parametrize('tau_c', [1e-9, 2e-11])
def test_redfield_dipolar_rates_match_solomon(tau_c):
    """A 1H-13C pair 1.1 Angstrom apart, no shielding anisotropy: the longitudinal block of the
    Redfield superoperator holds Solomon's rates (b = (mu0/4pi) gI gS hbar / r^3):
        rho_I = (b^2/10) (J(wI - wS) + 3 J(wI) + 6 J(wI + wS)),
        sigma = (b^2/10) (6 J(wI + wS) - J(wI - wS)),
    and the transverse rate of I+ is (b^2/20) (4 J(0) + J(wI - wS) + 3 J(wI) + 6 J(wS) + 6 J(wI + wS))."""
    sys = {'magnet': 14.1, 'isotopes': ['1H', '13C']}
    inter = {'coordinates': [np.zeros(3), np.array([0.3, -0.5, 1.1 * math.sqrt(1 - 0.34 / 1.21)])],
             'relaxation': ['redfield'], 'tau_c': [tau_c], 'rlx_keep': 'kite', 'equilibrium': 'zero'}
    ss = S.basis(S.create(sys, inter), {'formalism': 'sphten-liouv', 'approximation': 'none'})
    R = S.relaxation_terms(ss)['R0'].toarray()
    iz, sz = S.state(ss, 'Lz', '1H'), S.state(ss, 'Lz', '13C')
    ip = S.state(ss, 'L+', '1H')
    nz = lambda v: v / np.linalg.norm(v)
    rho = -np.vdot(nz(iz), R @ nz(iz)).real
    sig = -np.vdot(nz(iz), R @ nz(sz)).real
    r2 = -np.vdot(nz(ip), R @ nz(ip)).real
    wI, wS = abs(ss['basefrqs'][0]), abs(ss['basefrqs'][1])
    r = 1.1e-10
    b = 1e-7 * S.GAMMAS['1H'] * S.GAMMAS['13C'] * S.HBAR / r ** 3
    J = lambda w: tau_c / (1 + (w * tau_c) ** 2)
    rho_ref = b ** 2 / 10 * (J(wI - wS) + 3 * J(wI) + 6 * J(wI + wS))
    sig_ref = b ** 2 / 10 * (6 * J(wI + wS) - J(wI - wS))
    r2_ref = b ** 2 / 20 * (4 * J(0) + J(wI - wS) + 3 * J(wI) + 6 * J(wS) + 6 * J(wI + wS))
    print(f'tau_c {tau_c:g}: rho {rho:.6g} vs {rho_ref:.6g}, sigma {sig:.6g} vs {sig_ref:.6g}, '
          f'R2 {r2:.6g} vs {r2_ref:.6g}')
    assert abs(rho / rho_ref - 1) < 1e-4 and abs(sig / sig_ref - 1) < 1e-4 and abs(r2 / r2_ref - 1) < 1e-4


def test_redfield_validity_condition():
    """relaxation.m:126-134: a correlation time longer than the fastest relaxation time is an
    error (a slow-tumbling CSA spin at 10 ms); an axial pair of times errs only when both are."""
    with pytest.raises(ValueError, match='validity'):
        S.relaxation_terms(_one_spin([-2000.0, -500.0, 2500.0], 1e-2))
    R = S.relaxation_terms(_one_spin([-20.0, -5.0, 25.0], [1e-9, 2e-9]))['R0']
    assert R.nnz > 0


def test_redfield_two_species_take_their_own_correlation_times():
    """inter.chem.parts with one tau_c per species (corrfun.m:49-54): an uncoupled 1H and 13C,
    each with a shielding tensor, in two species tumbling at 1 ns and 50 ps: each spin's rates are
    the one-spin CSA rates at its own correlation time; one species for both would not be."""
    eig = [[-20.0, -5.0, 25.0], [40.0, -10.0, -30.0]]
    sys = {'magnet': 14.1, 'isotopes': ['1H', '13C']}
    base = {'zeeman': {'eigs': eig, 'euler': [[0.4, 1.1, -0.7], [1.3, 0.2, 2.0]]}, 'relaxation': ['redfield'],
            'rlx_keep': 'secular', 'equilibrium': 'zero'}
    ss = S.basis(S.create(sys, dict(base, tau_c=[1e-9, 5e-11], chem={'parts': [[0], [1]]})),
                 {'formalism': 'sphten-liouv', 'approximation': 'none'})
    R = S.relaxation_terms(ss)['R0'].toarray()
    for k, (iso, tc) in enumerate((('1H', 1e-9), ('13C', 5e-11))):
        lz = S.state(ss, 'Lz', iso)
        r1 = -np.vdot(lz, R @ lz).real / np.vdot(lz, lz).real
        w0 = abs(ss['basefrqs'][k])
        ref = w0 ** 2 * sum((e * 1e-6) ** 2 for e in eig[k]) * tc / (1 + (w0 * tc) ** 2) / 5
        assert abs(r1 / ref - 1) < 1e-4, (iso, r1, ref)
    one = S.basis(S.create(sys, dict(base, tau_c=[1e-9])), {'formalism': 'sphten-liouv', 'approximation': 'none'})
    R1 = S.relaxation_terms(one)['R0'].toarray()
    lz = S.state(one, 'Lz', '13C')
    assert abs(np.vdot(lz, R1 @ lz).real / np.vdot(lz, R @ lz).real - 1) > 0.1
    with pytest.raises(ValueError, match='one entry per chemical species'):
        S.create(sys, dict(base, tau_c=[1e-9], chem={'parts': [[0], [1]]}))


def test_redfield_dynamic_frequency_shifts():
    """relaxation.m:533-552: rlx_dfs = 'keep' leaves the imaginary part (the dynamic frequency
    shift) of the integral; 'ignore' (the default) takes the real part only."""
    Rk = S.relaxation_terms(_one_spin([-20.0, -5.0, 25.0], 1e-9, keep='labframe', dfs='keep'))['R0'].toarray()
    Ri = S.relaxation_terms(_one_spin([-20.0, -5.0, 25.0], 1e-9, keep='labframe'))['R0'].toarray()
    assert np.abs(Rk.imag).max() > 1e-3 * np.abs(Rk.real).max()
    assert np.abs(Ri.imag).max() == 0.0 and np.allclose(Rk.real, Ri.real, rtol=0, atol=1e-12)
