"""The oracle pinned against mathematics and against itself: closed-form
two-spin FIDs, three independent restatements (C++ faithful, C++ union
reduction, numpy/scipy), scipy's expm, the reference's own invariants
(tests/kernel/test_ctx_powder_average.m, test_ctx_powder_crystal.m,
test_dynamic_propagation_frontends.m:39-58, test_ctx_singlerot_acquire.m:58-60),
and a planted mistake that the checks must catch."""
import math

import numpy as np
import pytest
import scipy.linalg as sla

import pyoracle as O
from spinach_amd import grids, spinsys as S, workloads


@pytest.fixture(scope='module')
def two_spin():
    return workloads.static_2spin(npoints=128, n_orient=12)


@pytest.fixture(scope='module')
def four_spin():
    rng = np.random.default_rng(11)
    w = workloads.random_system(rng, 4, level=3)
    w['grid'] = grids.haar(5, rng)
    w['dt'] = 5e-6
    w['nsteps'] = 200
    return w


def test_expm_matches_scipy():
    rng = np.random.default_rng(0)
    for n, scale in ((4, 1e-3), (9, 0.3), (16, 3.0), (40, 30.0), (60, 300.0)):
        a = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) * scale / n
        e = O.refcpu_expm(a)
        r = sla.expm(a)
        assert np.linalg.norm(e - r) / np.linalg.norm(r) < 1e-11, (n, scale)


def test_chop_taylor_propagator_vs_expm():
    # the reference's check: one-spin L = comm(Lz), dt = 2.5e-4, Taylor branch
    # forced by small_matrix = 2, against expm with test_close tolerances
    # 1e-10 + 1e-10*max(1,|ref|) (test_dynamic_propagation_frontends.m:39-52)
    ss = S.basis(S.create({'magnet': 14.1, 'isotopes': ['1H']}, {}), {'approximation': 'none'})
    L = S.operator(ss, 'Lz', 0).toarray()
    dt = 2.5e-4
    p_obs = O.refcpu_propagator(L, dt, small_matrix=2.0)
    p_ref = sla.expm(-1j * L * dt)
    assert np.linalg.norm(p_obs - p_ref, 2) <= 1e-10 + 1e-10 * max(1.0, np.linalg.norm(p_ref, 2))
    # a dense generator with |A|_1 ~ 5 (three squarings): the 1e-10 chop of
    # every Taylor term and squaring bounds the element error by a few 1e-9
    rng = np.random.default_rng(1)
    n = 24
    h = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    h = (h + h.conj().T) * 1e4
    p_taylor = O.refcpu_propagator(h, 1e-5, small_matrix=2.0)
    p_exact = sla.expm(-1j * 1e-5 * h)
    assert np.abs(p_taylor - p_exact).max() < 5e-9
    assert np.abs(O.refcpu_propagator(h, 1e-5, small_matrix=200.0) - p_exact).max() < 1e-12


def test_two_spin_closed_form_all_restatements(two_spin):
    w = two_spin
    args = (w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'])
    cf = O.closed_form_2spin(w['ss'], w['grid'], w['dt'], w['nsteps'])
    fids = [O.refcpu_powder(*args, variant=v) for v in (0, 1, 2)]
    fids.append(O.np_powder(*args))
    fids.append(O.np_powder(*args, reduce=False))
    for f in fids:
        # input rounding: the 3.8e9 rad/s carrier cancels in FP64 inside the
        # reference's own conversion (hamiltonian.m:99-100, mat2sphten.m)
        assert O.rel_err(f, cf) < 5e-9
        assert O.rel_err(f, fids[0]) < 1e-12


def test_two_spin_full_length_closed_form():
    w = workloads.static_2spin(n_orient=16)
    f = O.refcpu_powder(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'])
    cf = O.closed_form_2spin(w['ss'], w['grid'], w['dt'], w['nsteps'])
    assert O.rel_err(f, cf) < 5e-8
    assert abs(f[0] - 0.25) < 1e-14  # f(0) = Tr(L- Lx)/N = 1/4


def test_planted_sign_error_is_caught(two_spin):
    # a wrong sign of the isotropic generator must fail the closed form
    w = two_spin
    f = O.refcpu_powder(-w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'])
    cf = O.closed_form_2spin(w['ss'], w['grid'], w['dt'], w['nsteps'])
    assert O.rel_err(f, cf) > 1e-2
    # and so must a wrong projection index in the Wigner contraction
    # (Q{2}{k,m} paired with D(k,-m))
    qm = list(w['Q'])
    for k in range(5):
        row = [w['Q'][9 + 5 * k + m] for m in range(5)]
        for m in range(5):
            qm[9 + 5 * k + m] = row[4 - m]
    f2 = O.refcpu_powder(w['I'], qm, w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'])
    assert O.rel_err(f2, cf) > 1e-2


def test_reduction_variants_agree(four_spin):
    w = four_spin
    args = (w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'])
    f0 = O.refcpu_powder(*args, variant=0)
    f2 = O.refcpu_powder(*args, variant=2)
    fn = O.np_powder(*args, reduce=False)
    assert O.rel_err(f2, f0) < 1e-11
    assert O.rel_err(fn, f0) < 1e-11


def test_weighted_sum_equals_explicit_sum(four_spin):
    w = four_spin
    fid, per = O.refcpu_powder(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'],
                               per_orientation=True)
    ref = w['grid']['weights'][0] * per[0]
    for k in range(1, len(per)):
        ref = ref + w['grid']['weights'][k] * per[k]
    assert np.abs(fid - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
    # f(0) = coil' * rho0 for every orientation
    assert np.abs(per[:, 0] - np.vdot(w['coil'], w['rho0'])).max() < 1e-13


def test_single_crystal_equals_single_orientation(four_spin):
    w = four_spin
    g = grids.single_crystal(0.3, 1.2, 2.0)
    f = O.refcpu_powder(w['I'], w['Q'], g, w['rho0'], w['coil'], w['dt'], w['nsteps'])
    h = O.assemble_dense(w['I'], w['Q'], (0.3, 1.2, 2.0))
    P = sla.expm(-1j * w['dt'] * h)
    r = w['rho0'].astype(complex)
    ref = []
    for _ in range(w['nsteps'] + 1):
        ref.append(np.vdot(w['coil'], r))
        r = P @ r
    assert O.rel_err(f, np.array(ref)) < 1e-11


def test_hermitian_generator_conserves_norm(four_spin):
    w = four_spin
    h = O.assemble_dense(w['I'], w['Q'], (0.1, 0.7, 1.9))
    assert np.abs(h - h.conj().T).max() == 0.0
    P = O.refcpu_expm(-1j * w['dt'] * h)
    r = P @ w['rho0']
    assert abs(np.linalg.norm(r) - np.linalg.norm(w['rho0'])) < 1e-13


def test_large_subspace_uses_chop_taylor_and_stays_close():
    # a subspace above small_matrix=200 takes the chop-Taylor branch in the
    # faithful variant; its 1e-10 quantisation moves the FID by far less
    # than 1e-6 against exact expm (variant 1)
    rng = np.random.default_rng(4)
    w = workloads.random_system(rng, 6, level=3)
    g = grids.haar(2, rng)
    args = (w['I'], w['Q'], g, w['rho0'], w['coil'], 5e-6, 256)
    st = O.union_states(w['I'], w['Q'], w['coil'])
    f0 = O.refcpu_powder(*args, variant=0)
    f1 = O.refcpu_powder(*args, variant=1)
    if len(st) >= 200:
        assert O.rel_err(f0, f1) < 1e-6
    else:
        assert O.rel_err(f0, f1) < 1e-12
