"""Host side of the crystal() and singlerot() contexts, Spinach grid files,
the Hamiltonian cache and the hp_acquire pulse (spinach_amd/contexts.py,
grids.py, spinsys.py, powder.py), checked without a GPU.

The Fokker-Planck construction of singlerot (kernel/contexts/singlerot.m:
150-241) is pinned by the exact MAS signal of one spin with a chemical shift
tensor: under spinning its offset is a trigonometric polynomial of the rotor
phase, omega(phi) = sum_{|m|<=2} c_m e^{i m phi}, so the rotor-phase average of
exp(-i int omega dt) is known in closed form; the Fokker-Planck FID must
converge to it spectrally in max_rank, and its complex conjugate (spinning
backwards) must not match.
"""
import math
import os

import numpy as np
import pytest
import scipy.io as sio
import scipy.linalg as sla

import pyoracle as O
from spinach_amd import contexts as Cx
from spinach_amd import grids, workloads
from spinach_amd import powder as P
from spinach_amd import spinsys as S

HERE = os.path.dirname(os.path.abspath(__file__))


# ------------------------------------------------------------------ fourdif

@pytest.mark.parametrize('N', [5, 8, 9, 16, 33])
def test_fourdif_differentiates_trigonometric_polynomials_exactly(N):
    x, D = Cx.fourdif(N)
    assert np.allclose(x, 2 * np.pi * np.arange(N) / N)
    assert np.allclose(D, -D.T, atol=1e-14)
    for k in range(1, (N - 1) // 2 + 1):
        assert np.allclose(D @ np.sin(k * x), k * np.cos(k * x), atol=1e-11)
        assert np.allclose(D @ np.cos(k * x), -k * np.sin(k * x), atol=1e-11)
    assert np.allclose(D @ np.ones(N), 0, atol=1e-12)
    # a sign error or the other parity's kernel fails
    assert not np.allclose(-D @ np.sin(x), np.cos(x), atol=1e-6)


def test_fourdif_parity_kernels_differ():
    _, D = Cx.fourdif(7)
    h = 2 * np.pi / 7
    wrong = np.array([[0 if j == k else 0.5 * (-1) ** (j - k) / math.tan((j - k) * h / 2) for k in range(7)]
                      for j in range(7)])
    x = h * np.arange(7)
    assert not np.allclose(wrong @ np.sin(3 * x), 3 * np.cos(3 * x), atol=1e-6)
    assert np.allclose(D @ np.sin(3 * x), 3 * np.cos(3 * x), atol=1e-11)


# ------------------------------------------------------- Fokker-Planck

def _one_spin(eigs=(-8.0, -1.0, 9.0), euler=(0.3, 0.7, 1.1), formalism='sphten-liouv'):
    ss = S.create({'magnet': 14.1, 'isotopes': ['1H']}, {'zeeman': {'eigs': [list(eigs)], 'euler': [list(euler)]}})
    bas = {'formalism': formalism}
    if formalism == 'sphten-liouv':
        bas['approximation'] = 'none'
    return S.basis(ss, bas)


def test_fokker_planck_blocks_are_the_rotor_frame_hamiltonians():
    """rate = 0: G(Omega) = blkdiag(H_1..H_N) with H_n the Hamiltonian at
    D_comp,n = D_lab2rot D_rotor(phi_n) D(Omega) (rotor_stack.m 'rotor' frame),
    which the oracle composes from expm of the angular momentum matrices."""
    w = workloads.static_2spin(npoints=8, n_orient=2)
    I, Q = S.hamiltonian(S.assume(w['ss'], 'nmr'))
    axis, mr = (1.0, 1.0, 1.0), 2
    Ifp, qfp = Cx.fokker_planck(I, Q, axis, 0.0, mr)
    N, dim = 2 * mr + 1, I.shape[0]
    eul = (0.4, 1.2, 2.1)
    G = O.assemble_dense(Ifp, qfp, eul)
    phases = 2 * np.pi * np.arange(N) / N
    for n in range(N):
        Hn = O.rotor_hamiltonian(I, S.q_list(Q), axis, eul, phases[n], 'rotor')
        blk = G[n * dim:(n + 1) * dim, n * dim:(n + 1) * dim]
        assert np.allclose(blk, (Hn + Hn.conj().T) / 2, atol=1e-9 * np.abs(Hn).max())
        for m in range(N):
            if m != n:
                assert np.abs(G[n * dim:(n + 1) * dim, m * dim:(m + 1) * dim]).max() == 0
    # the spinning term: off-diagonal blocks 1i*2*pi*rate*d_dphi[n,m]*1, G Hermitian
    rate = 1234.0
    Ifp2, qfp2 = Cx.fokker_planck(I, Q, axis, rate, mr)
    G2 = O.assemble_dense(Ifp2, qfp2, eul)
    _, d = Cx.fourdif(N)
    assert np.allclose(G2[0:dim, dim:2 * dim], 1j * 2 * np.pi * rate * d[0, 1] * np.eye(dim), atol=1e-9)
    assert np.allclose(G2, G2.conj().T, atol=1e-6)


def _analytic_mas(eul, rate, axis, t, n_phase=400):
    """Rotor-phase average of exp(-i int_0^t omega(phi0 + w t') dt') for one 1H
    spin, omega(phi) = H(phi)[0,0] - H(phi)[1,1] from the Hilbert rotor
    Hamiltonian ('rotor' frame) and its Fourier coefficients."""
    ssh = _one_spin(formalism='zeeman-hilb')
    Ih, Qh = S.hamiltonian(S.assume(ssh, 'nmr'))
    ql = S.q_list(Qh)
    K = 64
    ph = 2 * np.pi * np.arange(K) / K
    om = np.array([(lambda H: (H[0, 0] - H[1, 1]).real)(O.rotor_hamiltonian(Ih, ql, axis, eul, p, 'rotor'))
                   for p in ph])
    c = np.fft.fft(om) / K
    ms = np.fft.fftfreq(K, 1.0 / K)
    w = 2 * np.pi * rate
    p0 = 2 * np.pi * np.arange(n_phase) / n_phase
    Phi = np.zeros((n_phase, len(t)))
    for cm, m in zip(c, ms):
        if abs(cm) < 1e-12 * np.abs(c).max():
            continue
        if m == 0:
            Phi += cm.real * t[None, :]
        else:
            Phi += np.real(cm * np.exp(1j * m * p0)[:, None] * (np.exp(1j * m * w * t)[None, :] - 1) / (1j * m * w))
    return np.mean(np.exp(-1j * Phi), axis=0)


def fp_one_spin_case(max_rank, n_or=3):
    ss = _one_spin()
    I, Q = S.hamiltonian(S.assume(ss, 'nmr'))
    rate, axis = 2000.0, (1.0, 1.0, 1.0)
    rho0, coil = S.state(ss, 'L+', '1H'), S.state(ss, 'L+', '1H')
    g = grids.haar(n_or, np.random.default_rng(5))
    sweep, npts = 8 * rate, 64
    Ifp, qfp = Cx.fokker_planck(I, Q, axis, rate, max_rank)
    N = 2 * max_rank + 1
    r0, c0 = np.kron(np.ones(N) / N, rho0), np.kron(np.ones(N), coil)
    t = np.arange(npts) / sweep
    exact = [_analytic_mas((g['alphas'][o], g['betas'][o], g['gammas'][o]), rate, axis, t) * np.vdot(coil, rho0)
             for o in range(n_or)]
    return dict(I=Ifp, Q=qfp, grid=g, rho0=r0, coil=c0, dt=1 / sweep, nsteps=npts - 1, exact=np.array(exact),
                ss=ss, rate=rate, axis=axis)


def test_fokker_planck_converges_to_exact_mas_signal():
    errs = []
    for mr in (8, 16):
        c = fp_one_spin_case(mr)
        _, per = O.np_powder(c['I'], c['Q'], c['grid'], c['rho0'], c['coil'], c['dt'], c['nsteps'],
                             per_orientation=True)
        e = max(O.rel_err(per[o], c['exact'][o]) for o in range(3))
        eb = min(O.rel_err(per[o], np.conj(c['exact'][o])) for o in range(3))
        print(f'max_rank {mr}: Fokker-Planck vs exact MAS {e:.2e}, vs the backward-spinning signal {eb:.2e}')
        errs.append(e)
        assert eb > 1e-2
    assert errs[0] < 1e-2 and errs[1] < 1e-9


# ------------------------------------------------------------ grid files

def test_spinach_grid_files_load_by_name_and_path(tmp_path, monkeypatch):
    fix = np.load(os.path.join(HERE, 'golden', 'grids.npz'))
    gdir = tmp_path / 'kernel' / 'grids'
    gdir.mkdir(parents=True)
    for name in ('leb_2ang_rank_5', 'rep_3ang_100pts'):
        sio.savemat(str(gdir / (name + '.mat')), {k: fix[f'{name}__{k}'].reshape(-1, 1)
                                                   for k in ('alphas', 'betas', 'gammas', 'weights')})
    monkeypatch.setenv('SPINACH_ROOT', str(tmp_path))
    g = grids.load_grid('leb_2ang_rank_5')
    assert len(g['weights']) == 14 and abs(g['weights'].sum() - 1) < 1e-12
    for k in ('alphas', 'betas', 'gammas', 'weights'):
        assert np.array_equal(g[k], fix[f'leb_2ang_rank_5__{k}'])
    g2 = grids.load_grid(str(gdir / 'rep_3ang_100pts.mat'))
    assert np.array_equal(g2['gammas'], fix['rep_3ang_100pts__gammas'])
    assert P._grid('rep_3ang_100pts')['alphas'].shape == (100,)
    sio.savemat(str(gdir / 'broken.mat'), {'alphas': np.zeros(3), 'betas': np.zeros(3), 'gammas': np.zeros(2),
                                           'weights': np.ones(3)})
    with pytest.raises(ValueError):
        grids.load_grid('broken')
    monkeypatch.delenv('SPINACH_ROOT')
    with pytest.raises(ValueError, match='SPINACH_ROOT'):
        grids.load_grid('leb_2ang_rank_5')


# ------------------------------------------------------------ ham cache

def test_hamiltonian_cache_returns_equal_copies_keyed_by_inputs():
    w = workloads.static_2spin(npoints=8, n_orient=2)
    ss = dict(S.assume(w['ss'], 'nmr'))
    ss['enable'] = ['ham_cache']
    I1, Q1 = S.hamiltonian(ss)
    I2, Q2 = S.hamiltonian(ss)
    I0, Q0 = S.hamiltonian(S.assume(w['ss'], 'nmr'))
    assert (I1 != I0).nnz == 0 and (I2 != I0).nnz == 0
    I2.data[:] = 0  # the cache hands out copies
    I3, _ = S.hamiltonian(ss)
    assert (I3 != I0).nnz == 0
    other = dict(ss)
    other['magnet'] = 9.4
    other['basefrqs'] = [b * 9.4 / 14.1 for b in ss['basefrqs']]
    assert S._ham_key(other) != S._ham_key(ss)


# ------------------------------------------------------------ hp_acquire

def test_hp_acquire_pulse_turns_lz_into_lx():
    """hp_acquire.m:47: step(pulse_op, rho0, pulse_angle) with pulse_op = Ly
    (commutation superoperator) and angle pi/2 takes Lz to Lx."""
    w = workloads.static_2spin(npoints=8, n_orient=2)
    ss = w['ss']
    rho = P._hp_pulse({'rho0': S.state(ss, 'Lz', '1H'), 'pulse_op': S.operator(ss, 'Ly', '1H'),
                       'pulse_angle': math.pi / 2})
    assert np.allclose(rho, S.state(ss, 'Lx', '1H'), atol=1e-12)
    # an echo needs its operator and angle (hp_acquire.m:118-124)
    with pytest.raises(ValueError, match='echo_oper'):
        P._hp_pulse({'rho0': rho, 'pulse_op': S.operator(ss, 'Ly', '1H'), 'pulse_angle': 1.0, 'echo_time': 1e-3})


# --------------------------------------------------------------- crystal

@pytest.mark.parametrize('bad', [None, [0, 0], [0, np.nan, 0]])
def test_crystal_validates_the_orientation(bad):
    w = workloads.static_2spin(npoints=8, n_orient=2)
    params = {'rho0': w['rho0'], 'coil': w['coil'], 'sweep': 1e5, 'npoints': 8}
    if bad is not None:
        params['orientation'] = bad
    with pytest.raises(ValueError, match='orientation'):
        Cx.crystal(w['ss'], P.acquire, params, 'nmr')


# ------------------------------------------------------------------ gridfree (SLE)

def sle_dense_fid(H, R, rho0, coil, dt, nsteps):
    """FID of the stochastic Liouville equation by dense expm stepping on the states reachable
    from rho0 (connected component of the generator's pattern): f(n) = coil' exp(-i L dt)^n
    rho0, L = H + iR (evolution.m's definition)."""
    import scipy.sparse as sps
    import scipy.sparse.csgraph as cg
    L = sps.csr_matrix(H, dtype=complex)
    if R is not None:
        L = L + 1j * sps.csr_matrix(R)
    _, lab = cg.connected_components((abs(L) + abs(L).T) > 0, directed=False)
    keep = np.isin(lab, np.unique(lab[np.abs(rho0) > 0]))
    Ls = L[keep][:, keep].toarray()
    P_ = sla.expm(-1j * dt * Ls)
    x, c = np.asarray(rho0, complex)[keep], np.asarray(coil, complex)[keep]
    out = []
    for _ in range(int(nsteps) + 1):
        out.append(np.vdot(c, x))
        x = P_ @ x
    return np.array(out)


def sle_one_spin_case(max_rank, **kw):
    """One 1H with a chemical shift tensor; rho0 = coil = L+ (the SLE state kron(e_000, .))."""
    ss, I, q = _one_spin_csa_nested()
    H, R, spc = Cx.sle_generator(I, q, max_rank, **kw)
    e0 = np.zeros(spc)
    e0[0] = 1.0
    lp = S.state(ss, 'L+', '1H')
    return ss, I, q, H, R, np.kron(e0, lp), np.kron(e0, lp)


def _one_spin_csa_nested():
    sys = {'magnet': 14.1, 'isotopes': ['1H']}
    inter = {'zeeman': {'eigs': [[-20.0, -5.0, 25.0]], 'euler': [[0.3, 0.9, -0.4]]}}
    ss = S.basis(S.create(sys, inter), {'formalism': 'sphten-liouv', 'approximation': 'none'})
    I, Q = S.hamiltonian(S.assume(ss, 'nmr'))
    return ss, I, Q


def test_clebsch_gordan_matches_the_library_and_is_orthogonal():
    from sympy.physics.wigner import clebsch_gordan as scg
    for L1 in range(4):
        for L2 in range(5):
            for L in range(abs(L1 - L2), L1 + L2 + 1):
                for M1 in range(-L1, L1 + 1):
                    for M2 in range(-L2, L2 + 1):
                        if abs(M1 + M2) <= L:
                            assert abs(Cx.clebsch_gordan(L, M1 + M2, L1, M1, L2, M2)
                                       - float(scg(L1, L2, L, M1, M2, M1 + M2))) < 1e-15
    # sum over M1 + M2 = M of <2 M1 3 M2|L M><2 M1 3 M2|L' M> = delta(L, L')
    for M in range(-5, 6):
        for L in range(max(1, abs(M)), 6):
            for Lp in range(max(1, abs(M)), 6):
                s = sum(Cx.clebsch_gordan(L, M, 2, m1, 3, M - m1) * Cx.clebsch_gordan(Lp, M, 2, m1, 3, M - m1)
                        for m1 in range(-2, 3))
                assert abs(s - (L == Lp)) < 1e-14


def test_sle_operators_represent_rotations_and_multiplication():
    """sle_operators.m: Lx, Ly, Lz satisfy [Lx, Ly] = i Lz and L^2 = L(L+1) on rank L; the
    multiplication operators give <e_000|D^2_ab D^2_cd|e_000> = int D^2_ab D^2_cd dOmega/8pi^2
    = (-1)^(c'-d') delta(a', -c') delta(b', -d') / 5 (a' = 2 - a, ...), and e_000 times D^r_ab
    is e_r,a',b' / sqrt(2r+1)."""
    Lx, Ly, Lz, D, sb = Cx.sle_operators(4, [1, 2])
    assert Lx.shape[0] == 5 * 9 * 11 // 3
    assert abs(Lx @ Ly - Ly @ Lx - 1j * Lz).max() < 1e-13
    L2 = (Lx @ Lx + Ly @ Ly + Lz @ Lz).toarray()
    assert np.abs(L2 - np.diag(sb[:, 0] * (sb[:, 0] + 1.0))).max() < 1e-12
    e0 = np.zeros(Lx.shape[0])
    e0[0] = 1.0
    for a in range(5):
        for b in range(5):
            v = D[2][(a, b)] @ e0
            assert abs(v[Cx.sle_index(2, 2 - a, 2 - b)] - 1 / math.sqrt(5)) < 1e-15
            assert abs(np.abs(v).sum() - 1 / math.sqrt(5)) < 1e-15
            for c in range(5):
                for d in range(5):
                    val = e0 @ (D[2][(a, b)] @ (D[2][(c, d)] @ e0))
                    ref = (-1.0) ** ((2 - c) - (2 - d)) / 5 if (a + c == 4 and b + d == 4) else 0.0
                    assert abs(val - ref) < 1e-15
    # multiplication operators commute where the truncation does not reach (ranks <= 4 - 2 - 2)
    C = (D[2][(0, 1)] @ D[2][(3, 2)] - D[2][(3, 2)] @ D[2][(0, 1)]).toarray()
    low = sb[:, 0] == 0
    assert np.abs(C[np.ix_(low, low)]).max() < 1e-15


def test_gridfree_static_converges_to_the_powder_average():
    """Without spinning or diffusion the SLE FID is the orientation average of the static FID:
    the error against a 20,000-point powder average falls spectrally with max_rank (short
    times: |omega t| <= 3)."""
    ss, I, Q = _one_spin_csa_nested()
    lp = S.state(ss, 'L+', '1H')
    dt, nst = 1e-6, 30
    ref = O.np_powder(I, S.q_list(Q), grids.zcwn(20000), lp, lp, dt, nst, reduce=False)
    errs = []
    for mr in (2, 4, 6):
        _, _, _, H, R, r0, c0 = sle_one_spin_case(mr)
        f = sle_dense_fid(H, R, r0, c0, dt, nst)
        assert abs(f[0] - np.vdot(lp, lp)) < 1e-14
        errs.append(O.rel_err(f, ref))
    assert errs[0] > 10 * errs[1] > 100 * errs[2] and errs[2] < 1e-4, errs


@pytest.mark.parametrize('bad, match', [
    ({'grid': 'zcwn_10'}, 'spherical integration'),
    ({'max_rank': -1}, 'max_rank'),
    ({'rate': 1000.0}, 'requires parameters.axis'),
    ({'tau_c': -1e-9}, 'tau_c must be positive'),
    ({'tau_c': np.ones((3, 3))}, 'symmetric positive definite'),
])
def test_gridfree_validates_its_inputs(bad, match):
    ss, _, _ = _one_spin_csa_nested()
    lp = S.state(ss, 'L+', '1H')
    params = {'rho0': lp, 'coil': lp, 'sweep': 1e5, 'npoints': 4, 'max_rank': 2}
    params.update(bad)
    with pytest.raises(ValueError, match=match):
        Cx.gridfree(ss, P.acquire, params, 'nmr')
