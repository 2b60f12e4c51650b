"""Host-side mirrors of two more Spinach contexts that run on the same GPU
path as powder():

crystal(spin_system, acquire, parameters, assumptions)
    kernel/contexts/crystal.m:31-110: one orientation (parameters.orientation),
    H = I + orientation(Q, orientation), (H+H')/2, then the pulse sequence:
    the library's powder call on a one-point grid of unit weight.

singlerot(spin_system, acquire, parameters, assumptions)
    kernel/contexts/singlerot.m:74-316: magic-angle spinning by the
    Fokker-Planck formalism in Liouville space. Per orientation the generator
    is G = blkdiag(H_1..H_N) + 1i*M on N = 2*max_rank+1 rotor phases
    phi_n = 2*pi*n/N (derivatives/fourdif.m), M = 2*pi*rate*kron(d_dphi, 1),
    H_n = I + sum_{r,k,m} D_comp,n(k,m) Q{r}{k,m} with
    D_comp,n = D_lab2rot * D_rotor(phi_n) * D_mol2rot(Omega) (singlerot.m:227-241).
    Because D_rotor is diagonal, D_comp,n(k,m) = sum_a D_lab2rot(k,a)
    e^{-i a' phi_n} D_mol2rot(a,m) (a' the projection of row a), so
        G(Omega) = I_FP + sum_{r,a,m} D^r(Omega)(a,m) Q_FP{r}{a,m},
        I_FP = kron(1_N, I) + 1i*2*pi*rate*kron(d_dphi, 1),
        Q_FP{r}{a,m} = sum_n E_nn (x) e^{-i a' phi_n} sum_k D_lab2rot(k,a) Q{r}{k,m}:
    an ordinary powder problem on matrices of dimension N * dim whose
    orientation coefficients are the plain Wigner matrices of the grid, which
    the library propagates as it does the static problem. rho0 and coil are
    projected into rotor space as singlerot.m:100-113 does.

gridfree(spin_system, acquire, parameters, assumptions)
    kernel/contexts/gridfree.m:65-232: the stochastic Liouville equation. The
    orientation is a coordinate: states live in the product of the Wigner
    functions e_LMN = sqrt(2L+1) D^L_MN(Omega), L <= max_rank (orthonormal
    under dOmega/8pi^2; kernel/cache/sle_operators.m), and the spin space. The
    generator is H = kron(1, I) + sum_{r,k,m} kron(D{r}{k,m}, Q{r}{k,m}) with
    D{r}{k,m} the multiplication by D^r_km in that basis (products of two
    Clebsch-Gordan coefficients), plus 2*pi*rate*kron(axis . L, 1) for
    spinning and R = -kron(Rd, 1) for rotational diffusion (tau_c);
    rho0 = kron(e_000, rho0), coil = kron(e_000, coil). One problem, no grid:
    the library runs it as a crystal of that generator (a one-point grid of
    unit weight, no anisotropic slots) with R as its orientation-independent
    relaxation superoperator.
"""
from __future__ import annotations

import functools
from fractions import Fraction

import math

import numpy as np
import scipy.sparse as sps

from . import _native
from . import powder as P
from . import spinsys as S


def fourdif(N):
    """First-derivative Fourier differentiation matrix on the periodic grid
    x_j = 2*pi*j/N (derivatives/fourdif.m with m = 1): the derivative of the
    trigonometric interpolant, D[j,k] = (-1)^(j-k)/2 * csc((j-k)h/2) for odd N
    and (-1)^(j-k)/2 * cot((j-k)h/2) for even N (j != k, h = 2*pi/N), zero
    diagonal. Returns (x, D)."""
    if not (isinstance(N, (int, np.integer)) and N >= 1):
        raise ValueError('N must be a positive integer')
    h = 2 * math.pi / N
    x = h * np.arange(N)
    D = np.zeros((N, N))
    for j in range(N):
        for k in range(N):
            if j != k:
                t = (j - k) * h / 2
                f = 1.0 / math.sin(t) if N % 2 else math.cos(t) / math.sin(t)
                D[j, k] = 0.5 * (-1) ** (j - k) * f
    return x, D


def _rotor_angles(axis):
    ax = np.asarray(axis, float)
    if ax.shape != (3,) or not np.all(np.isfinite(ax)) or np.linalg.norm(ax) == 0:
        raise ValueError('parameters.axis must be a row vector of three real numbers.')
    phi = math.atan2(ax[1], ax[0])                              # cart2sph azimuth
    theta = math.pi / 2 - math.atan2(ax[2], math.hypot(ax[0], ax[1]))
    return phi, theta


def fokker_planck(I, Qnested, axis, rate, max_rank):
    """I_FP and the 34-slot Q_FP list of the single-rotor Fokker-Planck
    generator (see the module docstring); dimension (2*max_rank+1) * dim."""
    N = 2 * int(max_rank) + 1
    phases, d_dphi = fourdif(N)
    rphi, rtheta = _rotor_angles(axis)
    dim = I.shape[0]
    eye_s = sps.identity(dim, format='csr', dtype=complex)
    I_fp = sps.kron(sps.identity(N, format='csr'), sps.csr_matrix(I, dtype=complex), format='csr')
    I_fp = I_fp + 1j * 2 * math.pi * float(rate) * sps.kron(sps.csr_matrix(d_dphi), eye_s, format='csr')
    q_fp = []
    for r in (1, 2):
        sz = 2 * r + 1
        Dl = S.wigner(r, rphi, rtheta, 0.0)  # D_lab2rot (singlerot.m:229)
        for a in range(sz):
            proj = r - a  # projection of row/column a (descending m)
            for m in range(sz):
                acc = None
                for k in range(sz):
                    q = Qnested[r - 1][k][m]
                    if q is None or q.nnz == 0 or Dl[k, a] == 0:
                        continue
                    term = Dl[k, a] * sps.csr_matrix(q, dtype=complex)
                    acc = term if acc is None else acc + term
                if acc is None:
                    q_fp.append(None)
                    continue
                phase = np.exp(-1j * proj * phases)  # D_rotor(phi_n) diagonal entry
                blk = sps.kron(sps.diags(phase), acc, format='csr')
                blk.eliminate_zeros()
                q_fp.append(blk if blk.nnz else None)
    return I_fp.tocsr(), q_fp


def _grid(parameters):
    g = parameters['grid']
    if g == 'single_crystal':
        return P._grid(g), True
    return P._grid(g), False


def singlerot(spin_system, pulse_sequence, parameters, assumptions):
    """Fokker-Planck MAS powder FID (kernel/contexts/singlerot.m) for
    pulse_sequence == acquire in the sphten-liouv formalism. parameters: rho0,
    coil, sweep, npoints, rate (Hz), axis (3), max_rank, grid (dict, name or
    'single_crystal'), optional sum_up, dead_time, offset, mode. Returns
    (answer, grid)."""
    if pulse_sequence is not P.acquire:
        raise NotImplementedError('only the acquire() sequence runs on this path')
    for f in ('rate', 'axis', 'max_rank'):
        if f not in parameters:
            raise ValueError(f'parameters.{f} subfield must be present.')
    mr = parameters['max_rank']
    if not (isinstance(mr, (int, np.integer)) and mr >= 0):
        raise ValueError('parameters.max_rank must be a positive real integer.')
    if not np.isscalar(parameters['rate']) or not np.isreal(parameters['rate']):
        raise ValueError('parameters.rate must be a real number.')
    P._grumble(parameters)
    ss = S.assume(spin_system, assumptions)
    I, Q = S.hamiltonian(ss)
    I = S.frqoffset(ss, I, parameters.get('offset'))
    I_fp, q_fp = fokker_planck(I, Q, parameters['axis'], parameters['rate'], mr)
    N = 2 * mr + 1
    grid, single = _grid(parameters)
    # rotor-space projections (singlerot.m:100-113)
    space = np.zeros(N)
    if single:
        space[0] = 1.0
    else:
        space[:] = 1.0 / N
    rho0 = np.kron(space, np.asarray(parameters['rho0'], complex))
    coil = np.kron(np.ones(N), np.asarray(parameters['coil'], complex))
    params = dict(parameters, rho0=rho0, coil=coil, grid=grid)
    # relaxation (singlerot.m:137, 176-177): R = relaxation(spin_system) without angles,
    # kron(1_N, R) in the rotor space
    R = S.relaxation_isotropic(ss)
    relax = None if R is None else {'R0': sps.kron(sps.identity(N, format='csr'), R, format='csr'), 'diag': None}
    return P._acquire_on_grid(ss, I_fp, q_fp, params, relax=relax)


def crystal(spin_system, pulse_sequence, parameters, assumptions):
    """Single-orientation FID (kernel/contexts/crystal.m) for pulse_sequence ==
    acquire: parameters.orientation = [alpha, beta, gamma] plus the powder()
    parameters other than grid. Returns the FID."""
    if pulse_sequence is not P.acquire:
        raise NotImplementedError('only the acquire() sequence runs on this path')
    o = np.asarray(parameters.get('orientation', None), dtype=float) if 'orientation' in parameters else None
    if o is None:
        raise ValueError('system orientation must be specified in parameters.orientation variable.')
    if o.shape != (3,) or not np.all(np.isfinite(o)):
        raise ValueError('parameters.orientation variable must be a row vector with three real finite numbers.')
    grid = {'alphas': o[:1].copy(), 'betas': o[1:2].copy(), 'gammas': o[2:3].copy(), 'weights': np.ones(1)}
    params = dict(parameters, grid=grid)
    params.pop('sum_up', None)
    P._grumble(params)
    ss = S.assume(spin_system, assumptions)
    I, Q = S.hamiltonian(ss)
    I = S.frqoffset(ss, I, parameters.get('offset'))
    fid, _ = P._acquire_on_grid(ss, I, S.q_list(Q), params)
    return fid


# ------------------------------------------------------------------ gridfree
@functools.lru_cache(maxsize=None)
def clebsch_gordan(L, M, L1, M1, L2, M2):
    """<L1 M1 L2 M2 | L M> for integer ranks by Racah's formula in exact rational arithmetic
    (kernel/utilities/clebsch_gordan.m computes the same coefficient); 0 for inadmissible
    indices."""
    if M != M1 + M2 or not (abs(L1 - L2) <= L <= L1 + L2) or abs(M) > L or abs(M1) > L1 or abs(M2) > L2:
        return 0.0
    f = math.factorial
    pre = Fraction((2 * L + 1) * f(L + L1 - L2) * f(L - L1 + L2) * f(L1 + L2 - L), f(L1 + L2 + L + 1))
    pre *= f(L + M) * f(L - M) * f(L1 - M1) * f(L1 + M1) * f(L2 - M2) * f(L2 + M2)
    acc = Fraction(0)
    for k in range(0, L1 + L2 - L + 1):
        dens = (L1 + L2 - L - k, L1 - M1 - k, L2 + M2 - k, L - L2 + M1 + k, L - L1 - M2 + k)
        if min(dens) < 0:
            continue
        term = Fraction(1, f(k) * f(dens[0]) * f(dens[1]) * f(dens[2]) * f(dens[3]) * f(dens[4]))
        acc += -term if k % 2 else term
    # sqrt(pre) * acc with the sign of acc: exact until the final square root
    return math.copysign(math.sqrt(float(pre * acc * acc)), float(acc))


def sle_index(L, M, N):
    """Position of e_LMN in the Wigner-function basis (0-based): ranks ascending, then M and N
    descending (sle_operators.m:87-95)."""
    return L * (4 * L * L + 6 * (L - M) + 5) // 3 - M - N


def sle_operators(max_rank, int_ranks):
    """kernel/cache/sle_operators.m: (Lx, Ly, Lz, D, space_basis) on the Wigner functions
    e_LMN, L = 0..max_rank. Lz = diag(M); L+ raises M with sqrt(L(L+1) - M(M+1)); D[r][(m, n)]
    (m, n = 0..2r, projections r - m, r - n) maps e_{L2 M2 N2} to
        sum_L <r M1 L2 M2|L M><r N1 L2 N2|L N> sqrt((2 L2 + 1)/(2 L + 1)) e_{L, M1+M2, N1+N2},
    L from |L2 - r| to min(L2 + r, max_rank): multiplication by D^r_{M1 N1} in the basis."""
    if not (isinstance(max_rank, (int, np.integer)) and max_rank >= 1):
        raise ValueError('max_rank must be a real positive integer.')
    space = [(L, M, N) for L in range(max_rank + 1) for M in range(L, -L - 1, -1) for N in range(L, -L - 1, -1)]
    dim = len(space)
    assert dim == (1 + max_rank) * (1 + 2 * max_rank) * (3 + 2 * max_rank) // 3
    rows, cols, vals = [], [], []
    for (L, M, N) in space:
        if M < L:
            rows.append(sle_index(L, M + 1, N))
            cols.append(sle_index(L, M, N))
            vals.append(math.sqrt(L * (L + 1) - M * (M + 1)))
    Lp = sps.csr_matrix((vals, (rows, cols)), shape=(dim, dim))
    Lx = ((Lp + Lp.T) / 2).astype(complex)
    Ly = (Lp - Lp.T) / 2j
    Lz = sps.diags(np.array([M for (_, M, _) in space], float)).astype(complex).tocsr()
    D = {}
    for r in int_ranks:
        D[r] = {}
        for m in range(2 * r + 1):
            for n in range(2 * r + 1):
                M1, N1 = r - m, r - n
                rows, cols, vals = [], [], []
                for (L2, M2, N2) in space:
                    for L in range(abs(L2 - r), min(L2 + r, max_rank) + 1):
                        Mx, Nx = M1 + M2, N1 + N2
                        if abs(Mx) > L or abs(Nx) > L:
                            continue
                        v = (clebsch_gordan(L, Mx, r, M1, L2, M2) * clebsch_gordan(L, Nx, r, N1, L2, N2)
                             * math.sqrt((2 * L2 + 1) / (2 * L + 1)))
                        if v != 0.0:
                            rows.append(sle_index(L, Mx, Nx))
                            cols.append(sle_index(L2, M2, N2))
                            vals.append(v)
                D[r][(m, n)] = sps.csr_matrix((vals, (rows, cols)), shape=(dim, dim))
    return Lx.tocsr(), Ly.tocsr(), Lz, D, np.array(space)


def sle_generator(I, Qnested, max_rank, rate=0.0, axis=None, tau_c=None):
    """gridfree.m:98-189: (H, R) of the stochastic Liouville equation, dimension
    spc_dim * dim with the space index outer (kron(space, spin))."""
    int_ranks = [r for r in (1, 2) if any(q is not None and q.nnz for row in Qnested[r - 1] for q in row)]
    Lx, Ly, Lz, D, _ = sle_operators(int(max_rank), int_ranks)
    spc = Lx.shape[0]
    dim = I.shape[0]
    eye_sp = sps.identity(spc, format='csr', dtype=complex)
    eye_s = sps.identity(dim, format='csr', dtype=complex)
    H = sps.kron(eye_sp, sps.csr_matrix(I, dtype=complex), format='csr')
    for r in int_ranks:
        for k in range(2 * r + 1):
            for m in range(2 * r + 1):
                q = Qnested[r - 1][k][m]
                if q is not None and q.nnz:
                    H = H + sps.kron(D[r][(k, m)], sps.csr_matrix(q, dtype=complex), format='csr')
    if rate:
        ax = np.asarray(axis, float)
        ax = ax / np.linalg.norm(ax)
        Hr = 2 * math.pi * float(rate) * (ax[0] * Lx + ax[1] * Ly + ax[2] * Lz)
        H = H + sps.kron(Hr, eye_s, format='csr')
    R = None
    if tau_c is not None:
        tc = np.asarray(tau_c, float)
        if tc.ndim == 0:
            Rd = (1.0 / (6.0 * float(tc))) * (Lx @ Lx + Ly @ Ly + Lz @ Lz)
        else:
            dt_ = np.linalg.inv(6.0 * tc)
            Ls = (Lx, Ly, Lz)
            Rd = sum(dt_[a, b] * (Ls[a] @ Ls[b]) for a in range(3) for b in range(3))
        R = -sps.kron(sps.csr_matrix(Rd), eye_s, format='csr')
    H.eliminate_zeros()
    return H.tocsr(), R, spc


def _gridfree_grumble(ss, parameters):
    """gridfree.m:262-330."""
    if 'rframes' in parameters:
        raise ValueError('numerical rotating frame transformation is not supported by SLE formalism.')
    if 'grid' in parameters:
        raise ValueError('spherical integration is a part of the SLE formalism, this parameter is unnecessary.')
    if ss.get('formalism') != 'sphten-liouv':
        raise ValueError('this function is only available in Liouville space.')
    if 'max_rank' not in parameters:
        raise ValueError('parameters.max_rank subfield must be present.')
    mr = parameters['max_rank']
    if not (isinstance(mr, (int, np.integer)) and mr >= 0):
        raise ValueError('parameters.max_rank must be a positive real integer.')
    rate = parameters.get('rate')
    if rate is not None and not (np.isscalar(rate) and np.isreal(rate) and np.isfinite(rate)):
        raise ValueError('parameters.rate must be a finite real number.')
    if 'axis' in parameters:
        ax = np.asarray(parameters['axis'], float)
        if ax.shape != (3,) or not np.all(np.isfinite(ax)) or np.linalg.norm(ax) == 0:
            raise ValueError('parameters.axis must be a row vector of three finite real numbers with a non-zero norm.')
    elif rate:
        raise ValueError('a non-zero parameters.rate requires parameters.axis to be specified.')
    if 'tau_c' in parameters:
        tc = np.asarray(parameters['tau_c'])
        if not np.isrealobj(tc):
            raise ValueError('parameters.tau_c must have real elements.')
        if tc.size not in (1, 9):
            raise ValueError('parameters.tau_c must be a scalar or a 3x3 matrix.')
        if tc.size == 1 and float(tc) <= 0:
            raise ValueError('parameters.tau_c must be positive.')
        if tc.size == 9:
            tc = tc.reshape(3, 3)
            if (np.linalg.norm(tc - tc.T) > 1e-12 * np.linalg.norm(tc)
                    or np.any(np.linalg.eigvalsh((tc + tc.T) / 2) <= 0)):
                raise ValueError('3x3 parameters.tau_c must be symmetric positive definite.')


def gridfree(spin_system, pulse_sequence, parameters, assumptions):
    """The stochastic-Liouville-equation FID (kernel/contexts/gridfree.m) for pulse_sequence ==
    acquire: parameters rho0, coil, sweep, npoints, max_rank, optional rate (Hz) and axis
    (spinning), tau_c (s, scalar or 3x3: rotational diffusion), offset, mode. No grid: the
    orientation is part of the state. Returns the FID."""
    if pulse_sequence is not P.acquire:
        raise NotImplementedError('only the acquire() sequence runs on this path')
    _gridfree_grumble(spin_system, parameters)
    mr = int(parameters['max_rank'])
    if mr < 1:
        raise ValueError('max_rank must be a real positive integer.')
    ss = S.assume(spin_system, assumptions)
    I, Q = S.hamiltonian(ss)
    I = S.frqoffset(ss, I, parameters.get('offset'))
    H, R, spc = sle_generator(I, Q, mr, parameters.get('rate', 0.0), parameters.get('axis'),
                              parameters.get('tau_c'))
    # spin relaxation (gridfree.m:95-96, 105): relaxation(spin_system) without angles, kron(1, R)
    Rs = S.relaxation_isotropic(ss)
    if Rs is not None:
        Rk = sps.kron(sps.identity(spc, format='csr'), Rs, format='csr')
        R = Rk if R is None else (R + Rk).tocsr()
    e0 = np.zeros(spc)
    e0[0] = 1.0
    grid = {'alphas': np.zeros(1), 'betas': np.zeros(1), 'gammas': np.zeros(1), 'weights': np.ones(1)}
    params = dict(parameters, grid=grid, rho0=np.kron(e0, np.asarray(parameters['rho0'], complex)),
                  coil=np.kron(e0, np.asarray(parameters['coil'], complex)))
    for k in ('sum_up', 'max_rank', 'rate', 'axis', 'tau_c'):
        params.pop(k, None)
    P._grumble(params)
    mode = params.get('mode', 'fp64')
    plan = _native.plan_powder(H, [None] * 34, grid, params['rho0'], params['coil'], 1.0 / params['sweep'],
                               int(params['npoints']) - 1, tols=ss['tols'], mode=mode,
                               relax=None if R is None else {'R0': R, 'diag': None})
    try:
        if params.get('dead_time'):
            plan.set_dead_time(params['dead_time'])
        plan.execute()
        return plan.result()
    finally:
        plan.close()
