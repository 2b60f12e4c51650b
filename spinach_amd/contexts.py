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
"""
from __future__ import annotations

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
    if spin_system.get('rlx') is not None:
        raise NotImplementedError('relaxation in the Fokker-Planck MAS context is not supported on this path')
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
    return P._acquire_on_grid(ss, I_fp, q_fp, params)


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
