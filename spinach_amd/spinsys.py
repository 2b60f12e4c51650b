"""Host-side spin-system setup for the powder path: create, assume('nmr'),
basis (sphten-liouv), hamiltonian -> (I, Q), state.

This is the part of Spinach that runs once per `powder` call on the host and
produces the inputs that cross the C-ABI boundary: the isotropic generator I,
the 34 anisotropic components Q{rank}{k,m}, the initial state and the coil.
It is restricted to spin-1/2 particles (1H, 13C, 15N, 19F, 31P, ...) and to
the sphten-liouv formalism, which is what the powder workloads use.

Reference files restated here (paths relative to the Spinach root):
  kernel/create.m:448-466          carrier frequencies basefrq = -gamma*B0
  kernel/create.m:740-830          Zeeman eigs/euler -> S*diag(eigs)*S', ppm -> rad/s
  kernel/create.m:1050-1070        scalar couplings 2*pi*J*eye(3)
  kernel/utilities/dipolar.m:150-177  point dipolar tensors (0.5 double-counting factor)
  kernel/assume.m:89-137           'nmr': secular Zeeman, secular homonuclear, 'zz' heteronuclear
  kernel/basis.m:188-240,440-466   IK-0 / none subgraph enumeration, sortrows of the state list
  kernel/cache/ist_product_table.m:100-125  structure coefficients of normalised ISTs
  kernel/operators/irr_sph_ten.m:54-75      single-spin irreducible spherical tensors
  kernel/operators/superop.m:36-186         product superoperators from structure tables
  kernel/utilities/human2opspec.m:130-150   'L+','L-','Lx','Ly','Lz' -> IST coefficients
  kernel/hamiltonian.m:44-792      descriptor tables and I / Q{r}{k,m} assembly
  kernel/state.m:94-153            rho = operator(...,'left') * unit
  kernel/conventions/transforms/mat2sphten.m:41-65, euler2dcm.m
  kernel/utilities/wigner.m:39-77  rank-1 and rank-2 Wigner D matrices
  kernel/orientation.m:29-59       H = sum_r sum_{k,m} D(k,m) Q{r}{k,m}, then (H+H')/2
"""
from __future__ import annotations

import itertools
import math
import os

import numpy as np
import scipy.sparse as sps

# Magnetogyric ratios, rad/(s*T), multiplicity 2 only (kernel/spin.m)
GAMMAS = {
    '1H': 2.6752218744e8,
    '13C': 6.728284e7,
    '15N': -2.71261804e7,
    '19F': 2.51662e8,
    '31P': 1.08394e8,
}
# Fundamental constants (kernel/utilities/tolerances.m:348-352)
HBAR = 6.62607015e-34 / (2 * math.pi)
MU0 = 1.25663706127e-6

DEFAULT_TOLS = dict(
    inter_cutoff=1e-10,   # tolerances.m:49
    liouv_zero=1e-10,     # tolerances.m:68
    prop_chop=1e-10,      # tolerances.m:106
    subs_drop=1e-10,      # tolerances.m:125
    zte_nsteps=32,        # tolerances.m:182
    zte_tol=1e-24,        # tolerances.m:201
    zte_maxden=0.5,       # tolerances.m:216
    krylov_tol=10000,     # tolerances.m:246
    merge_dim=1000,       # tolerances.m:268
    dense_matrix=0.15,    # tolerances.m:279
    small_matrix=200,     # tolerances.m:290
    prox_cutoff=100.0,    # tolerances.m:235
    rlx_integration=1e-4,  # tolerances.m:296-311
    rlx_zero=1e-10,       # tolerances.m:74-90
)


# ---------------------------------------------------------------- transforms
def euler2dcm(alp, bet, gam):
    """Active ZYZ direction cosine matrix R = Rz(alp) Ry(bet) Rz(gam)
    (kernel/conventions/transforms/euler2dcm.m:40-58)."""
    ca, sa = math.cos(alp), math.sin(alp)
    cb, sb = math.cos(bet), math.sin(bet)
    cg, sg = math.cos(gam), math.sin(gam)
    ra = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1.0]])
    rb = np.array([[cb, 0, sb], [0, 1.0, 0], [-sb, 0, cb]])
    rg = np.array([[cg, -sg, 0], [sg, cg, 0], [0, 0, 1.0]])
    return ra @ rb @ rg


def mat2sphten(m):
    """Rank 0/1/2 spherical tensor components of a 3x3 interaction matrix
    (kernel/conventions/transforms/mat2sphten.m:41-65). Components are
    ordered [+1,0,-1] and [+2,+1,0,-1,-2]."""
    m = np.asarray(m, dtype=float)
    r0 = np.trace(m) / 3
    r1 = np.zeros(3, complex)
    r1[0] = -0.5 * (m[2, 0] - m[0, 2] - 1j * (m[2, 1] - m[1, 2]))
    r1[1] = (1j / math.sqrt(2)) * (m[0, 1] - m[1, 0])
    r1[2] = -0.5 * (m[2, 0] - m[0, 2] + 1j * (m[2, 1] - m[1, 2]))
    r2 = np.zeros(5, complex)
    r2[0] = 0.5 * (m[0, 0] - m[1, 1] - 1j * (m[0, 1] + m[1, 0]))
    r2[1] = -0.5 * (m[0, 2] + m[2, 0] - 1j * (m[1, 2] + m[2, 1]))
    r2[2] = (1 / math.sqrt(6)) * (2 * m[2, 2] - m[0, 0] - m[1, 1])
    r2[3] = 0.5 * (m[0, 2] + m[2, 0] + 1j * (m[1, 2] + m[2, 1]))
    r2[4] = 0.5 * (m[0, 0] - m[1, 1] + 1j * (m[0, 1] + m[1, 0]))
    return r0, r1, r2


def wigner(l, alp, bet, gam):
    """Wigner D^l(alp,bet,gam) = exp(-i Lz alp) exp(-i Ly bet) exp(-i Lz gam)
    with rows and columns in descending m (kernel/utilities/wigner.m:1-3).
    Element (m',m) = exp(-i m' alp) d^l_{m'm}(bet) exp(-i m gam); the small-d
    matrix is evaluated from the Wigner sum formula."""
    ms = np.arange(l, -l - 1, -1)
    d = np.zeros((2 * l + 1, 2 * l + 1))
    c, s = math.cos(bet / 2), math.sin(bet / 2)
    f = math.factorial
    for i, mp in enumerate(ms):
        for j, m in enumerate(ms):
            pref = math.sqrt(f(l + mp) * f(l - mp) * f(l + m) * f(l - m))
            acc = 0.0
            for k in range(max(0, m - mp), min(l + m, l - mp) + 1):
                acc += ((-1) ** (mp - m + k)) * (c ** (2 * l + m - mp - 2 * k)) * \
                       (s ** (mp - m + 2 * k)) / (f(l + m - k) * f(k) * f(mp - m + k) * f(l - mp - k))
            d[i, j] = pref * acc
    ea = np.exp(-1j * ms * alp)
    eg = np.exp(-1j * ms * gam)
    return ea[:, None] * d * eg[None, :]


# ------------------------------------------------------ single-spin algebra
def _pauli_half():
    lp = np.array([[0, 1], [0, 0]], complex)
    lm = lp.T.copy()
    lz = np.diag([0.5, -0.5]).astype(complex)
    return lp, lm, lz


def irr_sph_ten_half():
    """Single-spin-1/2 ISTs [E, T(1,+1), T(1,0), T(1,-1)] built as in
    kernel/operators/irr_sph_ten.m:54-75: top state (-1)^k 2^(-k/2) L+^k,
    then Racah lowering T(k,q-1) = [L-, T(k,q)] / sqrt((k+q)(k-q+1))."""
    lp, lm, _ = _pauli_half()
    t = [np.eye(2, dtype=complex)]
    top = -(2 ** -0.5) * lp
    t.append(top)
    k = 1
    for n in range(2, 2 * k + 2):
        q = k - n + 2
        prev = t[-1]
        t.append((lm @ prev - prev @ lm) / math.sqrt((k + q) * (k - q + 1)))
    return t


def ist_product_tables_half():
    """Left/right structure coefficients of the spin-1/2 IST algebra:
    T{n}*T{m}/|T{m}| = sum_k left[n,m,k] T{k}/|T{k}|, and likewise with the
    product order reversed (kernel/cache/ist_product_table.m:100-125)."""
    t = irr_sph_ten_half()
    norms = [math.sqrt(np.trace(x.conj().T @ x).real) for x in t]
    left = np.zeros((4, 4, 4))
    right = np.zeros((4, 4, 4))
    for k in range(4):
        for m in range(4):
            for n in range(4):
                tk = t[k] / norms[k]
                tn = t[n] / norms[n]
                tm = t[m] / norms[m]
                left[n, m, k] = (norms[n] * np.trace(tk.conj().T @ (tn @ tm))).real
                right[n, m, k] = (norms[n] * np.trace(tk.conj().T @ (tm @ tn))).real
    left[np.abs(left) < 1e-14] = 0.0
    right[np.abs(right) < 1e-14] = 0.0
    return left, right


_LPST, _RPST = ist_product_tables_half()

# human2opspec for spin-1/2 (kernel/utilities/human2opspec.m:130-150):
# operator name -> list of (IST linear index, coefficient)
_HUMAN = {
    'E': [(0, 1.0)],
    'L+': [(1, -math.sqrt(2))],
    'L-': [(3, math.sqrt(2))],
    'Lx': [(1, -math.sqrt(2) / 2), (3, math.sqrt(2) / 2)],
    'Ly': [(1, -math.sqrt(2) / 2j), (3, -math.sqrt(2) / 2j)],
    'Lz': [(2, 1.0)],
}


# ------------------------------------------------------------------ create
def create(sys, inter):
    """Build the spin-system structure (kernel/create.m). Supported fields:
    sys.magnet, sys.isotopes, sys.tols, sys.enable ('krylov', 'gpu', 'ham_cache'); inter.zeeman.{eigs,euler,matrix,scalar} (ppm),
    inter.coupling.{scalar (Hz), matrix (Hz)}, inter.coordinates (Angstrom).
    Every field uses 0-based spin numbers; couplings are dicts {(n,k): value}.
    """
    isotopes = list(sys['isotopes'])
    nspins = len(isotopes)
    for iso in isotopes:
        if iso not in GAMMAS:
            raise ValueError(f'isotope {iso} is not a supported spin-1/2 nucleus')
    tols = dict(DEFAULT_TOLS)
    tols.update(sys.get('tols', {}))
    magnet = float(sys['magnet'])
    gammas = np.array([GAMMAS[i] for i in isotopes])
    basefrqs = -gammas * magnet                       # create.m:466

    zee = inter.get('zeeman', {})
    zmat = [np.zeros((3, 3)) for _ in range(nspins)]
    have_zeeman = bool(zee)
    for n, eigs in enumerate(zee.get('eigs', [None] * nspins)):
        if eigs is None or max(abs(np.asarray(eigs, float))) == 0:
            continue
        eul = zee.get('euler', [None] * nspins)[n]
        s = np.eye(3) if eul is None else euler2dcm(*eul)
        comp = s @ np.diag(np.asarray(eigs, float)) @ s.T
        zmat[n] = zmat[n] + (comp + comp.T) / 2      # create.m:746-756
    for n, m in enumerate(zee.get('matrix', [None] * nspins)):
        if m is not None:
            zmat[n] = zmat[n] + np.asarray(m, float)
    for n, v in enumerate(zee.get('scalar', [None] * nspins)):
        if v is not None and v != 0:
            zmat[n] = zmat[n] + np.eye(3) * float(v)
    # ppm -> rad/s for nuclei (create.m:808); magnet frequencies if no Zeeman
    zmat = [(np.eye(3) + 1e-6 * zmat[n]) * basefrqs[n] if have_zeeman
            else np.eye(3) * basefrqs[n] for n in range(nspins)]

    coup = [[None] * nspins for _ in range(nspins)]

    def _add(n, k, m):
        coup[n][k] = m if coup[n][k] is None else coup[n][k] + m

    ic = inter.get('coupling', {})
    for (n, k), m in ic.get('matrix', {}).items():
        if np.linalg.norm(np.asarray(m, float), 2) > tols['inter_cutoff']:
            _add(n, k, 2 * math.pi * np.asarray(m, float))
    for (n, k), j in ic.get('scalar', {}).items():
        if abs(j) > tols['inter_cutoff']:
            _add(n, k, 2 * math.pi * np.eye(3) * float(j))   # create.m:1063-1069

    coords = inter.get('coordinates')
    if coords is not None:                                # utilities/dipolar.m
        coords = [None if c is None else np.asarray(c, float) for c in coords]
        for n in range(nspins):
            for k in range(nspins):
                if n == k or coords[n] is None or coords[k] is None:
                    continue
                dv = coords[k] - coords[n]
                dist = float(np.linalg.norm(dv))
                if dist >= tols['prox_cutoff']:
                    continue
                if dist < 0.5:
                    raise ValueError(f'collision detected between spin {n} and spin {k}')
                o = dv / dist
                a = 0.5 * gammas[n] * gammas[k] * HBAR * MU0 / (4 * math.pi * (dist * 1e-10) ** 3)
                d = a * (np.eye(3) - 3 * np.outer(o, o))
                d = d - np.eye(3) * np.trace(d) / 3
                _add(n, k, d)

    rlx = _relaxation_spec(inter, nspins)
    # chemical species (create.m: inter.chem.parts; one species of every spin by default)
    parts = [list(range(nspins))]
    if inter.get('chem', {}).get('parts') is not None:
        parts = [sorted(int(x) for x in pt) for pt in inter['chem']['parts']]
        flat = sorted(x for pt in parts for x in pt)
        if flat != list(range(nspins)):
            raise ValueError('inter.chem.parts must assign every spin to exactly one species.')
    if rlx is not None and rlx['tau_c'] is not None and len(rlx['tau_c']) != len(parts):
        raise ValueError('inter.tau_c must have one entry per chemical species.')

    enable = list(sys.get('enable', []))  # sys.enable switches (create.m:1556); 'krylov' is read by powder()
    for e in enable:
        if e not in ('krylov', 'gpu', 'ham_cache'):
            raise ValueError(f'sys.enable option {e!r} is not supported on this path')
    return dict(isotopes=isotopes, nspins=nspins, magnet=magnet, gammas=gammas, enable=enable,
                basefrqs=basefrqs, zeeman=zmat, coupling=coup, tols=tols, rlx=rlx, parts=parts,
                zeeman_strength=['full'] * nspins,
                coupling_strength=[['strong'] * nspins for _ in range(nspins)])


# -------------------------------------------------------------- relaxation
_RLX_THEORIES = ('t1_t2', 'damp', 'redfield')
_RLX_KEEP = ('diagonal', 'kite', 'secular', 'labframe')


def _rate_spec(v, what):
    """A relaxation rate: a real scalar, or a real 3x3 tensor whose rate at an
    orientation is ort*T*ort' (rlx_t1_t2.m, relaxation.m 'damp')."""
    a = np.asarray(v, dtype=complex)
    if a.shape == () and np.isfinite(a) and a.imag == 0:
        return float(a.real)
    if a.shape == (3, 3) and np.all(np.isfinite(a)) and np.all(a.imag == 0):
        return a.real.copy()
    raise ValueError(f'{what} must be a real scalar or a real 3x3 matrix.')


def _relaxation_spec(inter, nspins):
    """inter.relaxation, inter.r1_rates, inter.r2_rates, inter.damp_rate,
    inter.rlx_keep, inter.equilibrium (kernel/create.m:1134-1300). Returns None
    without relaxation theories."""
    theories = list(inter.get('relaxation', []))
    if not theories:
        return None
    for t in theories:
        if t not in _RLX_THEORIES:
            raise NotImplementedError(f'relaxation theory {t!r} is not supported on this path')
    keep = inter.get('rlx_keep')
    if keep not in _RLX_KEEP:
        raise ValueError('inter.rlx_keep must be one of ' + ', '.join(_RLX_KEEP) + '.')
    eq = inter.get('equilibrium', 'zero')
    if eq != 'zero':
        raise NotImplementedError("only inter.equilibrium = 'zero' (relaxation to the unit state) is supported")
    dfs = inter.get('rlx_dfs', 'ignore')                      # create.m:1181-1185
    if dfs not in ('keep', 'ignore'):
        raise ValueError('invalid value of the inter.rlx_dfs parameter.')
    spec = dict(theories=theories, keep=keep, equilibrium=eq, r1=None, r2=None, damp=None, tau_c=None, dfs=dfs)
    if 't1_t2' in theories:
        for f in ('r1_rates', 'r2_rates'):
            if f not in inter or len(inter[f]) != nspins:
                raise ValueError(f'inter.{f} must have one entry per spin for the t1_t2 theory.')
        spec['r1'] = [_rate_spec(v, f'R1 rate of spin {n}') for n, v in enumerate(inter['r1_rates'])]
        spec['r2'] = [_rate_spec(v, f'R2 rate of spin {n}') for n, v in enumerate(inter['r2_rates'])]
    if 'damp' in theories:
        if 'damp_rate' not in inter:
            raise ValueError('inter.damp_rate must be specified for the damp theory.')
        spec['damp'] = _rate_spec(inter['damp_rate'], 'inter.damp_rate')
    if 'redfield' in theories:
        if 'tau_c' not in inter:
            raise ValueError('Redfield theory requires rotational correlation time(s) in inter.tau_c.')
        spec['tau_c'] = []
        for t in list(inter['tau_c']):
            tc = np.atleast_1d(np.asarray(t, float))
            if tc.ndim != 1 or tc.size not in (1, 2, 3) or not np.all(np.isfinite(tc)):
                raise ValueError('inter.tau_c entries must hold one, two or three correlation times.')
            if not np.any(tc):
                raise ValueError('zero rotational correlation times are not supported.')   # relaxation.m:102-107
            spec['tau_c'].append(tc)
    return spec


def _ort(euler):
    """ort = [0 0 1]*euler2dcm(alpha, beta, gamma) (rlx_t1_t2.m, relaxation.m)."""
    return np.array([0.0, 0.0, 1.0]) @ euler2dcm(*euler)


def _rate_at(spec, euler):
    if np.isscalar(spec):
        return spec
    if euler is None:
        raise ValueError('Euler angles must be specified with anisotropic relaxation rates.')
    o = _ort(euler)
    return float(o @ spec @ o)


def relaxation(ss, euler=None):
    """The relaxation superoperator R at one orientation (kernel/relaxation.m
    for the t1_t2 and damp theories, sphten-liouv; zero without theories), as
    a sparse diagonal matrix: state i decays at the sum over its non-unit spins
    of R1 (longitudinal, T_{1,0}) or R2 (transverse, T_{1,+-1}) of that spin
    (rlx_t1_t2.m), plus the damping rate on every state but the unit state."""
    states = ss['basis']
    if states is None:
        raise NotImplementedError('relaxation is only available in the sphten-liouv formalism')
    dim = len(states)
    rates = np.zeros(dim)
    spec = ss.get('rlx')
    if spec is not None:
        if spec['r1'] is not None:
            r1 = np.array([_rate_at(v, euler) for v in spec['r1']])
            r2 = np.array([_rate_at(v, euler) for v in spec['r2']])
            lon = states == 2
            tra = (states == 1) | (states == 3)
            rates += lon.astype(float) @ r1 + tra.astype(float) @ r2
        if spec['damp'] is not None:
            unit = np.all(states == 0, axis=1)
            d = spec['damp']
            if euler is None and not np.isscalar(d):
                d = float(np.trace(d)) / 3          # relaxation.m:652-655: the isotropic damping rate
            rates += np.where(unit, 0.0, _rate_at(d, euler))
    return sps.diags(-rates, 0, shape=(dim, dim), format='csc')


def relaxation_isotropic(ss):
    """relaxation(spin_system) without Euler angles (kernel/relaxation.m as singlerot.m:137
    calls it): scalar T1/T2 rates (rate tensors are an error, as in rlx_t1_t2.m), the mean of a
    damping tensor's diagonal, and the Redfield term. Returns a sparse matrix, or None without
    relaxation theories."""
    spec = ss.get('rlx')
    if spec is None:
        return None
    R = sps.csr_matrix(relaxation(ss, None), dtype=complex)
    if spec['tau_c'] is not None:
        R = R + _redfield_R0(ss)
    return R.tocsr()


def _clean_up_dense(A, tol):
    """clean_up.m:48-51 on a dense array: A <- tol*round(A/tol), real and imaginary parts
    rounded half away from zero (MATLAB's round)."""
    def rnd(x):
        return np.sign(x) * np.floor(np.abs(x) + 0.5)
    A = np.asarray(A) / tol
    if np.iscomplexobj(A):
        return tol * (rnd(A.real) + 1j * rnd(A.imag))
    return tol * rnd(A)


def corrfun(tau_c, n, k, m, p, q):
    """kernel/utilities/corrfun.m:34-118 for one chemical species: (weights, rates) of the
    rotational correlation function of the rank-n Wigner functions (k, m) and (p, q) (0-based
    rows, projections n - index) as a sum of exponentials. One correlation time: isotropic
    diffusion, D = 1/(6 tau_c), weight delta_kp delta_mq / (2n+1), rate -n(n+1)D; two (axial,
    equatorial): rate -(n(n+1) D_eq + m'^2 (D_ax - D_eq)) with m' = n - m; three (rhombic, n = 2
    only): five exponentials with the weights (1/5) delta_kp h(j,m) h(j,q)."""
    tc = np.atleast_1d(np.asarray(tau_c, float))
    if tc.size == 1:
        D = 1.0 / (6.0 * tc[0])
        return [float(k == p and m == q) / (2 * n + 1)], [-n * (n + 1) * D]
    if tc.size == 2:
        Dax, Deq = 1.0 / (6.0 * tc[0]), 1.0 / (6.0 * tc[1])
        return ([float(k == p and m == q) / (2 * n + 1)],
                [-(n * (n + 1) * Deq + (n - m) ** 2 * (Dax - Deq))])
    Dxx, Dyy, Dzz = 1.0 / (6.0 * tc)
    mean = (Dxx + Dyy + Dzz) / 3
    if min(abs(Dxx - Dyy), abs(Dyy - Dzz), abs(Dzz - Dxx)) < 1e-6 * mean:
        raise ValueError('the three rotational correlation times must be different.')
    if n != 2:
        raise ValueError('rhombic rotational diffusion only implemented for n=2.')
    delta = math.sqrt(Dxx ** 2 + Dyy ** 2 + Dzz ** 2 - Dxx * Dyy - Dxx * Dzz - Dyy * Dzz)
    rates = [-(4 * Dxx + Dyy + Dzz), -(Dxx + 4 * Dyy + Dzz), -(Dxx + Dyy + 4 * Dzz),
             -(2 * Dxx + 2 * Dyy + 2 * Dzz - 2 * delta), -(2 * Dxx + 2 * Dyy + 2 * Dzz + 2 * delta)]
    lp = math.sqrt(2 / 3) * (Dxx + Dyy - 2 * Dzz + 2 * delta) / (Dxx - Dyy)
    lm = math.sqrt(2 / 3) * (Dxx + Dyy - 2 * Dzz - 2 * delta) / (Dxx - Dyy)
    r2 = 1 / math.sqrt(2)
    h = np.zeros((5, 5))
    h[0, 1], h[0, 3], h[1, 1], h[1, 3], h[2, 4], h[2, 0] = r2, r2, -r2, r2, r2, -r2
    h[3, 0], h[3, 2], h[3, 4] = np.array([1, lm, 1]) / math.sqrt(2 + lm ** 2)
    h[4, 0], h[4, 2], h[4, 4] = np.array([1, lp, 1]) / math.sqrt(2 + lp ** 2)
    return [float(k == p) * h[j, m] * h[j, q] / 5 for j in range(5)], rates


def _expmint(A, B, C, T):
    """kernel/utilities/expmint.m:41-63: int_0^T exp(-iAt) B exp(iCt) dt by the block
    exponential of aux = [[-A, iB], [0, -C]]: with E = exp(-i aux T), the integral is
    E11' * E12 (Van Loan), cleaned at prop_chop."""
    d = A.shape[0]
    aux = np.zeros((2 * d, 2 * d), complex)
    aux[:d, :d] = -A
    aux[:d, d:] = 1j * B
    aux[d:, d:] = -C
    import scipy.linalg as sla
    E = sla.expm(-1j * T * aux)
    return E[:d, :d].conj().T @ E[:d, d:]


def redfield_superoperator(ss):
    """The Redfield term of kernel/relaxation.m:96-137 (kernel/includes/redfield_integral_serial.m)
    over the chemical species, before the dynamic-shift and retention policies: over the
    lab-frame Hamiltonian's rotational components Q{n}{k,m} and each exponential (w, rate) of
    corrfun, R -= w * Q{n}{k,m} * int_0^T exp(-iBt) Q{n}{p,q}' exp(i(B - i rate)t) dt with
    B = clean_up(L0, rlx_integration/T), the on-shell kernel at zero shift, T =
    -1.5 log(1/rlx_integration)/rate, the block of states outside the species zeroed; every term
    cleaned at 1e-2*rlx_zero. Dense FP64 (the integrals by scipy's expm)."""
    spec = ss['rlx']
    tols = ss['tols']
    L0, Q = hamiltonian(assume(ss, 'labframe'))
    L0 = L0.toarray()
    dim = L0.shape[0]
    R = np.zeros((dim, dim), complex)
    parts = ss.get('parts') or [list(range(ss['nspins']))]
    for s_, part in enumerate(parts):
        # the species' states: a non-unit state on one of its spins (corrfun.m:52-54)
        outside = ~(ss['basis'][:, part].sum(axis=1) > 0)
        R += _redfield_species(L0, Q, spec['tau_c'][s_], np.ix_(outside, outside), tols)
    return R


def _redfield_species(L0, Q, tau, blk, tols):
    """One species' share of the Redfield sum (corrfun's weights and rates for its correlation
    time, the block of states outside the species zeroed in every operand)."""
    dim = L0.shape[0]
    R = np.zeros((dim, dim), complex)
    for n in (1, 2):
        sz = 2 * n + 1
        comps = Q[n - 1]
        for k in range(sz):
            for m in range(sz):
                A = comps[k][m]
                if A is None or abs(A).sum() == 0:
                    continue
                A = A.toarray()
                for p in range(sz):
                    for q in range(sz):
                        Cq = comps[p][q]
                        if Cq is None or abs(Cq).sum() == 0:
                            continue
                        weights, rates = corrfun(tau, n, k, m, p, q)
                        for w, rate in zip(weights, rates):
                            if w == 0:
                                continue
                            T = -1.5 * (1.0 / rate) * math.log(1.0 / tols['rlx_integration'])
                            B = _clean_up_dense(L0, tols['rlx_integration'] / abs(T))
                            C = Cq.toarray().conj().T
                            Dm = B - 1j * rate * np.eye(dim)
                            Ab = A.copy()
                            for X in (Ab, B, C, Dm):
                                X[blk] = 0
                            Ri = -w * Ab @ _expmint(B, C, Dm, T)
                            R += _clean_up_dense(Ri, 1e-2 * tols['rlx_zero'])
    return R


def _redfield_R0(ss):
    """The orientation-independent part the library takes as R0: the Redfield term with the
    dynamic-shift policy (relaxation.m:533-552), the retention policy (relaxation.m:556-634) and
    the validity check against the correlation times (relaxation.m:126-134)."""
    spec = ss['rlx']
    R = redfield_superoperator(ss)
    rmax = np.abs(np.diag(R)).max()
    for tc in spec['tau_c']:
        # MATLAB's `if` on an array holds when every element does: the shortest time decides
        if rmax > 0 and (1.0 / rmax) < np.min(tc):
            raise ValueError('T1,2>>tau_c validity condition violation in Redfield theory')
    if spec['dfs'] == 'ignore':
        R = R.real.astype(complex)
    states = ss['basis']
    M = np.where(states == 1, 1, np.where(states == 3, -1, 0))   # projection of each spin's T(1,m)
    keep = spec['keep']
    if keep == 'diagonal':
        R = np.diag(np.diag(R))
        unit = np.all(states == 0, axis=1)
        R[unit, unit] = 0.0
    elif keep == 'kite':
        lon = np.all(M == 0, axis=1)
        mask = np.outer(lon, lon) | np.eye(len(R), dtype=bool)
        R = np.where(mask, R, 0.0)
    elif keep == 'secular':
        f = M @ np.asarray(ss['basefrqs'], float)
        R = np.where(f[:, None] == f[None, :], R, 0.0)
    return sps.csr_matrix(R)


def relaxation_terms(ss):
    """The library's form of relaxation(ss, euler) for every orientation
    (sp_relax): R(Omega) = diag(g0 + sum_p q_p(Omega) g_p) with the monomials
    q = (ox^2, oy^2, oz^2, 2 ox oy, 2 ox oz, 2 oy oz) of ort; a tensor rate T
    contributes (Txx, Tyy, Tzz, (Txy+Tyx)/2, (Txz+Tzx)/2, (Tyz+Tzy)/2) to
    g1..g6, a scalar rate to g0. Returns None without relaxation theories."""
    spec = ss.get('rlx')
    if spec is None:
        return None
    states = ss['basis']
    if states is None:
        raise NotImplementedError('relaxation is only available in the sphten-liouv formalism')
    dim = len(states)
    g = np.zeros((7, dim))

    def add(mask, rate):
        if np.isscalar(rate):
            g[0, mask] -= rate
        else:
            t = rate
            comps = (t[0, 0], t[1, 1], t[2, 2], 0.5 * (t[0, 1] + t[1, 0]), 0.5 * (t[0, 2] + t[2, 0]),
                     0.5 * (t[1, 2] + t[2, 1]))
            for p, cval in enumerate(comps):
                g[1 + p, mask] -= cval

    if spec['r1'] is not None:
        for n in range(ss['nspins']):
            add(states[:, n] == 2, spec['r1'][n])
            add((states[:, n] == 1) | (states[:, n] == 3), spec['r2'][n])
    if spec['damp'] is not None:
        add(~np.all(states == 0, axis=1), spec['damp'])
    R0 = _redfield_R0(ss) if spec['tau_c'] is not None else None
    return {'R0': R0, 'diag': g}


def assume(ss, assumptions):
    """Interaction strengths under the 'nmr' assumption set (kernel/assume.m:89-137), or
    'labframe' (assume.m:55-66: full Zeeman interactions including the carrier, strong
    couplings; the Hamiltonian Redfield theory integrates over). Returns a modified copy."""
    if assumptions == 'labframe':
        ss = dict(ss)
        n = ss['nspins']
        ss['zeeman_strength'] = ['full'] * n
        ss['coupling_strength'] = [['strong'] * n for _ in range(n)]
        return ss
    if assumptions != 'nmr':
        raise NotImplementedError("only the 'nmr' and 'labframe' assumption sets are supported on this path")
    ss = dict(ss)
    n = ss['nspins']
    ss['zeeman_strength'] = ['secular'] * n
    iso = ss['isotopes']
    ss['coupling_strength'] = [['secular' if iso[a] == iso[b] else 'zz' for b in range(n)]
                               for a in range(n)]
    return ss


# ------------------------------------------------------------------- basis
def basis(ss, bas):
    """sphten-liouv basis (kernel/basis.m). approximation 'none' gives the
    complete 4^n basis; 'IK-0' with `level` keeps every state of correlation
    order <= level (all spins form one chemical part). Optional
    `projections` keeps only the listed total coherence orders plus the unit
    state. Rows are sorted lexicographically (basis.m:459-465)."""
    form = bas.get('formalism', 'sphten-liouv')
    if form == 'zeeman-hilb':
        ss = dict(ss)
        ss['formalism'] = 'zeeman-hilb'
        ss['basis'] = None
        ss['basis_keys'] = None
        return ss
    if form != 'sphten-liouv':
        raise NotImplementedError(f'formalism {form} is not supported')
    n = ss['nspins']
    appr = bas.get('approximation', 'none')
    if appr == 'none':
        states = np.array(list(itertools.product(range(4), repeat=n)), dtype=np.int8)
    elif appr == 'IK-0':
        level = int(bas['level'])
        if level > n:
            raise ValueError('IK-0 level exceeds the number of spins')
        rows = []
        for order in range(level + 1):
            for spins in itertools.combinations(range(n), order):
                for vals in itertools.product((1, 2, 3), repeat=order):
                    r = [0] * n
                    for s, v in zip(spins, vals):
                        r[s] = v
                    rows.append(r)
        states = np.array(rows, dtype=np.int8).reshape(-1, n)
    else:
        raise NotImplementedError(f'basis approximation {appr} is not supported')
    if 'projections' in bas:
        m = np.where(states == 1, 1, np.where(states == 3, -1, 0)).sum(axis=1)
        keep = np.zeros(len(states), bool)
        keep |= np.all(states == 0, axis=1)
        for p in np.atleast_1d(bas['projections']):
            keep |= (m == p)
        states = states[keep]
    keys = _keys(states)
    order = np.argsort(keys, kind='stable')
    states = states[order]
    keys = keys[order]
    ss = dict(ss)
    ss['basis'] = states
    ss['basis_keys'] = keys
    ss['formalism'] = 'sphten-liouv'
    return ss


def _keys(states):
    n = states.shape[1]
    w = (4 ** np.arange(n - 1, -1, -1)).astype(np.int64)
    return states.astype(np.int64) @ w


# --------------------------------------------------------------- operators
def _side_superop(ss, opspec, side):
    """Left or right product superoperator of one IST product (opspec maps
    spin -> IST index) in the basis; returns COO triples (row=result state,
    col=source state) as in kernel/operators/superop.m:62-170."""
    states = ss['basis']
    keys = ss['basis_keys']
    n = ss['nspins']
    dim = len(keys)
    tab = _LPST if side == 'left' else _RPST
    active = sorted(opspec)
    if not active:
        idx = np.arange(dim)
        return idx, idx, np.ones(dim)
    rows, cols, vals = [], [], []
    src = states[:, active].astype(np.int64)       # input digits on active spins
    pts = [tab[opspec[a]] for a in active]          # [m, k] per active spin
    place = [4 ** (n - 1 - a) for a in active]
    for ks in itertools.product(range(4), repeat=len(active)):
        coeff = np.ones(dim)
        newkey = keys.copy()
        for i, k in enumerate(ks):
            coeff = coeff * pts[i][src[:, i], k]
            newkey = newkey + (k - src[:, i]) * place[i]
        nz = coeff != 0
        if not nz.any():
            continue
        pos = np.searchsorted(keys, newkey[nz])
        pos_c = np.minimum(pos, dim - 1)
        found = keys[pos_c] == newkey[nz]
        colidx = np.nonzero(nz)[0][found]
        rows.append(pos_c[found])
        cols.append(colidx)
        vals.append(coeff[nz][found])
    if not rows:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0)
    return np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)


def operator(ss, operators, spins, operator_type='comm'):
    """Superoperator of a (product) operator (kernel/operator.m:71-210).
    `operators` is a name or list of names, `spins` an isotope name, 'all',
    an int, or a list of ints (0-based) matching `operators`. Types: 'comm',
    'left', 'right', 'acomm'. Returns scipy CSR complex128. In the
    zeeman-hilb formalism the Hilbert-space operator itself is returned and
    the type is ignored (kernel/operator.m:125-160)."""
    terms = _human2opspec(ss, operators, spins)
    if ss.get('formalism') == 'zeeman-hilb':
        return _hilbert_operator(ss, terms)
    dim = len(ss['basis_keys'])
    r_all, c_all, v_all = [], [], []
    for opspec, coeff in terms:
        if operator_type in ('left', 'right'):
            sides = [(operator_type, 1.0)]
        elif operator_type == 'comm':
            sides = [('left', 1.0), ('right', -1.0)]
        elif operator_type == 'acomm':
            sides = [('left', 1.0), ('right', 1.0)]
        else:
            raise ValueError('unknown operator type')
        for side, sgn in sides:
            r, c, v = _side_superop(ss, opspec, side)
            r_all.append(r)
            c_all.append(c)
            v_all.append(sgn * coeff * v.astype(complex))
    m = sps.coo_matrix((np.concatenate(v_all), (np.concatenate(r_all), np.concatenate(c_all))),
                       shape=(dim, dim)).tocsr()
    m.sum_duplicates()
    m.eliminate_zeros()
    return m


def _hilbert_operator(ss, terms):
    """Sum of coeff * kron(T_{s1}, T_{s2}, ...) over the terms, with the
    un-normalised single-spin ISTs of kernel/operators/irr_sph_ten.m and the
    spin ordering of kron (spin 0 leftmost)."""
    t = irr_sph_ten_half()
    n = ss['nspins']
    out = None
    for spec, coeff in terms:
        m = sps.csr_matrix(np.array([[coeff]], complex))
        for k in range(n):
            m = sps.kron(m, sps.csr_matrix(t[spec.get(k, 0)]), format='csr')
        out = m if out is None else out + m
    out = out.tocsr()
    out.eliminate_zeros()
    return out


def _human2opspec(ss, operators, spins):
    """Expand a human operator specification into (opspec, coeff) terms
    (kernel/utilities/human2opspec.m:55-150)."""
    if isinstance(operators, str):
        if isinstance(spins, str):
            if spins == 'all':
                nums = list(range(ss['nspins']))
            else:
                nums = [i for i, s in enumerate(ss['isotopes']) if s == spins]
            if not nums:
                raise ValueError('no such spins in the system.')
        elif isinstance(spins, (int, np.integer)):
            nums = [int(spins)]
        else:
            nums = [int(s) for s in spins]
        out = []
        for s in nums:
            out += _human2opspec(ss, [operators], [s])
        return out
    terms = [({}, 1.0 + 0j)]
    for name, s in zip(operators, spins):
        if name not in _HUMAN:
            raise ValueError(f'unrecognized operator {name}')
        new = []
        for spec, c in terms:
            for idx, cc in _HUMAN[name]:
                d = dict(spec)
                if idx != 0:
                    d[int(s)] = idx
                new.append((d, c * cc))
        terms = new
    return terms


def state(ss, states, spins):
    """Liouville-space state vector: left-side product superoperator applied
    to the unit state e_1 (kernel/state.m:94-153). Returns complex ndarray."""
    if ss.get('formalism') == 'zeeman-hilb':
        return operator(ss, states, spins).toarray()   # state.m:208-211
    dim = len(ss['basis_keys'])
    unit = np.zeros(dim, complex)
    unit[0] = 1.0
    return operator(ss, states, spins, 'left') @ unit


# ------------------------------------------------------------- hamiltonian
def _clean_up(m, tol):
    """tol*round(A/tol) quantisation with zero dropping (kernel/utilities/clean_up.m:70)."""
    m = m.tocsr().copy()
    d = m.data
    m.data = tol * (np.round((1 / tol) * d.real) + 1j * np.round((1 / tol) * d.imag))
    m.eliminate_zeros()
    return m


_HAM_CACHE = {}


def _ham_key(ss):
    """Digest of everything I and Q depend on: isotopes, magnet, carrier
    frequencies, interaction tensors, interaction strengths (the 'nmr'
    assumptions), the basis and the tolerances that drop terms (the cache key
    of hamiltonian.m:600-611 hashes the descriptor, isotopes and basis)."""
    import hashlib
    h = hashlib.sha256()
    h.update(repr((ss['isotopes'], ss['magnet'], ss.get('formalism'), ss['zeeman_strength'],
                   ss['coupling_strength'], ss['tols']['liouv_zero'], ss['tols']['inter_cutoff'])).encode())
    h.update(np.asarray(ss['basefrqs'], float).tobytes())
    for z in ss['zeeman']:
        h.update(np.asarray(z if z is not None else np.zeros((3, 3)), float).tobytes())
    for row in ss['coupling']:
        for c in row:
            h.update(b'-' if c is None else np.asarray(c, float).tobytes())
    b = ss.get('basis')
    h.update(b'hilbert' if b is None else np.ascontiguousarray(np.asarray(b, np.int64)).tobytes())
    return h.hexdigest()


def hamiltonian(ss):
    """Isotropic part I and irreducible components Q[r][k][m] (r=1,2 as
    Q[0], Q[1]; k,m descending projection) of the Liouvillian commutation
    superoperator (kernel/hamiltonian.m:44-792). Returns (I, Q) with scipy
    CSR complex128 matrices; empty components are None. With 'ham_cache' in
    sys.enable the result is kept per process under a digest of its inputs
    and returned from there on a repeated call (hamiltonian.m:600-640)."""
    if 'ham_cache' in ss.get('enable', []):
        key = _ham_key(ss)
        if key not in _HAM_CACHE:
            _HAM_CACHE[key] = _hamiltonian(ss)
        I, Q = _HAM_CACHE[key]
        return I.copy(), [[[q.copy() if q is not None else None for q in row] for row in rank] for rank in Q]
    return _hamiltonian(ss)


def _hamiltonian(ss):
    n = ss['nspins']
    tol = ss['tols']['liouv_zero']
    cut = ss['tols']['inter_cutoff']
    # Descriptor rows: (ops, spins, isotropic, ist_coeff[8], irr_comp[8])
    descr = []
    s23 = math.sqrt(2 / 3)
    for a in range(n):                                    # hamiltonian.m:77-240
        zm = ss['zeeman'][a]
        strength = ss['zeeman_strength'][a]
        if strength == 'secular':
            ziso = np.trace(zm) / 3 - ss['basefrqs'][a]
        elif strength == 'full':
            ziso = np.trace(zm) / 3
        else:
            ziso = 0.0
        _, lam, phi = mat2sphten(zm)
        aniso = (np.linalg.norm(lam) > tol) or (np.linalg.norm(phi) > tol)
        irr = np.concatenate([lam, phi]) if aniso else np.zeros(8, complex)
        ist_lz = np.zeros(8, complex)
        if aniso and strength == 'secular':
            ist_lz[5] = s23                               # T(2,0) on Lz
        elif aniso and strength == 'full':
            ist_lz[5] = s23
            descr.append((('L+',), (a,), 0.0, _vec8({0: -0.5, 4: -0.5}), irr))
            descr.append((('L-',), (a,), 0.0, _vec8({2: -0.5, 6: 0.5}), irr))
        iso_lz = ziso if abs(ziso) > tol else 0.0
        descr.append((('Lz',), (a,), iso_lz, ist_lz, irr))
    # pair list in column-major find() order (hamiltonian.m:262-266)
    pairs = [(l, s) for s in range(n) for l in range(n)
             if l != s and ss['coupling'][l][s] is not None
             and np.linalg.norm(ss['coupling'][l][s], 2) > 2 * math.pi * cut]
    for l, s in pairs:                                    # hamiltonian.m:290-575
        cm = ss['coupling'][l][s]
        strength = ss['coupling_strength'][l][s]
        ciso = np.trace(cm) / 3
        iso_terms = {}
        if abs(ciso) > 2 * math.pi * cut:
            if strength in ('strong', 'secular'):
                iso_terms = {('Lz', 'Lz'): ciso, ('L+', 'L-'): ciso / 2, ('L-', 'L+'): ciso / 2}
            elif strength in ('z*', '*z', 'zz'):
                iso_terms = {('Lz', 'Lz'): ciso}
        _, lam, phi = mat2sphten(cm)
        aniso = (np.linalg.norm(lam) > tol) or (np.linalg.norm(phi) > tol)
        irr = np.concatenate([lam, phi]) if aniso else np.zeros(8, complex)
        ist = {}
        if aniso:
            if strength == 'secular':
                ist = {('Lz', 'Lz'): {5: s23},
                       ('L+', 'L-'): {1: -math.sqrt(1 / 8), 5: -s23 / 4},
                       ('L-', 'L+'): {1: math.sqrt(1 / 8), 5: -s23 / 4}}
            elif strength == 'zz':
                ist = {('Lz', 'Lz'): {5: s23}}
            elif strength == 'strong':
                ist = {('Lz', 'Lz'): {5: s23},
                       ('L+', 'L-'): {1: -math.sqrt(1 / 8), 5: -s23 / 4},
                       ('L-', 'L+'): {1: math.sqrt(1 / 8), 5: -s23 / 4},
                       ('Lz', 'L+'): {0: 0.5, 4: -0.5},
                       ('L+', 'Lz'): {0: -0.5, 4: -0.5},
                       ('Lz', 'L-'): {2: 0.5, 6: 0.5},
                       ('L-', 'Lz'): {2: -0.5, 6: 0.5},
                       ('L+', 'L+'): {3: 0.5},
                       ('L-', 'L-'): {7: 0.5}}
            elif strength != 'ignore':
                raise NotImplementedError(f'coupling strength {strength} is not supported')
        for ops in set(iso_terms) | set(ist):
            descr.append((ops, (l, s), iso_terms.get(ops, 0.0),
                          _vec8(ist.get(ops, {})), irr))
    # operators (hamiltonian.m:673-690), assembly of I (:718-738) and Q (:741-792): natively
    # (sp_ham_build) in the sphten-liouv basis; in Python for zeeman-hilb, or with
    # SPINACH_PY_BUILDER=1 (the restatement the native builder is tested against)
    if ss.get('formalism') != 'zeeman-hilb' and os.environ.get('SPINACH_PY_BUILDER', '0') != '1':
        from . import _native
        return _native.build_hamiltonian(n, ss['basis'], descr, tol)
    dim = 2 ** n if ss.get('formalism') == 'zeeman-hilb' else len(ss['basis_keys'])
    opers = []
    for ops, sp_, iso, ist, irr in descr:
        if abs(iso) < tol and (np.sum(np.abs(irr)) < tol or np.sum(np.abs(ist)) < tol):
            opers.append(None)
            continue
        opers.append(operator(ss, list(ops), list(sp_), 'comm'))
    I = sps.csr_matrix((dim, dim), dtype=complex)
    for (ops, sp_, iso, ist, irr), op in zip(descr, opers):
        if op is not None and abs(iso) > tol:
            I = I + iso * op
    I = _clean_up(I, tol)
    Q = [[[None] * 3 for _ in range(3)], [[None] * 5 for _ in range(5)]]
    for r in (1, 2):
        off = 0 if r == 1 else 3
        for k in range(2 * r + 1):
            for m in range(2 * r + 1):
                acc = None
                for (ops, sp_, iso, ist, irr), op in zip(descr, opers):
                    if op is None:
                        continue
                    coeff = ist[off + k] * irr[off + m]
                    if abs(coeff) > tol:
                        acc = coeff * op if acc is None else acc + coeff * op
                if acc is not None:
                    acc = _clean_up(acc, tol)
                    if acc.nnz:
                        Q[r - 1][k][m] = acc
    return I, Q


def _vec8(d):
    v = np.zeros(8, complex)
    for i, x in d.items():
        v[i] = x
    return v


def frqoffset(ss, I, offsets):
    """Add 2*pi*offset*Lz for every spin of each channel (kernel/frqoffset.m).
    `offsets` maps isotope name -> offset in Hz."""
    for iso, off in (offsets or {}).items():
        if off != 0:
            I = I + 2 * math.pi * off * operator(ss, 'Lz', iso, 'comm')
    return I


def orientation(Q, euler):
    """Anisotropic part at one orientation, then (H+H')/2 (kernel/orientation.m:29-59)."""
    h = None
    for r in (1, 2):
        d = wigner(r, *euler)
        for k in range(2 * r + 1):
            for m in range(2 * r + 1):
                q = Q[r - 1][k][m]
                if q is not None and q.nnz > 0:
                    h = d[k, m] * q if h is None else h + d[k, m] * q
    if h is None:
        return None
    return (h + h.conj().T) / 2


def q_list(Q):
    """Flatten Q into the 34-slot order used at the C-ABI: rank 1 (k,m)
    row-major, then rank 2 (k,m) row-major; empty slots are None."""
    out = []
    for r in (1, 2):
        for k in range(2 * r + 1):
            for m in range(2 * r + 1):
                out.append(Q[r - 1][k][m])
    return out
