"""Host-side mirror of Spinach's powder context for the time-domain
acquisition path:

    fid = powder(spin_system, acquire, parameters, 'nmr')

kernel/contexts/powder.m:99-309 builds I and Q once, loads the grid, runs the
pulse sequence at every orientation and forms the weighted sum. Here the
orientation loop and the sum run on the GPU through libspinach_powder
(sp_powder_fid); only the host setup (hamiltonian, frqoffset, states) runs in
Python, as the MATLAB host code does in the reference.
"""
from __future__ import annotations

import numpy as np

from . import _native
from . import grids as _grids
from . import spinsys as S


def acquire(spin_system, parameters, H, R, K):
    """Marker for experiments/acquire.m. The powder context dispatches this
    sequence to the GPU library; calling it directly is not supported."""
    raise NotImplementedError('acquire() runs inside powder() on the GPU')


def hp_acquire(spin_system, parameters, H, R, K):
    """Marker for experiments/hp_acquire.m: an ideal pulse
    rho0 <- step(pulse_op, rho0, pulse_angle), then acquire(). powder()
    applies the pulse on the host (it does not depend on the orientation) and
    acquires on the GPU. With parameters.echo_time (and echo_oper, echo_angle)
    every orientation's state evolves for echo_time, takes the echo pulse and
    evolves for echo_time again before the acquisition (hp_acquire.m:51-58):
    the evolutions run on the GPU (sp_plan_propagate), the echo pulse on the
    host."""
    raise NotImplementedError('hp_acquire() runs inside powder() on the GPU')


def cpmg(spin_system, parameters, H, R, K):
    """Marker for experiments/cpmg.m: an ideal pi/2 pulse (pulse_op), an
    acquisition of npoints, then nloops of [ideal pi pulse, acquisition of
    2*npoints]; parameters: rho0, coil, pulse_op, timestep, npoints, nloops,
    grid. powder() runs the acquisitions and the between-pulse propagation of
    every orientation's state on the GPU and the orientation-independent pulses
    on the host."""
    raise NotImplementedError('cpmg() runs inside powder() on the GPU')


def sp_acquire(spin_system, parameters, H, R, K):
    """Marker for experiments/sp_acquire.m: a soft pulse on parameters.spins[0]
    (pulse_frq, pulse_pwr, pulse_dur, pulse_phi, pulse_rnk; the Fokker-Planck
    phase-space formulation of kernel/pulses/shaped_pulse_af.m) under each
    orientation's own Hamiltonian, then acquire()."""
    raise NotImplementedError('sp_acquire() runs inside powder() on the GPU')


_REQUIRED = ('rho0', 'coil', 'sweep', 'npoints', 'grid')


def _hp_pulse(parameters):
    """hp_acquire.m:47-50: rho = step(pulse_op, rho0, pulse_angle), the exact
    exponential of the pulse superoperator applied to rho0 (step() sums its
    Taylor series to machine precision)."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spl
    for f in ('pulse_op', 'pulse_angle'):
        if f not in parameters:
            raise ValueError(f'parameters.{f} subfield must be present.')
    _hp_echo_grumble(parameters)
    ang = parameters['pulse_angle']
    if not (np.isscalar(ang) and np.isreal(ang) and np.isfinite(ang)):
        raise ValueError('parameters.pulse_angle should be a real number.')
    op = sps.csr_matrix(parameters['pulse_op'], dtype=complex)
    rho = np.asarray(parameters['rho0'], complex).ravel()
    if op.shape != (rho.size, rho.size):
        raise ValueError('parameters.pulse_op must be a square operator of the state dimension.')
    return spl.expm_multiply(-1j * float(ang) * op, rho)


def _hp_echo_grumble(parameters):
    """hp_acquire.m:112-130: the echo fields come together and are real scalars."""
    if 'echo_time' in parameters:
        et = parameters['echo_time']
        if not (np.isscalar(et) and np.isreal(et) and np.isfinite(et) and et > 0):
            raise ValueError('parameters.echo_time must be a positive real scalar.')
        if 'echo_oper' not in parameters:
            raise ValueError('echo pulse operator must be specified in parameters.echo_oper variable.')
        if 'echo_angle' not in parameters:
            raise ValueError('echo pulse angle must be specified in parameters.echo_angle variable.')
        ea = parameters['echo_angle']
        if not (np.isscalar(ea) and np.isreal(ea) and np.isfinite(ea)):
            raise ValueError('parameters.echo_angle must be a real scalar.')
    elif 'echo_oper' in parameters or 'echo_angle' in parameters:
        raise ValueError('echo time must be specified in parameters.echo_time variable.')


def _grumble(parameters):
    """Input validation in the spirit of powder.m:340-428, plus the acquire.m
    options this path does not run (decouple.m, homodecoupling), which must
    not be dropped silently."""
    for f in _REQUIRED:
        if f not in parameters:
            raise ValueError(f'parameters.{f} subfield must be present.')
    if parameters.get('rframes'):
        raise NotImplementedError('rotating frames are not supported on this path')
    if parameters.get('needs'):
        raise NotImplementedError('parameters.needs is not supported on this path')
    if callable(parameters['rho0']):
        raise NotImplementedError('orientation-dependent rho0 is not supported on this path')
    if parameters.get('decouple'):
        raise NotImplementedError('decoupling (acquire.m:51) is not supported on this path')
    if 'homodec_oper' in parameters:
        raise NotImplementedError('homodecoupling (acquire.m:54-62) is not supported on this path')
    if parameters['sweep'] <= 0:
        raise ValueError('parameters.sweep must be positive.')
    if int(parameters['npoints']) < 1:
        raise ValueError('parameters.npoints must be a positive integer.')
    if 'dead_time' in parameters:
        td = parameters['dead_time']
        if not (np.isscalar(td) and np.isfinite(td) and td >= 0):
            raise ValueError('parameters.dead_time must be a non-negative real scalar.')


def _grid(g):
    if isinstance(g, dict):
        return {k: np.asarray(g[k], np.float64) for k in ('alphas', 'betas', 'gammas', 'weights')}
    if isinstance(g, str):
        if g.startswith('zcwn_'):
            return _grids.zcwn(int(g.split('_')[1]))
        if g == 'single_crystal':
            return _grids.single_crystal()
        return _grids.load_grid(g)   # a Spinach grid file (powder.m:173-179)
    raise ValueError(f'unknown grid specification {g!r}')


def _acquire_on_grid(ss, I, qlist, parameters, relax='orientation'):
    """acquire() at every orientation of parameters['grid'] and the weighted
    sum (powder.m:226-294) on the GPU: dt = 1/sweep, nsteps = npoints-1
    (acquire.m:70-71), the dead time (acquire.m:65-67), spin_system.tols and
    sys.enable = ['krylov'] (evolution.m:531-533). Returns (answer, grid)."""
    grid = parameters['grid']
    dt = 1.0 / parameters['sweep']
    nsteps = int(parameters['npoints']) - 1
    sum_up = parameters.get('sum_up', True)
    mode = parameters.get('mode', 'fp64')
    if 'krylov' in ss.get('enable', []):
        mode = mode + '+krylov'
    # R(Omega) at each orientation (powder.m:263) in the library's form, unless the caller
    # passes its own (a context whose generator lives in a larger space)
    if isinstance(relax, str):
        relax = S.relaxation_terms(ss)
    plan = _native.plan_powder(I, qlist, grid, parameters['rho0'], parameters['coil'], dt, nsteps,
                               tols=ss['tols'], mode=mode, relax=relax)
    try:
        if parameters.get('dead_time'):
            plan.set_dead_time(parameters['dead_time'])
        if parameters.get('initial_states') is not None:
            plan.set_initial_states(parameters['initial_states'])
        plan.execute()
        if sum_up:
            fid = plan.result()
        else:
            fid, per = plan.result(per_orientation=True)
    finally:
        plan.close()
    if sum_up:
        return fid, grid
    return [per[k] for k in range(per.shape[0])], grid


def _closure(I, qlist, extra, vec):
    """Full-space indices of the states reachable from the support of vec
    under the union sparsity pattern of I, the Q slots and the extra operators."""
    import scipy.sparse as sps
    import scipy.sparse.csgraph as cg
    u = abs(sps.csr_matrix(I))
    for q in list(qlist) + list(extra):
        if q is not None:
            u = u + abs(sps.csr_matrix(q))
    _, lab = cg.connected_components((u + u.T) > 0, directed=False)
    hit = np.unique(lab[np.abs(np.asarray(vec).ravel()) > 0])
    return np.nonzero(np.isin(lab, hit))[0]


def _hp_echo_on_grid(ss, I, qlist, parameters):
    """hp_acquire.m:51-64 at every orientation: the pulsed rho0 (already in parameters) evolves
    for echo_time under the orientation's generator, takes step(echo_oper, ., echo_angle),
    evolves for echo_time again, and acquire() runs from those per-orientation states. The
    evolutions run on a plan whose reduction is seeded by every state the echo pulse and L can
    reach (sp_plan_propagate); the acquisition on a plan with per-orientation initial states."""
    import scipy.linalg as sla
    import scipy.sparse as sps
    if parameters.get('dead_time'):
        raise NotImplementedError('hp_acquire takes no dead time')
    grid = parameters['grid']
    n_or = len(grid['weights'])
    dim = I.shape[0]
    eop = sps.csr_matrix(parameters['echo_oper'], dtype=complex)
    rho = np.asarray(parameters['rho0'], complex).ravel()
    coil = np.asarray(parameters['coil'], complex).ravel()
    if eop.shape != (dim, dim):
        raise ValueError('parameters.echo_oper must be a square operator of the state dimension.')
    st = _closure(I, qlist, [eop], np.abs(rho) + np.abs(coil))
    seed = np.zeros(dim, complex)
    seed[st] = 1.0
    U = sla.expm(-1j * float(parameters['echo_angle']) * eop[st][:, st].toarray())
    mode = parameters.get('mode', 'fp64')
    tau = float(parameters['echo_time'])
    relax = S.relaxation_terms(ss)              # L = H + iR(Omega) in both evolutions and the acquisition
    prop = _native.plan_powder(I, qlist, grid, rho, seed, tau, 1, tols=ss['tols'], mode=mode, relax=relax)
    acq = _native.plan_powder(I, qlist, grid, rho, coil, 1.0 / parameters['sweep'], int(parameters['npoints']) - 1,
                              tols=ss['tols'], mode=mode, relax=relax)
    try:
        prop.set_initial_states(np.tile(rho, (n_or, 1)))
        states = prop.propagate(tau, 1, dim)
        states[:, st] = states[:, st] @ U.T
        prop.set_initial_states(states)
        states = prop.propagate(tau, 1, dim)
        acq.set_initial_states(states)
        acq.execute()
        fid, per = acq.result(per_orientation=True)
    finally:
        prop.close()
        acq.close()
    if parameters.get('sum_up', True):
        return fid, grid
    return [per[k] for k in range(n_or)], grid


def _cpmg_on_grid(ss, I, qlist, parameters):
    """experiments/cpmg.m:36-57 at every orientation: rho = step(pulse_op, rho0,
    pi/2); fid = coil'*trajectory of npoints; then nloops times: the last state
    takes step(pulse_op, ., pi) and a trajectory of 2*npoints follows. The
    acquisitions run as plans with per-orientation initial states; the last
    state of each segment comes from sp_plan_propagate on a plan whose
    reduction is seeded by every state the pulses and L can reach."""
    import scipy.linalg as sla
    import scipy.sparse as sps
    import scipy.sparse.linalg as spl
    for f in ('nloops', 'timestep', 'npoints', 'pulse_op', 'rho0', 'coil'):
        if f not in parameters:
            raise ValueError(f'parameters.{f} must be specified.')
    nl, npts, dt = parameters['nloops'], parameters['npoints'], parameters['timestep']
    if not (isinstance(nl, (int, np.integer)) and nl >= 1):
        raise ValueError('parameters.nloops should be a positive integer.')
    if not (isinstance(npts, (int, np.integer)) and npts >= 1):
        raise ValueError('parameters.npoints should be a positive integer.')
    if not (np.isscalar(dt) and np.isreal(dt) and dt > 0):
        raise ValueError('parameters.timestep should be a positive real number.')
    grid = parameters['grid']
    n_or = len(grid['weights'])
    dim = I.shape[0]
    pop = sps.csr_matrix(parameters['pulse_op'], dtype=complex)
    rho = spl.expm_multiply(-1j * (np.pi / 2) * pop, np.asarray(parameters['rho0'], complex).ravel())
    coil = np.asarray(parameters['coil'], complex).ravel()
    st = _closure(I, qlist, [pop], np.abs(rho) + np.abs(coil))
    seed = np.zeros(dim, complex)
    seed[st] = 1.0
    U = sla.expm(-1j * np.pi * pop[st][:, st].toarray())    # step(pulse_op, ., pi) on the reachable states
    mode = parameters.get('mode', 'fp64')
    relax = S.relaxation_terms(ss)              # L = H + iR(Omega) between and during the acquisitions
    prop = _native.plan_powder(I, qlist, grid, rho, seed, dt, 1, tols=ss['tols'], mode=mode, relax=relax)
    plans = [_native.plan_powder(I, qlist, grid, rho, coil, dt, n - 1, tols=ss['tols'], mode=mode, relax=relax)
             for n in (npts, 2 * npts)]
    try:
        states = np.tile(rho, (n_or, 1))
        pers = []
        for seg in range(nl + 1):
            plan, n = (plans[0], npts) if seg == 0 else (plans[1], 2 * npts)
            if seg > 0:
                states[:, st] = states[:, st] @ U.T
            plan.set_initial_states(states)
            plan.execute()
            pers.append(plan.result(per_orientation=True)[1])
            if seg < nl:
                prop.set_initial_states(states)
                states = prop.propagate(dt, n - 1, dim)
    finally:
        for p in plans + [prop]:
            p.close()
    per = np.concatenate(pers, axis=1)
    w = np.asarray(grid['weights'])
    fid = w[0] * per[0]
    for o in range(1, n_or):
        fid = fid + w[o] * per[o]
    if parameters.get('sum_up', True):
        return fid, grid
    return [per[k] for k in range(n_or)], grid


def _sp_acquire_on_grid(ss, I, qlist, parameters):
    """experiments/sp_acquire.m:38-54 at every orientation: the soft pulse of
    kernel/pulses/shaped_pulse_af.m (one constant slice) in its Fokker-Planck
    phase space of N = 2*pulse_rnk+1 phases phi_n = 2*pi*n/N + pulse_phi,
        F(Omega) = kron(1_N, H(Omega)) + pwr*(kron(diag(cos phi), Ex) + kron(diag(sin phi), Ey))
                   + 2i*pi*frq*kron(d/dphi, 1),
    rho <- sum over the phase blocks of exp(-i F(Omega) dur) kron(e_1, rho0),
    with Ep = operator(L+) on parameters.spins[0], Ex = (Ep+Ep')/2,
    Ey = (Ep-Ep')/2i and frq = pulse_frq - offset; then acquire() from that
    state. F(Omega) is an ordinary powder problem (the slots kron(1_N, Q)), so
    the library propagates it (sp_plan_propagate) and the acquisition plan
    starts from the folded states (sp_plan_set_initial_states)."""
    import scipy.sparse as sps
    from . import contexts as Cx
    for f in ('spins', 'pulse_frq', 'pulse_pwr', 'pulse_dur', 'pulse_phi', 'pulse_rnk'):
        if f not in parameters:
            raise ValueError(f'parameters.{f} must be specified.')
    for f in ('pulse_frq', 'pulse_pwr', 'pulse_dur', 'pulse_phi'):
        v = parameters[f]
        if not (np.isscalar(v) and np.isreal(v) and np.isfinite(v)):
            raise ValueError(f'parameters.{f} must be a real scalar.')
    if parameters['pulse_pwr'] < 0:
        raise ValueError('parameters.pulse_pwr must be a non-negative real scalar.')
    if parameters['pulse_dur'] <= 0:
        raise ValueError('parameters.pulse_dur must be a positive real scalar.')
    rnk = parameters['pulse_rnk']
    if not (isinstance(rnk, (int, np.integer)) and rnk >= 0):
        raise ValueError('parameters.pulse_rnk must be a non-negative integer.')
    spin = parameters['spins'][0]
    off = (parameters.get('offset') or {}).get(spin, 0.0)
    frq = float(parameters['pulse_frq']) - float(off)
    ep = S.operator(ss, 'L+', spin)
    ex, ey = (ep + ep.conj().T) / 2, (ep - ep.conj().T) / 2j
    N = 2 * int(rnk) + 1
    phases, d_dphi = Cx.fourdif(N)
    phases = phases + float(parameters['pulse_phi'])
    one = sps.identity(N, format='csr')
    dim = I.shape[0]
    I_fp = (sps.kron(one, sps.csr_matrix(I, dtype=complex))
            + float(parameters['pulse_pwr']) * (sps.kron(sps.diags(np.cos(phases)), ex)
                                               + sps.kron(sps.diags(np.sin(phases)), ey))
            + 2j * np.pi * frq * sps.kron(sps.csr_matrix(d_dphi), sps.identity(dim, format='csr')))
    q_fp = [None if q is None else sps.kron(one, sps.csr_matrix(q, dtype=complex), format='csr') for q in qlist]
    proj = np.zeros(N)
    proj[0] = 1.0
    rho_fp = np.kron(proj, np.asarray(parameters['rho0'], complex).ravel())
    seed = (np.abs(rho_fp) > 0).astype(complex)   # the reduction keeps what the pulse can reach from rho0
    grid = parameters['grid']
    n_or = len(grid['weights'])
    dt_acq = 1.0 / parameters['sweep']
    k = max(1, int(np.ceil(parameters['pulse_dur'] / dt_acq)))
    dur = float(parameters['pulse_dur'])
    mode = parameters.get('mode', 'fp64')
    # relaxation (sp_acquire.m: L = H + 1i*R, kron(1_N, L) in the phase space): R0 and the
    # per-orientation diagonal planes replicated over the phase blocks
    rl = S.relaxation_terms(ss)
    relax_fp = None
    if rl is not None:
        relax_fp = {'R0': None if rl['R0'] is None else sps.kron(one, rl['R0'], format='csr'),
                    'diag': None if rl['diag'] is None else np.tile(rl['diag'], (1, N))}
    plan = _native.plan_powder(I_fp.tocsr(), q_fp, grid, rho_fp, seed, dur / k, k, tols=ss['tols'], mode=mode,
                               relax=relax_fp)
    try:
        st_fp = plan.propagate(dur / k, k, N * dim)
    finally:
        plan.close()
    states = st_fp.reshape(n_or, N, dim).sum(axis=1)
    return _acquire_on_grid(ss, I, qlist, dict(parameters, initial_states=states))


def powder(spin_system, pulse_sequence, parameters, assumptions):
    """Powder-averaged FID (kernel/contexts/powder.m:99-309) for
    pulse_sequence == acquire. parameters: rho0, coil, sweep (Hz), npoints,
    grid (dict, 'zcwn_<n>', 'single_crystal' or a Spinach grid file name),
    optional offset (dict isotope -> Hz), dead_time (s, acquire.m:65-67),
    sum_up (default True), mode ('fp64' | 'fp32'). spin_system.tols is passed
    to the library and spin_system.sys.enable = ['krylov'] selects time
    stepping (evolution.m:531-533).

    Returns (answer, grid) like the reference: answer is the weighted sum, or
    a list of per-orientation FIDs when sum_up is False."""
    if pulse_sequence is hp_acquire:
        parameters = dict(parameters, rho0=_hp_pulse(parameters))
    elif pulse_sequence is cpmg:
        # cpmg takes timestep and npoints itself (checked there); sweep is not one of its parameters
        parameters = dict(parameters, sweep=1.0)
    elif pulse_sequence not in (acquire, sp_acquire):
        raise NotImplementedError('only the acquire(), hp_acquire(), sp_acquire() and cpmg() sequences run on '
                                  'this path')
    _grumble(parameters)
    ss = S.assume(spin_system, assumptions)                 # powder.m:151
    I, Q = S.hamiltonian(ss)                                # powder.m:155
    I = S.frqoffset(ss, I, parameters.get('offset'))        # powder.m:161
    params = dict(parameters, grid=_grid(parameters['grid']))  # powder.m:173-179
    if pulse_sequence is cpmg:
        return _cpmg_on_grid(ss, I, S.q_list(Q), params)
    if pulse_sequence is hp_acquire and 'echo_time' in params:
        return _hp_echo_on_grid(ss, I, S.q_list(Q), params)
    if pulse_sequence is sp_acquire:
        return _sp_acquire_on_grid(ss, I, S.q_list(Q), params)
    return _acquire_on_grid(ss, I, S.q_list(Q), params)
