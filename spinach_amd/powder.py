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
    acquires on the GPU; the echo variant (parameters.echo_time) is not
    supported."""
    raise NotImplementedError('hp_acquire() runs inside powder() on the GPU')


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
    if 'echo_time' in parameters:
        raise NotImplementedError('the echo variant of hp_acquire (hp_acquire.m:51-58) is not supported on this path')
    ang = parameters['pulse_angle']
    if not (np.isscalar(ang) and np.isreal(ang) and np.isfinite(ang)):
        raise ValueError('parameters.pulse_angle should be a real number.')
    op = sps.csr_matrix(parameters['pulse_op'], dtype=complex)
    rho = np.asarray(parameters['rho0'], complex).ravel()
    if op.shape != (rho.size, rho.size):
        raise ValueError('parameters.pulse_op must be a square operator of the state dimension.')
    return spl.expm_multiply(-1j * float(ang) * op, rho)


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


def _acquire_on_grid(ss, I, qlist, parameters):
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
    # R(Omega) at each orientation (powder.m:263) in the library's form
    relax = S.relaxation_terms(ss)
    plan = _native.plan_powder(I, qlist, grid, parameters['rho0'], parameters['coil'], dt, nsteps,
                               tols=ss['tols'], mode=mode, relax=relax)
    try:
        if parameters.get('dead_time'):
            plan.set_dead_time(parameters['dead_time'])
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
    elif pulse_sequence is not acquire:
        raise NotImplementedError('only the acquire() and hp_acquire() sequences run on this path')
    _grumble(parameters)
    ss = S.assume(spin_system, assumptions)                 # powder.m:151
    I, Q = S.hamiltonian(ss)                                # powder.m:155
    I = S.frqoffset(ss, I, parameters.get('offset'))        # powder.m:161
    params = dict(parameters, grid=_grid(parameters['grid']))  # powder.m:173-179
    return _acquire_on_grid(ss, I, S.q_list(Q), params)
