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


_REQUIRED = ('rho0', 'coil', 'sweep', 'npoints', 'grid')


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
    raise ValueError(f'unknown grid specification {g!r}')


def powder(spin_system, pulse_sequence, parameters, assumptions):
    """Powder-averaged FID (kernel/contexts/powder.m:99-309) for
    pulse_sequence == acquire. parameters: rho0, coil, sweep (Hz), npoints,
    grid (dict or 'zcwn_<n>' / 'single_crystal'), optional offset (dict
    isotope -> Hz), dead_time (s, acquire.m:65-67), sum_up (default True),
    mode ('fp64' | 'fp32'). spin_system.tols is passed to the library and
    spin_system.sys.enable = ['krylov'] selects time stepping
    (evolution.m:531-533).

    Returns (answer, grid) like the reference: answer is the weighted sum, or
    a list of per-orientation FIDs when sum_up is False."""
    if pulse_sequence is not acquire:
        raise NotImplementedError('only the acquire() sequence runs on this path')
    _grumble(parameters)
    ss = S.assume(spin_system, assumptions)                 # powder.m:151
    I, Q = S.hamiltonian(ss)                                # powder.m:155
    I = S.frqoffset(ss, I, parameters.get('offset'))        # powder.m:161
    grid = _grid(parameters['grid'])                        # powder.m:173-179
    dt = 1.0 / parameters['sweep']                          # acquire.m:70-71
    nsteps = int(parameters['npoints']) - 1
    sum_up = parameters.get('sum_up', True)
    mode = parameters.get('mode', 'fp64')
    if 'krylov' in ss.get('enable', []):
        mode = mode + '+krylov'
    plan = _native.plan_powder(I, S.q_list(Q), grid, parameters['rho0'], parameters['coil'], dt, nsteps,
                               tols=ss['tols'], mode=mode)
    try:
        if parameters.get('dead_time'):
            plan.set_dead_time(parameters['dead_time'])     # acquire.m:65-67
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
