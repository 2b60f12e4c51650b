"""Learned powder interpolation with the plain method as acceptance check and fallback.

The plain GPU method computes FIDs on a coarse training grid and on held-out
orientations; a Fourier-feature MLP (BF16 MFMA kernels, libspinach_powder) is
trained on the coarse FIDs; its relative error on the held-out orientations
decides whether the fine-grid powder average comes from the network or from
the plain method. No counterpart exists in the reference.
"""
from __future__ import annotations

import time

import numpy as np

from . import _native


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def plain_per_orientation(w, grid, mode='fp64'):
    """Per-orientation FIDs of the plain method (unweighted), and the device time."""
    plan = _native.plan_powder(w['I'], w['Q'], grid, w['rho0'], w['coil'], w['dt'], w['nsteps'], mode=mode)
    st = plan.execute()
    fid, per = plan.result(per_orientation=True)
    plan.close()
    return per, st['elapsed_s']


def learned_powder(w, train_grid, heldout_grid, fine_grid, tol=1e-4, steps=3000, batch=4096, lr=1e-3,
                   width=256, depth=4, ka=3, kt=22, seed=0):
    """Returns (fid, info). info holds the held-out error, the decision
    ('surrogate' or 'fallback'), and the timings of both paths."""
    npts = w['nsteps'] + 1
    t_max = w['nsteps'] * w['dt']
    t0 = time.perf_counter()
    train_fids, t_train_plain = plain_per_orientation(w, train_grid)
    held_fids, _ = plain_per_orientation(w, heldout_grid)
    # the network sees FIDs scaled to unit peak; the scale is restored after inference
    scale = float(np.abs(train_fids).max())
    # time harmonics in steps of 1/(2 kt dt) up to the Nyquist frequency 1/(2 dt): the FID's band
    net = _native.SurrogateNet(t_max, width=width, depth=depth, ka=ka, kt=kt, seed=seed, t_feat=kt * w['dt'])
    t1 = time.perf_counter()
    trace = net.train(train_grid, train_fids / scale, w['dt'], steps, batch=batch, lr=lr)
    t_train = time.perf_counter() - t1
    _, held_pred, _ = net.powder(heldout_grid, npts, w['dt'], per_orientation=True)
    held_err = rel_err(held_pred * scale, held_fids)
    info = dict(heldout_rel_err=held_err, final_loss=float(trace[-1]) if len(trace) else None,
                train_seconds=t_train, train_orientations=len(train_grid['weights']),
                heldout_orientations=len(heldout_grid['weights']), fine_orientations=len(fine_grid['weights']))
    fid_s, _, ms_s = net.powder(fine_grid, npts, w['dt'])
    info['surrogate_fine_ms'] = ms_s
    plan = _native.plan_powder(w['I'], w['Q'], fine_grid, w['rho0'], w['coil'], w['dt'], w['nsteps'])
    st = plan.execute()
    fid_p = plan.result()
    plan.close()
    info['plain_fine_ms'] = st['elapsed_s'] * 1e3
    info['plain_stats'] = st
    info['speedup_surrogate_vs_plain'] = info['plain_fine_ms'] / ms_s if ms_s > 0 else None
    info['surrogate_fine_rel_err_vs_plain'] = rel_err(fid_s * scale, fid_p)
    if held_err <= tol:
        info['decision'] = 'surrogate'
        fid = fid_s * scale
    else:
        info['decision'] = 'fallback'
        fid = fid_p
    info['total_seconds'] = time.perf_counter() - t0
    net.close()
    return fid, info
