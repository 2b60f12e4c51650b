"""GPU checks of the learned powder interpolation: BF16 MFMA gradients against
the FP64 reference, training, and the acceptance/fallback decision."""
import numpy as np
import pytest

import mlp_reference as M
import pyoracle as O
from spinach_amd import _native, grids, surrogate, workloads

pytestmark = pytest.mark.gpu


def test_gradients_match_fp64_reference():
    rng = np.random.default_rng(0)
    width, depth, ka, kt = 64, 2, 3, 22
    net = _native.SurrogateNet(2e-3, width=width, depth=depth, ka=ka, kt=kt, seed=3)
    p = (rng.normal(size=net.nparam) * 0.2).astype(np.float32)
    net.set_params(p)
    n = 256
    a, b, g = rng.uniform(0, 6, n), rng.uniform(0, 3, n), rng.uniform(0, 6, n)
    t = rng.integers(0, 200, n).astype(np.int32)
    tgt = rng.normal(size=(n, 2)).astype(np.float32)
    loss, grad = net.grad(a, b, g, t, 1e-5, tgt)
    x = M.features(a, b, g, t, 1e-5, 2e-3, ka, kt)
    rloss, rgrad = M.loss_and_grad(p.astype(np.float64), x, tgt.astype(np.float64), width, depth)
    e = np.linalg.norm(grad - rgrad) / np.linalg.norm(rgrad)
    cos = float(grad @ rgrad / (np.linalg.norm(grad) * np.linalg.norm(rgrad)))
    print(f'surrogate grad: loss {loss:.6f} vs {rloss:.6f}, rel grad err {e:.2e}, cosine {cos:.6f}', flush=True)
    # BF16 operands (8-bit mantissa) with FP32 accumulation
    assert abs(loss - rloss) < 2e-2 * rloss
    assert e < 5e-2 and cos > 0.998
    net.close()


def test_training_reduces_loss_on_smooth_target():
    rng = np.random.default_rng(1)
    g = grids.haar(32, rng)
    npts, dt = 128, 1e-5
    t = np.arange(npts) * dt
    fids = np.array([np.cos(b) * np.exp(-t / 5e-4) * np.exp(-1j * 2e3 * np.cos(c) * t)
                     for b, c in zip(g['betas'], g['gammas'])])
    net = _native.SurrogateNet(npts * dt, width=128, depth=3, seed=2)
    trace = net.train(g, fids, dt, 400, batch=1024, lr=2e-3)
    print(f'loss {trace[0]:.4e} -> {trace[-1]:.4e}', flush=True)
    assert trace[-1] < 0.1 * trace[0]
    _, per, _ = net.powder(g, npts, dt, per_orientation=True)
    assert O.rel_err(per, fids) < 0.3
    net.close()


def test_acceptance_check_falls_back_to_plain_method():
    w = workloads.static_8spin(nsteps=256, n_orient=4000)
    g = w['grid']
    tr, ho, fine = grids.subset(g, np.arange(24)), grids.subset(g, np.arange(24, 40)), grids.subset(g, np.arange(40, 104))
    for gg in (tr, ho, fine):
        gg['weights'] = np.full(len(gg['weights']), 1.0 / len(gg['weights']))
    fid, info = surrogate.learned_powder(w, tr, ho, fine, tol=1e-4, steps=100, batch=1024, width=64, depth=2)
    print('learned powder info', info, flush=True)
    ref = O.refcpu_powder(w['I'], w['Q'], fine, w['rho0'], w['coil'], w['dt'], w['nsteps'])
    if info['decision'] == 'fallback':
        assert info['heldout_rel_err'] > 1e-4
        assert O.rel_err(fid, ref) < 1e-10
    else:
        assert info['heldout_rel_err'] <= 1e-4
