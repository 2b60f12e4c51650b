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
    net = _native.SurrogateNet(2e-3, width=width, depth=depth, ka=ka, kt=kt, seed=3, t_feat=kt * 1e-5)
    p = (rng.normal(size=net.nparam) * 0.2).astype(np.float32)
    net.set_params(p)
    n = 256
    a, b, g = rng.uniform(0, 6, n), rng.uniform(0, 3, n), rng.uniform(0, 6, n)
    t = rng.integers(0, 200, n).astype(np.int32)
    tgt = rng.normal(size=(n, 2)).astype(np.float32)
    loss, grad = net.grad(a, b, g, t, 1e-5, tgt)
    x = M.features(a, b, g, t, 1e-5, 2e-3, ka, kt, t_feat=kt * 1e-5)
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


def _band_limited(g, npts, dt, kt):
    """Three lines at 3, 9 and 17 steps of the feature frequency spacing
    1/(2 kt dt) (13.6, 40.9 and 77.3 kHz at dt = 5 us: the band of the 8-spin
    FIDs), smooth orientation-dependent amplitudes, 2 ms decay."""
    t = np.arange(npts) * dt
    nu = np.array([3, 9, 17]) / (2 * kt * dt)
    b, c = np.asarray(g['betas'])[:, None], np.asarray(g['gammas'])[:, None]
    amp = [0.25 * (3 * np.cos(b) ** 2 - 1) + 0.3, 0.4 * np.sin(b) * np.cos(c), 0.2 * np.cos(2 * c) * np.sin(b) ** 2]
    return sum(a * np.exp(-2j * np.pi * f * t)[None, :] for a, f in zip(amp, nu)) * np.exp(-t / 2e-3)[None, :]


def test_network_learns_a_band_limited_signal_at_the_real_sweep_width():
    """With harmonics spanning the band up to the Nyquist frequency the network
    represents a signal whose lines sit anywhere in the band and whose
    amplitudes depend smoothly on the orientation: held-out error and the
    fine-grid powder average against the exact one (measured: loss down
    900-2400-fold, 5.8-7.4e-2 held out, 5.3-9.5e-2 on the average; the
    trajectory depends on the summation order of the FP32 gradients)."""
    rng = np.random.default_rng(4)
    dt, npts, kt = 5e-6, 1024, 22
    tr, ho, fine = grids.haar(300, rng), grids.haar(200, rng), grids.haar(4000, rng)
    net = _native.SurrogateNet(npts * dt, width=256, depth=4, kt=kt, seed=5, t_feat=kt * dt)
    trace = net.train(tr, _band_limited(tr, npts, dt, kt), dt, 3000, batch=8192, lr=1e-3)
    _, ph, _ = net.powder(ho, npts, dt, per_orientation=True)
    e_ho = O.rel_err(ph, _band_limited(ho, npts, dt, kt))
    fid, _, ms = net.powder(fine, npts, dt)
    exact = (np.asarray(fine['weights'])[:, None] * _band_limited(fine, npts, dt, kt)).sum(axis=0)
    e_avg = O.rel_err(fid, exact)
    print(f'band-limited target: loss {trace[0]:.3e} -> {trace[-1]:.3e}, held-out {e_ho:.2e}, '
          f'fine-grid average {e_avg:.2e} ({ms:.1f} ms for 4000 x 1024 points)', flush=True)
    # BF16 weights and activations (8-bit mantissa) leave a few per cent; an untrained network is at ~1
    assert np.median(trace[-100:]) < 2e-3 * trace[0]
    assert e_ho < 0.1 and e_avg < 0.1
    net.close()


def test_acceptance_check_on_the_8spin_early_window():
    """On the 8-spin system the per-orientation FIDs decorrelate across
    orientations within ~16 steps, and 78% of the powder FID's energy lies in
    the first 8. There the network trained on 300 orientations and evaluated
    on a 20000-orientation grid is compared with the 300-orientation plain
    average (the method without it), both against the plain method on the
    fine grid (measured: 4.4e-2 against 4.2e-2, held-out 0.43). The network
    does not gain here, so the acceptance check must keep the plain method:
    the returned FID is the plain fine-grid average itself."""
    w = dict(workloads.static_8spin(n_orient=4000))
    w['nsteps'] = 8
    rng = np.random.default_rng(11)
    tr, ho, fine = grids.haar(300, rng), grids.haar(500, rng), grids.haar(20000, rng)
    fid, info = surrogate.learned_powder(w, tr, ho, fine, tol=1e-4, steps=3000, batch=2700, lr=1e-3)
    plain_fine, _ = surrogate.plain_per_orientation(w, fine)
    plain_tr, _ = surrogate.plain_per_orientation(w, tr)
    ref = (np.asarray(fine['weights'])[:, None] * plain_fine).sum(axis=0)
    coarse = (np.asarray(tr['weights'])[:, None] * plain_tr).sum(axis=0)
    e_net, e_coarse = info['surrogate_fine_rel_err_vs_plain'], O.rel_err(coarse, ref)
    print(f'8-spin first 9 points: held-out {info["heldout_rel_err"]:.2e}; fine-grid average with the network '
          f'{e_net:.2e}, 300-orientation plain average {e_coarse:.2e}; decision {info["decision"]}', flush=True)
    assert info['heldout_rel_err'] > 1e-4 and info['decision'] == 'fallback'
    assert O.rel_err(fid, ref) < 1e-12
    assert np.isfinite(e_net) and e_net < 1.0
