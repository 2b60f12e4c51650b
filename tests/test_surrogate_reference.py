"""The FP64 reference of the surrogate network, pinned by central finite
differences; a planted mistake in the backward pass must fail the check."""
import numpy as np

import mlp_reference as M


def _setup(seed=0, n=7, width=8, depth=2, ka=1, kt=2):
    rng = np.random.default_rng(seed)
    F = M.n_features(ka, kt)
    x = M.features(rng.uniform(0, 6, n), rng.uniform(0, 3, n), rng.uniform(0, 6, n), rng.integers(0, 100, n), 1e-5,
                   1e-3, ka, kt)
    npar = sum(fi * fo + fo for fi, fo in M.shapes(F, width, depth))
    p = rng.normal(size=npar) * 0.3
    tgt = rng.normal(size=(n, 2))
    return p, x, tgt, width, depth


def test_reference_gradient_matches_finite_differences():
    p, x, tgt, width, depth = _setup()
    loss, g = M.loss_and_grad(p, x, tgt, width, depth)
    rng = np.random.default_rng(1)
    for i in rng.choice(len(p), 40, replace=False):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, x, tgt, width, depth)[0] - M.loss_and_grad(pm, x, tgt, width, depth)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 + 1e-6 * abs(fd)


def test_features_layout():
    x = M.features(np.array([0.3]), np.array([1.1]), np.array([2.0]), np.array([50]), 1e-5, 1e-3, 2, 3)
    assert x.shape[1] == 32
    assert abs(x[0, 0] - np.cos(0.3)) < 1e-15 and abs(x[0, 3] - np.sin(0.6)) < 1e-15
    cb, sb = np.cos(1.1), np.sin(1.1)
    assert abs(x[0, 12] - 0.5 * (3 * cb * cb - 1)) < 1e-15 and abs(x[0, 16] - sb * sb * np.sin(4.0)) < 1e-15
    tau = 50 * 1e-5 / 1e-3
    assert abs(x[0, 17] - np.cos(np.pi * tau)) < 1e-15
    assert x[0, 23] == tau and x[0, 24] == 1.0 and not x[0, 25:].any()
    # harmonics on their own time scale
    xf = M.features(np.array([0.3]), np.array([1.1]), np.array([2.0]), np.array([50]), 1e-5, 1e-3, 2, 3, t_feat=3e-5)
    assert abs(xf[0, 17] - np.cos(np.pi * 50 / 3)) < 1e-13 and xf[0, 23] == tau


def test_planted_backward_mistake_is_caught():
    p, x, tgt, width, depth = _setup(2)
    _, g = M.loss_and_grad(p, x, tgt, width, depth)
    # drop the tanh derivative: the gradient must move away from the true one
    F = x.shape[1]
    layers = M.unpack(p, F, width, depth)
    hs = [x]
    for l, (W, b) in enumerate(layers):
        z = hs[-1] @ W.T + b
        hs.append(np.tanh(z) if l < depth else z)
    n = x.shape[0]
    dz = 2 * (hs[-1] - tgt) / n
    W, b = layers[depth]
    wrong_dz_prev = dz @ W  # missing (1 - h^2)
    W0, _ = layers[depth - 1]
    wrong = wrong_dz_prev.T @ hs[depth - 1]
    off = sum(fi * fo + fo for fi, fo in M.shapes(F, width, depth)[:depth - 1])
    true_block = g[off:off + W0.size].reshape(W0.shape)
    assert np.linalg.norm(wrong - true_block) > 1e-3 * np.linalg.norm(true_block)
