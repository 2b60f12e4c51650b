"""FP64 numpy restatement of the surrogate network (features, forward, loss,
backward). TEST INFRASTRUCTURE ONLY: the checker for the BF16 MFMA kernels of
spinach_amd/csrc/surrogate.hip; pinned itself by central finite differences.

Parameter layout (matches the device): for each layer l (fan-in -> fan-out),
W_l (fan_out x fan_in, row-major) followed by b_l (fan_out); hidden layers use
tanh, the output layer (2 units) is linear; loss = mean over points of
|y - target|^2 summed over the two outputs.
"""
import math

import numpy as np


def n_features(ka, kt):
    nf = 6 * ka + 5 + 2 * kt + 2
    return (nf + 31) // 32 * 32


def features(alphas, betas, gammas, tidx, dt, t_max, ka, kt, t_feat=None):
    """t_max scales the linear time feature, t_feat (default t_max) the
    harmonics cos/sin(pi m t / t_feat)."""
    t_feat = t_max if t_feat is None else t_feat
    F = n_features(ka, kt)
    n = len(alphas)
    x = np.zeros((n, F))
    c = 0
    for ang in (alphas, betas, gammas):
        for k in range(1, ka + 1):
            x[:, c] = np.cos(k * ang)
            x[:, c + 1] = np.sin(k * ang)
            c += 2
    cb, sb, g = np.cos(betas), np.sin(betas), np.asarray(gammas)
    for v in (0.5 * (3 * cb * cb - 1), sb * cb * np.cos(g), sb * cb * np.sin(g), sb * sb * np.cos(2 * g),
              sb * sb * np.sin(2 * g)):
        x[:, c] = v
        c += 1
    tau = np.asarray(tidx) * dt / t_max
    sf = np.asarray(tidx) * dt / t_feat
    for m in range(1, kt + 1):
        x[:, c] = np.cos(math.pi * m * sf)
        x[:, c + 1] = np.sin(math.pi * m * sf)
        c += 2
    x[:, c] = tau
    x[:, c + 1] = 1.0
    return x


def shapes(F, width, depth):
    out = []
    for l in range(depth + 1):
        fi = F if l == 0 else width
        fo = 2 if l == depth else width
        out.append((fi, fo))
    return out


def unpack(p, F, width, depth):
    layers = []
    off = 0
    for fi, fo in shapes(F, width, depth):
        W = p[off:off + fi * fo].reshape(fo, fi)
        off += fi * fo
        b = p[off:off + fo]
        off += fo
        layers.append((W, b))
    return layers


def loss_and_grad(p, x, target, width, depth):
    F = x.shape[1]
    layers = unpack(np.asarray(p, np.float64), F, width, depth)
    hs = [x]
    for l, (W, b) in enumerate(layers):
        z = hs[-1] @ W.T + b
        hs.append(np.tanh(z) if l < depth else z)
    y = hs[-1]
    n = x.shape[0]
    d = y - target
    loss = float((d * d).sum() / n)
    dz = 2 * d / n
    grads = []
    for l in range(depth, -1, -1):
        W, b = layers[l]
        gW = dz.T @ hs[l]
        gb = dz.sum(axis=0)
        grads.append((gW, gb))
        if l > 0:
            dh = dz @ W
            dz = dh * (1 - hs[l] ** 2)
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat
