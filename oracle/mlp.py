"""CPU oracle of the learned sizing model (test infrastructure only).

The sizing model has no counterpart in the reference, so this module
restates the model the library defines. That definition is the header
comment of sesmesh_amd/csrc/dev_mlp.hpp and the smg_model_* section of
include/sesmesh_gpu.h. The restatement is in numpy float64:

* the per-brick features (sphere counts in 4 shells x 16 directions);
* the 64-128-128-2 ReLU network;
* its loss (BCE on occupancy + 0.5 occupied (y1 - log(1 + load))^2, mean
  over the rows) and the analytic gradients of that loss.

With ``bf16=True`` the operands are rounded where the device rounds them:
the inputs, W1 and W2, the layer-1 activations, and the back-propagated
dH2 and dH1 feeding the matrix-core products. Accumulation stays float64,
so the device differs only by FP32 accumulation order and the roundings
that order moves.

How each part is pinned (tests/test_oracle_mlp.py):
* bf16 rounding matches torch's float32 -> bfloat16 conversion;
* the gradients match central finite differences of the float64 loss;
* the feature bins of single spheres placed at known offsets are asserted.
Only the tests use this module; the product never imports it.
"""
import numpy as np

N_IN, N_H, N_OUT = 64, 128, 2
P_W1 = 0
P_B1 = P_W1 + N_IN * N_H
P_W2 = P_B1 + N_H
P_B2 = P_W2 + N_H * N_H
P_W3 = P_B2 + N_H
P_B3 = P_W3 + N_H * N_OUT
N_PARAMS = P_B3 + N_OUT  # 25090


def bf16(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def unpack(P):
    P = np.asarray(P, np.float64)
    return (P[P_W1:P_B1].reshape(N_IN, N_H), P[P_B1:P_W2], P[P_W2:P_B2].reshape(N_H, N_H), P[P_B2:P_W3],
            P[P_W3:P_B3].reshape(N_H, N_OUT), P[P_B3:N_PARAMS])


def _r(x, on):
    return bf16(np.asarray(x, np.float32)).astype(np.float64) if on else np.asarray(x, np.float64)


def forward(P, X, bf16_ops=True):
    """Outputs (n x 2) and the activations the backward pass needs."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    Xb = _r(X, bf16_ops)
    W1b, W2b = _r(W1, bf16_ops), _r(W2, bf16_ops)
    h1 = _r(np.maximum(Xb @ W1b + b1, 0.0), bf16_ops)
    h2 = np.maximum(h1 @ W2b + b2, 0.0)
    y = h2 @ W3 + b3
    return y, (Xb, W1b, W2b, h1, h2)


def loss_and_grads(P, X, occ, target, lam=1.0, bf16_ops=True):
    """(bce, mse) mean loss terms and the gradient of their sum w.r.t. P."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    n = len(X)
    y, (Xb, W1b, W2b, h1, h2) = forward(P, X, bf16_ops)
    t0 = np.asarray(occ, np.float64)
    t1 = np.asarray(target, np.float64)
    z0, z1 = y[:, 0], y[:, 1]
    bce = np.maximum(z0, 0) - z0 * t0 + np.log1p(np.exp(-np.abs(z0)))
    e = z1 - t1
    mse = 0.5 * lam * t0 * e * e
    dy = np.stack([(1.0 / (1.0 + np.exp(-z0)) - t0) / n, lam * t0 * e / n], axis=1)
    g = np.zeros(N_PARAMS)
    g[P_W3:P_B3] = (h2.T @ dy).ravel()
    g[P_B3:] = dy.sum(0)
    dh2 = (dy @ W3.T) * (h2 > 0)
    g[P_B2:P_W3] = dh2.sum(0)
    dh2b = _r(dh2, bf16_ops)
    g[P_W2:P_B2] = (h1.T @ dh2b).ravel()
    dh1 = (dh2b @ W2b.T) * (h1 > 0)
    g[P_B1:P_W2] = dh1.sum(0)
    g[P_W1:P_B1] = (Xb.T @ _r(dh1, bf16_ops)).ravel()
    return (bce.mean(), mse.mean()), g


def feature_bins(d, L):
    """Bin (shell * 16 + direction) of offsets d (n x 3) within reach L."""
    dist = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    shell = np.minimum(np.floor(4.0 * dist / L), 3).astype(np.int64)
    octant = (d[:, 0] >= 0) * 1 + (d[:, 1] >= 0) * 2 + (d[:, 2] >= 0) * 4
    ad = np.abs(d)
    polar = ((ad[:, 2] > ad[:, 0]) & (ad[:, 2] > ad[:, 1])) * 8
    return shell * 16 + octant + polar, dist


def features(centers, probe, max_expanded, brick, spacing, keys):
    """Features (n x 64 float32) of bricks `keys` (n x 3 lattice brick keys)."""
    centers = np.asarray(centers, np.float64)
    L = 0.5 * np.sqrt(3.0) * brick * spacing + 2.0 * probe + max_expanded
    out = np.zeros((len(keys), N_IN), np.float32)
    for i, k in enumerate(np.asarray(keys, np.int64)):
        c = (k.astype(np.float64) + 0.5) * brick * spacing
        bins, dist = feature_bins(centers - c, L)
        cnt = np.bincount(bins[dist < L], minlength=N_IN)
        out[i] = np.sqrt(cnt.astype(np.float32))
    return out
