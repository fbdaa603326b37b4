"""Pins the CPU oracle of the learned sizing model (oracle/mlp.py), which the
GPU tests then use as the checker of the device's forward pass, gradients
and features.

* bf16 rounding: equal to torch's float32 -> bfloat16 conversion (round to
  nearest even) on random and tie values;
* analytic gradients: equal to central finite differences of the float64
  loss (no bf16 rounding), so a dropped term, a wrong sign or a transposed
  product fails;
* feature bins: single spheres at chosen offsets land in the bins the
  definition names.
"""
import numpy as np
import pytest
import torch

from oracle import mlp


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 3, 5000), rng.normal(0, 1e-3, 1000)]).astype(np.float32)
    # Exact ties: 1 + 2^-8 (even below), 1 + 3 * 2^-8 (odd below) in bf16 units.
    ties = np.array([1 + 2**-8, 1 + 3 * 2**-8, -(1 + 2**-8), 2.0 ** -130], np.float32)
    x = np.concatenate([x, ties])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mlp.bf16(x).view(np.uint32), want.view(np.uint32))
    assert mlp.bf16(np.float32(1 + 2**-8)) == 1.0 and mlp.bf16(np.float32(1 + 3 * 2**-8)) == np.float32(1 + 2**-6)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    n = 12
    P = rng.normal(0, 0.2, mlp.N_PARAMS)
    X = rng.uniform(0, 3, (n, 64))
    occ = rng.integers(0, 2, n)
    tgt = rng.uniform(0, 6, n)

    def loss(p):
        (a, b), _ = mlp.loss_and_grads(p, X, occ, tgt, bf16_ops=False)
        return a + b

    _, g = mlp.loss_and_grads(P, X, occ, tgt, bf16_ops=False)
    picks = np.concatenate([rng.choice(np.arange(lo, hi), k, replace=False)
                            for lo, hi, k in [(mlp.P_W1, mlp.P_B1, 20), (mlp.P_B1, mlp.P_W2, 8),
                                              (mlp.P_W2, mlp.P_B2, 20), (mlp.P_B2, mlp.P_W3, 8),
                                              (mlp.P_W3, mlp.P_B3, 12), (mlp.P_B3, mlp.N_PARAMS, 2)]])
    h = 1e-6
    for i in picks:
        e = np.zeros(mlp.N_PARAMS)
        e[i] = h
        fd = (loss(P + e) - loss(P - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_gradient_of_a_wrong_formula_is_caught():
    # The finite-difference check is sensitive: flipping the sign of the
    # load term's gradient must fail it.
    rng = np.random.default_rng(3)
    P = rng.normal(0, 0.2, mlp.N_PARAMS)
    X = rng.uniform(0, 3, (6, 64))
    occ, tgt = np.ones(6), rng.uniform(0, 6, 6)
    (a, b), g = mlp.loss_and_grads(P, X, occ, tgt, bf16_ops=False)
    e = np.zeros(mlp.N_PARAMS)
    e[mlp.P_B3 + 1] = 1e-6
    (a1, b1), _ = mlp.loss_and_grads(P + e, X, occ, tgt, bf16_ops=False)
    (a0, b0), _ = mlp.loss_and_grads(P - e, X, occ, tgt, bf16_ops=False)
    fd = ((a1 + b1) - (a0 + b0)) / 2e-6
    assert abs(fd - g[mlp.P_B3 + 1]) < 1e-4 * abs(fd)
    assert abs(fd + g[mlp.P_B3 + 1]) > 1e-2 * abs(fd)


@pytest.mark.parametrize("offset,shell,octant,polar", [
    ((1.0, 2.0, 0.5), 1, 7, 0), ((-3.0, 0.5, 6.0), 3, 6, 8), ((0.2, -0.1, -7.9), 3, 1, 8),
    ((-5.0, -0.5, 0.1), 2, 4, 0), ((0.3, 0.2, -0.1), 0, 3, 0)])
def test_feature_bins_of_single_spheres(offset, shell, octant, polar):
    # Brick 10 at 0.25 A, probe 1.4, largest expanded radius 3.2: L = 0.5 sqrt(3) 2.5 + 2.8 + 3.2.
    L = 0.5 * np.sqrt(3.0) * 10 * 0.25 + 2 * 1.4 + 3.2
    key = np.array([[2, -1, 3]])
    c = (key[0] + 0.5) * 10 * 0.25
    f = mlp.features([c + np.array(offset)], 1.4, 3.2, 10, 0.25, key)[0]
    bin_ = shell * 16 + octant + polar
    assert np.linalg.norm(offset) < L
    assert f[bin_] == 1.0 and f.sum() == 1.0
    far = mlp.features([c + np.array([L + 0.01, 0, 0])], 1.4, 3.2, 10, 0.25, key)[0]
    assert far.sum() == 0.0
