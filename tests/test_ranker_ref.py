"""Pins the FP64 ranker reference (oracle/mlp_ref.py) that the GPU ranker
kernels are checked against: gradients by central finite differences, the
loss by closed forms, and BF16 rounding by hand-worked ties and by torch's own
float32 -> bfloat16 conversion (round to nearest even)."""
import numpy as np
import pytest
import torch

import mlp_ref


def test_layout():
    assert mlp_ref.P == 64 * 24 + 64 + 64 + 1 == 1665
    flat = np.arange(mlp_ref.P, dtype=np.float64)
    w1, b1, w2, b2 = mlp_ref.split(flat)
    assert w1.shape == (64, 24) and w1[1, 0] == 24 and b1[0] == 1536 and w2[0] == 1600 and b2 == 1664


def test_loss_closed_forms():
    X = np.random.default_rng(0).normal(size=(7, 24))
    flat = np.zeros(mlp_ref.P)
    loss, g = mlp_ref.loss_and_grad(flat, X, np.ones(7))
    assert loss == pytest.approx(np.log(2.0))          # z = 0: softplus(0) = log 2
    assert g[-1] == pytest.approx(-0.5)                # dL/db2 = sigmoid(0) - 1
    flat[-1] = 3.0                                     # constant logit 3
    loss, g = mlp_ref.loss_and_grad(flat, X, np.zeros(7))
    assert loss == pytest.approx(np.log1p(np.exp(3.0)))
    assert g[-1] == pytest.approx(1 / (1 + np.exp(-3.0)))
    assert np.all(g[:-1] == 0.0)                       # all hidden units dead (zero input weights, zero bias)


def test_forward_by_hand():
    flat = np.zeros(mlp_ref.P)
    flat[0] = 2.0                 # w1[0,0]
    flat[1536] = -1.0              # b1[0]
    flat[1600] = 0.5               # w2[0]
    flat[-1] = 0.25                # b2
    X = np.zeros((3, 24))
    X[:, 0] = [0.0, 1.0, 3.0]      # a0 = 2 x - 1 -> relu 0, 1, 5
    np.testing.assert_allclose(mlp_ref.forward(flat, X), [0.25, 0.75, 2.75])


def test_gradient_finite_differences():
    rng = np.random.default_rng(1)
    flat = rng.normal(0, 0.3, mlp_ref.P)
    X = rng.normal(size=(40, 24))
    y = (rng.random(40) < 0.4).astype(np.float64)
    _, g = mlp_ref.loss_and_grad(flat, X, y)
    eps = 1e-6
    for i in list(rng.choice(mlp_ref.P, 60, replace=False)) + [0, 1535, 1536, 1599, 1600, 1663, 1664]:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += eps
        fm[i] -= eps
        num = (mlp_ref.loss_and_grad(fp, X, y)[0] - mlp_ref.loss_and_grad(fm, X, y)[0]) / (2 * eps)
        assert g[i] == pytest.approx(num, rel=1e-5, abs=1e-8), i


def test_bf16_round_ties_and_torch():
    one = np.float32(1.0)
    cases = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5, 0.0], np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, 0.0], np.float32)
    np.testing.assert_array_equal(mlp_ref.bf16_round(cases), want)
    assert mlp_ref.bf16_round(np.array([one]))[0] == one
    r = np.random.default_rng(2).normal(0, 10, 100000).astype(np.float32)
    t = torch.from_numpy(r).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(mlp_ref.bf16_round(r), t)
