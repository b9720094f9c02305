"""CPU pins of the MLP oracle (oracle/mlp_ref.py).

BF16 rounding against torch's own float32 -> bfloat16 conversion (round to
nearest even), and the backward pass against central finite differences of
the exact-arithmetic network, so that a dropped term, a wrong transpose or a
wrong ReLU mask in the restatement fails here before it is used to check the
GPU kernels.
"""
import numpy as np
import torch

from oracle import mlp_ref


def test_bf16_matches_torch_rounding():
    g = np.random.default_rng(3)
    x = np.concatenate([
        g.standard_normal(5000).astype(np.float32) * 10.0 ** g.integers(-30, 30, 5000),
        np.array([0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, np.inf, -np.inf,
                  3.4e38, 1e-40, -1e-41, 65504.0], np.float32)]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(mlp_ref.bf16(x), ref)
    assert np.isnan(mlp_ref.bf16(np.array([np.nan], np.float32)))[0]


def test_backward_matches_finite_differences():
    shapes = [(5, 7), (5, 5), (5, 5), (3, 5)]
    p = mlp_ref.init_params(shapes, seed=1)
    p[[35, 36, 37]] = [0.1, -0.2, 0.05]  # non-zero biases
    g = np.random.default_rng(2)
    x = g.standard_normal((6, 7))
    y = g.integers(0, 3, 6)
    loss, grad = mlp_ref.loss_grad(p, shapes, x, y, rounding=False)
    eps = 1e-6
    fd = np.empty_like(grad)
    for k in range(len(p)):
        q = p.copy()
        q[k] += eps
        lp, _ = mlp_ref.loss_grad(q, shapes, x, y, rounding=False)
        q[k] -= 2 * eps
        lm, _ = mlp_ref.loss_grad(q, shapes, x, y, rounding=False)
        fd[k] = (lp - lm) / (2 * eps)
    assert np.abs(fd - grad).max() <= 1e-7 * max(1.0, np.abs(grad).max())
    # the loss itself: mean cross-entropy from the logits
    z, _ = mlp_ref.forward(p, shapes, x, rounding=False)
    lse = np.log(np.exp(z).sum(axis=1))
    assert abs(loss - np.mean(lse - z[np.arange(6), y])) < 1e-12


def test_rounded_network_is_close_to_exact():
    shapes = [(128, 150), (128, 128), (128, 128), (3, 128)]
    p = mlp_ref.init_params(shapes, seed=4)
    g = np.random.default_rng(5)
    x = g.standard_normal((64, 150))
    y = g.integers(0, 3, 64)
    l1, g1 = mlp_ref.loss_grad(p, shapes, x, y, rounding=True)
    l0, g0 = mlp_ref.loss_grad(p, shapes, x, y, rounding=False)
    assert abs(l1 - l0) < 1e-2 * abs(l0)
    assert np.linalg.norm(g1 - g0) < 0.1 * np.linalg.norm(g0)  # BF16 keeps about 3 digits
