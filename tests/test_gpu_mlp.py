"""GPU MLP kernels (BF16 MFMA, FP32 accumulation) against the CPU oracle.

The oracle rounds the same operands to BF16 and accumulates in FP32, so the
two differ only by the order of the FP32 sums (and the BF16 roundings that
order occasionally flips); a wrong MFMA lane layout, transpose or stride
gives O(1) differences.  Tolerances (written in the asserts): logits within
1e-2 of their scale, loss within 1e-3 relative, gradient within 2e-2
relative L2.
"""
import numpy as np
import pytest

from dacti_amd.mlp import MLP
from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def _net(n_in=150, seed=0):
    print(f"[mlp] create n_in={n_in}", flush=True)
    return MLP(n_in, 128, 3, 3, seed=seed)


def test_forward_matches_oracle():
    m = _net()
    p = m.get_params()
    g = np.random.default_rng(1)
    p = (p + 0.05 * g.standard_normal(p.shape)).astype(np.float32)  # non-zero biases too
    m.set_params(p)
    np.testing.assert_array_equal(m.get_params(), p)
    x = g.standard_normal((257, 150)).astype(np.float32)
    z = m.forward(x)
    zr, _ = mlp_ref.forward(p, m.layer_shapes(), x)
    assert z.shape == (257, 3)
    assert np.abs(z - zr).max() <= 1e-2 * np.abs(zr).max()
    m.close()


def test_loss_grad_matches_oracle():
    m = _net(n_in=70, seed=3)
    p = m.get_params()
    g = np.random.default_rng(2)
    x = g.standard_normal((200, 70)).astype(np.float32)
    y = g.integers(0, 3, 200).astype(np.int32)
    loss, grad = m.loss_grad(x, y)
    lr, gr = mlp_ref.loss_grad(p, m.layer_shapes(), x, y)
    assert abs(loss - lr) <= 1e-3 * abs(lr)
    assert np.linalg.norm(grad - gr) <= 2e-2 * np.linalg.norm(gr)
    m.close()


def test_training_learns_a_linear_teacher():
    g = np.random.default_rng(7)
    T = g.standard_normal((64, 3))
    x = g.standard_normal((24000, 64)).astype(np.float32)
    y = (x @ T).argmax(axis=1).astype(np.int32)
    m = _net(n_in=64, seed=5)
    l0, _ = m.loss_grad(x[:2048], y[:2048])
    loss = m.train(x[:20000], y[:20000], epochs=8, batch=512, lr=2e-3, seed=1)
    acc = (m.predict(x[20000:]) == y[20000:]).mean()
    ms, fl = m.train_stats()
    print(f"[mlp] loss {l0:.3f} -> {loss:.3f}, held-out accuracy {acc:.3f}, {fl / ms / 1e9:.2f} TFLOP/s", flush=True)
    assert loss < 0.5 * l0 and acc > 0.9
    m.close()


def test_adam_steps_match_reference():
    # every sample identical, so every batch has the same mean gradient and
    # the trajectory does not depend on the shuffle: 2 epochs of 4 batches
    # (the first step eager, the others replayed from the step's graph) are
    # 8 Adam steps on the oracle's gradient.  Adam (torch/Kingma-Ba form, the
    # one dacti_mlp_train documents): m = 0.9 m + 0.1 g, v = 0.999 v + 0.001
    # g^2, p -= lr (m / (1 - 0.9^t)) / (sqrt(v / (1 - 0.999^t)) + 1e-8).
    m = _net(n_in=40, seed=11)
    g = np.random.default_rng(4)
    p0 = (m.get_params() + 0.05 * g.standard_normal(m.get_params().shape)).astype(np.float32)
    m.set_params(p0)
    x = np.repeat(g.standard_normal((1, 40)).astype(np.float32), 256, axis=0)
    y = np.full(256, 2, dtype=np.int32)
    lr = 1e-3
    m.train(x, y, epochs=2, batch=64, lr=lr, seed=3)
    p_gpu = m.get_params().astype(np.float64)
    p = p0.astype(np.float64)
    mo = np.zeros_like(p)
    v = np.zeros_like(p)
    for t in range(1, 9):
        _, gr = mlp_ref.loss_grad(p.astype(np.float32), m.layer_shapes(), x[:16], y[:16])
        mo = 0.9 * mo + 0.1 * gr
        v = 0.999 * v + 0.001 * gr * gr
        p = p - lr * (mo / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
    d_gpu, d_ref = p_gpu - p0, p - p0
    print(f"[mlp] Adam: |update| {np.linalg.norm(d_ref):.3g}, |gpu - ref| {np.linalg.norm(d_gpu - d_ref):.3g}", flush=True)
    assert np.linalg.norm(d_gpu - d_ref) <= 0.1 * np.linalg.norm(d_ref)
    m.close()
