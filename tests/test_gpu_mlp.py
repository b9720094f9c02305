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
