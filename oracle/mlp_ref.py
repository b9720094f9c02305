"""CPU restatement of the refinement-predictor MLP (TEST INFRASTRUCTURE).

Only tests/ may use this module, as the checker of the GPU kernels in
dacti_amd/csrc/mlp.hip; nothing in the product path imports it.  The
network has no counterpart in the reference (whose refinement is the
weighted-Laplacian estimator, integrator/error_estimators.hpp:10-47).

forward/backward follow the GPU's arithmetic: every GEMM operand is rounded
to BF16 (round to nearest, ties to even, on the FP32 bit pattern) and the
products are accumulated in FP32; with rounding=False and float64 inputs the
same code is the exact-arithmetic network, which tests pin against central
finite differences.
"""
import numpy as np


def bf16(a):
    """FP32 -> BF16 (round to nearest even), returned as FP32 values."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def unpack(params, shapes):
    """Flat parameters -> [(W, b)], W of shape (out, in)."""
    out, k = [], 0
    for o, i in shapes:
        W = params[k:k + o * i].reshape(o, i)
        k += o * i
        b = params[k:k + o]
        k += o
        out.append((W, b))
    assert k == len(params)
    return out


def forward(params, shapes, x, rounding=True):
    """Returns (logits, activations): activations[l] is the input of layer l."""
    dt = np.float32 if rounding else np.float64
    r = bf16 if rounding else (lambda a: a)
    layers = unpack(np.asarray(params, dt), shapes)
    a = np.asarray(x, dt)
    acts = []
    for l, (W, b) in enumerate(layers):
        acts.append(a)
        z = (r(a) @ r(W).T).astype(dt) + b
        a = z if l == len(layers) - 1 else np.maximum(z, 0)
    return a, acts


def loss_grad(params, shapes, x, y, rounding=True):
    """Mean softmax cross-entropy and its gradient (flat, same layout)."""
    dt = np.float32 if rounding else np.float64
    r = bf16 if rounding else (lambda a: a)
    layers = unpack(np.asarray(params, dt), shapes)
    z, acts = forward(params, shapes, x, rounding)
    n = len(y)
    mx = z.max(axis=1, keepdims=True)
    e = np.exp(z - mx)
    s = e.sum(axis=1, keepdims=True)
    loss = float(np.mean((mx[:, 0] + np.log(s[:, 0])) - z[np.arange(n), y]))
    dz = (e / s).astype(dt)
    dz[np.arange(n), y] -= 1
    dz = (dz / n).astype(dt)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        gW = (r(dz).T @ r(acts[l])).astype(dt)
        gb = dz.sum(axis=0).astype(dt)
        grads[l] = (gW, gb)
        if l > 0:
            da = (r(dz) @ r(W)).astype(dt)
            dz = np.where(acts[l] > 0, da, 0).astype(dt)
    flat = np.concatenate([np.concatenate([g.ravel(), b]) for g, b in grads])
    return loss, flat


def init_params(shapes, seed=0):
    g = np.random.default_rng(seed)
    parts = []
    for o, i in shapes:
        parts.append(g.uniform(-1 / np.sqrt(i), 1 / np.sqrt(i), size=o * i))
        parts.append(np.zeros(o))
    return np.concatenate(parts)
