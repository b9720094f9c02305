"""Refinement-predictor MLP on the GPU (ctypes binding of include/dacti_mlp.h).

n_in -> n_hidden (ReLU) x n_layers -> n_out logits, softmax cross-entropy;
FP32 master parameters, BF16 MFMA GEMMs with FP32 accumulation.  The flat
parameter layout is, per layer, W (out x in, row-major) then b (out).
"""
import ctypes as C

import numpy as np

from . import _lib

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)

SIGNATURES = {
    "dacti_mlp_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "dacti_mlp_n_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "dacti_mlp_get_params": (C.c_int, [C.c_void_p, _F]),
    "dacti_mlp_set_params": (C.c_int, [C.c_void_p, _F]),
    "dacti_mlp_forward": (C.c_int, [C.c_void_p, _F, C.c_int64, _F]),
    "dacti_mlp_forward_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "dacti_mlp_loss_grad": (C.c_int, [C.c_void_p, _F, _I, C.c_int64, _F, _F]),
    "dacti_mlp_train": (C.c_int, [C.c_void_p, _F, _I, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _F]),
    "dacti_mlp_train_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "dacti_mlp_destroy": (C.c_int, [C.c_void_p]),
    "dacti_mlp_last_error": (C.c_char_p, [C.c_void_p]),
}


def _load():
    lib = _lib.load()
    if not getattr(lib, "_mlp_bound", False):
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib._mlp_bound = True
    return lib


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        assert a.shape == shape, (a.shape, shape)
    return a


class MLP:
    def __init__(self, n_in, n_hidden=128, n_layers=3, n_out=3, seed=0):
        self.lib = _load()
        self.n_in, self.n_hidden, self.n_layers, self.n_out = n_in, n_hidden, n_layers, n_out
        h = C.c_void_p()
        rc = self.lib.dacti_mlp_create(n_in, n_hidden, n_layers, n_out, seed, C.byref(h))
        if rc != 0:
            raise RuntimeError(f"dacti_mlp_create failed ({rc}): {self.lib.dacti_mlp_last_error(None).decode()}")
        self.h = h
        n = C.c_int64()
        self._check(self.lib.dacti_mlp_n_params(self.h, C.byref(n)), "n_params")
        self.n_params = n.value

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.lib.dacti_mlp_last_error(self.h).decode()}")

    def layer_shapes(self):
        dims = [self.n_in] + [self.n_hidden] * self.n_layers + [self.n_out]
        return [(dims[l + 1], dims[l]) for l in range(len(dims) - 1)]

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        self._check(self.lib.dacti_mlp_get_params(self.h, p.ctypes.data_as(_F)), "get_params")
        return p

    def set_params(self, p):
        p = _f32(p, (self.n_params,))
        self._check(self.lib.dacti_mlp_set_params(self.h, p.ctypes.data_as(_F)), "set_params")

    def forward(self, x):
        x = _f32(x)
        n = x.shape[0]
        out = np.empty((n, self.n_out), np.float32)
        self._check(self.lib.dacti_mlp_forward(self.h, x.ctypes.data_as(_F), n, out.ctypes.data_as(_F)), "forward")
        return out

    def forward_noalloc(self, x_dev, n, logits_dev):
        """forward() on device arrays (addresses as ints) without any
        allocation: x_dev holds n x n_in FP32 values, logits_dev receives
        n x n_out FP32 values."""
        self._check(self.lib.dacti_mlp_forward_noalloc(self.h, C.c_void_p(x_dev), n, C.c_void_p(logits_dev)),
                    "forward_noalloc")

    def predict(self, x):
        return self.forward(x).argmax(axis=1)

    def loss_grad(self, x, y):
        x = _f32(x)
        y = np.ascontiguousarray(y, dtype=np.int32)
        loss = C.c_float()
        g = np.empty(self.n_params, np.float32)
        self._check(self.lib.dacti_mlp_loss_grad(self.h, x.ctypes.data_as(_F), y.ctypes.data_as(_I), x.shape[0],
                                                 C.byref(loss), g.ctypes.data_as(_F)), "loss_grad")
        return float(loss.value), g

    def train(self, x, y, epochs=10, batch=1024, lr=1e-3, seed=0):
        x = _f32(x)
        y = np.ascontiguousarray(y, dtype=np.int32)
        loss = C.c_float()
        self._check(self.lib.dacti_mlp_train(self.h, x.ctypes.data_as(_F), y.ctypes.data_as(_I), x.shape[0], epochs,
                                             batch, lr, seed, C.byref(loss)), "train")
        return float(loss.value)

    def train_stats(self):
        ms, fl = C.c_double(), C.c_double()
        self._check(self.lib.dacti_mlp_train_stats(self.h, C.byref(ms), C.byref(fl)), "train_stats")
        return ms.value, fl.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.dacti_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
