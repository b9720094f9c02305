"""dacti_mlp_forward_noalloc (the entry point that allocates nothing)
against dacti_mlp_forward: equal bit for bit, on two calls in a row (the
second smaller, so stale scratch contents from the first would show), plus
its argument checks.  Device buffers come from the HIP runtime the library
itself uses."""
import ctypes as C

import numpy as np
import pytest

from dacti_amd.mlp import MLP

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return hip


H2D, D2H = 1, 2


def _dev(hip, nbytes):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), nbytes) == 0
    return p.value


def test_forward_noalloc_equals_forward():
    hip = _hip()
    m = MLP(66, 128, 3, 3, seed=5)
    g = np.random.default_rng(7)
    p = m.get_params()
    m.set_params((p + 0.05 * g.standard_normal(p.shape)).astype(np.float32))
    # 20000 rows: three 8192-row chunks, the last one ragged; then 300 rows
    for n in (20000, 300):
        print(f"[no-alloc] forward_noalloc n={n}", flush=True)
        x = g.standard_normal((n, 66)).astype(np.float32)
        want = m.forward(x)
        dx, dz = _dev(hip, x.nbytes), _dev(hip, n * 3 * 4)
        try:
            assert hip.hipMemcpy(dx, x.ctypes.data, x.nbytes, H2D) == 0
            assert hip.hipMemset(dz, 0xff, n * 3 * 4) == 0
            m.forward_noalloc(dx, n, dz)
            got = np.empty((n, 3), np.float32)
            assert hip.hipMemcpy(got.ctypes.data, dz, got.nbytes, D2H) == 0
        finally:
            hip.hipFree(dx)
            hip.hipFree(dz)
        np.testing.assert_array_equal(got, want)
    with pytest.raises(RuntimeError, match="invalid arguments"):
        m.forward_noalloc(0, 10, 0)
    m.forward_noalloc(0, 0, 0)  # nothing to do
    m.close()
