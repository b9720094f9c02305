"""Device arrays for the predictor's device-resident path: hipMalloc'd
buffers with their shape and dtype, copies to and from numpy, and freeing.
The HIP runtime is the one libdacti_gpu.so links (libamdhip64)."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from . import _lib
        _lib.load()  # the engine's library brings the HIP runtime it was linked with
        h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


H2D, D2H, D2D = 1, 2, 3


class DeviceArray:
    """An uninitialised device array of `shape` and numpy `dtype`."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        if self.nbytes:
            rc = hip().hipMalloc(C.byref(self.ptr), self.nbytes)
            if rc != 0:
                raise MemoryError(f"hipMalloc of {self.nbytes} bytes failed ({rc})")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes and hip().hipMemcpy(d.ptr, a.ctypes.data, d.nbytes, H2D) != 0:
            raise RuntimeError("hipMemcpy to the device failed")
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes and hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) != 0:
            raise RuntimeError("hipMemcpy from the device failed")
        return out

    def __len__(self):
        return self.shape[0]

    @property
    def address(self):
        return self.ptr.value or 0

    def free(self):
        if self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
