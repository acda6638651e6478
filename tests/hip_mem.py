"""Device buffers for tests of the SAQGPU_DEVICE_PTRS paths, allocated through the same HIP
runtime libsaqgpu.so links (/opt/rocm), so no second runtime is started in the process."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipStreamSynchronize.argtypes = [C.c_void_p]
        L.hipStreamDestroy.argtypes = [C.c_void_p]
        _hip = L
    return _hip


class DevArray:
    """A device copy of a numpy array (or zeros of a shape and dtype)."""

    def __init__(self, a=None, shape=None, dtype=None):
        if a is not None:
            a = np.ascontiguousarray(a)
            shape, dtype = a.shape, a.dtype
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(1, self.nbytes)) == 0
        if a is not None:
            assert hip().hipMemcpy(self.p, a.ctypes.data, self.nbytes, 1) == 0  # host to device
        else:
            z = np.zeros(self.shape, self.dtype)
            assert hip().hipMemcpy(self.p, z.ctypes.data, self.nbytes, 1) == 0

    def data_ptr(self):
        return self.p.value

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.p, self.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()
