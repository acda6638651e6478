"""Segment plan of SAQ (SaqDataMaker::analyze_plan, saqlib/quantization/saq_data.hpp:103-227),
computed by the host library (saqgpu_make_plan)."""
import ctypes as C

import numpy as np

from . import lib, SaqGpuError, MAX_SEG


def make_plan(var, num_dim, avg_bits, enable_segmentation=True, seg_eqseg=0):
    L = lib()
    f = L.saqgpu_make_plan
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int32, C.c_float, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
    v = np.zeros(num_dim, np.float32)
    v[:] = np.asarray(var, np.float32)[:num_dim]
    dims = np.zeros(MAX_SEG, np.int32)
    bits = np.zeros(MAX_SEG, np.int32)
    n = C.c_int32(0)
    rc = f(v.ctypes.data, num_dim, avg_bits, int(enable_segmentation), seg_eqseg, dims.ctypes.data,
           bits.ctypes.data, C.byref(n))
    if rc != 0:
        raise SaqGpuError(f"saqgpu_make_plan failed ({rc})")
    return [(int(dims[i]), int(bits[i])) for i in range(n.value)]
