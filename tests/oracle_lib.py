"""ctypes bindings for the checker libraries (test infrastructure only).

``liboracle.so`` is the oracle (CPU restatement of the reference hot path);
``_ref/libsaqref.so`` is the reference's own glog-free headers compiled with its
release flags.  Both are built by ``oracle/Makefile``.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsaqref.so")

MAX_SEG = 64
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
i64 = C.c_int64


class Plan(C.Structure):
    _fields_ = [("num_seg", C.c_int32), ("num_dim", C.c_int32), ("use_compact_layout", C.c_int32),
                ("seg_dim", C.c_int32 * MAX_SEG), ("seg_bits", C.c_int32 * MAX_SEG)]


class Cluster(C.Structure):
    _fields_ = [("num_vec", C.c_int64), ("short_factors", C.c_void_p), ("short_code", C.c_void_p),
                ("long_code", C.c_void_p), ("long_factors", C.c_void_p), ("ids", C.c_void_p),
                ("centroid_rot", C.c_void_p)]


class Sizes(C.Structure):
    _fields_ = [("num_blocks", C.c_int64), ("short_factor_floats", C.c_int64),
                ("short_code_bytes", C.c_int64), ("long_code_bytes", C.c_int64),
                ("long_stride", C.c_int64), ("long_factor_floats", C.c_int64),
                ("centroid_floats", C.c_int64)]


class LutScalars(C.Structure):
    _fields_ = [("delta", C.c_float), ("sum_q", C.c_float), ("sum_vl_lut", C.c_float),
                ("q_l2sqr", C.c_float), ("q_l2norm", C.c_float)]


class CaqCfg(C.Structure):
    _fields_ = [("rd_lmt", C.c_int32), ("eps", C.c_float), ("rotate", C.c_int32), ("ori_qb", C.c_int32)]


class CodeInfo(C.Structure):
    _fields_ = [("o_l2norm", C.c_float), ("ip_cent_oa", C.c_float), ("fac_rescale", C.c_float),
                ("fac_error", C.c_float), ("ip_o_oa", C.c_double), ("oa_l2sqr", C.c_double),
                ("delta", C.c_double), ("v_mi", C.c_double), ("rounds", C.c_int32), ("status", C.c_int32)]


class SingleLayout(C.Structure):
    _fields_ = [("fac_off", C.c_int64), ("code_off", C.c_int64), ("long_off", C.c_int64),
                ("exfac_off", C.c_int64), ("bytes", C.c_int64), ("seg_code_off", C.c_int64 * MAX_SEG),
                ("seg_long_off", C.c_int64 * MAX_SEG)]


class Metrics(C.Structure):
    _fields_ = [("fast_bitsum", C.c_uint64), ("acc_bitsum", C.c_uint64), ("total_comp_cnt", C.c_uint64)]


_orc = None
_ref = None


CPUBASE_SO = os.path.join(ROOT, "oracle", "libcpubase.so")
_cpub = None


def host_has_avx512():
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    return all(f in flags for f in ("avx512f", "avx512bw", "avx512vl", "avx512dq"))


def cpubase():
    """The oracle with its fast scan on AVX-512 (the benchmark's CPU baseline), or None where the
    host has no AVX-512 or the library is not built."""
    global _cpub
    if _cpub is None:
        if not host_has_avx512() or not os.path.exists(CPUBASE_SO):
            return None
        _cpub = _bind(C.CDLL(CPUBASE_SO))
    return _cpub


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run make -C oracle")
        _orc = _bind(C.CDLL(ORACLE_SO))
    return _orc


def _bind(L):
    L.orc_eigen_sum.restype = C.c_float
    L.orc_eigen_sum.argtypes = [f32p, i64]
    L.orc_eigen_sqnorm.restype = C.c_float
    L.orc_eigen_sqnorm.argtypes = [f32p, i64]
    L.orc_eigen_dot.restype = C.c_float
    L.orc_eigen_dot.argtypes = [f32p, f32p, i64]
    L.orc_eigen_diff_sqnorm.restype = C.c_float
    L.orc_eigen_diff_sqnorm.argtypes = [f32p, f32p, i64]
    L.orc_eigen_var_q2.restype = C.c_float
    L.orc_eigen_var_q2.argtypes = [f32p, f32p, i64]
    L.orc_gemv_rowvec.argtypes = [f32p, f32p, i64, f32p]
    L.orc_tree16.restype = C.c_float
    L.orc_tree16.argtypes = [f32p]
    L.orc_pack_lut.argtypes = [i64, f32p, f32p]
    L.orc_pack_codes.argtypes = [i64, u8p, i64, u8p]
    L.orc_lut_prepare.argtypes = [i64, f32p, u16p, C.POINTER(LutScalars)]
    L.orc_accumulate.argtypes = [u8p, u16p, i64, u32p]
    L.orc_cvt_u16.restype = C.c_uint16
    L.orc_cvt_u16.argtypes = [C.c_float]
    L.orc_compute_ip.restype = C.c_float
    L.orc_compute_ip.argtypes = [C.c_int, f32p, u8p, i64]
    L.orc_compact_code.argtypes = [C.c_int, u16p, i64, u8p]
    L.orc_code_at.restype = C.c_int
    L.orc_code_at.argtypes = [C.c_int, u8p, i64, i64]
    L.orc_caq_encode.restype = C.c_int
    L.orc_caq_encode.argtypes = [f32p, i64, C.c_int, C.POINTER(CaqCfg), C.c_void_p, u16p,
                                 C.POINTER(CodeInfo)]
    L.orc_rotate_rows.argtypes = [f32p, i64, i64, C.c_void_p, f32p, f32p]
    L.orc_quantize_segment.restype = C.c_int
    L.orc_quantize_segment.argtypes = [C.POINTER(Plan), C.c_int, f32p, i64, f32p, C.c_void_p,
                                       C.POINTER(CaqCfg), C.c_void_p, C.c_void_p]
    L.orc_cluster_sizes.argtypes = [C.POINTER(Plan), i64, C.POINTER(Sizes)]
    L.orc_plan_dp.restype = C.c_int
    L.orc_plan_dp.argtypes = [f32p, i64, C.c_float, i32p, i32p]
    L.orc_plan_eqseg.restype = C.c_int
    L.orc_plan_eqseg.argtypes = [i64, C.c_float, C.c_int, i32p, i32p]
    L.orc_probe.argtypes = [f32p, i64, i64, f32p, i64, C.c_int, u32p, C.c_void_p]
    L.orc_search.restype = C.c_int
    L.orc_search.argtypes = [C.POINTER(Plan), f32p, C.c_void_p, C.c_void_p, u32p, i64, f32p, i64,
                             C.c_float, C.c_int, u32p, C.c_void_p, C.c_void_p]
    L.orc_search_batch.restype = C.c_int
    L.orc_search_batch.argtypes = [C.POINTER(Plan), f32p, C.c_void_p, C.c_void_p, u32p, i64, f32p,
                                   i64, i64, C.c_float, C.c_int, C.c_int, u32p, C.c_void_p,
                                   C.c_void_p]
    L.orc_estimate.restype = C.c_int
    L.orc_estimate.argtypes = [C.POINTER(Plan), f32p, C.c_void_p, C.c_void_p, u32p, i64, f32p,
                               C.c_float, C.c_int, u32p, f32p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_decode_cluster.argtypes = [C.POINTER(Plan), C.c_void_p, C.c_int, i32p, f32p, f32p]
    L.orc_estimate_lanes.restype = i64
    L.orc_estimate_lanes.argtypes = [C.POINTER(Plan), f32p, C.c_void_p, C.c_void_p, u32p, i64, f32p,
                                     C.c_float, f32p, f32p, f32p]
    L.orc_single_layout.argtypes = [C.POINTER(Plan), C.POINTER(SingleLayout)]
    L.orc_single_transpose_bin.argtypes = [u16p, i64, u64p]
    L.orc_single_warmup_ip.restype = C.c_float
    L.orc_single_warmup_ip.argtypes = [u64p, u64p, C.c_float, C.c_float, i64]
    L.orc_single_mask_ip.restype = C.c_float
    L.orc_single_mask_ip.argtypes = [f32p, u64p, i64]
    L.orc_single_quantize.restype = C.c_int64
    L.orc_single_quantize.argtypes = [C.POINTER(Plan), C.c_void_p, f32p, i64, C.POINTER(CaqCfg), u8p, i64]
    L.orc_single_estimate.argtypes = [C.POINTER(Plan), f32p, C.c_void_p, u8p, i64, i64, f32p, C.c_float,
                                      C.c_int, f32p, f32p, f32p]
    L.orc_search_batch_timed.argtypes = [C.POINTER(Plan), f32p, C.c_void_p, C.c_void_p, u32p, i64, C.c_void_p,
                                         f32p, i64, i64, C.c_float, C.c_int, u32p, C.c_void_p,
                                         np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")]
    return L


def ref():
    """The reference's own kernels, or None where they are not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            return None
        try:
            L = C.CDLL(REF_SO)
        except OSError:
            return None
        L.ref_pack_codes.argtypes = [i64, u8p, i64, u8p]
        L.ref_pack_lut.argtypes = [i64, f32p, f32p]
        L.ref_accumulate.argtypes = [u8p, u16p, i64, i32p]
        L.ref_compute_ip.restype = C.c_float
        L.ref_compute_ip.argtypes = [C.c_int, f32p, u8p, i64]
        L.ref_compact_code.argtypes = [C.c_int, u16p, i64, u8p]
        L.ref_eigen_sum.restype = C.c_float
        L.ref_eigen_sum.argtypes = [f32p, i64]
        L.ref_eigen_sqnorm.restype = C.c_float
        L.ref_eigen_sqnorm.argtypes = [f32p, i64]
        L.ref_eigen_dot_rowcol.restype = C.c_float
        L.ref_eigen_dot_rowcol.argtypes = [f32p, f32p, i64]
        L.ref_eigen_diff_sqnorm.restype = C.c_float
        L.ref_eigen_diff_sqnorm.argtypes = [f32p, i64, i64, f32p, i64]
        L.ref_eigen_var_q2.restype = C.c_float
        L.ref_eigen_var_q2.argtypes = [f32p, i64, i64, f32p, i64]
        L.ref_eigen_gemv.argtypes = [f32p, f32p, i64, f32p]
        L.ref_eigen_colwise_mean_var.argtypes = [f32p, i64, i64, f32p, f32p]
        L.ref_eigen_gemm.argtypes = [f32p, i64, i64, f32p, f32p]
        L.ref_lut_quantize.argtypes = [f32p, i64, C.c_float, C.c_float, u16p]
        L.ref_lut_delta.restype = C.c_float
        L.ref_lut_delta.argtypes = [C.c_float, C.c_float]
        L.ref_code_floor.argtypes = [f32p, i64, C.c_double, C.c_double, C.c_int, f32p]
        L.ref_config_layout.argtypes = [C.c_void_p]
        L.ref_new_transpose_bin.argtypes = [u16p, i64, u64p]
        L.ref_warmup_ip_x0_q.restype = C.c_float
        L.ref_warmup_ip_x0_q.argtypes = [u64p, u64p, C.c_float, C.c_float, i64]
        L.ref_mask_ip_x0_q.restype = C.c_float
        L.ref_mask_ip_x0_q.argtypes = [f32p, u64p, i64]
        L.ref_single_delta.restype = C.c_float
        L.ref_single_delta.argtypes = [C.c_float, C.c_float]
        _ref = L
    return _ref


def make_plan(dims, bits, num_dim, compact=False):
    p = Plan()
    p.num_seg = len(dims)
    p.num_dim = num_dim
    p.use_compact_layout = 1 if compact else 0
    for i, (d, b) in enumerate(zip(dims, bits)):
        p.seg_dim[i] = int(d)
        p.seg_bits[i] = int(b)
    return p
