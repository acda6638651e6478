"""saq_amd -- MI355X (gfx950) implementation of SAQ's IVF scan and CAQ encoder.

Python binding of the C ABI in include/saqgpu.h (libsaqgpu.so, built in-tree by
saq_amd/build.py).  There is no CPU fallback: if the library is missing or the
GPU call fails, the call raises.

The Index class mirrors the reference's saqlib::IVF surface (construct/search/
estimate/save/load, saqlib/index/ivf.hpp:84-100) on top of the C ABI.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsaqgpu.so")
MAX_SEG = 64
DEVICE_PTRS = 1
USE_GRAPH = 2  # include/saqgpu.h SAQGPU_USE_GRAPH


class SaqGpuError(RuntimeError):
    pass


class _Desc(C.Structure):
    _fields_ = [("num_data", C.c_int64), ("num_dim", C.c_int32), ("num_cen", C.c_int32),
                ("num_seg", C.c_int32), ("use_compact_layout", C.c_int32),
                ("seg_dim", C.c_void_p), ("seg_bits", C.c_void_p), ("rotations", C.c_void_p),
                ("variance", C.c_void_p), ("centroids", C.c_void_p), ("cluster_sizes", C.c_void_p),
                ("device", C.c_int32)]


class ClusterSizes(C.Structure):
    _fields_ = [("num_blocks", C.c_int64), ("short_factor_floats", C.c_int64),
                ("short_code_bytes", C.c_int64), ("long_code_bytes", C.c_int64),
                ("long_factor_floats", C.c_int64), ("centroid_floats", C.c_int64)]


class CaqCfg(C.Structure):
    _fields_ = [("caq_adj_rd_lmt", C.c_int32), ("caq_adj_eps", C.c_float), ("caq_ori_qB", C.c_int32)]


class SearchParams(C.Structure):
    _fields_ = [("topk", C.c_int32), ("nprobe", C.c_int32), ("searcher_vars_bound_m", C.c_float),
                ("dist_type", C.c_int32), ("nprobe_per_query", C.c_void_p)]


class BatchedParams(C.Structure):
    _fields_ = [("topk", C.c_int32), ("nprobe", C.c_int32), ("nprobe_per_query", C.c_void_p),
                ("precision", C.c_int32), ("query_tile", C.c_int32), ("exact_probe", C.c_int32)]


PREC_INT8 = 1
PREC_INT8X2 = 2
PREC_BF16 = 3
SAQGPU_NOALLOC_SCRATCH_FLOATS = 1 << 24  # include/saqgpu.h

_lib = None

_SIGS = {
    "saqgpu_last_error": (C.c_char_p, []),
    "saqgpu_index_create": (C.c_int, [C.POINTER(_Desc), C.POINTER(C.c_void_p)]),
    "saqgpu_index_destroy": (None, [C.c_void_p]),
    "saqgpu_index_search_handle": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "saqgpu_cluster_sizes_of": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(ClusterSizes)]),
    "saqgpu_index_set_cluster": (C.c_int, [C.c_void_p, C.c_int32] + [C.c_void_p] * 6),
    "saqgpu_index_get_cluster": (C.c_int, [C.c_void_p, C.c_int32] + [C.c_void_p] * 6),
    "saqgpu_encode": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(CaqCfg), C.c_int32,
                                C.c_void_p, C.c_void_p]),
    "saqgpu_encode_residuals": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                          C.POINTER(CaqCfg)]),
    "saqgpu_probe": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p, C.c_int32,
                               C.c_void_p]),
    "saqgpu_probe_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p,
                                       C.c_void_p]),
    "saqgpu_search": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(SearchParams),
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "saqgpu_estimate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_float, C.c_int32,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "saqgpu_prepare_batched": (C.c_int, [C.c_void_p]),
    "saqgpu_search_batched": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(BatchedParams),
                                        C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "saqgpu_autotune_batched": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(BatchedParams),
                                          C.c_int32, C.c_void_p, C.POINTER(C.c_int32)]),
    "saqgpu_single_layout": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_void_p]),
    "saqgpu_single_quantize": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(CaqCfg), C.c_void_p,
                                         C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]),
    "saqgpu_single_estimate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
                                         C.c_float, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                         C.c_void_p]),
    "saqgpu_index_device_bytes": (C.c_int64, [C.c_void_p]),
    "saqgpu_timing": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "saqgpu_make_plan": (C.c_int, [C.c_void_p, C.c_int32, C.c_float, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                   C.c_void_p]),
    "saqgpu_column_mean_var": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]),
    "saqgpu_encode_one_centroid": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                             C.POINTER(CaqCfg), C.c_int32, C.POINTER(C.c_int64), C.c_void_p]),
}


def lib():
    """The loaded C ABI.  Raises if the gfx950 library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SaqGpuError(f"{LIB_PATH} is missing: build it with python saq_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def column_mean_var(data):
    """Column means and variances of a host n x d float32 matrix in the reference's Eigen order
    (saqgpu_column_mean_var)."""
    data = np.ascontiguousarray(data, np.float32)
    n, d = data.shape
    mean, var = np.zeros(d, np.float32), np.zeros(d, np.float32)
    _check(lib().saqgpu_column_mean_var(data.ctypes.data, n, d, mean.ctypes.data, var.ctypes.data))
    return mean, var


def _check(rc):
    if rc != 0:
        raise SaqGpuError(f"saqgpu error {rc}: {lib().saqgpu_last_error().decode()}")


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return a.data_ptr()  # torch tensor on the device


class Index:
    """Device-resident SAQ IVF index (the state of saqlib::IVF, ivf.hpp:27-121)."""

    def __init__(self, num_dim, plan, centroids, cluster_sizes, variance, rotations=None,
                 use_compact_layout=False, device=0):
        self.num_dim = int(num_dim)
        self.plan = [(int(d), int(b)) for d, b in plan]
        self.centroids = np.ascontiguousarray(centroids, dtype=np.float32)
        self.num_cen = self.centroids.shape[0]
        self.cluster_sizes = np.ascontiguousarray(cluster_sizes, dtype=np.int64)
        self.num_data = int(self.cluster_sizes.sum())
        self.D_pad = sum(d for d, _ in self.plan)
        self.variance = np.zeros(self.D_pad, np.float32)
        self.variance[:len(variance)] = variance
        self.rotations = None
        if rotations is not None:
            self.rotations = np.ascontiguousarray(np.concatenate([np.asarray(r, np.float32).ravel()
                                                                  for r in rotations]))
        dims = np.array([d for d, _ in self.plan], np.int32)
        bits = np.array([b for _, b in self.plan], np.int32)
        self._keep = (dims, bits)
        desc = _Desc(self.num_data, self.num_dim, self.num_cen, len(self.plan), int(use_compact_layout),
                     dims.ctypes.data, bits.ctypes.data, _ptr(self.rotations), self.variance.ctypes.data,
                     self.centroids.ctypes.data, self.cluster_sizes.ctypes.data, device)
        h = C.c_void_p()
        _check(lib().saqgpu_index_create(C.byref(desc), C.byref(h)))
        self.h = h

    def search_handle(self):
        """A handle on this index's device data with its own workspace and lock
        (saqgpu_index_search_handle): searches on different handles run concurrently.  Close
        every handle before the index."""
        h = C.c_void_p()
        _check(lib().saqgpu_index_search_handle(self.h, C.byref(h)))
        other = Index.__new__(Index)
        other.__dict__.update({k: v for k, v in self.__dict__.items() if k != "h"})
        other.h = h
        other._parent = self  # the index outlives its handles
        return other

    def close(self):
        if getattr(self, "h", None):
            lib().saqgpu_index_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cluster_sizes_of(self, cid):
        s = ClusterSizes()
        _check(lib().saqgpu_cluster_sizes_of(self.h, cid, C.byref(s)))
        return s

    def set_cluster(self, cid, clu):
        """Uploads a cluster given in the reference SaqCluData layout (dict of arrays)."""
        _check(lib().saqgpu_index_set_cluster(
            self.h, cid, _ptr(clu.get("short_factors")), _ptr(clu["short_code"]), _ptr(clu["long_code"]),
            _ptr(clu["long_factors"]), _ptr(clu["ids"]), _ptr(clu["centroid_rot"])))

    def get_cluster(self, cid):
        s = self.cluster_sizes_of(cid)
        n = int(self.cluster_sizes[cid])
        out = {
            "short_factors": np.zeros(max(1, s.short_factor_floats), np.float32),
            "short_code": np.zeros(max(1, s.short_code_bytes), np.uint8),
            "long_code": np.zeros(max(1, s.long_code_bytes), np.uint8),
            "long_factors": np.zeros(max(1, s.long_factor_floats), np.float32),
            "ids": np.zeros(max(1, n), np.uint32),
            "centroid_rot": np.zeros(self.D_pad, np.float32),
        }
        _check(lib().saqgpu_index_get_cluster(self.h, cid, *[_ptr(out[k]) for k in
                                                              ("short_factors", "short_code", "long_code",
                                                               "long_factors", "ids", "centroid_rot")]))
        out["short_factors"] = out["short_factors"][:s.short_factor_floats]
        out["short_code"] = out["short_code"][:s.short_code_bytes]
        out["long_code"] = out["long_code"][:s.long_code_bytes]
        out["long_factors"] = out["long_factors"][:s.long_factor_floats]
        out["ids"] = out["ids"][:n]
        return out

    def encode(self, data, cluster_ids, rd_lmt=6, eps=1e-8, stream=None, ori_qb=0, centroid=None):
        """IVF::construct's quantisation on the GPU (data: host ndarray or device tensor).
        centroid: one centroid (num_dim floats, host) that every list is quantised against
        (IVF::construct's use_1_centroid), else each list's own."""
        cfg = CaqCfg(rd_lmt, eps, ori_qb)
        nfail = C.c_int64(0)
        dev = not isinstance(data, np.ndarray)
        if centroid is None:
            _check(lib().saqgpu_encode(self.h, _ptr(data), _ptr(cluster_ids), int(data.shape[0]), C.byref(cfg),
                                       DEVICE_PTRS if dev else 0, C.byref(nfail), stream))
        else:
            cen = np.ascontiguousarray(centroid, np.float32).ravel()
            _check(lib().saqgpu_encode_one_centroid(self.h, _ptr(data), _ptr(cluster_ids), int(data.shape[0]),
                                                    _ptr(cen), C.byref(cfg), DEVICE_PTRS if dev else 0,
                                                    C.byref(nfail), stream))
        return int(nfail.value)

    def encode_residuals(self, cid, seg, o, centroid_seg, rd_lmt=6, eps=1e-8, ori_qb=0):
        cfg = CaqCfg(rd_lmt, eps, ori_qb)
        _check(lib().saqgpu_encode_residuals(self.h, cid, seg, _ptr(np.ascontiguousarray(o, np.float32)),
                                             _ptr(np.ascontiguousarray(centroid_seg, np.float32)), C.byref(cfg)))

    def probe(self, queries, nprobe, dist_type=0, out=None, stream=None):
        dev = not isinstance(queries, np.ndarray)
        nq = int(queries.shape[0])
        if out is None:
            if dev:
                import torch
                out = torch.empty((nq, nprobe), dtype=torch.int32, device=queries.device)
            else:
                out = np.zeros((nq, nprobe), np.uint32)
        _check(lib().saqgpu_probe(self.h, _ptr(queries), nq, nprobe, dist_type, _ptr(out),
                                  DEVICE_PTRS if dev else 0, stream))
        return out

    def probe_noalloc(self, queries, nprobe, out, dist_type=0, stream=None):
        """saqgpu_probe_noalloc: device tensors only (queries nq x num_dim, out nq x nprobe int32);
        queued on stream, no device allocation."""
        _check(lib().saqgpu_probe_noalloc(self.h, _ptr(queries), int(queries.shape[0]), nprobe, dist_type,
                                          _ptr(out), stream))
        return out

    def search(self, queries, topk, nprobe, probes=None, m=4.0, dist_type=0, out_ids=None, out_dists=None,
               metrics=None, stream=None, want_dists=True, want_metrics=True, nprobe_per_query=None):
        """IVF::search for a batch (queries: nq x num_dim).  Returns (ids, dists, metrics)."""
        dev = not isinstance(queries, np.ndarray)
        nq = int(queries.shape[0])
        npq = nprobe_per_query
        if npq is not None and not dev:
            npq = np.ascontiguousarray(npq, np.int32)
        p = SearchParams(topk, nprobe, m, dist_type, _ptr(npq))
        if not dev:
            queries = np.ascontiguousarray(queries, np.float32)
            if probes is not None:
                probes = np.ascontiguousarray(probes, np.uint32)
            out_ids = np.zeros((nq, topk), np.uint32) if out_ids is None else out_ids
            if want_dists and out_dists is None:
                out_dists = np.zeros((nq, topk), np.float32)
            if want_metrics and metrics is None:
                metrics = np.zeros((nq, 3), np.uint64)
        _check(lib().saqgpu_search(self.h, _ptr(queries), nq, _ptr(probes), C.byref(p), _ptr(out_ids),
                                   _ptr(out_dists), _ptr(metrics), DEVICE_PTRS if dev else 0, stream))
        return out_ids, out_dists, metrics

    def prepare_batched(self):
        _check(lib().saqgpu_prepare_batched(self.h))

    def _bparams(self, topk, nprobe, nprobe_per_query, precision, query_tile, dev, exact_probe=False):
        npq = None
        if nprobe_per_query is not None and not dev:
            npq = np.ascontiguousarray(nprobe_per_query, np.int32)
        elif nprobe_per_query is not None:
            npq = nprobe_per_query
        p = BatchedParams(topk, nprobe, _ptr(npq), precision, query_tile, int(bool(exact_probe)))
        return p, npq

    def search_batched(self, queries, topk, nprobe, probes=None, nprobe_per_query=None, precision=PREC_INT8,
                       query_tile=0, out_ids=None, out_dists=None, stream=None, want_dists=True, exact_probe=False,
                       graph=False):
        """Batched int8-MFMA search (saqgpu_search_batched).  Returns (ids, dists).  graph: replay a
        captured HIP graph while the shape and device buffers stay the same (device tensors and a
        stream only)."""
        dev = not isinstance(queries, np.ndarray)
        nq = int(queries.shape[0])
        p, keep = self._bparams(topk, nprobe, nprobe_per_query, precision, query_tile, dev, exact_probe)
        if not dev:
            queries = np.ascontiguousarray(queries, np.float32)
            if probes is not None:
                probes = np.ascontiguousarray(probes, np.uint32)
            out_ids = np.zeros((nq, topk), np.uint32) if out_ids is None else out_ids
            if want_dists and out_dists is None:
                out_dists = np.zeros((nq, topk), np.float32)
        _check(lib().saqgpu_search_batched(self.h, _ptr(queries), nq, _ptr(probes), C.byref(p), _ptr(out_ids),
                                           _ptr(out_dists), (DEVICE_PTRS if dev else 0) | (USE_GRAPH if graph else 0),
                                           stream))
        del keep
        return out_ids, out_dists

    def autotune_batched(self, queries, topk, nprobe, probes=None, nprobe_per_query=None, precision=PREC_INT8,
                         stream=None, exact_probe=False):
        """Times every query tile on this batch; the fastest becomes the default.  Returns it."""
        dev = not isinstance(queries, np.ndarray)
        if not dev:
            queries = np.ascontiguousarray(queries, np.float32)
            if probes is not None:
                probes = np.ascontiguousarray(probes, np.uint32)
        p, keep = self._bparams(topk, nprobe, nprobe_per_query, precision, 0, dev, exact_probe)
        best = C.c_int32(0)
        _check(lib().saqgpu_autotune_batched(self.h, _ptr(queries), int(queries.shape[0]), _ptr(probes), C.byref(p),
                                             DEVICE_PTRS if dev else 0, stream, C.byref(best)))
        del keep
        return best.value

    def estimate(self, query, nprobe, probes=None, m=4.0, dist_type=0):
        """IVF::estimate for one query (host array of num_dim floats): every vector of the
        nprobe probed lists in probe order.  Returns (ids, acc, fast, vars)."""
        query = np.ascontiguousarray(query, np.float32).ravel()
        if probes is not None:
            probes = np.ascontiguousarray(probes, np.uint32)[:nprobe]
            total = int(sum(int(self.cluster_sizes[c]) for c in probes))
        else:  # the library probes (FlatInitializer's lists); room for the nprobe largest lists
            if getattr(self, "_sizes_desc", None) is None:
                self._sizes_desc = np.cumsum(np.sort(np.asarray(self.cluster_sizes, np.int64))[::-1])
            total = int(self._sizes_desc[min(nprobe, len(self._sizes_desc)) - 1])
        ids = np.empty(max(1, total), np.uint32)
        acc, fast, var = (np.empty(max(1, total), np.float32) for _ in range(3))
        n = C.c_int64(0)
        _check(lib().saqgpu_estimate(self.h, query.ctypes.data, _ptr(probes), nprobe, m, dist_type,
                                     ids.ctypes.data, acc.ctypes.data, fast.ctypes.data, var.ctypes.data,
                                     C.byref(n)))
        k = n.value
        return ids[:k], acc[:k], fast[:k], var[:k]

    def single_layout(self):
        """SaqSingleDataWrapper record size for this plan and its offsets: (bytes, code section,
        long-code section, ExFactor section, [(short-code, long-code offset) per segment])."""
        nb = C.c_int64(0)
        off = np.zeros(3 + 2 * len(self.plan), np.int64)
        _check(lib().saqgpu_single_layout(self.h, C.byref(nb), off.ctypes.data))
        return (int(nb.value), int(off[0]), int(off[1]), int(off[2]),
                [(int(off[3 + 2 * s]), int(off[4 + 2 * s])) for s in range(len(self.plan))])

    def single_quantize(self, data, rd_lmt=6, eps=1e-8, stride=None, out=None, stream=None, ori_qb=0):
        """SAQuantizerSingle::quantize of every row of data into SaqSingleDataWrapper records
        `stride` bytes apart (default: the record size rounded up to 64).  Host ndarray in, uint8
        ndarray out; or device tensors with `out` given.  Returns (records, stride, n_check_failed)."""
        nbytes = self.single_layout()[0]
        stride = stride or (nbytes + 63) // 64 * 64
        dev = not isinstance(data, np.ndarray)
        n = int(data.shape[0])
        if not dev:
            data = np.ascontiguousarray(data, np.float32)
            out = np.zeros(max(1, n) * stride, np.uint8) if out is None else out
        cfg = CaqCfg(rd_lmt, eps, ori_qb)
        nfail = C.c_int64(0)
        _check(lib().saqgpu_single_quantize(self.h, _ptr(data), n, C.byref(cfg), _ptr(out), stride,
                                            DEVICE_PTRS if dev else 0, C.byref(nfail), stream))
        return out, stride, int(nfail.value)

    def single_estimate(self, queries, records, nrec, stride, m=4.0, dist_type=0, out=None, stream=None,
                        want_all=True):
        """SaqSingleEstimator for every (query, record): returns (vars, fast, acc), each nq x nrec
        (vars/fast None when want_all is False).  Host ndarrays, or device tensors with `out` =
        (vars or None, fast or None, acc) given."""
        dev = not isinstance(queries, np.ndarray)
        nq = int(queries.shape[0])
        if not dev:
            queries = np.ascontiguousarray(queries, np.float32)
            shp = (nq, nrec)
            out = tuple(np.zeros(shp, np.float32) if (want_all or i == 2) else None for i in range(3))
        ov, of, oa = out
        _check(lib().saqgpu_single_estimate(self.h, _ptr(queries), nq, _ptr(records), nrec, stride, m, dist_type,
                                            _ptr(ov), _ptr(of), _ptr(oa), DEVICE_PTRS if dev else 0, stream))
        return ov, of, oa

    def timing(self, enable):
        """Starts (enable=True) or stops event timing of the scan kernel; returns (ms, launches)
        recorded since the last call."""
        ms = C.c_double(0)
        n = C.c_int64(0)
        _check(lib().saqgpu_timing(self.h, int(enable), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def device_bytes(self):
        return int(lib().saqgpu_index_device_bytes(self.h))
