"""Small seeded SAQ indexes in the reference layout, built by the oracle encoder.

Test infrastructure: the GPU path is checked against the oracle on exactly these
inputs.  Cluster membership lists follow IVF::construct (ids pushed in increasing
order per cluster, saqlib/index/ivf.hpp:149-159).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from oracle_lib import CaqCfg, Cluster, Metrics, SingleLayout, Sizes, make_plan, oracle
from saq_amd import data as sd


@dataclass
class Case:
    num_dim: int
    plan: list
    base: np.ndarray
    queries: np.ndarray
    centroids: np.ndarray
    cluster_ids: np.ndarray
    variance: np.ndarray
    rotations: list
    clusters: list = field(default_factory=list)  # dicts of reference arrays
    compact: bool = False
    rd_lmt: int = 6
    eps: float = 1e-8

    @property
    def D_pad(self):
        return sum(d for d, _ in self.plan)

    @property
    def K(self):
        return self.centroids.shape[0]

    def sizes(self):
        return np.bincount(self.cluster_ids.astype(np.int64), minlength=self.K).astype(np.int64)

    def orc_plan(self):
        return make_plan([d for d, _ in self.plan], [b for _, b in self.plan], self.num_dim, self.compact)


def make_case(n=3000, d=128, K=12, nq=24, plan=None, avg_bits=4.0, seed=1, rotate=True, compact=False,
              comps=16, tau=8, rd_lmt=6, eps=1e-8, cluster_ids=None, encode=True):
    base, queries = sd.mixture(n, d, comps, tau, seed, queries=nq)
    mean, pcs, var = sd.pca_fit(base)
    base = sd.pca_apply(base, mean, pcs)
    queries = sd.pca_apply(queries, mean, pcs)
    if cluster_ids is None:
        cent, ids = sd.kmeans(base, K, iters=4, seed=0)
    else:
        ids = np.asarray(cluster_ids, np.uint32)
        cent = np.stack([base[ids == c].mean(0) if np.any(ids == c) else np.zeros(d, np.float32)
                         for c in range(K)]).astype(np.float32)
    D_pad = (d + 63) // 64 * 64
    if plan is None:
        vpad = np.zeros(D_pad, np.float32)
        vpad[:d] = var
        dims = np.zeros(64, np.int32)
        bits = np.zeros(64, np.int32)
        ns = oracle().orc_plan_dp(vpad, D_pad, avg_bits, dims, bits)
        plan = [(int(dims[i]), int(bits[i])) for i in range(ns)]
    rots = [sd.random_rotation(ds, seed * 100 + i) if rotate else None for i, (ds, _) in enumerate(plan)]
    case = Case(d, plan, base, queries, cent, ids, var, rots, compact=compact, rd_lmt=rd_lmt, eps=eps)
    if encode:
        encode_oracle(case)
    return case


def encode_oracle(case, ori_qb=0, one_centroid=None):
    """IVF::construct with the oracle encoder: per cluster, per segment (saq_quantizer.hpp:34-60).
    ori_qb: caq_ori_qB; one_centroid: the centroid every list is quantised against
    (use_1_centroid), else each list's own."""
    O = oracle()
    plan = case.orc_plan()
    cfg = CaqCfg(case.rd_lmt, case.eps, 1, ori_qb)
    case.clusters = []
    for c in range(case.K):
        members = np.nonzero(case.cluster_ids == c)[0]
        n = len(members)
        sz = Sizes()
        O.orc_cluster_sizes(C.byref(plan), n, C.byref(sz))
        clu = {
            "short_factors": np.zeros(max(1, sz.short_factor_floats), np.float32),
            "short_code": np.zeros(max(1, sz.short_code_bytes), np.uint8),
            "long_code": np.zeros(max(1, sz.long_code_bytes), np.uint8),
            "long_factors": np.zeros(max(1, sz.long_factor_floats), np.float32),
            "ids": members.astype(np.uint32) if n else np.zeros(1, np.uint32),
            "centroid_rot": np.zeros(case.D_pad, np.float32),
            "num_vec": n,
        }
        out = C.c_void_p * 7
        off = 0
        for s, (ds, b) in enumerate(case.plan):
            X = np.zeros((max(1, n), ds), np.float32)
            cen = np.zeros(ds, np.float32)
            cp = max(0, min(ds, case.num_dim - off))
            if n:
                X[:n, :cp] = case.base[members, off:off + cp]
            cen[:cp] = (case.centroids[c] if one_centroid is None else one_centroid)[off:off + cp]
            co = _cluster_out(clu)
            P = case.rotations[s]
            O.orc_quantize_segment(C.byref(plan), s, X, n, cen, P.ctypes.data if P is not None else None,
                                   C.byref(cfg), C.cast(C.pointer(co), C.c_void_p), None)
            off += ds
        clu["short_factors"] = clu["short_factors"][:sz.short_factor_floats]
        clu["short_code"] = clu["short_code"][:sz.short_code_bytes]
        clu["long_code"] = clu["long_code"][:sz.long_code_bytes]
        clu["long_factors"] = clu["long_factors"][:sz.long_factor_floats]
        case.clusters.append(clu)
    return case


class ClusterOut(C.Structure):
    _fields_ = [("num_vec", C.c_int64), ("short_factors", C.c_void_p), ("short_code", C.c_void_p),
                ("long_code", C.c_void_p), ("long_factors", C.c_void_p), ("ids", C.c_void_p),
                ("centroid_rot", C.c_void_p)]


def _cluster_out(clu):
    return ClusterOut(clu["num_vec"], clu["short_factors"].ctypes.data, clu["short_code"].ctypes.data,
                      clu["long_code"].ctypes.data, clu["long_factors"].ctypes.data, clu["ids"].ctypes.data,
                      clu["centroid_rot"].ctypes.data)


def orc_clusters(case):
    arr = (Cluster * case.K)()
    for c, clu in enumerate(case.clusters):
        arr[c] = Cluster(clu["num_vec"], clu["short_factors"].ctypes.data if clu["short_factors"].size else None,
                         clu["short_code"].ctypes.data if clu["short_code"].size else None,
                         clu["long_code"].ctypes.data if clu["long_code"].size else None,
                         clu["long_factors"].ctypes.data if clu["long_factors"].size else None,
                         clu["ids"].ctypes.data, clu["centroid_rot"].ctypes.data)
    return arr


def orc_rot_ptrs(case):
    arr = (C.c_void_p * len(case.plan))()
    for i, P in enumerate(case.rotations):
        arr[i] = P.ctypes.data if P is not None else None
    return arr


def orc_probes(case, nprobe, dist_type=0):
    O = oracle()
    pr = np.zeros((case.queries.shape[0], nprobe), np.uint32)
    for i, q in enumerate(case.queries):
        O.orc_probe(case.centroids, case.K, case.num_dim, np.ascontiguousarray(q), nprobe, dist_type, pr[i], None)
    return pr


def orc_search(case, probes, topk, m=4.0, dist_type=0, nthreads=8):
    O = oracle()
    plan = case.orc_plan()
    nq, nprobe = probes.shape
    ids = np.zeros((nq, topk), np.uint32)
    dists = np.zeros((nq, topk), np.float32)
    mets = (Metrics * nq)()
    vpad = np.zeros(case.D_pad, np.float32)
    vpad[:case.num_dim] = case.variance
    clus = orc_clusters(case)
    rots = orc_rot_ptrs(case)
    O.orc_search_batch(C.byref(plan), vpad, rots, clus, np.ascontiguousarray(probes, np.uint32), nprobe,
                       np.ascontiguousarray(case.queries, np.float32), nq, topk, m, dist_type, nthreads, ids,
                       dists.ctypes.data, C.cast(mets, C.c_void_p))
    met = np.array([[x.fast_bitsum, x.acc_bitsum, x.total_comp_cnt] for x in mets], np.uint64)
    return ids, dists, met


def single_layout(case):
    lay = SingleLayout()
    oracle().orc_single_layout(C.byref(case.orc_plan()), C.byref(lay))
    return lay


def orc_single_quantize(case, data, stride=None):
    """SAQuantizerSingle::quantize of every row of data into records of `stride` bytes."""
    lay = single_layout(case)
    stride = stride or (lay.bytes + 63) // 64 * 64
    data = np.ascontiguousarray(data, np.float32)
    n = data.shape[0]
    recs = np.zeros(max(1, n) * stride, np.uint8)
    cfg = CaqCfg(case.rd_lmt, case.eps, 1 if case.rotations[0] is not None else 0)
    nfail = oracle().orc_single_quantize(C.byref(case.orc_plan()), C.cast(orc_rot_ptrs(case), C.c_void_p),
                                         data if n else np.zeros(1, np.float32), n, C.byref(cfg), recs, stride)
    return recs[:n * stride], stride, nfail


def orc_single_estimate(case, recs, stride, query, m=4.0, dist_type=0):
    nrec = len(recs) // stride
    out = np.zeros((3, max(1, nrec)), np.float32)
    vpad = np.zeros(case.D_pad, np.float32)
    vpad[:case.num_dim] = case.variance
    oracle().orc_single_estimate(C.byref(case.orc_plan()), vpad, C.cast(orc_rot_ptrs(case), C.c_void_p),
                                 recs if nrec else np.zeros(1, np.uint8), nrec, stride,
                                 np.ascontiguousarray(query, np.float32), m, dist_type, out[0], out[1], out[2])
    return out[0, :nrec], out[1, :nrec], out[2, :nrec]


def gpu_index(case, device=0):
    from saq_amd import Index
    rots = case.rotations if case.rotations[0] is not None else None
    ix = Index(case.num_dim, case.plan, case.centroids, case.sizes(), case.variance, rotations=rots,
               use_compact_layout=case.compact, device=device)
    for c, clu in enumerate(case.clusters):
        ix.set_cluster(c, clu)
    return ix
