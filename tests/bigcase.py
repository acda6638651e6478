"""Full-size workloads for the GPU tests (test infrastructure).

The datasets of the benchmark configurations are generated, PCA-rotated and clustered on the
GPU with the benchmark's own generators (bench.py: make_data, pca, rotate_inplace, kmeans,
exact_topk), encoded into a device index by the library, and read back list by list
(saqgpu_index_get_cluster) for the oracle, which then searches, encodes or decodes exactly
the lists a check needs.
"""
import ctypes as C
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import bench as B
import saq_fixtures as F
from oracle_lib import CaqCfg, Cluster, Metrics, Sizes, make_plan, oracle
from saq_amd import Index
from saq_amd import data as sdata
from saq_amd.plan import make_plan as gpu_make_plan

THREADS = min(16, os.cpu_count() or 1)


def log(msg):
    print(f"[bigcase] {msg}", flush=True)


class Dataset:
    """Base (device), queries (device and host), PCA variance, k-means lists."""

    def __init__(self, torch, n, dim, comps, tau, zipf, nlist, nq, kmeans_iters=6, gt_k=0, seed=1):
        dev = torch.device("cuda:0")
        self.torch, self.dev = torch, dev
        log(f"dataset N={n} D={dim} comps={comps} tau={tau} zipf={zipf} nlist={nlist} nq={nq}")
        base, queries = B.make_data(torch, n, dim, comps, tau, zipf, nq, seed, dev)
        mean, pcs, var = B.pca(torch, base)
        self.base = B.rotate_inplace(torch, base, mean, pcs)
        self.queries = ((queries - mean) @ pcs).contiguous()
        self.queries_np = self.queries.cpu().numpy()
        cent, cid = B.kmeans(torch, self.base, nlist, kmeans_iters, 0)
        self.centroids = cent.cpu().numpy()
        self.sizes = torch.bincount(cid, minlength=nlist).cpu().numpy().astype(np.int64)
        self.cid32 = cid.to(torch.int32).contiguous()
        self.cid_np = cid.cpu().numpy().astype(np.uint32)
        del cid
        self.var = var
        self.n, self.dim, self.nlist, self.nq = n, dim, nlist, nq
        self.gt = B.exact_topk(torch, self.base, self.queries, gt_k) if gt_k else None
        self._members = None
        log(f"lists: mean {self.sizes.mean():.0f} max {self.sizes.max()} empty {(self.sizes == 0).sum()}")

    def members(self, c):
        """Ids of list c in increasing order (IVF::construct's bucketing, ivf.hpp:149-159)."""
        if self._members is None:
            order = np.argsort(self.cid_np, kind="stable")
            starts = np.concatenate([[0], np.cumsum(self.sizes)])
            self._members = (order.astype(np.uint32), starts)
        order, starts = self._members
        return order[starts[c]:starts[c + 1]]


class Built:
    """A device index of a dataset at bits per dimension, with the oracle's view of it."""

    def __init__(self, ds, bits, rd_lmt=6, eps=1e-8):
        self.ds = ds
        self.plan = gpu_make_plan(ds.var, ds.dim, bits)
        self.rots = [sdata.random_rotation(d, 1000 + i) for i, (d, _) in enumerate(self.plan)]
        self.ix = Index(ds.dim, self.plan, ds.centroids, ds.sizes, ds.var, rotations=self.rots)
        log(f"encode B={bits} plan {self.plan}")
        self.nfail = self.ix.encode(ds.base, ds.cid32, rd_lmt=rd_lmt, eps=eps)
        self.rd_lmt, self.eps = rd_lmt, eps
        self.D_pad = sum(d for d, _ in self.plan)
        self.vpad = np.zeros(self.D_pad, np.float32)
        self.vpad[:ds.dim] = ds.var
        self.oplan = make_plan([d for d, _ in self.plan], [b for _, b in self.plan], ds.dim)
        self.rp = (C.c_void_p * len(self.plan))(*[r.ctypes.data for r in self.rots])
        self.arrays = {}

    def close(self):
        self.ix.close()

    def lists(self, cids):
        """The reference-layout arrays of lists cids, read back from the device (cached)."""
        for c in sorted(set(int(x) for x in cids)):
            if c not in self.arrays:
                a = self.ix.get_cluster(c)
                a["num_vec"] = int(self.ds.sizes[c])
                self.arrays[c] = a
        return self.arrays

    def orc_clusters(self, cids):
        arrs = self.lists(cids)
        out = (Cluster * self.ds.nlist)()
        for c, a in arrs.items():
            out[c] = Cluster(a["num_vec"], a["short_factors"].ctypes.data if a["short_factors"].size else None,
                             a["short_code"].ctypes.data if a["short_code"].size else None,
                             a["long_code"].ctypes.data if a["long_code"].size else None,
                             a["long_factors"].ctypes.data if a["long_factors"].size else None,
                             a["ids"].ctypes.data if a["ids"].size else None, a["centroid_rot"].ctypes.data)
        return out

    def orc_probes(self, qidx, nprobe):
        O = oracle()
        pr = np.zeros((len(qidx), nprobe), np.uint32)
        for i, q in enumerate(qidx):
            O.orc_probe(self.ds.centroids, self.ds.nlist, self.ds.dim, np.ascontiguousarray(self.ds.queries_np[q]),
                        nprobe, 0, pr[i], None)
        return pr

    def orc_search(self, qidx, probes, topk, npq=None, m=4.0):
        """IVF::search of the queries qidx with probe lists probes (rows), by the oracle on
        the device index's lists; per-query nprobe npq (optional)."""
        O = oracle()
        nq = len(qidx)
        ids = np.zeros((nq, topk), np.uint32)
        dists = np.zeros((nq, topk), np.float32)
        mets = (Metrics * nq)()
        need = set()
        for i in range(nq):
            need.update(probes[i, :(npq[i] if npq is not None else probes.shape[1])].tolist())
        clus = self.orc_clusters(need)
        qs = np.ascontiguousarray(self.ds.queries_np[np.asarray(qidx)])
        groups = [(np.arange(nq), probes.shape[1])] if npq is None else \
            [(np.nonzero(npq == v)[0], int(v)) for v in np.unique(npq)]
        for rows, npb in groups:
            sub_ids = np.zeros((len(rows), topk), np.uint32)
            sub_d = np.zeros((len(rows), topk), np.float32)
            sub_m = (Metrics * len(rows))()
            O.orc_search_batch(C.byref(self.oplan), self.vpad, self.rp, clus,
                               np.ascontiguousarray(probes[rows, :npb], np.uint32), npb, np.ascontiguousarray(qs[rows]),
                               len(rows), topk, m, 0, THREADS, sub_ids, sub_d.ctypes.data, C.cast(sub_m, C.c_void_p))
            ids[rows] = sub_ids
            dists[rows] = sub_d
            for j, r in enumerate(rows):
                mets[r] = sub_m[j]
        met = np.array([[x.fast_bitsum, x.acc_bitsum, x.total_comp_cnt] for x in mets], np.uint64)
        return ids, dists, met

    def orc_encode_list(self, c):
        """The oracle encoder's reference-layout arrays of list c (IVF::construct, one list)."""
        O = oracle()
        members = self.ds.members(c)
        n = len(members)
        sz = Sizes()
        O.orc_cluster_sizes(C.byref(self.oplan), n, C.byref(sz))
        clu = {"short_factors": np.zeros(max(1, sz.short_factor_floats), np.float32),
               "short_code": np.zeros(max(1, sz.short_code_bytes), np.uint8),
               "long_code": np.zeros(max(1, sz.long_code_bytes), np.uint8),
               "long_factors": np.zeros(max(1, sz.long_factor_floats), np.float32),
               "ids": members.astype(np.uint32) if n else np.zeros(1, np.uint32),
               "centroid_rot": np.zeros(self.D_pad, np.float32), "num_vec": n}
        torch = self.ds.torch
        X_all = self.ds.base[torch.from_numpy(members.astype(np.int64)).to(self.ds.dev)].cpu().numpy() if n \
            else np.zeros((0, self.ds.dim), np.float32)
        cfg = CaqCfg(self.rd_lmt, self.eps, 1)
        off = 0
        for s, (d, b) in enumerate(self.plan):
            X = np.zeros((max(1, n), d), np.float32)
            cp = max(0, min(d, self.ds.dim - off))
            X[:n, :cp] = X_all[:, off:off + cp]
            cen = np.zeros(d, np.float32)
            cen[:cp] = self.ds.centroids[c, off:off + cp]
            co = F._cluster_out(clu)
            O.orc_quantize_segment(C.byref(self.oplan), s, X, n, cen, self.rots[s].ctypes.data, C.byref(cfg),
                                   C.cast(C.pointer(co), C.c_void_p), None)
            off += d
        for k, f in (("short_factors", "short_factor_floats"), ("short_code", "short_code_bytes"),
                     ("long_code", "long_code_bytes"), ("long_factors", "long_factor_floats")):
            clu[k] = clu[k][:getattr(sz, f)]
        return clu

    def tables(self, cids):
        """Per list: ids and per segment (full codes, |o|, rescale) decoded from the device
        index's lists by the oracle (the batched model's inputs)."""
        O = oracle()
        arrs = self.lists(cids)
        clus = self.orc_clusters(cids)
        out = {}

        def one(c):
            n = arrs[c]["num_vec"]
            segs = []
            for s, (d, _) in enumerate(self.plan):
                codes = np.zeros(max(1, n) * d, np.int32)
                on = np.zeros(max(1, n), np.float32)
                rs = np.zeros(max(1, n), np.float32)
                if n:
                    O.orc_decode_cluster(C.byref(self.oplan), C.cast(C.pointer(clus[c]), C.c_void_p), s, codes, on, rs)
                segs.append((codes[:n * d].reshape(n, d).astype(np.int64), on[:n], rs[:n]))
            return c, (arrs[c]["ids"][:n], segs)

        with ThreadPoolExecutor(THREADS) as ex:
            for c, t in ex.map(one, sorted(set(int(x) for x in cids))):
                out[c] = t
        return out

    def rotated_queries(self, qidx):
        O = oracle()
        qr = np.zeros((len(qidx), self.D_pad), np.float32)
        for i, q in enumerate(qidx):
            off = 0
            for s, (d, _) in enumerate(self.plan):
                x = np.zeros(d, np.float32)
                cp = max(0, min(d, self.ds.dim - off))
                x[:cp] = self.ds.queries_np[q, off:off + cp]
                O.orc_gemv_rowvec(x, self.rots[s], d, qr[i, off:off + d])
                off += d
        return qr

    def model_view(self):
        """The attributes batched_model.model_search reads from a case."""
        class V:
            pass
        v = V()
        v.plan = self.plan
        v.clusters = {c: {"centroid_rot": a["centroid_rot"]} for c, a in self.arrays.items()}
        return v


def check_batched(res, alld, ids, dists, tol=1e-5):
    """Batched results against the model (tests/batched_model.py): every returned distance equals
    the model's distance of that vector to tol relative, rows are sorted by (distance, id), and
    the IDs are the model's top-k except at model ties within tolerance of the k-th distance."""
    for i in range(ids.shape[0]):
        want = res[i]
        n = len(want)
        got_ids = ids[i, :n].tolist()
        got_d = dists[i, :n]
        for u, dv in zip(got_ids, got_d.tolist()):
            assert u in alld[i], f"query {i}: id {u} not in a probed list"
            ref = alld[i][u]
            assert abs(dv - ref) <= tol * max(1.0, abs(ref)), f"query {i} id {u}: {dv} vs model {ref}"
        pairs = list(zip(got_d.tolist(), got_ids))
        assert pairs == sorted(pairs), f"query {i}: not sorted by (distance, id)"
        if n:
            kth = want[-1][1]
            for u in set(got_ids) ^ {u for u, _ in want}:
                assert abs(alld[i][u] - kth) <= 2 * tol * max(1.0, abs(kth)), f"query {i}: id {u} differs off a tie"


def rows_well_formed(ids, dists):
    """Every row sorted by (distance, id) with distinct ids (the properties checked on every query
    of a full-size batch)."""
    d = dists.astype(np.float64)
    assert np.all(np.isfinite(d))
    asc = (np.diff(d, axis=1) > 0) | ((np.diff(d, axis=1) == 0) & (np.diff(ids.astype(np.int64), axis=1) > 0))
    assert asc.all(), f"{int((~asc).any(1).sum())} rows not sorted by (distance, id)"
    srt = np.sort(ids, axis=1)
    assert (np.diff(srt.astype(np.int64), axis=1) != 0).all(), "an id repeats within a row"


def recall(found, gt, k):
    return float(np.mean([len(set(f[:k].tolist()) & set(g[:k].tolist())) / k for f, g in zip(found, gt)]))
