"""Pins of the oracle's search control flow (CPU, no GPU).

The reference's searcher (saqlib/quantization/saq_searcher.hpp) needs glog and gflags, which
this image lacks, so it cannot be compiled here.  Its control flow is instead pinned by
checks that a plausible mistake in the oracle's restatement would fail:

* one segment (scanCluster, saq_searcher.hpp:176-214): an independent Python model of the
  block loop -- mask taken once per block from the fast estimate against the k-th distance,
  refines in lane order, the last block cut to n % 32 lanes (so a list whose size is a
  multiple of 32 never scores its last block) -- fed with the oracle's own per-vector fast
  and accurate estimates (IVF::estimate, index/ivf.hpp:313-362) and the ResultPool model
  (utils/pool.hpp:12-61) must give the oracle search's IDs and distances bit for bit;
* duplicate vectors (equal distances) come out newest first, the pool's tie rule;
* with k at least the number of probed vectors nothing can be pruned, so every probed
  vector comes back with its accurate distance (multi-segment plans included);
* pruning only ever loses candidates: the i-th returned distance is never below the i-th
  smallest accurate estimate.
"""
import ctypes as C

import numpy as np
import pytest

import saq_fixtures as F
from oracle_lib import Metrics, oracle

FLT_MAX = float(np.finfo(np.float32).max)


def estimate(case, probe, q, m=4.0):
    O = oracle()
    n = int(sum(case.clusters[c]["num_vec"] for c in probe))
    ids = np.zeros(max(1, n), np.uint32)
    acc = np.zeros(max(1, n), np.float32)
    fast = np.zeros(max(1, n), np.float32)
    var = np.zeros(max(1, n), np.float32)
    vpad = np.zeros(case.D_pad, np.float32)
    vpad[:case.num_dim] = case.variance
    met = Metrics()
    w = O.orc_estimate(C.byref(case.orc_plan()), vpad, F.orc_rot_ptrs(case), F.orc_clusters(case),
                       np.ascontiguousarray(probe, np.uint32), len(probe), np.ascontiguousarray(q), m, 0, ids,
                       acc, fast.ctypes.data, var.ctypes.data, C.cast(C.pointer(met), C.c_void_p))
    assert w == n
    return ids[:n], acc[:n], fast[:n]


def pool_insert(pool, k, u, d):
    """utils/pool.hpp:36-44: reject when full and d > tail; new entries go before equal ones."""
    if len(pool) == k and d > pool[-1][1]:
        return
    pos = 0
    while pos < len(pool) and pool[pos][1] < d:
        pos += 1
    pool.insert(pos, (u, d))
    del pool[k:]


def scan_model(case, probe, q, k):
    ids, acc, fast = estimate(case, probe, q)
    pool, off = [], 0
    for c in probe:
        n = int(case.clusters[c]["num_vec"])
        nb = (n + 31) // 32
        distk = pool[-1][1] if len(pool) == k else FLT_MAX
        for b in range(nb):
            cnt = n % 32 if b == nb - 1 else 32
            mask = [j for j in range(cnt) if fast[off + 32 * b + j] < distk]
            for j in mask:
                v = off + 32 * b + j
                pool_insert(pool, k, int(ids[v]), float(acc[v]))
                distk = pool[-1][1] if len(pool) == k else FLT_MAX
        off += n
    return pool


def skewed_ids(n, K, seed, sizes):
    """Cluster assignment with chosen list sizes (multiples of 32 included)."""
    assert sum(sizes) == n and len(sizes) == K
    ids = np.concatenate([np.full(s, c, np.uint32) for c, s in enumerate(sizes)])
    return np.random.default_rng(seed).permutation(ids)


@pytest.mark.parametrize("nq", [6, 24])
@pytest.mark.parametrize("topk", [1, 5, 10, 40])
@pytest.mark.parametrize("bits", [1, 4, 7])
def test_single_segment_search_matches_block_model(topk, bits, nq):
    # (bits 4, k 40, 6 queries) holds a block where an insert lowers the k-th distance below a
    # later lane's fast estimate: re-testing lanes against the updated distance fails there
    sizes = [64, 96, 37, 160, 33, 10]  # 64, 96 and 160 are multiples of 32
    case = F.make_case(n=sum(sizes), d=128, K=len(sizes), nq=nq, plan=[(128, bits)], seed=11 + bits,
                       cluster_ids=skewed_ids(sum(sizes), len(sizes), 3, sizes))
    probes = F.orc_probes(case, len(sizes))
    ids, dists, _ = F.orc_search(case, probes, topk)
    for i in range(case.queries.shape[0]):
        pool = scan_model(case, probes[i], case.queries[i], topk)
        want_ids = [u for u, _ in pool] + [0] * (topk - len(pool))
        assert list(ids[i]) == want_ids, f"query {i}"
        got = dists[i][:len(pool)]
        assert np.array_equal(got.view(np.uint32), np.array([d for _, d in pool], np.float32).view(np.uint32))


def test_last_block_of_a_multiple_of_32_list_is_never_scored():
    sizes = [64, 32]
    case = F.make_case(n=96, d=128, K=2, nq=4, plan=[(128, 4)], seed=5, cluster_ids=skewed_ids(96, 2, 1, sizes))
    probes = F.orc_probes(case, 2)
    ids, _, _ = F.orc_search(case, probes, 96)
    for c in range(2):
        members = case.clusters[c]["ids"][:sizes[c]]
        skipped = set(members[-32:].tolist())
        for row in ids:
            assert not skipped & set(row.tolist())
            # the other vectors of the list are all there (k covers everything)
            assert set(members[:-32].tolist()) <= set(row.tolist())


def test_equal_distances_come_out_newest_first():
    n, d = 70, 128
    case = F.make_case(n=n, d=d, K=1, nq=3, plan=[(128, 4)], seed=9, cluster_ids=np.zeros(n, np.uint32),
                       encode=False)
    case.base[40:50] = case.base[5]  # ten copies of vector 5: identical codes and estimates
    F.encode_oracle(case)
    probes = F.orc_probes(case, 1)
    ids, dists, _ = F.orc_search(case, probes, n)
    dup = {5} | set(range(40, 50))
    for row, dr in zip(ids, dists):
        pos = [i for i, u in enumerate(row) if u in dup]
        assert len(pos) == 11 and pos == list(range(pos[0], pos[0] + 11))
        assert len(set(dr[pos].view(np.uint32).tolist())) == 1
        # scan order is 5, 40, ..., 49: each later copy goes before the earlier ones
        assert list(row[pos]) == [49, 48, 47, 46, 45, 44, 43, 42, 41, 40, 5]


@pytest.mark.parametrize("plan", [[(64, 4), (64, 2)], [(128, 7), (64, 3), (64, 0)], [(64, 1), (128, 5)]])
def test_unpruned_search_returns_every_vector_with_its_accurate_distance(plan):
    d = sum(x for x, _ in plan)
    case = F.make_case(n=300, d=d, K=3, nq=4, plan=plan, seed=21, tau=16)
    probes = F.orc_probes(case, 3)
    k = 300
    ids, dists, _ = F.orc_search(case, probes, k)
    for i in range(case.queries.shape[0]):
        eid, eacc, _ = estimate(case, probes[i], case.queries[i])
        assert sorted(ids[i].tolist()) == sorted(eid.tolist())
        ref = dict(zip(eid.tolist(), eacc.tolist()))
        got = dists[i]
        assert np.all(np.diff(got) >= 0)
        # stage 3 accumulates fast + sum_s (acc_s - fast_s): equal to sum_s acc_s up to rounding
        for u, dv in zip(ids[i].tolist(), got.tolist()):
            assert abs(dv - ref[u]) <= 1e-5 * max(1.0, abs(ref[u]))


@pytest.mark.parametrize("plan", [[(128, 4)], [(64, 6), (64, 3)]])
def test_pruning_only_loses_candidates(plan):
    case = F.make_case(n=1500, d=128, K=8, nq=10, plan=plan, seed=23)
    probes = F.orc_probes(case, 4)
    k = 20
    ids, dists, _ = F.orc_search(case, probes, k)
    for i in range(case.queries.shape[0]):
        eid, eacc, _ = estimate(case, probes[i], case.queries[i])
        best = np.sort(eacc)[:k]
        filled = np.count_nonzero(dists[i] != 0) or k
        assert np.all(dists[i][:filled] >= best[:filled] - 1e-5 * np.abs(best[:filled]))
        assert set(ids[i].tolist()) <= set(eid.tolist())
