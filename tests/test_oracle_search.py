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
  smallest accurate estimate;
* several segments (searchCluster, saq_searcher.hpp:68-173): an independent float32 model of the
  three stages, fed the oracle's per-segment variance, fast and accurate estimates of every lane
  (padded lanes of a last block included), must give the oracle search's IDs, distances and
  QueryRuntimeMetrics bit for bit, while its cases skip blocks at stage 1, stop stage 2 early and
  insert partial stage-3 sums; the model with a planted mistake (distk fixed per block, or the
  complete instead of the partial sum inserted) no longer matches.
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
    """utils/pool.hpp:17-30: reject when full and d > tail; new entries go before equal ones."""
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


# ---- searchCluster (saq_searcher.hpp:68-173), several segments
def lanes(case, probe, q, m=4.0):
    """The oracle's per-segment estimates of every lane of every probed block:
    (vars, fast, acc), each blocks x S x 32 float32."""
    O = oracle()
    S = len(case.plan)
    nb = int(sum((case.clusters[c]["num_vec"] + 31) // 32 for c in probe))
    out = [np.zeros(max(1, nb) * S * 32, np.float32) for _ in range(3)]
    vpad = np.zeros(case.D_pad, np.float32)
    vpad[:case.num_dim] = case.variance
    w = O.orc_estimate_lanes(C.byref(case.orc_plan()), vpad, F.orc_rot_ptrs(case), F.orc_clusters(case),
                             np.ascontiguousarray(probe, np.uint32), len(probe), np.ascontiguousarray(q), m, *out)
    assert w == nb
    return [o[:nb * S * 32].reshape(nb, S, 32) for o in out]


def search_cluster_model(case, probe, q, k, stats, mistake=None):
    """saq_searcher.hpp:87-161 in float32, lane by lane, from the per-segment estimates alone:
    stage 1 sums the variance estimates in segment order and skips the block when the minimum
    over all 32 lanes (padded ones included) exceeds distk; stage 2 swaps, segment by segment
    (bits > 0 only), the variance estimate for the fast one and stops when the minimum exceeds
    distk; stage 3 (only when the last minimum is <= distk) takes lanes whose running sum is below
    distk (re-read after every insert) up to the list's end, adds acc_s - cd_s per segment, stops
    adding once the sum reaches distk, and inserts that (possibly partial) sum.
    mistake (to show the comparison has power): "block_distk" keeps the block's starting distk
    through stage 3, "full_sum" inserts the complete sum instead of the partial one."""
    vars_, fast, acc = lanes(case, probe, q)
    S = len(case.plan)
    bits = [b for _, b in case.plan]
    f32 = np.float32
    pool = []
    fast_bits = acc_bits = comp = 0
    blk_g = 0
    for c in probe:
        n = int(case.clusters[c]["num_vec"])
        ids = case.clusters[c]["ids"]
        nb = (n + 31) // 32
        distk = f32(pool[-1][1]) if len(pool) == k else f32(FLT_MAX)
        for b in range(nb):
            V, Fa, A = vars_[blk_g], fast[blk_g], acc[blk_g]
            blk_g += 1
            curr = np.zeros(32, np.float32)
            for s in range(S):
                curr = curr + V[s]
            mi = curr.min()
            if mi > distk:
                stats["stage1_skip"] += 1
                continue
            cd = V.copy()
            for s in range(S):
                if bits[s] == 0:
                    continue
                curr = curr - cd[s]
                cd[s] = Fa[s]
                fast_bits += 32 * case.plan[s][0]
                curr = curr + cd[s]
                mi = curr.min()
                if mi > distk:
                    stats["stage2_break"] += 1
                    break
            if mi > distk:
                continue
            for j in range(32):
                if not curr[j] < distk:
                    continue
                idx = 32 * b + j
                if idx >= n:
                    break
                a = f32(curr[j])
                for s in range(S):
                    a = f32(a + f32(A[s, j] - cd[s, j]))
                    if bits[s] > 0:
                        acc_bits += case.plan[s][0] * (bits[s] - 1)
                    if a >= distk and mistake != "full_sum":
                        if s < S - 1:
                            stats["partial"] += 1
                        break
                pool_insert(pool, k, int(ids[idx]), float(a))
                if mistake != "block_distk":
                    distk = f32(pool[-1][1]) if len(pool) == k else f32(FLT_MAX)
        comp += nb * 32
    return pool, (fast_bits, acc_bits, comp)


@pytest.mark.parametrize("plan", [[(64, 4), (64, 2)], [(128, 7), (64, 3), (64, 0)], [(64, 1), (64, 5), (128, 2)],
                                  [(64, 0), (64, 6), (64, 0), (64, 3)]])
@pytest.mark.parametrize("topk", [1, 5, 20])
def test_multi_segment_search_matches_block_model(plan, topk):
    d = sum(x for x, _ in plan)
    sizes = [96, 37, 160, 64, 33, 10, 250]
    case = F.make_case(n=sum(sizes), d=d, K=len(sizes), nq=12, plan=plan, seed=31 + topk, tau=16,
                       cluster_ids=skewed_ids(sum(sizes), len(sizes), 7, sizes))
    probes = F.orc_probes(case, len(sizes))
    ids, dists, met = F.orc_search(case, probes, topk)
    stats = {"stage1_skip": 0, "stage2_break": 0, "partial": 0}
    for i in range(case.queries.shape[0]):
        pool, mets = search_cluster_model(case, probes[i], case.queries[i], topk, stats)
        want_ids = [u for u, _ in pool] + [0] * (topk - len(pool))
        assert list(ids[i]) == want_ids, f"query {i}"
        got = dists[i][:len(pool)]
        assert np.array_equal(got.view(np.uint32), np.array([d for _, d in pool], np.float32).view(np.uint32))
        assert tuple(int(x) for x in met[i]) == mets, f"query {i} metrics"
    assert stats["partial"] > 0, "no partial stage-3 sum was inserted"


@pytest.mark.parametrize("mistake", [None, "block_distk", "full_sum"])
def test_multi_segment_model_branches_and_power(mistake):
    """One configuration that skips blocks at stage 1, stops stage 2 early and inserts partial
    stage-3 sums; the model with a planted mistake no longer reproduces the oracle."""
    plan = [(128, 7), (64, 3), (64, 0)]
    sizes = [96, 37, 160, 64, 33, 10, 250]
    case = F.make_case(n=sum(sizes), d=256, K=len(sizes), nq=12, plan=plan, seed=32, tau=16,
                       cluster_ids=skewed_ids(sum(sizes), len(sizes), 7, sizes))
    probes = F.orc_probes(case, len(sizes))
    ids, dists, met = F.orc_search(case, probes, 5)
    stats = {"stage1_skip": 0, "stage2_break": 0, "partial": 0}
    same = True
    for i in range(case.queries.shape[0]):
        pool, mets = search_cluster_model(case, probes[i], case.queries[i], 5, stats, mistake)
        same &= [u for u, _ in pool] == list(ids[i]) and \
            np.array_equal(np.array([d for _, d in pool], np.float32), dists[i]) and \
            tuple(int(x) for x in met[i]) == mets
    if mistake is None:
        assert same and all(v > 0 for v in stats.values()), stats
    else:
        assert not same
