"""Parity of the gfx950 scan (saqgpu_search through the C ABI) against the oracle.

Bar: candidate IDs bit-exact, distances bit-exact (they are the same fp32
operations in the same order), QueryRuntimeMetrics counters identical.
Cases cover the reference's branches: one segment (scanCluster) and several
(searchCluster), 0-bit and 1-bit segments, 9+-bit extra codes, no rotation,
the compact long-code layout, topk larger than the probed set, lists whose
size is a multiple of 32 (the last block scanCluster skips) and empty lists.
"""
import numpy as np
import pytest

import saq_fixtures as F

pytestmark = pytest.mark.gpu


def run_case(case, topk, nprobe, m=4.0, dist_type=0):
    probes = F.orc_probes(case, nprobe, dist_type=dist_type)
    ids_o, d_o, met_o = F.orc_search(case, probes, topk, m=m, dist_type=dist_type)
    ix = F.gpu_index(case)
    ids_g, d_g, met_g = ix.search(case.queries, topk, nprobe, probes=probes, m=m, dist_type=dist_type)
    bad = np.nonzero((ids_o != ids_g).any(1))[0]
    assert len(bad) == 0, f"id mismatch in queries {bad[:5]}: oracle {ids_o[bad[0]][:10]} gpu {ids_g[bad[0]][:10]}"
    assert np.array_equal(d_o.view(np.uint32), d_g.view(np.uint32)), "distance bits differ"
    assert np.array_equal(met_o, met_g), f"metrics differ: {met_o[:3]} vs {met_g[:3]}"
    ix.close()
    return ids_g


@pytest.mark.parametrize("d,bits", [(128, 4.0), (128, 1.0), (128, 8.0)])
@pytest.mark.parametrize("topk", [10, 100])
def test_single_segment(d, bits, topk):
    case = F.make_case(n=3000, d=d, K=12, nq=24, plan=[(128, int(bits))], seed=3)
    run_case(case, topk, 5)


@pytest.mark.parametrize("avg_bits", [4.0, 1.5, 2.0])
def test_dp_plan_multi_segment(avg_bits):
    case = F.make_case(n=4000, d=256, K=16, nq=32, avg_bits=avg_bits, seed=5, tau=32)
    assert len(case.plan) >= 1
    run_case(case, 10, 6)
    run_case(case, 100, 6)


@pytest.mark.parametrize("plan", [[(64, 1), (64, 5), (128, 3)], [(128, 10), (128, 0)], [(64, 12), (64, 7), (64, 6),
                                                                                       (64, 2)]])
def test_mixed_bit_plans(plan):
    case = F.make_case(n=2500, d=256, K=10, nq=20, plan=plan, seed=7, tau=32)
    run_case(case, 20, 4)


def test_no_rotation_and_padded_dim():
    case = F.make_case(n=2000, d=200, K=8, nq=16, plan=[(128, 4), (128, 2)], seed=9, rotate=False)
    run_case(case, 10, 3)


def test_compact_layout():
    case = F.make_case(n=2000, d=256, K=8, nq=16, plan=[(128, 5), (128, 3)], seed=11, compact=True)
    run_case(case, 10, 3)


def test_multiple_of_32_lists_and_empty_lists():
    # sizes 64, 96, 0, 32, 45, ... : scanCluster never scores the last block of a list whose
    # size is a multiple of 32 (saq_searcher.hpp:186); the replay must do the same
    K = 6
    sizes = [64, 96, 0, 32, 45, 263]
    ids = np.concatenate([np.full(s, c) for c, s in enumerate(sizes)]).astype(np.uint32)
    n = int(sum(sizes))
    for plan in ([(128, 4)], [(64, 3), (64, 2)]):
        case = F.make_case(n=n, d=128, K=K, nq=12, plan=plan, seed=13, cluster_ids=ids)
        run_case(case, 10, 6)


def test_topk_exceeds_probed_vectors():
    case = F.make_case(n=600, d=128, K=8, nq=8, plan=[(128, 4)], seed=15)
    run_case(case, 1024, 2)


@pytest.mark.parametrize("plan", [[(128, 4)], [(64, 5), (64, 3)], [(64, 9), (64, 0)]])
@pytest.mark.parametrize("topk", [10, 100])
def test_many_slices(plan, topk):
    """nprobe 40 of 48 lists: the scan runs the slices [0,1) [1,3) [3,12) [12,41) of its 41 work
    items, each refining the vectors whose fast estimate beats the k-th distance at its start."""
    case = F.make_case(n=6000, d=128, K=48, nq=24, plan=plan, seed=23)
    run_case(case, topk, 40)


@pytest.mark.parametrize("topk", [1, 10])
def test_refine_early_stop(topk):
    """A plan whose first segment carries most of the distance (64 dims at 9 bits, then 64 x 1
    and 64 x 2) over 24 of 32 lists: a candidate's running sum can reach the slice's starting
    k-th distance after segment 0, and the refine then skips its later segments; IDs, distances
    and metrics must still equal the oracle's."""
    case = F.make_case(n=6000, d=192, K=32, nq=24, plan=[(64, 9), (64, 1), (64, 2)], seed=31, tau=8)
    run_case(case, topk, 24)


@pytest.mark.parametrize("plan", [[(128, 4)], [(64, 5), (64, 3)]])
def test_query_groups(plan):
    """1100 queries: the scan splits them into 4 groups, each on its own stream with its own
    share of the workspace; every group equals the oracle."""
    case = F.make_case(n=12000, d=128, K=40, nq=1100, plan=plan, seed=29)
    run_case(case, 10, 20)


def test_scan_workspace_growth(monkeypatch):
    """Block-info and candidate space far below what the slices need (16 records): the
    search is repeated with the space the device asked for, and still equals the oracle."""
    monkeypatch.setenv("SAQGPU_SLICE_CAP", "16")
    for plan in ([(64, 4), (64, 2)], [(128, 4)]):
        case = F.make_case(n=4000, d=128, K=24, nq=16, plan=plan, seed=25)
        run_case(case, 10, 20)


@pytest.mark.parametrize("plan", [[(128, 4)], [(64, 5), (64, 3)]])
def test_inner_product_search(plan):
    """IP: an unfilled pool's distk is -FLT_MAX (pool.hpp:32-34), so nothing is inserted and
    every result is 0; the metrics still count the scanned blocks."""
    case = F.make_case(n=2000, d=128, K=8, nq=12, plan=plan, seed=27)
    ids = run_case(case, 10, 4, dist_type=1)
    assert not ids.any()


def test_nprobe_per_query():
    """Query i scans the first nprobe_per_query[i] lists of its probe list: the same as an
    IVF::search call with that nprobe."""
    import dataclasses
    case = F.make_case(n=3000, d=128, K=16, nq=40, plan=[(64, 5), (64, 3)], seed=21)
    probes = F.orc_probes(case, 8)
    npq = np.random.default_rng(5).integers(1, 9, 40).astype(np.int32)
    ix = F.gpu_index(case)
    ids, d, met = ix.search(case.queries, 10, 8, probes=probes, nprobe_per_query=npq)
    for v in np.unique(npq):
        rows = np.nonzero(npq == v)[0]
        sub = dataclasses.replace(case, queries=case.queries[rows])
        ids_o, d_o, met_o = F.orc_search(sub, probes[rows, :v], 10)
        assert np.array_equal(ids[rows], ids_o), f"nprobe {v}"
        assert np.array_equal(d[rows].view(np.uint32), d_o.view(np.uint32))
        assert np.array_equal(met[rows], met_o)
    ix.close()


def test_gpu_probe_matches_oracle():
    case = F.make_case(n=3000, d=200, K=40, nq=32, plan=[(128, 4), (128, 4)], seed=17)
    ix = F.gpu_index(case)
    for nprobe in (1, 7, 40):
        want = F.orc_probes(case, nprobe)
        got = ix.probe(case.queries, nprobe)
        assert np.array_equal(want, got)
    # the search with GPU-chosen probes equals the search with the oracle's probes
    ids_a, _, _ = ix.search(case.queries, 10, 7)
    ids_b, _, _ = ix.search(case.queries, 10, 7, probes=F.orc_probes(case, 7))
    assert np.array_equal(ids_a, ids_b)
    ix.close()


@pytest.mark.parametrize("path", ["gemm", "tile"])
@pytest.mark.parametrize("d", [32, 40, 48, 64, 100, 960])
@pytest.mark.parametrize("dist_type", [0, 1])
def test_gpu_probe_matches_oracle_across_dims(d, dist_type, path, monkeypatch):
    """Both exact coarse quantisers give the oracle's lists for every D: the GEMM filter
    (candidates within the fp32 error bound, recomputed in Eigen's order) and the register-tiled
    probe (Eigen's order for every pair: the full 32-sum streams, the trailing 16-element packet
    (D mod 32 >= 16) and the scalar tail)."""
    monkeypatch.setenv("SAQGPU_PROBE", path)
    dp = (d + 63) // 64 * 64
    case = F.make_case(n=1500, d=d, K=300, nq=70, plan=[(dp, 2)], seed=19 + d)
    ix = F.gpu_index(case)
    for nprobe in (1, 33, 300):
        got = ix.probe(case.queries, nprobe, dist_type=dist_type)
        want = F.orc_probes(case, nprobe, dist_type=dist_type)
        assert np.array_equal(want, got), f"d={d} nprobe={nprobe}"
    ix.close()


@pytest.mark.parametrize("nq", [12, 80])
@pytest.mark.parametrize("path", ["gemm", "tile"])
@pytest.mark.parametrize("dist_type", [0, 1])
def test_gpu_probe_exact_ties(dist_type, path, nq, monkeypatch):
    """Exact distance ties come out in std::partial_sort's order (the oracle calls the same
    libstdc++ algorithm the reference does): duplicated centroids tie in pairs, and a zero query
    ties every centroid under IP, so the radix select has to reach the id bytes and the heap
    selection decides the order.  Up to 64 queries the GEMM path computes every exact distance
    on the whole chip before the tie pass, above it each flagged query's workgroup does."""
    monkeypatch.setenv("SAQGPU_PROBE", path)
    case = F.make_case(n=1500, d=96, K=300, nq=nq, plan=[(128, 2)], seed=23)
    case.centroids[150:] = case.centroids[:150]
    case.queries[0] = 0.0
    case.queries[5] = case.centroids[7]
    ix = F.gpu_index(case)
    for nprobe in (1, 2, 149, 150, 151, 256, 257, 300):
        got = ix.probe(case.queries, nprobe, dist_type=dist_type)
        want = F.orc_probes(case, nprobe, dist_type=dist_type)
        assert np.array_equal(want, got), f"nprobe={nprobe}"
    ix.close()


@pytest.mark.parametrize("dist_type", [0, 1])
def test_gpu_probe_candidate_fallback(dist_type, monkeypatch):
    """1500 identical centroids: every key ties, so the GEMM filter's candidate set exceeds its
    1024 places and the kernel recomputes all K in Eigen's order; the order is std::partial_sort's."""
    monkeypatch.setenv("SAQGPU_PROBE", "gemm")
    case = F.make_case(n=3000, d=64, K=1500, nq=6, plan=[(64, 2)], seed=29, encode=False)
    case.centroids[:] = case.centroids[0]
    from saq_amd import Index
    ix = Index(case.num_dim, case.plan, case.centroids, np.zeros(case.K, np.int64), case.variance)
    for nprobe in (1, 20, 1024):
        got = ix.probe(case.queries, nprobe, dist_type=dist_type)
        want = F.orc_probes(case, nprobe, dist_type=dist_type)
        assert np.array_equal(want, got), f"nprobe={nprobe}"
    ix.close()


@pytest.mark.parametrize("dist_type", [0, 1])
def test_gpu_probe_more_than_4096_lists(dist_type, monkeypatch):
    """K = 4500: the GEMM filter reads its keys again in every pass (they no longer fit 16 per
    thread); lists equal the oracle's."""
    monkeypatch.setenv("SAQGPU_PROBE", "gemm")
    from oracle_lib import oracle
    from saq_amd import Index
    O = oracle()
    rng = np.random.default_rng(31)
    K, d, nq = 4500, 96, 8
    cent = rng.standard_normal((K, d)).astype(np.float32)
    queries = rng.standard_normal((nq, d)).astype(np.float32)
    var = np.ones(128, np.float32)
    ix = Index(d, [(128, 2)], cent, np.zeros(K, np.int64), var)
    for nprobe in (1, 40):
        got = ix.probe(queries, nprobe, dist_type=dist_type)
        want = np.zeros((nq, nprobe), np.uint32)
        for i, q in enumerate(queries):
            O.orc_probe(cent, K, d, np.ascontiguousarray(q), nprobe, dist_type, want[i], None)
        assert np.array_equal(want, got), f"nprobe={nprobe}"
    ix.close()


def test_cluster_layout_round_trip():
    for plan, compact in (([(256, 4)], False), ([(64, 1), (64, 5), (128, 3)], False), ([(128, 5), (128, 0)], True)):
        case = F.make_case(n=1500, d=256, K=6, nq=4, plan=plan, seed=19, compact=compact)
        ix = F.gpu_index(case)
        for c in range(case.K):
            got = ix.get_cluster(c)
            ref = case.clusters[c]
            for key in ("short_factors", "short_code", "long_code", "long_factors", "centroid_rot"):
                assert np.array_equal(got[key], ref[key][:got[key].size]), (plan, c, key)
            assert np.array_equal(got["ids"], ref["ids"][:ref["num_vec"]])
        ix.close()


def test_search_handles_run_concurrently_and_agree():
    """Concurrent IVF::search callers (test_qps.cpp:77-84): four search handles, each searched from
    its own host thread on its own stream, return what the index returns; a handle cannot change
    the index, and the index refuses changes while handles are open."""
    import ctypes as C
    import threading
    from hip_mem import hip
    from saq_amd import SaqGpuError
    case = F.make_case(n=6000, d=128, K=24, nq=64, plan=[(64, 5), (64, 3)], seed=91)
    ix = F.gpu_index(case)
    want = ix.search(case.queries, 10, 8)
    H = hip()
    hs, streams = [], []
    for _ in range(4):
        hs.append(ix.search_handle())
        st = C.c_void_p()
        assert H.hipStreamCreate(C.byref(st)) == 0
        streams.append(st)
    out, errs = [None] * 4, []

    def run(k):
        try:
            out[k] = [hs[k].search(case.queries, 10, 8, stream=streams[k].value) for _ in range(3)]
        except Exception as e:  # reported below
            errs.append(e)

    threads = [threading.Thread(target=run, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errs, errs
    for k in range(4):
        for ids, d, met in out[k]:
            assert np.array_equal(ids, want[0])
            assert np.array_equal(d.view(np.uint32), want[1].view(np.uint32))
            assert np.array_equal(met, want[2])
    with pytest.raises(SaqGpuError, match="search handle"):
        hs[0].encode(case.base, case.cluster_ids)
    with pytest.raises(SaqGpuError, match="open search handles"):
        ix.encode(case.base, case.cluster_ids)
    for h, st in zip(hs, streams):
        h.close()
        assert H.hipStreamDestroy(st) == 0
    ix.encode(case.base, case.cluster_ids)
    again = ix.search(case.queries, 10, 8)
    assert np.array_equal(again[0], want[0])
    ix.close()


@pytest.mark.parametrize("dist_type", [0, 1])
def test_gpu_bf16x3_probe_select_more_than_4096_lists(dist_type, monkeypatch):
    """The batched scan's coarse quantiser at K = 5000 (keys read from global memory): the lists
    a query's threads' minimum keys bound, then the exact select over them in LDS.  Centroids
    c_j = R (t_j, 0, ...) with t_j a permutation of a 0.1 grid and queries R (x, 0, ...), x 0.03
    off the grid (R a random rotation), keep every two keys a query selects among at least 0.003
    apart (L2: 0.07^2 - 0.03^2; IP: 0.1 |x|), ten times the bf16x3 GEMM's error bound
    (2^-16 of 2 |x| |t|) and more, so the lists equal the float64 order.  Identical centroids (every key ties, more candidates
    than LDS holds) take the full select and come out in id order."""
    monkeypatch.setenv("SAQGPU_PROBE", "bf16x3")
    from saq_amd import Index
    rng = np.random.default_rng(37)
    K, d, nq = 5000, 64, 6
    R, _ = np.linalg.qr(rng.standard_normal((d, d)))
    t = (rng.permutation(K) - K // 2) * 0.1
    x = np.array([0.03, -0.47, 0.83, 0.13, -0.97, 0.53])
    cent = np.zeros((K, d)); cent[:, 0] = t
    qs = np.zeros((nq, d)); qs[:, 0] = x
    cent = (cent @ R.T).astype(np.float32)
    qs = (qs @ R.T).astype(np.float32)
    var = np.ones(64, np.float32)
    ix = Index(d, [(64, 2)], cent, np.zeros(K, np.int64), var)
    for nprobe in (1, 37, 256):
        got = ix.probe(qs, nprobe, dist_type=dist_type)
        key = (t[None, :] - x[:, None]) ** 2 if dist_type == 0 else -(t[None, :] * x[:, None])
        want = np.argsort(key, axis=1, kind="stable")[:, :nprobe].astype(np.uint32)
        assert np.array_equal(want, got), f"nprobe={nprobe}"
    ix.close()
    same = np.repeat(cent[:1], K, axis=0)
    ix = Index(d, [(64, 2)], same, np.zeros(K, np.int64), var)
    for nprobe in (1, 100, 256):
        got = ix.probe(qs, nprobe, dist_type=dist_type)
        assert np.array_equal(got, np.tile(np.arange(nprobe, dtype=np.uint32), (nq, 1))), f"nprobe={nprobe}"
    ix.close()
