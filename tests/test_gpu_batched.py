"""The batched int8-MFMA scan (saqgpu_search_batched) against the numpy model of the same
estimator (tests/batched_model.py), and against the plain scan's recall.

Bar: every returned distance equals the model's distance of that vector to 1e-5 relative
(BF16 queries: plus the model's bound for the float32 sums of the matrix cores); the returned
(distance, id) list is sorted and its IDs are the model's top-k except where the model's own
distances tie within that tolerance at the k-th place; results do not depend on the query
tile; recall@10 against exact FP64 neighbours is at least the plain scan's (int8x2).
"""
import numpy as np
import pytest

import batched_model as BM
import saq_fixtures as F
from saq_amd import PREC_BF16, PREC_INT8, PREC_INT8X2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def model(case, probes, topk, qs, npq=None):
    tabs = BM.vector_tables(case)
    qr = BM.rotated_queries(case)
    return BM.model_search(case, tabs, qr, probes, topk, qs=qs, npq=npq, with_err=True)


def check(case, res, alld, alle, ids, dists, topk):
    for i in range(ids.shape[0]):
        want = res[i]
        n = len(want)
        got_ids = ids[i, :n].tolist()
        got_d = dists[i, :n]
        assert np.all(ids[i, n:] == 0) and np.all(dists[i, n:] == 0), f"query {i}: unfilled slots"
        for u, dv in zip(got_ids, got_d.tolist()):
            assert u in alld[i], f"query {i}: id {u} not in a probed list"
            ref = alld[i][u]
            tol = TOL * max(1.0, abs(ref)) + alle[i][u]
            assert abs(dv - ref) <= tol, f"query {i} id {u}: {dv} vs model {ref} (bound {tol})"
        pairs = list(zip(got_d.tolist(), got_ids))
        assert pairs == sorted(pairs), f"query {i}: not sorted by (distance, id)"
        if n:
            kth, kid = want[-1][1], want[-1][0]
            for u in set(got_ids) ^ {u for u, _ in want}:
                tol = 2 * TOL * max(1.0, abs(kth)) + 2 * (alle[i][u] + alle[i][kid])
                assert abs(alld[i][u] - kth) <= tol, f"query {i}: id {u} differs off a tie"


def run(case, topk, nprobe, qs, npq=None, query_tile=0):
    probes = F.orc_probes(case, nprobe)
    res, alld, alle = model(case, probes, topk, qs, npq)
    ix = F.gpu_index(case)
    print(f"batched case plan={case.plan} k={topk} nprobe={nprobe} prec={qs} tile={query_tile}", flush=True)
    ids, d = ix.search_batched(case.queries, topk, nprobe, probes=probes, nprobe_per_query=npq,
                               precision=qs, query_tile=query_tile)
    check(case, res, alld, alle, ids, d, topk)
    ix.close()
    return ids, d


@pytest.mark.parametrize("plan", [[(128, 4)], [(128, 7)], [(128, 8)], [(128, 12)], [(64, 1), (64, 3)],
                                  [(64, 9), (64, 5), (64, 2), (64, 0)]])
@pytest.mark.parametrize("qs", [PREC_INT8, PREC_INT8X2, PREC_BF16])
def test_batched_matches_model(plan, qs):
    d = sum(x for x, _ in plan)
    case = F.make_case(n=2500, d=d, K=10, nq=24, plan=plan, seed=51, tau=32)
    run(case, 10, 4, qs)


def test_dp_plan_and_padded_dim():
    case = F.make_case(n=3000, d=200, K=12, nq=20, avg_bits=4.0, seed=53, tau=32)
    run(case, 10, 5, PREC_INT8X2)


@pytest.mark.parametrize("topk", [1, 65, 100, 128, 200])
def test_topk_sizes_and_unfilled(topk):
    case = F.make_case(n=1200, d=128, K=8, nq=12, plan=[(64, 6), (64, 3)], seed=55)
    run(case, topk, 2, PREC_INT8X2)


def test_query_tiles_give_identical_results():
    # 4 lists and 200 queries: several full tiles per list at every tile size
    case = F.make_case(n=3000, d=128, K=4, nq=200, plan=[(128, 5)], seed=57)
    for prec in (PREC_INT8, PREC_INT8X2, PREC_BF16):
        out = [run(case, 10, 2, prec, query_tile=t) for t in (16, 32, 64, 128)]
        for ids, d in out[1:]:
            assert np.array_equal(ids, out[0][0])
            assert np.array_equal(d.view(np.uint32), out[0][1].view(np.uint32))


@pytest.mark.parametrize("topk,tiles,prec", [(10, (16, 64, 128), PREC_INT8), (100, (16, 32, 64), PREC_INT8),
                                             (128, (16, 64), PREC_INT8X2), (10, (16, 128), PREC_BF16)])
def test_long_lists_split_into_vector_ranges(topk, tiles, prec):
    """Lists of ~4000 vectors: each (list, query tile) becomes 16 work tiles of 256-vector
    ranges with candidate lists of their own, which the merge reads back per pair."""
    case = F.make_case(n=8000, d=128, K=2, nq=150, plan=[(64, 6), (64, 3)], seed=63)
    out = [run(case, topk, 2, prec, query_tile=t) for t in tiles]
    for ids, d in out[1:]:
        assert np.array_equal(out[0][0], ids)
        assert np.array_equal(out[0][1].view(np.uint32), d.view(np.uint32))


def test_uneven_nprobe_per_query():
    case = F.make_case(n=2000, d=128, K=16, nq=40, plan=[(64, 4), (64, 4)], seed=59)
    npq = np.random.default_rng(3).integers(1, 9, 40).astype(np.int32)
    run(case, 10, 8, PREC_INT8X2, npq=npq)


def test_empty_lists_and_gpu_probe():
    ids_in = np.zeros(900, np.uint32)
    ids_in[300:600] = 2
    ids_in[600:] = 4  # lists 1 and 3 are empty
    case = F.make_case(n=900, d=128, K=5, nq=10, plan=[(128, 4)], seed=61, cluster_ids=ids_in)
    ids_a, d_a = run(case, 20, 5, PREC_INT8X2)
    ix = F.gpu_index(case)
    # GPU coarse quantiser in the reference's order: the same lists, the same results
    ids_b, d_b = ix.search_batched(case.queries, 20, 5, precision=PREC_INT8X2, exact_probe=True)
    assert np.array_equal(ids_a, ids_b) and np.array_equal(d_a.view(np.uint32), d_b.view(np.uint32))
    ix.close()


def test_fast_probe_selects_the_same_lists_without_near_ties():
    """The GEMM-based coarse quantiser differs from the exact one only where two centroid
    distances tie within fp32 rounding; on this data no query has such a tie at its nprobe-th
    list, so every result equals the exact-probe result."""
    case = F.make_case(n=4000, d=200, K=24, nq=64, plan=[(128, 5), (128, 3)], seed=67)
    ix = F.gpu_index(case)
    for nprobe in (1, 4, 11):
        a, da = ix.search_batched(case.queries, 10, nprobe, exact_probe=True)
        b, db = ix.search_batched(case.queries, 10, nprobe, exact_probe=False)
        assert np.array_equal(a, b), f"nprobe={nprobe}"
        assert np.array_equal(da.view(np.uint32), db.view(np.uint32))
    ix.close()


def test_autotune_keeps_results():
    case = F.make_case(n=3000, d=128, K=6, nq=100, plan=[(128, 4)], seed=63)
    probes = F.orc_probes(case, 3)
    ix = F.gpu_index(case)
    a, _ = ix.search_batched(case.queries, 10, 3, probes=probes)
    best = ix.autotune_batched(case.queries, 10, 3, probes=probes)
    assert best in (16, 32, 64, 128)
    b, _ = ix.search_batched(case.queries, 10, 3, probes=probes)
    assert np.array_equal(a, b)
    ix.close()


def test_recall_not_below_plain_scan():
    case = F.make_case(n=20000, d=128, K=64, nq=100, plan=[(128, 4)], seed=65, comps=32, tau=16)
    X = case.base.astype(np.float64)
    Q = case.queries.astype(np.float64)
    dd = (Q * Q).sum(1)[:, None] + (X * X).sum(1)[None, :] - 2 * Q @ X.T
    gt = np.argsort(dd, axis=1)[:, :10]
    ix = F.gpu_index(case)
    plain, _, _ = ix.search(case.queries, 10, 16)
    bat, _ = ix.search_batched(case.queries, 10, 16, precision=PREC_INT8X2)
    r_plain = np.mean([len(set(plain[i]) & set(gt[i])) / 10 for i in range(100)])
    r_bat = np.mean([len(set(bat[i]) & set(gt[i])) / 10 for i in range(100)])
    print(f"recall@10 plain {r_plain:.3f} batched {r_bat:.3f}")
    assert r_bat >= r_plain
    ix.close()


def test_graph_replays_equal_eager_calls():
    """SAQGPU_USE_GRAPH: the first call captures the batched step into a HIP graph; later calls with
    the same buffers replay it.  Each replay -- with new query contents written into the same device
    buffer and eager work on the GPU in between -- equals an eager call on the same queries, for the
    GEMM coarse quantiser (a library GEMM inside the graph) and the exact one."""
    import ctypes as C
    from hip_mem import DevArray, hip
    case = F.make_case(n=4000, d=200, K=24, nq=64, plan=[(128, 5), (128, 3)], seed=71)
    ix = F.gpu_index(case)
    H = hip()
    stream = C.c_void_p()
    assert H.hipStreamCreate(C.byref(stream)) == 0
    rng = np.random.default_rng(5)
    for exact in (False, True):
        qbuf = DevArray(case.queries)
        ids = DevArray(shape=(64, 10), dtype=np.uint32)
        dd = DevArray(shape=(64, 10), dtype=np.float32)
        for rep in range(4):
            q = case.queries[rng.permutation(64)].copy()
            assert H.hipMemcpy(C.c_void_p(qbuf.data_ptr()), q.ctypes.data, q.nbytes, 1) == 0
            print(f"graph replay exact={exact} rep={rep}", flush=True)
            ix.search_batched(qbuf, 10, 11, exact_probe=exact, out_ids=ids, out_dists=dd, stream=stream.value,
                              graph=True)
            assert H.hipStreamSynchronize(stream) == 0
            got_i, got_d = ids.numpy(), dd.numpy()
            want_i, want_d = ix.search_batched(q, 10, 11, exact_probe=exact)
            assert np.array_equal(got_i, want_i), f"exact={exact} rep={rep}"
            assert np.array_equal(got_d.view(np.uint32), want_d.view(np.uint32))
            ix.search(case.queries, 10, 4)  # other GPU work between replays
        qbuf.free()
        ids.free()
        dd.free()
    assert H.hipStreamDestroy(stream) == 0
    ix.close()
