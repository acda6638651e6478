"""The benchmark's CPU baseline (oracle/libcpubase.so: the oracle compiled with its fast scan on
AVX-512 byte shuffles) returns the oracle's IDs, distance bits and metrics, and its timed
batch search (test_qps's loop, one query per thread at a time) returns the same IDs."""
import ctypes as C

import numpy as np
import pytest

import saq_fixtures as F
from oracle_lib import Metrics, cpubase, oracle


def search(L, case, probes, topk):
    plan = case.orc_plan()
    nq, nprobe = probes.shape
    ids = np.zeros((nq, topk), np.uint32)
    dists = np.zeros((nq, topk), np.float32)
    mets = (Metrics * nq)()
    vpad = np.zeros(case.D_pad, np.float32)
    vpad[:case.num_dim] = case.variance
    L.orc_search_batch(C.byref(plan), vpad, F.orc_rot_ptrs(case), F.orc_clusters(case),
                       np.ascontiguousarray(probes, np.uint32), nprobe, np.ascontiguousarray(case.queries), nq, topk,
                       4.0, 0, 4, ids, dists.ctypes.data, C.cast(mets, C.c_void_p))
    return ids, dists, np.array([[x.fast_bitsum, x.acc_bitsum, x.total_comp_cnt] for x in mets], np.uint64)


@pytest.mark.parametrize("plan", [[(128, 4)], [(128, 1)], [(64, 9), (64, 5), (64, 0), (64, 2)], None])
@pytest.mark.parametrize("topk", [10, 100])
def test_avx512_baseline_equals_oracle(plan, topk):
    L = cpubase()
    if L is None:
        pytest.skip("host without AVX-512")
    d = 256 if plan is None else sum(x for x, _ in plan)
    case = F.make_case(n=4000, d=d, K=12, nq=24, plan=plan, avg_bits=4.0, seed=51, tau=32)
    probes = F.orc_probes(case, 6)
    want = search(oracle(), case, probes, topk)
    got = search(L, case, probes, topk)
    assert np.array_equal(want[0], got[0])
    assert np.array_equal(want[1].view(np.uint32), got[1].view(np.uint32))
    assert np.array_equal(want[2], got[2])
    # the timed loop (dynamic query order) gives the same IDs
    ids = np.zeros((case.queries.shape[0], topk), np.uint32)
    lat = np.zeros(case.queries.shape[0], np.float64)
    vpad = np.zeros(case.D_pad, np.float32)
    vpad[:case.num_dim] = case.variance
    L.orc_search_batch_timed(C.byref(case.orc_plan()), vpad, F.orc_rot_ptrs(case), F.orc_clusters(case),
                             np.ascontiguousarray(probes, np.uint32), 6, None, np.ascontiguousarray(case.queries),
                             case.queries.shape[0], topk, 4.0, 3, ids, None, lat)
    assert np.array_equal(want[0], ids) and (lat > 0).all()
