"""IVF::estimate on the GPU (saqgpu_estimate) against the oracle's estimate: the accurate,
fast and variance estimates of every vector of the probed lists, in probe order, bit for bit
(the same fp32/fp64 operations in the same order)."""
import ctypes as C

import numpy as np
import pytest

import saq_fixtures as F
from oracle_lib import Metrics, oracle

pytestmark = pytest.mark.gpu


def oracle_estimate(case, probe, q, m, dist_type):
    O = oracle()
    n = int(sum(case.clusters[c]["num_vec"] for c in probe))
    ids = np.zeros(max(1, n), np.uint32)
    acc, fast, var = (np.zeros(max(1, n), np.float32) for _ in range(3))
    vpad = np.zeros(case.D_pad, np.float32)
    vpad[:case.num_dim] = case.variance
    met = Metrics()
    w = O.orc_estimate(C.byref(case.orc_plan()), vpad, F.orc_rot_ptrs(case), F.orc_clusters(case),
                       np.ascontiguousarray(probe, np.uint32), len(probe), np.ascontiguousarray(q), m, dist_type,
                       ids, acc, fast.ctypes.data, var.ctypes.data, C.cast(C.pointer(met), C.c_void_p))
    assert w == n
    return ids[:n], acc[:n], fast[:n], var[:n]


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.mark.parametrize("kw", [dict(d=128, plan=[(128, 4)]), dict(d=128, plan=[(128, 1)]),
                                dict(d=256, avg_bits=4.0, tau=32), dict(d=256, plan=[(64, 9), (64, 5), (64, 0), (64, 2)]),
                                dict(d=200, plan=[(128, 7), (128, 3)], rotate=False),
                                dict(d=256, plan=[(128, 6), (128, 3)], compact=True)])
@pytest.mark.parametrize("dist_type", [0, 1])
def test_estimate_matches_oracle(kw, dist_type):
    case = F.make_case(n=2000, K=8, nq=6, seed=71, **kw)
    ix = F.gpu_index(case)
    probes = F.orc_probes(case, 4, dist_type)
    for i, q in enumerate(case.queries):
        print(f"estimate case {kw} dt={dist_type} query {i}", flush=True)
        g = ix.estimate(q, 4, probes=probes[i], m=4.0, dist_type=dist_type)
        o = oracle_estimate(case, probes[i], q, 4.0, dist_type)
        assert np.array_equal(g[0], o[0])
        for k in (1, 2, 3):
            assert same(g[k], o[k]), f"query {i} output {k}"
    ix.close()


def test_estimate_gpu_probe_and_empty_lists():
    ids_in = np.zeros(600, np.uint32)
    ids_in[200:] = 2  # list 1 empty
    case = F.make_case(n=600, d=128, K=3, nq=3, plan=[(64, 4), (64, 2)], seed=73, cluster_ids=ids_in)
    ix = F.gpu_index(case)
    for q in case.queries:
        g = ix.estimate(q, 3)  # GPU coarse quantiser (exact order)
        o = oracle_estimate(case, ix.probe(q[None, :], 3)[0], q, 4.0, 0)
        assert np.array_equal(g[0], o[0]) and all(same(g[k], o[k]) for k in (1, 2, 3))
        assert len(g[0]) == 600
    ix.close()
