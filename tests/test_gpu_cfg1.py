"""The plain GPU scan at the reference CPU workload's full size (N = 100,000 base + 1,000 query
vectors, D = 128, a 64-component mixture with per-dim std 1/sqrt(1 + i/8), PCA, IVF nlist = 256,
SAQ B = 4, k = 10, nprobe = 32) against the oracle: every list's encoded bytes, every query's
probe list, IDs, distance bits and QueryRuntimeMetrics, for all 1,000 queries.  Recall@10
against exact FP64 neighbours is reported.  The data is generated on the GPU (torch, imported
here at collection so that the process runs one HIP runtime)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import bigcase as BC  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ds():
    d = BC.Dataset(torch, 100_000, 128, 64, 8, 0.0, 256, 1000, kmeans_iters=10, gt_k=10)
    yield d
    del d.base
    torch.cuda.empty_cache()


def test_cfg1_full_size_plain_scan_equals_oracle(ds):
    b = BC.Built(ds, 4.0)
    assert b.plan == [(128, 4)]  # D = 128, B = 4: one 128-d segment (saq_data.hpp:152)
    assert b.nfail == 0
    BC.log("cfg1: encoder bytes of all 256 lists vs the oracle encoder")
    got = b.lists(range(ds.nlist))
    for c in range(ds.nlist):
        want = b.orc_encode_list(c)
        for k in ("short_factors", "short_code", "long_code", "long_factors", "ids", "centroid_rot"):
            assert np.array_equal(np.asarray(got[c][k]).view(np.uint8), np.asarray(want[k]).view(np.uint8)), \
                f"list {c} {k}"
    BC.log("cfg1: coarse quantiser, 1000 queries, nprobe 32")
    pr_g = b.ix.probe(ds.queries_np, 32)
    pr_o = b.orc_probes(range(ds.nq), 32)
    assert np.array_equal(pr_g, pr_o)
    BC.log("cfg1: search, 1000 queries, k 10, nprobe 32")
    ids_g, d_g, met_g = b.ix.search(ds.queries_np, 10, 32)
    ids_o, d_o, met_o = b.orc_search(np.arange(ds.nq), pr_o, 10)
    bad = np.nonzero((ids_g != ids_o).any(1))[0]
    assert len(bad) == 0, f"ids differ in queries {bad[:5]}"
    assert np.array_equal(d_g.view(np.uint32), d_o.view(np.uint32))
    assert np.array_equal(met_g, met_o)
    r = BC.recall(ids_g, ds.gt, 10)
    BC.log(f"cfg1: recall@10 {r:.4f} (GPU = oracle for all 1000 queries)")
    assert r > 0.5
    b.close()
