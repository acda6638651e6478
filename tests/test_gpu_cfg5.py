"""The serving configuration at full size: N = 5,000,000, D = 1536 (512-component mixture, std
decay 1/sqrt(1 + i/64), PCA, nlist = 8192), B = 1.5, the index resident on the device.  Requests
of 1..64 queries with k in {10, 100} and one nprobe in [16, 128] each are batched as the serving
loop batches them (k = the batch's largest, nprobe per query) onto the batched scan; sampled
queries of every batch equal the batched estimator's numpy model, every row is well formed, and
the plain scan of sampled queries equals the oracle bit for bit at their nprobe."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import batched_model as BM  # noqa: E402
import bigcase as BC  # noqa: E402
from saq_amd import PREC_INT8  # noqa: E402

pytestmark = pytest.mark.gpu

_built = {}


@pytest.fixture(scope="module")
def ds():
    d = BC.Dataset(torch, 5_000_000, 1536, 512, 64, 0.0, 8192, 4096, kmeans_iters=6)
    yield d
    for b in _built.values():
        b.close()
    _built.clear()
    del d.base
    torch.cuda.empty_cache()


def built(ds):
    if "b" not in _built:
        _built["b"] = BC.Built(ds, 1.5)
    return _built["b"]


def requests(seed, nreq):
    rng = np.random.default_rng(seed)
    nqs = rng.integers(1, 65, nreq)
    ks = np.where(rng.random(nreq) < 0.5, 10, 100)
    nps = rng.integers(16, 129, nreq)
    starts = rng.integers(0, 4096 - 64, nreq)
    qidx = np.concatenate([np.arange(s, s + n) for s, n in zip(starts, nqs)])
    npq = np.concatenate([np.full(n, p, np.int32) for n, p in zip(nqs, nps)])
    return qidx, npq, int(ks.max())


@pytest.mark.parametrize("seed", [5, 6])
def test_cfg5_served_batches_equal_model(ds, seed):
    b = built(ds)
    qidx, npq, k = requests(seed, 12)
    BC.log(f"cfg5: batch of {len(qidx)} queries from 12 requests, k {k}, nprobe {npq.min()}..{npq.max()}")
    q = np.ascontiguousarray(ds.queries_np[qidx])
    probes = b.ix.probe(q, int(npq.max()))
    ids, dists = b.ix.search_batched(q, k, int(npq.max()), probes=probes, nprobe_per_query=npq, precision=PREC_INT8)
    BC.rows_well_formed(ids, dists)
    pick = np.array([0, len(qidx) // 2, len(qidx) - 1])
    need = np.concatenate([probes[i, :npq[i]] for i in pick])
    tabs = b.tables(need)
    qr = b.rotated_queries(qidx[pick])
    res, alld = BM.model_search(b.model_view(), tabs, qr, probes[pick], k, qs=1, npq=npq[pick])
    BC.check_batched(res, alld, ids[pick], dists[pick])


def test_cfg5_plain_scan_equals_oracle(ds):
    b = built(ds)
    qidx = np.arange(0, 4096, 512)
    npq = np.random.default_rng(9).integers(16, 129, len(qidx)).astype(np.int32)
    for k in (10, 100):
        BC.log(f"cfg5: plain search of {len(qidx)} queries, k {k}")
        ids_g, d_g, met_g = b.ix.search(ds.queries_np[qidx], k, 128, nprobe_per_query=npq)
        pr_o = b.orc_probes(qidx, 128)
        ids_o, d_o, met_o = b.orc_search(qidx, pr_o, k, npq=npq)
        assert np.array_equal(ids_g, ids_o)
        assert np.array_equal(d_g.view(np.uint32), d_o.view(np.uint32))
        assert np.array_equal(met_g, met_o)
