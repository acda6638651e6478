"""The north star's configuration at full size: N = 10,000,000, D = 1024, a 2,048-component
mixture with Zipf(1.1) weights (PCA, nlist = 16384, heavily uneven lists), 10,000 queries with
nprobe drawn per query from U[32, 256], k = 10.

* plain scan, B = 4: all 10,000 queries searched; a sample equals the oracle bit for bit (IDs,
  distances, metrics) with the same per-query nprobe;
* batched int8-MFMA scan, B = 4 and B = 2: all 10,000 queries with the exact probe lists; every
  row is sorted by (distance, id) without repeats; sampled queries equal the numpy model of the
  batched estimator (tests/batched_model.py) fed with codes the oracle decodes from the index;
  recall@10 against exact FP64 neighbours is no lower than the plain scan's;
* encoder (the precision-and-encoder configuration's codes, same N and D): at B = 4 and B = 8,
  whole lists equal the oracle encoder's bytes.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import batched_model as BM  # noqa: E402
import bigcase as BC  # noqa: E402
from saq_amd import PREC_INT8  # noqa: E402

pytestmark = pytest.mark.gpu

_built = {}
_plain = {}
SAMPLE = [0, 1, 2, 3, 4, 5, 6, 7, 4096, 9999]


@pytest.fixture(scope="module")
def ds():
    d = BC.Dataset(torch, 10_000_000, 1024, 2048, 32, 1.1, 16384, 10000, kmeans_iters=6, gt_k=10)
    d.npq = np.random.default_rng(3).integers(32, 257, d.nq).astype(np.int32)
    yield d
    for b in _built.values():
        b.close()
    _built.clear()
    _plain.clear()
    del d.base
    torch.cuda.empty_cache()


def built(ds, bits):
    if bits not in _built:
        _built[bits] = BC.Built(ds, bits)
    return _built[bits]


def plain(ds, bits):
    b = built(ds, bits)
    if bits not in _plain:
        BC.log(f"cfg3 B={bits}: plain search of {ds.nq} queries, nprobe U[32, 256]")
        _plain[bits] = b.ix.search(ds.queries_np, 10, 256, nprobe_per_query=ds.npq)
    return _plain[bits]


def test_cfg3_plain_scan_equals_oracle(ds):
    b = built(ds, 4.0)
    assert b.nfail == 0
    ids_g, d_g, met_g = plain(ds, 4.0)
    qidx = np.array(SAMPLE)
    pr_o = b.orc_probes(qidx, 256)
    assert np.array_equal(b.ix.probe(ds.queries_np[qidx], 256), pr_o)
    BC.log(f"cfg3: oracle search of {len(qidx)} sampled queries")
    ids_o, d_o, met_o = b.orc_search(qidx, pr_o, 10, npq=ds.npq[qidx])
    assert np.array_equal(ids_g[qidx], ids_o)
    assert np.array_equal(d_g[qidx].view(np.uint32), d_o.view(np.uint32))
    assert np.array_equal(met_g[qidx], met_o)
    BC.log(f"cfg3 B=4: plain recall@10 {BC.recall(ids_g, ds.gt, 10):.4f}")


@pytest.mark.parametrize("bits", [4.0, 2.0])
def test_cfg3_batched_full_size(ds, bits):
    b = built(ds, bits)
    BC.log(f"cfg3 B={bits}: exact probe lists of {ds.nq} queries")
    probes = b.ix.probe(ds.queries_np, 256)
    BC.log(f"cfg3 B={bits}: batched search of {ds.nq} queries (int8)")
    ids, dists = b.ix.search_batched(ds.queries_np, 10, 256, probes=probes, nprobe_per_query=ds.npq,
                                     precision=PREC_INT8)
    BC.rows_well_formed(ids, dists)
    qidx = np.array(SAMPLE[:4])
    need = np.concatenate([probes[q, :ds.npq[q]] for q in qidx])
    BC.log(f"cfg3 B={bits}: model of {len(qidx)} sampled queries over {len(set(need.tolist()))} lists")
    tabs = b.tables(need)
    qr = b.rotated_queries(qidx)
    res, alld = BM.model_search(b.model_view(), tabs, qr, probes[qidx], 10, qs=1, npq=ds.npq[qidx])
    BC.check_batched(res, alld, ids[qidx], dists[qidx])
    r_b = BC.recall(ids, ds.gt, 10)
    r_p = BC.recall(plain(ds, bits)[0], ds.gt, 10)
    BC.log(f"cfg3 B={bits}: recall@10 batched {r_b:.4f} plain {r_p:.4f}")
    assert r_b >= r_p


@pytest.mark.parametrize("bits", [4.0, 8.0])
def test_cfg4_encoder_lists_equal_oracle(ds, bits):
    b = built(ds, bits)
    assert b.nfail == 0
    sizes = ds.sizes
    nz = np.nonzero(sizes)[0]
    order = nz[np.argsort(sizes[nz], kind="stable")]
    picks = {int(order[0]), int(order[len(order) // 2]), int(order[int(len(order) * 0.9)])}
    m32 = [int(c) for c in order if sizes[c] % 32 == 0 and 256 <= sizes[c] <= 2048]
    if m32:
        picks.add(m32[0])
    for c in sorted(picks):
        BC.log(f"cfg4 B={bits}: list {c} ({sizes[c]} vectors, D 1024) vs the oracle encoder")
        got = b.lists([c])[c]
        want = b.orc_encode_list(c)
        for k in ("short_factors", "short_code", "long_code", "long_factors", "ids", "centroid_rot"):
            assert np.array_equal(np.asarray(got[k]).view(np.uint8), np.asarray(want[k]).view(np.uint8)), \
                f"list {c} {k}"
