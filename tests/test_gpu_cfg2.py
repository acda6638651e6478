"""The plain GPU drop-in at N = 1,000,000, D = 960 (256-component mixture, std decay
1/sqrt(1 + i/32), PCA, nlist = 4096) for B = 1.5 (mixed-bit segments) and
B = 4, k = 10 and 100, nprobe 16, 64 and 256, against the oracle on a 64-query subset: probe
lists, IDs, distance bits and QueryRuntimeMetrics.  Encoded bytes of a sample of lists (the
largest, median ones, the smallest non-empty one) equal the oracle encoder's."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import bigcase as BC  # noqa: E402

pytestmark = pytest.mark.gpu

NQ = 64
_built = {}


@pytest.fixture(scope="module")
def ds():
    d = BC.Dataset(torch, 1_000_000, 960, 256, 32, 0.0, 4096, 1000, kmeans_iters=8)
    yield d
    for b in _built.values():
        b.close()
    _built.clear()
    del d.base
    torch.cuda.empty_cache()


def built(ds, bits):
    if bits not in _built:
        _built[bits] = BC.Built(ds, bits)
    return _built[bits]


def sample_lists(sizes):
    nz = np.nonzero(sizes)[0]
    order = nz[np.argsort(sizes[nz], kind="stable")]
    return sorted({int(order[-1]), int(order[len(order) // 2]), int(order[len(order) // 2 + 1]), int(order[0])})


@pytest.mark.parametrize("bits", [1.5, 4.0])
def test_cfg2_encoder_bytes(ds, bits):
    b = built(ds, bits)
    assert b.nfail == 0
    if bits == 1.5:  # fractional B: segments of different widths, 0-bit ones included
        assert len({x for _, x in b.plan}) >= 2, b.plan
    for c in sample_lists(ds.sizes):
        BC.log(f"cfg2 B={bits}: list {c} ({ds.sizes[c]} vectors) vs the oracle encoder")
        got = b.lists([c])[c]
        want = b.orc_encode_list(c)
        for k in ("short_factors", "short_code", "long_code", "long_factors", "ids", "centroid_rot"):
            assert np.array_equal(np.asarray(got[k]).view(np.uint8), np.asarray(want[k]).view(np.uint8)), \
                f"list {c} {k}"


@pytest.mark.parametrize("bits", [1.5, 4.0])
@pytest.mark.parametrize("topk", [10, 100])
@pytest.mark.parametrize("nprobe", [16, 64, 256])
def test_cfg2_search_equals_oracle(ds, bits, topk, nprobe):
    b = built(ds, bits)
    qidx = np.arange(NQ)
    BC.log(f"cfg2 B={bits} k={topk} nprobe={nprobe}: {NQ} queries")
    pr_o = b.orc_probes(qidx, nprobe)
    assert np.array_equal(b.ix.probe(ds.queries_np[:NQ], nprobe), pr_o)
    ids_g, d_g, met_g = b.ix.search(ds.queries_np[:NQ], topk, nprobe)
    ids_o, d_o, met_o = b.orc_search(qidx, pr_o, topk)
    bad = np.nonzero((ids_g != ids_o).any(1))[0]
    assert len(bad) == 0, f"ids differ in queries {bad[:5]}"
    assert np.array_equal(d_g.view(np.uint32), d_o.view(np.uint32))
    assert np.array_equal(met_g, met_o)
