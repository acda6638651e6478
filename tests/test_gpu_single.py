"""Single-vector records on the GPU (SAQuantizerSingle / SaqSingleEstimator, the reference's
unit_test/ut_single_estimator.cpp path) against the oracle, bit for bit: the records' bytes and
the three estimates of every (query, record) pair."""
import numpy as np
import pytest

import saq_fixtures as F

pytestmark = pytest.mark.gpu

PLANS = [
    (200, None, True, 300),                           # the DP plan, ragged num_dim, rotated
    (256, [(64, 0), (64, 1), (128, 5)], True, 300),   # 0-bit, 1-bit and a 5-bit (IP4 + IP1) segment
    (128, [(128, 9)], False, 300),                    # 9 bits (IP8 + 256 IP1), no rotation
    (256, [(64, 3), (192, 7)], True, 300),
    (1536, [(1536, 13)], True, 100),                  # records too large to stage: read in place
]


def _case(d, plan, rotate, n=300, seed=7):
    return F.make_case(n=n, d=d, K=4, nq=5, plan=plan, avg_bits=4.0, seed=seed, rotate=rotate, encode=False)


@pytest.mark.parametrize("d,plan,rotate,n", PLANS)
def test_single_records_match_oracle(d, plan, rotate, n):
    case = _case(d, plan, rotate, n=n)
    ix = F.gpu_index(case)
    try:
        nbytes, code_off, long_off, exf_off, segs = ix.single_layout()
        lay = F.single_layout(case)
        assert (nbytes, code_off, long_off, exf_off) == (lay.bytes, lay.code_off, lay.long_off, lay.exfac_off)
        for s, (c, l) in enumerate(segs):
            assert (c, l) == (lay.seg_code_off[s], lay.seg_long_off[s])
        want, stride, nf0 = F.orc_single_quantize(case, case.base)
        got, stride2, nf1 = ix.single_quantize(case.base, rd_lmt=case.rd_lmt, eps=case.eps, stride=stride)
        assert stride2 == stride and nf0 == nf1
        assert np.array_equal(got[:len(want)], want)
        for dist_type in (0, 1):
            v, f, a = ix.single_estimate(case.queries, got, case.base.shape[0], stride, 4.0, dist_type)
            for qi, q in enumerate(case.queries):
                wv, wf, wa = F.orc_single_estimate(case, want, stride, q, 4.0, dist_type)
                assert np.array_equal(a[qi].view(np.uint32), wa.view(np.uint32)), (dist_type, qi)
                assert np.array_equal(f[qi].view(np.uint32), wf.view(np.uint32)), (dist_type, qi)
                assert np.array_equal(v[qi].view(np.uint32), wv.view(np.uint32)), (dist_type, qi)
    finally:
        ix.close()


def test_single_edge_cases():
    case = _case(200, None, True, n=40)
    ix = F.gpu_index(case)
    try:
        nbytes = ix.single_layout()[0]
        # no rows, one row, a wider stride than the record
        got, stride, nf = ix.single_quantize(case.base[:0])
        assert nf == 0
        want, stride, _ = F.orc_single_quantize(case, case.base[:1], stride=nbytes + 8 - nbytes % 8 + 64)
        got, _, _ = ix.single_quantize(case.base[:1], stride=stride)
        assert np.array_equal(got[:stride], want)
        v, f, a = ix.single_estimate(case.queries[:0], got, 1, stride)
        assert a.shape == (0, 1)
        v, f, a = ix.single_estimate(case.queries, got, 0, stride)
        assert a.shape == (case.queries.shape[0], 0)
        with pytest.raises(Exception):
            ix.single_quantize(case.base[:2], stride=nbytes - 8)
        with pytest.raises(Exception):
            ix.single_estimate(case.queries, got, 1, stride + 4)
    finally:
        ix.close()


def test_single_device_pointers_and_size():
    """Device buffers in and out at D = 1024 with 20k records: the same bytes as the host call,
    and the reference test's expectations on the estimates (ut_single_estimator.cpp:180-190)."""
    from hip_mem import DevArray
    case = F.make_case(n=20000, d=1024, K=4, nq=4, avg_bits=4.0, seed=11, rotate=True, encode=False)
    ix = F.gpu_index(case)
    bufs = []
    try:
        host, stride, _ = ix.single_quantize(case.base)
        n, nq = case.base.shape[0], case.queries.shape[0]
        data = DevArray(case.base)
        recs = DevArray(shape=(n * stride,), dtype=np.uint8)
        q = DevArray(case.queries)
        outs = tuple(DevArray(shape=(nq, n), dtype=np.float32) for _ in range(3))
        bufs += [data, recs, q, *outs]
        ix.single_quantize(data, stride=stride, out=recs)
        assert np.array_equal(recs.numpy(), host[:n * stride])
        ix.single_estimate(q, recs, n, stride, out=outs)
        v, f, a = (o.numpy() for o in outs)
        hv, hf, ha = ix.single_estimate(case.queries, host, n, stride)
        assert np.array_equal(a, ha) and np.array_equal(f, hf) and np.array_equal(v, hv)
        for qi, qq in enumerate(case.queries):
            true = ((case.base.astype(np.float64) - qq) ** 2).sum(1)
            assert (a[qi] >= 0).all() and (f[qi] >= 0).all() and (v[qi] >= 0).all()
            assert np.mean(np.abs(a[qi] - true) / true) < 0.2
        # a sample of records against the oracle at full dimension
        sel = np.arange(0, n, 97)
        sub = host.reshape(n, stride)[sel].reshape(-1)
        want, _, _ = F.orc_single_quantize(case, case.base[sel], stride=stride)
        assert np.array_equal(sub, want)
        wv, wf, wa = F.orc_single_estimate(case, sub, stride, case.queries[0])
        assert np.array_equal(a[0, sel], wa) and np.array_equal(f[0, sel], wf) and np.array_equal(v[0, sel], wv)
    finally:
        for b in bufs:
            b.free()
        ix.close()
