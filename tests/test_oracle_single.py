"""The oracle's single-vector quantiser and estimator (SAQuantizerSingle, SaqSingleEstimator),
pinned three ways:

- the reference's own bit-plane kernels (utils/space.hpp, compiled from the reference tree into
  oracle/_ref) give the same bits and floats on random inputs;
- closed forms: the plane product equals the integer sum of code bits times quantised query,
  the short code unpacks to the MSB of each code, the accurate estimate equals
  |o|^2 + |q|^2 - 2 * rescale * <oa, q> with oa the dequantised code;
- the reference test's own expectations (unit_test/ut_single_estimator.cpp:180-190): every
  estimate is non-negative for L2 and the accurate one is within 20% on average.
"""
import ctypes as C

import numpy as np
import pytest

import saq_fixtures as F
from oracle_lib import CaqCfg, CodeInfo, oracle, ref


def _ref_or_skip():
    R = ref()
    if R is None:  # built only where the reference tree is present
        pytest.skip("reference kernels (oracle/_ref) not built here")
    return R


def _bits_msb_first(words, D):
    """Bit 63 - k of word b is dim 64 b + k."""
    out = np.zeros(D, np.uint8)
    for d in range(D):
        out[d] = (int(words[d // 64]) >> (63 - d % 64)) & 1
    return out


@pytest.mark.parametrize("D", [64, 128, 192, 512, 1024])
def test_bit_planes_match_reference(D):
    R = _ref_or_skip()
    O = oracle()
    rng = np.random.default_rng(D)
    for _ in range(4):
        q = rng.integers(0, 256, D).astype(np.uint16)
        t_ref = np.zeros(D // 64 * 8, np.uint64)
        t_orc = np.zeros_like(t_ref)
        R.ref_new_transpose_bin(q, D, t_ref)
        O.orc_single_transpose_bin(q, D, t_orc)
        assert np.array_equal(t_ref, t_orc)
        x = rng.integers(0, 2**63, D // 64, dtype=np.uint64) * 2 + rng.integers(0, 2, D // 64, dtype=np.uint64)
        delta, vl = np.float32(rng.uniform(1e-3, 1)), np.float32(rng.uniform(-2, 0))
        a = R.ref_warmup_ip_x0_q(x, t_ref, delta, vl, D)
        b = O.orc_single_warmup_ip(x, t_orc, delta, vl, D)
        assert np.float32(a).tobytes() == np.float32(b).tobytes()
        # closed form: sum over dims of bit * q, and the popcount
        bits = _bits_msb_first(x, D)
        ip = int((bits.astype(np.int64) * q.astype(np.int64)).sum())
        ppc = int(bits.sum())
        want = np.float32(np.float32(delta * np.float32(ip)) + np.float32(vl * np.float32(ppc)))
        assert np.float32(b) == want
        qf = rng.standard_normal(D).astype(np.float32)
        m_ref = R.ref_mask_ip_x0_q(qf, x, D)
        m_orc = O.orc_single_mask_ip(qf, x, D)
        assert np.float32(m_ref).tobytes() == np.float32(m_orc).tobytes()
        assert abs(m_orc - float(qf.astype(np.float64) @ bits)) < 1e-4 * (1 + np.abs(qf).sum())


def test_query_delta_matches_reference():
    R = _ref_or_skip()
    rng = np.random.default_rng(5)
    for _ in range(2000):
        vl, vr = sorted(rng.standard_normal(2).astype(np.float32))
        want = np.float32(np.float64(np.float32(vr - vl)) / 255.99)
        assert np.float32(R.ref_single_delta(vl, vr)) == want


def _case(d=200, plan=None, rotate=True, n=300, seed=3):
    return F.make_case(n=n, d=d, K=4, nq=3, plan=plan, avg_bits=4.0, seed=seed, rotate=rotate, encode=False)


def _decode(case, recs, stride, i):
    """Per segment: o_l2norm, ip_cent_oa, codes (MSB from the short code, low bits from the long
    code), ExFactor."""
    O = oracle()
    lay = F.single_layout(case)
    rec = recs[i * stride:(i + 1) * stride]
    fac = rec[lay.fac_off:lay.fac_off + 8 * len(case.plan)].view(np.float32)
    exf = rec[lay.exfac_off:lay.exfac_off + 8 * len(case.plan)].view(np.float32)
    out = []
    for s, (D, b) in enumerate(case.plan):
        codes = None
        if b > 0:
            raw = rec[lay.seg_code_off[s]:lay.seg_code_off[s] + D // 8]
            # bytes of every 8 reversed (quantizer.hpp:185), each byte MSB-first
            byts = raw.reshape(-1, 8)[:, ::-1].reshape(-1)
            msb = np.unpackbits(byts).astype(np.int64)
            codes = msb << (b - 1)
            if b > 1:
                lc = np.ascontiguousarray(rec[lay.seg_long_off[s]:lay.seg_long_off[s] + D * (b - 1) // 8])
                codes = codes + np.array([O.orc_code_at(b - 1, lc, D, d) for d in range(D)], np.int64)
        out.append((fac[2 * s], fac[2 * s + 1], codes, exf[2 * s], exf[2 * s + 1]))
    return out


@pytest.mark.parametrize("plan", [None, [(64, 0), (64, 1), (128, 5)], [(128, 9)], [(64, 3), (192, 7)]])
def test_single_quantize_layout_round_trip(plan):
    case = _case(d=256 if plan else 200, plan=plan)
    O = oracle()
    recs, stride, nfail = F.orc_single_quantize(case, case.base[:40])
    lay = F.single_layout(case)
    S = len(case.plan)
    code_b = sum(D // 8 for D, b in case.plan if b > 0)
    long_b = sum(D * (b - 1) // 8 for D, b in case.plan if b > 1)
    r64 = lambda x: (x + 63) // 64 * 64
    assert lay.bytes == r64(8 * S) + r64(code_b) + long_b + 8 * S  # single_data.hpp:287-320
    cfg = CaqCfg(case.rd_lmt, case.eps, 1)
    for i in range(40):
        dec = _decode(case, recs, stride, i)
        off = 0
        for s, (D, b) in enumerate(case.plan):
            x = np.zeros(D, np.float32)
            c = min(D, case.num_dim - off)
            x[:c] = case.base[i, off:off + c]
            o = np.zeros(D, np.float32)
            O.orc_gemv_rowvec(x, case.rotations[s], D, o)
            code = np.zeros(D, np.uint16)
            info = CodeInfo()
            O.orc_caq_encode(o, D, b, C.byref(cfg), None, code, C.byref(info))
            onorm, ipc, codes, resc, err = dec[s]
            assert onorm == np.float32(info.o_l2norm)
            assert abs(onorm - np.linalg.norm(o.astype(np.float64))) < 1e-5 * (1 + onorm)
            if b == 0:
                assert ipc == 0 and resc == 0 and err == 0
            else:
                assert ipc == 0.0
                assert np.array_equal(codes, code.astype(np.int64))
                assert resc == np.float32(info.fac_rescale) and err == np.float32(info.fac_error)
            off += D


@pytest.mark.parametrize("dist_type", [0, 1])
@pytest.mark.parametrize("plan", [None, [(64, 0), (64, 1), (128, 5)], [(256, 9)]])
def test_single_estimate_closed_forms(plan, dist_type):
    case = _case(d=256 if plan else 200, plan=plan)
    O = oracle()
    recs, stride, _ = F.orc_single_quantize(case, case.base[:60])
    for q in case.queries:
        vars_, fast, acc = F.orc_single_estimate(case, recs, stride, q, 4.0, dist_type)
        for i in range(60):
            dec = _decode(case, recs, stride, i)
            tot_acc = 0.0
            off = 0
            for s, (D, b) in enumerate(case.plan):
                x = np.zeros(D, np.float32)
                c = min(D, case.num_dim - off)
                x[:c] = q[off:off + c]
                qr = np.zeros(D, np.float32)
                O.orc_gemv_rowvec(x, case.rotations[s], D, qr)
                onorm, _, codes, resc, _ = dec[s]
                q2 = float(qr.astype(np.float64) @ qr)
                if b == 0:
                    ip = 0.0
                else:
                    delta = 2.0 / (1 << b)
                    oa = (codes + 0.5) * delta - 1.0
                    ip = float(resc) * float(oa @ qr.astype(np.float64))
                tot_acc += ip if dist_type == 1 else float(onorm) ** 2 + q2 - 2 * ip
                off += D
            assert abs(acc[i] - tot_acc) <= 1e-4 * (1 + abs(tot_acc))
        if dist_type == 0:
            true = ((case.base[:60].astype(np.float64) - q) ** 2).sum(1)
            assert (acc >= 0).all() and (fast >= 0).all() and (vars_ >= 0).all()
            if plan is None:  # the plan the DP picks, as in the reference test
                assert np.mean(np.abs(acc - true) / true) < 0.2  # ut_single_estimator.cpp:190


def test_single_estimate_fast_formula():
    """compFastDist of a 1-segment plan against the closed form on the quantised query."""
    case = _case(d=128, plan=[(128, 4)], rotate=False)
    recs, stride, _ = F.orc_single_quantize(case, case.base[:30])
    q = case.queries[0].astype(np.float32)
    _, fast, _ = F.orc_single_estimate(case, recs, stride, q, 4.0, 0)
    vl, vr = q.min(), q.max()
    delta = np.float32(np.float64(np.float32(vr - vl)) / 255.99)
    qsq = np.floor((q - vl) / delta).astype(np.int64)
    for i in range(30):
        onorm, _, codes, _, _ = _decode(case, recs, stride, i)[0]
        msb = (codes >> 3) & 1
        tmp = float(delta) * float(msb @ qsq) + (float(vl) + 0.5 * float(delta)) * float(msb.sum())
        ip1 = (tmp - (0.5 * q.astype(np.float64).sum() - 0.58 * np.linalg.norm(q.astype(np.float64)))) \
            * 5.0 / np.sqrt(128.0) * float(onorm)
        want = max(0.0, float(q.astype(np.float64) @ q) + float(onorm) ** 2 - ip1)
        assert abs(fast[i] - want) <= 1e-4 * (1 + want)


def test_single_quantize_empty_and_unrotated():
    case = _case(d=128, plan=[(128, 4)], rotate=False)
    recs, stride, nfail = F.orc_single_quantize(case, np.zeros((0, 128), np.float32))
    assert len(recs) == 0 and nfail == 0
    recs, stride, _ = F.orc_single_quantize(case, case.base[:5])
    v, f, a = F.orc_single_estimate(case, recs, stride, case.queries[0])
    assert v.shape == (5,) and np.isfinite(a).all()
