"""The oracle's cluster estimator (CaqCluEstimator + Lut, quantization/caq/caq_estimator.hpp:27-218,
fastscan/lut.hpp:20-127) pinned by closed forms on the IVF layout.

For every lane of every block of a list, per segment, the oracle's variance, fast and accurate
estimates (orc_estimate_lanes) are compared with formulas evaluated in float64 from independent
inputs: the codes re-derived by encoding the rotated residuals (and checked to be the bytes the
list stores, through the packers pinned against the reference's compiled code), the stored
factors, and the rotated query.  With r = q_rot - c_rot, oa = (c + 1/2) * 2/2^b - 1 the
dequantised code, msb the top code bit and T = (D/4) * delta_lut / 2 the fast-scan LUT's total
rounding (each of the D/4 u16 tables is off by at most half a step):

  vars = max(0, |o|^2 + |r|^2 - 2 m sqrt(sum var_i x_i^2))                 (caq_estimator.hpp:116-136)
  fast = max(0, |o|^2 + |r|^2 - (<msb, r> - (0.5 sum r - 0.58 |r|)) 5/sqrt(D) |o|), within 5/sqrt(D) |o| T
                                                                           (lut.hpp:98-118, :148-179)
  acc  = |o|^2 + |r|^2 - 2 rescale <oa, r>, within 2 |rescale| T          (lut.hpp:121-126, :190-217)

and for inner product (r = q_rot): vars = <q,c> - bound, fast = fst/2 + <q,c>, acc = rescale <oa,q>
+ <q,c>, summed over segments.  A wrong constant (0.58, 5 = 4/0.8, 2/2^b), sign, dropped term
(sum r, |r|^2) or factor (rescale, |o|) moves an estimate by far more than these bounds.
"""
import ctypes as C

import numpy as np
import pytest

import saq_fixtures as F
from oracle_lib import CaqCfg, CodeInfo, LutScalars, Metrics, oracle


def rd16(x):
    return (x + 15) // 16 * 16


def segment_codes(case, c, s):
    """Rotated residuals and full codes of list c's members in segment s, from the oracle encoder,
    after checking that they are what the list stores (short code via pack_codes, long code via
    the CodeHelper compaction)."""
    O = oracle()
    clu = case.clusters[c]
    n = clu["num_vec"]
    D, b = case.plan[s]
    off = sum(d for d, _ in case.plan[:s])
    X = np.zeros((max(1, n), D), np.float32)
    cp = max(0, min(D, case.num_dim - off))
    members = clu["ids"][:n]
    X[:n, :cp] = case.base[members, off:off + cp]
    crot = np.ascontiguousarray(clu["centroid_rot"][off:off + D])
    P = case.rotations[s]
    Ores = np.zeros((max(1, n), D), np.float32)
    O.orc_rotate_rows(X, n, D, P.ctypes.data if P is not None else None, crot, Ores)
    codes = np.zeros((n, D), np.uint16)
    cfg = CaqCfg(case.rd_lmt, case.eps, 1 if P is not None else 0)
    info = CodeInfo()
    for i in range(n):
        O.orc_caq_encode(np.ascontiguousarray(Ores[i]), D, b, C.byref(cfg), None, codes[i], C.byref(info))
    if b > 0 and n:
        S = len(case.plan)
        single = S == 1
        code_off = 256 if single else 4 * sum(d for d, bb in case.plan[:s] if bb > 0)
        per_block = (256 if single else 0) + 4 * sum(d for d, bb in case.plan if bb > 0)
        msb = ((codes >> (b - 1)) & 1).astype(np.uint8)
        sb = np.packbits(msb, axis=1, bitorder="big")  # MSB-first bytes (cluster_packer.hpp:124-139)
        nb = (n + 31) // 32
        blocks = np.zeros(nb * D * 4, np.uint8)
        O.orc_pack_codes(D, np.ascontiguousarray(sb), n, blocks)
        stored = clu["short_code"].reshape(nb, per_block)[:, code_off:code_off + D * 4].reshape(-1)
        assert np.array_equal(stored, blocks), f"list {c} segment {s}: short codes differ"
        lb = D * (b - 1) // 8
        if lb:
            compact = single or case.compact
            low = (codes & ((1 << (b - 1)) - 1)).astype(np.uint16)
            for i in range(n):
                want = np.zeros(lb, np.uint8)
                O.orc_compact_code(b - 1, np.ascontiguousarray(low[i]), D, want)
                if compact:
                    start = sum(rd16(d * (bb - 1) // 8 * n) for d, bb in case.plan[:s] if bb > 0) + i * lb
                else:
                    stride = sum(rd16(d * (bb - 1) // 8) for d, bb in case.plan if bb > 0)
                    start = i * stride + sum(rd16(d * (bb - 1) // 8) for d, bb in case.plan[:s] if bb > 0)
                assert np.array_equal(clu["long_code"][start:start + lb], want), f"list {c} seg {s} vec {i}"
    return Ores[:n], codes


def o_l2norm(case, c, s, n):
    clu = case.clusters[c]
    S = len(case.plan)
    nb = (n + 31) // 32
    if S == 1:
        per_block = 256 + 4 * case.plan[0][0]
        rows = clu["short_code"].reshape(nb, per_block)[:, :128].copy().view(np.float32)
    else:
        rows = clu["short_factors"].reshape(nb, S, 2, 32)[:, s, 0, :]
    return rows.reshape(-1)[:n]


def query_segments(case, q, m=4.0):
    """Per segment: x (the padded query segment), the rotated query and the variance bound."""
    O = oracle()
    out = []
    off = 0
    for s, (D, b) in enumerate(case.plan):
        x = np.zeros(D, np.float32)
        cp = max(0, min(D, case.num_dim - off))
        x[:cp] = q[off:off + cp]
        P = case.rotations[s]
        qr = np.zeros(D, np.float32)
        if P is not None:
            O.orc_gemv_rowvec(x, P, D, qr)
        else:
            qr[:] = x
        var = np.zeros(D, np.float64)
        var[:cp] = case.variance[off:off + cp]
        bound = m * np.sqrt(float((var * x.astype(np.float64) ** 2).sum()))
        out.append((x, qr, bound, off))
        off += D
    return out


CASES = [dict(d=128, plan=[(128, 4)]), dict(d=128, plan=[(128, 1)]), dict(d=128, plan=[(128, 9)], rotate=False),
         dict(d=256, plan=[(64, 9), (64, 5), (64, 0), (64, 2)]), dict(d=200, avg_bits=4.0, tau=32),
         dict(d=256, plan=[(128, 3), (128, 12)], compact=True)]


@pytest.mark.parametrize("kw", CASES)
def test_lane_estimates_closed_forms(kw):
    kw = dict(kw)
    d = kw.pop("d")
    case = F.make_case(n=700, d=d, K=3, nq=3, seed=41, **kw)
    O = oracle()
    S = len(case.plan)
    import test_oracle_search as TS
    for q in case.queries:
        segq = query_segments(case, q)
        for c in range(case.K):
            n = case.clusters[c]["num_vec"]
            if n == 0:
                continue
            vars_, fast, acc = TS.lanes(case, [c], q)
            for s, (D, b) in enumerate(case.plan):
                x, qr, bound, off = segq[s]
                crot = case.clusters[c]["centroid_rot"][off:off + D]
                r = (qr - crot).astype(np.float32)  # the residual in fp32, as the reference forms it
                r64 = r.astype(np.float64)
                r2 = float(r64 @ r64)
                on = o_l2norm(case, c, s, n).astype(np.float64)
                v = vars_[:, s, :].reshape(-1)[:n].astype(np.float64)
                want_v = np.maximum(0.0, on ** 2 + r2 - 2 * bound)
                assert np.all(np.abs(v - want_v) <= 1e-5 * (on ** 2 + r2 + 2 * bound) + 1e-6)
                f = fast[:, s, :].reshape(-1)[:n].astype(np.float64)
                a = acc[:, s, :].reshape(-1)[:n].astype(np.float64)
                if b == 0:
                    assert np.array_equal(f, v)
                    assert np.all(np.abs(a - (on ** 2 + r2)) <= 1e-5 * (on ** 2 + r2) + 1e-6)
                    continue
                _, codes = segment_codes(case, c, s)
                sc = LutScalars()
                lut = np.zeros(D // 4 * 16, np.uint16)
                O.orc_lut_prepare(D, r, lut, C.byref(sc))
                T = (D / 4) * float(sc.delta) / 2 * 1.01
                msb = ((codes >> (b - 1)) & 1).astype(np.float64)
                ip1 = msb @ r64
                g = 5.0 / np.sqrt(D)
                fst = (ip1 - (0.5 * r64.sum() - 0.58 * np.sqrt(r2))) * g * on
                want_f = np.maximum(0.0, on ** 2 + r2 - fst)
                scale = on ** 2 + r2 + np.abs(fst)
                assert np.all(np.abs(f - want_f) <= g * on * T + 2e-5 * scale + 1e-6), f"fast, list {c} seg {s}"
                resc = case.clusters[c]["long_factors"].reshape(n, S, 2)[:, s, 0].astype(np.float64)
                oa = (codes.astype(np.float64) + 0.5) * (2.0 / (1 << b)) - 1.0
                ipo = resc * (oa @ r64)
                want_a = on ** 2 + r2 - 2 * ipo
                scale = on ** 2 + r2 + 2 * np.abs(ipo)
                assert np.all(np.abs(a - want_a) <= 2 * np.abs(resc) * T + 2e-5 * scale + 1e-6), \
                    f"acc, list {c} seg {s}"


@pytest.mark.parametrize("kw", CASES[3:5])
def test_inner_product_estimates_closed_forms(kw):
    """IVF::estimate with DistType::IP: the residual is the rotated query itself, and each segment
    adds <q, c> (caq_estimator.hpp:98-101, :121-125, :161-165, :211-212)."""
    kw = dict(kw)
    d = kw.pop("d")
    case = F.make_case(n=500, d=d, K=2, nq=2, seed=43, **kw)
    O = oracle()
    S = len(case.plan)
    for q in case.queries:
        segq = query_segments(case, q)
        for c in range(case.K):
            n = case.clusters[c]["num_vec"]
            ids = np.zeros(n, np.uint32)
            acc = np.zeros(n, np.float32)
            fast = np.zeros(n, np.float32)
            var = np.zeros(n, np.float32)
            vpad = np.zeros(case.D_pad, np.float32)
            vpad[:case.num_dim] = case.variance
            met = Metrics()
            O.orc_estimate(C.byref(case.orc_plan()), vpad, F.orc_rot_ptrs(case), F.orc_clusters(case),
                           np.array([c], np.uint32), 1, np.ascontiguousarray(q), 4.0, 1, ids, acc,
                           fast.ctypes.data, var.ctypes.data, C.cast(C.pointer(met), C.c_void_p))
            want_a = np.zeros(n)
            want_v = np.zeros(n)
            want_f = np.zeros(n)
            tol_a = np.zeros(n)
            tol_f = np.zeros(n)
            for s, (D, b) in enumerate(case.plan):
                x, qr, bound, off = segq[s]
                q64 = qr.astype(np.float64)
                crot = case.clusters[c]["centroid_rot"][off:off + D].astype(np.float64)
                qc = float(q64 @ crot)
                want_v += qc - bound
                on = o_l2norm(case, c, s, n).astype(np.float64)
                if b == 0:
                    want_a += qc
                    want_f += qc - bound
                    continue
                _, codes = segment_codes(case, c, s)
                sc = LutScalars()
                lut = np.zeros(D // 4 * 16, np.uint16)
                O.orc_lut_prepare(D, qr, lut, C.byref(sc))
                T = (D / 4) * float(sc.delta) / 2 * 1.01
                msb = ((codes >> (b - 1)) & 1).astype(np.float64)
                g = 5.0 / np.sqrt(D)
                fst = (msb @ q64 - (0.5 * q64.sum() - 0.58 * np.sqrt(q64 @ q64))) * g * on
                want_f += fst * 0.5 + qc
                tol_f += 0.5 * g * on * T + 2e-5 * (np.abs(fst) + abs(qc))
                resc = case.clusters[c]["long_factors"].reshape(n, S, 2)[:, s, 0].astype(np.float64)
                oa = (codes.astype(np.float64) + 0.5) * (2.0 / (1 << b)) - 1.0
                ipo = resc * (oa @ q64)
                want_a += ipo + qc
                tol_a += np.abs(resc) * T + 2e-5 * (np.abs(ipo) + abs(qc))
            assert np.all(np.abs(var - want_v) <= 1e-5 * (np.abs(want_v) + 1) * S)
            assert np.all(np.abs(fast - want_f) <= tol_f + 1e-5)
            assert np.all(np.abs(acc - want_a) <= tol_a + 1e-5)


def decode(case, c, s):
    """orc_decode_cluster: full codes, |o| and rescale of list c's segment s."""
    O = oracle()
    n = case.clusters[c]["num_vec"]
    D = case.plan[s][0]
    codes = np.zeros(max(1, n) * D, np.int32)
    on = np.zeros(max(1, n), np.float32)
    rs = np.zeros(max(1, n), np.float32)
    clus = F.orc_clusters(case)
    O.orc_decode_cluster(C.byref(case.orc_plan()), C.cast(C.pointer(clus[c]), C.c_void_p), s, codes, on, rs)
    return codes[:n * D].reshape(n, D), on[:n], rs[:n]


@pytest.mark.parametrize("kw", CASES)
def test_decode_cluster_round_trip(kw):
    """The bulk decoder the full-size GPU tests use reads back exactly the codes the encoder
    produced (whose packing is checked against the list's bytes in segment_codes)."""
    kw = dict(kw)
    d = kw.pop("d")
    case = F.make_case(n=300, d=d, K=3, nq=1, seed=47, **kw)
    S = len(case.plan)
    for c in range(case.K):
        n = case.clusters[c]["num_vec"]
        for s in range(S):
            codes, on, rs = decode(case, c, s)
            _, want = segment_codes(case, c, s)
            assert np.array_equal(codes, want.astype(np.int32))
            assert np.array_equal(on, o_l2norm(case, c, s, n))
            assert np.array_equal(rs, case.clusters[c]["long_factors"].reshape(n, S, 2)[:, s, 0])
