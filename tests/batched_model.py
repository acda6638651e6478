"""Numpy model of the batched int8-MFMA scan (test infrastructure).

The batched path computes, for every vector of every probed list, the accurate SAQ estimate
(caq_estimator.hpp:190-217) from the full codes, with the residual query r = q_rot - c_rot
quantised per (query, list, segment) to int8 (or two int8 terms, or bfloat16).  This model
restates that in numpy from independent inputs: the oracle's rotated query (orc_gemv_rowvec),
the oracle encoder's codes and factors (orc_caq_encode on the oracle's rotated residuals) and
the rotated centroids.  The integer dot products are exact; the fp32 epilogue takes the same
terms in the same order as the kernel, with its fused multiply-adds (fma32), but with its own
order for the sums of r and r^2, so model distances match to about 1e-6 relative, not bitwise.
BF16 dot products are summed by the matrix cores in float32: see model_search for their bound.
"""
import ctypes as C

import numpy as np

from oracle_lib import CaqCfg, CodeInfo, Sizes, oracle
import saq_fixtures as F


def vector_tables(case):
    """Per list: ids, and per segment the full codes (n x D_s), o_l2norm (n) and rescale (n)."""
    O = oracle()
    plan = case.orc_plan()
    cfg = CaqCfg(case.rd_lmt, case.eps, 1)
    out = []
    for c in range(case.K):
        members = np.nonzero(case.cluster_ids == c)[0]
        n = len(members)
        sz = Sizes()
        O.orc_cluster_sizes(C.byref(plan), n, C.byref(sz))
        scratch = {
            "short_factors": np.zeros(max(1, sz.short_factor_floats), np.float32),
            "short_code": np.zeros(max(1, sz.short_code_bytes), np.uint8),
            "long_code": np.zeros(max(1, sz.long_code_bytes), np.uint8),
            "long_factors": np.zeros(max(1, sz.long_factor_floats), np.float32),
            "ids": np.zeros(max(1, n), np.uint32), "centroid_rot": np.zeros(case.D_pad, np.float32),
            "num_vec": n,
        }
        segs = []
        off = 0
        for s, (ds, b) in enumerate(case.plan):
            X = np.zeros((max(1, n), ds), np.float32)
            cen = np.zeros(ds, np.float32)
            cp = max(0, min(ds, case.num_dim - off))
            if n:
                X[:n, :cp] = case.base[members, off:off + cp]
            cen[:cp] = case.centroids[c, off:off + cp]
            Orot = np.zeros((max(1, n), ds), np.float32)
            P = case.rotations[s]
            co = F._cluster_out(scratch)
            O.orc_quantize_segment(C.byref(plan), s, X, n, cen, P.ctypes.data if P is not None else None,
                                   C.byref(cfg), C.cast(C.pointer(co), C.c_void_p), Orot.ctypes.data)
            codes = np.zeros((n, ds), np.int64)
            norm = np.zeros(n, np.float32)
            resc = np.zeros(n, np.float32)
            for i in range(n):
                code = np.zeros(ds, np.uint16)
                info = CodeInfo()
                O.orc_caq_encode(np.ascontiguousarray(Orot[i]), ds, b, C.byref(cfg), cen.ctypes.data, code,
                                 C.byref(info))
                codes[i] = code
                norm[i] = info.o_l2norm
                resc[i] = info.fac_rescale
            segs.append((codes, norm, resc))
            off += ds
        out.append((members.astype(np.uint32), segs))
    return out


def rotated_queries(case):
    O = oracle()
    nq = case.queries.shape[0]
    qr = np.zeros((nq, case.D_pad), np.float32)
    for i in range(nq):
        off = 0
        for s, (ds, _) in enumerate(case.plan):
            qs = np.zeros(ds, np.float32)
            cp = max(0, min(ds, case.num_dim - off))
            qs[:cp] = case.queries[i, off:off + cp]
            P = case.rotations[s]
            if P is None:
                qr[i, off:off + ds] = qs
            else:
                o = np.zeros(ds, np.float32)
                O.orc_gemv_rowvec(qs, np.ascontiguousarray(P, np.float32), ds, o)
                qr[i, off:off + ds] = o
            off += ds
    return qr


def f32(x):
    return np.float32(x)


def fma32(a, b, c):
    """float32 fused multiply-add: the product of two float32 values is exact in float64, so one
    float64 add and one rounding to float32 (double rounding aside) give fmaf's result."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(np.float32)


def bf16_rne(x):
    """float32 values rounded to bfloat16 (round to nearest, ties to even), as float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def model_search(case, tables, qrot, probes, topk, qs=1, npq=None, with_err=False):
    """Per query: the topk smallest (distance, id) pairs of the batched estimator, and the
    distance of every probed vector (dict id -> distance).

    qs = 3 (BF16 queries): r / amax rounded to bfloat16 times the unshifted codes, summed in
    float32 by the matrix cores in an order of their own.  The model sums exactly (float64) and,
    with with_err, also returns per vector a bound on the distance's deviation from it: the
    float32 summation bound 2 n u sum |v code| (n dimensions, u = 2^-24, a factor 2 for an
    accumulation that truncates) carried through |a1| and |2 rescale|."""
    nq = qrot.shape[0]
    results, alld, allerr = [], [], []
    offs = np.cumsum([0] + [d for d, _ in case.plan])
    for qi in range(nq):
        lim = probes.shape[1] if npq is None else int(npq[qi])
        cand, cerr = {}, {}
        for c in probes[qi, :lim]:
            ids, segs = tables[c]
            n = len(ids)
            if n == 0:
                continue
            crot = case.clusters[c]["centroid_rot"]
            q2 = f32(0)
            consts = []
            for s, (ds, b) in enumerate(case.plan):
                r = (qrot[qi, offs[s]:offs[s] + ds] - crot[offs[s]:offs[s] + ds]).astype(np.float32)
                q2 = f32(q2 + f32(np.sum(r.astype(np.float64) ** 2)))
                if b == 0:
                    consts.append(None)
                    continue
                amax = f32(np.max(np.abs(r)))
                sqd = f32(2.0 / (1 << b))
                sr = f32(np.sum(r.astype(np.float64)))
                if qs == 3:
                    inv = f32(f32(1) / amax) if amax > 0 else f32(0)
                    v = bf16_rne((r * inv).astype(np.float32))
                    consts.append((v.astype(np.float64), None, f32(sqd * amax), f32(0), f32(f32(-1.0 + sqd / 2) * sr)))
                    continue
                s1 = f32(amax / f32(127))
                inv1 = f32(f32(127) / amax) if amax > 0 else f32(0)
                s2 = f32(s1 / f32(254))
                inv2 = f32(f32(254 * 127) / amax) if amax > 0 else f32(0)
                v1 = np.clip(np.rint((r * inv1).astype(np.float32)), -127, 127)
                v2 = np.zeros_like(v1)
                if qs == 2:
                    rem = (r - (v1.astype(np.float32) * s1).astype(np.float32)).astype(np.float32)
                    v2 = np.clip(np.rint((rem * inv2).astype(np.float32)), -127, 127)
                a1, a2 = f32(sqd * s1), f32(sqd * s2)
                cst = f32(f32(-1.0 + sqd / 2) * sr)
                if b >= 8:  # the code shift's correction, folded into the constant as pair prep does
                    cst = fma32(a1, f32(128 * int(v1.sum())), cst)
                    if qs == 2:
                        cst = fma32(a2, f32(128 * int(v2.sum())), cst)
                consts.append((v1.astype(np.int64), v2.astype(np.int64), a1, a2, cst))
            o2 = np.zeros(n, np.float32)
            for s in range(len(case.plan)):
                o2 = (o2 + segs[s][1] * segs[s][1]).astype(np.float32)
            dist = (o2 + q2).astype(np.float32)
            err = np.zeros(n, np.float64)
            for s, (ds, b) in enumerate(case.plan):
                if b == 0:
                    continue
                v1, v2, a1, a2, kq = consts[s]
                codes, _, resc = segs[s]
                if qs == 3:
                    # unshifted codes; 9+ bits as the low byte and the high part
                    lo = codes & 255 if b >= 9 else codes
                    hi = codes >> 8
                    ip = fma32(a1, (lo @ v1).astype(np.float32), kq)
                    mag = np.abs(lo) @ np.abs(v1)
                    if b >= 9:
                        ip = fma32(f32(256) * a1, (hi @ v1).astype(np.float32), ip)
                        mag = mag + 256 * (np.abs(hi) @ np.abs(v1))
                    m2r = (f32(-2) * resc).astype(np.float32)
                    dist = fma32(m2r, ip, dist)
                    err += np.abs(m2r.astype(np.float64)) * abs(float(a1)) * 2 * ds * 2.0 ** -24 * mag
                    continue
                # the kernel's operands: codes of 8 bits shifted by -128; 9+ bits as a shifted low
                # byte plane and a high plane
                lo = (codes & 255) - 128 if b >= 9 else (codes - 128 if b == 8 else codes)
                hi = codes >> 8
                ip = fma32(a1, (lo @ v1).astype(np.float32), kq)
                if b >= 9:
                    ip = fma32(f32(256) * a1, (hi @ v1).astype(np.float32), ip)
                if qs == 2:
                    ip = fma32(a2, (lo @ v2).astype(np.float32), ip)
                    if b >= 9:
                        ip = fma32(f32(256) * a2, (hi @ v2).astype(np.float32), ip)
                dist = fma32((f32(-2) * resc).astype(np.float32), ip, dist)
            for u, d, e in zip(ids.tolist(), dist.tolist(), err.tolist()):
                cand[u] = d
                cerr[u] = e
        order = sorted(cand.items(), key=lambda t: (t[1], t[0]))[:topk]
        results.append(order)
        alld.append(cand)
        allerr.append(cerr)
    if with_err:
        return results, alld, allerr
    return results, alld
