"""The oracle's CAQ code adjustment (CAQEncoder::code_adjustment, quantization/caq/caq_encoder.hpp:57-142)
against an exact-rational model of the reference's source.

The model runs the same greedy coordinate ascent with every quantity an exact rational
(Python Fraction; all inputs are dyadic: float32 data, float64 delta and v_mi): the starting
codes and running <o, oa> / |oa|^2 as encode() forms them (:187-211, its float32 products
code_j * o_j taken exactly), per dimension first the ++ moves and then the -- moves while
(ip^2 + eps |oa|_0^2) |oa'|^2 < ip'^2 |oa|^2, adj_cnt counted only when a code changed, the round
limit checked before the exact recompute of <o, oa> and |oa|^2 (so the last limited round keeps
its running values), and no recompute after it.  The oracle evaluates the same steps in float64.
The two can differ only where a comparison sits within float64 rounding of a tie: every vector
whose final codes differ must have met a decision whose exact relative margin is below 1e-11,
and such vectors are counted and listed.  A dropped recompute, swapped ++/-- order, an off-by-one
round limit or eps applied to the running |oa|^2 changes codes of vectors far from any tie.
"""
import ctypes as C
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle_lib import CaqCfg, CodeInfo, oracle

HALF = Fr(1, 2)


def start_state(o, bits):
    """encode() (caq_encoder.hpp:172-211) up to the adjustment: v_mx, delta, the floor codes of the
    release build (the hoisted float reciprocal, pinned against the compiled reference
    elsewhere), and exact <o, oa>, |oa|^2 from its closed forms."""
    O = oracle()
    cm = (1 << bits) - 1
    v_mx = float(max(np.float32(o.max()), np.float32(-o.min())))
    v_mi = -v_mx
    delta = (v_mx - v_mi) / (cm + 1)
    finv = np.float32(1.0) / np.float32(delta)
    cf = np.minimum(np.floor((o - np.float32(v_mi)) * finv), np.float32(cm)).astype(np.float32)
    codes = cf.astype(np.int64)
    vec_sum = float(O.orc_eigen_sum(o, len(o)))
    prods = (cf * o).astype(np.float32)  # the float32 products code_f[j] * o[j]
    ip_o_code = sum((Fr(float(p)) for p in prods), Fr(0))
    d, vmi = Fr(delta), Fr(v_mi)
    ip = ip_o_code * d + (vmi + HALF * d) * Fr(vec_sum)
    c2 = int((codes * codes).sum())
    cs = int(codes.sum())
    ln = d * d * c2 + (d * d + 2 * d * vmi) * cs + (Fr(1, 4) * d * d + d * vmi + vmi * vmi) * len(o)
    return codes, d, vmi, ip, ln


def exact_adjust(o, bits, rd_lmt, eps):
    codes, d, vmi, ip, ln = start_state(o, bits)
    codes = [int(c) for c in codes]
    cm = (1 << bits) - 1
    O = [Fr(float(x)) for x in o]
    re_eps = Fr(float(np.float32(eps))) * ln
    min_margin = float("inf")
    if not rd_lmt or ln == 0:
        return codes, ip, ln, min_margin, 0
    adj, rnd = 1, 1
    while adj:
        assert ip >= 0
        adj = 0
        for j in range(len(O)):
            cur = 0
            oa = (codes[j] + HALF) * d + vmi
            c = codes[j]
            tmp = ln - oa * oa
            ipd = d * O[j]
            for step in (1, -1):
                while (c < cm) if step > 0 else (c > 0):
                    nq = oa + step * d
                    nl = tmp + nq * nq
                    ni = ip + step * ipd
                    lhs = (ip * ip + re_eps) * nl
                    rhs = ni * ni * ln
                    if rhs != 0:
                        min_margin = min(min_margin, float(abs(lhs - rhs) / abs(rhs)))
                    if lhs >= rhs:
                        break
                    c += step
                    ip, oa, ln = ni, nq, nl
                    cur += 1
            if codes[j] != c:
                codes[j] = c
                adj += cur
        if rd_lmt and rnd >= rd_lmt:
            break
        qs = [(c + HALF) * d + vmi for c in codes]
        ip = sum((q * x for q, x in zip(qs, O)), Fr(0))
        ln = sum((q * q for q in qs), Fr(0))
        rnd += 1
    return codes, ip, ln, min_margin, rnd  # the round counter at exit (one past the last round unless limited)


def vectors(D, n, seed):
    rng = np.random.default_rng(seed)
    out = [rng.standard_normal(D).astype(np.float32) * np.float32(rng.uniform(0.1, 3.0)) for _ in range(n)]
    # repeated values and a spike: exact ties between moves of equal dims
    v = np.repeat(rng.standard_normal(D // 8), 8).astype(np.float32)
    v[3] = 5.0
    out.append(v)
    out.append((rng.integers(-3, 4, D) * 0.25).astype(np.float32))
    return out


@pytest.mark.parametrize("bits", [1, 2, 4, 8])
@pytest.mark.parametrize("rd_lmt,eps", [(1, 0.0), (2, 1e-8), (6, 1e-8), (6, 0.0), (6, 2e-3), (3, 0.02)])
def test_adjustment_matches_exact_model(bits, rd_lmt, eps):
    O = oracle()
    D = 64
    cfg = CaqCfg(rd_lmt, eps, 0)
    ties = []
    n_vec = 0
    for i, o in enumerate(vectors(D, 24, 100 * bits + rd_lmt)):
        want, ip, ln, margin, rounds = exact_adjust(o, bits, rd_lmt, eps)
        code = np.zeros(D, np.uint16)
        info = CodeInfo()
        O.orc_caq_encode(np.ascontiguousarray(o), D, bits, C.byref(cfg), None, code, C.byref(info))
        n_vec += 1
        if list(code.astype(int)) != want:
            assert margin < 1e-11, f"vector {i}: codes differ away from a tie (min margin {margin:.3g})"
            ties.append(i)
            continue
        assert info.rounds == rounds, f"vector {i}: {info.rounds} rounds, model {rounds}"
        # the oracle's running values (recomputed in float64 between rounds) track the exact ones;
        # encode_and_fac reports them scaled by 1 / v_mx (rescale_vmx_to1, caq_encoder.hpp:29-40)
        v_mx = float(max(np.float32(o.max()), np.float32(-o.min())))
        assert abs(info.ip_o_oa - float(ip) / v_mx) <= 1e-9 * abs(float(ip) / v_mx) + 1e-12
        assert abs(info.oa_l2sqr - float(ln) / v_mx ** 2) <= 1e-9 * abs(float(ln)) / v_mx ** 2 + 1e-12
    assert len(ties) <= n_vec // 10, f"too many tie mismatches: {ties}"
    if ties:
        print(f"bits {bits} rd_lmt {rd_lmt} eps {eps}: tie mismatches at vectors {ties}")


def test_adjustment_model_has_power():
    """Two plausible mistakes in a copy of the model (no exact recompute between rounds; -- tried
    before ++) each change codes of vectors far from any tie."""
    O = oracle()
    D, bits = 64, 4
    cfg = CaqCfg(6, 1e-8, 0)
    differs = 0
    for o in vectors(D, 24, 7):
        code = np.zeros(D, np.uint16)
        info = CodeInfo()
        O.orc_caq_encode(np.ascontiguousarray(o), D, bits, C.byref(cfg), None, code, C.byref(info))
        want = exact_adjust(o, bits, 6, 1e-8)[0]
        assert list(code.astype(int)) == want
        # one round only differs from the oracle's six wherever a second round moves a code
        one = exact_adjust(o, bits, 1, 1e-8)[0]
        differs += one != want
    assert differs > 0
