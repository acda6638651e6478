"""The encoder options around the CAQ encode and the build-time column statistics (CPU only).

* caq_ori_qB (CAQEncoder, caq_encoder.hpp:144-169, 217-219): the codes are found with the code
  maximum of ori_qB bits and then shifted down by ori_qB - b (downUpSample), delta scaled by
  2^(ori_qB - b) and <o, oa>, |oa|^2 recomputed.  Pinned by the identity with the oracle's own
  ori_qB-bit encode (whose path the other encoder tests pin): codes >> (ori_qB - b), the same
  v_mi, delta times 2^shift, and the recomputed sums against a float64 recomputation (to 1e-12:
  the test rebuilds delta from its rescaled value).
* saqgpu_column_mean_var (the library's host code): the means of data.colwise().mean() equal
  the reference's compiled Eigen bit for bit (oracle/_ref, built with its release flags); the
  variances, whose sums that -Ofast build reassociates, agree to 1e-5 relative; closed forms
  (constant columns, one row) hold exactly.
"""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import CaqCfg, CodeInfo, oracle, ref


def encode(o, bits, ori_qb=0, rd_lmt=6, eps=1e-8):
    O = oracle()
    code = np.zeros(len(o), np.uint16)
    info = CodeInfo()
    rc = O.orc_caq_encode(np.ascontiguousarray(o, np.float32), len(o), bits, C.byref(CaqCfg(rd_lmt, eps, 0, ori_qb)),
                          None, code, C.byref(info))
    return rc, code, info


@pytest.mark.parametrize("D", [64, 128])
@pytest.mark.parametrize("bits,qb", [(1, 2), (1, 8), (3, 5), (4, 8), (7, 9)])
@pytest.mark.parametrize("rd_lmt", [0, 6])
def test_ori_qb_is_the_high_bit_encode_shifted_down(D, bits, qb, rd_lmt):
    rng = np.random.default_rng(D * 100 + bits * 10 + qb)
    for t in range(20):
        o = (rng.standard_normal(D) * rng.uniform(0.1, 3)).astype(np.float32)
        rc, code, info = encode(o, bits, qb, rd_lmt)
        rc_h, code_h, info_h = encode(o, qb, 0, rd_lmt)
        assert rc == rc_h == 0
        shift = qb - bits
        assert np.array_equal(code, code_h >> shift)
        assert code.max() <= (1 << bits) - 1
        assert info.v_mi == info_h.v_mi
        assert info.delta == info_h.delta * (1 << shift)
        # <o, oa> and |oa|^2 of the shifted codes, recomputed in order (after rescale_vmx_to1)
        v_mx = float(np.max(np.abs(o)))
        delta = info.delta * v_mx
        q = (code.astype(np.float64) + 0.5) * delta - v_mx
        ip = 0.0
        oa = 0.0
        for j in range(D):
            ip += q[j] * float(o[j])
            oa += q[j] * q[j]
        sr = 1.0 / v_mx  # (delta here is rebuilt from the rescaled one: equal to ~1e-15)
        assert info.ip_o_oa == pytest.approx(ip * sr, rel=1e-12, abs=1e-12)
        assert info.oa_l2sqr == pytest.approx(oa * (sr * sr), rel=1e-12)


def test_ori_qb_off_is_the_plain_encode():
    o = np.random.default_rng(3).standard_normal(128).astype(np.float32)
    _, a, ia = encode(o, 4, 0)
    _, b, ib = encode(o, 4, 0)
    assert np.array_equal(a, b) and ia.delta == ib.delta


def _col_stats(X):
    import saq_amd
    return saq_amd.column_mean_var(X)


@pytest.mark.parametrize("n,d", [(1, 5), (2, 16), (5, 33), (37, 20), (101, 35), (1000, 70), (4099, 130)])
def test_column_means_equal_compiled_eigen(n, d):
    R = ref()
    if R is None:
        pytest.skip("oracle/_ref is not built (the reference tree is absent)")
    X = (np.random.default_rng(n + d).standard_normal((n, d)) * 3 + 1).astype(np.float32)
    mean, var = _col_stats(X)
    rm, rv = np.zeros(d, np.float32), np.zeros(d, np.float32)
    R.ref_eigen_colwise_mean_var(X, n, d, rm, rv)
    assert np.array_equal(mean.view(np.uint32), rm.view(np.uint32))
    # the compiled variance sums are reassociated by -Ofast (up to ~15 ulps apart at n = 4099)
    assert np.all(np.abs(var - rv) <= 1e-5 * np.abs(rv))


def test_column_stats_closed_forms():
    X = np.tile(np.array([[1.5, -2.0, 0.0, 3.25] * 5], np.float32), (77, 1))
    mean, var = _col_stats(X)
    assert np.array_equal(mean, X[0]) and np.all(var == 0)
    one = np.arange(40, dtype=np.float32)[None, :]
    mean, var = _col_stats(one)
    assert np.array_equal(mean, one[0]) and np.all(var == 0)
    # a planted mistake (the plain division instead of the reciprocal product) is visible
    X = (np.random.default_rng(5).standard_normal((37, 20)) * 3 + 1).astype(np.float32)
    mean, _ = _col_stats(X)
    div = np.array([np.float32(np.float32(sum_ordered(X[:, j], j < 16)) / np.float32(37)) for j in range(20)])
    rcp = np.array([np.float32(np.float32(sum_ordered(X[:, j], j < 16)) * (np.float32(1) / np.float32(37)))
                    for j in range(20)])
    assert np.array_equal(mean.view(np.uint32), rcp.view(np.uint32))
    assert not np.array_equal(mean.view(np.uint32), div.view(np.uint32))


def sum_ordered(col, packet):
    """Eigen's column sum order (PartialReduxEvaluator.h:100-121 for packet columns, else row
    after row), in float32."""
    F = np.float32
    n = len(col)
    p = F(col[0])
    i = 1
    if packet:
        size4 = (n - 1) & ~3
        while i < size4:
            p = F(p + F(F(col[i] + col[i + 1]) + F(col[i + 2] + col[i + 3])))
            i += 4
    while i < n:
        p = F(p + col[i])
        i += 1
    return p
