"""Pins the oracle before it is trusted as the checker.

Three kinds of pin, chosen so that a dropped term, a wrong index or a
transposed operand in the oracle fails at least one of them:

* against the reference's own kernels compiled from its sources with its
  release flags (oracle/_ref): fast-scan packing and accumulation, LUT
  construction, every extra-bit layout and inner product, Eigen's reductions
  and GEMV.  Bit-exact.
* against closed forms computed independently in numpy (exact integer LUT
  sums, code round trips, float32 sums that must equal float64 within a bound).
* against properties the mathematics fixes (ResultPool against a sorted-list
  model; the plan DP against brute-force enumeration; the code adjustment's
  local optimality; the encoder's factor identities).
"""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import (CaqCfg, CodeInfo, LutScalars, Sizes, make_plan, oracle, ref)

needs_ref = pytest.mark.skipif(ref() is None, reason="reference kernels (oracle/_ref) not built here")
# every reduction on the path runs over a multiple of 64 elements (segments are 64-aligned);
# for ragged tails -Ofast lets GCC reassociate the scalar tail, so those sizes are not pinned.
SIZES = [16, 32, 48, 64, 96, 128, 192, 960, 1024, 1536]


def rng(seed):
    return np.random.default_rng(seed)


# ------------------------------------------------------------------ reductions
@needs_ref
@pytest.mark.parametrize("n", SIZES)
def test_eigen_reductions_match_reference(n):
    O, R = oracle(), ref()
    g = rng(n)
    for _ in range(20):
        a = (g.standard_normal(n) * g.uniform(0.01, 100)).astype(np.float32)
        b = g.standard_normal(n).astype(np.float32)
        var = np.abs(g.standard_normal(n + 64)).astype(np.float32)
        assert np.float32(O.orc_eigen_sum(a, n)) == np.float32(R.ref_eigen_sum(a, n))
        assert np.float32(O.orc_eigen_sqnorm(a, n)) == np.float32(R.ref_eigen_sqnorm(a, n))
        assert np.float32(O.orc_eigen_dot(a, b, n)) == np.float32(R.ref_eigen_dot_rowcol(a, b, n))
        rows = np.stack([b, a]).astype(np.float32)
        assert np.float32(O.orc_eigen_diff_sqnorm(a, b, n)) == np.float32(
            R.ref_eigen_diff_sqnorm(rows, 2, 1, b, n))
        off = 64 if n % 64 == 0 else 0
        assert np.float32(O.orc_eigen_var_q2(np.ascontiguousarray(var[off:off + n]), a, n)) == np.float32(
            R.ref_eigen_var_q2(var, n + 64, off, a, n))


def test_reduction_tree_is_a_sum():
    # closed form: the 16-lane tree is a float32 summation, within n*eps of the float64 sum
    O = oracle()
    g = rng(1)
    for n in SIZES + [5, 17]:
        a = g.standard_normal(n).astype(np.float32)
        got = O.orc_eigen_sum(a, n)
        exact = float(np.sum(a.astype(np.float64)))
        assert abs(got - exact) <= n * 1.2e-7 * float(np.sum(np.abs(a))) + 1e-30
    # a wrong lane pairing breaks this exact case: only lane 3 is large
    v = np.zeros(16, np.float32)
    v[3] = 1e8
    v[11] = 1.0
    v[7] = -1e8
    # tree: b3 = v3+v11 = 1e8 (1 lost), b7 = v7+v15 = -1e8 -> c3 = b3+b7 = 0
    assert O.orc_tree16(v) == 0.0


@needs_ref
@pytest.mark.parametrize("D", [64, 128, 192, 960, 1024])
def test_gemv_matches_reference(D):
    O, R = oracle(), ref()
    g = rng(D)
    P = np.linalg.qr(g.standard_normal((D, D)))[0].astype(np.float32)
    for _ in range(3):
        q = (g.standard_normal(D) * 3).astype(np.float32)
        a = np.zeros(D, np.float32)
        b = np.zeros(D, np.float32)
        O.orc_gemv_rowvec(q, P, D, a)
        R.ref_eigen_gemv(q, P, D, b)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert np.allclose(a, q.astype(np.float64) @ P.astype(np.float64), atol=1e-4)


# ------------------------------------------------------------------ fast scan
@needs_ref
@pytest.mark.parametrize("D", [64, 128, 960])
def test_pack_lut_matches_reference(D):
    O, R = oracle(), ref()
    q = rng(D).standard_normal(D).astype(np.float32)
    a = np.zeros(D * 4, np.float32)
    b = np.zeros(D * 4, np.float32)
    O.orc_pack_lut(D, q, a)
    R.ref_pack_lut(D, q, b)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # closed form: entry j of codebook m is the sum of q[4m+t] over the bits of j (bit 3 = t=0)
    lut = a.reshape(-1, 16).astype(np.float64)
    for m in range(D // 4):
        for j in range(16):
            want = sum(float(q[4 * m + t]) for t in range(4) if j & (8 >> t))
            assert abs(lut[m, j] - want) < 1e-5


def raw_short_codes(g, n, D):
    return g.integers(0, 256, size=(n, D // 8), dtype=np.uint8)


def nibble(short, m):
    """Nibble of codebook m (dims 4m..4m+3, first dim in bit 3) from MSB-first bytes."""
    byte = short[:, m // 2]
    return (byte >> 4) if m % 2 == 0 else (byte & 15)


@needs_ref
@pytest.mark.parametrize("n", [32, 20, 64, 33])
@pytest.mark.parametrize("D", [64, 128, 960])
def test_pack_codes_and_accumulate_match_reference(n, D):
    O, R = oracle(), ref()
    g = rng(n * D)
    sc = raw_short_codes(g, n, D)
    nb = (n + 31) // 32
    a = np.zeros(nb * 4 * D, np.uint8)
    b = np.zeros(nb * 4 * D, np.uint8)
    O.orc_pack_codes(D, sc, n, a)
    R.ref_pack_codes(D, np.concatenate([sc, np.zeros((nb * 32 - n, D // 8), np.uint8)]), nb * 32, b)
    assert np.array_equal(a, b)
    lut16 = g.integers(0, 65536, size=D * 4, dtype=np.uint16)
    for blk in range(nb):
        s1 = np.zeros(32, np.uint32)
        s2 = np.zeros(32, np.int32)
        codes = np.ascontiguousarray(a[blk * 4 * D:(blk + 1) * 4 * D])
        O.orc_accumulate(codes, lut16, D, s1)
        R.ref_accumulate(codes, lut16, D, s2)
        assert np.array_equal(s1.astype(np.int64), s2.astype(np.int64))
        # closed form: exact integer sum over codebooks
        rows = np.zeros((32, D // 8), np.uint8)
        take = sc[blk * 32:min(n, blk * 32 + 32)]
        rows[:len(take)] = take
        lt = lut16.reshape(-1, 16).astype(np.int64)
        want = np.zeros(32, np.int64)
        for m in range(D // 4):
            want += lt[m, nibble(rows, m)]
        assert np.array_equal(s1.astype(np.int64), want)


@pytest.mark.parametrize("D", [64, 960])
def test_lut_prepare_closed_form(D):
    O = oracle()
    g = rng(7)
    r = (g.standard_normal(D) * 2).astype(np.float32)
    lut16 = np.zeros(D * 4, np.uint16)
    sc = LutScalars()
    O.orc_lut_prepare(D, r, lut16, C.byref(sc))
    lf = np.zeros(D * 4, np.float32)
    O.orc_pack_lut(D, r, lf)
    vl, vr = float(lf.min()), float(lf.max())
    assert abs(sc.delta - (vr - vl) / 65535.99) <= 1e-6 * abs(sc.delta)
    rec = lut16.astype(np.float64) * sc.delta + vl
    assert np.all(rec <= lf + 1e-4 * (vr - vl)) and np.all(lf - rec <= sc.delta * 1.001 + 1e-5)
    assert lut16.max() <= 65535 and lut16.max() >= 65534
    assert abs(sc.sum_q - r.astype(np.float64).sum()) < 1e-3
    assert abs(sc.q_l2sqr - (r.astype(np.float64) ** 2).sum()) < 1e-3 * sc.q_l2sqr
    assert abs(sc.sum_vl_lut - (vl + 0.5 * sc.delta) * (D // 4)) < 1e-4 * abs(sc.sum_vl_lut) + 1e-6


@needs_ref
def test_fast_math_sensitivity_of_quantisation():
    """The reference ships -Ofast builds: its LUT quantisation and code rounding multiply by a
    hoisted reciprocal.  The oracle (and the GPU) reproduce that bit for bit."""
    O, R = oracle(), ref()
    g = rng(11)
    mism_lut = mism_delta = mism_code = 0
    total_lut = total_code = 0
    for t in range(400):
        D = 128 if t % 2 else 960
        r = (g.standard_normal(D) * g.uniform(0.1, 10)).astype(np.float32)
        lf = np.zeros(D * 4, np.float32)
        O.orc_pack_lut(D, r, lf)
        vl, vr = np.float32(lf.min()), np.float32(lf.max())
        sc = LutScalars()
        l16 = np.zeros(D * 4, np.uint16)
        O.orc_lut_prepare(D, r, l16, C.byref(sc))
        mism_delta += int(np.float32(R.ref_lut_delta(vl, vr)) != np.float32(sc.delta))
        l16r = np.zeros(D * 4, np.uint16)
        R.ref_lut_quantize(lf, D * 4, vl, np.float32(sc.delta), l16r)
        mism_lut += int(np.sum(l16r != l16))
        total_lut += D * 4
        bits = int(g.integers(1, 9))
        cm = (1 << bits) - 1
        vmx = float(max(r.max(), -r.min()))
        delta = 2 * vmx / (cm + 1)
        cf = np.zeros(D, np.float32)
        R.ref_code_floor(r, D, -vmx, delta, cm, cf)
        code = np.zeros(D, np.uint16)
        info = CodeInfo()
        cfg = CaqCfg(0, 1e-8, 0)
        O.orc_caq_encode(r, D, bits, C.byref(cfg), None, code, C.byref(info))
        mism_code += int(np.sum(cf.astype(np.uint16) != code))
        total_code += D
    # the oracle restates the release build: 1/delta hoisted and multiplied (-freciprocal-math)
    print(f"oracle vs -Ofast reference: delta {mism_delta}/400, lut {mism_lut}/{total_lut}, "
          f"codes {mism_code}/{total_code}")
    assert mism_delta == 0 and mism_lut == 0 and mism_code == 0


# ------------------------------------------------------------------ extra-bit codes
@needs_ref
@pytest.mark.parametrize("k", list(range(1, 13)))
@pytest.mark.parametrize("D", [64, 128, 960])
def test_compact_and_compute_ip_match_reference(k, D):
    O, R = oracle(), ref()
    g = rng(k * 1000 + D)
    for _ in range(4):
        raw = g.integers(0, 1 << k, size=D, dtype=np.uint16)
        nbytes = D * k // 8
        a = np.zeros(nbytes + 64, np.uint8)
        b = np.zeros(nbytes + 64, np.uint8)
        O.orc_compact_code(k, raw, D, a)
        R.ref_compact_code(k, raw, D, b)
        assert np.array_equal(a[:nbytes], b[:nbytes])
        dec = np.array([O.orc_code_at(k, a, D, d) for d in range(D)])
        assert np.array_equal(dec, raw.astype(np.int64))
        q = (g.standard_normal(D) * 3).astype(np.float32)
        x = np.float32(O.orc_compute_ip(k, q, a, D))
        y = np.float32(R.ref_compute_ip(k, q, a, D))
        assert x.view(np.uint32) == y.view(np.uint32)
        assert abs(float(x) - float(q.astype(np.float64) @ raw.astype(np.float64))) < 1e-3 * (1 << k) * D ** 0.5


# ------------------------------------------------------------------ ResultPool (via search)
def pool_model(k, items):
    """Sorted-list model of utils/pool.hpp:12-61: ties go before older equal entries."""
    lst = []
    for u, d in items:
        if len(lst) == k and d > lst[-1][1]:
            continue
        pos = sum(1 for _, x in lst if x < d)
        lst.insert(pos, (u, d))
        lst = lst[:k]
    return lst


# ------------------------------------------------------------------ plan DP
def brute_plan(var, D_pad, avg_bits):
    nbf = 64
    tot = int(np.float32(avg_bits) * np.float32(D_pad) + np.float32(nbf))
    chunks = D_pad // 64
    maxseg = chunks if avg_bits < 2 else chunks // 2
    csum = [float(np.float32(var[64 * i:64 * i + 64].sum())) for i in range(chunks)]
    best = [np.inf]

    def rec(i, ns, used, err):
        if i == chunks:
            best[0] = min(best[0], err)
            return
        if ns == maxseg:
            return
        # zero-bit tail segment
        rec(chunks, ns + 1, used, err + sum(csum[i:]))
        for j in range(1, chunks - i + 1):
            vs = sum(csum[i:i + j])
            for b in range(1, 14):
                nb = used + b * j * 64 + nbf
                if nb > tot:
                    break
                rec(i + j, ns + 1, nb, err + vs / (1 << b))
    rec(0, 0, 0, 0.0)
    return best[0], tot


@pytest.mark.parametrize("D_pad,avg", [(128, 4.0), (192, 1.5), (256, 2.0), (256, 3.3), (192, 0.8)])
def test_plan_dp_against_brute_force(D_pad, avg):
    O = oracle()
    g = rng(D_pad)
    var = (1.0 / np.sqrt(1 + np.arange(D_pad) / 8) * g.uniform(0.5, 1.5, D_pad)).astype(np.float32) ** 2
    dims = np.zeros(64, np.int32)
    bits = np.zeros(64, np.int32)
    n = O.orc_plan_dp(var, D_pad, avg, dims, bits)
    dims, bits = dims[:n], bits[:n]
    assert dims.sum() == D_pad and np.all(dims % 64 == 0)
    best, tot = brute_plan(var, D_pad, avg)
    used = sum(int(d) * int(b) + 64 for d, b in zip(dims, bits) if b)
    assert used <= tot
    obj, off = 0.0, 0
    for d, b in zip(dims, bits):
        vs = sum(float(np.float32(var[off + 64 * i:off + 64 * i + 64].sum())) for i in range(d // 64))
        obj += vs / (1 << int(b))
        off += int(d)
    assert obj <= best * 1.01 + 1e-12


def test_plan_eqseg():
    O = oracle()
    dims = np.zeros(64, np.int32)
    bits = np.zeros(64, np.int32)
    n = O.orc_plan_eqseg(192, 4.0, 2, dims, bits)
    assert list(dims[:n]) == [128, 64] and list(bits[:n]) == [4, 4]
    n = O.orc_plan_eqseg(1024, 3.0, 4, dims, bits)
    assert list(dims[:n]) == [256] * 4 and list(bits[:n]) == [3] * 4


# ------------------------------------------------------------------ encoder
@pytest.mark.parametrize("bits", [1, 2, 4, 8, 9])
def test_caq_encode_properties(bits):
    O = oracle()
    g = rng(bits)
    D = 64
    cfg = CaqCfg(1000, 0.0, 0)
    for _ in range(5):
        o = (g.standard_normal(D) * g.uniform(0.1, 5)).astype(np.float32)
        code = np.zeros(D, np.uint16)
        info = CodeInfo()
        O.orc_caq_encode(o, D, bits, C.byref(cfg), None, code, C.byref(info))
        assert info.status == 0
        cm = (1 << bits) - 1
        assert code.max() <= cm
        vmx = float(max(o.max(), -o.min()))
        delta = 2 * vmx / (cm + 1)
        oa = (code.astype(np.float64) + 0.5) * delta - vmx
        o64 = o.astype(np.float64)

        def cos(x):
            return float(o64 @ x / np.linalg.norm(x) / np.linalg.norm(o64))
        # floor initialisation, then greedy ascent: never worse than the rounding start
        c0 = np.minimum(np.floor((o - np.float32(-vmx)) / np.float32(delta)), cm)
        oa0 = (c0.astype(np.float64) + 0.5) * delta - vmx
        assert cos(oa) >= cos(oa0) - 1e-12
        # the loop stopped (rounds under the limit) => no single +-1 move improves the cosine
        for j in range(D):
            for s in (-1, 1):
                c = int(code[j]) + s
                if 0 <= c <= cm:
                    t = oa.copy()
                    t[j] = (c + 0.5) * delta - vmx
                    assert cos(t) <= cos(oa) + 1e-9
        # factor identities (values are after rescale_vmx_to1)
        assert abs(info.o_l2norm - np.linalg.norm(o64)) < 1e-5 * np.linalg.norm(o64)
        ip_scaled = float(o64 @ oa) / vmx
        assert abs(info.ip_o_oa - ip_scaled) < 1e-6 * abs(ip_scaled)
        assert abs(info.fac_rescale * info.ip_o_oa - float(o64 @ o64)) < 1e-4 * float(o64 @ o64)


def test_caq_encode_zero_bits_and_zero_vector():
    O = oracle()
    D = 64
    cfg = CaqCfg(6, 1e-8, 0)
    code = np.zeros(D, np.uint16)
    info = CodeInfo()
    o = np.ones(D, np.float32)
    O.orc_caq_encode(o, D, 0, C.byref(cfg), None, code, C.byref(info))
    assert info.fac_rescale == 0 and abs(info.o_l2norm - 8.0) < 1e-6
    z = np.zeros(D, np.float32)
    O.orc_caq_encode(z, D, 4, C.byref(cfg), None, code, C.byref(info))
    assert info.o_l2norm == 0 and info.status == 0


def test_cluster_sizes_layout():
    O = oracle()
    s = Sizes()
    p = make_plan([128], [4], 128)
    O.orc_cluster_sizes(C.byref(p), 70, C.byref(s))
    assert s.num_blocks == 3 and s.short_factor_floats == 0
    assert s.short_code_bytes == 3 * (256 + 4 * 128)
    assert s.long_code_bytes == (70 * 128 * 3 // 8 + 15) // 16 * 16
    p = make_plan([128, 64, 64], [5, 2, 0], 256)
    O.orc_cluster_sizes(C.byref(p), 33, C.byref(s))
    assert s.short_factor_floats == 2 * 3 * 64
    assert s.short_code_bytes == 2 * (4 * 128 + 4 * 64)
    assert s.long_stride == 64 + 16 and s.long_code_bytes == 80 * 33


# ------------------------------------------------------------------ index file format
def test_index_file_config_layout_matches_reference():
    """The raw QuantizeConfig / QuantSingleConfig bytes IVF::save writes (saq_data.hpp:30,
    quantizer_data.hpp:29): the offsets tests/index_file.py and saq_amd/host/saqlib_gpu.hpp
    use against the compiled reference config.h."""
    import index_file as IF
    R = ref()
    if R is None or not hasattr(R, "ref_config_layout"):
        pytest.skip("reference harness not built")
    out = np.zeros(16, np.int64)
    R.ref_config_layout(out.ctypes.data)
    assert out[0] == IF.CFG_BYTES and out[6] == IF.SINGLE_CFG_BYTES and out[5] == IF.SINGLE_OFF
    assert [out[1], out[2], out[3], out[4]] == [IF.CFG_FIELDS[k][0] for k in
                                                ("avg_bits", "seg_eqseg", "enable_segmentation", "use_compact_layout")]
    assert [out[7], out[8], out[9], out[10], out[11], out[12]] == [
        IF.SINGLE_FIELDS[k][0] for k in ("quant_type", "random_rotation", "use_fastscan", "caq_adj_rd_lmt",
                                         "caq_adj_eps", "caq_ori_qB")]
    assert out[13] == 4 and out[14] == 1  # int-sized enum; use_fastscan defaults to true


# ---- std::partial_sort's tie order (FlatInitializer::centroids_distances, initializer.hpp:52-76)
def _adjust(h, hole, n, v, less):
    top, child = hole, hole
    while child < (n - 1) // 2:
        child = 2 * (child + 1)
        if less(h[child], h[child - 1]):
            child -= 1
        h[hole] = h[child]
        hole = child
    if n % 2 == 0 and child == (n - 2) // 2:
        child = 2 * (child + 1)
        h[hole] = h[child - 1]
        hole = child - 1
    parent = (hole - 1) // 2
    while hole > top and less(h[parent], v):
        h[hole] = h[parent]
        hole = parent
        parent = (hole - 1) // 2
    h[hole] = v


def stl_partial_sort(items, m, less):
    """GCC 11 libstdc++'s std::partial_sort (bits/stl_algo.h __partial_sort = __heap_select +
    __sort_heap; bits/stl_heap.h __make_heap / __pop_heap / __adjust_heap / __push_heap), the
    library the reference builds with: the first m of items in its order."""
    h = list(items[:m])
    if m >= 2:
        for parent in range((m - 2) // 2, -1, -1):
            _adjust(h, parent, m, h[parent], less)
    for x in items[m:]:
        if less(x, h[0]):
            _adjust(h, 0, m, x, less)
    for last in range(m - 1, 0, -1):
        v = h[last]
        h[last] = h[0]
        _adjust(h, 0, last, v, less)
    return h


@pytest.mark.parametrize("dist_type", [0, 1])
def test_probe_order_is_libstdcxx_partial_sort(dist_type):
    """The oracle's coarse quantiser orders exact distance ties as std::partial_sort does (the
    reference's toolchain): Candidate's operator< for L2, std::greater<> through operator> =
    !(a < b) for IP (defines.hpp:52-54).  Duplicated centroids make every query tie."""
    O = oracle()
    rng = np.random.default_rng(5)
    K, d = 300, 16
    cent = rng.integers(-2, 3, (K, d)).astype(np.float32)  # small integer grid: many exact ties
    cent[150:] = cent[:150]
    for qi in range(6):
        q = rng.integers(-2, 3, d).astype(np.float32)
        dist = np.array([O.orc_eigen_dot(q, np.ascontiguousarray(c), d) if dist_type else
                         O.orc_eigen_diff_sqnorm(np.ascontiguousarray(c), q, d) for c in cent], np.float32)
        items = [(i, float(dist[i])) for i in range(K)]
        less = (lambda a, b: not (a[1] < b[1])) if dist_type else (lambda a, b: a[1] < b[1])
        for m in (1, 2, 7, 150, 151, 299, 300):
            want = [i for i, _ in stl_partial_sort(items, m, less)]
            got = np.zeros(m, np.uint32)
            O.orc_probe(cent, K, d, q, m, dist_type, got, None)
            assert list(got) == want, f"query {qi} nprobe {m}"
