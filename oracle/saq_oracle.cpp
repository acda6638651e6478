// saq_oracle.cpp -- TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path.
//
// A from-scratch CPU restatement of the reference SAQ hot path (howarlii/SAQ).
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it.  Every function cites the reference file:line it restates (paths relative
// to the reference's saqlib/ unless stated otherwise).
//
// Arithmetic contract: IEEE single/double, reference source order, no FP
// contraction (the reference builds with -ffp-contract=off, CMakeLists.txt:90).
// std::fmaf appears exactly where the reference calls _mm512_fmadd_ps or an
// Eigen pmadd.  Reductions follow the vectorised order the reference executes:
// Eigen's redux (third/Eigen/src/Core/Redux.h:227-272) with AVX512 predux, and
// GCC's _mm512_reduce_add_ps, which share one 16->8->4->2->1 tree.
//
// How each part is pinned (tests/test_oracle_*.py): the reductions, GEMV, LUT packing and
// quantisation, fast-scan accumulate, code packers and compute_ip bit for bit against the
// reference's own headers compiled in oracle/_ref; the cluster and single-vector estimators by
// float64 closed forms; scanCluster / searchCluster by independent block models; the CAQ
// adjustment by an exact-rational model; the plan DP by brute force.  Parity unpinned: the
// encoder's float64 sums under the reference's -Ofast (this file keeps source order, and no
// -Ofast encoder can be built here: its headers include glog) and the rotation GEMM order of
// IVF::construct (orc_rotate_rows follows the GEMV order).
#include "saq_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>
#if defined(ORC_AVX512)
#include <immintrin.h>
#endif

namespace {

constexpr int64_t kBlock = 32;  // KFastScanSize, defines.hpp:15
constexpr int kPos[16] = {3, 3, 2, 3, 1, 3, 2, 3, 0, 3, 2, 3, 1, 3, 2, 3};  // fastscan.hpp:22-39
constexpr int kPerm0[16] = {0, 8, 1, 9, 2, 10, 3, 11, 4, 12, 5, 13, 6, 14, 7, 15};  // fastscan.hpp:41-42

inline int64_t rd_up(int64_t x, int64_t y) { return (x + y - 1) / y * y; }  // utils/tools.hpp:8-14

// Eigen redux over n elements produced by elem(i) (Redux.h:227-272): two 16-lane
// packet accumulators stride 32, one trailing packet, predux tree, scalar tail.
// The first aligned index is 0 for every call site on the path: plain vectors
// are 64-byte aligned and expressions report no direct access
// (DenseCoeffsBase.h:643-647).
template <class F>
inline float eigen_redux(int64_t n, F elem) {
    if (n < 16) {
        float r = elem(0);
        for (int64_t i = 1; i < n; ++i) r = r + elem(i);
        return r;
    }
    const int64_t aligned = n / 16 * 16, aligned2 = n / 32 * 32;
    float p0[16], p1[16];
    for (int l = 0; l < 16; ++l) p0[l] = elem(l);
    if (aligned > 16) {
        for (int l = 0; l < 16; ++l) p1[l] = elem(16 + l);
        for (int64_t idx = 32; idx < aligned2; idx += 32)
            for (int l = 0; l < 16; ++l) {
                p0[l] = p0[l] + elem(idx + l);
                p1[l] = p1[l] + elem(idx + 16 + l);
            }
        for (int l = 0; l < 16; ++l) p0[l] = p0[l] + p1[l];
        if (aligned > aligned2)
            for (int l = 0; l < 16; ++l) p0[l] = p0[l] + elem(aligned2 + l);
    }
    float r = orc_tree16(p0);
    for (int64_t i = aligned; i < n; ++i) r = r + elem(i);
    return r;
}

inline float max0(float v) { return (0.0f > v) ? 0.0f : v; }  // _mm512_max_ps(zero, v)

// ResultPool (utils/pool.hpp:12-61).
struct Pool {
    bool greater;
    int64_t cap, size = 0;
    std::vector<uint32_t> ids;
    std::vector<float> dist;
    Pool(int64_t k, bool g) : greater(g), cap(k), ids(k + 1, 0), dist(k + 1, 0.0f) {}
    int64_t bsearch(float d) const {  // pool.hpp:51-60
        int64_t lo = 0, len = size;
        while (len > 1) {
            int64_t half = len >> 1;
            len -= half;
            lo += (dist[lo + half - 1] < d) * half;
        }
        return (lo < size && dist[lo] < d) ? lo + 1 : lo;
    }
    void insert(uint32_t u, float d) {  // pool.hpp:17-30
        if (greater) d = -d;
        if (size == cap && d > dist[size - 1]) return;
        int64_t lo = bsearch(d);
        std::memmove(&ids[lo + 1], &ids[lo], (size - lo) * sizeof(uint32_t));
        ids[lo] = u;
        std::memmove(&dist[lo + 1], &dist[lo], (size - lo) * sizeof(float));
        dist[lo] = d;
        size += (size < cap);
    }
    float distk_raw() const { return size == cap ? dist[size - 1] : FLT_MAX; }  // pool.hpp:46-49
    float distk() const { return greater ? -distk_raw() : distk_raw(); }
};

// Per-segment layout offsets of SaqCluData (cluster_data.hpp:206-294).
struct Layout {
    int nseg;
    int64_t num_vec, num_blocks;
    bool single;
    int64_t fac_per_block;          // floats per block in short_factors (multi)
    int64_t code_per_block;         // bytes per block in short_code
    int64_t fac_off[ORC_MAX_SEG];   // multi: float offset of segment in a block row
    int64_t code_off[ORC_MAX_SEG];  // byte offset of segment codes within a block
    int64_t long_bytes[ORC_MAX_SEG];
    int64_t long_off[ORC_MAX_SEG];  // compact: region start; non-compact: slot start
    int64_t long_stride;            // non-compact per-vector stride (0 for compact)
    int64_t long_total;
    bool compact;
    int64_t cen_off[ORC_MAX_SEG];
};

Layout make_layout(const orc_plan_t *p, int64_t n) {
    Layout L{};
    L.nseg = p->num_seg;
    L.num_vec = n;
    L.num_blocks = (n + kBlock - 1) / kBlock;
    L.single = (p->num_seg == 1);
    L.compact = L.single || p->use_compact_layout;  // cluster_data.hpp:211-212
    int64_t fac = 0, code = 0, lpv = 0, ltot = 0, cen = 0;
    for (int s = 0; s < L.nseg; ++s) {
        const int64_t D = p->seg_dim[s], b = p->seg_bits[s];
        L.cen_off[s] = cen;
        cen += D;
        L.fac_off[s] = fac;
        fac += 2 * kBlock;  // kNumShortFactors * 32
        L.code_off[s] = code;
        code += b ? D * kBlock / 8 : 0;
        L.long_bytes[s] = b ? D * (b - 1) / 8 : 0;
        if (L.compact) {
            L.long_off[s] = ltot;
            ltot += rd_up(L.long_bytes[s] * n, 16);
        } else {
            L.long_off[s] = lpv;
            lpv += rd_up(L.long_bytes[s], 16);
        }
    }
    if (L.single) {
        // one interleaved block: [32 o_l2norm][32 ip_cent_oa][codes] (cluster_data.hpp:233-252)
        L.fac_per_block = 0;
        L.code_off[0] = 2 * kBlock * sizeof(float);
        L.code_per_block = 2 * kBlock * sizeof(float) + code;
    } else {
        L.fac_per_block = fac;
        L.code_per_block = code;
    }
    L.long_stride = L.compact ? 0 : lpv;
    L.long_total = L.compact ? ltot : lpv * n;
    return L;
}

inline const float *o_l2norm_row(const Layout &L, const orc_cluster_t &c, int s, int64_t blk) {
    if (L.single) return reinterpret_cast<const float *>(c.short_code + blk * L.code_per_block);
    return c.short_factors + blk * L.fac_per_block + L.fac_off[s];
}
inline const uint8_t *block_code(const Layout &L, const orc_cluster_t &c, int s, int64_t blk) {
    return c.short_code + blk * L.code_per_block + L.code_off[s];
}
inline const uint8_t *long_code(const Layout &L, const orc_cluster_t &c, int s, int64_t v) {
    if (L.compact) return c.long_code + L.long_off[s] + v * L.long_bytes[s];
    return c.long_code + L.long_off[s] + v * L.long_stride;
}

// CaqCluEstimator per segment (quantization/caq/caq_estimator.hpp:27-218) with
// its Lut (quantization/fastscan/lut.hpp:20-127).
struct SegEst {
    int64_t D;
    int bits;
    bool ip;
    std::vector<float> qrot;  // query_data_ (rotated query segment)
    std::vector<float> r;     // Lut::query_ (residual, or the query for IP)
    std::vector<uint16_t> lut16;
#if defined(ORC_AVX512)
    std::vector<uint8_t> lutpk;  // per 8-dim column: lo bytes of its 2 codebooks' LUTs, then hi bytes
#endif
    orc_lut_scalars_t sc{};
    float bound = 0, ip_q_c = 0, q_l2sqr = 0;
    float sumq_cb = 0, rescale_sqrtD = 0;
    float ipxb[kBlock];
    uint64_t fast_bits = 0, acc_bits = 0;

    void prepare(const float *c_rot) {  // caq_estimator.hpp:94-106
        r.resize(D);
        if (ip) {
            ip_q_c = orc_eigen_dot(qrot.data(), c_rot, D);
            for (int64_t i = 0; i < D; ++i) r[i] = qrot[i];
        } else {
            for (int64_t i = 0; i < D; ++i) r[i] = qrot[i] - c_rot[i];
        }
        lut16.resize(D / 4 * 16);
        orc_lut_prepare(D, r.data(), lut16.data(), &sc);
#if defined(ORC_AVX512)
        lutpk.resize(D / 8 * 64);
        for (int64_t i = 0; i < D / 8; ++i)
            for (int h = 0; h < 2; ++h)
                for (int e = 0; e < 16; ++e) {
                    const uint16_t v = lut16[(2 * i + h) * 16 + e];
                    lutpk[i * 64 + 16 * h + e] = (uint8_t)(v & 0xFF);
                    lutpk[i * 64 + 32 + 16 * h + e] = (uint8_t)(v >> 8);
                }
#endif
        q_l2sqr = sc.q_l2sqr;
        // lut.hpp:104-105: set1_ps(0.5 * sum_q_ - const_bound * q_l2norm_), set1_ps(4 / 0.8f * one_over_sqrtD_)
        sumq_cb = (float)(0.5 * (double)sc.sum_q - (double)(0.58f * sc.q_l2norm));
        const float one_over_sqrtD = (float)(1.0 / (double)std::sqrt((float)D));  // lut.hpp:47
        rescale_sqrtD = (4 / 0.8f) * one_over_sqrtD;
    }
    void vars_est(const float *x, float *out) const {  // caq_estimator.hpp:116-136
        if (ip) {
            for (int j = 0; j < kBlock; ++j) out[j] = ip_q_c - bound;
            return;
        }
        const float f = q_l2sqr - 2 * bound;
        for (int j = 0; j < kBlock; ++j) out[j] = max0(x[j] * x[j] + f);
    }
    void fast(const float *x, const uint8_t *codes, float *out) {  // caq_estimator.hpp:148-179
        if (bits == 0) {
            vars_est(x, out);
            return;
        }
        uint32_t S[kBlock];
#if defined(ORC_AVX512)
        accumulate_avx512(codes, lutpk.data(), D, S);
        fast_epilogue_avx512(S, x, out);
        fast_bits += kBlock * D;
        return;
#else
        orc_accumulate(codes, lut16.data(), D, S);
#endif
        for (int j = 0; j < kBlock; ++j) {  // lut.hpp:107-118
            float t = (float)(int32_t)S[j] * sc.delta;
            t = t + sc.sum_vl_lut;
            ipxb[j] = t;
            float f = (t - sumq_cb) * rescale_sqrtD;
            f = f * x[j];
            if (ip)
                out[j] = f * 0.5f + ip_q_c;
            else
                out[j] = max0(x[j] * x[j] + (q_l2sqr - f));
        }
        fast_bits += kBlock * D;
    }
#if defined(ORC_AVX512)
    // The exact integer sums of accumulate_hacc with byte shuffles: per 8-dim column one 32-byte load
    // (two codebooks x 32 vectors), its low and high nibbles index the LUTs' low and high bytes
    // (vpshufb per 16-byte lane = per codebook), the bytes summed in u16 lanes and folded into u32
    // every 128 columns (128 x 255 < 2^16).  Lane j of a half stands for vector kPerm0[j] (low
    // nibbles) or kPerm0[j] + 16 (high nibbles).
    static void accumulate_avx512(const uint8_t *codes, const uint8_t *lutpk, int64_t D, uint32_t *S) {
        const __m256i nib = _mm256_set1_epi8(0x0F);
        __m512i t32[4] = {_mm512_setzero_si512(), _mm512_setzero_si512(), _mm512_setzero_si512(),
                          _mm512_setzero_si512()};  // (lo, hi) x (low-nibble vectors, high-nibble vectors)
        const int64_t cols = D / 8;
        for (int64_t c0 = 0; c0 < cols; c0 += 128) {
            __m512i a_lo = _mm512_setzero_si512(), a_hi = _mm512_setzero_si512();
            __m512i b_lo = _mm512_setzero_si512(), b_hi = _mm512_setzero_si512();
            const int64_t c1 = std::min(cols, c0 + 128);
            for (int64_t i = c0; i < c1; ++i) {
                const __m256i cw = _mm256_loadu_si256(reinterpret_cast<const __m256i *>(codes + 32 * i));
                const __m256i tlo = _mm256_loadu_si256(reinterpret_cast<const __m256i *>(lutpk + 64 * i));
                const __m256i thi = _mm256_loadu_si256(reinterpret_cast<const __m256i *>(lutpk + 64 * i + 32));
                const __m256i il = _mm256_and_si256(cw, nib), ih = _mm256_and_si256(_mm256_srli_epi16(cw, 4), nib);
                a_lo = _mm512_add_epi16(a_lo, _mm512_cvtepu8_epi16(_mm256_shuffle_epi8(tlo, il)));
                a_hi = _mm512_add_epi16(a_hi, _mm512_cvtepu8_epi16(_mm256_shuffle_epi8(thi, il)));
                b_lo = _mm512_add_epi16(b_lo, _mm512_cvtepu8_epi16(_mm256_shuffle_epi8(tlo, ih)));
                b_hi = _mm512_add_epi16(b_hi, _mm512_cvtepu8_epi16(_mm256_shuffle_epi8(thi, ih)));
            }
            const __m512i *acc[4] = {&a_lo, &a_hi, &b_lo, &b_hi};
            for (int k = 0; k < 4; ++k) {
                const __m512i v = *acc[k];
                // u16 lanes 0..15: codebook 2i, 16..31: codebook 2i+1; both belong to the same vectors
                const __m512i w = _mm512_add_epi32(_mm512_cvtepu16_epi32(_mm512_castsi512_si256(v)),
                                                   _mm512_cvtepu16_epi32(_mm512_extracti64x4_epi64(v, 1)));
                t32[k] = _mm512_add_epi32(t32[k], w);
            }
        }
        alignas(64) uint32_t lo_a[16], hi_a[16], lo_b[16], hi_b[16];
        _mm512_store_si512(reinterpret_cast<__m512i *>(lo_a), t32[0]);
        _mm512_store_si512(reinterpret_cast<__m512i *>(hi_a), t32[1]);
        _mm512_store_si512(reinterpret_cast<__m512i *>(lo_b), t32[2]);
        _mm512_store_si512(reinterpret_cast<__m512i *>(hi_b), t32[3]);
        for (int j = 0; j < 16; ++j) {
            S[kPerm0[j]] = lo_a[j] + (hi_a[j] << 8);
            S[kPerm0[j] + 16] = lo_b[j] + (hi_b[j] << 8);
        }
    }
    // lut.hpp:107-118 + caq_estimator.hpp:161-176 on 16 lanes, the same IEEE operations in the same
    // order as the scalar form (no contraction).
    void fast_epilogue_avx512(const uint32_t *S, const float *x, float *out) {
        const __m512 delta = _mm512_set1_ps(sc.delta), vl = _mm512_set1_ps(sc.sum_vl_lut);
        const __m512 sq = _mm512_set1_ps(sumq_cb), rs = _mm512_set1_ps(rescale_sqrtD);
        const __m512 q2 = _mm512_set1_ps(q_l2sqr), qc = _mm512_set1_ps(ip_q_c), half = _mm512_set1_ps(0.5f);
        for (int h = 0; h < 2; ++h) {
            const __m512i s = _mm512_loadu_si512(reinterpret_cast<const void *>(S + 16 * h));
            __m512 t = _mm512_mul_ps(_mm512_cvtepi32_ps(s), delta);
            t = _mm512_add_ps(t, vl);
            _mm512_storeu_ps(ipxb + 16 * h, t);
            const __m512 xv = _mm512_loadu_ps(x + 16 * h);
            __m512 f = _mm512_mul_ps(_mm512_sub_ps(t, sq), rs);
            f = _mm512_mul_ps(f, xv);
            __m512 o;
            if (ip)
                o = _mm512_add_ps(_mm512_mul_ps(f, half), qc);
            else
                o = _mm512_max_ps(_mm512_setzero_ps(), _mm512_add_ps(_mm512_mul_ps(xv, xv), _mm512_sub_ps(q2, f)));
            _mm512_storeu_ps(out + 16 * h, o);
        }
    }
#endif
    float accurate(float x, const uint8_t *lc, int64_t j, const float *rescale_ptr) {  // :190-217
        const float o_l2sqr = x * x;
        if (bits == 0) return ip ? ip_q_c : o_l2sqr + q_l2sqr;
        const float sqd = (float)(2.0 / (1 << bits));
        const double ex = (double)orc_compute_ip(bits - 1, r.data(), lc, D);  // lut.hpp:121-126
        const float t = (float)(((double)ipxb[j] + ex * (double)sqd) + (-1.0 + (double)(sqd / 2)) * (double)sc.sum_q);
        const float ipoq = (*rescale_ptr) * t;
        acc_bits += D * (bits - 1);
        if (ip) return ipoq + ip_q_c;
        return o_l2sqr + q_l2sqr - 2 * ipoq;
    }
};

struct QueryState {
    std::vector<SegEst> seg;
};

void init_query(const orc_plan_t *plan, const float *variance, const float *const *P,
                const float *query, float m, int dist_type, QueryState &qs) {
    // SaqEstimatorBase ctor (saq_estimator.hpp:28-48) + CaqCluEstimator ctor (caq_estimator.hpp:56-68)
    qs.seg.resize(plan->num_seg);
    int64_t off = 0;
    for (int s = 0; s < plan->num_seg; ++s) {
        SegEst &e = qs.seg[s];
        e.D = plan->seg_dim[s];
        e.bits = plan->seg_bits[s];
        e.ip = (dist_type == 1);
        std::vector<float> cq(e.D, 0.0f);
        const int64_t copy = std::min<int64_t>(e.D, plan->num_dim - off);
        for (int64_t i = 0; i < copy; ++i) cq[i] = query[off + i];
        const float vars2 = orc_eigen_var_q2(variance + off, cq.data(), e.D);
        e.bound = std::sqrt(vars2) * m;  // setPruneBound(sqrt(vars2)), caq_estimator.hpp:80-82
        e.qrot.resize(e.D);
        if (P && P[s])
            orc_gemv_rowvec(cq.data(), P[s], e.D, e.qrot.data());
        else
            e.qrot = cq;
        off += e.D;
    }
}

// SAQSearcher::scanCluster (saq_searcher.hpp:176-214).
void scan_cluster(const Layout &L, const orc_cluster_t &c, QueryState &qs, Pool &pool, orc_metrics_t &mt) {
    SegEst &e = qs.seg[0];
    e.prepare(c.centroid_rot);
    float distk = pool.distk();
    float est[kBlock];
    for (int64_t blk = 0; blk < L.num_blocks; ++blk) {
        const int64_t cnt = (blk == L.num_blocks - 1) ? L.num_vec % kBlock : kBlock;
        const float *x = o_l2norm_row(L, c, 0, blk);
        e.fast(x, block_code(L, c, 0, blk), est);
        uint32_t mask = 0;
        for (int j = 0; j < kBlock; ++j) mask |= (uint32_t)(est[j] < distk) << j;
        mask = (uint32_t)(mask & ((1ull << cnt) - 1));
        while (mask) {
            const int j = __builtin_ctz(mask);
            mask -= 1u << j;
            const int64_t idx = kBlock * blk + j;
            const float d = e.accurate(x[j], long_code(L, c, 0, idx), j, &c.long_factors[(idx * L.nseg) * 2]);
            pool.insert(c.ids[idx], d);
            distk = pool.distk();
        }
    }
    mt.total_comp_cnt += L.num_blocks * kBlock;
}

// SAQSearcher::searchCluster<enable_var=true> (saq_searcher.hpp:68-173).
void search_cluster(const Layout &L, const orc_cluster_t &c, QueryState &qs, Pool &pool, orc_metrics_t &mt) {
    const int S = L.nseg;
    for (int s = 0; s < S; ++s) qs.seg[s].prepare(c.centroid_rot + L.cen_off[s]);
    float distk = pool.distk();
    std::vector<float> cd(S * kBlock);
    float curr[kBlock];
    auto bmin = [](const float *v) {
        float m = v[0];
        for (int j = 1; j < kBlock; ++j) m = std::min(m, v[j]);
        return m;
    };
    for (int64_t blk = 0; blk < L.num_blocks; ++blk) {
        for (int j = 0; j < kBlock; ++j) curr[j] = 0.0f;
        const int64_t begin = blk * kBlock;
        float mi = FLT_MAX;
        for (int s = 0; s < S; ++s) {  // stage 1
            qs.seg[s].vars_est(o_l2norm_row(L, c, s, blk), &cd[s * kBlock]);
            for (int j = 0; j < kBlock; ++j) curr[j] = curr[j] + cd[s * kBlock + j];
        }
        mi = bmin(curr);
        if (mi > distk) continue;
        for (int s = 0; s < S; ++s) {  // stage 2
            if (qs.seg[s].bits == 0) continue;
            for (int j = 0; j < kBlock; ++j) curr[j] = curr[j] - cd[s * kBlock + j];
            qs.seg[s].fast(o_l2norm_row(L, c, s, blk), block_code(L, c, s, blk), &cd[s * kBlock]);
            for (int j = 0; j < kBlock; ++j) curr[j] = curr[j] + cd[s * kBlock + j];
            mi = bmin(curr);
            if (mi > distk) break;
        }
        if (mi <= distk) {  // stage 3
            for (int j = 0; j < kBlock; ++j) {
                if (curr[j] < distk) {
                    const int64_t idx = begin + j;
                    if (idx >= L.num_vec) break;
                    float acc = curr[j];
                    for (int s = 0; s < S; ++s) {
                        const float x = o_l2norm_row(L, c, s, blk)[j];
                        const float a = qs.seg[s].accurate(x, long_code(L, c, s, idx), j,
                                                           &c.long_factors[(idx * S + s) * 2]);
                        acc += a - cd[s * kBlock + j];
                        if (acc >= distk) break;
                    }
                    pool.insert(c.ids[idx], acc);
                    distk = pool.distk();
                }
            }
        }
    }
    mt.total_comp_cnt += L.num_blocks * kBlock;
}

}  // namespace

extern "C" {

float orc_tree16(const float *a) {
    // predux<Packet16f> with AVX512DQ (arch/AVX512/PacketMath.h:978-983 -> AVX/PacketMath.h:798-801
    // -> SSE/PacketMath.h:987-997); GCC's _mm512_reduce_add_ps computes the same tree.
    float b[8], c[4];
    for (int i = 0; i < 8; ++i) b[i] = a[i] + a[i + 8];
    for (int i = 0; i < 4; ++i) c[i] = b[i] + b[i + 4];
    const float d0 = c[0] + c[2], d1 = c[1] + c[3];
    return d0 + d1;
}

float orc_eigen_sum(const float *x, int64_t n) {
    return eigen_redux(n, [&](int64_t i) { return x[i]; });
}
float orc_eigen_sqnorm(const float *x, int64_t n) {  // squaredNorm = cwiseAbs2().sum()
    return eigen_redux(n, [&](int64_t i) { return x[i] * x[i]; });
}
float orc_eigen_dot(const float *a, const float *b, int64_t n) {  // dot = binaryExpr(conj_prod).sum()
    return eigen_redux(n, [&](int64_t i) { return a[i] * b[i]; });
}
float orc_eigen_diff_sqnorm(const float *a, const float *b, int64_t n) {  // (a - b).squaredNorm()
    return eigen_redux(n, [&](int64_t i) {
        const float d = a[i] - b[i];
        return d * d;
    });
}
float orc_eigen_var_q2(const float *var, const float *q, int64_t n) {
    // (var.array() * q.array().square()).sum(), saq_estimator.hpp:42
    return eigen_redux(n, [&](int64_t i) { return var[i] * (q[i] * q[i]); });
}

void orc_gemv_rowvec(const float *q, const float *P, int64_t D, float *out) {
    // q (1xD) * P (DxD row-major): Eigen evaluates Pᵀ·qᵀ with the column-major GEMV kernel
    // (third/Eigen/src/Core/products/GeneralMatrixVector.h:147-252): per block of block_cols
    // input dims an FMA chain from zero, then res = fma(c, alpha=1, res).
    const int64_t bc = D < 128 ? D : (D * (int64_t)sizeof(float) < 32000 ? 16 : 4);
    for (int64_t j = 0; j < D; ++j) out[j] = 0.0f;
    for (int64_t j2 = 0; j2 < D; j2 += bc) {
        const int64_t jend = std::min(j2 + bc, D);
        for (int64_t o = 0; o < D; ++o) {
            float c = 0.0f;
            for (int64_t i = j2; i < jend; ++i) c = std::fmaf(P[i * D + o], q[i], c);
            out[o] = std::fmaf(c, 1.0f, out[o]);
        }
    }
}

void orc_pack_lut(int64_t D, const float *q, float *lut) {  // fastscan.hpp:236-248
    for (int64_t i = 0; i < D / 4; ++i) {
        lut[0] = 0;
        for (int j = 1; j < 16; ++j) lut[j] = lut[j - (j & -j)] + q[kPos[j]];
        lut += 16;
        q += 4;
    }
}

void orc_pack_codes(int64_t D, const uint8_t *codes, int64_t num, uint8_t *blocks) {
    // fastscan.hpp:75-112; absent rows of a short final batch read as zero (get_column :45-57)
    const int64_t num_rd = (num + 31) & ~int64_t(31), cols = D / 8;
    for (int64_t row = 0; row < num_rd; row += kBlock) {
        for (int64_t i = 0; i < cols; ++i) {
            uint8_t col[32];
            for (int j = 0; j < 32; ++j) col[j] = (row + j < num) ? codes[(row + j) * cols + i] : 0;
            for (int j = 0; j < 16; ++j) {
                const int v = kPerm0[j];
                blocks[j] = (uint8_t)((col[v] >> 4) | ((col[v + 16] >> 4) << 4));
                blocks[j + 16] = (uint8_t)((col[v] & 15) | ((col[v + 16] & 15) << 4));
            }
            blocks += 32;
        }
    }
}

uint16_t orc_cvt_u16(float x) {
    // static_cast<uint16_t>(float) as GCC emits it on x86-64: cvttss2si to int32 (0x80000000 when
    // NaN or out of range), then the low 16 bits.
    int32_t v;
    if (x != x || !(x > -2147483904.0f && x < 2147483648.0f))
        v = INT32_MIN;
    else
        v = (int32_t)x;
    return (uint16_t)v;
}

void orc_lut_prepare(int64_t D, const float *r, uint16_t *lut16, orc_lut_scalars_t *sc) {
    // Lut::prepare, lut.hpp:60-87 (use_highacc_ = true)
    sc->sum_q = orc_eigen_sum(r, D);
    sc->q_l2sqr = orc_eigen_sqnorm(r, D);
    sc->q_l2norm = std::sqrt(sc->q_l2sqr);
    const int64_t T = D / 4 * 16;
    std::vector<float> lf(T);
    orc_pack_lut(D, r, lf.data());
    float vl = lf[0], vr = lf[0];
    for (int64_t i = 1; i < T; ++i) {
        vl = std::min(vl, lf[i]);
        vr = std::max(vr, lf[i]);
    }
    sc->delta = (float)((double)(vr - vl) / ((1 << 16) - 0.01));
    // lut.hpp:74-77 divides by the loop-invariant delta_; the reference's release build (-Ofast,
    // i.e. -freciprocal-math) hoists 1/delta_ and multiplies.  Restated as built, which the
    // reference kernels compiled with those flags confirm entry for entry.
    const float inv = 1.0f / sc->delta;
    for (int64_t i = 0; i < T; ++i) lut16[i] = orc_cvt_u16((lf[i] - vl) * inv);
    sc->sum_vl_lut = (float)(((double)vl + 0.5 * (double)sc->delta) * (double)(float)(T / 16));
}

void orc_accumulate(const uint8_t *codes, const uint16_t *lut16, int64_t D, uint32_t *S) {
    // accumulate_hacc (fastscan_highacc.hpp:66-139) computes the exact integer sum
    // S_v = sum_m LUT16[m][nibble_m(v)]; codebook m occupies bytes 16m..16m+15 of the block,
    // byte j holding vector kPerm0[j] (low nibble) and kPerm0[j]+16 (high nibble).
    for (int v = 0; v < kBlock; ++v) S[v] = 0;
    for (int64_t m = 0; m < D / 4; ++m) {
        const uint8_t *cb = codes + 16 * m;
        const uint16_t *t = lut16 + 16 * m;
        for (int j = 0; j < 16; ++j) {
            S[kPerm0[j]] += t[cb[j] & 15];
            S[kPerm0[j] + 16] += t[cb[j] >> 4];
        }
    }
}

int orc_code_at(int k, const uint8_t *y, int64_t D, int64_t d) {
    // Decoders of the CodeHelper<k> compacted layouts (utils/code_helper.hpp:95-591).
    const int64_t c64 = d / 64, r = d % 64, b = r % 16, part = r / 16;
    switch (k) {
    case 0: return 0;
    case 1: return (y[d / 8] >> (d % 8)) & 1;
    case 2: return (y[16 * c64 + b] >> (2 * part)) & 3;
    case 3: return ((y[24 * c64 + b] >> (2 * part)) & 3) | (((y[24 * c64 + 16 + r % 8] >> (r / 8)) & 1) << 2);
    case 4: {
        const int64_t c32 = d / 32, r32 = d % 32;
        const uint8_t v = y[16 * c32 + r32 % 16];
        return r32 < 16 ? (v & 15) : (v >> 4);
    }
    case 5: return 2 * orc_code_at(4, y, D, d) + orc_code_at(1, y + D * 4 / 8, D, d);
    case 6:
    case 7: {
        const int64_t base = (k == 6 ? 48 : 56) * c64;
        const uint8_t c1 = y[base + b], c2 = y[base + 16 + b], c3 = y[base + 32 + b];
        int v;
        if (part == 0) v = c1 & 0x3F;
        else if (part == 1) v = c2 & 0x3F;
        else if (part == 2) v = ((c1 >> 2) & 0x30) | (c3 & 0x0F);
        else v = ((c2 >> 2) & 0x30) | ((c3 >> 4) & 0x0F);
        if (k == 7) v |= ((y[base + 48 + r % 8] >> (r / 8)) & 1) << 6;
        return v;
    }
    case 8: return y[d];
    default: return y[d] | (orc_code_at(k - 8, y + D, D, d) << 8);
    }
}

float orc_compute_ip(int k, const float *q, const uint8_t *y, int64_t D) {
    // CodeHelper<k>::compute_ip (code_helper.hpp:82-92, 105-591): 16 lanes, lane l takes the
    // dims = l (mod 16) in increasing order, then _mm512_reduce_add_ps.
    if (k == 0) return 0;
    if (k > 8) {  // :85-88
        const float lo = orc_compute_ip(8, q, y, D);
        return lo + 256 * orc_compute_ip(k - 8, q, y + D, D);
    }
    if (k == 5) return 2 * orc_compute_ip(4, q, y, D) + orc_compute_ip(1, q, y + D * 4 / 8, D);  // :372-374
    float acc[16];
    for (int l = 0; l < 16; ++l) acc[l] = 0.0f;
    for (int64_t i = 0; i < D; i += 16) {
        for (int l = 0; l < 16; ++l) {
            const int c = orc_code_at(k, y, D, i + l);
            if (k == 1)
                acc[l] = acc[l] + (c ? q[i + l] : 0.0f);  // masked load + add, :111-121
            else
                acc[l] = std::fmaf(q[i + l], (float)c, acc[l]);  // _mm512_fmadd_ps
        }
    }
    return orc_tree16(acc);
}

void orc_compact_code(int k, const uint16_t *raw, int64_t D, uint8_t *out) {
    // CodeHelper<k>::compacted_code16 / compacted_code8 (code_helper.hpp:58-80, 95-568).
    if (k == 0) return;
    if (k > 8) {
        for (int64_t i = 0; i < D; ++i) out[i] = (uint8_t)(raw[i] & 0xFF);
        std::vector<uint16_t> hi(D);
        for (int64_t i = 0; i < D; ++i) hi[i] = raw[i] >> 8;
        orc_compact_code(k - 8, hi.data(), D, out + D);
        return;
    }
    auto lo = [&](int64_t d) { return (uint8_t)raw[d]; };
    switch (k) {
    case 1:
        for (int64_t i = 0; i < D; i += 8) {
            uint8_t v = 0;
            for (int j = 0; j < 8; ++j) v |= (uint8_t)((lo(i + j) & 1) << j);
            out[i / 8] = v;
        }
        return;
    case 2:
    case 3: {
        const int64_t stride = k == 2 ? 16 : 24;
        for (int64_t c = 0; c < D / 64; ++c) {
            uint8_t *o = out + stride * c;
            for (int b = 0; b < 16; ++b) {
                uint8_t v = 0;
                for (int p = 0; p < 4; ++p) v |= (uint8_t)((lo(64 * c + 16 * p + b) & 3) << (2 * p));
                o[b] = v;
            }
            if (k == 3) {
                for (int kk = 0; kk < 8; ++kk) {
                    uint8_t v = 0;
                    for (int j = 0; j < 8; ++j) v |= (uint8_t)(((lo(64 * c + 8 * j + kk) >> 2) & 1) << j);
                    o[16 + kk] = v;
                }
            }
        }
        return;
    }
    case 4:
        for (int64_t c = 0; c < D / 32; ++c)
            for (int b = 0; b < 16; ++b)
                out[16 * c + b] = (uint8_t)((lo(32 * c + b) & 15) | ((lo(32 * c + 16 + b) & 15) << 4));
        return;
    case 5: {
        std::vector<uint16_t> hi(D), b1(D);
        for (int64_t i = 0; i < D; ++i) {
            hi[i] = raw[i] >> 1;
            b1[i] = raw[i] & 1;
        }
        orc_compact_code(4, hi.data(), D, out);
        orc_compact_code(1, b1.data(), D, out + D * 4 / 8);
        return;
    }
    case 6:
    case 7: {
        const int64_t stride = k == 6 ? 48 : 56;
        for (int64_t c = 0; c < D / 64; ++c) {
            uint8_t *o = out + stride * c;
            for (int b = 0; b < 16; ++b) {
                const uint8_t v0 = lo(64 * c + b), v1 = lo(64 * c + 16 + b);
                const uint8_t v2 = lo(64 * c + 32 + b), v3 = lo(64 * c + 48 + b);
                o[b] = (uint8_t)((v0 & 0x3F) | ((v2 << 2) & 0xC0));
                o[16 + b] = (uint8_t)((v1 & 0x3F) | ((v3 << 2) & 0xC0));
                o[32 + b] = (uint8_t)((v2 & 0x0F) | ((v3 & 0x0F) << 4));
            }
            if (k == 7) {
                for (int kk = 0; kk < 8; ++kk) {
                    uint8_t v = 0;
                    for (int j = 0; j < 8; ++j) v |= (uint8_t)(((lo(64 * c + 8 * j + kk) >> 6) & 1) << j);
                    o[48 + kk] = v;
                }
            }
        }
        return;
    }
    case 8:
        for (int64_t i = 0; i < D; ++i) out[i] = lo(i);
        return;
    }
}

int orc_caq_encode(const float *o, int64_t D, int bits, const orc_caq_cfg_t *cfg, const float *centroid,
                   uint16_t *code, orc_code_info_t *info) {
    // CAQEncoder::encode + code_adjustment + encode_and_fac (caq/caq_encoder.hpp:57-142, 172-247);
    // with caq_ori_qB the codes are found at ori_qb bits (the code maximum, :163-168) and shifted
    // down to `bits` by downUpSample (:144-157)
    std::memset(info, 0, sizeof(*info));
    double v_mx = 0, v_mi = 0, delta = 0, ip = 0, oa_l2sqr = 0;
    const int code_max = (1 << (cfg->ori_qb ? cfg->ori_qb : bits)) - 1;
    if (bits > 0) {
        float mx = o[0], mn = o[0];
        for (int64_t j = 1; j < D; ++j) {
            mx = std::max(mx, o[j]);
            mn = std::min(mn, o[j]);
        }
        v_mx = std::max(mx, -mn);
        v_mi = -v_mx;
        delta = (v_mx - v_mi) / (code_max + 1);
        const double vec_sum = orc_eigen_sum(o, D);
        // caq_encoder.hpp:192 evaluates ((o - v_mi) / delta).floor().min(code_max) in float; the
        // release build hoists the reciprocal of the loop-invariant divisor (as for the LUT above).
        const float fvmi = (float)v_mi, finv = 1.0f / (float)delta;
        double ip_o_code = 0;
        uint64_t code_l2sqr = 0;
        uint32_t code_sum = 0;
        for (int64_t j = 0; j < D; ++j) {
            float cf = 0.0f;
            if (delta) cf = std::min(std::floor((o[j] - fvmi) * finv), (float)code_max);
            ip_o_code += cf * o[j];
            const int c = (int)cf;
            code[j] = (uint16_t)c;
            code_l2sqr += (uint64_t)(c * c);
            code_sum += (uint32_t)c;
        }
        ip = ip_o_code * delta + (v_mi + 0.5 * delta) * vec_sum;
        oa_l2sqr = delta * delta * (double)code_l2sqr + (delta * delta + 2 * delta * v_mi) * (double)code_sum;
        oa_l2sqr += (0.25 * delta * delta + delta * v_mi + v_mi * v_mi) * (double)D;

        if (cfg->rd_lmt && oa_l2sqr != 0) {  // code_adjustment, :57-142
            const double re_eps = (double)cfg->eps * oa_l2sqr;
            int round = 1;
            for (int adj_cnt = 1; adj_cnt; round++) {
                if (!(ip >= 0)) {
                    info->status = 1;  // CHECK_GE(ip_resi_oa, 0) would abort
                    break;
                }
                adj_cnt = 0;
                for (int64_t j = 0; j < D; ++j) {
                    int cur = 0;
                    const float oj = o[j];
                    double oa = (code[j] + 0.5) * delta + v_mi;
                    uint16_t c = code[j];
                    const double tmp = oa_l2sqr - oa * oa;
                    const double ip_delta = delta * oj;
                    while (c < code_max) {
                        const double nq = oa + delta, nl = tmp + nq * nq, ni = ip + ip_delta;
                        if ((ip * ip + re_eps) * nl >= ni * ni * oa_l2sqr) break;
                        c++;
                        ip = ni;
                        oa = nq;
                        oa_l2sqr = nl;
                        cur++;
                    }
                    while (c > 0) {
                        const double nq = oa - delta, nl = tmp + nq * nq, ni = ip - ip_delta;
                        if ((ip * ip + re_eps) * nl >= ni * ni * oa_l2sqr) break;
                        c--;
                        ip = ni;
                        oa = nq;
                        oa_l2sqr = nl;
                        cur++;
                    }
                    if (code[j] != c) {
                        code[j] = c;
                        adj_cnt += cur;
                    }
                }
                if (cfg->rd_lmt && round >= cfg->rd_lmt) break;
                double chk_oa = 0, chk_ip = 0;
                for (int64_t j = 0; j < D; ++j) {
                    const double q = (code[j] + 0.5) * delta + v_mi;
                    chk_ip += q * o[j];
                    chk_oa += q * q;
                }
                oa_l2sqr = chk_oa;
                ip = chk_ip;
            }
            info->rounds = round;
        }
        if (cfg->ori_qb) {  // downUpSample, :144-157
            const int shift = cfg->ori_qb - bits;
            delta *= 1 << shift;
            ip = 0;
            oa_l2sqr = 0;
            for (int64_t j = 0; j < D; ++j) {
                code[j] = (uint16_t)(code[j] >> shift);
                const double q = (code[j] + 0.5) * delta + v_mi;
                ip += q * o[j];
                oa_l2sqr += q * q;
            }
        }
    }
    const double o_l2sqr = orc_eigen_sqnorm(o, D);
    const double o_l2norm = std::sqrt(o_l2sqr);
    double fac_rescale = ip ? o_l2sqr / ip : 0;
    const double fac_error =
        o_l2sqr * (double)1.9f * std::sqrt((((o_l2sqr * oa_l2sqr) / (ip * ip)) - 1) / (double)(D - 1));
    if (v_mx) {  // CaqCode::rescale_vmx_to1, :29-40
        const double sr = 1.0 / v_mx;
        delta *= sr;
        ip *= sr;
        oa_l2sqr *= sr * sr;
        fac_rescale *= v_mx;
        v_mi *= sr;
        v_mx *= sr;
    }
    info->o_l2norm = (float)o_l2norm;
    info->fac_rescale = (float)fac_rescale;
    info->fac_error = (float)fac_error;
    info->ip_o_oa = ip;
    info->oa_l2sqr = oa_l2sqr;
    info->delta = delta;
    info->v_mi = v_mi;
    info->ip_cent_oa = 0;
    if (bits >= 1 && centroid) {  // get_oa() then centroid->dot(oa), :42-45, :241
        std::vector<float> oa(D);
        const float fd = (float)delta, fm = (float)v_mi;
        for (int64_t j = 0; j < D; ++j) oa[j] = (float)code[j] * fd + fm;
        info->ip_cent_oa = orc_eigen_dot(centroid, oa.data(), D);
    }
    return info->status;
}

void orc_rotate_rows(const float *X, int64_t n, int64_t D, const float *P, const float *c_rot, float *O) {
    // O = X*P - c_rot (quantizer.hpp:45-48), each row computed in the GEMV order above.
    std::vector<float> t(D);
    for (int64_t i = 0; i < n; ++i) {
        if (P)
            orc_gemv_rowvec(X + i * D, P, D, t.data());
        else
            std::memcpy(t.data(), X + i * D, D * sizeof(float));
        for (int64_t j = 0; j < D; ++j) O[i * D + j] = t[j] - c_rot[j];
    }
}

void orc_cluster_sizes(const orc_plan_t *plan, int64_t n, orc_cluster_sizes_t *out) {
    const Layout L = make_layout(plan, n);
    out->num_blocks = L.num_blocks;
    out->short_factor_floats = L.single ? 0 : L.fac_per_block * L.num_blocks;
    out->short_code_bytes = L.code_per_block * L.num_blocks;
    out->long_code_bytes = L.long_total;
    out->long_stride = L.long_stride;
    out->long_factor_floats = n * L.nseg * 2;
    int64_t c = 0;
    for (int s = 0; s < L.nseg; ++s) c += plan->seg_dim[s];
    out->centroid_floats = c;
}

int orc_quantize_segment(const orc_plan_t *plan, int seg, const float *X, int64_t n, const float *centroid,
                         const float *P, const orc_caq_cfg_t *cfg, orc_cluster_out_t *clu, float *O_out) {
    // QuantizerCluster::quantize (quantizer.hpp:40-73) + ClusterPacker (cluster_packer.hpp:59-139).
    // X: n x D_s rows already sliced / zero-padded for this segment (saq_quantizer.hpp:40-58).
    // Output arrays must be zero-initialised by the caller (padded lanes stay zero).
    const Layout L = make_layout(plan, n);
    const int64_t D = plan->seg_dim[seg];
    const int bits = plan->seg_bits[seg];
    float *crot = clu->centroid_rot + L.cen_off[seg];
    if (P)
        orc_gemv_rowvec(centroid, P, D, crot);
    else
        std::memcpy(crot, centroid, D * sizeof(float));
    std::vector<float> O(n * D);
    orc_rotate_rows(X, n, D, P, crot, O.data());
    if (O_out) std::memcpy(O_out, O.data(), n * D * sizeof(float));

    std::vector<uint8_t> shortc(L.num_blocks * kBlock * (D / 8), 0);
    std::vector<float> fac_norm(L.num_blocks * kBlock, 0.0f), fac_ipc(L.num_blocks * kBlock, 0.0f);
    std::vector<uint16_t> code(D), lowc(D);
    const int short_bit = bits ? 1 << (bits - 1) : 0;
    int status = 0;
    for (int64_t i = 0; i < n; ++i) {
        orc_code_info_t info;
        status |= orc_caq_encode(&O[i * D], D, bits, cfg, centroid, code.data(), &info);
        fac_norm[i] = info.o_l2norm;
        if (bits == 0) continue;
        fac_ipc[i] = info.ip_cent_oa;
        uint8_t *sc = &shortc[i * (D / 8)];  // pack_short_codes, cluster_packer.hpp:124-139
        for (int64_t j = 0; j < D / 8; ++j) {
            uint8_t byte = 0;
            for (int t = 0; t < 8; ++t) byte |= (code[8 * j + t] & short_bit) ? (uint8_t)(0x80 >> t) : 0;
            sc[j] = byte;
        }
        float *ef = &clu->long_factors[(i * L.nseg + seg) * 2];
        ef[0] = info.fac_rescale;
        ef[1] = info.fac_error;
        for (int64_t j = 0; j < D; ++j) lowc[j] = (uint16_t)(code[j] & (short_bit - 1));
        uint8_t *lc = L.compact ? clu->long_code + L.long_off[seg] + i * L.long_bytes[seg]
                                : clu->long_code + L.long_off[seg] + i * L.long_stride;
        orc_compact_code(bits - 1, lowc.data(), D, lc);
    }
    for (int64_t blk = 0; blk < L.num_blocks; ++blk) {  // finalize_and_store, :84-116
        float *fn;
        if (L.single)
            fn = reinterpret_cast<float *>(clu->short_code + blk * L.code_per_block);
        else
            fn = clu->short_factors + blk * L.fac_per_block + L.fac_off[seg];
        if (bits)
            orc_pack_codes(D, &shortc[blk * kBlock * (D / 8)], kBlock,
                           clu->short_code + blk * L.code_per_block + L.code_off[seg]);
        std::memcpy(fn, &fac_norm[blk * kBlock], kBlock * sizeof(float));
        std::memcpy(fn + kBlock, &fac_ipc[blk * kBlock], kBlock * sizeof(float));
    }
    return status;
}

int orc_plan_eqseg(int64_t D_pad, float avg_bits, int num_segs, int32_t *dims, int32_t *bits) {
    // SaqDataMaker::equal_segmentation, saq_data.hpp:133-145
    const int b = (int)(size_t)avg_bits;
    int64_t d = 0;
    int cnt = 0;
    for (int i = 0; i < num_segs && d < D_pad; i++) {
        const int64_t t = rd_up((D_pad - d) / (num_segs - i), 64);
        dims[cnt] = (int32_t)t;
        bits[cnt] = b;
        cnt++;
        d += t;
    }
    return cnt;
}

int orc_plan_dp(const float *var, int64_t D_pad, float avg_bits, int32_t *dims, int32_t *bits) {
    // SaqDataMaker::dynamic_programming, saq_data.hpp:147-227
    const size_t nbf = 2 * sizeof(float) * 8;
    const size_t tot_bits = (size_t)(avg_bits * (float)D_pad + (float)nbf);
    const size_t max_segs = avg_bits < 2 ? D_pad / 64 : D_pad / 64 / 2;
    const double lmt = std::numeric_limits<double>::max();
    const size_t i_end = D_pad / 64;
    const size_t NB = tot_bits + 1, NI = i_end + 1;
    std::vector<double> ff((max_segs + 1) * NI * NB, lmt);
    std::vector<size_t> fs((max_segs + 1) * NI * NB, 0);
    auto at = [&](size_t ns, size_t i, size_t b) { return (ns * NI + i) * NB + b; };
    size_t ans_ns = 0, ans_i = i_end, ans_b = tot_bits;
    ff[at(0, 0, 0)] = 0;
    for (size_t ns = 0; ns <= max_segs; ns++)
        for (size_t i = 0; i <= i_end; i++)
            for (size_t ub = 0; ub <= tot_bits; ++ub) {
                const double fcur = ff[at(ns, i, ub)];
                if (!(fcur < lmt)) continue;
                if (i == i_end) {
                    if (fcur * (1.01) < ff[at(ans_ns, ans_i, ans_b)]) {
                        ans_ns = ns;
                        ans_i = i;
                        ans_b = ub;
                    }
                    continue;
                }
                if (ns == max_segs) continue;
                double var_sum = 0;
                for (size_t j = 1; (i + j) * 64 <= (size_t)D_pad; j++) {
                    var_sum += orc_eigen_sum(var + (i + j - 1) * 64, 64);
                    for (size_t b = 1; b <= 13; ++b) {
                        const size_t bn = ub + b * j * 64 + nbf;
                        if (bn > tot_bits) break;
                        const double v = var_sum / (1 << b);
                        const size_t t = at(ns + 1, i + j, bn);
                        if (ff[t] > fcur + v) {
                            ff[t] = fcur + v;
                            fs[t] = (i << 4) + b;
                        }
                    }
                }
                const size_t t0 = at(ns + 1, i_end, ub);
                if (fcur + var_sum < ff[t0]) {
                    ff[t0] = fcur + var_sum;
                    fs[t0] = (i << 4) + 0;
                }
            }
    std::vector<std::pair<int32_t, int32_t>> plan;
    size_t ns = ans_ns, i = ans_i, B = ans_b;
    while (i > 0) {
        const size_t sec = fs[at(ns, i, B)];
        const size_t pi = sec >> 4, cb = sec & 0xf, dl = (i - pi) * 64;
        plan.emplace_back((int32_t)dl, (int32_t)cb);
        ns--;
        i = pi;
        if (cb) B -= cb * dl + nbf;
    }
    std::reverse(plan.begin(), plan.end());
    for (size_t k = 0; k < plan.size(); ++k) {
        dims[k] = plan[k].first;
        bits[k] = plan[k].second;
    }
    return (int)plan.size();
}

void orc_probe(const float *cent, int64_t K, int64_t D, const float *q, int64_t nprobe, int dist_type,
               uint32_t *probe, float *probe_dist) {
    // FlatInitializer::centroids_distances (index/initializer.hpp:52-76)
    struct Cand {
        uint32_t id;
        float distance;
    };
    std::vector<Cand> cd(K);
    for (int64_t i = 0; i < K; ++i) {
        cd[i].id = (uint32_t)i;
        cd[i].distance = dist_type == 1 ? orc_eigen_dot(q, cent + i * D, D)
                                        : orc_eigen_diff_sqnorm(cent + i * D, q, D);
    }
    if (dist_type == 1)
        std::partial_sort(cd.begin(), cd.begin() + nprobe, cd.end(),
                          [](const Cand &a, const Cand &b) { return !(a.distance < b.distance); });
    else
        std::partial_sort(cd.begin(), cd.begin() + nprobe, cd.end(),
                          [](const Cand &a, const Cand &b) { return a.distance < b.distance; });
    for (int64_t i = 0; i < nprobe; ++i) {
        probe[i] = cd[i].id;
        if (probe_dist) probe_dist[i] = cd[i].distance;
    }
}

int orc_search(const orc_plan_t *plan, const float *variance, const float *const *P, const orc_cluster_t *clusters,
               const uint32_t *probe, int64_t nprobe, const float *query, int64_t topk, float m, int dist_type,
               uint32_t *out_ids, float *out_dists, orc_metrics_t *metrics) {
    // IVF::search (index/ivf.hpp:275-311) with an explicit probe order.
    QueryState qs;
    init_query(plan, variance, P, query, m, dist_type, qs);
    Pool pool(topk, dist_type == 1);
    orc_metrics_t mt{0, 0, 0};
    for (int64_t i = 0; i < nprobe; ++i) {
        const orc_cluster_t &c = clusters[probe[i]];
        const Layout L = make_layout(plan, c.num_vec);
        if (L.num_blocks == 0) continue;
        if (plan->num_seg == 1)
            scan_cluster(L, c, qs, pool, mt);
        else
            search_cluster(L, c, qs, pool, mt);
    }
    for (int64_t i = 0; i < topk; ++i) {
        out_ids[i] = pool.ids[i];
        // ResultPool::get (pool.hpp:38): (greater ? -1 : 1) * stored distance, so an unfilled
        // IP slot reads -0.0
        if (out_dists) out_dists[i] = (pool.greater ? -1.0f : 1.0f) * pool.dist[i];
    }
    for (auto &e : qs.seg) {
        mt.fast_bitsum += e.fast_bits;
        mt.acc_bitsum += e.acc_bits;
    }
    if (metrics) *metrics = mt;
    return 0;
}

int orc_search_batch(const orc_plan_t *plan, const float *variance, const float *const *P,
                     const orc_cluster_t *clusters, const uint32_t *probes, int64_t nprobe, const float *queries,
                     int64_t nq, int64_t topk, float m, int dist_type, int nthreads, uint32_t *out_ids,
                     float *out_dists, orc_metrics_t *metrics) {
    // Queries distributed over threads like test_qps's detach_loop (src/test_qps.cpp:77-84).
    if (nthreads < 1) nthreads = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t)
        th.emplace_back([=]() {
            for (int64_t i = t; i < nq; i += nthreads)
                orc_search(plan, variance, P, clusters, probes + i * nprobe, nprobe, queries + i * plan->num_dim, topk,
                           m, dist_type, out_ids + i * topk, out_dists ? out_dists + i * topk : nullptr,
                           metrics ? metrics + i : nullptr);
        });
    for (auto &x : th) x.join();
    return 0;
}

int orc_search_batch_timed(const orc_plan_t *plan, const float *variance, const float *const *P,
                           const orc_cluster_t *clusters, const uint32_t *probes, int64_t nprobe,
                           const int32_t *npq, const float *queries, int64_t nq, int64_t topk, float m,
                           int nthreads, uint32_t *out_ids, float *out_dists, double *lat_ms) {
    // test_qps's timing loop (src/test_qps.cpp:77-84, 115-123): threads take the next query from a
    // shared counter and time each IVF::search call; lat_ms[i] = query i's wall time.  npq (may be
    // null): per-query nprobe, the first npq[i] entries of its probe row.
    if (nthreads < 1) nthreads = 1;
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t)
        th.emplace_back([&]() {
            for (int64_t i; (i = next.fetch_add(1)) < nq;) {
                const auto t0 = std::chrono::steady_clock::now();
                orc_search(plan, variance, P, clusters, probes + i * nprobe, npq ? npq[i] : nprobe,
                           queries + i * plan->num_dim, topk, m, 0, out_ids + i * topk,
                           out_dists ? out_dists + i * topk : nullptr, nullptr);
                lat_ms[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            }
        });
    for (auto &x : th) x.join();
    return 0;
}

int orc_estimate(const orc_plan_t *plan, const float *variance, const float *const *P,
                 const orc_cluster_t *clusters, const uint32_t *probe, int64_t nprobe, const float *query, float m,
                 int dist_type, uint32_t *out_ids, float *out_acc, float *out_fast, float *out_vars,
                 orc_metrics_t *metrics) {
    // IVF::estimate (index/ivf.hpp:313-362) with SaqCluEstimator (saq_estimator.hpp:110-176): every
    // vector of every probed list, no pruning.  Returns the number of vectors written.
    QueryState qs;
    init_query(plan, variance, P, query, m, dist_type, qs);
    const int S = plan->num_seg;
    int64_t w = 0;
    orc_metrics_t mt{0, 0, 0};
    std::vector<float> fast(kBlock), vars(kBlock), cd(kBlock);
    for (int64_t pi = 0; pi < nprobe; ++pi) {
        const orc_cluster_t &c = clusters[probe[pi]];
        const Layout L = make_layout(plan, c.num_vec);
        for (int s = 0; s < S; ++s) qs.seg[s].prepare(c.centroid_rot + L.cen_off[s]);
        for (int64_t v = 0; v < c.num_vec; ++v) {
            const int64_t blk = v / kBlock, j = v % kBlock;
            if (j == 0) {
                for (int l = 0; l < kBlock; ++l) fast[l] = vars[l] = 0.0f;
                for (int s = 0; s < S; ++s) {
                    qs.seg[s].fast(o_l2norm_row(L, c, s, blk), block_code(L, c, s, blk), cd.data());
                    for (int l = 0; l < kBlock; ++l) fast[l] = fast[l] + cd[l];
                }
                for (int s = 0; s < S; ++s) {
                    qs.seg[s].vars_est(o_l2norm_row(L, c, s, blk), cd.data());
                    for (int l = 0; l < kBlock; ++l) vars[l] = vars[l] + cd[l];
                }
            }
            float acc = 0;
            for (int s = 0; s < S; ++s)
                acc += qs.seg[s].accurate(o_l2norm_row(L, c, s, blk)[j], long_code(L, c, s, v), j,
                                          &c.long_factors[(v * S + s) * 2]);
            out_ids[w] = c.ids[v];
            out_acc[w] = acc;
            if (out_fast) out_fast[w] = fast[j];
            if (out_vars) out_vars[w] = vars[j];
            ++w;
        }
    }
    for (auto &e : qs.seg) {
        mt.fast_bitsum += e.fast_bits;
        mt.acc_bitsum += e.acc_bits;
    }
    if (metrics) *metrics = mt;
    return (int)w;
}

void orc_decode_cluster(const orc_plan_t *plan, const orc_cluster_t *c, int seg, int32_t *codes, float *o_l2norm,
                        float *rescale) {
    // The full b-bit codes of one segment of a list, read back from its SaqCluData arrays: the
    // MSB from the fast-scan blocks (inverse of pack_codes, fastscan.hpp:75-112, bits MSB-first per
    // byte as cluster_packer.hpp:124-139 stores them), the low b-1 bits from the CodeHelper<b-1>
    // long code; plus |o| (short factors) and the ExFactor rescale of every vector.
    const Layout L = make_layout(plan, c->num_vec);
    const int64_t D = plan->seg_dim[seg];
    const int bits = plan->seg_bits[seg];
    for (int64_t v = 0; v < c->num_vec; ++v) {
        const int64_t blk = v / kBlock, j = v % kBlock;
        o_l2norm[v] = o_l2norm_row(L, *c, seg, blk)[j];
        rescale[v] = c->long_factors[(v * L.nseg + seg) * 2];
        if (bits == 0) {
            for (int64_t d = 0; d < D; ++d) codes[v * D + d] = 0;
            continue;
        }
        const uint8_t *b = block_code(L, *c, seg, blk);
        int jpos = 0;  // the byte lane of vector j within a 16-byte half (kPerm0)
        while (kPerm0[jpos] != (j & 15)) ++jpos;
        const uint8_t *lc = long_code(L, *c, seg, v);
        for (int64_t i = 0; i < D / 8; ++i) {
            const uint8_t hi = b[32 * i + jpos], lo = b[32 * i + 16 + jpos];
            const uint8_t byte = j < 16 ? (uint8_t)(((hi & 15) << 4) | (lo & 15)) : (uint8_t)((hi & 0xF0) | (lo >> 4));
            for (int t = 0; t < 8; ++t) {
                const int64_t d = 8 * i + t;
                const int msb = (byte >> (7 - t)) & 1;
                codes[v * D + d] = (msb << (bits - 1)) | (bits > 1 ? orc_code_at(bits - 1, lc, D, d) : 0);
            }
        }
    }
}

int64_t orc_estimate_lanes(const orc_plan_t *plan, const float *variance, const float *const *P,
                           const orc_cluster_t *clusters, const uint32_t *probe, int64_t nprobe, const float *query,
                           float m, float *out_vars, float *out_fast, float *out_acc) {
    // The per-segment estimates searchCluster combines (saq_searcher.hpp:95-158), for every lane of
    // every block of the probed lists, padded lanes of a last block included: out_*[(blk * S + s) * 32
    // + lane] with blk counted over the probed lists in order.  varsEstDist and compFastDist of every
    // segment (a 0-bit segment's fast estimate is its variance estimate, caq_estimator.hpp:149-152);
    // compAccurateDist of every real vector (0 for padded lanes).  L2 only.  Returns the block count.
    QueryState qs;
    init_query(plan, variance, P, query, m, 0, qs);
    const int S = plan->num_seg;
    int64_t w = 0;
    for (int64_t pi = 0; pi < nprobe; ++pi) {
        const orc_cluster_t &c = clusters[probe[pi]];
        const Layout L = make_layout(plan, c.num_vec);
        for (int s = 0; s < S; ++s) qs.seg[s].prepare(c.centroid_rot + L.cen_off[s]);
        for (int64_t blk = 0; blk < L.num_blocks; ++blk, ++w) {
            for (int s = 0; s < S; ++s) {
                float *v = out_vars + (w * S + s) * kBlock, *f = out_fast + (w * S + s) * kBlock;
                float *a = out_acc + (w * S + s) * kBlock;
                const float *x = o_l2norm_row(L, c, s, blk);
                qs.seg[s].vars_est(x, v);
                qs.seg[s].fast(x, block_code(L, c, s, blk), f);
                for (int j = 0; j < kBlock; ++j) {
                    const int64_t idx = blk * kBlock + j;
                    a[j] = idx < c.num_vec ? qs.seg[s].accurate(x[j], long_code(L, c, s, idx), j,
                                                                 &c.long_factors[(idx * S + s) * 2])
                                           : 0.0f;
                }
            }
        }
    }
    return w;
}

// ---------------------------------------------------------------- single-vector path

void orc_single_layout(const orc_plan_t *plan, orc_single_layout_t *out) {
    // SaqSingleDataWrapper ctor + set_memory_base (quantization/single_data.hpp:180-285): the four
    // sections and each segment's place in them; bytes = calculate_memory_size (:287-320).
    std::memset(out, 0, sizeof(*out));
    const int S = plan->num_seg;
    int64_t code = 0, lng = 0;
    for (int s = 0; s < S; ++s) {
        const int64_t D = rd_up(plan->seg_dim[s], 64), b = plan->seg_bits[s];
        out->seg_code_off[s] = b > 0 ? code : -1;
        out->seg_long_off[s] = b > 1 ? lng : -1;
        if (b > 0) code += D / 8;
        if (b > 1) lng += D * (b - 1) / 8;
    }
    out->fac_off = 0;
    out->code_off = rd_up((int64_t)S * 2 * 4, 64);
    out->long_off = out->code_off + rd_up(code, 64);
    out->exfac_off = out->long_off + lng;
    out->bytes = out->exfac_off + (int64_t)S * 8;
    for (int s = 0; s < S; ++s) {
        if (out->seg_code_off[s] >= 0) out->seg_code_off[s] += out->code_off;
        if (out->seg_long_off[s] >= 0) out->seg_long_off[s] += out->long_off;
    }
}

void orc_single_transpose_bin(const uint16_t *q, int64_t D, uint64_t *tq) {
    // utils::new_transpose_bin (utils/space.hpp:266-293) with b_query = 8: the 8 bit planes of
    // each 64 dims, plane p at tq[p], dim k of the block at bit 63 - k.  The reference shifts
    // 32-bit lanes by 16 - 8, which equals a 16-bit shift because every value is below 256.
    for (int64_t b = 0; b < D / 64; ++b)
        for (int p = 0; p < 8; ++p) {
            uint64_t w = 0;
            for (int k = 0; k < 64; ++k) w |= (uint64_t)((q[64 * b + k] >> p) & 1u) << (63 - k);
            tq[8 * b + p] = w;
        }
}

float orc_single_warmup_ip(const uint64_t *x, const uint64_t *tq, float delta, float vl, int64_t D) {
    // utils::warmup_ip_x0_q (utils/space.hpp:126-226): integer sums, so any order gives the same
    // ip and popcount; then delta * ip + vl * popcount in float.
    uint64_t ip = 0, ppc = 0;
    for (int64_t b = 0; b < D / 64; ++b) {
        ppc += (uint64_t)__builtin_popcountll(x[b]);
        for (int p = 0; p < 8; ++p) ip += (uint64_t)__builtin_popcountll(x[b] & tq[8 * b + p]) << p;
    }
    return (delta * (float)ip) + (vl * (float)ppc);
}

float orc_single_mask_ip(const float *q, const uint64_t *x, int64_t D) {
    // utils::mask_ip_x0_q (utils/space.hpp:228-265): bits reversed so bit k is dim k; 16 lanes,
    // lane l takes dims 64 b + 16 m + l (m = 0..3) masked (+0.0 when clear), block by block; then
    // _mm512_reduce_add_ps.  The source adds the four masked loads of a block to the sum one after
    // another; the release build (-Ofast, CMakeLists.txt:92) reassociates that chain into
    // sum + ((m0 + m1) + (m2 + m3)), which is the order the reference executes and the one kept
    // here (checked bit for bit against the compiled reference kernel).
    float acc[16];
    for (int l = 0; l < 16; ++l) acc[l] = 0.0f;
    for (int64_t b = 0; b < D / 64; ++b) {
        const uint64_t w = x[b];
        for (int l = 0; l < 16; ++l) {
            float mv[4];
            for (int m = 0; m < 4; ++m) {
                const int k = 16 * m + l;
                mv[m] = ((w >> (63 - k)) & 1u) ? q[64 * b + k] : 0.0f;
            }
            acc[l] = acc[l] + ((mv[0] + mv[1]) + (mv[2] + mv[3]));
        }
    }
    return orc_tree16(acc);
}

int64_t orc_single_quantize(const orc_plan_t *plan, const float *const *P, const float *data, int64_t n,
                            const orc_caq_cfg_t *cfg, uint8_t *records, int64_t stride) {
    // SAQuantizerSingle::quantize (quantization/saq_quantizer.hpp:78-94) -> QuantizerSingle::quantize
    // (quantizer.hpp:105-124): per segment the row slice (zero-padded past num_dim) times P_s (a
    // row-vector GEMV), CAQEncoder::encode_and_fac without a centroid, then
    // store_quantization_result (:134-160): o_l2norm; for bits > 0 ip_cent_oa = 0, the MSB codes
    // MSB-first with the bytes of every 8 reversed (:166-188), the ExFactor and, for bits > 1, the
    // CodeHelper<bits-1> compacted low bits (:194-207).  A 1-bit segment packs the quantised
    // codes (encode_and_fac hands none over below 2 bits, caq_encoder.hpp:238-244).
    orc_single_layout_t L;
    orc_single_layout(plan, &L);
    int64_t nfail = 0;
    std::vector<float> x, o;
    std::vector<uint16_t> code, low;
    for (int64_t i = 0; i < n; ++i) {
        uint8_t *rec = records + i * stride;
        std::memset(rec, 0, (size_t)L.bytes);
        float *fac = reinterpret_cast<float *>(rec + L.fac_off);
        float *exf = reinterpret_cast<float *>(rec + L.exfac_off);
        int64_t off = 0;
        for (int s = 0; s < plan->num_seg; ++s) {
            const int64_t D = plan->seg_dim[s];
            const int bits = plan->seg_bits[s];
            x.assign(D, 0.0f);
            const int64_t copy = std::min<int64_t>(D, plan->num_dim - off);
            for (int64_t j = 0; j < copy; ++j) x[j] = data[i * plan->num_dim + off + j];
            o.resize(D);
            if (P && P[s])
                orc_gemv_rowvec(x.data(), P[s], D, o.data());
            else
                o = x;
            code.assign(D, 0);
            orc_code_info_t info;
            if (orc_caq_encode(o.data(), D, bits, cfg, nullptr, code.data(), &info)) ++nfail;
            fac[2 * s] = info.o_l2norm;
            off += D;
            if (bits == 0) continue;
            fac[2 * s + 1] = 0.0f;
            const uint16_t sb = (uint16_t)(1u << (bits - 1));
            uint8_t *sc = rec + L.seg_code_off[s];
            for (int64_t j = 0; j < D / 8; ++j) {
                uint8_t byte = 0;
                for (int t = 0; t < 8; ++t)
                    if (code[8 * j + t] & sb) byte |= (uint8_t)(0x80u >> t);
                sc[j + 7 - 2 * (j % 8)] = byte;
            }
            exf[2 * s] = info.fac_rescale;
            exf[2 * s + 1] = info.fac_error;
            if (bits > 1) {
                low.resize(D);
                for (int64_t j = 0; j < D; ++j) low[j] = code[j] & (uint16_t)(sb - 1);
                orc_compact_code(bits - 1, low.data(), D, rec + L.seg_long_off[s]);
            }
        }
    }
    return nfail;
}

void orc_single_estimate(const orc_plan_t *plan, const float *variance, const float *const *P,
                         const uint8_t *records, int64_t nrec, int64_t stride, const float *query, float m,
                         int dist_type, float *out_vars, float *out_fast, float *out_acc) {
    // SaqSingleEstimator<Any> (quantization/saq_estimator.hpp:248-294) over CaqSingleEstimator
    // (caq/caq_estimator.hpp:220-369, 450-497): the per-segment query of SaqEstimatorBase (:28-48),
    // prepare (:292-308), varsEstDist (:310-315), compFastDist (:317-340), compAccurateDist
    // (:342-368); the segment estimates are summed in segment order from 0.
    orc_single_layout_t L;
    orc_single_layout(plan, &L);
    struct Seg {
        int64_t D;
        int bits;
        float bound, q_l2sqr, q_l2norm, sum_q, delta, vlh, fconst;
        std::vector<float> qr;
        std::vector<uint64_t> tq;
    };
    const bool ip = dist_type == 1;
    std::vector<Seg> sg(plan->num_seg);
    int64_t off = 0;
    for (int s = 0; s < plan->num_seg; ++s) {
        Seg &e = sg[s];
        e.D = plan->seg_dim[s];
        e.bits = plan->seg_bits[s];
        std::vector<float> cq(e.D, 0.0f);
        const int64_t copy = std::min<int64_t>(e.D, plan->num_dim - off);
        for (int64_t i = 0; i < copy; ++i) cq[i] = query[off + i];
        e.bound = std::sqrt(orc_eigen_var_q2(variance + off, cq.data(), e.D)) * m;
        e.qr.resize(e.D);
        if (P && P[s])
            orc_gemv_rowvec(cq.data(), P[s], e.D, e.qr.data());
        else
            e.qr = cq;
        off += e.D;
        e.q_l2sqr = orc_eigen_sqnorm(e.qr.data(), e.D);
        e.q_l2norm = std::sqrt(e.q_l2sqr);
        e.sum_q = orc_eigen_sum(e.qr.data(), e.D);
        float vl = e.qr[0], vr = e.qr[0];
        for (int64_t i = 1; i < e.D; ++i) {
            vl = std::min(vl, e.qr[i]);
            vr = std::max(vr, e.qr[i]);
        }
        e.delta = (float)((double)(vr - vl) / ((1 << 8) - 0.01));
        std::vector<uint16_t> qsq(e.D);
        for (int64_t i = 0; i < e.D; ++i) {
            const float t = (e.qr[i] - vl) / e.delta;  // NaN when delta is 0: the x86 cast gives 0
            qsq[i] = (t == t) ? (uint16_t)(int32_t)t : 0;
        }
        e.tq.resize(e.D / 64 * 8);
        orc_single_transpose_bin(qsq.data(), e.D, e.tq.data());
        e.vlh = (float)((double)vl + 0.5 * (double)e.delta);
        const float one_over_sqrtD = (float)(1.0 / (double)std::sqrt((float)e.D));
        e.fconst = (4.0f / 0.8f) * one_over_sqrtD;  // 4 / est_error * one_over_sqrtD_
    }
    for (int64_t r = 0; r < nrec; ++r) {
        const uint8_t *rec = records + r * stride;
        const float *fac = reinterpret_cast<const float *>(rec + L.fac_off);
        const float *exf = reinterpret_cast<const float *>(rec + L.exfac_off);
        float rv = 0.0f, rf = 0.0f, ra = 0.0f;
        for (int s = 0; s < plan->num_seg; ++s) {
            const Seg &e = sg[s];
            const float o = fac[2 * s];
            float v;
            if (ip) {
                v = 0.0f - e.bound;  // ip_q_c_ is never set for single records
            } else {
                const float t = o * o + e.q_l2sqr - 2 * e.bound;
                v = (0.0f < t) ? t : 0.0f;  // std::max(0.0f, t)
            }
            float f = v, a;
            const float o2 = o * o;
            if (e.bits == 0) {
                a = ip ? 0.0f : o2 + e.q_l2sqr;
            } else {
                const uint64_t *x = reinterpret_cast<const uint64_t *>(rec + L.seg_code_off[s]);
                const float tmp = orc_single_warmup_ip(x, e.tq.data(), e.delta, e.vlh, e.D);
                const double cst = 0.5 * (double)e.sum_q - (double)(0.58f * e.q_l2norm);
                const float q1 = (float)((((double)tmp - cst) * (double)e.fconst) * (double)o);
                if (ip) {
                    f = (float)((double)q1 * 0.5);
                } else {
                    const float t = e.q_l2sqr + o2 - q1;
                    f = (t < 0.0f) ? 0.0f : t;  // std::max(t, 0.0f)
                }
                const float ipo1 = orc_single_mask_ip(e.qr.data(), x, e.D);
                const double ex = e.bits > 1 ? (double)orc_compute_ip(e.bits - 1, e.qr.data(), rec + L.seg_long_off[s], e.D)
                                             : 0.0;
                const double cd = 2.0 / (1 << e.bits);
                const float tt = (float)((double)ipo1 + ex * cd + (-1.0 + cd / 2) * (double)e.sum_q);
                const float ipoq = exf[2 * s] * tt;
                a = ip ? ipoq + 0.0f : o2 + e.q_l2sqr - 2 * ipoq;
            }
            rv += v;
            rf += f;
            ra += a;
        }
        if (out_vars) out_vars[r] = rv;
        if (out_fast) out_fast[r] = rf;
        if (out_acc) out_acc[r] = ra;
    }
}

}  // extern "C"
