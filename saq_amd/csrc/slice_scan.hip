// slice_scan.hip -- the plain SAQ IVF scan on gfx950, run as slices of the probe sequence.
//
// Reproduces IVF::search (saqlib/index/ivf.hpp:275-311) bit for bit.  The reference walks a
// query's probed lists one block at a time and refines a vector only when its fast estimate
// beats the current k-th distance (scanCluster / searchCluster, saq_searcher.hpp:68-214), so
// every decision depends on all earlier ones.  The k-th distance only ever decreases.  Split
// the probe sequence into slices [p0, p1) (1, 3, 12, 48, ... lists): every vector the
// reference refines inside a slice has a fast estimate below the k-th distance T reached at
// the slice's start.  Each slice runs as three kernels:
//
//   list_scan_kernel   one workgroup per (query, probed list of the slice), all independent:
//                      the list's u16 LUTs (CaqCluEstimator::prepare, caq_estimator.hpp:94-106),
//                      the fast scan of every block (one thread per vector, exact integer LUT
//                      sums, the reference's fp32 epilogue), the per-block pruning minima of
//                      searchCluster, and the superset {fast estimate < T} of the vectors the
//                      reference may refine, compacted per 8-block chunk.
//   refine_kernel      one workgroup per chunk with candidates: the extra-bit inner products
//                      (CodeHelper<b-1>::compute_ip, 16 lanes per item in the reference's lane
//                      order) from long records staged in LDS, the accurate distances
//                      (compAccurateDist, caq_estimator.hpp:190-217) and searchCluster's
//                      stage-3 running sums, written as one record per candidate.
//   replay_kernel      one wavefront per query replays the slice in the reference's order: the
//                      per-block masks of scanCluster (fixed at block start, the N2 last-block
//                      skip), the stage 1/2/3 decisions of searchCluster and ResultPool inserts
//                      (utils/pool.hpp:12-61), reading only block info and candidate records.
//                      Its pool is the next slice's T.
#include <algorithm>
#include <cfloat>

#include "common.hpp"
#include "estimators.hpp"
#include "slice_scan.hpp"

#pragma clang fp contract(off)

namespace saqgpu {

constexpr int kLT = 256;               // workgroup of the list scan and the refine
constexpr int kChunk = kLT / kBlock;   // 8 blocks (256 vectors, one per thread) per chunk

// Cross-lane moves without the LDS crossbar: readlane of a wave-uniform lane, and the DPP
// whole-wave shift (lane i <- lane i - 1, lane 0 <- fill).
__device__ __forceinline__ float rdlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ uint32_t rdlane(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }
__device__ __forceinline__ unsigned long long rdlane64(unsigned long long v, int l) {
    const uint32_t lo = rdlane((uint32_t)v, l), hi = rdlane((uint32_t)(v >> 32), l);
    return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t wave_shr1(uint32_t v, uint32_t fill) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x138 /* wave_shr:1 */, 0xF, 0xF, false);
}

// The minimum over each 32-lane half of the wave, in every lane of it: DPP quad permutations
// (lane ^ 1, lane ^ 2), the half-row and row mirrors (which pair the quads and then the half-rows
// once each holds its minimum), then lane ^ 16 by ds_swizzle.  A compare and select per step:
// the values are finite, and fminf would canonicalise every moved operand first.
__device__ __forceinline__ float min32(float v) {
    float t = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));  // [1,0,3,2]
    v = t < v ? t : v;
    t = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));  // [2,3,0,1]
    v = t < v ? t : v;
    t = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = t < v ? t : v;
    t = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x140, 0xF, 0xF, false));  // row_mirror
    v = t < v ? t : v;
    t = __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));  // lane ^ 16
    return t < v ? t : v;
}

// ---------------------------------------------------------------- list scan

struct LSmem {
    SegDesc *segs;
    int8_t *gseg, *bseq;  // group -> segment; quantised segments in order
    float *r, *sc, *cbmin, *cbmax;
    uint16_t *lut;
    float *xs, *ipq;      // [S][256]
    int *ctl;             // [0] block-info base, [1] candidate base, [4..8) wave counts
};

__device__ __host__ inline size_t ls_layout(unsigned char *base, const IndexView &ix, LSmem *mp) {
    const int S = ix.nseg;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        off = (off + 15) & ~size_t(15);
        unsigned char *p = base ? base + off : nullptr;
        off += bytes;
        return p;
    };
    LSmem m{};
    m.segs = reinterpret_cast<SegDesc *>(take((size_t)S * sizeof(SegDesc)));
    m.gseg = reinterpret_cast<int8_t *>(take(ix.ngroups));
    m.bseq = reinterpret_cast<int8_t *>(take(S));
    m.r = reinterpret_cast<float *>(take(ix.D_pad * 4));
    m.sc = reinterpret_cast<float *>(take((size_t)S * SC_N * 4));
    m.cbmin = reinterpret_cast<float *>(take((size_t)ix.ngroups * 16 * 4));
    m.cbmax = reinterpret_cast<float *>(take((size_t)ix.ngroups * 16 * 4));
    // code groups rounded up to 8, the extra groups' tables zero (the scan sums 8 groups at a time)
    m.lut = reinterpret_cast<uint16_t *>(take((size_t)((ix.ngroups + 7) & ~7) * 256 * 2));
    m.xs = reinterpret_cast<float *>(take((size_t)S * kLT * 4));
    m.ipq = reinterpret_cast<float *>(take((size_t)S * kLT * 4));
    m.ctl = reinterpret_cast<int *>(take(16 * 4));
    if (mp) *mp = m;
    return off + 16;
}

size_t list_scan_smem_bytes(const IndexView &ix) { return ls_layout(nullptr, ix, nullptr); }

// One segment of searchCluster's stage-3 running sum (saq_searcher.hpp:149-156): the accurate
// distance av of segment s (compAccurateDist, caq_estimator.hpp:190-217) replaces its fast
// estimate cd.  x, ipq, ex, resc: |o|, the fast-scan inner product, the extra-bit inner product
// (0 below 2 bits) and the rescale factor; sc: the segment's scalars.
template <int DT>
__device__ __forceinline__ float stage3_step(const SegDesc &sd, const float *sc, int S, float x, float ipq, float ex,
                                             float resc, float acc) {
    float av, cd;
    if (sd.bits == 0) {
        av = DT == 1 ? sc[SC_IPQC] : x * x + sc[SC_QL2];
        cd = DT == 1 ? sc[SC_VFAC] : max0(x * x + sc[SC_VFAC]);
    } else {
        float f = (ipq - sc[SC_SUMQCB]) * sd.rescale_sqrtD;
        f = f * x;
        cd = DT == 1 ? f * 0.5f + sc[SC_IPQC] : max0(x * x + (sc[SC_QL2] - f));
        const float sqd = sd.sqd;
        // Lut::getExtIP (lut.hpp:121-126)
        const float t = (float)(((double)ipq + (double)ex * (double)sqd) + (-1.0 + (double)(sqd / 2)) * (double)sc[SC_SUMQ]);
        const float ipoq = resc * t;
        av = DT == 1 ? ipoq + sc[SC_IPQC] : x * x + sc[SC_QL2] - 2 * ipoq;
    }
    return S == 1 ? av : acc + (av - cd);
}

// The candidate's record: its accurate distance's running sum after every segment.  cur: its
// fast estimate; x(s), ipq(s), ex(s), resc(s) as for stage3_step.
template <int DT, class X, class I, class E, class R>
__device__ __forceinline__ void write_record(uint32_t *rec, uint32_t pid, int ln, float cur, const SegDesc *segs,
                                             const float *scs, int S, X x_of, I ipq_of, E ex_of, R resc_of) {
    float acc = cur;
    rec[0] = pid;
    rec[1] = (uint32_t)ln;
    rec[2] = __float_as_uint(cur);
    for (int s = 0; s < S; ++s) {
        const int bits = segs[s].bits;
        acc = stage3_step<DT>(segs[s], scs + s * SC_N, S, x_of(s), bits > 0 ? ipq_of(s) : 0.0f,
                              bits >= 2 ? ex_of(s) : 0.0f, bits > 0 ? resc_of(s) : 0.0f, acc);
        rec[3 + s] = __float_as_uint(acc);
    }
}

template <int DT>
__global__ __launch_bounds__(kLT) void list_scan_kernel(SliceArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const IndexView &ix = a.ix;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int L = a.p1 - a.p0;
    const int64_t pair = a.q0 * L + blockIdx.x;
    const int64_t q = pair / L;
    const int item = a.p0 + (int)(pair - q * L);
    const int p = item_list(item);
    const int S = ix.nseg, NG = ix.ngroups, K = a.topk;
    const int NPq = a.npq ? a.npq[q] : a.nprobe;
    const uint32_t cid = p < NPq ? a.probes[q * a.nprobe + p] : 0xFFFFFFFFu;
    const int n = cid < (uint32_t)ix.num_cen ? ix.clu_n[cid] : 0;  // bad ids: rejected on the host
    const int nb = (n + kBlock - 1) / kBlock;
    const int b0 = item_first_block(item, nb), b1 = item_end_block(item, nb);
    if (b1 <= b0) {
        if (tid == 0) a.pair_blk[pair] = kNoPair;
        return;  // uniform per workgroup
    }
    const bool single = S == 1;
    const int bst = binfo_stride(S), pst = pre_stride(S), dstr = desc_stride(S);
    LSmem m;
    ls_layout(smem, ix, &m);
    const int slot = blockIdx.x % kSlots;
    if (tid == 0) {
        const uint32_t need = (uint32_t)(b1 - b0) * (uint32_t)bst;
        const uint32_t loc = atomicAdd(&a.ctr[CTR_BINFO + slot], need);
        const bool ok = (uint64_t)loc + need <= a.binfo_cap;
        if (!ok) a.ctr[CTR_OVERFLOW] = 1;
        const uint32_t base = (uint32_t)slot * a.binfo_cap + loc;
        m.ctl[0] = ok ? (int)base : -1;
        a.pair_blk[pair] = ok ? base : kNoPair;
    }
    for (int i = tid; i < S; i += kLT) m.segs[i] = ix.segs[i];
    unsigned long long endmask = 0;
    int nbq = 0;
    for (int s = 0; s < S; ++s) {
        const SegDesc sd = ix.segs[s];
        if (sd.bits == 0) continue;
        endmask |= 1ull << (sd.group_off + sd.ngroups - 1);
        for (int g = tid; g < sd.ngroups; g += kLT) m.gseg[sd.group_off + g] = (int8_t)s;
        if (tid == 0) m.bseq[nbq] = (int8_t)s;
        ++nbq;
    }
    __syncthreads();
    const int binfo_base = m.ctl[0];
    if (binfo_base < 0) return;  // out of block-info space: the host grows it and searches again
    // the k-th distance at the start of this slice (ResultPool::distk, pool.hpp:32-34)
    float T;
    {
        const int size = a.first ? 0 : a.pool_size[q];
        const float raw = size == K ? a.pool_d[q * K + K - 1] : FLT_MAX;
        T = DT == 1 ? -raw : raw;
    }
    // zero tables for the code groups past NG up to a multiple of 8 (prepare_list's barriers order
    // these stores before the scan)
    const int NG8 = (NG + 7) & ~7;
    for (int i = tid; i < (NG8 - NG) * 128; i += kLT) reinterpret_cast<uint32_t *>(m.lut + NG * 256)[i] = 0u;
    // CaqCluEstimator::prepare of every segment (caq_estimator.hpp:94-106)
    prepare_list<DT>(m.segs, S, m.gseg, NG, a.qrot + q * ix.D_pad, ix.crot + (int64_t)cid * ix.D_pad, a.bound + q * S,
                     m.r, m.sc, m.cbmin, m.cbmax, m.lut, ix.D_pad);
    const uint8_t *blk0 = ix.blocks + ix.clu_blk[cid] * ix.block_stride;
    uint32_t *binfo = a.binfo + binfo_base - (int64_t)b0 * bst;  // indexed by the list's block
    unsigned long long *cnt64 = reinterpret_cast<unsigned long long *>(a.ctr + CTR_CAND) + slot;
    const int v = tid, bl = v >> 5, ln = v & 31;

    for (int st = b0; st < b1; st += kChunk) {
        const int nbs = min(kChunk, b1 - st);
        const bool active = bl < nbs;
        const int idx = (st + bl) * kBlock + ln;
        // ---- fast scan: exact integer LUT sums (accumulate_hacc, fastscan_highacc.hpp:66-139),
        // one segment ending at every set bit of endmask
        if (active) {
            const uint8_t *rec = blk0 + (int64_t)(st + bl) * ix.block_stride;
            const float *fac = reinterpret_cast<const float *>(rec);
            // this lane's 8-byte word of code group g at cw[32 g]
            const uint2 *cw = reinterpret_cast<const uint2 *>(rec + ix.fac_floats * 4) + ln;
            uint32_t acc = 0;
            int bk = 0;
            // the sums of code groups g0 .. g0 + 7 from their words ww
            auto sum8 = [&](const uint2 (&ww)[8], int g0) {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int g = g0 + u;
                    // byte b of xh / xl is twice the high / low nibble of code byte b: one byte
                    // extract and add per LUT address (the integer sum is order-free)
                    const uint8_t *t = reinterpret_cast<const uint8_t *>(m.lut + g * 256);
                    uint32_t xh = (ww[u].x >> 3) & 0x1E1E1E1Eu, xl = (ww[u].x << 1) & 0x1E1E1E1Eu;
                    uint32_t yh = (ww[u].y >> 3) & 0x1E1E1E1Eu, yl = (ww[u].y << 1) & 0x1E1E1E1Eu;
                    // keep the four masked words as they are (no instruction): otherwise the mask is
                    // folded back into a shift and an and per nibble
                    asm("" : "+v"(xh), "+v"(xl), "+v"(yh), "+v"(yl));
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        acc += *reinterpret_cast<const uint16_t *>(t + (2 * b) * 32 + __builtin_amdgcn_ubfe(xh, 8 * b, 8));
                        acc += *reinterpret_cast<const uint16_t *>(t + (2 * b + 1) * 32 + __builtin_amdgcn_ubfe(xl, 8 * b, 8));
                    }
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        acc += *reinterpret_cast<const uint16_t *>(t + (2 * b + 8) * 32 + __builtin_amdgcn_ubfe(yh, 8 * b, 8));
                        acc += *reinterpret_cast<const uint16_t *>(t + (2 * b + 9) * 32 + __builtin_amdgcn_ubfe(yl, 8 * b, 8));
                    }
                    if ((endmask >> g) & 1ull) {
                        const int s = m.bseq[bk++];
                        const float *sc = m.sc + s * SC_N;
                        float t2 = (float)(int32_t)acc * sc[SC_DELTA];  // lut.hpp:108-109
                        t2 = t2 + sc[SC_SUMVL];
                        m.ipq[s * kLT + v] = t2;
                        acc = 0;
                    }
                }
            };
            // the code words of 8 groups at a time in two register sets, the next 8 loading while
            // these are summed.  Groups past NG (up to NG8 + 7) read the next records or the array's
            // tail padding (kBlockTailPad): their tables are zero, or they are never summed.
            uint2 wa[8], wb[8];
            const uint2 *cp = cw;
#pragma unroll
            for (int u = 0; u < 8; ++u) wa[u] = cp[32 * u];
            {  // |o| of every segment, loaded after the first code words: one wait covers both
                float xv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) xv[u] = u < S ? fac[u * kBlock + ln] : 0.0f;
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (u < S) m.xs[u * kLT + v] = xv[u];
                for (int s = 8; s < S; ++s) m.xs[s * kLT + v] = fac[s * kBlock + ln];
            }
            for (int g0 = 0; g0 < NG8; g0 += 16) {
                cp += 256;
#pragma unroll
                for (int u = 0; u < 8; ++u) wb[u] = cp[32 * u];
                sum8(wa, g0);
                if (g0 + 8 >= NG8) break;
                cp += 256;
#pragma unroll
                for (int u = 0; u < 8; ++u) wa[u] = cp[32 * u];
                sum8(wb, g0 + 8);
            }
        }
        // ---- per-vector estimates, block minima and the candidate test against T.  Lanes of
        // blocks past the chunk's end read stale |o| and sums: their block's results are never
        // stored or flagged, and the block minima never mix two blocks.
        bool flag = false;
        float cur = 0.0f;
        if (2 * wave >= nbs) {
            // both of this wave's blocks lie past the chunk's end: no estimates, no candidates
        } else if (single) {
            const float *sc = m.sc;
            const float x = m.xs[v];
            if (m.segs[0].bits == 0) {
                cur = DT == 1 ? sc[SC_VFAC] : max0(x * x + sc[SC_VFAC]);
            } else {
                const float ipq = m.ipq[v];
                float f = (ipq - sc[SC_SUMQCB]) * m.segs[0].rescale_sqrtD;
                f = f * x;
                cur = DT == 1 ? f * 0.5f + sc[SC_IPQC] : max0(x * x + (sc[SC_QL2] - f));
            }
            // scanCluster scores n % 32 vectors of the last block: none when n % 32 == 0 (N2)
            const int cnt = (st + bl == nb - 1) ? n % kBlock : kBlock;
            flag = active && ln < cnt && cur < T;
        } else {
            // stage 1: sum of the variance estimates (saq_searcher.hpp:95-110)
            for (int s = 0; s < S; ++s) {
                const float x = m.xs[s * kLT + v];
                const float vd = DT == 1 ? m.sc[s * SC_N + SC_VFAC] : max0(x * x + m.sc[s * SC_N + SC_VFAC]);
                cur = cur + vd;
            }
            const float m1 = min32(cur);
            float M = m1;
            // the block-info words 2 .. S + 3 (m1, M, the segments' minima), one per lane of the
            // block and stored together; with more than 30 segments each minimum is stored alone
            const bool wide = S > 30;
            uint32_t bw = ln == 0 ? __float_as_uint(m1) : 0u;
            // stage 2: each quantised segment's variance estimate replaced by its fast
            // estimate, in segment order (saq_searcher.hpp:112-132)
            for (int s = 0; s < S; ++s) {
                if (m.segs[s].bits == 0) continue;
                const float *sc = m.sc + s * SC_N;
                const float x = m.xs[s * kLT + v];
                const float vd = DT == 1 ? sc[SC_VFAC] : max0(x * x + sc[SC_VFAC]);
                const float ipq = m.ipq[s * kLT + v];
                float f = (ipq - sc[SC_SUMQCB]) * m.segs[s].rescale_sqrtD;
                f = f * x;
                const float fd = DT == 1 ? f * 0.5f + sc[SC_IPQC] : max0(x * x + (sc[SC_QL2] - f));
                cur = cur - vd;
                cur = cur + fd;
                const float ms = min32(cur);
                M = ms > M ? ms : M;
                if (wide) {
                    if (active && ln == 0) binfo[(st + bl) * bst + 4 + s] = __float_as_uint(ms);
                } else if (ln == 2 + s) {
                    bw = __float_as_uint(ms);
                }
            }
            if (ln == 1) bw = __float_as_uint(M);
            if (active && (wide ? ln < 2 : ln < S + 2)) binfo[(st + bl) * bst + 2 + ln] = bw;
            // a block is refined only if m1 <= distk and M <= distk, a vector only if
            // cur < distk (saq_searcher.hpp:105-158); distk <= T throughout the slice
            flag = active && idx < n && m1 <= T && M <= T && cur < T;
        }
        // ---- ordered compaction of the candidates (wave-major = block order, lane order)
        const unsigned long long bal = __ballot(flag);
        if (lane == 0) m.ctl[4 + wave] = __popcll(bal);
        __syncthreads();
        int before = 0, total = 0;
        for (int w = 0; w < 4; ++w) {
            const int c = m.ctl[4 + w];
            if (w < wave) before += c;
            total += c;
        }
        // the chunk's candidates: `total` contiguous records and one descriptor per cmax of them
        // (a refine wavefront takes one descriptor in one round)
        const int ndesc = (total + a.cmax - 1) >> __builtin_ctz(a.cmax);  // cmax: a power of two
        if (tid == 0 && total) {
            const unsigned long long old =
                atomicAdd(cnt64, ((unsigned long long)ndesc << 32) | (unsigned long long)total);
            const uint32_t cl = (uint32_t)old, dl = (uint32_t)(old >> 32);
            const bool ok = (uint64_t)cl + total <= a.cand_cap && (uint64_t)dl + ndesc <= a.desc_cap;
            if (!ok) a.ctr[CTR_OVERFLOW] = 1;
            m.ctl[1] = ok ? (int)(slot * a.cand_cap + cl) : -1;
            m.ctl[2] = (int)dl;
        }
        __syncthreads();
        const int cbase = total ? m.ctl[1] : 0;
        if (total) {
            // every descriptor below the slot's counter is read by the refine kernel: those whose
            // records did not fit carry no candidates.  Head words, then the list's scalars (the
            // refine kernel's estimator state).
            const uint32_t dl = (uint32_t)m.ctl[2];
            const int per = kDescHead + kDescSc * S;
            // wave w writes descriptors w, w + 4, ..., a lane per word
            for (int k = wave; k < ndesc; k += 4) {
                if (dl + k >= a.desc_cap) continue;
                uint32_t *dd = a.desc + ((int64_t)slot * a.desc_cap + dl + k) * dstr;
                for (int w = lane; w < per; w += 64) {
                    uint32_t val;
                    if (w == 0) {
                        val = (uint32_t)q;
                    } else if (w == 1) {
                        val = cid;
                    } else if (w == 2) {
                        val = cbase >= 0 ? (uint32_t)(cbase + k * a.cmax) : 0u;
                    } else if (w == 3) {
                        val = cbase >= 0 ? (uint32_t)min(a.cmax, total - k * a.cmax) : 0u;
                    } else {
                        // SC_SUMQCB, SC_QL2, SC_SUMQ, SC_VFAC, SC_IPQC: consecutive scalars
                        static_assert(SC_QL2 == SC_SUMQCB + 1 && SC_SUMQ == SC_SUMQCB + 2 &&
                                      SC_VFAC == SC_SUMQCB + 3 && SC_IPQC == SC_SUMQCB + 4 && kDescSc == 5, "");
                        const int t = w - kDescHead, s = t / kDescSc, j = t - s * kDescSc;
                        val = __float_as_uint(m.sc[s * SC_N + SC_SUMQCB + j]);
                    }
                    dd[w] = val;
                }
            }
        }
        if (active && ln == 0) {
            const uint32_t half = lane < 32 ? (uint32_t)bal : (uint32_t)(bal >> 32);
            const int boff = before + (lane < 32 ? 0 : __popc((uint32_t)bal));
            binfo[(st + bl) * bst + 0] = (uint32_t)(cbase + boff);
            binfo[(st + bl) * bst + 1] = cbase < 0 ? 0u : (uint32_t)__popc(half);
        }
        if (flag && cbase >= 0) {
            const int rank = before + __popcll(bal & ((1ull << lane) - 1));
            uint32_t *pr = a.pre + (int64_t)(cbase + rank) * pst;
            pr[0] = (uint32_t)(st * kBlock + v);
            pr[1] = __float_as_uint(cur);
            for (int s = 0; s < S; ++s) pr[2 + s] = __float_as_uint(m.segs[s].bits ? m.ipq[s * kLT + v] : 0.0f);
        }
        __syncthreads();  // xs / ipq / ctl are rewritten by the next chunk
    }
}

// ---------------------------------------------------------------- refine

// Refine of the chunks the list scan handed over: one wavefront per chunk descriptor, each
// with its own LDS, no workgroup barriers.  Every global read is an LDS DMA issued in a batch
// (the query and centroid rows, the round's pre-records, then the long records and the
// per-candidate PIDs, |o| and rescale factors), so a round waits for memory twice.
struct RWmem {
    float *r, *cbuf, *sc;
    uint8_t *lbuf;        // cmax long records
    float *exv;           // [cmax][S]
    uint32_t *pre;        // [cmax][pre_stride]
    float *side;          // [cmax][2 S + 1]: PID, per segment |o|, per segment rescale
    int *ord;             // [cmax] the candidates still refined, in order
};

constexpr int kRW = 4;  // wavefronts of a refine workgroup

__device__ __host__ inline size_t rw_layout(unsigned char *base, const IndexView &ix, int cmax, RWmem *mp) {
    const int S = ix.nseg;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        off = (off + 15) & ~size_t(15);
        unsigned char *p = base ? base + off : nullptr;
        off += bytes;
        return p;
    };
    RWmem m{};
    // the long records; before them the buffer holds the centroid row
    m.lbuf = take(std::max<size_t>((size_t)cmax * ix.long_stride, (size_t)ix.D_pad * 4));
    m.cbuf = reinterpret_cast<float *>(m.lbuf);
    m.r = reinterpret_cast<float *>(take(ix.D_pad * 4));
    m.sc = reinterpret_cast<float *>(take((size_t)S * SC_N * 4));
    m.exv = reinterpret_cast<float *>(take((size_t)cmax * S * 4));
    m.pre = reinterpret_cast<uint32_t *>(take((size_t)cmax * pre_stride(S) * 4));
    m.side = reinterpret_cast<float *>(take((size_t)cmax * (2 * S + 1) * 4 + 256));
    m.ord = reinterpret_cast<int *>(take((size_t)cmax * 4));
    if (mp) *mp = m;
    return (off + 15) & ~size_t(15);
}

size_t refine_smem_bytes(const IndexView &ix, int cmax) {
    return kRW * rw_layout(nullptr, ix, cmax, nullptr) + 16 + (size_t)ix.nseg * sizeof(SegDesc) + 16;
}

// 4-byte LDS DMA of n consecutive dwords by one wavefront (lane l -> dst + l of each 64)
__device__ __forceinline__ void dma_dwords(const uint32_t *src, uint32_t *dst, int n, int lane) {
    for (int i0 = 0; i0 < n; i0 += 64)
        if (i0 + lane < n) __builtin_amdgcn_global_load_lds(src + i0 + lane, dst + i0, 4, 0, 0);
}

// Persistent over the slice's chunk descriptors (wavefront w takes w, w + all wavefronts, ...).
template <int DT>
__global__ __launch_bounds__(kRW * 64) void refine_kernel(SliceArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const IndexView &ix = a.ix;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int S = ix.nseg, CM = a.cmax;
    const int rst = cand_stride(S), pst = pre_stride(S), sst = 2 * S + 1, dstr = desc_stride(S);
    const size_t wbytes = rw_layout(nullptr, ix, CM, nullptr);
    SegDesc *segs = reinterpret_cast<SegDesc *>(smem + kRW * wbytes + 16);
    RWmem m;
    rw_layout(smem + wave * wbytes, ix, CM, &m);
    for (int i = tid; i < S; i += kRW * 64) segs[i] = ix.segs[i];
    __syncthreads();
    // descriptors of every slot, numbered slot by slot: lane s holds slot s's count and the
    // inclusive prefix
    const uint32_t cnt_s = lane < kSlots
        ? min((uint32_t)(reinterpret_cast<const unsigned long long *>(a.ctr + CTR_CAND)[lane] >> 32), a.desc_cap)
        : 0u;
    uint32_t incl = cnt_s;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    const uint32_t ndesc = rdlane(incl, 63);
    const int chunks = (int)(ix.long_stride / 16);
    const int grp = lane >> 4, gl = lane & 15;
    for (uint32_t g = blockIdx.x * kRW + wave; g < ndesc; g += gridDim.x * kRW) {
        const int slot = __popcll(__ballot(incl <= g));
        const uint32_t local = g - (rdlane(incl, slot) - rdlane(cnt_s, slot));
        const uint32_t *dd = a.desc + ((int64_t)slot * a.desc_cap + local) * dstr;
        const int64_t q = dd[0];
        const uint32_t cid = dd[1];
        const uint32_t cb = dd[2];
        const int total = (int)dd[3];
        if (total == 0) continue;  // uniform per wavefront
        const int64_t vec0 = ix.clu_vec[cid];
        const int64_t blk0 = ix.clu_blk[cid];
        // the k-th distance at the start of the slice, as the list scan took it: the replay's
        // distk never exceeds it, so once a candidate's running sum reaches it stage 3 stops at
        // or before that segment and the later segments are never read
        float T;
        {
            const int size = a.first ? 0 : a.pool_size[q];
            const float raw = size == a.topk ? a.pool_d[q * a.topk + a.topk - 1] : FLT_MAX;
            T = DT == 1 ? -raw : raw;
        }
        // the residual and the per-segment scalars of this (query, list): the values the list
        // scan derived (CaqCluEstimator::prepare, caq_estimator.hpp:94-106)
        dma_dwords(reinterpret_cast<const uint32_t *>(a.qrot + q * ix.D_pad), reinterpret_cast<uint32_t *>(m.r),
                   ix.D_pad, lane);
        dma_dwords(reinterpret_cast<const uint32_t *>(ix.crot + (int64_t)cid * ix.D_pad),
                   reinterpret_cast<uint32_t *>(m.cbuf), ix.D_pad, lane);
        dma_dwords(a.pre + (int64_t)cb * pst, m.pre, min(CM, total) * pst, lane);
        if (lane < kDescSc * S) {  // the list scan's scalars of this (query, list)
            constexpr int kSc[kDescSc] = {SC_SUMQCB, SC_QL2, SC_SUMQ, SC_VFAC, SC_IPQC};
            const int s = lane / kDescSc, j = lane - s * kDescSc;
            m.sc[s * SC_N + kSc[j]] = __uint_as_float(dd[kDescHead + lane]);
        }
        dma_wait();
        __builtin_amdgcn_wave_barrier();
        if (DT != 1)
            for (int i = lane; i < ix.D_pad; i += 64) m.r[i] = m.r[i] - m.cbuf[i];
        __builtin_amdgcn_wave_barrier();
        for (int c0 = 0; c0 < total; c0 += CM) {
            const int nc = min(CM, total - c0);
            if (c0 > 0) {  // the round's pre-records (the first round's came with the rows)
                dma_dwords(a.pre + (int64_t)(cb + c0) * pst, m.pre, nc * pst, lane);
                dma_wait();
                __builtin_amdgcn_wave_barrier();
            }
            // long records, and per candidate its PID, |o| and rescale of every segment
            const int nchunk = nc * chunks;
            for (int off = 0; off < nchunk; off += 64) {
                const int i = off + lane;
                if (i < nchunk) {
                    const int c = i / chunks, ch = i - c * chunks;
                    const int64_t gvec = vec0 + (int64_t)m.pre[c * pst];
                    __builtin_amdgcn_global_load_lds(ix.long_codes + gvec * ix.long_stride + ch * 16,
                                                     m.lbuf + (size_t)off * 16, 16, 0, 0);
                }
            }
            const int nside = nc * sst;
            for (int i0 = 0; i0 < nside; i0 += 64) {
                const int i = i0 + lane;
                if (i < nside) {
                    const int c = i / sst, j = i - c * sst;
                    const int vl = (int)m.pre[c * pst];
                    const int64_t gvec = vec0 + vl;
                    const void *src;
                    if (j == 0)
                        src = ix.ids + gvec;
                    else if (j <= S)
                        src = ix.blocks + (blk0 + (vl >> 5)) * ix.block_stride + ((j - 1) * kBlock + (vl & 31)) * 4;
                    else
                        src = ix.exfac + (gvec * S + (j - 1 - S)) * 2;
                    __builtin_amdgcn_global_load_lds(src, m.side + i0, 4, 0, 0);
                }
            }
            for (int i = lane; i < nc * S; i += 64) m.exv[i] = 0.0f;  // segments never refined read 0
            dma_wait();
            __builtin_amdgcn_wave_barrier();
            // extra-bit inner products, segment by segment (one code path per wavefront step), of
            // the candidates whose running sum (lane c) has not reached T
            unsigned long long live = __ballot(lane < nc);
            float run = lane < nc ? __uint_as_float(m.pre[lane * pst + 1]) : 0.0f;
            for (int s = 0; s < S; ++s) {
                const SegDesc &sd = segs[s];
                if (sd.bits >= 2) {
                    if ((live >> lane) & 1ull) m.ord[__popcll(live & ((1ull << lane) - 1))] = lane;
                    __builtin_amdgcn_wave_barrier();
                    const int nl = __popcll(live);
                    for (int k = grp; k < nl; k += 4) {
                        const int c = m.ord[k];
                        const float ex = ex_ip_group(sd.bits - 1, m.r + sd.doff,
                                                     m.lbuf + c * ix.long_stride + sd.long_off, sd.dim, gl);
                        if (gl == 0) m.exv[c * S + s] = ex;
                    }
                    __builtin_amdgcn_wave_barrier();
                }
                if (S > 1 && s + 1 < S) {
                    if ((live >> lane) & 1ull) {
                        const float *sd2 = m.side + lane * sst;
                        run = stage3_step<DT>(sd, m.sc + s * SC_N, S, sd2[1 + s],
                                              sd.bits > 0 ? __uint_as_float(m.pre[lane * pst + 2 + s]) : 0.0f,
                                              m.exv[lane * S + s], sd.bits > 0 ? sd2[1 + S + s] : 0.0f, run);
                    }
                    live &= ~__ballot(((live >> lane) & 1ull) && run >= T);
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (lane < nc) {
                const int c = lane;
                const uint32_t *pr = m.pre + c * pst;
                const float *sd = m.side + c * sst;
                write_record<DT>(
                    a.cand + (int64_t)(cb + c0 + c) * rst, __float_as_uint(sd[0]), (int)(pr[0] & 31),
                    __uint_as_float(pr[1]), segs, m.sc, S, [&](int s) { return sd[1 + s]; },
                    [&](int s) { return __uint_as_float(pr[2 + s]); }, [&](int s) { return m.exv[c * S + s]; },
                    [&](int s) { return sd[1 + S + s]; });
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// ---------------------------------------------------------------- replay

// The replay's ResultPool lives in registers: entry 64 r + lane is row r of lane `lane`
// (KR rows cover topk <= 64 KR).
template <int KR>
struct RegPool {
    float d[KR];
    uint32_t id[KR];
    // entry i (the same value in every lane; i wave-uniform)
    __device__ __forceinline__ float dist_at(int i) const {
        float v = 0.0f;
#pragma unroll
        for (int r = 0; r < KR; ++r)
            if (r == (i >> 6)) v = rdlane(d[r], i & 63);
        return v;
    }
};

constexpr int kMeta = 64;  // lists whose metadata the replay holds at once (one per lane)

template <int DT, int KR>
__global__ __launch_bounds__(64) void replay_kernel(SliceArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const IndexView &ix = a.ix;
    const int lane = threadIdx.x;
    const int64_t q = blockIdx.x;
    const int S = ix.nseg, K = a.topk;
    const int L = a.p1 - a.p0;
    const bool single = S == 1;
    const int bst = binfo_stride(S), rst = cand_stride(S);
    // LDS: the current chunk's candidate records (<= 256) and two block-info windows (64
    // blocks each, double-buffered)
    uint32_t *crec = reinterpret_cast<uint32_t *>(smem);
    uint32_t *winfo = crec + kLT * rst;

    int size = a.first ? 0 : a.pool_size[q];
    RegPool<KR> pool;
#pragma unroll
    for (int r = 0; r < KR; ++r) {
        const int i = r * 64 + lane;
        pool.d[r] = (i < size) ? a.pool_d[q * K + i] : 0.0f;
        pool.id[r] = (i < size) ? a.pool_id[q * K + i] : 0u;
    }
    // per segment (lane s): the prefix sums of scanned and refined bits, and which segments
    // carry codes
    unsigned long long fastcum = 0, cumbits = 0;
    unsigned long long qmask = 0;
    {
        unsigned long long fc = 0, cb = 0;
        for (int s = 0; s < S; ++s) {
            const SegDesc sd = ix.segs[s];
            if (sd.bits > 0) {
                fc += (unsigned long long)kBlock * sd.dim;
                cb += (unsigned long long)sd.dim * (sd.bits - 1);
                qmask |= 1ull << s;
            }
            if (s == lane) {
                fastcum = fc;
                cumbits = cb;
            }
        }
    }
    // running metrics: the same value in every lane
    unsigned long long fast_bits = 0, acc_bits = 0, comp = 0;
    if (!a.first) {
        fast_bits = a.met[q * 3 + 0];
        acc_bits = a.met[q * 3 + 1];
        comp = a.met[q * 3 + 2];
    }
    const int sbits0 = ix.segs[0].bits, sdim0 = ix.segs[0].dim;

    float tail = size == K ? pool.dist_at(K - 1) : FLT_MAX;  // raw (negated for IP)
    float dk = DT == 1 ? -tail : tail;
    // ResultPool::insert (pool.hpp:17-30): reject when full and d > tail; the new entry goes
    // before every entry >= d (ties: new before old), the rest move up one, the (k+1)-th drops
    auto insert = [&](uint32_t u, float d) {
        if (DT == 1) d = -d;
        if (size == K && d > tail) return;
        int lo = 0;
#pragma unroll
        for (int r = 0; r < KR; ++r) lo += __popcll(__ballot(r * 64 + lane < size && pool.d[r] < d));
        uint32_t carry_d = 0, carry_i = 0;
#pragma unroll
        for (int r = 0; r < KR; ++r) {
            const int i = r * 64 + lane;
            const uint32_t db = __float_as_uint(pool.d[r]);
            const uint32_t up_d = wave_shr1(db, carry_d);
            const uint32_t up_i = wave_shr1(pool.id[r], carry_i);
            carry_d = rdlane(db, 63);
            carry_i = rdlane(pool.id[r], 63);
            if (i > lo) {
                pool.d[r] = __uint_as_float(up_d);
                pool.id[r] = up_i;
            } else if (i == lo) {
                pool.d[r] = d;
                pool.id[r] = u;
            }
        }
        size += (size < K);
        tail = size == K ? pool.dist_at(K - 1) : FLT_MAX;
        dk = DT == 1 ? -tail : tail;
    };

    const int NPq = a.npq ? a.npq[q] : a.nprobe;
    const int pend = min(a.p1, NPq + 1);  // items of the first NPq lists
    // items [mbase, mbase + 64): lane j holds item mbase + j's block-info base, its blocks and
    // its list's blocks (0 for item 1, which continues list 0)
    int mbase = a.p0;
    uint32_t m_base = kNoPair;
    int m_nb = 0, m_lnb = 0;
    auto load_meta = [&](int pb) {
        mbase = pb;
        const int it = pb + lane;
        m_base = kNoPair;
        m_nb = m_lnb = 0;
        if (it < pend) {
            m_base = a.pair_blk[q * L + (it - a.p0)];
            if (m_base != kNoPair) {
                const int nb = (ix.clu_n[a.probes[q * a.nprobe + item_list(it)]] + kBlock - 1) / kBlock;
                m_nb = item_end_block(it, nb) - item_first_block(it, nb);
                m_lnb = it == 1 ? 0 : nb;
            }
        }
    };
    // block info of the window (item p, blocks w0..w0+63) -> LDS, by LDS DMA (4 bytes a lane)
    auto fetch_window = [&](int p, int w0, uint32_t *dst) {
        const uint32_t base = rdlane(m_base, p - mbase);
        const int nb = (int)rdlane((uint32_t)m_nb, p - mbase);
        const int ndw = min(64, nb - w0) * bst;
        const uint32_t *src = a.binfo + base + (int64_t)w0 * bst;
        for (int i0 = 0; i0 < ndw; i0 += 64)
            if (i0 + lane < ndw) __builtin_amdgcn_global_load_lds(src + i0 + lane, dst + i0, 4, 0, 0);
    };
    // first window at or after (p, w0) of a non-empty item of the slice (p == pend: none)
    auto seek_window = [&](int &p, int &w0) {
        for (;;) {
            if (p >= pend) return;
            if (p - mbase >= kMeta) load_meta(p);
            const int nb = (int)rdlane((uint32_t)m_nb, p - mbase);
            if (rdlane(m_base, p - mbase) != kNoPair && w0 < nb) return;
            ++p;
            w0 = 0;
        }
    };
    // The candidate records of a chunk (8 blocks, written together by one list-scan step) are
    // contiguous: one LDS DMA brings them all.  Returns the LDS record index of block b's
    // first candidate.  off / csum: the lanes' block info (csum: its chunk's total).
    int loaded = -1;
    uint32_t loaded_base = 0;
    auto block_records = [&](int b, uint32_t off, uint32_t csum) -> int {
        const int ch = b >> 3;
        const uint32_t cb = rdlane(off, ch * 8);
        if (ch != loaded) {
            const int ndw = (int)rdlane(csum, ch * 8) * rst;
            const uint32_t *src = a.cand + (int64_t)cb * rst;
            for (int i0 = 0; i0 < ndw; i0 += 64)
                if (i0 + lane < ndw) __builtin_amdgcn_global_load_lds(src + i0 + lane, crec + i0, 4, 0, 0);
            dma_wait();
            __builtin_amdgcn_wave_barrier();
            loaded = ch;
            loaded_base = cb;
        }
        return (int)(rdlane(off, b) - loaded_base);
    };

    load_meta(a.p0);
    int cp = a.p0, cw = 0;
    seek_window(cp, cw);
    int buf = 0;
    if (cp < pend) fetch_window(cp, cw, winfo);
    dma_wait();
    __builtin_amdgcn_wave_barrier();
    unsigned long long fb_lane = 0;  // fast bits summed per lane, reduced at the end
    while (cp < pend) {
        const int nb = (int)rdlane((uint32_t)m_nb, cp - mbase);
        const uint32_t *wi = winfo + buf * 64 * bst;
        if (cw == 0) {
            // once per list: total_comp_cnt, and scanCluster's scanned bits
            const int lnb = (int)rdlane((uint32_t)m_lnb, cp - mbase);
            comp += (unsigned long long)lnb * kBlock;
            if (single && sbits0 > 0) fast_bits += (unsigned long long)lnb * kBlock * sdim0;
        }
        // the next window's block info streams in while this one is replayed
        int np = cp, nw = cw + 64;
        seek_window(np, nw);
        if (np < pend) fetch_window(np, nw, winfo + (buf ^ 1) * 64 * bst);
        const int blk = cw + lane;
        const bool valid = blk < nb;
        uint32_t cnt = 0, off = 0;
        if (valid) {
            off = wi[lane * bst + 0];
            cnt = wi[lane * bst + 1];
        }
        uint32_t csum = cnt;
        for (int o = 1; o < 8; o <<= 1) csum += __shfl_xor(csum, o, 8);
        loaded = -1;
        if (single) {
            unsigned long long hb = __ballot(cnt > 0);
            while (hb) {
                const int b = __builtin_ctzll(hb);
                hb &= hb - 1;
                const int bc = (int)rdlane(cnt, b);
                const uint32_t *br = crec + block_records(b, off, csum) * rst;
                float cv = FLT_MAX, av = 0.0f;
                uint32_t cid = 0;
                if (lane < bc) {
                    cid = br[lane * rst + 0];
                    cv = __uint_as_float(br[lane * rst + 2]);
                    av = __uint_as_float(br[lane * rst + 3]);
                }
                // scanCluster: the block's mask is fixed at its start (saq_searcher.hpp:190-195)
                unsigned long long msk = __ballot(lane < bc && cv < dk);
                if (sbits0 > 0) acc_bits += (unsigned long long)__popcll(msk) * sdim0 * (sbits0 - 1);
                while (msk) {
                    const int j = __builtin_ctzll(msk);
                    msk &= msk - 1;
                    insert(rdlane(cid, j), rdlane(av, j));
                }
            }
        } else {
            const float m1 = valid ? __uint_as_float(wi[lane * bst + 2]) : FLT_MAX;
            const float M = valid ? __uint_as_float(wi[lane * bst + 3]) : FLT_MAX;
            int cur = 0;
            for (;;) {
                // blocks cur.. see the same distk until the next insert
                const bool pass1 = valid && lane >= cur && m1 <= dk;
                const unsigned long long cbm = __ballot(pass1 && M <= dk && cnt > 0);
                const int stop = cbm ? __builtin_ctzll(cbm) : 64;
                // stage 2 stops at the first quantised segment whose block minimum exceeds
                // distk (saq_searcher.hpp:112-132); the shuffle runs in every lane (a disabled
                // source lane would read as 0)
                const bool counted = pass1 && lane <= stop;
                int sb = S - 1;
                if (counted) {
                    for (int s = 0; s < S; ++s) {
                        if (((qmask >> s) & 1ull) && __uint_as_float(wi[lane * bst + 4 + s]) > dk) {
                            sb = s;
                            break;
                        }
                    }
                }
                const unsigned long long fc = __shfl(fastcum, sb);
                if (counted) fb_lane += fc;
                if (!cbm) break;
                const int bc = (int)rdlane(cnt, stop);
                const uint32_t *br = crec + block_records(stop, off, csum) * rst;
                float cv = FLT_MAX;
                uint32_t cid = 0;
                if (lane < bc) {
                    cid = br[lane * rst + 0];
                    cv = __uint_as_float(br[lane * rst + 2]);
                }
                // stage 3 (saq_searcher.hpp:134-158): distk re-read after every insert
                unsigned long long rem = bc >= 64 ? ~0ull : ((1ull << bc) - 1);
                for (;;) {
                    const unsigned long long hit = __ballot(lane < bc && cv < dk) & rem;
                    if (!hit) break;
                    const int j = __builtin_ctzll(hit);
                    rem &= ~((2ull << j) - 1);
                    const float pv = lane < S ? __uint_as_float(br[j * rst + 3 + lane]) : 0.0f;
                    const unsigned long long ge = __ballot(lane < S && pv >= dk);
                    const int sb = ge ? __builtin_ctzll(ge) : S - 1;
                    acc_bits += rdlane64(cumbits, sb);
                    insert(rdlane(cid, j), rdlane(pv, sb));
                }
                cur = stop + 1;
            }
        }
        dma_wait();  // the next window's block info
        __builtin_amdgcn_wave_barrier();
        cp = np;
        cw = nw;
        buf ^= 1;
    }
    // per-lane fast bits to the total
    for (int o = 32; o >= 1; o >>= 1) fb_lane += __shfl_xor(fb_lane, o);
    fast_bits += fb_lane;
    if (a.last) {
#pragma unroll
        for (int r = 0; r < KR; ++r) {
            const int i = r * 64 + lane;
            if (i >= K) continue;
            // unfilled entries: id 0, distance (IP: -1 *) 0 as ResultPool::get returns them
            a.out_ids[q * K + i] = i < size ? pool.id[r] : 0u;
            if (a.out_dists) {
                const float dv = i < size ? pool.d[r] : 0.0f;
                a.out_dists[q * K + i] = DT == 1 ? -dv : dv;
            }
        }
        if (a.out_metrics && lane == 0) {
            a.out_metrics[q * 3 + 0] = fast_bits;
            a.out_metrics[q * 3 + 1] = acc_bits;
            a.out_metrics[q * 3 + 2] = comp;
        }
    } else {
#pragma unroll
        for (int r = 0; r < KR; ++r) {
            const int i = r * 64 + lane;
            if (i >= K) continue;
            a.pool_d[q * K + i] = pool.d[r];
            a.pool_id[q * K + i] = pool.id[r];
        }
        if (lane == 0) {
            a.pool_size[q] = size;
            a.met[q * 3 + 0] = fast_bits;
            a.met[q * 3 + 1] = acc_bits;
            a.met[q * 3 + 2] = comp;
        }
    }
}

size_t replay_smem_bytes(const IndexView &ix) {
    const int bst = binfo_stride(ix.nseg), rst = cand_stride(ix.nseg);
    return (size_t)(kLT * rst + 2 * 64 * bst) * 4 + 16;
}

// Counters of a slice: block-info and candidate space restart at zero; the most any slot of
// any slice asked for is kept (the host sizes the buffers from it).  all: a new search (also
// clears the overflow flag).  One wavefront, lane s = slot s.
__global__ void slice_reset_kernel(uint32_t *ctr, int all) {
    const int s = threadIdx.x;
    unsigned long long *c64 = reinterpret_cast<unsigned long long *>(ctr + CTR_CAND);
    uint32_t b = 0, c = 0, d = 0;
    if (s < kSlots) {
        b = ctr[CTR_BINFO + s];
        c = (uint32_t)c64[s];
        d = (uint32_t)(c64[s] >> 32);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        b = max(b, (uint32_t)__shfl_xor(b, o));
        c = max(c, (uint32_t)__shfl_xor(c, o));
        d = max(d, (uint32_t)__shfl_xor(d, o));
    }
    __syncthreads();
    if (s == 0) {
        if (all) {
            ctr[CTR_OVERFLOW] = ctr[CTR_MAX_BINFO] = ctr[CTR_MAX_CAND] = ctr[CTR_MAX_DESC] = 0;
        } else {
            ctr[CTR_MAX_BINFO] = max(ctr[CTR_MAX_BINFO], b);
            ctr[CTR_MAX_CAND] = max(ctr[CTR_MAX_CAND], c);
            ctr[CTR_MAX_DESC] = max(ctr[CTR_MAX_DESC], d);
        }
    }
    if (s < kSlots) {
        ctr[CTR_BINFO + s] = 0;
        c64[s] = 0;
    }
}

// ---------------------------------------------------------------- launchers

hipError_t launch_slice_reset(uint32_t *ctr, int all, hipStream_t st) {
    hipLaunchKernelGGL(slice_reset_kernel, dim3(1), dim3(64), 0, st, ctr, all);
    return hipGetLastError();
}

hipError_t launch_list_scan(const SliceArgs &args, int dist_type, size_t smem, hipStream_t st) {
    const void *fn = dist_type == 1 ? (const void *)list_scan_kernel<1> : (const void *)list_scan_kernel<0>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    const int L = args.p1 - args.p0;
    // one workgroup per (query, list); launches split by query keep grid * 256 threads in range
    const int64_t qstep = std::max<int64_t>(1, ((int64_t)1 << 22) / L);
    for (int64_t q0 = 0; q0 < args.nq; q0 += qstep) {
        SliceArgs a = args;
        a.q0 = q0;
        const int64_t nqc = std::min<int64_t>(qstep, args.nq - q0);
        if (dist_type == 1)
            hipLaunchKernelGGL(list_scan_kernel<1>, dim3((unsigned)(nqc * L)), dim3(kLT), smem, st, a);
        else
            hipLaunchKernelGGL(list_scan_kernel<0>, dim3((unsigned)(nqc * L)), dim3(kLT), smem, st, a);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_refine(const SliceArgs &a, int dist_type, size_t smem, int grid, hipStream_t st) {
    const void *fn = dist_type == 1 ? (const void *)refine_kernel<1> : (const void *)refine_kernel<0>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    if (dist_type == 1)
        hipLaunchKernelGGL(refine_kernel<1>, dim3(grid), dim3(kRW * 64), smem, st, a);
    else
        hipLaunchKernelGGL(refine_kernel<0>, dim3(grid), dim3(kRW * 64), smem, st, a);
    return hipGetLastError();
}

template <int DT>
static void launch_replay_dt(const SliceArgs &a, size_t smem, hipStream_t st) {
    const dim3 g((unsigned)a.nq), b(64);
    if (a.topk <= 64)
        hipLaunchKernelGGL((replay_kernel<DT, 1>), g, b, smem, st, a);
    else if (a.topk <= 128)
        hipLaunchKernelGGL((replay_kernel<DT, 2>), g, b, smem, st, a);
    else if (a.topk <= 256)
        hipLaunchKernelGGL((replay_kernel<DT, 4>), g, b, smem, st, a);
    else
        hipLaunchKernelGGL((replay_kernel<DT, 16>), g, b, smem, st, a);
}

hipError_t launch_replay(const SliceArgs &a, int dist_type, hipStream_t st) {
    if (a.topk > 1024) return hipErrorInvalidValue;
    const size_t smem = replay_smem_bytes(a.ix);
    if (dist_type == 1)
        launch_replay_dt<1>(a, smem, st);
    else
        launch_replay_dt<0>(a, smem, st);
    return hipGetLastError();
}

}  // namespace saqgpu
