// slice_scan.hip -- the plain SAQ IVF scan on gfx950, run as slices of the probe sequence.
//
// Reproduces IVF::search (saqlib/index/ivf.hpp:275-311) bit for bit.  The reference walks a
// query's probed lists one block at a time and refines a vector only when its fast estimate
// beats the current k-th distance (scanCluster / searchCluster, saq_searcher.hpp:68-214), so
// every decision depends on all earlier ones.  The k-th distance only ever decreases.  Split
// the probe sequence into slices [p0, p1) (1, 3, 12, 48, ... lists): every vector the
// reference refines inside a slice has a fast estimate below the k-th distance T reached at
// the slice's start.  Each slice runs as two kernels:
//
//   list_scan_kernel   one workgroup per (query, probed list of the slice), all independent:
//                      the list's u16 LUTs (CaqCluEstimator::prepare, caq_estimator.hpp:94-106),
//                      the fast scan of every block (one thread per vector, exact integer LUT
//                      sums, the reference's fp32 epilogue), the per-block pruning minima of
//                      searchCluster, and -- for the superset {fast estimate < T} -- the
//                      extra-bit inner products and accurate distances (compAccurateDist,
//                      caq_estimator.hpp:190-217) with searchCluster's stage-3 running sums.
//                      It writes one block-info record per block and one record per candidate.
//   replay_kernel      one wavefront per query replays the slice in the reference's order: the
//                      per-block masks of scanCluster (fixed at block start, the N2 last-block
//                      skip), the stage 1/2/3 decisions of searchCluster and ResultPool inserts
//                      (utils/pool.hpp:12-61), reading only block info and candidate records.
//                      Its pool is the next slice's T.
//
// The streaming kernel carries the bytes (short codes, factors, the superset's long codes) and
// has no sequential dependency; the replay reads a few bytes per block.
#include <algorithm>
#include <cfloat>

#include "common.hpp"
#include "estimators.hpp"
#include "slice_scan.hpp"

#pragma clang fp contract(off)

namespace saqgpu {

constexpr int kLT = 256;               // list-scan workgroup: one thread per vector of a chunk
constexpr int kChunk = kLT / kBlock;   // 8 blocks per chunk
// ---------------------------------------------------------------- list scan

struct LSmem {
    SegDesc *segs;
    int8_t *gseg, *bseq, *rseq;  // group -> segment; quantised segments; segments with extra bits
    float *r, *sc, *cbmin, *cbmax;
    uint16_t *lut;
    float *xs, *ipq, *curr;  // [S][256], [S][256], [256]
    int *candv;              // chunk-local thread of each candidate, in (block, lane) order
    float *exv;              // [cmax][S] extra-bit inner products of a round of candidates
    int *ctl;                // [0] block-info base, [1] candidate base, [2] segments with extra bits,
                             // [4..8) wave counts
};

__device__ __host__ inline size_t ls_layout(unsigned char *base, const IndexView &ix, int cmax, LSmem *mp) {
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
    m.rseq = reinterpret_cast<int8_t *>(take(S));
    m.r = reinterpret_cast<float *>(take(ix.D_pad * 4));
    m.sc = reinterpret_cast<float *>(take((size_t)S * SC_N * 4));
    m.cbmin = reinterpret_cast<float *>(take((size_t)ix.ngroups * 16 * 4));
    m.cbmax = reinterpret_cast<float *>(take((size_t)ix.ngroups * 16 * 4));
    m.lut = reinterpret_cast<uint16_t *>(take((size_t)ix.ngroups * 256 * 2));
    m.xs = reinterpret_cast<float *>(take((size_t)S * kLT * 4));
    m.ipq = reinterpret_cast<float *>(take((size_t)S * kLT * 4));
    m.curr = reinterpret_cast<float *>(take(kLT * 4));
    m.candv = reinterpret_cast<int *>(take(kLT * 4));
    m.exv = reinterpret_cast<float *>(take((size_t)cmax * S * 4));
    m.ctl = reinterpret_cast<int *>(take(16 * 4));
    if (mp) *mp = m;
    return off + 16;
}

size_t list_scan_smem_bytes(const IndexView &ix, int cmax) { return ls_layout(nullptr, ix, cmax, nullptr); }

template <int DT>
__global__ __launch_bounds__(kLT) void list_scan_kernel(SliceArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const IndexView &ix = a.ix;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int L = a.p1 - a.p0;
    const int64_t pair = a.q0 * L + blockIdx.x;
    const int64_t q = pair / L;
    const int p = a.p0 + (int)(pair - q * L);
    const int S = ix.nseg, NG = ix.ngroups, K = a.topk, CM = a.cmax;
    const int NPq = a.npq ? a.npq[q] : a.nprobe;
    const uint32_t cid = p < NPq ? a.probes[q * a.nprobe + p] : 0xFFFFFFFFu;
    const int n = cid < (uint32_t)ix.num_cen ? ix.clu_n[cid] : 0;  // bad ids: rejected on the host
    if (n == 0) {
        if (tid == 0) a.pair_blk[pair] = kNoPair;
        return;  // uniform per workgroup
    }
    const int nb = (n + kBlock - 1) / kBlock;
    const bool single = S == 1;
    const int bst = binfo_stride(S), rst = cand_stride(S);
    LSmem m;
    ls_layout(smem, ix, CM, &m);
    if (tid == 0) {
        const uint32_t need = (uint32_t)nb * (uint32_t)bst;
        const uint32_t base = atomicAdd(&a.ctr[0], need);
        const bool ok = (uint64_t)base + need <= a.binfo_cap;
        if (!ok) a.ctr[2] = 1;
        m.ctl[0] = ok ? (int)base : -1;
        a.pair_blk[pair] = ok ? base : kNoPair;
    }
    for (int i = tid; i < S; i += kLT) m.segs[i] = ix.segs[i];
    uint32_t endmask = 0;
    int nbq = 0, nrs = 0;
    for (int s = 0; s < S; ++s) {
        const SegDesc sd = ix.segs[s];
        if (sd.bits >= 2) {
            if (tid == 0) m.rseq[nrs] = (int8_t)s;
            ++nrs;
        }
        if (sd.bits == 0) continue;
        endmask |= 1u << (sd.group_off + sd.ngroups - 1);
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
    // CaqCluEstimator::prepare of every segment (caq_estimator.hpp:94-106)
    prepare_list<DT>(m.segs, S, m.gseg, NG, a.qrot + q * ix.D_pad, ix.crot + (int64_t)cid * ix.D_pad, a.bound + q * S,
                     m.r, m.sc, m.cbmin, m.cbmax, m.lut, ix.D_pad);
    const int64_t vec0 = ix.clu_vec[cid];
    const uint8_t *blk0 = ix.blocks + ix.clu_blk[cid] * ix.block_stride;
    uint32_t *binfo = a.binfo + binfo_base;
    const int v = tid, bl = v >> 5, ln = v & 31;

    for (int st = 0; st < nb; st += kChunk) {
        const int nbs = min(kChunk, nb - st);
        const bool active = bl < nbs;
        const int idx = (st + bl) * kBlock + ln;
        // ---- fast scan: exact integer LUT sums (accumulate_hacc, fastscan_highacc.hpp:66-139),
        // one segment ending at every set bit of endmask
        if (active) {
            const uint8_t *rec = blk0 + (int64_t)(st + bl) * ix.block_stride;
            const float *fac = reinterpret_cast<const float *>(rec);
            for (int s = 0; s < S; ++s) m.xs[s * kLT + v] = fac[s * kBlock + ln];
            const uint8_t *codes = rec + ix.fac_floats * 4 + ln * 8;
            uint32_t acc = 0;
            int bk = 0;
            for (int g0 = 0; g0 < NG; g0 += 8) {
                uint2 w[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (g0 + u < NG) w[u] = *reinterpret_cast<const uint2 *>(codes + (g0 + u) * 256);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int g = g0 + u;
                    if (g >= NG) break;
                    const uint16_t *t = m.lut + g * 256;
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const uint32_t byte = (w[u].x >> (8 * b)) & 0xFF;
                        acc += t[(2 * b) * 16 + (byte >> 4)];
                        acc += t[(2 * b + 1) * 16 + (byte & 15)];
                    }
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const uint32_t byte = (w[u].y >> (8 * b)) & 0xFF;
                        acc += t[(2 * b + 8) * 16 + (byte >> 4)];
                        acc += t[(2 * b + 9) * 16 + (byte & 15)];
                    }
                    if ((endmask >> g) & 1u) {
                        const int s = m.bseq[bk++];
                        const float *sc = m.sc + s * SC_N;
                        float t2 = (float)(int32_t)acc * sc[SC_DELTA];  // lut.hpp:495-497
                        t2 = t2 + sc[SC_SUMVL];
                        m.ipq[s * kLT + v] = t2;
                        acc = 0;
                    }
                }
            }
        }
        // ---- per-vector estimates, block minima and the candidate test against T
        bool flag = false;
        {
            float cur = 0.0f;
            if (single) {
                const float *sc = m.sc;
                const float x = active ? m.xs[v] : 0.0f;
                if (m.segs[0].bits == 0) {
                    cur = DT == 1 ? sc[SC_VFAC] : max0(x * x + sc[SC_VFAC]);
                } else {
                    const float ipq = active ? m.ipq[v] : 0.0f;
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
                    const float x = active ? m.xs[s * kLT + v] : 0.0f;
                    const float vd = DT == 1 ? m.sc[s * SC_N + SC_VFAC] : max0(x * x + m.sc[s * SC_N + SC_VFAC]);
                    cur = cur + vd;
                }
                float m1 = cur;
                for (int o = 16; o >= 1; o >>= 1) m1 = fminf(m1, __shfl_xor(m1, o, 32));
                float M = m1;
                // stage 2: each quantised segment's variance estimate replaced by its fast
                // estimate, in segment order (saq_searcher.hpp:112-132)
                for (int s = 0; s < S; ++s) {
                    if (m.segs[s].bits == 0) continue;
                    const float *sc = m.sc + s * SC_N;
                    const float x = active ? m.xs[s * kLT + v] : 0.0f;
                    const float vd = DT == 1 ? sc[SC_VFAC] : max0(x * x + sc[SC_VFAC]);
                    const float ipq = active ? m.ipq[s * kLT + v] : 0.0f;
                    float f = (ipq - sc[SC_SUMQCB]) * m.segs[s].rescale_sqrtD;
                    f = f * x;
                    const float fd = DT == 1 ? f * 0.5f + sc[SC_IPQC] : max0(x * x + (sc[SC_QL2] - f));
                    cur = cur - vd;
                    cur = cur + fd;
                    float ms = cur;
                    for (int o = 16; o >= 1; o >>= 1) ms = fminf(ms, __shfl_xor(ms, o, 32));
                    M = fmaxf(M, ms);
                    if (active && ln == 0) binfo[(st + bl) * bst + 4 + s] = __float_as_uint(ms);
                }
                if (active && ln == 0) {
                    binfo[(st + bl) * bst + 2] = __float_as_uint(m1);
                    binfo[(st + bl) * bst + 3] = __float_as_uint(M);
                }
                // a block is refined only if m1 <= distk and M <= distk, a vector only if
                // cur < distk (saq_searcher.hpp:105-158); distk <= T throughout the slice
                flag = active && idx < n && m1 <= T && M <= T && cur < T;
            }
            if (active) m.curr[v] = cur;
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
        const int rank = before + __popcll(bal & ((1ull << lane) - 1));
        if (flag) m.candv[rank] = v;
        if (tid == 0 && total) {
            const uint32_t cb = atomicAdd(&a.ctr[1], (uint32_t)total);
            const bool ok = (uint64_t)cb + total <= a.cand_cap;
            if (!ok) a.ctr[2] = 1;
            m.ctl[1] = ok ? (int)cb : -1;
        }
        __syncthreads();
        const int cbase = total ? m.ctl[1] : 0;
        if (active && ln == 0) {
            const uint32_t half = lane < 32 ? (uint32_t)bal : (uint32_t)(bal >> 32);
            const int boff = before + (lane < 32 ? 0 : __popc((uint32_t)bal));
            binfo[(st + bl) * bst + 0] = (uint32_t)(cbase + boff);
            binfo[(st + bl) * bst + 1] = cbase < 0 ? 0u : (uint32_t)__popc(half);
        }
        // ---- refine the candidates, up to cmax at a time: one thread per (candidate, segment
        // with extra bits) computes the extra-bit inner product from the long code in HBM
        if (cbase >= 0) {
            for (int c0 = 0; c0 < total; c0 += CM) {
                const int nc = min(CM, total - c0);
                for (int item = tid; item < nc * nrs; item += kLT) {
                    const int c = item / nrs, s = m.rseq[item - c * nrs];
                    const SegDesc &sd = m.segs[s];
                    const int64_t gvec = vec0 + (int64_t)(st * kBlock + m.candv[c0 + c]);
                    const uint8_t *y = ix.long_codes + gvec * ix.long_stride + sd.long_off;
                    m.exv[c * S + s] = ex_ip_thread(sd.bits - 1, m.r + sd.doff, y, sd.dim);
                }
                __syncthreads();
                // accurate distance of every segment (compAccurateDist, caq_estimator.hpp:190-217)
                // and searchCluster's stage-3 running sum (saq_searcher.hpp:149-156)
                if (tid < nc) {
                    const int c = tid, vv = m.candv[c0 + c];
                    const int64_t gvec = vec0 + (int64_t)(st * kBlock + vv);
                    uint32_t *rec = a.cand + (int64_t)(cbase + c0 + c) * rst;
                    float acc = m.curr[vv];
                    rec[0] = ix.ids[gvec];
                    rec[1] = (uint32_t)(vv & 31);
                    rec[2] = __float_as_uint(acc);
                    for (int s = 0; s < S; ++s) {
                        const int bits = m.segs[s].bits;
                        const float *sc = m.sc + s * SC_N;
                        const float x = m.xs[s * kLT + vv];
                        float av, cd;
                        if (bits == 0) {
                            av = DT == 1 ? sc[SC_IPQC] : x * x + sc[SC_QL2];
                            cd = DT == 1 ? sc[SC_VFAC] : max0(x * x + sc[SC_VFAC]);
                        } else {
                            const float ipq = m.ipq[s * kLT + vv];
                            float f = (ipq - sc[SC_SUMQCB]) * m.segs[s].rescale_sqrtD;
                            f = f * x;
                            cd = DT == 1 ? f * 0.5f + sc[SC_IPQC] : max0(x * x + (sc[SC_QL2] - f));
                            const float sqd = m.segs[s].sqd;
                            const float ex = bits >= 2 ? m.exv[c * S + s] : 0.0f;
                            // Lut::getExtIP (lut.hpp:508-513)
                            const float t = (float)(((double)ipq + (double)ex * (double)sqd) +
                                                    (-1.0 + (double)(sqd / 2)) * (double)sc[SC_SUMQ]);
                            const float ipoq = ix.exfac[(gvec * S + s) * 2] * t;
                            av = DT == 1 ? ipoq + sc[SC_IPQC] : x * x + sc[SC_QL2] - 2 * ipoq;
                        }
                        if (single)
                            acc = av;
                        else
                            acc += av - cd;
                        rec[3 + s] = __float_as_uint(acc);
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- replay

// The replay's ResultPool lives in registers: entry 64 r + lane is row r of lane `lane`
// (KR rows cover topk <= 64 KR).
template <int KR>
struct RegPool {
    float d[KR];
    uint32_t id[KR];
    // entry i (the same value in every lane)
    __device__ __forceinline__ float dist_at(int i) const {
        float v = 0.0f;
#pragma unroll
        for (int r = 0; r < KR; ++r)
            if (r == (i >> 6)) v = __shfl(d[r], i & 63);
        return v;
    }
};

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
    uint32_t *crec = reinterpret_cast<uint32_t *>(smem);  // the current chunk's candidate records (<= 256)
    unsigned long long *tabs = reinterpret_cast<unsigned long long *>(crec + kLT * rst);  // kLT * rst dwords: 8-byte aligned
    unsigned long long *fastcum = tabs, *cumbits = tabs + S;  // per segment prefix sums of bits

    int size = a.first ? 0 : a.pool_size[q];
    RegPool<KR> pool;
#pragma unroll
    for (int r = 0; r < KR; ++r) {
        const int i = r * 64 + lane;
        pool.d[r] = (i < size) ? a.pool_d[q * K + i] : 0.0f;
        pool.id[r] = (i < size) ? a.pool_id[q * K + i] : 0u;
    }
    if (lane == 0) {
        unsigned long long fc = 0, cb = 0;
        for (int s = 0; s < S; ++s) {
            const SegDesc sd = ix.segs[s];
            if (sd.bits > 0) {
                fc += (unsigned long long)kBlock * sd.dim;
                cb += (unsigned long long)sd.dim * (sd.bits - 1);
            }
            fastcum[s] = fc;
            cumbits[s] = cb;
        }
    }
    // running metrics: the same value in every lane
    unsigned long long fast_bits = 0, acc_bits = 0, comp = 0;
    if (!a.first) {
        fast_bits = a.met[q * 3 + 0];
        acc_bits = a.met[q * 3 + 1];
        comp = a.met[q * 3 + 2];
    }
    __syncthreads();
    const int sbits0 = ix.segs[0].bits, sdim0 = ix.segs[0].dim;

    float tail = size == K ? pool.dist_at(K - 1) : FLT_MAX;  // raw (negated for IP)
    float dk = DT == 1 ? -tail : tail;
    // ResultPool::insert (pool.hpp:233-246): reject when full and d > tail; the new entry goes
    // before every entry >= d (ties: new before old), the rest move up one, the (k+1)-th drops
    auto insert = [&](uint32_t u, float d) {
        if (DT == 1) d = -d;
        if (size == K && d > tail) return;
        int lo = 0;
#pragma unroll
        for (int r = 0; r < KR; ++r) lo += __popcll(__ballot(r * 64 + lane < size && pool.d[r] < d));
        float carry_d = 0.0f;
        uint32_t carry_i = 0;
#pragma unroll
        for (int r = 0; r < KR; ++r) {
            const int i = r * 64 + lane;
            const float up_d = __shfl_up(pool.d[r], 1);
            const uint32_t up_i = __shfl_up(pool.id[r], 1);
            const float last_d = __shfl(pool.d[r], 63);
            const uint32_t last_i = __shfl(pool.id[r], 63);
            if (i > lo) {
                pool.d[r] = lane == 0 ? carry_d : up_d;
                pool.id[r] = lane == 0 ? carry_i : up_i;
            } else if (i == lo) {
                pool.d[r] = d;
                pool.id[r] = u;
            }
            carry_d = last_d;
            carry_i = last_i;
        }
        size += (size < K);
        tail = size == K ? pool.dist_at(K - 1) : FLT_MAX;
        dk = DT == 1 ? -tail : tail;
    };
    // The candidate records of a chunk (8 blocks, written together by one list-scan step) are
    // contiguous: one LDS DMA brings them all.  Returns the LDS record index of the block's
    // first candidate.  off/cnt/csum are the lanes' block info (csum: its chunk's total).
    int loaded = -1;        // chunk (window-relative) whose records are in LDS
    uint32_t loaded_base = 0;
    auto block_records = [&](int b, uint32_t off, uint32_t csum) -> int {
        const int ch = b >> 3;
        const uint32_t cb = __shfl(off, ch * 8);
        if (ch != loaded) {
            const int ndw = (int)__shfl(csum, ch * 8) * rst;
            const uint32_t *src = a.cand + (int64_t)cb * rst;
            for (int i0 = 0; i0 < ndw; i0 += 64)
                if (i0 + lane < ndw) __builtin_amdgcn_global_load_lds(src + i0 + lane, crec + i0, 4, 0, 0);
            dma_wait();
            __builtin_amdgcn_wave_barrier();
            loaded = ch;
            loaded_base = cb;
        }
        return (int)(__shfl(off, b) - loaded_base);
    };

    const int NPq = a.npq ? a.npq[q] : a.nprobe;
    const int pend = min(a.p1, NPq);
    unsigned long long fb_lane = 0;  // fast bits summed per lane, reduced at the end
    for (int p = a.p0; p < pend; ++p) {
        const uint32_t base = a.pair_blk[q * L + (p - a.p0)];
        if (base == kNoPair) continue;  // empty list (or block-info overflow: searched again)
        const uint32_t cid = a.probes[q * a.nprobe + p];
        const int n = ix.clu_n[cid];
        const int nb = (n + kBlock - 1) / kBlock;
        comp += (unsigned long long)nb * kBlock;  // total_comp_cnt
        const uint32_t *bi = a.binfo + base;
        if (single) {
            if (sbits0 > 0) fast_bits += (unsigned long long)nb * kBlock * sdim0;
            for (int w0 = 0; w0 < nb; w0 += 64) {
                const int blk = w0 + lane;
                uint32_t cnt = 0, off = 0;
                if (blk < nb) {
                    off = bi[blk * bst + 0];
                    cnt = bi[blk * bst + 1];
                }
                uint32_t csum = cnt;
                for (int o = 1; o < 8; o <<= 1) csum += __shfl_xor(csum, o, 8);
                loaded = -1;
                unsigned long long hb = __ballot(cnt > 0);
                while (hb) {
                    const int b = __builtin_ctzll(hb);
                    hb &= hb - 1;
                    const int bc = __shfl((int)cnt, b);
                    const uint32_t *br = crec + block_records(b, off, csum) * rst;
                    // scanCluster: the block's mask is fixed at its start (saq_searcher.hpp:190-195)
                    const float cv = lane < bc ? __uint_as_float(br[lane * rst + 2]) : FLT_MAX;
                    unsigned long long msk = __ballot(lane < bc && cv < dk);
                    if (sbits0 > 0) acc_bits += (unsigned long long)__popcll(msk) * sdim0 * (sbits0 - 1);
                    while (msk) {
                        const int j = __builtin_ctzll(msk);
                        msk &= msk - 1;
                        insert(br[j * rst + 0], __uint_as_float(br[j * rst + 3]));
                    }
                }
            }
        } else {
            for (int w0 = 0; w0 < nb; w0 += 64) {
                const int blk = w0 + lane;
                const bool valid = blk < nb;
                uint32_t cnt = 0, off = 0;
                float m1 = FLT_MAX, M = FLT_MAX;
                if (valid) {
                    off = bi[blk * bst + 0];
                    cnt = bi[blk * bst + 1];
                    m1 = __uint_as_float(bi[blk * bst + 2]);
                    M = __uint_as_float(bi[blk * bst + 3]);
                }
                uint32_t csum = cnt;
                for (int o = 1; o < 8; o <<= 1) csum += __shfl_xor(csum, o, 8);
                loaded = -1;
                int cur = 0;
                for (;;) {
                    // blocks cur.. see the same distk until the next insert
                    const bool pass1 = valid && lane >= cur && m1 <= dk;
                    const unsigned long long cb = __ballot(pass1 && M <= dk && cnt > 0);
                    const int stop = cb ? __builtin_ctzll(cb) : 64;
                    if (pass1 && lane <= stop) {
                        // stage 2 stops at the first quantised segment whose block minimum
                        // exceeds distk (saq_searcher.hpp:112-132)
                        int sb = S - 1;
                        for (int s = 0; s < S; ++s) {
                            if (ix.segs[s].bits > 0 && __uint_as_float(bi[blk * bst + 4 + s]) > dk) {
                                sb = s;
                                break;
                            }
                        }
                        fb_lane += fastcum[sb];
                    }
                    if (!cb) break;
                    const int bc = __shfl((int)cnt, stop);
                    const uint32_t *br = crec + block_records(stop, off, csum) * rst;
                    // stage 3 (saq_searcher.hpp:134-158): distk re-read after every insert
                    const float cv = lane < bc ? __uint_as_float(br[lane * rst + 2]) : FLT_MAX;
                    unsigned long long rem = bc >= 64 ? ~0ull : ((1ull << bc) - 1);
                    for (;;) {
                        const unsigned long long hit = __ballot(lane < bc && cv < dk) & rem;
                        if (!hit) break;
                        const int j = __builtin_ctzll(hit);
                        rem &= ~((2ull << j) - 1);
                        const float pv = lane < S ? __uint_as_float(br[j * rst + 3 + lane]) : 0.0f;
                        const unsigned long long ge = __ballot(lane < S && pv >= dk);
                        const int sb = ge ? __builtin_ctzll(ge) : S - 1;
                        acc_bits += cumbits[sb];
                        insert(br[j * rst + 0], __shfl(pv, sb));
                    }
                    cur = stop + 1;
                }
            }
        }
    }
    // per-lane fast bits to lane 0's total
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
    return (size_t)kLT * cand_stride(ix.nseg) * 4 + 8 + (size_t)ix.nseg * 16 + 16;
}

// counters of a slice: block-info and candidate space restart at zero, the most either
// reached is kept in ctr[3], ctr[4] (the host sizes the buffers from them); ctr[2] is the
// overflow flag of the whole search.
__global__ void slice_reset_kernel(uint32_t *ctr, int all) {
    if (threadIdx.x == 0) {
        if (all) {
            ctr[2] = ctr[3] = ctr[4] = 0;
        } else {
            ctr[3] = max(ctr[3], ctr[0]);
            ctr[4] = max(ctr[4], ctr[1]);
        }
        ctr[0] = ctr[1] = 0;
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
