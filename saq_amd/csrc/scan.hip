// scan.hip -- the plain SAQ IVF scan on gfx950.
//
// Reproduces IVF::search (saqlib/index/ivf.hpp:275-311) for a batch of queries:
//   query_prep_kernel  per (query, segment): the rotated query q_s * P_s in the order of
//                      Eigen's column-major GEMV (saq_estimator.hpp:28-48, caq_estimator.hpp:56-82)
//   query_bound_kernel per (query, segment), one wavefront: the variance bound
//                      m * sqrt(sum var * q^2) (caq_estimator.hpp:80-82)
//   scan_kernel        one workgroup per query walks its probe list in order.  Per probed
//                      cluster it builds every segment's u16 fast-scan LUT in LDS
//                      (Lut::prepare, lut.hpp:60-87), then streams the cluster in steps of
//                      8 blocks (256 vectors, one lane per vector):
//                        1. fast scan: exact integer LUT sums from 8-byte code loads, then the
//                           reference's fp32 epilogue (lut.hpp:89-119, caq_estimator.hpp:116-179)
//                        2. speculative refine: every vector that can still be refined under
//                           the current k-th distance gets its extra-bit inner product
//                           (CodeHelper<b-1>::compute_ip) from 16 lanes, in the reference's
//                           16-lane order and reduction tree, and its stage-3 running sum
//                      The block records of the next step (or of the next probed list), the next
//                      list's rotated centroid and the candidates' long codes move from HBM to LDS
//                      by LDS DMA (global_load_lds_dwordx4), issued ahead of use so that the
//                      transfers overlap the replay of the current step.
//                        3. exact-order replay by one wavefront: the block/segment pruning of
//                           scanCluster / searchCluster (saq_searcher.hpp:68-214) and
//                           ResultPool insertion (utils/pool.hpp:12-61) in reference order.
//                      The k-th distance only decreases, so every vector the replay refines is
//                      in the speculative set; IDs and distances match the sequential
//                      reference bit for bit.
#include <cfloat>

#include "common.hpp"
#include "estimators.hpp"

#pragma clang fp contract(off)

namespace saqgpu {

constexpr int kThreads = 256;
constexpr int kStepBlocks = 8;            // 256 vectors per step

struct PrepArgs {
    IndexView ix;
    const float *queries;  // nq x num_dim
    int64_t nq;
    const float *variance; // D_pad
    const float *rot;      // concatenated P_s or nullptr
    const int64_t *rot_off;
    float m;
    float *qrot;   // nq x D_pad
    float *bound;  // nq x nseg
};

constexpr int kPrepQ = 8;

// grid (ceil(nq/8), nseg, ceil(D_s/256)): 256 outputs x 8 queries per workgroup.
__global__ __launch_bounds__(256) void query_prep_kernel(PrepArgs a) {
    __shared__ __align__(16) float qs[kPrepQ][256];
    const IndexView &ix = a.ix;
    const int s = blockIdx.y;
    const SegDesc sd = ix.segs[s];
    const int D = sd.dim;
    const int j = blockIdx.z * 256 + threadIdx.x;
    if (blockIdx.z * 256 >= D) return;
    const int64_t q0 = (int64_t)blockIdx.x * kPrepQ;
    const float *P = a.rot ? a.rot + a.rot_off[s] : nullptr;
    // Eigen GEMV block of input dims (GeneralMatrixVector.h:147): all of D below 128, else 16.
    const int bc = D < 128 ? D : (D * 4 < 32000 ? 16 : 4);
    float acc[kPrepQ], c[kPrepQ];
#pragma unroll
    for (int t = 0; t < kPrepQ; ++t) acc[t] = c[t] = 0.0f;
    for (int i0 = 0; i0 < D; i0 += 256) {
        __syncthreads();
        for (int t = 0; t < kPrepQ; ++t) {
            const int64_t q = q0 + t;
            const int d = sd.doff + i0 + threadIdx.x;
            float v = 0.0f;
            if (q < a.nq && i0 + (int)threadIdx.x < D && d < ix.num_dim) v = a.queries[q * ix.num_dim + d];
            qs[t][threadIdx.x] = v;
        }
        __syncthreads();
        if (j < D && bc == 16 && P) {
            // blocks of 16 input dims (256-dim chunks hold whole blocks): the block's 16 rows of P
            // are loaded two blocks ahead (the last block's rows again past the chunk's end, so
            // the loads are unconditional), then the block's FMA chain per query, then res += c
            const int iend = min(256, D - i0);
            float pa[16], pb[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) pa[u] = P[(int64_t)(i0 + u) * D + j];
#pragma unroll
            for (int u = 0; u < 16; ++u) pb[u] = P[(int64_t)(i0 + min(16, iend - 16) + u) * D + j];
            for (int ib = 0; ib < iend; ib += 16) {
                float pr[16];
                const int inext = min(ib + 32, iend - 16);
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    pr[u] = pa[u];
                    pa[u] = pb[u];
                    pb[u] = P[(int64_t)(i0 + inext + u) * D + j];
                }
                // each query's 16 inputs of the block in four 16-byte LDS reads, then its chain
#pragma unroll
                for (int t = 0; t < kPrepQ; ++t) {
                    const float4 *xv = reinterpret_cast<const float4 *>(&qs[t][ib]);
                    float x[16];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const float4 f = xv[v];
                        x[4 * v] = f.x;
                        x[4 * v + 1] = f.y;
                        x[4 * v + 2] = f.z;
                        x[4 * v + 3] = f.w;
                    }
#pragma unroll
                    for (int u = 0; u < 16; ++u) c[t] = __fmaf_rn(pr[u], x[u], c[t]);
                }
#pragma unroll
                for (int t = 0; t < kPrepQ; ++t) {
                    acc[t] = acc[t] + c[t];  // pmadd(c, alpha = 1, res)
                    c[t] = 0.0f;
                }
            }
        } else if (j < D) {
            const int iend = min(256, D - i0);
            for (int ii = 0; ii < iend; ++ii) {
                const int i = i0 + ii;
                const float p = P ? P[(int64_t)i * D + j] : (i == j ? 1.0f : 0.0f);
#pragma unroll
                for (int t = 0; t < kPrepQ; ++t) c[t] = __fmaf_rn(p, qs[t][ii], c[t]);
                if ((i + 1) % bc == 0 || i + 1 == D) {
#pragma unroll
                    for (int t = 0; t < kPrepQ; ++t) {
                        acc[t] = acc[t] + c[t];  // pmadd(c, alpha = 1, res)
                        c[t] = 0.0f;
                    }
                }
            }
        }
    }
    if (j < D) {
        for (int t = 0; t < kPrepQ; ++t) {
            const int64_t q = q0 + t;
            if (q >= a.nq) continue;
            const int d = sd.doff + j;
            // no rotator: the zero-padded query slice itself (saq_estimator.hpp:35-40)
            a.qrot[q * ix.D_pad + d] = P ? acc[t] : (d < ix.num_dim ? a.queries[q * ix.num_dim + d] : 0.0f);
        }
    }
}

// grid (nq, nseg), one wavefront each: bound = sqrt((var.segment * q^2).sum()) * m of one query's
// segment (saq_estimator.hpp:42-44), Eigen redux order
__global__ __launch_bounds__(64) void query_bound_kernel(PrepArgs a) {
    const IndexView &ix = a.ix;
    const int64_t q = blockIdx.x;
    const int s = blockIdx.y;
    const SegDesc sd = ix.segs[s];
    const float *qq = a.queries + q * ix.num_dim;
    const float v2 = eigen_redux_wave(sd.dim, [&](int i) {
        const int d = sd.doff + i;
        const float x = d < ix.num_dim ? qq[d] : 0.0f;
        return a.variance[d] * (x * x);
    });
    if (threadIdx.x == 0) a.bound[q * ix.nseg + s] = sqrtf(v2) * a.m;
}

struct ScanArgs {
    IndexView ix;
    const float *qrot;
    const float *bound;
    const uint32_t *probes;
    int32_t nprobe;          // probe-list stride
    const int32_t *npq;      // lists scanned per query (nullptr: nprobe)
    int32_t topk;
    int64_t nq;
    uint32_t *out_ids;
    float *out_dists;
    uint64_t *out_metrics;
    unsigned long long *prof;  // optional per-query phase clocks (measurement only)
    int32_t cmax;              // candidates per refine round (bounded by LDS for their records)
};

// control block indices
enum { C_PSIZE = 0, C_POS, C_MASK, C_CURSOR, C_N = 8 };

struct Smem {
    float *r, *qr, *crbuf;
    uint16_t *lut;
    float *cbmin, *cbmax, *sc, *xs, *ipq, *curr, *binfo, *accd, *exr, *exv;
    int *cand;
    uint32_t *cid;
    int *wcnt;
    int8_t *gseg, *candidx, *bseq;
    int *sbits;
    float *sresc, *ssqd;
    uint8_t *lbuf, *blkbuf;
    float *pd;
    uint32_t *pid;
    int *ctl;
    unsigned long long *met;
    SegDesc *segs;
};

__device__ __host__ inline size_t carve_layout(unsigned char *base, const IndexView &ix, int topk, int cmax,
                                                Smem *mp) {
    const int S = ix.nseg;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        off = (off + 15) & ~size_t(15);
        unsigned char *p = base + off;
        off += bytes;
        return p;
    };
    Smem m{};
    m.blkbuf = take((size_t)kStepBlocks * ix.block_stride);
    m.lbuf = take((size_t)cmax * ix.long_stride);
    m.r = reinterpret_cast<float *>(take(ix.D_pad * 4));
    m.qr = reinterpret_cast<float *>(take(ix.D_pad * 4));
    m.crbuf = reinterpret_cast<float *>(take(ix.D_pad * 4));
    m.lut = reinterpret_cast<uint16_t *>(take((size_t)ix.ngroups * 256 * 2));
    m.cbmin = reinterpret_cast<float *>(take((size_t)ix.ngroups * 16 * 4));
    m.cbmax = reinterpret_cast<float *>(take((size_t)ix.ngroups * 16 * 4));
    m.sc = reinterpret_cast<float *>(take((size_t)S * SC_N * 4));
    m.xs = reinterpret_cast<float *>(take((size_t)S * kThreads * 4));
    m.ipq = reinterpret_cast<float *>(take((size_t)S * kThreads * 4));
    m.curr = reinterpret_cast<float *>(take(kThreads * 4));
    m.binfo = reinterpret_cast<float *>(take((size_t)kStepBlocks * (S + 2) * 4));
    m.accd = reinterpret_cast<float *>(take((size_t)cmax * S * 4));
    m.exr = reinterpret_cast<float *>(take((size_t)cmax * S * 4));
    m.exv = reinterpret_cast<float *>(take((size_t)cmax * S * 4));
    m.cand = reinterpret_cast<int *>(take((size_t)cmax * 4));
    m.cid = reinterpret_cast<uint32_t *>(take((size_t)cmax * 4));
    m.wcnt = reinterpret_cast<int *>(take(16));
    m.gseg = reinterpret_cast<int8_t *>(take(ix.ngroups));
    m.candidx = reinterpret_cast<int8_t *>(take(kThreads));
    m.bseq = reinterpret_cast<int8_t *>(take(S));
    m.sbits = reinterpret_cast<int *>(take((size_t)S * 4));
    m.sresc = reinterpret_cast<float *>(take((size_t)S * 4));
    m.ssqd = reinterpret_cast<float *>(take((size_t)S * 4));
    m.pd = reinterpret_cast<float *>(take((size_t)(topk + 1) * 4));
    m.pid = reinterpret_cast<uint32_t *>(take((size_t)(topk + 1) * 4));
    m.ctl = reinterpret_cast<int *>(take(C_N * 4));
    m.met = reinterpret_cast<unsigned long long *>(take(4 * 8));
    m.segs = reinterpret_cast<SegDesc *>(take((size_t)S * sizeof(SegDesc)));
    if (mp) *mp = m;
    return off + 16;
}

size_t scan_smem_bytes(const IndexView &ix, int topk, int cmax) {
    return carve_layout(nullptr, ix, topk, cmax, nullptr);
}


template <int DT>
__device__ __forceinline__ float pool_distk(const Smem &m, int K) {
    const int size = m.ctl[C_PSIZE];
    const float raw = size == K ? m.pd[size - 1] : FLT_MAX;
    return DT == 1 ? -raw : raw;
}

template <int DT>
__global__ __launch_bounds__(kThreads) void scan_kernel(ScanArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const IndexView &ix = a.ix;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t q = blockIdx.x;
    const int S = ix.nseg, K = a.topk, NG = ix.ngroups, CM = a.cmax;
    Smem mm;
    carve_layout(smem, ix, K, CM, &mm);
    const Smem m = mm;
    for (int i = tid; i < S; i += kThreads) {
        const SegDesc sd = ix.segs[i];
        m.segs[i] = sd;
        m.sbits[i] = sd.bits;
        m.sresc[i] = sd.rescale_sqrtD;
        m.ssqd[i] = sd.sqd;
    }
    for (int i = tid; i <= K; i += kThreads) {
        m.pd[i] = 0.0f;
        m.pid[i] = 0;
    }
    if (tid < C_N) m.ctl[tid] = 0;
    if (tid < 4) m.met[tid] = 0;
    const float *qrot = a.qrot + q * ix.D_pad;
    for (int i = tid; i < ix.D_pad; i += kThreads) m.qr[i] = qrot[i];
    // group -> segment table (groups of one segment are contiguous), segment ends as a mask,
    // the quantised segments in order, and the prefix sums of refine bits (one per lane)
    uint32_t endmask = 0;
    int nbq = 0;
    for (int s = 0; s < S; ++s) {
        const SegDesc sd = ix.segs[s];
        if (sd.bits == 0) continue;
        endmask |= 1u << (sd.group_off + sd.ngroups - 1);
        for (int g = tid; g < sd.ngroups; g += kThreads) m.gseg[sd.group_off + g] = (int8_t)s;
        if (tid == 0) m.bseq[nbq] = (int8_t)s;
        ++nbq;
    }
    unsigned long long cumbits = 0, fastcum = 0;
    for (int s = 0; s <= min(lane, S - 1); ++s) {
        const SegDesc sd = ix.segs[s];
        if (sd.bits > 0) {
            cumbits += (unsigned long long)sd.dim * (sd.bits - 1);
            fastcum += (unsigned long long)kBlock * sd.dim;
        }
    }
    unsigned long long tph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    unsigned long long tlast = a.prof ? wall_clock64() : 0;
    auto mark = [&](int ph) {
        if (a.prof && tid == 0) {
            const unsigned long long t = wall_clock64();
            tph[ph] += t - tlast;
            tlast = t;
        }
    };
    const bool single = (S == 1);
    const float *bound = a.bound + q * S;
    const int ncb = NG * 16;
    const int v = tid, bl = v >> 5, ln = v & 31;

    const int NPq = a.npq ? a.npq[q] : a.nprobe;
    auto next_valid = [&](int i) {
        for (; i < NPq; ++i) {
            const uint32_t c = a.probes[q * a.nprobe + i];
            if (c < (uint32_t)ix.num_cen && ix.clu_n[c] > 0) break;  // bad ids are rejected on the host
        }
        return i;
    };
    // LDS DMA of the block records of one step, and of a list's rotated centroid
    auto fetch_step = [&, m](uint32_t cid, int st) {
        const int nbn = (ix.clu_n[cid] + kBlock - 1) / kBlock;
        const int nbs = min(kStepBlocks, nbn - st);
        dma_to_lds(ix.blocks + (ix.clu_blk[cid] + st) * ix.block_stride, m.blkbuf, (int)(nbs * ix.block_stride),
                   wave, lane);
    };
    auto fetch_crot = [&, m](uint32_t cid) {
        dma_to_lds(reinterpret_cast<const uint8_t *>(ix.crot + (int64_t)cid * ix.D_pad),
                   reinterpret_cast<uint8_t *>(m.crbuf), ix.D_pad * 4, wave, lane);
    };

    int pi = next_valid(0);
    if (pi < NPq) {
        const uint32_t c0 = a.probes[q * a.nprobe + pi];
        fetch_crot(c0);
        fetch_step(c0, 0);
    }
    while (pi < NPq) {
        const uint32_t cid = a.probes[q * a.nprobe + pi];
        const int n = ix.clu_n[cid];
        const int nb = (n + kBlock - 1) / kBlock;
        const int64_t vec0 = ix.clu_vec[cid];
        const int pn = next_valid(pi + 1);
        dma_wait();
        __syncthreads();
        mark(0);

        // ---- CaqCluEstimator::prepare for every segment (caq_estimator.hpp:94-106)
        prepare_list<DT>(m.segs, S, m.gseg, NG, m.qr, m.crbuf, bound, m.r, m.sc, m.cbmin, m.cbmax, m.lut, ix.D_pad);
        mark(2);
        if (tid == 0) m.met[2] += (unsigned long long)nb * kBlock;  // total_comp_cnt

        // ---- blocks, 8 at a time; the block records are already in LDS (blkbuf)
        for (int st = 0; st < nb; st += kStepBlocks) {
            const int nbs = min(kStepBlocks, nb - st);
            const bool active = bl < nbs;
            if (st > 0) {
                dma_wait();
                __syncthreads();
            }
            if (active) {
                const uint8_t *rec = m.blkbuf + bl * ix.block_stride;
                const float *fac = reinterpret_cast<const float *>(rec);
                for (int s = 0; s < S; ++s) m.xs[s * kThreads + v] = fac[s * kBlock + ln];
                const uint8_t *codes = rec + ix.fac_floats * 4 + ln * 8;
                // exact integer LUT sums (accumulate_hacc, fastscan_highacc.hpp:66-139), one segment
                // ending at every set bit of endmask
                uint32_t acc = 0;
                int bk = 0;
#pragma unroll 2
                for (int g = 0; g < NG; ++g) {
                    const uint2 w = *reinterpret_cast<const uint2 *>(codes + g * 256);
                    const uint16_t *t = m.lut + g * 256;
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const uint32_t byte = (w.x >> (8 * b)) & 0xFF;
                        acc += t[(2 * b) * 16 + (byte >> 4)];
                        acc += t[(2 * b + 1) * 16 + (byte & 15)];
                    }
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const uint32_t byte = (w.y >> (8 * b)) & 0xFF;
                        acc += t[(2 * b + 8) * 16 + (byte >> 4)];
                        acc += t[(2 * b + 9) * 16 + (byte & 15)];
                    }
                    if ((endmask >> g) & 1u) {
                        const int s = m.bseq[bk++];
                        const float *sc = m.sc + s * SC_N;
                        float t2 = (float)(int32_t)acc * sc[SC_DELTA];  // lut.hpp:495-497
                        t2 = t2 + sc[SC_SUMVL];
                        m.ipq[s * kThreads + v] = t2;
                        acc = 0;
                    }
                }
            }
            // per-vector estimates (distk independent) and per-block minima
            {
                float cur = 0.0f;
                if (single) {
                    const float *sc = m.sc;
                    const float x = active ? m.xs[v] : 0.0f;
                    if (m.sbits[0] == 0) {
                        cur = DT == 1 ? sc[SC_VFAC] : max0(x * x + sc[SC_VFAC]);
                    } else {
                        const float ipq = active ? m.ipq[v] : 0.0f;
                        float f = (ipq - sc[SC_SUMQCB]) * m.sresc[0];
                        f = f * x;
                        cur = DT == 1 ? f * 0.5f + sc[SC_IPQC] : max0(x * x + (sc[SC_QL2] - f));
                    }
                    if (active) m.curr[v] = cur;
                } else {
                    for (int s = 0; s < S; ++s) {
                        const float x = active ? m.xs[s * kThreads + v] : 0.0f;
                        const float vd = DT == 1 ? m.sc[s * SC_N + SC_VFAC] : max0(x * x + m.sc[s * SC_N + SC_VFAC]);
                        cur = cur + vd;
                    }
                    float m1 = cur;
                    for (int o = 16; o >= 1; o >>= 1) m1 = fminf(m1, __shfl_xor(m1, o, 32));
                    float M = m1;
                    for (int s = 0; s < S; ++s) {
                        if (m.sbits[s] == 0) continue;
                        const float *sc = m.sc + s * SC_N;
                        const float x = active ? m.xs[s * kThreads + v] : 0.0f;
                        const float vd = DT == 1 ? sc[SC_VFAC] : max0(x * x + sc[SC_VFAC]);
                        const float ipq = active ? m.ipq[s * kThreads + v] : 0.0f;
                        float f = (ipq - sc[SC_SUMQCB]) * m.sresc[s];
                        f = f * x;
                        const float fd = DT == 1 ? f * 0.5f + sc[SC_IPQC] : max0(x * x + (sc[SC_QL2] - f));
                        cur = cur - vd;
                        cur = cur + fd;
                        float ms = cur;
                        for (int o = 16; o >= 1; o >>= 1) ms = fminf(ms, __shfl_xor(ms, o, 32));
                        M = fmaxf(M, ms);
                        if (active && ln == 0) m.binfo[bl * (S + 2) + 2 + s] = ms;
                    }
                    if (active) {
                        m.curr[v] = cur;
                        if (ln == 0) {
                            m.binfo[bl * (S + 2) + 0] = m1;
                            m.binfo[bl * (S + 2) + 1] = M;
                        }
                    }
                }
            }
            if (tid == 0) {
                m.ctl[C_POS] = 0;
                m.ctl[C_MASK] = 0;
                m.ctl[C_CURSOR] = 0;  // no block in progress
            }
            __syncthreads();
            mark(3);
            // the step's block records are consumed: fetch the next work unit's
            if (st + kStepBlocks < nb) {
                fetch_step(cid, st + kStepBlocks);
            } else if (pn < NPq) {
                const uint32_t nc = a.probes[q * a.nprobe + pn];
                fetch_crot(nc);
                fetch_step(nc, 0);
            }

            // ---- refine / replay rounds
            for (;;) {
                const float distk = pool_distk<DT>(m, K);
                const int pos = m.ctl[C_POS];
                const int pblk = pos >> 5;
                const bool inprog = m.ctl[C_CURSOR] != 0;  // block pblk was started and left mid-way
                const uint32_t inmask = (uint32_t)m.ctl[C_MASK];  // single: its remaining mask
                bool flag = false;
                if (active && v >= pos) {
                    const int idx = (st + bl) * kBlock + ln;
                    if (single) {
                        if (bl == pblk && inprog) {
                            flag = (inmask >> ln) & 1;
                        } else {
                            const int cnt = (st + bl == nb - 1) ? n % kBlock : kBlock;
                            flag = (ln < cnt) && (m.curr[v] < distk);
                        }
                    } else {
                        const bool blk_ok = (bl == pblk && inprog) || (m.binfo[bl * (S + 2) + 1] <= distk);
                        flag = blk_ok && idx < n && (m.curr[v] < distk);
                    }
                }
                // ordered compaction of the speculative candidates (wave-major, lane order)
                const unsigned long long bal = __ballot(flag);
                if (lane == 0) m.wcnt[wave] = __popcll(bal);
                __syncthreads();
                int before = 0, total = 0;
                for (int w = 0; w < 4; ++w) {
                    if (w < wave) before += m.wcnt[w];
                    total += m.wcnt[w];
                }
                const int rank = before + __popcll(bal & ((1ull << lane) - 1));
                const bool take = flag && rank < CM;
                if (take) m.cand[rank] = v;
                if (active) m.candidx[v] = take ? (int8_t)rank : (int8_t)-1;
                const int ncand = min(total, CM);
                __syncthreads();
                mark(4);
                if (ncand) {
                    // long-code records of the candidates by LDS DMA; ids and rescale factors
                    const int chunks = (int)(ix.long_stride / 16);
                    const int nchunk = ncand * chunks;
                    for (int off = wave * 64; off < nchunk; off += kThreads) {
                        const int i = off + lane;
                        if (i < nchunk) {
                            const int c = i / chunks, ch = i - c * chunks;
                            const int64_t gvec = vec0 + (int64_t)(st * kBlock + m.cand[c]);
                            __builtin_amdgcn_global_load_lds(ix.long_codes + gvec * ix.long_stride + ch * 16,
                                                             m.lbuf + (size_t)off * 16, 16, 0, 0);
                        }
                    }
                    uint32_t my_id = 0;
                    if (tid < ncand) my_id = ix.ids[vec0 + (int64_t)(st * kBlock + m.cand[tid])];
                    const int nr = ncand * S;
                    float rs = 0.0f;
                    if (tid < nr) {
                        const int c = tid / S, s = tid - c * S;
                        const int64_t gvec = vec0 + (int64_t)(st * kBlock + m.cand[c]);
                        rs = ix.exfac[(gvec * S + s) * 2];
                    }
                    for (int i = tid + kThreads; i < nr; i += kThreads) {
                        const int c = i / S, s = i - c * S;
                        const int64_t gvec = vec0 + (int64_t)(st * kBlock + m.cand[c]);
                        m.exr[i] = ix.exfac[(gvec * S + s) * 2];
                    }
                    if (tid < nr) m.exr[tid] = rs;
                    if (tid < ncand) m.cid[tid] = my_id;
                    dma_wait();
                    __syncthreads();
                    // extra-bit inner products (CodeHelper<b-1>::compute_ip), 16 lanes per item
                    const int grp = tid >> 4, gl = tid & 15;
                    for (int item = grp; item < ncand * S; item += kThreads / 16) {
                        const int c = item / S, s = item - c * S;
                        const int bits = m.sbits[s];
                        if (bits < 2) continue;
                        const SegDesc &sd = m.segs[s];
                        const uint8_t *y = m.lbuf + c * ix.long_stride + sd.long_off;
                        const float ex = ex_ip_group(bits - 1, m.r + sd.doff, y, sd.dim, gl);
                        if (gl == 0) m.exv[item] = ex;
                    }
                    __syncthreads();
                    // per candidate: the accurate distance of every segment (compAccurateDist,
                    // caq_estimator.hpp:190-217) and the running sum of searchCluster's stage 3
                    // (saq_searcher.hpp:149-156), which does not depend on the k-th distance --
                    // only where the replay stops it does
                    if (tid < ncand) {
                        const int c = tid, vv = m.cand[c];
                        float acc = m.curr[vv];
                        for (int s = 0; s < S; ++s) {
                            const int bits = m.sbits[s];
                            const float *sc = m.sc + s * SC_N;
                            const float x = m.xs[s * kThreads + vv];
                            float av, cd;
                            if (bits == 0) {
                                av = DT == 1 ? sc[SC_IPQC] : x * x + sc[SC_QL2];
                                cd = DT == 1 ? sc[SC_VFAC] : max0(x * x + sc[SC_VFAC]);
                            } else {
                                const float ipq = m.ipq[s * kThreads + vv];
                                float f = (ipq - sc[SC_SUMQCB]) * m.sresc[s];
                                f = f * x;
                                cd = DT == 1 ? f * 0.5f + sc[SC_IPQC] : max0(x * x + (sc[SC_QL2] - f));
                                const float sqd = m.ssqd[s];
                                const float ex = bits >= 2 ? m.exv[c * S + s] : 0.0f;
                                // Lut::getExtIP (lut.hpp:508-513)
                                const float t = (float)(((double)ipq + (double)ex * (double)sqd) +
                                                        (-1.0 + (double)(sqd / 2)) * (double)sc[SC_SUMQ]);
                                const float ipoq = m.exr[c * S + s] * t;
                                av = DT == 1 ? ipoq + sc[SC_IPQC] : x * x + sc[SC_QL2] - 2 * ipoq;
                            }
                            if (single)
                                acc = av;
                            else
                                acc += av - cd;
                            m.accd[c * S + s] = acc;
                        }
                    }
                }
                __syncthreads();
                mark(5);
                // exact-order replay by wavefront 0 (LDS and registers only)
                if (wave == 0) {
                    const int lastv = total > CM ? m.cand[CM - 1] : kThreads;
                    int size = m.ctl[C_PSIZE];
                    float tail = size == K ? m.pd[K - 1] : FLT_MAX;  // raw (negated for IP)
                    float dk = DT == 1 ? -tail : tail;
                    int cur_pos = pos;
                    bool prog = inprog;
                    uint32_t mask_state = inmask;
                    bool need_more = false;
                    unsigned long long fast_bits = 0, acc_bits = 0;
                    auto insert = [&, m](uint32_t u, float d) {  // ResultPool::insert (pool.hpp:233-246)
                        if (DT == 1) d = -d;
                        if (size == K && d > tail) return;
                        int lo = 0;
                        for (int base = 0; base < size; base += 64) {
                            const int i = base + lane;
                            lo += __popcll(__ballot((i < size) && (m.pd[i] < d)));
                        }
                        for (int hi = size; hi > lo; hi -= 64) {
                            const int idx = hi - lane;
                            float dv = 0.0f;
                            uint32_t iv = 0;
                            const bool act = idx > lo;
                            if (act) {
                                dv = m.pd[idx - 1];
                                iv = m.pid[idx - 1];
                            }
                            __builtin_amdgcn_wave_barrier();
                            if (act) {
                                m.pd[idx] = dv;
                                m.pid[idx] = iv;
                            }
                            __builtin_amdgcn_wave_barrier();
                        }
                        if (lane == 0) {
                            m.pd[lo] = d;
                            m.pid[lo] = u;
                        }
                        __builtin_amdgcn_wave_barrier();
                        size += (size < K);
                        tail = size == K ? m.pd[K - 1] : FLT_MAX;
                        dk = DT == 1 ? -tail : tail;
                    };
                    for (int b = pblk; b < nbs && !need_more; ++b) {
                        const int j0 = (b == pblk) ? (pos & 31) : 0;
                        const bool fresh = !(b == pblk && prog);
                        const float my_curr = lane < 32 ? m.curr[b * 32 + lane] : FLT_MAX;
                        if (single) {
                            uint32_t msk;
                            if (fresh) {
                                fast_bits += (unsigned long long)kBlock * m.segs[0].dim * (m.sbits[0] ? 1 : 0);
                                const int cnt = (st + b == nb - 1) ? n % kBlock : kBlock;
                                msk = (uint32_t)__ballot(lane < 32 && my_curr < dk) & (uint32_t)((1ull << cnt) - 1);
                            } else {
                                msk = mask_state;
                            }
                            while (msk) {
                                const int j = __builtin_ctz(msk);
                                const int vv = b * 32 + j;
                                if (vv > lastv) {
                                    need_more = true;
                                    cur_pos = vv;
                                    mask_state = msk;
                                    prog = true;
                                    break;
                                }
                                msk &= msk - 1;
                                const int c = m.candidx[vv];
                                if (m.sbits[0]) acc_bits += (unsigned long long)m.segs[0].dim * (m.sbits[0] - 1);
                                insert(m.cid[c], m.accd[c * S]);
                            }
                            if (!need_more) prog = false;
                        } else {
                            if (fresh) {
                                // stage 1 (skip on the variance minimum), then stage 2: the fast
                                // estimate of each quantised segment in turn, stopping at the first
                                // whose block minimum exceeds dk (saq_searcher.hpp:95-132)
                                const float *bi = m.binfo + b * (S + 2);
                                const float bv = lane < S + 2 ? bi[lane] : 0.0f;
                                if (__shfl(bv, 0) > dk) continue;
                                const bool brk = lane >= 2 && lane < S + 2 && m.sbits[lane - 2] > 0 && bv > dk;
                                const unsigned long long bm = __ballot(brk);
                                const int sb = bm ? __builtin_ctzll(bm) - 2 : S - 1;
                                fast_bits += (unsigned long long)__shfl(fastcum, sb);
                                if (__shfl(bv, 1) > dk) continue;  // stage 2 broke: no refine
                            }
                            const int nvalid = min(kBlock, n - (st + b) * kBlock);
                            uint32_t rem = (uint32_t)((1ull << nvalid) - 1) & ~((1u << j0) - 1u);
                            for (;;) {
                                const uint32_t hit = (uint32_t)__ballot(lane < 32 && my_curr < dk) & rem;
                                if (!hit) break;
                                const int j = __builtin_ctz(hit);
                                const int vv = b * 32 + j;
                                if (vv > lastv) {
                                    need_more = true;
                                    cur_pos = vv;
                                    prog = true;
                                    break;
                                }
                                rem &= ~((2u << j) - 1u);
                                const int c = m.candidx[vv];
                                // the first segment at which the running sum reaches dk ends the loop
                                const float pv = lane < S ? m.accd[c * S + lane] : 0.0f;
                                const unsigned long long ge = __ballot(lane < S && pv >= dk);
                                const int sb = ge ? __builtin_ctzll(ge) : S - 1;
                                acc_bits += (unsigned long long)__shfl(cumbits, sb);
                                insert(m.cid[c], __shfl(pv, sb));
                            }
                            if (!need_more) prog = false;
                        }
                    }
                    if (lane == 0) {
                        m.met[0] += fast_bits;
                        m.met[1] += acc_bits;
                        m.ctl[C_PSIZE] = size;
                        m.ctl[C_POS] = need_more ? cur_pos : kThreads;
                        m.ctl[C_MASK] = need_more ? (int)mask_state : 0;
                        m.ctl[C_CURSOR] = need_more && prog ? 1 : 0;
                    }
                }
                __syncthreads();
                mark(6);
                if (m.ctl[C_POS] >= kThreads) break;
                __syncthreads();
            }
        }
        pi = pn;
    }
    __syncthreads();
    for (int i = tid; i < K; i += kThreads) {
        a.out_ids[q * K + i] = m.pid[i];
        if (a.out_dists) a.out_dists[q * K + i] = DT == 1 ? -m.pd[i] : m.pd[i];
    }
    if (a.out_metrics && tid < 3) a.out_metrics[q * 3 + tid] = m.met[tid];
    if (a.prof && tid == 0)
        for (int i = 0; i < 8; ++i) a.prof[q * 8 + i] = tph[i];
}

}  // namespace saqgpu

// ---------------------------------------------------------------- launchers
namespace saqgpu {

hipError_t launch_query_prep(const IndexView &ix, const float *queries, int64_t nq, const float *variance,
                             const float *rot, const int64_t *rot_off, float m, float *qrot, float *bound,
                             int max_dim, hipStream_t st) {
    PrepArgs a{ix, queries, nq, variance, rot, rot_off, m, qrot, bound};
    if (nq == 0) return hipSuccess;
    dim3 grid((unsigned)((nq + kPrepQ - 1) / kPrepQ), (unsigned)ix.nseg, (unsigned)((max_dim + 255) / 256));
    hipLaunchKernelGGL(query_prep_kernel, grid, dim3(256), 0, st, a);
    hipLaunchKernelGGL(query_bound_kernel, dim3((unsigned)nq, (unsigned)ix.nseg), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_scan(const IndexView &ix, const float *qrot, const float *bound, const uint32_t *probes,
                       int nprobe, const int32_t *npq, int topk, int64_t nq, int dist_type, uint32_t *out_ids,
                       float *out_dists,
                       uint64_t *out_metrics, size_t smem, hipStream_t st, unsigned long long *prof, int cmax) {
    ScanArgs a{ix, qrot, bound, probes, nprobe, npq, topk, nq, out_ids, out_dists, out_metrics, prof, cmax};
    if (ix.ngroups > 32 || ix.nseg > 32 || cmax > 64 || cmax < 1) return hipErrorInvalidValue;
    const void *fn = dist_type == 1 ? (const void *)scan_kernel<1> : (const void *)scan_kernel<0>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    if (dist_type == 1)
        hipLaunchKernelGGL(scan_kernel<1>, dim3((unsigned)nq), dim3(kThreads), smem, st, a);
    else
        hipLaunchKernelGGL(scan_kernel<0>, dim3((unsigned)nq), dim3(kThreads), smem, st, a);
    return hipGetLastError();
}

}  // namespace saqgpu
