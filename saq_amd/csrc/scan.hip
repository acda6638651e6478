// scan.hip -- the plain SAQ IVF scan on gfx950.
//
// Reproduces IVF::search (saqlib/index/ivf.hpp:275-311) for a batch of queries:
//   query_prep_kernel  per (query, segment): the rotated query q_s * P_s in the order of
//                      Eigen's column-major GEMV and the variance bound
//                      m * sqrt(sum var * q^2) (saq_estimator.hpp:28-48, caq_estimator.hpp:56-82)
//   scan_kernel        one workgroup per query walks its probe list in order.  Per probed
//                      cluster it builds every segment's u16 fast-scan LUT in LDS
//                      (Lut::prepare, lut.hpp:60-87), then streams the cluster in steps of
//                      8 blocks (256 vectors, one lane per vector):
//                        1. fast scan: exact integer LUT sums from 8-byte code loads, then the
//                           reference's fp32 epilogue (lut.hpp:89-119, caq_estimator.hpp:116-179)
//                        2. speculative refine: every vector that can still be refined under
//                           the current k-th distance gets its extra-bit inner product
//                           (CodeHelper<b-1>::compute_ip) from 16 lanes, in the reference's
//                           16-lane order and reduction tree
//                        3. exact-order replay by one wavefront: the block/segment pruning of
//                           scanCluster / searchCluster (saq_searcher.hpp:68-214) and
//                           ResultPool insertion (utils/pool.hpp:12-61) in reference order.
//                      The k-th distance only decreases, so every vector the replay refines is
//                      in the speculative set; IDs and distances match the sequential
//                      reference bit for bit.
#include <cfloat>

#include "common.hpp"

#pragma clang fp contract(off)

namespace saqgpu {

constexpr int kThreads = 256;
constexpr int kStepBlocks = 8;            // 256 vectors per step
constexpr int kCandMax = 64;              // candidates refined per replay round
constexpr int kPos[16] = {3, 3, 2, 3, 1, 3, 2, 3, 0, 3, 2, 3, 1, 3, 2, 3};  // fastscan.hpp:22-39

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
    __shared__ float qs[kPrepQ][256];
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
        if (j < D) {
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
    if (blockIdx.z == 0) {
        // bound = sqrt((var.segment * q^2).sum()) * m, Eigen redux order
        const int wave = threadIdx.x >> 6;
        for (int t = wave; t < kPrepQ; t += 4) {
            const int64_t q = q0 + t;
            if (q >= a.nq) continue;
            const float *qq = a.queries + q * ix.num_dim;
            const float v2 = eigen_redux_wave(D, [&](int i) {
                const int d = sd.doff + i;
                const float x = d < ix.num_dim ? qq[d] : 0.0f;
                return a.variance[d] * (x * x);
            });
            if ((threadIdx.x & 63) == 0) a.bound[q * ix.nseg + s] = sqrtf(v2) * a.m;
        }
    }
}

struct ScanArgs {
    IndexView ix;
    const float *qrot;
    const float *bound;
    const uint32_t *probes;
    int32_t nprobe;
    int32_t topk;
    int64_t nq;
    uint32_t *out_ids;
    float *out_dists;
    uint64_t *out_metrics;
};

// Per-segment scalars kept in LDS for the current (query, cluster).
enum {
    SC_DELTA = 0, SC_SUMVL, SC_SUMQCB, SC_QL2, SC_SUMQ, SC_VFAC, SC_IPQC, SC_VL, SC_INV, SC_N = 12
};

struct Smem {
    float *r;
    uint16_t *lut;
    float *cbmin, *cbmax, *sc, *xs, *ipq, *curr, *binfo, *accd;
    int *cand;
    float *pd;
    uint32_t *pid;
    int *ctl;
    unsigned long long *met;
    SegDesc *segs;
};

// control block indices
enum { C_PSIZE = 0, C_NCAND, C_POS, C_LASTV, C_MORE, C_MASK, C_CURSOR, C_DISTK, C_N = 16 };

__device__ __forceinline__ void carve(unsigned char *base, const IndexView &ix, int topk, Smem &m) {
    const int S = ix.nseg;
    float *p = reinterpret_cast<float *>(base);
    m.r = p;
    p += ix.D_pad;
    m.lut = reinterpret_cast<uint16_t *>(p);
    p += ix.ngroups * 16 * 16 / 2;
    m.cbmin = p;
    p += ix.ngroups * 16;
    m.cbmax = p;
    p += ix.ngroups * 16;
    m.sc = p;
    p += S * SC_N;
    m.xs = p;
    p += S * kThreads;
    m.ipq = p;
    p += S * kThreads;
    m.curr = p;
    p += kThreads;
    m.binfo = p;
    p += kStepBlocks * (S + 2);
    m.accd = p;
    p += kCandMax * S;
    m.cand = reinterpret_cast<int *>(p);
    p += kCandMax;
    m.pd = p;
    p += topk + 1;
    m.pid = reinterpret_cast<uint32_t *>(p);
    p += topk + 1;
    m.ctl = reinterpret_cast<int *>(p);
    p += C_N;
    // 8-byte alignment for the counters
    uintptr_t u = reinterpret_cast<uintptr_t>(p);
    u = (u + 15) & ~uintptr_t(15);
    m.met = reinterpret_cast<unsigned long long *>(u);
    m.segs = reinterpret_cast<SegDesc *>(m.met + 4);
}

size_t scan_smem_bytes(const IndexView &ix, int topk) {
    const int S = ix.nseg;
    size_t f = ix.D_pad + ix.ngroups * 16 * 16 / 2 + 2 * ix.ngroups * 16 + S * SC_N + 2 * S * kThreads + kThreads +
               kStepBlocks * (S + 2) + kCandMax * S + kCandMax + 2 * (topk + 1) + C_N;
    return f * 4 + 16 + 4 * 8 + S * sizeof(SegDesc) + 16;
}

// Lane accumulator of CodeHelper<k>::compute_ip (code_helper.hpp:105-591) for lane gl of a
// 16-lane group: dims gl, gl+16, gl+32, ... in increasing order, decoded from the
// reference's compacted layout k (1..8).
__device__ __forceinline__ float lane_acc(int k, const float *r, const uint8_t *y, int D, int gl) {
    float acc = 0.0f;
    switch (k) {
    case 1:
        for (int i = 0; i < D / 16; ++i) {
            const int bit = (y[2 * i + (gl >> 3)] >> (gl & 7)) & 1;
            acc = acc + (bit ? r[16 * i + gl] : 0.0f);
        }
        break;
    case 2:
    case 3: {
        const int stride = k == 2 ? 16 : 24;
        for (int c = 0; c < D / 64; ++c) {
            const int low = y[stride * c + gl];
            const int top = k == 3 ? y[stride * c + 16 + (gl & 7)] : 0;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                int code = (low >> (2 * p)) & 3;
                if (k == 3) code |= ((top >> (2 * p + (gl >> 3))) & 1) << 2;
                acc = __fmaf_rn(r[64 * c + 16 * p + gl], (float)code, acc);
            }
        }
        break;
    }
    case 4:
        for (int c = 0; c < D / 32; ++c) {
            const int b = y[16 * c + gl];
            acc = __fmaf_rn((float)(b & 15), r[32 * c + gl], acc);
            acc = __fmaf_rn((float)(b >> 4), r[32 * c + 16 + gl], acc);
        }
        break;
    case 6:
    case 7: {
        const int stride = k == 6 ? 48 : 56;
        for (int c = 0; c < D / 64; ++c) {
            const int base = stride * c;
            const int c1 = y[base + gl], c2 = y[base + 16 + gl], c3 = y[base + 32 + gl];
            const int top = k == 7 ? y[base + 48 + (gl & 7)] : 0;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                int code;
                if (p == 0) code = c1 & 0x3F;
                else if (p == 1) code = c2 & 0x3F;
                else if (p == 2) code = ((c1 >> 2) & 0x30) | (c3 & 0x0F);
                else code = ((c2 >> 2) & 0x30) | ((c3 >> 4) & 0x0F);
                if (k == 7) code |= ((top >> (2 * p + (gl >> 3))) & 1) << 6;
                acc = __fmaf_rn(r[64 * c + 16 * p + gl], (float)code, acc);
            }
        }
        break;
    }
    case 8:
        for (int i = 0; i < D / 16; ++i) acc = __fmaf_rn(r[16 * i + gl], (float)y[16 * i + gl], acc);
        break;
    default:
        break;
    }
    return acc;
}

// CodeHelper<k>::compute_ip for k = 1..12 by a 16-lane group; valid in every lane of it.
__device__ __forceinline__ float ex_ip_group(int k, const float *r, const uint8_t *y, int D, int gl) {
    if (k == 5) {  // 2 * IP4 + IP1 (code_helper.hpp:372-374)
        const float a = tree16_xor(lane_acc(4, r, y, D, gl));
        const float b = tree16_xor(lane_acc(1, r, y + D * 4 / 8, D, gl));
        return 2 * a + b;
    }
    if (k > 8) {  // IP8 + 256 * IP_{k-8} (code_helper.hpp:85-88)
        const float lo = tree16_xor(lane_acc(8, r, y, D, gl));
        float hi;
        if (k - 8 == 5) {
            const float a = tree16_xor(lane_acc(4, r, y + D, D, gl));
            const float b = tree16_xor(lane_acc(1, r, y + D + D * 4 / 8, D, gl));
            hi = 2 * a + b;
        } else {
            hi = tree16_xor(lane_acc(k - 8, r, y + D, D, gl));
        }
        return lo + 256 * hi;
    }
    return tree16_xor(lane_acc(k, r, y, D, gl));
}

// ResultPool::insert by one wavefront (utils/pool.hpp:233-246).  u, d uniform.
template <int DT>
__device__ __forceinline__ void pool_insert(Smem &m, int K, uint32_t u, float d, int lane) {
    if (DT == 1) d = -d;
    const int size = m.ctl[C_PSIZE];
    if (size == K && d > m.pd[size - 1]) return;
    int lo = 0;
    for (int base = 0; base < size; base += 64) {
        const int i = base + lane;
        const bool p = (i < size) && (m.pd[i] < d);
        lo += __popcll(__ballot(p));
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
        m.ctl[C_PSIZE] = size + (size < K);
    }
    __builtin_amdgcn_wave_barrier();
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
    const int S = ix.nseg, K = a.topk;
    Smem m;
    carve(smem, ix, K, m);
    for (int i = tid; i < S; i += kThreads) m.segs[i] = ix.segs[i];
    for (int i = tid; i <= K; i += kThreads) {
        m.pd[i] = 0.0f;
        m.pid[i] = 0;
    }
    if (tid < C_N) m.ctl[tid] = 0;
    if (tid < 4) m.met[tid] = 0;
    __syncthreads();
    const bool single = (S == 1);
    const float *qrot = a.qrot + q * ix.D_pad;
    const float *bound = a.bound + q * S;
    const int ncb = ix.ngroups * 16;

    for (int pi = 0; pi < a.nprobe; ++pi) {
        const uint32_t cid = a.probes[q * a.nprobe + pi];
        if (cid >= (uint32_t)ix.num_cen) continue;  // invalid ids are rejected on the host
        const int n = ix.clu_n[cid];
        if (n == 0) continue;
        const int nb = (n + kBlock - 1) / kBlock;
        const int64_t blk0 = ix.clu_blk[cid], vec0 = ix.clu_vec[cid];
        const float *crot = ix.crot + (int64_t)cid * ix.D_pad;

        // ---- CaqCluEstimator::prepare for every segment (caq_estimator.hpp:94-106)
        for (int i = tid; i < ix.D_pad; i += kThreads) m.r[i] = DT == 1 ? qrot[i] : qrot[i] - crot[i];
        __syncthreads();
        for (int s = wave; s < S; s += 4) {
            const SegDesc sd = m.segs[s];
            const float *rs = m.r + sd.doff;
            const float sum_q = eigen_redux_wave(sd.dim, [&](int i) { return rs[i]; });
            const float ql2 = eigen_redux_wave(sd.dim, [&](int i) { return rs[i] * rs[i]; });
            float ipqc = 0.0f;
            if (DT == 1) ipqc = eigen_redux_wave(sd.dim, [&](int i) { return qrot[sd.doff + i] * crot[sd.doff + i]; });
            if (lane == 0) {
                float *sc = m.sc + s * SC_N;
                const float ql2n = sqrtf(ql2);
                sc[SC_SUMQ] = sum_q;
                sc[SC_QL2] = ql2;
                sc[SC_SUMQCB] = (float)(0.5 * (double)sum_q - (double)(0.58f * ql2n));
                sc[SC_IPQC] = ipqc;
                sc[SC_VFAC] = DT == 1 ? ipqc - bound[s] : ql2 - 2 * bound[s];
            }
        }
        // LUT pass 1: per-codebook min / max of the fp32 table (fastscan.hpp:236-248)
        for (int c = tid; c < ncb; c += kThreads) {
            const int g = c >> 4;
            int s = 0;
            while (!(g >= m.segs[s].group_off && g < m.segs[s].group_off + m.segs[s].ngroups)) ++s;
            const float *rr = m.r + m.segs[s].doff + (g - m.segs[s].group_off) * 64 + (c & 15) * 4;
            float l[16];
            l[0] = 0;
            float mn = 0.0f, mx = 0.0f;
#pragma unroll
            for (int j = 1; j < 16; ++j) {
                l[j] = l[j - (j & -j)] + rr[kPos[j]];
                mn = fminf(mn, l[j]);
                mx = fmaxf(mx, l[j]);
            }
            m.cbmin[c] = mn;
            m.cbmax[c] = mx;
        }
        __syncthreads();
        for (int s = wave; s < S; s += 4) {
            const SegDesc sd = m.segs[s];
            if (sd.bits == 0) continue;
            float mn = FLT_MAX, mx = -FLT_MAX;
            for (int c = sd.group_off * 16 + lane; c < (sd.group_off + sd.ngroups) * 16; c += 64) {
                mn = fminf(mn, m.cbmin[c]);
                mx = fmaxf(mx, m.cbmax[c]);
            }
            for (int o = 32; o >= 1; o >>= 1) {
                mn = fminf(mn, __shfl_xor(mn, o));
                mx = fmaxf(mx, __shfl_xor(mx, o));
            }
            if (lane == 0) {
                float *sc = m.sc + s * SC_N;
                const float delta = (float)((double)(mx - mn) / ((1 << 16) - 0.01));
                sc[SC_DELTA] = delta;
                sc[SC_INV] = 1.0f / delta;  // the release build's hoisted reciprocal (lut.hpp:464)
                sc[SC_VL] = mn;
                sc[SC_SUMVL] = (float)(((double)mn + 0.5 * (double)delta) * (double)(float)(sd.dim / 4));
            }
        }
        __syncthreads();
        // LUT pass 2: quantise to u16 (lut.hpp:461-465)
        for (int c = tid; c < ncb; c += kThreads) {
            const int g = c >> 4;
            int s = 0;
            while (!(g >= m.segs[s].group_off && g < m.segs[s].group_off + m.segs[s].ngroups)) ++s;
            const float *rr = m.r + m.segs[s].doff + (g - m.segs[s].group_off) * 64 + (c & 15) * 4;
            const float vl = m.sc[s * SC_N + SC_VL], inv = m.sc[s * SC_N + SC_INV];
            float l[16];
            l[0] = 0;
#pragma unroll
            for (int j = 1; j < 16; ++j) l[j] = l[j - (j & -j)] + rr[kPos[j]];
            uint32_t *dst = reinterpret_cast<uint32_t *>(m.lut + c * 16);
#pragma unroll
            for (int j = 0; j < 16; j += 2)
                dst[j / 2] = cvt_u16((l[j] - vl) * inv) | (cvt_u16((l[j + 1] - vl) * inv) << 16);
        }
        __syncthreads();
        if (tid == 0) m.met[2] += (unsigned long long)nb * kBlock;  // total_comp_cnt

        // ---- blocks, 8 at a time
        for (int st = 0; st < nb; st += kStepBlocks) {
            const int nbs = min(kStepBlocks, nb - st);
            const int v = tid, bl = v >> 5, ln = v & 31;
            const bool active = bl < nbs;
            const uint8_t *rec = ix.blocks + (blk0 + st + bl) * ix.block_stride;
            if (active) {
                const float *fac = reinterpret_cast<const float *>(rec);
                for (int s = 0; s < S; ++s) m.xs[s * kThreads + v] = fac[s * kBlock + ln];
                const uint8_t *codes = rec + ix.fac_floats * 4;
                for (int s = 0; s < S; ++s) {
                    const SegDesc sd = m.segs[s];
                    if (sd.bits == 0) continue;
                    uint32_t acc = 0;
                    for (int g = sd.group_off; g < sd.group_off + sd.ngroups; ++g) {
                        const uint2 w = *reinterpret_cast<const uint2 *>(codes + (int64_t)g * 256 + ln * 8);
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
                    }
                    const float *sc = m.sc + s * SC_N;
                    float t = (float)(int32_t)acc * sc[SC_DELTA];  // lut.hpp:495-497
                    t = t + sc[SC_SUMVL];
                    m.ipq[s * kThreads + v] = t;
                }
            }
            // per-vector estimates (distk independent) and per-block minima
            {
                float cur = 0.0f;
                float m1 = FLT_MAX;
                if (single) {
                    const SegDesc sd = m.segs[0];
                    const float *sc = m.sc;
                    const float x = active ? m.xs[v] : 0.0f;
                    if (sd.bits == 0) {
                        cur = DT == 1 ? sc[SC_VFAC] : max0(x * x + sc[SC_VFAC]);
                    } else {
                        const float ipq = active ? m.ipq[v] : 0.0f;
                        float f = (ipq - sc[SC_SUMQCB]) * sd.rescale_sqrtD;
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
                    m1 = cur;
                    for (int o = 16; o >= 1; o >>= 1) m1 = fminf(m1, __shfl_xor(m1, o, 32));
                    float M = m1;
                    for (int s = 0; s < S; ++s) {
                        const SegDesc sd = m.segs[s];
                        if (sd.bits == 0) continue;
                        const float *sc = m.sc + s * SC_N;
                        const float x = active ? m.xs[s * kThreads + v] : 0.0f;
                        const float vd = DT == 1 ? sc[SC_VFAC] : max0(x * x + sc[SC_VFAC]);
                        const float ipq = active ? m.ipq[s * kThreads + v] : 0.0f;
                        float f = (ipq - sc[SC_SUMQCB]) * sd.rescale_sqrtD;
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
                // ordered compaction of flagged vectors (wave-major, lane order)
                const unsigned long long bal = __ballot(flag);
                __shared__ int wcnt[4];
                if (lane == 0) wcnt[wave] = __popcll(bal);
                __syncthreads();
                int before = 0, total = 0;
                for (int w = 0; w < 4; ++w) {
                    if (w < wave) before += wcnt[w];
                    total += wcnt[w];
                }
                const int rank = before + __popcll(bal & ((1ull << lane) - 1));
                if (flag && rank < kCandMax) m.cand[rank] = v;
                const int ncand = min(total, kCandMax);
                __syncthreads();
                // extra-bit inner products, 16 lanes per (candidate, segment)
                {
                    const int grp = tid >> 4, gl = tid & 15;
                    for (int item = grp; item < ncand * S; item += kThreads / 16) {
                        const int c = item / S, s = item - c * S;
                        const SegDesc sd = m.segs[s];
                        if (sd.bits < 2) continue;
                        const int vv = m.cand[c];
                        const int64_t gvec = vec0 + (int64_t)(st * kBlock + vv);
                        const uint8_t *y = ix.long_codes + gvec * ix.long_stride + sd.long_off;
                        const float ex = ex_ip_group(sd.bits - 1, m.r + sd.doff, y, sd.dim, gl);
                        if (gl == 0) {
                            const float *sc = m.sc + s * SC_N;
                            const float ipq = m.ipq[s * kThreads + vv];
                            const float x = m.xs[s * kThreads + vv];
                            // Lut::getExtIP (lut.hpp:508-513) and compAccurateDist (caq_estimator.hpp:190-217)
                            const float t = (float)(((double)ipq + (double)ex * (double)sd.sqd) +
                                                    (-1.0 + (double)(sd.sqd / 2)) * (double)sc[SC_SUMQ]);
                            const float resc = ix.exfac[(gvec * S + s) * 2];
                            const float ipoq = resc * t;
                            m.accd[c * S + s] = DT == 1 ? ipoq + sc[SC_IPQC] : x * x + sc[SC_QL2] - 2 * ipoq;
                        }
                    }
                }
                __syncthreads();
                // exact-order replay by wavefront 0
                if (wave == 0) {
                    const int lastv = total > kCandMax ? m.cand[kCandMax - 1] : kThreads;
                    int cur_pos = pos;
                    bool prog = inprog;
                    uint32_t mask_state = inmask;
                    float dk = distk;
                    int ci = 0;
                    bool need_more = false;
                    unsigned long long fast_bits = 0, acc_bits = 0;
                    for (int b = pblk; b < nbs && !need_more; ++b) {
                        const int j0 = (b == pblk) ? (pos & 31) : 0;
                        const bool fresh = !(b == pblk && prog);
                        if (single) {
                            uint32_t msk;
                            const SegDesc sd = m.segs[0];
                            if (fresh) {
                                fast_bits += (unsigned long long)kBlock * sd.dim * (sd.bits ? 1 : 0);
                                const int cnt = (st + b == nb - 1) ? n % kBlock : kBlock;
                                const bool lt = (lane < 32) && (m.curr[b * 32 + (lane & 31)] < dk);
                                msk = (uint32_t)__ballot(lt) & (uint32_t)((1ull << cnt) - 1);
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
                                const int64_t gvec = vec0 + (int64_t)(st * kBlock + vv);
                                float d;
                                const float x = m.xs[vv];
                                const float *sc = m.sc;
                                if (sd.bits == 0) {
                                    d = DT == 1 ? sc[SC_IPQC] : x * x + sc[SC_QL2];
                                } else if (sd.bits == 1) {
                                    const float t = (float)(((double)m.ipq[vv] + 0.0 * (double)sd.sqd) +
                                                            (-1.0 + (double)(sd.sqd / 2)) * (double)sc[SC_SUMQ]);
                                    const float ipoq = ix.exfac[gvec * 2] * t;
                                    d = DT == 1 ? ipoq + sc[SC_IPQC] : x * x + sc[SC_QL2] - 2 * ipoq;
                                } else {
                                    while (ci < ncand - 1 && m.cand[ci] < vv) ++ci;
                                    d = m.accd[ci * S];
                                    acc_bits += (unsigned long long)sd.dim * (sd.bits - 1);
                                }
                                pool_insert<DT>(m, K, ix.ids[gvec], d, lane);
                                dk = pool_distk<DT>(m, K);
                            }
                            if (!need_more) prog = false;
                        } else {
                            if (fresh) {
                                const float *bi = m.binfo + b * (S + 2);
                                if (bi[0] > dk) continue;  // stage 1 skip
                                for (int s = 0; s < S; ++s) {
                                    if (m.segs[s].bits == 0) continue;
                                    fast_bits += (unsigned long long)kBlock * m.segs[s].dim;
                                    if (bi[2 + s] > dk) break;
                                }
                                if (bi[1] > dk) continue;  // stage 2 broke: no refine
                            }
                            for (int j = j0; j < 32; ++j) {
                                const int vv = b * 32 + j;
                                if (!(m.curr[vv] < dk)) continue;
                                if ((st + b) * kBlock + j >= n) break;
                                if (vv > lastv) {
                                    need_more = true;
                                    cur_pos = vv;
                                    prog = true;
                                    break;
                                }
                                while (ci < ncand - 1 && m.cand[ci] < vv) ++ci;
                                const int64_t gvec = vec0 + (int64_t)(st * kBlock + vv);
                                float acc = m.curr[vv];
                                for (int s = 0; s < S; ++s) {
                                    const SegDesc sd = m.segs[s];
                                    const float *sc = m.sc + s * SC_N;
                                    const float x = m.xs[s * kThreads + vv];
                                    float av, cd;
                                    if (sd.bits == 0) {
                                        av = DT == 1 ? sc[SC_IPQC] : x * x + sc[SC_QL2];
                                        cd = DT == 1 ? sc[SC_VFAC] : max0(x * x + sc[SC_VFAC]);
                                    } else {
                                        const float ipq = m.ipq[s * kThreads + vv];
                                        float f = (ipq - sc[SC_SUMQCB]) * sd.rescale_sqrtD;
                                        f = f * x;
                                        cd = DT == 1 ? f * 0.5f + sc[SC_IPQC] : max0(x * x + (sc[SC_QL2] - f));
                                        if (sd.bits == 1) {
                                            const float t = (float)(((double)ipq + 0.0 * (double)sd.sqd) +
                                                                    (-1.0 + (double)(sd.sqd / 2)) * (double)sc[SC_SUMQ]);
                                            const float ipoq = ix.exfac[(gvec * S + s) * 2] * t;
                                            av = DT == 1 ? ipoq + sc[SC_IPQC] : x * x + sc[SC_QL2] - 2 * ipoq;
                                        } else {
                                            av = m.accd[ci * S + s];
                                            acc_bits += (unsigned long long)sd.dim * (sd.bits - 1);
                                        }
                                    }
                                    acc += av - cd;
                                    if (acc >= dk) break;
                                }
                                pool_insert<DT>(m, K, ix.ids[gvec], acc, lane);
                                dk = pool_distk<DT>(m, K);
                            }
                            if (!need_more) prog = false;
                        }
                    }
                    if (lane == 0) {
                        m.met[0] += fast_bits;
                        m.met[1] += acc_bits;
                        m.ctl[C_POS] = need_more ? cur_pos : kThreads;
                        m.ctl[C_MASK] = need_more ? (int)mask_state : 0;
                        m.ctl[C_CURSOR] = need_more && prog ? 1 : 0;
                    }
                }
                __syncthreads();
                if (m.ctl[C_POS] >= kThreads) break;
                __syncthreads();
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < K; i += kThreads) {
        a.out_ids[q * K + i] = m.pid[i];
        if (a.out_dists) a.out_dists[q * K + i] = DT == 1 ? -m.pd[i] : m.pd[i];
    }
    if (a.out_metrics && tid < 3) a.out_metrics[q * 3 + tid] = m.met[tid];
}

}  // namespace saqgpu

// ---------------------------------------------------------------- launchers
namespace saqgpu {

hipError_t launch_query_prep(const IndexView &ix, const float *queries, int64_t nq, const float *variance,
                             const float *rot, const int64_t *rot_off, float m, float *qrot, float *bound,
                             int max_dim, hipStream_t st) {
    PrepArgs a{ix, queries, nq, variance, rot, rot_off, m, qrot, bound};
    dim3 grid((unsigned)((nq + kPrepQ - 1) / kPrepQ), (unsigned)ix.nseg, (unsigned)((max_dim + 255) / 256));
    hipLaunchKernelGGL(query_prep_kernel, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_scan(const IndexView &ix, const float *qrot, const float *bound, const uint32_t *probes,
                       int nprobe, int topk, int64_t nq, int dist_type, uint32_t *out_ids, float *out_dists,
                       uint64_t *out_metrics, size_t smem, hipStream_t st) {
    ScanArgs a{ix, qrot, bound, probes, nprobe, topk, nq, out_ids, out_dists, out_metrics};
    if (dist_type == 1) {
        hipFuncSetAttribute((const void *)scan_kernel<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        hipLaunchKernelGGL(scan_kernel<1>, dim3((unsigned)nq), dim3(kThreads), smem, st, a);
    } else {
        hipFuncSetAttribute((const void *)scan_kernel<0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        hipLaunchKernelGGL(scan_kernel<0>, dim3((unsigned)nq), dim3(kThreads), smem, st, a);
    }
    return hipGetLastError();
}

}  // namespace saqgpu
