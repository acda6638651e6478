// single.hip -- the single-vector estimator on gfx950: SaqSingleEstimator over records in the
// SaqSingleDataWrapper layout (quantization/saq_estimator.hpp:248-294,
// caq/caq_estimator.hpp:220-369, 450-497, quantization/single_data.hpp:156-320).
//
//   single_prep_kernel  one wavefront per (query, segment): CaqEstimatorSingleImpl::prepare on
//                       the rotated segment query (:292-308): |q|^2 and sum_q in Eigen's redux
//                       order, min/max, the 8-bit quantised query and its 8 bit planes per 64
//                       dims (utils::new_transpose_bin, utils/space.hpp:266-293) taken by
//                       ballots, bit 63 - k = dim k.
//   single_est_kernel   a 16-lane group per (query, record), 16 records per workgroup: the
//                       record's short code against the planes by popcount (warmup_ip_x0_q,
//                       space.hpp:126-226), the masked sum of the query over the code's set bits
//                       (mask_ip_x0_q, :228-265) and the extra-bit inner product
//                       (CodeHelper<b-1>::compute_ip) on the 16-lane grouping the reference's
//                       AVX-512 kernels use, then varsEstDist / compFastDist / compAccurateDist
//                       summed over segments in order.
#include <cfloat>

#include "common.hpp"
#include "estimators.hpp"

#pragma clang fp contract(off)

namespace saqgpu {

// Per (query, segment) state of CaqEstimatorSingleImpl.
struct SingleQ {
    float q_l2sqr, q_l2norm, sum_q, delta, vlh, pad_[3];
};

struct SinglePrepArgs {
    IndexView ix;
    const float *qrot;  // nq x D_pad
    int64_t nq;
    SingleQ *qs;        // nq x nseg
    uint64_t *planes;   // nq x (D_pad / 64 * 8)
};

// grid (nq, nseg), 64 threads
__global__ __launch_bounds__(64) void single_prep_kernel(SinglePrepArgs a) {
    const IndexView &ix = a.ix;
    const int64_t q = blockIdx.x;
    const int s = blockIdx.y;
    const int lane = threadIdx.x;
    const SegDesc sd = ix.segs[s];
    const int D = sd.dim;
    const float *qr = a.qrot + q * ix.D_pad + sd.doff;
    const float l2 = eigen_redux_wave(D, [&](int i) { return qr[i] * qr[i]; });
    const float sum = eigen_redux_wave(D, [&](int i) { return qr[i]; });
    float vl = qr[lane], vr = qr[lane];
    for (int i = lane + 64; i < D; i += 64) {
        vl = fminf(vl, qr[i]);
        vr = fmaxf(vr, qr[i]);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        vl = fminf(vl, __shfl_xor(vl, o));
        vr = fmaxf(vr, __shfl_xor(vr, o));
    }
    // (q_vr_ - q_vl_) / ((1 << 8) - 0.01): a float difference divided in double, kept as float
    const float delta = (float)((double)(vr - vl) / ((1 << 8) - 0.01));
    uint64_t *pl = a.planes + q * (ix.D_pad / 64 * 8) + sd.doff / 64 * 8;
    for (int b = 0; b < D / 64; ++b) {
        // query_sq_ = ((q - vl) / delta).cast<uint16_t>(), as x86 truncates (NaN when delta is 0 -> 0)
        const uint32_t v = cvt_u16((qr[64 * b + lane] - vl) / delta);
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const uint64_t m = __builtin_bitreverse64(__ballot((v >> p) & 1u));
            if (lane == p) pl[8 * b + p] = m;
        }
    }
    if (lane == 0) {
        SingleQ st;
        st.q_l2sqr = l2;
        st.q_l2norm = sqrtf(l2);
        st.sum_q = sum;
        st.delta = delta;
        st.vlh = (float)((double)vl + 0.5 * (double)delta);  // q_vl_ + 0.5 * delta_ as a float argument
        st.pad_[0] = st.pad_[1] = st.pad_[2] = 0.0f;
        a.qs[q * ix.nseg + s] = st;
    }
}

constexpr size_t kSingleStageBytes = 48 * 1024;  // LDS of the staged form (records + 4 queries)

struct SingleEstArgs {
    IndexView ix;
    const float *qrot;       // nq x D_pad
    const float *bound;      // nq x nseg
    const SingleQ *qs;       // nq x nseg
    const uint64_t *planes;  // nq x (D_pad / 64 * 8)
    const uint8_t *recs;
    int64_t stride, nrec, nq;
    const int32_t *soff;     // per segment: short-code offset, long-code offset (-1: none)
    int32_t exf_off;
    int32_t dist_type;
    float *out_vars, *out_fast, *out_acc;  // nq x nrec each (vars, fast may be null)
};

constexpr int kSRec = 16;  // records per workgroup (one per 16-lane group)
constexpr int kSQ = 16;    // queries per workgroup of the staged form

// varsEstDist, compFastDist and compAccurateDist of record rec (any address space) for query q,
// by the 16-lane group of lane gl, summed over segments; qr and pl are the query's rotated
// vector and bit planes.  Valid in every lane of the group.
__device__ __forceinline__ float3 single_pair(const SingleEstArgs &a, const float *qr, const uint64_t *pl,
                                              const uint8_t *rec, int64_t q, int gl) {
    const IndexView &ix = a.ix;
    const float *fac = reinterpret_cast<const float *>(rec);
    const float *exf = reinterpret_cast<const float *>(rec + a.exf_off);
    const bool ip = a.dist_type == 1;
    float rv = 0.0f, rf = 0.0f, ra = 0.0f;
    for (int s = 0; s < ix.nseg; ++s) {
        const SegDesc sd = ix.segs[s];
        const SingleQ st = a.qs[q * ix.nseg + s];
        const float bound = a.bound[q * ix.nseg + s];
        const float o = fac[2 * s];
        float v;
        if (ip) {
            v = 0.0f - bound;  // ip_q_c_ - bound; ip_q_c_ stays 0 for single records
        } else {
            const float t = o * o + st.q_l2sqr - 2 * bound;
            v = (0.0f < t) ? t : 0.0f;  // std::max(0.0f, t)
        }
        float f = v, acc;
        const float o2 = o * o;
        if (sd.bits == 0) {
            acc = ip ? 0.0f : o2 + st.q_l2sqr;
        } else {
            const uint64_t *x = reinterpret_cast<const uint64_t *>(rec + a.soff[2 * s]);
            const uint64_t *ps = pl + sd.doff / 64 * 8;
            const float *qs = qr + sd.doff;
            const int nb = sd.dim / 64;
            // warmup_ip_x0_q: exact integer sums, any order
            uint32_t ipi = 0, ppc = 0;
            for (int b = gl; b < nb; b += 16) {
                const uint64_t w = x[b];
                ppc += (uint32_t)__popcll(w);
#pragma unroll
                for (int p = 0; p < 8; ++p) ipi += (uint32_t)__popcll(w & ps[8 * b + p]) << p;
            }
#pragma unroll
            for (int o2x = 8; o2x >= 1; o2x >>= 1) {
                ipi += __shfl_xor(ipi, o2x, 16);
                ppc += __shfl_xor(ppc, o2x, 16);
            }
            const float tmp = (st.delta * (float)ipi) + (st.vlh * (float)ppc);
            // (tmp - (0.5 * sum_q - 0.58f * |q|)) * (4 / 0.8f * 1/sqrt(D)) * |o|, in double
            const double cst = 0.5 * (double)st.sum_q - (double)(0.58f * st.q_l2norm);
            const float q1 = (float)((((double)tmp - cst) * (double)sd.rescale_sqrtD) * (double)o);
            if (ip) {
                f = (float)((double)q1 * 0.5);
            } else {
                const float t = st.q_l2sqr + o2 - q1;
                f = (t < 0.0f) ? 0.0f : t;  // std::max(t, 0.0f)
            }
            // mask_ip_x0_q: lane gl takes dims 64 b + 16 m + gl; the release build adds a
            // block's four masked values as (m0 + m1) + (m2 + m3)
            float mk = 0.0f;
            for (int b = 0; b < nb; ++b) {
                const uint64_t w = x[b];
                float mv[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int k = 16 * m + gl;
                    mv[m] = ((w >> (63 - k)) & 1u) ? qs[64 * b + k] : 0.0f;
                }
                mk = mk + ((mv[0] + mv[1]) + (mv[2] + mv[3]));
            }
            const float ipo1 = tree16_xor(mk);
            const double ex = sd.bits > 1 ? (double)ex_ip_group(sd.bits - 1, qs, rec + a.soff[2 * s + 1], sd.dim, gl)
                                          : 0.0;
            const double cd = 2.0 / (1 << sd.bits);
            const float tt = (float)((double)ipo1 + ex * cd + (-1.0 + cd / 2) * (double)st.sum_q);
            const float ipoq = exf[2 * s] * tt;
            acc = ip ? ipoq + 0.0f : o2 + st.q_l2sqr - 2 * ipoq;
        }
        rv += v;
        rf += f;
        ra += acc;
    }
    return make_float3(rv, rf, ra);
}

// grid nq * ceil(nrec / 16), 256 threads, records read in place (records too large to stage):
// workgroup w takes query w % nq of record tile w / nq, so the workgroups of one tile run back
// to back and its records come from HBM once and from L2 for the other queries.
__global__ __launch_bounds__(256) void single_est_kernel(SingleEstArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const IndexView &ix = a.ix;
    const int64_t q = (int64_t)blockIdx.x % a.nq, tile = (int64_t)blockIdx.x / a.nq;
    const int tid = threadIdx.x, grp = tid >> 4, gl = tid & 15;
    float *qr = reinterpret_cast<float *>(smem);
    uint64_t *pl = reinterpret_cast<uint64_t *>(smem + (size_t)ix.D_pad * 4);
    const int npl = ix.D_pad / 64 * 8;
    for (int i = tid; i < ix.D_pad; i += 256) qr[i] = a.qrot[q * ix.D_pad + i];
    for (int i = tid; i < npl; i += 256) pl[i] = a.planes[q * npl + i];
    __syncthreads();
    const int64_t r = tile * kSRec + grp;
    if (r >= a.nrec) return;
    const float3 e = single_pair(a, qr, pl, a.recs + r * a.stride, q, gl);
    if (gl == 0) {
        const int64_t i = q * a.nrec + r;
        if (a.out_vars) a.out_vars[i] = e.x;
        if (a.out_fast) a.out_fast[i] = e.y;
        a.out_acc[i] = e.z;
    }
}

constexpr int kQN = 4;  // queries evaluated together per record pass of the staged form

template <int N>
struct FQ {
    float v[N];
};

// lane_acc_k (the 16-lane accumulator of CodeHelper<K>::compute_ip) for N queries at once: each
// dim's code is decoded once and fed to every query's own FMA chain, in the same order.
template <int K, int N>
__device__ __forceinline__ FQ<N> lane_acc_n(const float *r, int rstride, const uint8_t *y, int D, int gl) {
    FQ<N> acc;
#pragma unroll
    for (int t = 0; t < N; ++t) acc.v[t] = 0.0f;
    const int n = D / 16;  // a multiple of 4
    for (int i0 = 0; i0 < n; i0 += 4) {
        int c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = lane_code<K>(y, i0 + u, gl);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int t = 0; t < N; ++t) {
                const float x = r[t * rstride + 16 * (i0 + u) + gl];
                acc.v[t] = K == 1 ? acc.v[t] + (c[u] ? x : 0.0f) : __fmaf_rn(x, (float)c[u], acc.v[t]);
            }
    }
    return acc;
}

template <int N>
__device__ __forceinline__ FQ<N> lane_acc_nk(int k, const float *r, int rstride, const uint8_t *y, int D, int gl) {
    switch (k) {
    case 1: return lane_acc_n<1, N>(r, rstride, y, D, gl);
    case 2: return lane_acc_n<2, N>(r, rstride, y, D, gl);
    case 3: return lane_acc_n<3, N>(r, rstride, y, D, gl);
    case 4: return lane_acc_n<4, N>(r, rstride, y, D, gl);
    case 6: return lane_acc_n<6, N>(r, rstride, y, D, gl);
    case 7: return lane_acc_n<7, N>(r, rstride, y, D, gl);
    default: return lane_acc_n<8, N>(r, rstride, y, D, gl);
    }
}

// ex_ip_group for N queries (query t's segment vector at r + t * rstride): k = 5 is
// 2 * IP4 + IP1, k > 8 is IP8 + 256 * IP_{k-8}, each part reduced by the 16-lane tree.
template <int N>
__device__ __forceinline__ FQ<N> ex_ip_n(int k, const float *r, int rstride, const uint8_t *y, int D, int gl) {
    FQ<N> out;
    const int klo = k > 8 ? 8 : k;
    if (klo == 5) {
        const FQ<N> a4 = lane_acc_nk<N>(4, r, rstride, y, D, gl);
        const FQ<N> a1 = lane_acc_nk<N>(1, r, rstride, y + D * 4 / 8, D, gl);
#pragma unroll
        for (int t = 0; t < N; ++t) out.v[t] = 2 * tree16_xor(a4.v[t]) + tree16_xor(a1.v[t]);
    } else {
        const FQ<N> a = lane_acc_nk<N>(klo, r, rstride, y, D, gl);
#pragma unroll
        for (int t = 0; t < N; ++t) out.v[t] = tree16_xor(a.v[t]);
    }
    if (k > 8) {
        const int kh = k - 8;
        FQ<N> hi;
        if (kh == 5) {
            const FQ<N> a4 = lane_acc_nk<N>(4, r, rstride, y + D, D, gl);
            const FQ<N> a1 = lane_acc_nk<N>(1, r, rstride, y + D + D * 4 / 8, D, gl);
#pragma unroll
            for (int t = 0; t < N; ++t) hi.v[t] = 2 * tree16_xor(a4.v[t]) + tree16_xor(a1.v[t]);
        } else {
            const FQ<N> a = lane_acc_nk<N>(kh, r, rstride, y + D, D, gl);
#pragma unroll
            for (int t = 0; t < N; ++t) hi.v[t] = tree16_xor(a.v[t]);
        }
#pragma unroll
        for (int t = 0; t < N; ++t) out.v[t] = out.v[t] + 256 * hi.v[t];
    }
    return out;
}

// single_pair for kQN queries q0 .. q0 + kQN - 1 (rows t of qr / pl, vectors D_pad and planes
// npl apart): the record's words and codes are read once for all of them.
struct Est3 {
    float v[kQN], f[kQN], a[kQN];
};

__device__ __forceinline__ Est3 single_pair_n(const SingleEstArgs &a, const float *qr, const uint64_t *pl, int npl,
                                             const uint8_t *rec, int64_t q0, int gl) {
    const IndexView &ix = a.ix;
    const float *fac = reinterpret_cast<const float *>(rec);
    const float *exf = reinterpret_cast<const float *>(rec + a.exf_off);
    const bool ip = a.dist_type == 1;
    Est3 r;
#pragma unroll
    for (int t = 0; t < kQN; ++t) r.v[t] = r.f[t] = r.a[t] = 0.0f;
    for (int s = 0; s < ix.nseg; ++s) {
        const SegDesc sd = ix.segs[s];
        const float o = fac[2 * s];
        const float o2 = o * o;
        SingleQ st[kQN];
        float v[kQN];
#pragma unroll
        for (int t = 0; t < kQN; ++t) {
            const int64_t q = min<int64_t>(q0 + t, a.nq - 1);
            st[t] = a.qs[q * ix.nseg + s];
            const float bound = a.bound[q * ix.nseg + s];
            if (ip) {
                v[t] = 0.0f - bound;
            } else {
                const float x = o * o + st[t].q_l2sqr - 2 * bound;
                v[t] = (0.0f < x) ? x : 0.0f;
            }
        }
        if (sd.bits == 0) {
#pragma unroll
            for (int t = 0; t < kQN; ++t) {
                r.v[t] += v[t];
                r.f[t] += v[t];
                r.a[t] += ip ? 0.0f : o2 + st[t].q_l2sqr;
            }
            continue;
        }
        const uint64_t *x = reinterpret_cast<const uint64_t *>(rec + a.soff[2 * s]);
        const int nb = sd.dim / 64;
        const int pof = sd.doff / 64 * 8;
        uint32_t ipi[kQN], ppc = 0;
#pragma unroll
        for (int t = 0; t < kQN; ++t) ipi[t] = 0;
        for (int b = gl; b < nb; b += 16) {
            const uint64_t w = x[b];
            ppc += (uint32_t)__popcll(w);
#pragma unroll
            for (int t = 0; t < kQN; ++t)
#pragma unroll
                for (int p = 0; p < 8; ++p) ipi[t] += (uint32_t)__popcll(w & pl[t * npl + pof + 8 * b + p]) << p;
        }
#pragma unroll
        for (int o2x = 8; o2x >= 1; o2x >>= 1) {
            ppc += __shfl_xor(ppc, o2x, 16);
#pragma unroll
            for (int t = 0; t < kQN; ++t) ipi[t] += __shfl_xor(ipi[t], o2x, 16);
        }
        float mk[kQN];
#pragma unroll
        for (int t = 0; t < kQN; ++t) mk[t] = 0.0f;
        const float *qs = qr + sd.doff;
        for (int b = 0; b < nb; ++b) {
            const uint64_t w = x[b];
            bool on[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) on[m] = (w >> (63 - (16 * m + gl))) & 1u;
#pragma unroll
            for (int t = 0; t < kQN; ++t) {
                float mv[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) mv[m] = on[m] ? qs[t * ix.D_pad + 64 * b + 16 * m + gl] : 0.0f;
                mk[t] = mk[t] + ((mv[0] + mv[1]) + (mv[2] + mv[3]));
            }
        }
        FQ<kQN> ex;
        if (sd.bits > 1)
            ex = ex_ip_n<kQN>(sd.bits - 1, qs, ix.D_pad, rec + a.soff[2 * s + 1], sd.dim, gl);
        const double cd = 2.0 / (1 << sd.bits);
#pragma unroll
        for (int t = 0; t < kQN; ++t) {
            const float tmp = (st[t].delta * (float)ipi[t]) + (st[t].vlh * (float)ppc);
            const double cst = 0.5 * (double)st[t].sum_q - (double)(0.58f * st[t].q_l2norm);
            const float q1 = (float)((((double)tmp - cst) * (double)sd.rescale_sqrtD) * (double)o);
            float f;
            if (ip) {
                f = (float)((double)q1 * 0.5);
            } else {
                const float y = st[t].q_l2sqr + o2 - q1;
                f = (y < 0.0f) ? 0.0f : y;
            }
            const float ipo1 = tree16_xor(mk[t]);
            const double e = sd.bits > 1 ? (double)ex.v[t] : 0.0;
            const float tt = (float)((double)ipo1 + e * cd + (-1.0 + cd / 2) * (double)st[t].sum_q);
            const float ipoq = exf[2 * s] * tt;
            r.v[t] += v[t];
            r.f[t] += f;
            r.a[t] += ip ? ipoq + 0.0f : o2 + st[t].q_l2sqr - 2 * ipoq;
        }
    }
    return r;
}

// grid ceil(nrec / 16) * ceil(nq / 16), 256 threads: the workgroup copies its 16 records to LDS
// once and takes up to 16 queries over them, 4 at a time (their rotated vectors and planes
// staged together, every record's codes decoded once for the 4); the record tile varies slowest
// so a tile's workgroups run back to back.
__global__ __launch_bounds__(256) void single_est_tile_kernel(SingleEstArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const IndexView &ix = a.ix;
    const int64_t nqb = (a.nq + kSQ - 1) / kSQ;
    const int64_t qb = (int64_t)blockIdx.x % nqb, tile = (int64_t)blockIdx.x / nqb;
    const int tid = threadIdx.x, grp = tid >> 4, gl = tid & 15;
    const int64_t r0 = tile * kSRec;
    const int nr = (int)min<int64_t>(kSRec, a.nrec - r0);
    const int npl = ix.D_pad / 64 * 8;
    uint8_t *lrec = smem;
    float *qr = reinterpret_cast<float *>(smem + (size_t)kSRec * a.stride);
    uint64_t *pl = reinterpret_cast<uint64_t *>(smem + (size_t)kSRec * a.stride + (size_t)kQN * ix.D_pad * 4);
    {  // the tile's records, 8 bytes a lane (stride is a multiple of 8)
        const uint2 *src = reinterpret_cast<const uint2 *>(a.recs + r0 * a.stride);
        uint2 *dst = reinterpret_cast<uint2 *>(lrec);
        const int nw = (int)(nr * a.stride / 8);
        for (int i = tid; i < nw; i += 256) dst[i] = src[i];
    }
    const int64_t qend = min<int64_t>(a.nq, (qb + 1) * kSQ);
    for (int64_t q0 = qb * kSQ; q0 < qend; q0 += kQN) {
        __syncthreads();
        for (int t = 0; t < kQN; ++t) {  // queries past nq repeat the last one (not written)
            const int64_t q = min<int64_t>(q0 + t, a.nq - 1);
            for (int i = tid; i < ix.D_pad; i += 256) qr[t * ix.D_pad + i] = a.qrot[q * ix.D_pad + i];
            for (int i = tid; i < npl; i += 256) pl[t * npl + i] = a.planes[q * npl + i];
        }
        __syncthreads();
        if (grp < nr) {
            const Est3 e = single_pair_n(a, qr, pl, npl, lrec + (size_t)grp * a.stride, q0, gl);
            if (gl == 0)
#pragma unroll
                for (int t = 0; t < kQN; ++t) {
                    if (q0 + t >= qend) break;
                    const int64_t i = (q0 + t) * a.nrec + r0 + grp;
                    if (a.out_vars) a.out_vars[i] = e.v[t];
                    if (a.out_fast) a.out_fast[i] = e.f[t];
                    a.out_acc[i] = e.a[t];
                }
        }
    }
}

hipError_t launch_single_estimate(const IndexView &ix, const float *qrot, const float *bound, int64_t nq,
                                  SingleQ *qs, uint64_t *planes, const uint8_t *recs, int64_t stride, int64_t nrec,
                                  const int32_t *soff, int32_t exf_off, int32_t dist_type, float *out_vars,
                                  float *out_fast, float *out_acc, hipStream_t st) {
    if (nq == 0) return hipSuccess;
    SinglePrepArgs p{ix, qrot, nq, qs, planes};
    hipLaunchKernelGGL(single_prep_kernel, dim3((unsigned)nq, (unsigned)ix.nseg), dim3(64), 0, st, p);
    if (nrec == 0) return hipGetLastError();
    SingleEstArgs a{ix, qrot, bound, qs, planes, recs, stride, nrec, nq, soff, exf_off, dist_type,
                    out_vars, out_fast, out_acc};
    const size_t qbytes = (size_t)ix.D_pad * 4 + (size_t)ix.D_pad / 64 * 8 * 8;
    const int64_t ntiles = (nrec + kSRec - 1) / kSRec;
    if ((size_t)kSRec * stride + kQN * qbytes <= kSingleStageBytes) {
        hipLaunchKernelGGL(single_est_tile_kernel, dim3((unsigned)(ntiles * ((nq + kSQ - 1) / kSQ))), dim3(256),
                           (size_t)kSRec * stride + kQN * qbytes, st, a);
    } else {
        hipLaunchKernelGGL(single_est_kernel, dim3((unsigned)(nq * ntiles)), dim3(256), qbytes, st, a);
    }
    return hipGetLastError();
}

size_t single_query_state_bytes() { return sizeof(SingleQ); }

}  // namespace saqgpu
