// estimators.hpp -- the reference's per-list estimator state and extra-bit inner products,
// shared by the plain scan (slice_scan.hip) and IVF::estimate (estimate.hip).
#pragma once
#include <cfloat>

#include "common.hpp"

namespace saqgpu {

constexpr int kPos[16] = {3, 3, 2, 3, 1, 3, 2, 3, 0, 3, 2, 3, 1, 3, 2, 3};  // fastscan.hpp:22-39

// Per-segment scalars kept in LDS for the current (query, cluster).
enum {
    SC_DELTA = 0, SC_SUMVL, SC_SUMQCB, SC_QL2, SC_SUMQ, SC_VFAC, SC_IPQC, SC_VL, SC_INV, SC_N = 12
};

// Code of dimension 16 i + gl in the reference's compacted layout K (1..8)
// (code_helper.hpp:95-591); each layout keeps the dims = gl (mod 16) of lane gl together.
template <int K>
__device__ __forceinline__ int lane_code(const uint8_t *y, int i, int gl) {
    if (K == 1) return (y[2 * i + (gl >> 3)] >> (gl & 7)) & 1;
    if (K == 2) return (y[16 * (i >> 2) + gl] >> (2 * (i & 3))) & 3;
    if (K == 3) {
        const int c = i >> 2, p = i & 3;
        return ((y[24 * c + gl] >> (2 * p)) & 3) | (((y[24 * c + 16 + (gl & 7)] >> (2 * p + (gl >> 3))) & 1) << 2);
    }
    if (K == 4) {
        const int b = y[16 * (i >> 1) + gl];
        return (i & 1) ? (b >> 4) : (b & 15);
    }
    if (K == 6 || K == 7) {
        const int c = i >> 2, p = i & 3;
        const int base = (K == 6 ? 48 : 56) * c;
        int code;
        if (p == 0) code = y[base + gl] & 0x3F;
        else if (p == 1) code = y[base + 16 + gl] & 0x3F;
        else if (p == 2) code = ((y[base + gl] >> 2) & 0x30) | (y[base + 32 + gl] & 0x0F);
        else code = ((y[base + 16 + gl] >> 2) & 0x30) | ((y[base + 32 + gl] >> 4) & 0x0F);
        if (K == 7) code |= ((y[base + 48 + (gl & 7)] >> (2 * p + (gl >> 3))) & 1) << 6;
        return code;
    }
    return y[16 * i + gl];  // K == 8
}

// Lane accumulator of CodeHelper<K>::compute_ip (code_helper.hpp:105-591) for lane gl of a
// 16-lane group: dims gl, gl+16, gl+32, ... in increasing order.  Codes and query values are
// read eight dims ahead so the LDS reads overlap; the FMA chain keeps the reference's order
// (a masked add for K = 1).
template <int K>
__device__ __forceinline__ float lane_acc_k(const float *r, const uint8_t *y, int D, int gl) {
    float acc = 0.0f;
    const int n = D / 16;  // a multiple of 4
    int i0 = 0;
    for (; i0 + 8 <= n; i0 += 8) {
        int c[8];
        float x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            c[u] = lane_code<K>(y, i0 + u, gl);
            x[u] = r[16 * (i0 + u) + gl];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = K == 1 ? acc + (c[u] ? x[u] : 0.0f) : __fmaf_rn(x[u], (float)c[u], acc);
    }
    for (; i0 < n; i0 += 4) {
        int c[4];
        float x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            c[u] = lane_code<K>(y, i0 + u, gl);
            x[u] = r[16 * (i0 + u) + gl];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = K == 1 ? acc + (c[u] ? x[u] : 0.0f) : __fmaf_rn(x[u], (float)c[u], acc);
    }
    return acc;
}

__device__ __forceinline__ float lane_acc(int k, const float *r, const uint8_t *y, int D, int gl) {
    switch (k) {
    case 1: return lane_acc_k<1>(r, y, D, gl);
    case 2: return lane_acc_k<2>(r, y, D, gl);
    case 3: return lane_acc_k<3>(r, y, D, gl);
    case 4: return lane_acc_k<4>(r, y, D, gl);
    case 6: return lane_acc_k<6>(r, y, D, gl);
    case 7: return lane_acc_k<7>(r, y, D, gl);
    case 8: return lane_acc_k<8>(r, y, D, gl);
    default: return 0.0f;
    }
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

constexpr int kPrepThreads = 256;  // the workgroup size prepare_list is written for

// CaqCluEstimator::prepare of every segment for one (query, list) (caq_estimator.hpp:94-106,
// lut.hpp:60-87): the residual r = q_rot - c_rot (q_rot itself for IP), each segment's scalars
// in sc[s * SC_N ..] (sum_q, |r|^2, the fast-estimate offset, the variance estimate, the u16
// LUT's delta, vl and reciprocal) and the quantised u16 LUT of every code group.  segs, gseg
// (group -> segment), qr, crbuf and the outputs live in LDS; bound is the query's per-segment
// variance bound.  Every thread of a 256-thread workgroup calls it; it ends with a barrier.
// One segment's scalars of CaqCluEstimator::prepare by one wavefront (all 64 lanes call it):
// sum_q, |r|^2, the fast-estimate offset, <r, c> (IP) and the variance estimate's constant.
// r and crbuf point at the whole padded vectors.
template <int DT>
__device__ __forceinline__ void segment_scalars(const SegDesc &sd, int s, const float *r, const float *crbuf,
                                                const float *bound, float *sc) {
    const int lane = threadIdx.x & 63;
    const float *rs = r + sd.doff;
    const float *cs = crbuf + sd.doff;
    const float sum_q = eigen_redux_wave(sd.dim, [rs](int i) { return rs[i]; });
    const float ql2 = eigen_redux_wave(sd.dim, [rs](int i) { return rs[i] * rs[i]; });
    float ipqc = 0.0f;
    if (DT == 1) ipqc = eigen_redux_wave(sd.dim, [rs, cs](int i) { return rs[i] * cs[i]; });
    if (lane == 0) {
        float *ss = sc + s * SC_N;
        const float ql2n = sqrtf(ql2);
        ss[SC_SUMQ] = sum_q;
        ss[SC_QL2] = ql2;
        ss[SC_SUMQCB] = (float)(0.5 * (double)sum_q - (double)(0.58f * ql2n));
        ss[SC_IPQC] = ipqc;
        ss[SC_VFAC] = DT == 1 ? ipqc - bound[s] : ql2 - 2 * bound[s];
    }
}

// The first half of prepare_list: the residual and every segment's scalars except the LUT's.
template <int DT>
__device__ void prepare_scalars(const SegDesc *segs, int S, const float *qr, const float *crbuf, const float *bound,
                                float *r, float *sc, int D_pad) {
    const int tid = threadIdx.x, wave = tid >> 6;
    // the residual; a thread's loads (up to 8 each of q and c) are all in flight at once
    float qv[8], cv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int d = tid + u * kPrepThreads;
        qv[u] = d < D_pad ? qr[d] : 0.0f;
        cv[u] = (DT != 1 && d < D_pad) ? crbuf[d] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int d = tid + u * kPrepThreads;
        if (d < D_pad) r[d] = DT == 1 ? qv[u] : qv[u] - cv[u];
    }
    for (int d = tid + 8 * kPrepThreads; d < D_pad; d += kPrepThreads) r[d] = DT == 1 ? qr[d] : qr[d] - crbuf[d];
    __syncthreads();
    for (int s = wave; s < S; s += 4) segment_scalars<DT>(segs[s], s, r, crbuf, bound, sc);
}

template <int DT>
__device__ void prepare_list(const SegDesc *segs, int S, const int8_t *gseg, int NG, const float *qr,
                             const float *crbuf, const float *bound, float *r, float *sc, float *cbmin,
                             float *cbmax, uint16_t *lut, int D_pad) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ncb = NG * 16;
    prepare_scalars<DT>(segs, S, qr, crbuf, bound, r, sc, D_pad);
    // LUT pass 1: per-codebook min / max of the fp32 table (fastscan.hpp:236-248)
    for (int c = tid; c < ncb; c += kPrepThreads) {
        const int g = c >> 4;
        const SegDesc &sd = segs[gseg[g]];
        const float *rr = r + sd.doff + (g - sd.group_off) * 64 + (c & 15) * 4;
        float l[16];
        l[0] = 0;
        float mn = 0.0f, mx = 0.0f;
#pragma unroll
        for (int j = 1; j < 16; ++j) {
            l[j] = l[j - (j & -j)] + rr[kPos[j]];
            mn = fminf(mn, l[j]);
            mx = fmaxf(mx, l[j]);
        }
        cbmin[c] = mn;
        cbmax[c] = mx;
    }
    __syncthreads();
    for (int s = wave; s < S; s += 4) {
        const SegDesc sd = segs[s];
        if (sd.bits == 0) continue;
        float mn = FLT_MAX, mx = -FLT_MAX;
        for (int c = sd.group_off * 16 + lane; c < (sd.group_off + sd.ngroups) * 16; c += 64) {
            mn = fminf(mn, cbmin[c]);
            mx = fmaxf(mx, cbmax[c]);
        }
        for (int o = 32; o >= 1; o >>= 1) {
            mn = fminf(mn, __shfl_xor(mn, o));
            mx = fmaxf(mx, __shfl_xor(mx, o));
        }
        if (lane == 0) {
            float *ss = sc + s * SC_N;
            const float delta = (float)((double)(mx - mn) / ((1 << 16) - 0.01));
            ss[SC_DELTA] = delta;
            ss[SC_INV] = 1.0f / delta;  // the release build's hoisted reciprocal (lut.hpp:74-77)
            ss[SC_VL] = mn;
            ss[SC_SUMVL] = (float)(((double)mn + 0.5 * (double)delta) * (double)(float)(sd.dim / 4));
        }
    }
    __syncthreads();
    // LUT pass 2: quantise to u16 (lut.hpp:74-77)
    for (int c = tid; c < ncb; c += kPrepThreads) {
        const int g = c >> 4;
        const int s = gseg[g];
        const SegDesc &sd = segs[s];
        const float *rr = r + sd.doff + (g - sd.group_off) * 64 + (c & 15) * 4;
        const float vl = sc[s * SC_N + SC_VL], inv = sc[s * SC_N + SC_INV];
        float l[16];
        l[0] = 0;
#pragma unroll
        for (int j = 1; j < 16; ++j) l[j] = l[j - (j & -j)] + rr[kPos[j]];
        uint32_t *dst = reinterpret_cast<uint32_t *>(lut + c * 16);
#pragma unroll
        for (int j = 0; j < 16; j += 2)
            dst[j / 2] = cvt_u16((l[j] - vl) * inv) | (cvt_u16((l[j + 1] - vl) * inv) << 16);
    }
    __syncthreads();
}

}  // namespace saqgpu
