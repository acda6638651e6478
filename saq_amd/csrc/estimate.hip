// estimate.hip -- IVF::estimate on gfx950 (saqlib/index/ivf.hpp:313-362).
//
// For one query and its probed lists: the accurate, fast and variance estimates of every
// vector, unpruned, in probe order (the reference's test_relative_error path).  One workgroup
// per 8-block piece (256 vectors, one thread per vector) of a probed list, so that a query's
// long lists spread over the chip: prepare_list builds the list's residual, per-segment scalars
// and u16 LUTs (CaqCluEstimator::prepare), then the piece's vectors get:
//   fast  = sum over segments, in order from 0, of the segment's fast estimate (its variance
//           estimate for 0-bit segments)          SaqCluEstimator::compFastDist (:148-158)
//   vars  = sum of the variance estimates         SaqCluEstimator::varsEstDist  (:127-137)
//   acc   = sum of the accurate estimates         SaqCluEstimator::compAccurateDist (:169-176)
// with the extra-bit inner products from 16-lane groups in the reference's lane order.
#include "common.hpp"
#include "estimators.hpp"

#pragma clang fp contract(off)

namespace saqgpu {

struct EstArgs {
    IndexView ix;
    const float *qrot;      // D_pad
    const float *bound;     // nseg
    const uint32_t *probes; // nprobe
    const int64_t *offs;    // output start of each probed list
    const int2 *pieces;     // per workgroup: (probe index, first block of its 8)
    uint32_t *out_ids;
    float *out_acc, *out_fast, *out_vars;
};

struct EstSmem {
    SegDesc *segs;
    int8_t *gseg, *bseq;
    float *r, *qr, *crbuf, *sc, *cbmin, *cbmax;
    uint16_t *lut;
    float *xs, *ipq, *exv;  // [S][256], [S][256], [256][S]
};

__device__ __host__ inline size_t est_layout(unsigned char *base, const IndexView &ix, EstSmem *mp) {
    const int S = ix.nseg;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        off = (off + 15) & ~size_t(15);
        unsigned char *p = base ? base + off : nullptr;
        off += bytes;
        return p;
    };
    EstSmem m{};
    m.segs = reinterpret_cast<SegDesc *>(take((size_t)S * sizeof(SegDesc)));
    m.gseg = reinterpret_cast<int8_t *>(take(ix.ngroups));
    m.bseq = reinterpret_cast<int8_t *>(take(S));
    m.r = reinterpret_cast<float *>(take(ix.D_pad * 4));
    m.qr = reinterpret_cast<float *>(take(ix.D_pad * 4));
    m.crbuf = reinterpret_cast<float *>(take(ix.D_pad * 4));
    m.sc = reinterpret_cast<float *>(take((size_t)S * SC_N * 4));
    m.cbmin = reinterpret_cast<float *>(take((size_t)ix.ngroups * 16 * 4));
    m.cbmax = reinterpret_cast<float *>(take((size_t)ix.ngroups * 16 * 4));
    m.lut = reinterpret_cast<uint16_t *>(take((size_t)ix.ngroups * 256 * 2));
    m.xs = reinterpret_cast<float *>(take((size_t)S * 256 * 4));
    m.ipq = reinterpret_cast<float *>(take((size_t)S * 256 * 4));
    m.exv = reinterpret_cast<float *>(take((size_t)S * 256 * 4));
    if (mp) *mp = m;
    return off + 16;
}

template <int DT>
__global__ __launch_bounds__(256) void estimate_kernel(EstArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const IndexView &ix = a.ix;
    const int tid = threadIdx.x;
    const int S = ix.nseg, NG = ix.ngroups;
    EstSmem m;
    est_layout(smem, ix, &m);
    const int2 piece = a.pieces[blockIdx.x];
    const uint32_t c = a.probes[piece.x];
    const int n = ix.clu_n[c];
    if (n == 0) return;  // uniform per workgroup
    for (int i = tid; i < S; i += 256) m.segs[i] = ix.segs[i];
    unsigned long long endmask = 0;
    int nbq = 0;
    for (int s = 0; s < S; ++s) {
        const SegDesc sd = ix.segs[s];
        if (sd.bits == 0) continue;
        endmask |= 1ull << (sd.group_off + sd.ngroups - 1);
        for (int g = tid; g < sd.ngroups; g += 256) m.gseg[sd.group_off + g] = (int8_t)s;
        if (tid == 0) m.bseq[nbq] = (int8_t)s;
        ++nbq;
    }
    for (int i = tid; i < ix.D_pad; i += 256) {
        m.qr[i] = a.qrot[i];
        m.crbuf[i] = ix.crot[(int64_t)c * ix.D_pad + i];
    }
    __syncthreads();
    prepare_list<DT>(m.segs, S, m.gseg, NG, m.qr, m.crbuf, a.bound, m.r, m.sc, m.cbmin, m.cbmax, m.lut, ix.D_pad);

    const int nb = (n + kBlock - 1) / kBlock;
    const int64_t vec0 = ix.clu_vec[c];
    const int v = tid, bl = v >> 5, ln = v & 31;
    {
        const int st = piece.y;
        const bool active = st + bl < nb;
        const int idx = (st + bl) * kBlock + ln;
        if (active) {
            const uint8_t *rec = ix.blocks + (ix.clu_blk[c] + st + bl) * ix.block_stride;
            const float *fac = reinterpret_cast<const float *>(rec);
            for (int s = 0; s < S; ++s) m.xs[s * 256 + v] = fac[s * kBlock + ln];
            const uint8_t *codes = rec + ix.fac_floats * 4 + ln * 8;
            // exact integer LUT sums (accumulate_hacc, fastscan_highacc.hpp:66-139)
            uint32_t acc = 0;
            int bk = 0;
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
                if ((endmask >> g) & 1ull) {
                    const int s = m.bseq[bk++];
                    const float *sc = m.sc + s * SC_N;
                    float t2 = (float)(int32_t)acc * sc[SC_DELTA];  // lut.hpp:108-109
                    t2 = t2 + sc[SC_SUMVL];
                    m.ipq[s * 256 + v] = t2;
                    acc = 0;
                }
            }
        }
        __syncthreads();
        // extra-bit inner products, one 16-lane group per (vector, segment with b >= 2)
        {
            const int grp = tid >> 4, gl = tid & 15;
            const int nvec = min(256, n - st * kBlock);
            for (int item = grp; item < nvec * S; item += 16) {
                const int vv = item / S, s = item - vv * S;
                const SegDesc &sd = m.segs[s];
                if (sd.bits < 2) continue;
                const uint8_t *y = ix.long_codes + (vec0 + st * kBlock + vv) * ix.long_stride + sd.long_off;
                const float ex = ex_ip_group(sd.bits - 1, m.r + sd.doff, y, sd.dim, gl);
                if (gl == 0) m.exv[vv * S + s] = ex;
            }
        }
        __syncthreads();
        if (active && idx < n) {
            float fast = 0.0f, vars = 0.0f, accd = 0.0f;
            const int64_t gv = vec0 + idx;
            for (int s = 0; s < S; ++s) {
                const SegDesc &sd = m.segs[s];
                const float *sc = m.sc + s * SC_N;
                const float x = m.xs[s * 256 + v];
                const float vd = DT == 1 ? sc[SC_VFAC] : max0(x * x + sc[SC_VFAC]);  // caq_estimator.hpp:116-136
                vars = vars + vd;
                float fd, ad;
                if (sd.bits == 0) {
                    fd = vd;
                    ad = DT == 1 ? sc[SC_IPQC] : x * x + sc[SC_QL2];
                } else {
                    const float ipq = m.ipq[s * 256 + v];
                    float f = (ipq - sc[SC_SUMQCB]) * sd.rescale_sqrtD;  // caq_estimator.hpp:148-179
                    f = f * x;
                    fd = DT == 1 ? f * 0.5f + sc[SC_IPQC] : max0(x * x + (sc[SC_QL2] - f));
                    // caq_estimator.hpp:190-217
                    const double ex = sd.bits >= 2 ? (double)m.exv[v * S + s] : 0.0;
                    const float t = (float)(((double)ipq + ex * (double)sd.sqd) +
                                            (-1.0 + (double)(sd.sqd / 2)) * (double)sc[SC_SUMQ]);
                    const float ipoq = ix.exfac[(gv * S + s) * 2] * t;
                    ad = DT == 1 ? ipoq + sc[SC_IPQC] : x * x + sc[SC_QL2] - 2 * ipoq;
                }
                fast = fast + fd;
                accd = accd + ad;
            }
            const int64_t o = a.offs[piece.x] + idx;
            a.out_ids[o] = ix.ids[gv];
            a.out_acc[o] = accd;
            if (a.out_fast) a.out_fast[o] = fast;
            if (a.out_vars) a.out_vars[o] = vars;
        }
        __syncthreads();
    }
}

size_t estimate_smem_bytes(const IndexView &ix) { return est_layout(nullptr, ix, nullptr); }

hipError_t launch_estimate(const IndexView &ix, const float *qrot, const float *bound, const uint32_t *probes,
                           const int64_t *offs, const int2 *pieces, int npieces, int dist_type, uint32_t *out_ids,
                           float *out_acc, float *out_fast, float *out_vars, hipStream_t st) {
    if (npieces < 1) return hipSuccess;
    EstArgs a{ix, qrot, bound, probes, offs, pieces, out_ids, out_acc, out_fast, out_vars};
    const size_t smem = estimate_smem_bytes(ix);
    const void *fn = dist_type == 1 ? (const void *)estimate_kernel<1> : (const void *)estimate_kernel<0>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    if (dist_type == 1)
        hipLaunchKernelGGL(estimate_kernel<1>, dim3(npieces), dim3(256), smem, st, a);
    else
        hipLaunchKernelGGL(estimate_kernel<0>, dim3(npieces), dim3(256), smem, st, a);
    return hipGetLastError();
}

}  // namespace saqgpu
