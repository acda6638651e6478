// query_prep.hip -- the per-query set-up of the scan on gfx950 (SaqEstimatorBase +
// CaqCluEstimator construction, saqlib/quantization/saq_estimator.hpp:28-48,
// caq/caq_estimator.hpp:56-82):
//   query_prep_kernel  per (query, segment): the rotated query q_s * P_s in the order of
//                      Eigen's column-major GEMV
//   query_bound_kernel per (query, segment), one wavefront: the variance bound
//                      m * sqrt(sum var * q^2) (caq_estimator.hpp:80-82)
// The encoder uses the same kernels for the rotated centroids c_s * P_s (quantizer.hpp:46).
#include <cfloat>

#include "common.hpp"
#include "estimators.hpp"

#pragma clang fp contract(off)

namespace saqgpu {

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

}  // namespace saqgpu

// ---------------------------------------------------------------- launcher
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

}  // namespace saqgpu
