// probe.hip -- the coarse quantiser on gfx950: FlatInitializer::centroids_distances
// (saqlib/index/initializer.hpp:52-76) for a batch of queries.
//
//   probe_dist_kernel    distance of every (query, centroid) pair in the order Eigen reduces
//                        (centroids_.row(i) - query).squaredNorm() / query.dot(row): two
//                        16-lane packet accumulators, trailing packet, predux tree, scalar tail.
//   probe_select_kernel  one workgroup per query selects the nprobe best centroids by an exact
//                        64-bit radix select on (distance key, centroid id), then sorts them
//                        ascending (descending similarity for IP).  Equal distances are ordered
//                        by id; std::partial_sort leaves the order of exact ties unspecified.
#include <cfloat>

#include "common.hpp"

#pragma clang fp contract(off)

namespace saqgpu {

// Eigen redux over n elements (any n) evaluated by one thread.
template <class F>
__device__ __forceinline__ float eigen_redux_thread(int n, F elem) {
    if (n < 16) {
        float r = elem(0);
        for (int i = 1; i < n; ++i) r = r + elem(i);
        return r;
    }
    const int aligned = n / 16 * 16, aligned2 = n / 32 * 32;
    float p0[16], p1[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) p0[l] = elem(l);
    if (aligned > 16) {
#pragma unroll
        for (int l = 0; l < 16; ++l) p1[l] = elem(16 + l);
        for (int idx = 32; idx < aligned2; idx += 32) {
#pragma unroll
            for (int l = 0; l < 16; ++l) {
                p0[l] = p0[l] + elem(idx + l);
                p1[l] = p1[l] + elem(idx + 16 + l);
            }
        }
#pragma unroll
        for (int l = 0; l < 16; ++l) p0[l] = p0[l] + p1[l];
        if (aligned > aligned2) {
#pragma unroll
            for (int l = 0; l < 16; ++l) p0[l] = p0[l] + elem(aligned2 + l);
        }
    }
    float b[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) b[i] = p0[i] + p0[i + 8];
    float c[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) c[i] = b[i] + b[i + 4];
    float r = (c[0] + c[2]) + (c[1] + c[3]);
    for (int i = aligned; i < n; ++i) r = r + elem(i);
    return r;
}

// grid (ceil(K/256), nq)
template <int DT>
__global__ __launch_bounds__(256) void probe_dist_kernel(const float *__restrict__ cent, int K, int D,
                                                         const float *__restrict__ queries, float *__restrict__ out) {
    extern __shared__ float qsh[];
    const int64_t q = blockIdx.y;
    for (int i = threadIdx.x; i < D; i += 256) qsh[i] = queries[q * D + i];
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= K) return;
    const float *row = cent + (int64_t)c * D;
    float d;
    if (DT == 1)
        d = eigen_redux_thread(D, [&](int i) { return qsh[i] * row[i]; });
    else
        d = eigen_redux_thread(D, [&](int i) {
            const float t = row[i] - qsh[i];
            return t * t;
        });
    out[q * K + c] = d;
}

// Register-tiled form of probe_dist_kernel for D >= 32.  Eigen's redux keeps 32 running sums
// (two 16-lane packets) over the first floor(D/32)*32 elements, sum j taking the elements
// i = j (mod 32) in increasing order; here lane j of a half-wave holds exactly that sum, for
// 8 queries x 16 centroids at once, with the rows staged through LDS in chunks of 128 dims.
// The combine (p0 + p1), the trailing 16-element packet, the halving tree and the scalar tail
// then follow Eigen's order across lanes.  A workgroup covers 32 queries x 32 centroids.
constexpr int kPT = 32;     // queries (and centroids) per workgroup tile
constexpr int kPCh = 128;   // dims per LDS chunk

// One step of the 8 x 16 register tile: element i of 8 query rows and 16 centroid rows.
template <int DT, bool kInit>
__device__ __forceinline__ void probe_step(float (&acc)[8][16], const float (*qs)[kPCh], const float (*cs)[kPCh],
                                           int qa, int cb, int i) {
    float qv[8], cv[16];
#pragma unroll
    for (int a = 0; a < 8; ++a) qv[a] = qs[qa + a][i];
#pragma unroll
    for (int b = 0; b < 16; ++b) cv[b] = cs[cb + b][i];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            float e;
            if (DT == 1) {
                e = qv[a] * cv[b];
            } else {
                const float d = cv[b] - qv[a];
                e = d * d;
            }
            acc[a][b] = kInit ? e : acc[a][b] + e;
        }
}

template <int DT, bool kVec>
__global__ __launch_bounds__(256) void probe_tile_kernel(const float *__restrict__ cent, int K, int D,
                                                         const float *__restrict__ queries, int64_t nq,
                                                         float *__restrict__ out) {
    __shared__ float qs[kPT][kPCh];
    __shared__ float cs[kPT][kPCh];
    const int tid = threadIdx.x, hw = tid >> 5, j = tid & 31;
    const int64_t q0 = (int64_t)blockIdx.y * kPT;
    const int c0 = blockIdx.x * kPT;
    const int qa = (hw >> 1) * 8;   // this half-wave's 8 queries within the tile
    const int cb = (hw & 1) * 16;   // and 16 centroids
    const int aligned = D / 16 * 16, aligned2 = D / 32 * 32;
    float acc[8][16];
    // chunk staging: thread tid moves float4 number tid + 256u (u < 4) of each 32 x 128 tile;
    // rows past nq or K repeat the last valid row (their results are not stored)
    float4 pq[4], pc[4];
    auto fetch = [&](int base) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e4 = tid + 256 * u, r = e4 >> 5, i = base + 4 * (e4 & 31);
            const int64_t q = min<int64_t>(q0 + r, nq - 1);
            const int c = min(c0 + r, K - 1);
            if (kVec) {
                pq[u] = i < aligned2 ? *reinterpret_cast<const float4 *>(queries + q * D + i) : float4{};
                pc[u] = i < aligned2 ? *reinterpret_cast<const float4 *>(cent + (int64_t)c * D + i) : float4{};
            } else {
                const float *qr = queries + q * D, *cr = cent + (int64_t)c * D;
                pq[u] = i < aligned2 ? float4{qr[i], qr[i + 1], qr[i + 2], qr[i + 3]} : float4{};
                pc[u] = i < aligned2 ? float4{cr[i], cr[i + 1], cr[i + 2], cr[i + 3]} : float4{};
            }
        }
    };
    fetch(0);
    for (int base = 0; base < aligned2; base += kPCh) {
        const int len = min(kPCh, aligned2 - base);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e4 = tid + 256 * u, r = e4 >> 5, c4 = e4 & 31;
            *reinterpret_cast<float4 *>(&qs[r][4 * c4]) = pq[u];
            *reinterpret_cast<float4 *>(&cs[r][4 * c4]) = pc[u];
        }
        __syncthreads();
        if (base + kPCh < aligned2) fetch(base + kPCh);  // in flight while this chunk is reduced
        int t = 0;
        if (base == 0) {
            probe_step<DT, true>(acc, qs, cs, qa, cb, j);
            t = 1;
        }
        for (; t < len / 32; ++t) probe_step<DT, false>(acc, qs, cs, qa, cb, 32 * t + j);
    }
    // per pair: combine, trailing packet, halving tree, tail (lane 0 ends with the result)
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            const int64_t q = q0 + qa + a;
            const int c = c0 + cb + b;
            const bool ok = q < nq && c < K;
            const float *qr = queries + (ok ? q : 0) * D;
            const float *cr = cent + (int64_t)(ok ? c : 0) * D;
            auto elem = [&](int i) {
                if (DT == 1) return qr[i] * cr[i];
                const float d = cr[i] - qr[i];
                return d * d;
            };
            float p0;
            if (aligned > 16) {
                const float hi = __shfl_down(acc[a][b], 16, 32);
                p0 = acc[a][b] + hi;
                if (aligned > aligned2 && j < 16) p0 = p0 + elem(aligned2 + j);
            } else {
                p0 = j < 16 ? elem(j) : 0.0f;
            }
            const float b8 = p0 + __shfl_down(p0, 8, 32);
            const float c4 = b8 + __shfl_down(b8, 4, 32);
            const float c2 = __shfl_down(c4, 2, 32), c1 = __shfl_down(c4, 1, 32), c3 = __shfl_down(c4, 3, 32);
            float r = (c4 + c2) + (c1 + c3);
            if (j == 0 && ok) {
                for (int i = aligned; i < D; ++i) r = r + elem(i);
                out[q * K + c] = r;
            }
        }
}

// ---- fp32-MFMA coarse quantiser (batched mode)
// dist[q][c] = |q|^2 + |c|^2 - 2 q.c (or q.c for IP) with q.c on v_mfma_f32_16x16x4_f32.  This is
// not FlatInitializer's reduction order, so a probe list can differ from the reference's where
// two centroid distances lie within fp32 rounding of each other; the batched scan, which is
// approximate by design, may use it.  A workgroup covers 64 queries x 64 centroids, each wave
// 32 x 32, with 32-dim chunks of both staged through LDS.
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kGT = 64, kGK = 32;

__global__ __launch_bounds__(256) void sqnorm_rows_kernel(const float *__restrict__ x, int64_t rows, int D,
                                                          float *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows) return;
    float s = 0.0f;
    for (int i = lane; i < D; i += 64) s += x[r * D + i] * x[r * D + i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) out[r] = s;
}

template <int DT>
__global__ __launch_bounds__(256) void probe_gemm_kernel(const float *__restrict__ cent, const float *__restrict__ cnorm,
                                                         int K, int D, const float *__restrict__ queries,
                                                         const float *__restrict__ qnorm, int64_t nq,
                                                         float *__restrict__ out) {
    __shared__ float As[kGT][kGK + 1];
    __shared__ float Bs[kGT][kGK + 1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t q0 = (int64_t)blockIdx.y * kGT;
    const int c0 = blockIdx.x * kGT;
    const int qb = (w >> 1) * 32, cbb = (w & 1) * 32;
    v4f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < D; k0 += kGK) {
        __syncthreads();
        for (int e = tid; e < kGT * kGK; e += 256) {
            const int r = e / kGK, k = e % kGK;
            const int64_t q = q0 + r;
            const int c = c0 + r;
            As[r][k] = (q < nq && k0 + k < D) ? queries[q * D + k0 + k] : 0.0f;
            Bs[r][k] = (c < K && k0 + k < D) ? cent[(int64_t)c * D + k0 + k] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kGK; kk += 4) {
            float av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = As[qb + a * 16 + lane % 16][kk + lane / 16];
#pragma unroll
            for (int b = 0; b < 2; ++b) bv[b] = Bs[cbb + b * 16 + lane % 16][kk + lane / 16];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t q = q0 + qb + a * 16 + 4 * (lane / 16) + r;
                const int c = c0 + cbb + b * 16 + lane % 16;
                if (q < nq && c < K) {
                    const float ip = acc[a][b][r];
                    out[q * K + c] = DT == 1 ? ip : (qnorm[q] + cnorm[c]) - 2.0f * ip;
                }
            }
}

constexpr int kSelMax = 4096;

// 64-bit key: ascending order = probe order.
template <int DT>
__device__ __forceinline__ unsigned long long sel_key(float d, int id) {
    uint32_t k = float_key(d);
    if (DT == 1) k = ~k;  // largest inner product first
    return ((unsigned long long)k << 32) | (uint32_t)id;
}

template <int DT>
__global__ __launch_bounds__(256) void probe_select_kernel(const float *__restrict__ dist, int K, int nprobe,
                                                           uint32_t *__restrict__ probes) {
    __shared__ unsigned int hist[256];
    __shared__ unsigned long long prefix_s;
    __shared__ int rank_s;
    __shared__ int cnt_s;
    __shared__ unsigned long long sel[kSelMax];
    const int64_t q = blockIdx.x;
    const float *dq = dist + q * K;
    const int tid = threadIdx.x;
    if (tid == 0) {
        prefix_s = 0;
        rank_s = nprobe;  // find the nprobe-th smallest key (1-based)
        cnt_s = 0;
    }
    // radix select of the nprobe-th smallest 64-bit key, 8 bits per pass from the top
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = 56 - 8 * pass;
        for (int i = tid; i < 256; i += 256) hist[i] = 0;
        __syncthreads();
        const unsigned long long pre = prefix_s;
        const unsigned long long hmask = pass == 0 ? 0ull : (~0ull << (shift + 8));
        for (int i = tid; i < K; i += 256) {
            const unsigned long long k = sel_key<DT>(dq[i], i);
            if ((k & hmask) == pre) atomicAdd(&hist[(k >> shift) & 0xFF], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            int r = rank_s;
            int b = 0;
            for (; b < 256; ++b) {
                if ((int)hist[b] >= r) break;
                r -= hist[b];
            }
            prefix_s = pre | ((unsigned long long)b << shift);
            rank_s = r;
        }
        __syncthreads();
    }
    const unsigned long long thr = prefix_s;  // exactly nprobe keys are <= thr
    for (int i = tid; i < K; i += 256) {
        const unsigned long long k = sel_key<DT>(dq[i], i);
        if (k <= thr) {
            const int p = atomicAdd(&cnt_s, 1);
            if (p < kSelMax) sel[p] = k;
        }
    }
    __syncthreads();
    int n = min(cnt_s, kSelMax);
    int P = 1;
    while (P < n) P <<= 1;
    for (int i = n + tid; i < P; i += 256) sel[i] = ~0ull;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < P; i += 256) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const unsigned long long a = sel[i], b = sel[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        sel[i] = b;
                        sel[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < nprobe && i < n; i += 256) probes[q * nprobe + i] = (uint32_t)(sel[i] & 0xFFFFFFFFu);
}

hipError_t launch_sqnorm_rows(const float *x, int64_t rows, int D, float *out, hipStream_t st) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(sqnorm_rows_kernel, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, st, x, rows, D, out);
    return hipGetLastError();
}

// The batched scan's coarse quantiser: fp32-MFMA distances (cnorm = |c|^2 per centroid, qnorm
// workspace of nq floats), then the same exact top-nprobe selection.
hipError_t launch_probe_fast(const float *centroids, const float *cnorm, int K, int D, const float *queries,
                             int64_t nq, int nprobe, int dist_type, float *qnorm, float *dist_ws, uint32_t *probes,
                             hipStream_t st) {
    if (nprobe > kSelMax || (nq + kGT - 1) / kGT > 65535) return hipErrorInvalidValue;
    hipError_t e = launch_sqnorm_rows(queries, nq, D, qnorm, st);
    if (e != hipSuccess) return e;
    dim3 g((unsigned)((K + kGT - 1) / kGT), (unsigned)((nq + kGT - 1) / kGT));
    if (dist_type == 1) {
        hipLaunchKernelGGL(probe_gemm_kernel<1>, g, dim3(256), 0, st, centroids, cnorm, K, D, queries, qnorm, nq, dist_ws);
        hipLaunchKernelGGL(probe_select_kernel<1>, dim3((unsigned)nq), dim3(256), 0, st, dist_ws, K, nprobe, probes);
    } else {
        hipLaunchKernelGGL(probe_gemm_kernel<0>, g, dim3(256), 0, st, centroids, cnorm, K, D, queries, qnorm, nq, dist_ws);
        hipLaunchKernelGGL(probe_select_kernel<0>, dim3((unsigned)nq), dim3(256), 0, st, dist_ws, K, nprobe, probes);
    }
    return hipGetLastError();
}

template <int DT>
static void probe_tile(bool vec, dim3 g, hipStream_t st, const float *centroids, int K, int D, const float *queries,
                       int64_t nq, float *dist_ws) {
    if (vec)
        hipLaunchKernelGGL((probe_tile_kernel<DT, true>), g, dim3(256), 0, st, centroids, K, D, queries, nq, dist_ws);
    else
        hipLaunchKernelGGL((probe_tile_kernel<DT, false>), g, dim3(256), 0, st, centroids, K, D, queries, nq, dist_ws);
}

hipError_t launch_probe(const float *centroids, int K, int D, const float *queries, int64_t nq, int nprobe,
                        int dist_type, float *dist_ws, uint32_t *probes, hipStream_t st) {
    if (nprobe > kSelMax) return hipErrorInvalidValue;
    const bool tiled = D >= 32 && (nq + kPT - 1) / kPT <= 65535;
    const bool vec = D % 4 == 0 && ((uintptr_t)centroids % 16) == 0 && ((uintptr_t)queries % 16) == 0;
    dim3 g1((unsigned)((K + 255) / 256), (unsigned)nq);
    dim3 gt((unsigned)((K + kPT - 1) / kPT), (unsigned)((nq + kPT - 1) / kPT));
    const size_t sm = (size_t)D * 4;
    if (!tiled && nq > 65535) return hipErrorInvalidValue;
    if (dist_type == 1) {
        if (tiled)
            probe_tile<1>(vec, gt, st, centroids, K, D, queries, nq, dist_ws);
        else
            hipLaunchKernelGGL(probe_dist_kernel<1>, g1, dim3(256), sm, st, centroids, K, D, queries, dist_ws);
        hipLaunchKernelGGL(probe_select_kernel<1>, dim3((unsigned)nq), dim3(256), 0, st, dist_ws, K, nprobe, probes);
    } else {
        if (tiled)
            probe_tile<0>(vec, gt, st, centroids, K, D, queries, nq, dist_ws);
        else
            hipLaunchKernelGGL(probe_dist_kernel<0>, g1, dim3(256), sm, st, centroids, K, D, queries, dist_ws);
        hipLaunchKernelGGL(probe_select_kernel<0>, dim3((unsigned)nq), dim3(256), 0, st, dist_ws, K, nprobe, probes);
    }
    return hipGetLastError();
}

}  // namespace saqgpu
