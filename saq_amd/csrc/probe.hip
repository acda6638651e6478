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

hipError_t launch_probe(const float *centroids, int K, int D, const float *queries, int64_t nq, int nprobe,
                        int dist_type, float *dist_ws, uint32_t *probes, hipStream_t st) {
    if (nprobe > kSelMax) return hipErrorInvalidValue;
    dim3 g1((unsigned)((K + 255) / 256), (unsigned)nq);
    const size_t sm = (size_t)D * 4;
    if (dist_type == 1) {
        hipLaunchKernelGGL(probe_dist_kernel<1>, g1, dim3(256), sm, st, centroids, K, D, queries, dist_ws);
        hipLaunchKernelGGL(probe_select_kernel<1>, dim3((unsigned)nq), dim3(256), 0, st, dist_ws, K, nprobe, probes);
    } else {
        hipLaunchKernelGGL(probe_dist_kernel<0>, g1, dim3(256), sm, st, centroids, K, D, queries, dist_ws);
        hipLaunchKernelGGL(probe_select_kernel<0>, dim3((unsigned)nq), dim3(256), 0, st, dist_ws, K, nprobe, probes);
    }
    return hipGetLastError();
}

}  // namespace saqgpu
