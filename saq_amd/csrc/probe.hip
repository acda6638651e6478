// probe.hip -- the coarse quantiser on gfx950: FlatInitializer::centroids_distances
// (saqlib/index/initializer.hpp:52-76) for a batch of queries.
//
//   probe_dist_kernel    distance of every (query, centroid) pair in the order Eigen reduces
//                        (centroids_.row(i) - query).squaredNorm() / query.dot(row): two
//                        16-lane packet accumulators, trailing packet, predux tree, scalar tail.
//   probe_select_kernel  one workgroup per query selects the nprobe best centroids by an exact
//                        64-bit radix select on (distance key, centroid id), then sorts them
//                        ascending (descending similarity for IP).  That is the reference's
//                        order whenever the distances involved are distinct; a query with two
//                        equal distances among its first nprobe + 1 is flagged for
//                        probe_heap_kernel.  Its kIP form reads q.c products of a library GEMM
//                        (the batched scan's approximate coarse quantiser) and ranks by
//                        |c|^2 - 2 q.c.
//   probe_exact_kernel   the exact lists from a library GEMM's q.c: the centroids within a proven
//                        fp32 error bound of the nprobe-th key, recomputed in Eigen's order
//                        (by half-waves), then an exact (distance, id) selection.
//   probe_heap_kernel    the flagged queries only: std::partial_sort's own order for exact ties,
//                        by running the heap selection libstdc++ (GCC 11, the reference's
//                        toolchain) implements -- make_heap over the first nprobe centroids,
//                        a pop_heap for every later one that compares below the top, then
//                        sort_heap -- over every centroid's distance in id order.
//   sqnorm_rows_kernel   |x|^2 per row (the centroid norms of the GEMM-based forms).
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

// eigen_redux_thread's result computed by the 32 lanes of a half-wave (l = lane % 32), all
// lanes returning it: lane l carries Eigen's running sum l (p0[l] for l < 16, p1[l - 16] above)
// over the elements l, l + 32, ... below floor(n/32)*32, so a row is read in coalesced runs; the
// combine p0 + p1, the trailing 16-element packet, the halving tree and the scalar tail follow
// in Eigen's order.
template <class F>
__device__ __forceinline__ float eigen_redux_half(int n, F elem, int l) {
    const int base = threadIdx.x & 32;
    if (n < 16) {
        float r = elem(0);
        for (int i = 1; i < n; ++i) r = r + elem(i);
        return r;
    }
    const int aligned = n / 16 * 16, aligned2 = n / 32 * 32;
    float p0;
    if (aligned > 16) {
        float acc = elem(l);
        for (int idx = 32; idx < aligned2; idx += 32) acc = acc + elem(idx + l);
        const float p1 = __shfl(acc, base + ((l + 16) & 31));
        p0 = acc + p1;
        if (aligned > aligned2 && l < 16) p0 = p0 + elem(aligned2 + l);
    } else {
        p0 = l < 16 ? elem(l) : 0.0f;
    }
    const float b = p0 + __shfl(p0, base + ((l + 8) & 31));   // b[i] = p0[i] + p0[i + 8], i < 8
    const float c = b + __shfl(b, base + ((l + 4) & 31));     // c[i] = b[i] + b[i + 4], i < 4
    const float c0 = __shfl(c, base), c1 = __shfl(c, base + 1), c2 = __shfl(c, base + 2), c3 = __shfl(c, base + 3);
    float r = (c0 + c2) + (c1 + c3);
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

// |x_r|^2 per row, one wave per row (the coarse quantiser's norms).
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

constexpr int kSelMax = 4096;

// 64-bit key: ascending order = probe order.
template <int DT>
__device__ __forceinline__ unsigned long long sel_key(float d, int id) {
    uint32_t k = float_key(d);
    if (DT == 1) k = ~k;  // largest inner product first
    return ((unsigned long long)k << 32) | (uint32_t)id;
}

// Inclusive prefix sum of v over the workgroup's 256 threads (4 waves); wsum: 4 ints of LDS.
__device__ __forceinline__ int block_scan256(int v, int *wsum) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o);
        if (lane >= o) v += u;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    for (int i = 0; i < w; ++i) v += wsum[i];
    __syncthreads();
    return v;
}

// LDS of block_radix_kth.
struct RadixLds {
    unsigned int hist[256];
    int wsum[4];
    unsigned long long prefix;
    int rank, done;
    int last_pass;  // the pass that ended the select: 4 and above only when float keys tie across the r-th
};

// The r-th smallest of the workgroup's distinct 64-bit keys (1 <= r <= their number), as the
// value thr with exactly r keys <= thr: a radix select by 8-bit digits from the top.  A pass
// whose bucket is taken whole ends the search, so distinct float keys take at most 4 passes
// and only exact ties reach the id bytes.  A thread's
// keys: kv[0 .. KPT) (centroids tid + 256 j below K) or, for KPT = 0, key(c) for c = tid,
// tid + 256, ... below K.
// pass0 > 0: every key shares its top 8 pass0 bits with pre0, and the search starts below them.
template <int KPT, class KeyFn>
__device__ __forceinline__ unsigned long long block_radix_kth(const unsigned long long (&kv)[KPT > 0 ? KPT : 1],
                                                              KeyFn key, int K, int r, RadixLds &sh, int pass0 = 0,
                                                              unsigned long long pre0 = 0) {
    const int tid = threadIdx.x;
    unsigned long long pre = pass0 > 0 ? pre0 & (~0ull << (64 - 8 * pass0)) : 0ull;
    for (int pass = pass0; pass < 8; ++pass) {
        const int shift = 56 - 8 * pass;
        sh.hist[tid] = 0;
        __syncthreads();
        const unsigned long long hmask = pass == 0 ? 0ull : (~0ull << (shift + 8));
        if constexpr (KPT > 0) {
#pragma unroll
            for (int j = 0; j < KPT; ++j)
                if (tid + 256 * j < K && (kv[j] & hmask) == pre) atomicAdd(&sh.hist[(kv[j] >> shift) & 0xFF], 1u);
        } else {
            for (int c0 = tid; c0 < K; c0 += 256 * 8) {
                unsigned long long kk[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) kk[u] = key(min(c0 + 256 * u, K - 1));  // loads issued together
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (c0 + 256 * u < K && (kk[u] & hmask) == pre) atomicAdd(&sh.hist[(kk[u] >> shift) & 0xFF], 1u);
            }
        }
        __syncthreads();
        const int h = (int)sh.hist[tid];
        const int incl = block_scan256(h, sh.wsum);
        const int excl = incl - h;
        if (excl < r && r <= incl) {  // exactly one thread: the bucket holding the r-th key
            sh.prefix = pre | ((unsigned long long)tid << shift);
            sh.rank = r - excl;
            sh.done = (r - excl) == h;
            sh.last_pass = pass;
        }
        __syncthreads();
        pre = sh.prefix;
        r = sh.rank;
        const bool done = sh.done;
        __syncthreads();
        if (done || pass == 7) return pass == 7 ? pre : (pre | ((1ull << shift) - 1));
    }
    return pre;  // not reached
}

// One workgroup per query.  Radix select of the nprobe-th smallest 64-bit key (distance key,
// centroid id), 8 bits per pass from the top; a pass whose chosen byte bucket is taken whole
// ends the search (every key below the bucket's upper end is selected), so distinct distances
// need at most 4 passes and the id bytes are only looked at for exact distance ties.  The
// selected keys are then ordered: by rank counting for up to 256 of them, else by a bitonic
// sort in LDS.
//
// kIP: dist holds inner products q.c and the key is formed on the fly: |c|^2 - 2 q.c for L2
// (|q|^2 is the same for every centroid of a query and is left out), q.c itself for IP.
// KPT > 0: K <= 256 KPT, the keys of a thread's centroids (tid + 256 j) stay in registers
// across the passes; KPT = 0 forms them again in every pass.
// ties (exact distances only, may be null): ties[q] = 1 when two of the query's nprobe + 1 smallest
// distances are equal, where std::partial_sort's order is its own (probe_heap_kernel).  A tie across
// the nprobe-th place is one the select resolves in the id bytes (a pass of 4 or more); one within
// the selection shows as equal float keys side by side once they are ordered.
template <int DT, bool kIP = false, int KPT = 0>
__global__ __launch_bounds__(256) void probe_select_kernel(const float *__restrict__ dist, int K, int nprobe,
                                                           uint32_t *__restrict__ probes,
                                                           const float *__restrict__ cnorm = nullptr,
                                                           int32_t *__restrict__ ties = nullptr) {
    __shared__ RadixLds rsh;
    __shared__ int cnt_s;
    __shared__ unsigned long long sel[kSelMax];
    __shared__ unsigned long long small[KPT > 0 ? 1 : 256];
    __shared__ unsigned long long thr_s, kmin_s;
    const int64_t q = blockIdx.x;
    const float *dq = dist + q * K;
    const int tid = threadIdx.x;
    auto key = [=](int i) {
        const float v = dq[i];
        return sel_key<DT>(kIP && DT == 0 ? cnorm[i] - 2.0f * v : v, i);
    };
    unsigned long long kv[KPT > 0 ? KPT : 1];
    kv[0] = 0;
    if constexpr (KPT > 0) {
#pragma unroll
        for (int j = 0; j < KPT; ++j) kv[j] = tid + 256 * j < K ? key(tid + 256 * j) : ~0ull;
    }
    if constexpr (KPT == 0) {
        // Candidates first (nprobe <= 256, no tie report): T = the nprobe-th smallest of the 256
        // threads' minimum keys has at least nprobe keys <= T, so the nprobe smallest keys all are
        // <= T.  Those keys (a few times nprobe for the distance spreads of coarse quantisers) go
        // to LDS in one more read, and the exact radix select runs over them there instead of
        // over all K keys in global memory.  More than kSelMax of them: the full select below.
        if (ties == nullptr && nprobe <= 256 && K >= 256) {
            constexpr int U = 16;  // keys a thread loads together
            unsigned long long m = ~0ull;
            for (int c0 = tid; c0 < K; c0 += 256 * U) {
                unsigned long long kk[U];
#pragma unroll
                for (int u = 0; u < U; ++u) kk[u] = key(min(c0 + 256 * u, K - 1));  // loads issued together
#pragma unroll
                for (int u = 0; u < U; ++u) m = c0 + 256 * u < K ? min(m, kk[u]) : m;
            }
            small[tid] = m;
            if (tid == 0) cnt_s = 0;
            __syncthreads();
            int rank = 0;
            for (int j = 0; j < 256; ++j) rank += small[j] < m;
            if (rank == nprobe - 1) thr_s = m;  // keys are distinct (ids): one thread
            if (rank == 0) kmin_s = m;
            __syncthreads();
            const unsigned long long T = thr_s, kmin = kmin_s;
            const int lane = tid & 63;
            for (int c0 = tid; c0 < K; c0 += 256 * U) {
                unsigned long long kk[U];
                int nt = 0;
#pragma unroll
                for (int u = 0; u < U; ++u) kk[u] = key(min(c0 + 256 * u, K - 1));
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (c0 + 256 * u >= K) kk[u] = ~0ull;  // above T: T is a key below K
                    nt += kk[u] <= T;
                }
                int incl = nt;  // one LDS atomic per wave and chunk: the wave's prefix sum of takes
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int v = __shfl_up(incl, o);
                    if (lane >= o) incl += v;
                }
                int b = 0;
                if (lane == 63 && incl) b = atomicAdd(&cnt_s, incl);
                int p = __shfl(b, 63) + incl - nt;
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (kk[u] <= T) {
                        if (p < kSelMax) sel[p] = kk[u];
                        ++p;
                    }
            }
            __syncthreads();
            const int n = cnt_s;
            if (n <= kSelMax) {
                // the candidates lie in [kmin, T]: their common top bytes need no pass
                const int pass0 = kmin == T ? 7 : __clzll(kmin ^ T) / 8;
                const unsigned long long thr =
                    block_radix_kth<0>(kv, [&](int i) { return sel[i]; }, n, nprobe, rsh, pass0, kmin);
                if (tid == 0) cnt_s = 0;
                __syncthreads();
                for (int i = tid; i < n; i += 256) {
                    const unsigned long long k = sel[i];
                    if (k <= thr) small[atomicAdd(&cnt_s, 1)] = k;  // exactly nprobe of them
                }
                __syncthreads();
                if (tid < nprobe) {
                    const unsigned long long k = small[tid];
                    int r = 0;
                    for (int j = 0; j < nprobe; ++j) r += small[j] < k;
                    probes[q * nprobe + r] = (uint32_t)(k & 0xFFFFFFFFu);
                }
                return;
            }
        }
    }
    const unsigned long long thr = block_radix_kth<KPT>(kv, key, K, nprobe, rsh);  // exactly nprobe keys <= thr
    if (tid == 0) cnt_s = 0;
    __syncthreads();
    if constexpr (KPT > 0) {
#pragma unroll
        for (int j = 0; j < KPT; ++j)
            if (tid + 256 * j < K && kv[j] <= thr) {
                const int p = atomicAdd(&cnt_s, 1);
                if (p < kSelMax) sel[p] = kv[j];
            }
    } else {
        for (int c0 = tid; c0 < K; c0 += 256 * 8) {
            unsigned long long kk[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) kk[u] = key(min(c0 + 256 * u, K - 1));
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (c0 + 256 * u < K && kk[u] <= thr) {
                    const int p = atomicAdd(&cnt_s, 1);
                    if (p < kSelMax) sel[p] = kk[u];
                }
        }
    }
    __syncthreads();
    const int n = min(cnt_s, kSelMax);
    const bool boundary_tie = rsh.last_pass >= 4;
    if (n <= 256) {
        bool tie = false;
        if (tid < n) {
            const unsigned long long k = sel[tid];
            int rank = 0, same = 0;
            for (int j = 0; j < n; ++j) {
                rank += sel[j] < k;
                same += (sel[j] >> 32) == (k >> 32);
            }
            tie = same > 1;
            if (rank < nprobe) probes[q * nprobe + rank] = (uint32_t)(k & 0xFFFFFFFFu);
        }
        if (ties) {
            tie = __syncthreads_or(tie);
            if (tid == 0) ties[q] = (tie || boundary_tie) ? 1 : 0;
        }
        return;
    }
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
    if (ties) {
        bool tie = false;
        for (int i = tid; i + 1 < n; i += 256) tie |= (sel[i] >> 32) == (sel[i + 1] >> 32);
        tie = __syncthreads_or(tie);
        if (tid == 0) ties[q] = (tie || boundary_tie) ? 1 : 0;
    }
}

hipError_t launch_sqnorm_rows(const float *x, int64_t rows, int D, float *out, hipStream_t st) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(sqnorm_rows_kernel, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, st, x, rows, D, out);
    return hipGetLastError();
}

// bfloat16 of x, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t bf16_rne(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Rows of x (D floats) as three bfloat16 parts of 3 D each, for a GEMM over K = 3 D whose
// products sum to hi.hi + hi.lo + lo.hi: hi = bf16(x), lo = bf16(x - hi).  Part order: centroid
// rows (second = 0) hi | lo | hi, query rows (second = 1) hi | hi | lo.
__global__ __launch_bounds__(256) void split_bf16_kernel(const float *__restrict__ x, int64_t rows, int D,
                                                         int second, uint16_t *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= rows * D) return;
    const int64_t r = e / D;
    const int i = (int)(e - r * D);
    const float v = x[e];
    const uint16_t hi = bf16_rne(v);
    const uint16_t lo = bf16_rne(v - __uint_as_float((uint32_t)hi << 16));
    uint16_t *o = out + r * 3 * D + i;
    o[0] = hi;
    o[D] = second ? hi : lo;
    o[2 * D] = second ? lo : hi;
}

hipError_t launch_split_bf16(const float *x, int64_t rows, int D, int second, uint16_t *out, hipStream_t st) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(split_bf16_kernel, dim3((unsigned)((rows * D + 255) / 256)), dim3(256), 0, st, x, rows, D,
                       second, out);
    return hipGetLastError();
}

// Top-nprobe selection over inner products ip (nq x K, from a library GEMM): keys are
// cnorm - 2 ip for L2 (cnorm = |c|^2), ip for IP.
hipError_t launch_probe_select_ip(const float *ip, const float *cnorm, int K, int64_t nq, int nprobe, int dist_type,
                                  uint32_t *probes, hipStream_t st) {
    if (nprobe > kSelMax) return hipErrorInvalidValue;
    if (nq == 0) return hipSuccess;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3((unsigned)nq), dim3(256), 0, st, ip, K, nprobe, probes, cnorm, (int32_t *)nullptr);
    };
    if (dist_type == 1)
        K <= 4096 ? go(probe_select_kernel<1, true, 16>) : go(probe_select_kernel<1, true, 0>);
    else
        K <= 4096 ? go(probe_select_kernel<0, true, 16>) : go(probe_select_kernel<0, true, 0>);
    return hipGetLastError();
}

// ---- exact coarse quantiser through a GEMM filter
// One workgroup per query.  key_c = |c|^2 - 2 q.c (L2) or -q.c (IP) from the fp32 GEMM products;
// the reference's distance d_c (Eigen order, fp32) differs from key_c + |q|^2 (L2) or -key_c (IP)
// by at most E_c = 2 (D + 2) u (|q| + |c|)^2 (u = 2^-24: the forward error bounds of the two
// fp32 sums, the squared differences and the final additions), used here with a factor 4 of
// margin.  If T is the nprobe-th smallest key and E the largest E_c, every centroid of the
// reference's top nprobe has key <= T + 2E: its distance is at most the nprobe-th smallest
// distance, which is at most the largest distance among the nprobe smallest keys.  Those
// candidates get the reference's distance (the same Eigen-order sum as probe_dist_kernel) and an
// exact (distance, id) selection; more than kExactCand candidates fall back to all K.
constexpr int kExactCand = 1024;

__device__ __forceinline__ float block_max256(float v, float *wtmp) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    if (lane == 0) wtmp[w] = v;
    __syncthreads();
    v = fmaxf(fmaxf(wtmp[0], wtmp[1]), fmaxf(wtmp[2], wtmp[3]));
    __syncthreads();
    return v;
}

// KPT > 0: K <= 256 KPT, and every thread keeps the keys of its centroids (tid + 256 j) in
// registers for the passes below; KPT = 0 forms them again in every pass.
// ties[q] as probe_select_kernel's (the candidates hold every centroid whose distance can reach the
// nprobe-th, ties at the boundary included); the all-K fallback always flags its query.
template <int DT, int KPT>
__global__ __launch_bounds__(256) void probe_exact_kernel(const float *__restrict__ ip, const float *__restrict__ cnorm,
                                                          const float *__restrict__ cent, int K, int D,
                                                          const float *__restrict__ queries, int nprobe,
                                                          uint32_t *__restrict__ probes, int32_t *__restrict__ ties) {
    extern __shared__ float qsh[];  // D floats
    __shared__ RadixLds rsh;
    __shared__ float wtmp[4];
    __shared__ int cnt_s;
    __shared__ unsigned long long sel[kExactCand];
    const int64_t q = blockIdx.x;
    const int tid = threadIdx.x;
    // |q|^2 summed in any order; its rounding, like that of |c|, is far inside the margin (the
    // norms are inflated by 1e-4 relative)
    float qn = 0.0f;
    for (int i = tid; i < D; i += 256) {
        const float x = queries[q * D + i];
        qsh[i] = x;
        qn += x * x;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) qn += __shfl_xor(qn, o);
    if ((tid & 63) == 0) wtmp[tid >> 6] = qn;
    __syncthreads();
    qn = (wtmp[0] + wtmp[1]) + (wtmp[2] + wtmp[3]);
    __syncthreads();
    const float qa = sqrtf(qn) * 1.0001f;
    const float scale = 8.0f * (float)(D + 2) * 5.9604645e-8f;  // 4 x 2 (D + 2) u
    const float *iq = ip + q * K;
    auto keyf = [&](int c) { return DT == 1 ? -iq[c] : cnorm[c] - 2.0f * iq[c]; };
    float kv[KPT > 0 ? KPT : 1];
    float emax = 0.0f;
    if constexpr (KPT > 0) {
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const int c = tid + 256 * j;
            kv[j] = c < K ? keyf(c) : 0.0f;
        }
    }
    for (int c = tid; c < K; c += 256) {
        const float cc = sqrtf(fmaxf(cnorm[c], 0.0f)) * 1.0001f;
        emax = fmaxf(emax, scale * (qa + cc) * (qa + cc));
    }
    emax = block_max256(emax, wtmp);
    // the nprobe-th smallest (key, id)
    const float *cn = cnorm;
    auto key64 = [=](int c) {
        const float kf = DT == 1 ? -iq[c] : cn[c] - 2.0f * iq[c];
        return ((unsigned long long)float_key(kf) << 32) | (uint32_t)c;
    };
    unsigned long long kk[KPT > 0 ? KPT : 1];
    kk[0] = 0;
    if constexpr (KPT > 0) {
#pragma unroll
        for (int j = 0; j < KPT; ++j)
            kk[j] = ((unsigned long long)float_key(kv[j]) << 32) | (uint32_t)(tid + 256 * j);
    }
    const unsigned long long thr = block_radix_kth<KPT>(kk, key64, K, nprobe, rsh);
    // T = the nprobe-th smallest key: the largest key <= thr
    float tkey = -INFINITY;
    if constexpr (KPT > 0) {
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const int c = tid + 256 * j;
            const unsigned long long k = ((unsigned long long)float_key(kv[j]) << 32) | (uint32_t)c;
            if (c < K && k <= thr) tkey = fmaxf(tkey, kv[j]);
        }
    } else {
        for (int c = tid; c < K; c += 256) {
            const float kf = keyf(c);
            const unsigned long long k = ((unsigned long long)float_key(kf) << 32) | (uint32_t)c;
            if (k <= thr) tkey = fmaxf(tkey, kf);
        }
    }
    tkey = block_max256(tkey, wtmp);
    const float lim = tkey + 2.0f * emax;
    if (tid == 0) cnt_s = 0;
    __syncthreads();
    if constexpr (KPT > 0) {
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const int c = tid + 256 * j;
            if (c < K && kv[j] <= lim) {
                const int p = atomicAdd(&cnt_s, 1);
                if (p < kExactCand) sel[p] = (uint32_t)c;
            }
        }
    } else {
        for (int c = tid; c < K; c += 256)
            if (keyf(c) <= lim) {
                const int p = atomicAdd(&cnt_s, 1);
                if (p < kExactCand) sel[p] = (uint32_t)c;
            }
    }
    __syncthreads();
    const bool all = cnt_s > kExactCand;  // too many candidates: every centroid
    const int n = all ? K : cnt_s;
    auto exact = [&](int c) {
        const float *row = cent + (int64_t)c * D;
        float d;
        if (DT == 1)
            d = eigen_redux_thread(D, [&](int i) { return qsh[i] * row[i]; });
        else
            d = eigen_redux_thread(D, [&](int i) {
                const float t = row[i] - qsh[i];
                return t * t;
            });
        return sel_key<DT>(d, c);
    };
    if (!all) {
        // the candidates' distances, two per wavefront (one per half-wave) at a time
        const int half = tid >> 5, l = tid & 31;
        for (int j0 = 0; j0 < n; j0 += 8) {
            const int j = j0 + half;
            const int c = j < n ? (int)sel[j] : 0;
            const float *row = cent + (int64_t)c * D;
            float d;
            if (DT == 1)
                d = eigen_redux_half(D, [&](int i) { return qsh[i] * row[i]; }, l);
            else
                d = eigen_redux_half(D, [&](int i) {
                    const float t = row[i] - qsh[i];
                    return t * t;
                }, l);
            __syncthreads();  // every half-wave has read its candidate id
            if (j < n && l == 0) sel[j] = sel_key<DT>(d, c);
            __syncthreads();
        }
        int P = 1;
        while (P < n) P <<= 1;
        for (int i = n + tid; i < P; i += 256) sel[i] = ~0ull;
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = tid; i < P; i += 256) {
                    const int ixj = i ^ j;
                    if (ixj > i) {
                        const unsigned long long x = sel[i], y = sel[ixj];
                        if ((x > y) == ((i & k) == 0)) {
                            sel[i] = y;
                            sel[ixj] = x;
                        }
                    }
                }
                __syncthreads();
            }
        for (int i = tid; i < nprobe; i += 256) probes[q * nprobe + i] = (uint32_t)(sel[i] & 0xFFFFFFFFu);
        bool tie = false;
        for (int i = tid; i < nprobe && i + 1 < n; i += 256) tie |= (sel[i] >> 32) == (sel[i + 1] >> 32);
        tie = __syncthreads_or(tie);
        if (tid == 0) ties[q] = tie ? 1 : 0;
        return;
    }
    // fallback: the nprobe smallest exact keys over all K by a second radix select
    const unsigned long long none[1] = {0};
    const unsigned long long ethr = block_radix_kth<0>(none, exact, K, nprobe, rsh);
    if (tid == 0) cnt_s = 0;
    __syncthreads();
    for (int c = tid; c < K; c += 256) {
        const unsigned long long k = exact(c);
        if (k <= ethr) {
            const int p = atomicAdd(&cnt_s, 1);
            if (p < kExactCand) sel[p] = k;
        }
    }
    __syncthreads();
    const int m = min(cnt_s, kExactCand);  // == nprobe <= kExactCand
    int P = 1;
    while (P < m) P <<= 1;
    for (int i = m + tid; i < P; i += 256) sel[i] = ~0ull;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < P; i += 256) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const unsigned long long x = sel[i], y = sel[ixj];
                    if ((x > y) == ((i & k) == 0)) {
                        sel[i] = y;
                        sel[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    for (int i = tid; i < nprobe; i += 256) probes[q * nprobe + i] = (uint32_t)(sel[i] & 0xFFFFFFFFu);
    if (tid == 0) ties[q] = 1;  // every key tied somewhere near the boundary: let the heap order them
}

// ---- std::partial_sort's order for exact ties
// The comparator of FlatInitializer::centroids_distances (initializer.hpp:61-70): Candidate's
// operator< (distance <) for L2 and std::greater<>, i.e. Candidate's operator> = !(a < b)
// (defines.hpp:52-54), for IP.
template <int DT>
__device__ __forceinline__ bool cand_less(float a, float b) {
    return DT == 1 ? !(a < b) : a < b;
}

// libstdc++'s __adjust_heap (with its __push_heap) on the heap hd/hi of len elements: the hole at
// `hole` sinks to a leaf along the larger children, then value (vd, vi) rises from there.
template <int DT>
__device__ void heap_adjust(float *hd, uint32_t *hi, int hole, int len, float vd, uint32_t vi) {
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (cand_less<DT>(hd[child], hd[child - 1])) child--;
        hd[hole] = hd[child];
        hi[hole] = hi[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        hd[hole] = hd[child - 1];
        hi[hole] = hi[child - 1];
        hole = child - 1;
    }
    int parent = (hole - 1) / 2;
    while (hole > top && cand_less<DT>(hd[parent], vd)) {
        hd[hole] = hd[parent];
        hi[hole] = hi[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    hd[hole] = vd;
    hi[hole] = vi;
}

// One workgroup per query; queries without a tie flag return at once.  recompute: dist holds GEMM
// products, replaced first by the reference's distances (Eigen order, by half-waves) of every
// centroid; otherwise dist already holds them.  Then one thread runs std::partial_sort (GCC 11
// libstdc++: __heap_select = __make_heap of [0, nprobe) + __pop_heap of every later element that
// compares below the top, then __sort_heap) over the candidates (id c, distance dist[c]) in id
// order.  The top of the heap only moves one way (down for L2; up for IP, whose comparator is
// !(a < b)), so a candidate that does not compare below the top at the start of its 256-candidate
// chunk never would later: the workgroup keeps those out, in id order, and one thread runs
// __pop_heap for the rest.
template <int DT>
__global__ __launch_bounds__(256) void probe_heap_kernel(float *__restrict__ dist, int recompute,
                                                         const float *__restrict__ cent, int K, int D,
                                                         const float *__restrict__ queries, int nprobe,
                                                         const int32_t *__restrict__ ties,
                                                         uint32_t *__restrict__ probes) {
    __shared__ float hd[kSelMax];
    __shared__ uint32_t hi[kSelMax];
    __shared__ float cd[256];
    __shared__ uint32_t ci[256];
    __shared__ int wcount[4];
    const int64_t q = blockIdx.x;
    if (!ties[q]) return;  // uniform per workgroup
    const int tid = threadIdx.x;
    float *row = dist + q * K;
    const float *qv = queries + q * D;
    if (recompute) {
        const int half = tid >> 5, l = tid & 31;
        for (int c0 = 0; c0 < K; c0 += 8) {
            const int c = min(c0 + half, K - 1);
            const float *cr = cent + (int64_t)c * D;
            float d;
            if (DT == 1)
                d = eigen_redux_half(D, [&](int i) { return qv[i] * cr[i]; }, l);
            else
                d = eigen_redux_half(D, [&](int i) {
                    const float t = cr[i] - qv[i];
                    return t * t;
                }, l);
            if (c0 + half < K && l == 0) row[c] = d;
        }
        __syncthreads();  // the row's writes are visible to the workgroup
    }
    // __make_heap of the first nprobe candidates
    for (int c = tid; c < nprobe; c += 256) {
        hd[c] = row[c];
        hi[c] = (uint32_t)c;
    }
    __syncthreads();
    if (tid == 0 && nprobe >= 2)
        for (int parent = (nprobe - 2) / 2;; --parent) {
            heap_adjust<DT>(hd, hi, parent, nprobe, hd[parent], hi[parent]);
            if (parent == 0) break;
        }
    __syncthreads();
    // __pop_heap(first, middle, i) for every later candidate that compares below the top
    const int lane = tid & 63, w = tid >> 6;
    for (int c0 = nprobe; c0 < K; c0 += 256) {
        const int c = c0 + tid;
        const float d = c < K ? row[c] : 0.0f;
        const bool pass = c < K && cand_less<DT>(d, hd[0]);
        const unsigned long long b = __ballot(pass);
        if (lane == 0) wcount[w] = __popcll(b);
        __syncthreads();
        int off = 0;
        for (int i = 0; i < w; ++i) off += wcount[i];
        const int n = wcount[0] + wcount[1] + wcount[2] + wcount[3];
        if (pass) {
            const int pos = off + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0));
            cd[pos] = d;
            ci[pos] = (uint32_t)c;
        }
        __syncthreads();
        if (tid == 0)
            for (int i = 0; i < n; ++i)
                if (cand_less<DT>(cd[i], hd[0])) heap_adjust<DT>(hd, hi, 0, nprobe, cd[i], ci[i]);
        __syncthreads();
    }
    if (tid == 0)
        for (int last = nprobe - 1; last >= 1; --last) {  // __sort_heap
            const float vd = hd[last];
            const uint32_t vi = hi[last];
            hd[last] = hd[0];
            hi[last] = hi[0];
            heap_adjust<DT>(hd, hi, 0, last, vd, vi);
        }
    __syncthreads();
    for (int i = tid; i < nprobe; i += 256) probes[q * nprobe + i] = hi[i];
}

// queries up to which a call computes every (query, centroid) distance before the tie pass
constexpr int64_t kHeapPrecomputeMax = 64;

// The exact coarse quantiser over GEMM products ip (nq x K) and centroid norms cnorm: probe
// lists equal to probe_tile + probe_select's.  nprobe <= kExactCand.
// ip is overwritten with exact distances in the rows of queries with ties; ties: nq ints.
hipError_t launch_probe_exact(float *ip, const float *cnorm, const float *centroids, int K, int D,
                              const float *queries, int64_t nq, int nprobe, int dist_type, uint32_t *probes,
                              int32_t *ties, hipStream_t st) {
    if (nprobe > kExactCand || (size_t)D * 4 > 48 * 1024) return hipErrorInvalidValue;
    if (nq == 0) return hipSuccess;
    const size_t sm = (size_t)D * 4;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3((unsigned)nq), dim3(256), sm, st, ip, cnorm, centroids, K, D, queries, nprobe,
                           probes, ties);
    };
    if (dist_type == 1)
        K <= 4096 ? go(probe_exact_kernel<1, 16>) : go(probe_exact_kernel<1, 0>);
    else
        K <= 4096 ? go(probe_exact_kernel<0, 16>) : go(probe_exact_kernel<0, 0>);
    // a query with ties needs every centroid's exact distance for std::partial_sort's order: for
    // a few queries they are computed by the whole chip beforehand (into ip, which the exact
    // kernel has finished with), for many by each flagged query's own workgroup
    const int recompute = nq > kHeapPrecomputeMax ? 1 : 0;
    if (!recompute) {
        const dim3 g1((unsigned)((K + 255) / 256), (unsigned)nq);
        if (dist_type == 1)
            hipLaunchKernelGGL(probe_dist_kernel<1>, g1, dim3(256), sm, st, centroids, K, D, queries, ip);
        else
            hipLaunchKernelGGL(probe_dist_kernel<0>, g1, dim3(256), sm, st, centroids, K, D, queries, ip);
    }
    if (dist_type == 1)
        hipLaunchKernelGGL(probe_heap_kernel<1>, dim3((unsigned)nq), dim3(256), 0, st, ip, recompute, centroids, K, D,
                           queries, nprobe, ties, probes);
    else
        hipLaunchKernelGGL(probe_heap_kernel<0>, dim3((unsigned)nq), dim3(256), 0, st, ip, recompute, centroids, K, D,
                           queries, nprobe, ties, probes);
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

// dist_ws: nq x K floats; ties: nq ints.
hipError_t launch_probe(const float *centroids, int K, int D, const float *queries, int64_t nq, int nprobe,
                        int dist_type, float *dist_ws, int32_t *ties, uint32_t *probes, hipStream_t st) {
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
        hipLaunchKernelGGL((probe_select_kernel<1, false>), dim3((unsigned)nq), dim3(256), 0, st, dist_ws, K, nprobe, probes,
                           nullptr, ties);
        hipLaunchKernelGGL(probe_heap_kernel<1>, dim3((unsigned)nq), dim3(256), 0, st, dist_ws, 0, centroids, K, D,
                           queries, nprobe, ties, probes);
    } else {
        if (tiled)
            probe_tile<0>(vec, gt, st, centroids, K, D, queries, nq, dist_ws);
        else
            hipLaunchKernelGGL(probe_dist_kernel<0>, g1, dim3(256), sm, st, centroids, K, D, queries, dist_ws);
        hipLaunchKernelGGL((probe_select_kernel<0, false>), dim3((unsigned)nq), dim3(256), 0, st, dist_ws, K, nprobe, probes,
                           nullptr, ties);
        hipLaunchKernelGGL(probe_heap_kernel<0>, dim3((unsigned)nq), dim3(256), 0, st, dist_ws, 0, centroids, K, D,
                           queries, nprobe, ties, probes);
    }
    return hipGetLastError();
}

}  // namespace saqgpu
