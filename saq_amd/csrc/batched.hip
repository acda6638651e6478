// batched.hip -- the batched, cluster-grouped SAQ scan on int8 MFMA (gfx950).
//
// The plain scan (slice_scan.hip) follows every query's probe sequence and reproduces the
// reference bit for bit.  This path serves many queries at once instead: the (query, list)
// pairs of a batch are grouped by list, and each work tile -- one list and up to 16*QTM of
// the queries that probe it -- turns the estimate of every vector of the list into a dense
// query x code contraction on v_mfma_i32_16x16x64_i8.  Each list's codes are then read from
// HBM once per query tile instead of once per query.
//
// Estimator.  For a segment with b > 0 bits the reference's accurate estimate is
// (caq_estimator.hpp:190-217, with c = msb * 2^(b-1) + low the full b-bit code):
//     ip_s = rescale * (sq_delta * sum_i r_i c_i + (-1 + sq_delta / 2) * sum_i r_i),
//     dist = sum_s (|o_s|^2 + |r_s|^2) - 2 * sum_{s: b>0} ip_s,        r = q_rot - c_rot.
// Here sum_i r_i c_i comes from the MFMA: r is quantised per (query, list, segment) to int8
// with an fp32 scale (one term, or two terms for int8x2: r ~ s1*q1 + s2*q2), codes are int8
// (b <= 7 as is, b = 8 shifted by -128, b >= 9 split into a high plane and a shifted low
// plane), and the shift is corrected through the per-query sum of the quantised r.  The
// fast-scan LUT stage is not needed: the full code gives the accurate estimate directly.
//
// Per-vector factors |o|^2 = sum_s o_l2norm_s^2, the ExFactor rescale of each quantised
// segment and the ids sit beside the codes in the same 16-vector record, so one tile streams
// one contiguous region per 16 vectors.
//
// Top-k.  Each tile keeps, per query, the k smallest (distance, id) pairs of the list in a
// wave-held sorted list (ties broken by id, so the result does not depend on how pairs are
// grouped into tiles); merge_kernel then merges each query's per-list lists.
#include <algorithm>
#include <cfloat>
#include <type_traits>

#include "batched.hpp"

namespace saqgpu {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------ layout conversion
// Low bits of the CodeHelper<k> compacted layouts (utils/code_helper.hpp:95-591) for
// k <= 8, one code.
__device__ int code_at_le8(int k, const uint8_t *y, int D, int d) {
    const int c64 = d / 64, r = d % 64, b = r % 16, part = r / 16;
    switch (k) {
    case 0: return 0;
    case 1: return (y[d / 8] >> (d % 8)) & 1;
    case 2: return (y[16 * c64 + b] >> (2 * part)) & 3;
    case 3: return ((y[24 * c64 + b] >> (2 * part)) & 3) | (((y[24 * c64 + 16 + r % 8] >> (r / 8)) & 1) << 2);
    case 4: {
        const uint8_t v = y[16 * (d / 32) + (d % 32) % 16];
        return (d % 32) < 16 ? (v & 15) : (v >> 4);
    }
    case 5: {  // 4-bit part, then a 1-bit part (code_helper.hpp:372-374)
        const uint8_t v = y[16 * (d / 32) + (d % 32) % 16];
        const int c4 = (d % 32) < 16 ? (v & 15) : (v >> 4);
        return 2 * c4 + ((y[D * 4 / 8 + d / 8] >> (d % 8)) & 1);
    }
    case 6:
    case 7: {
        const int base = (k == 6 ? 48 : 56) * c64;
        const uint8_t c1 = y[base + b], c2 = y[base + 16 + b], c3 = y[base + 32 + b];
        int v;
        if (part == 0) v = c1 & 0x3F;
        else if (part == 1) v = c2 & 0x3F;
        else if (part == 2) v = ((c1 >> 2) & 0x30) | (c3 & 0x0F);
        else v = ((c2 >> 2) & 0x30) | ((c3 >> 4) & 0x0F);
        if (k == 7) v |= ((y[base + 48 + r % 8] >> (r / 8)) & 1) << 6;
        return v;
    }
    default: return y[d];  // k == 8
    }
}

// k > 8: a plane of the low bytes, then the k-8 high bits (code_helper.hpp:85-88).
__device__ int code_at(int k, const uint8_t *y, int D, int d) {
    if (k > 8) return y[d] | (code_at_le8(k - 8, y + D, D, d) << 8);
    return code_at_le8(k, y, D, d);
}

// Full b-bit code of dimension d of segment s of vector vi (within list c), from the plain
// layout: the MSB from the block record's short code, the low bits from the long code.
__device__ int full_code(const IndexView &ix, const SegDesc &sd, int s_group_off, int c, int vi, int d) {
    const int64_t blk = ix.clu_blk[c] + vi / kBlock;
    const uint8_t *rec = ix.blocks + blk * ix.block_stride + (int64_t)ix.nseg * kBlock * 4;
    const uint8_t byte = rec[(int64_t)(s_group_off + d / 64) * 256 + (vi % kBlock) * 8 + (d / 8) % 8];
    const int msb = (byte >> (7 - d % 8)) & 1;
    if (sd.bits == 1) return msb;
    const uint8_t *lc = ix.long_codes + (ix.clu_vec[c] + vi) * ix.long_stride + sd.long_off;
    return (msb << (sd.bits - 1)) | code_at(sd.bits - 1, lc, sd.dim, d);
}

// grid (total mblocks, total 64-dim chunks of quantised segments), 64 threads: one lane's
// bytes of one chunk of one mblock.  Lane l holds vector l % 16, dims 16 * (l / 16) .. +16 of
// the chunk, in the order the MFMA's B operand takes them.
__global__ __launch_bounds__(64) void bq_build_codes_kernel(IndexView ix, BatchedView bv, const int32_t *mb_clu,
                                                            const int32_t *chunk_seg, const int32_t *chunk_kc,
                                                            uint8_t *out) {
    const int64_t mb = blockIdx.x;
    const int ch = blockIdx.y;
    const int l = threadIdx.x;
    const int c = mb_clu[mb];
    const int s = chunk_seg[ch], kc = chunk_kc[ch];
    const SegDesc sd = ix.segs[s];
    const BSeg bs = bv.segs[s];
    const int vi = (int)(mb - bv.clu_mb[c]) * kMB + l % kMB;
    const bool valid = vi < bv.clu_n[c];
    int code[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) code[t] = valid ? full_code(ix, sd, sd.group_off, c, vi, kc * 64 + 16 * (l / 16) + t) : 0;
    uint8_t *dst = out + mb * bv.mb_stride + bs.code_off + (int64_t)kc * bs.chunk_bytes;
    if (sd.bits <= 4) {
        // packed byte p: low nibble = dim (p < 4 ? p : p + 4), high nibble = that dim + 4
        uint8_t o[8];
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const int t = p < 4 ? p : p + 4;
            o[p] = (uint8_t)(code[t] | (code[t + 4] << 4));
        }
#pragma unroll
        for (int p = 0; p < 8; ++p) dst[l * 8 + p] = o[p];
    } else if (sd.bits <= 8) {
        const int off = sd.bits == 8 ? 128 : 0;
#pragma unroll
        for (int t = 0; t < 16; ++t) dst[l * 16 + t] = (uint8_t)(int8_t)(code[t] - off);
    } else {
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            dst[l * 16 + t] = (uint8_t)(int8_t)((code[t] & 255) - 128);
            dst[1024 + l * 16 + t] = (uint8_t)(code[t] >> 8);
        }
    }
}

// grid (total mblocks), 16 threads: |o|^2, -2 x the rescale of every quantised segment and the id.
__global__ __launch_bounds__(64) void bq_build_factors_kernel(IndexView ix, BatchedView bv, const int32_t *mb_clu,
                                                              uint8_t *out) {
    const int64_t mb = blockIdx.x;
    const int l = threadIdx.x;
    if (l >= kMB) return;
    const int c = mb_clu[mb];
    const int vi = (int)(mb - bv.clu_mb[c]) * kMB + l;
    const bool valid = vi < bv.clu_n[c];
    float *f = reinterpret_cast<float *>(out + mb * bv.mb_stride + bv.fac_off);
    uint32_t *idp = reinterpret_cast<uint32_t *>(f + kMB * (1 + bv.nqs));
    if (!valid) {
        f[l] = INFINITY;
        for (int q = 0; q < bv.nqs; ++q) f[kMB * (1 + q) + l] = 0.0f;
        idp[l] = 0xFFFFFFFFu;
        return;
    }
    const int64_t blk = ix.clu_blk[c] + vi / kBlock;
    const float *norms = reinterpret_cast<const float *>(ix.blocks + blk * ix.block_stride);
    const int64_t gv = ix.clu_vec[c] + vi;
    float o2 = 0.0f;
    for (int s = 0; s < bv.nseg; ++s) {
        const float x = norms[s * kBlock + vi % kBlock];
        o2 = o2 + x * x;
        const int q = bv.segs[s].qs;
        if (q >= 0) f[kMB * (1 + q) + l] = -2.0f * ix.exfac[(gv * bv.nseg + s) * 2];  // dist -= 2 rescale ip
    }
    f[l] = o2;
    idp[l] = ix.ids[gv];
}

// ------------------------------------------------------------------ grouping
// Zeroes the per-list pair counts and the tile counter of a launch.
// Also sets every query's distance bound to +inf (a key above every finite distance's).
__global__ void bq_zero_kernel(int32_t *counts, int K, int32_t *work, uint32_t *qbound, int64_t nq) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < K; i += gridDim.x * blockDim.x) counts[i] = 0;
    for (int64_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x)
        qbound[i] = 0xFFFFFFFFu;
    if (blockIdx.x == 0 && threadIdx.x == 0) *work = 0;
}

// (q, p) -> slot within its list's pair group
__global__ void bq_count_kernel(const uint32_t *probes, const int32_t *npq, int64_t nq, int NP, int32_t *counts,
                                int32_t *slot) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nq * NP) return;
    const int64_t q = i / NP;
    const int p = (int)(i % NP);
    const int lim = npq ? npq[q] : NP;
    if (p >= lim) {
        slot[i] = -1;
        return;
    }
    slot[i] = atomicAdd(&counts[probes[i]], 1);
}

// One workgroup.  The split size first: W = sum over lists of (query tiles x list vectors) is
// the launch's work in vector-tiles, and a list is cut into ranges of sv = max(kMinSplit,
// W / workers) vectors (a multiple of 64; workers = the tile kernel's resident workgroups, at most
// kSplitWaves), so no work tile holds more than about one resident workgroup's share and one long
// list cannot set the kernel's tail.  Then exclusive
// prefix sums, over the lists in the given order (largest first, so the longest tiles start
// first), of the pair counts, of the work tiles (query tiles x ranges) and of the candidate
// lists (pairs x ranges); bq_tiles_kernel writes the tiles themselves.  The bounds sum_c tiles_c <= npairs / QT
// + K + kSplitWaves and sum_c cands_c <= npairs + kSplitWaves * QT follow from
// sv >= W / kSplitWaves.
__global__ __launch_bounds__(1024) void bq_scan_kernel(const int32_t *counts, const int32_t *order,
                                                       const int32_t *clu_n, int K, int QT, int workers, int32_t *cstart,
                                                       int32_t *candstart, int32_t *nsplit, int32_t *ntiles,
                                                       int32_t *npairs, int32_t *sv_out, int32_t *tfirst) {
    // thread t holds lists order[t * per .. t * per + per - 1] (per <= kScanRounds)
    __shared__ int wsum[3][16];
    __shared__ long long wk_s[16];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int per = (K + 1023) / 1024;
    int oc[kScanRounds], cn[kScanRounds], nv[kScanRounds];
#pragma unroll
    for (int j = 0; j < kScanRounds; ++j) oc[j] = j < per && t * per + j < K ? order[t * per + j] : -1;
#pragma unroll
    for (int j = 0; j < kScanRounds; ++j) {
        cn[j] = oc[j] >= 0 ? counts[oc[j]] : 0;
        nv[j] = oc[j] >= 0 ? clu_n[oc[j]] : 0;
    }
    // per is uniform: the slots past it (all zero) are skipped, with their integer divisions
    long long wk = 0;
#pragma unroll
    for (int j = 0; j < kScanRounds; ++j) {
        if (j >= per) break;
        wk += (long long)((cn[j] + QT - 1) / QT) * nv[j];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) wk += __shfl_xor(wk, o);
    if (lane == 0) wk_s[wv] = wk;
    __syncthreads();
    long long W = 0;
    for (int i = 0; i < 16; ++i) W += wk_s[i];
    const long long want = (W + workers - 1) / workers;
    const int sv = (int)max((long long)kMinSplit, (want + 63) / 64 * 64);
    int ns[kScanRounds];
    int sc = 0, st = 0, sk = 0;
#pragma unroll
    for (int j = 0; j < kScanRounds; ++j) {
        ns[j] = 1;
        if (j >= per) continue;
        ns[j] = max(1, (nv[j] + sv - 1) / sv);
        sc += cn[j];
        st += (cn[j] + QT - 1) / QT * ns[j];
        sk += cn[j] * ns[j];
    }
    // exclusive prefix of the three per-thread sums over the workgroup
    int ic = sc, it = st, ik = sk;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int uc = __shfl_up(ic, o), ut = __shfl_up(it, o), uk = __shfl_up(ik, o);
        if (lane >= o) {
            ic += uc;
            it += ut;
            ik += uk;
        }
    }
    if (lane == 63) {
        wsum[0][wv] = ic;
        wsum[1][wv] = it;
        wsum[2][wv] = ik;
    }
    __syncthreads();
    int bc = 0, bt = 0, bk = 0, tc = 0, tt = 0, tk = 0;
    for (int i = 0; i < 16; ++i) {
        const bool before = i < wv;
        bc += before ? wsum[0][i] : 0;
        bt += before ? wsum[1][i] : 0;
        bk += before ? wsum[2][i] : 0;
        tc += wsum[0][i];
        tt += wsum[1][i];
        tk += wsum[2][i];
    }
    int rc = bc + ic - sc, rt = bt + it - st, rk = bk + ik - sk;
#pragma unroll
    for (int j = 0; j < kScanRounds; ++j) {
        if (j >= per) break;
        if (oc[j] < 0) continue;
        const int c = oc[j], n = cn[j];
        cstart[c] = rc;
        candstart[c] = rk;
        nsplit[c] = ns[j];
        tfirst[t * per + j] = rt;
        rt += (n + QT - 1) / QT * ns[j];
        rc += n;
        rk += n * ns[j];
    }
    if (t == 0) {
        *ntiles = tt;
        *npairs = tc;
        *sv_out = sv;
    }
}

// Every work tile, one thread each: every query tile of QT of a list's pairs times every vector
// range of the list, in that order from the list's first tile; the candidate list of (pair slot i,
// range r) is candstart[c] + r * n + i.  The list holding tile rt is the last one in the scan
// order whose first tile is at or before rt (a list without tiles shares its successor's first
// tile, and the search passes over it).
__global__ void bq_tiles_kernel(const int32_t *order, const int32_t *tfirst, int K, int QT, const int32_t *counts,
                                const int32_t *cstart, const int32_t *candstart, const int32_t *nsplit,
                                const int32_t *ntiles, const int32_t *sv_p, int32_t *tile_clu, int32_t *tile_first,
                                int32_t *tile_cnt, int32_t *tile_v0, int32_t *tile_cand) {
    const int nt = *ntiles, sv = *sv_p;
    for (int rt = blockIdx.x * blockDim.x + threadIdx.x; rt < nt; rt += gridDim.x * blockDim.x) {
        int lo = 0, hi = K;  // the first list whose first tile is after rt
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (tfirst[mid] <= rt)
                lo = mid + 1;
            else
                hi = mid;
        }
        const int jj = lo - 1;  // >= 0: tfirst[0] = 0
        const int c = order[jj], n = counts[c], ns = nsplit[c];
        const int local = rt - tfirst[jj], qi = local / ns, r = local - qi * ns, i = qi * QT;
        tile_clu[rt] = c;
        tile_first[rt] = cstart[c] + i;
        tile_cnt[rt] = min(QT, n - i);
        tile_v0[rt] = r * sv;
        tile_cand[rt] = candstart[c] + r * n + i;
    }
}

__global__ void bq_scatter_kernel(const uint32_t *probes, const int32_t *slot, int64_t n, const int32_t *cstart,
                                  int32_t *pairs) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n || slot[i] < 0) return;
    pairs[cstart[probes[i]] + slot[i]] = (int32_t)i;
}

// ------------------------------------------------------------------ wave-held top-k
__device__ __forceinline__ bool dless(float a, uint32_t ai, float b, uint32_t bi) {
    return a < b || (a == b && ai < bi);
}

// Sorted list of up to 64 * KR (distance, id) pairs; element i sits in register i / 64 of
// lane i % 64.
template <int KR>
struct WaveTopK {
    float d[KR];
    uint32_t id[KR];
    int n, k;
    float kd;
    uint32_t kid;

    __device__ void refresh() {
        if (n < k) {
            kd = INFINITY;
            kid = 0xFFFFFFFFu;
        } else {
            const int r = (k - 1) >> 6, ln = (k - 1) & 63;
            float x = d[0];
            uint32_t xi = id[0];
#pragma unroll
            for (int j = 1; j < KR; ++j)
                if (j == r) {
                    x = d[j];
                    xi = id[j];
                }
            kd = __shfl(x, ln);
            kid = __shfl(xi, ln);
        }
    }
    __device__ void load(const float *sd, const uint32_t *si, int cnt, int kk) {
        const int lane = threadIdx.x & 63;
        k = kk;
        n = cnt;
#pragma unroll
        for (int j = 0; j < KR; ++j) {
            const int i = j * 64 + lane;
            d[j] = i < n ? sd[i] : INFINITY;
            id[j] = i < n ? si[i] : 0xFFFFFFFFu;
        }
        refresh();
    }
    __device__ void store(float *sd, uint32_t *si) const {
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int j = 0; j < KR; ++j) {
            const int i = j * 64 + lane;
            if (i < n) {
                sd[i] = d[j];
                si[i] = id[j];
            }
        }
    }
    // x must be less than the current k-th pair
    __device__ void insert(float x, uint32_t xi) {
        const int lane = threadIdx.x & 63;
        int pos = 0;
#pragma unroll
        for (int j = 0; j < KR; ++j) {
            const int i = j * 64 + lane;
            pos += __popcll(__ballot(i < n && dless(d[j], id[j], x, xi)));
        }
        float nd[KR];
        uint32_t ni[KR];
#pragma unroll
        for (int j = 0; j < KR; ++j) {
            float up = __shfl_up(d[j], 1);
            uint32_t upi = __shfl_up(id[j], 1);
            if (j > 0) {
                const float w = __shfl(d[j - 1], 63);
                const uint32_t wi = __shfl(id[j - 1], 63);
                if (lane == 0) {
                    up = w;
                    upi = wi;
                }
            }
            const int i = j * 64 + lane;
            nd[j] = i > pos ? up : (i == pos ? x : d[j]);
            ni[j] = i > pos ? upi : (i == pos ? xi : id[j]);
        }
#pragma unroll
        for (int j = 0; j < KR; ++j) {
            d[j] = nd[j];
            id[j] = ni[j];
        }
        n = min(n + 1, k);
        refresh();
    }
    // offers one candidate per lane
    __device__ void offer(float x, uint32_t xi) {
        unsigned long long m = __ballot(dless(x, xi, kd, kid));
        while (m) {
            const int j = __builtin_ctzll(m);
            m &= m - 1;
            const float bx = __shfl(x, j);
            const uint32_t bi = __shfl(xi, j);
            if (dless(bx, bi, kd, kid)) insert(bx, bi);
        }
    }
};

// ------------------------------------------------------------------ buffered selection (k <= 64)
// A wave-wide sorted list of 64 (distance, id) pairs (lane l holds the l-th smallest; unused
// places hold (INF, 0xFFFFFFFF)) absorbs up to 64 candidates at once: the candidates are
// sorted by a bitonic network across the lanes, reversed and taken pairwise-smaller against
// the list (the 64 smallest of both, as a bitonic sequence), then a bitonic merge sorts them.
// The value of lane (lane ^ J): DPP quad permutations for J = 1, 2 (a VALU operand modifier),
// ds_swizzle for J = 4, 8, 16 (no address operand), ds_bpermute for J = 32.
template <int J>
__device__ __forceinline__ int xor_lane(int v) {
    if constexpr (J == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    else if constexpr (J == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    else if constexpr (J < 32) return __builtin_amdgcn_ds_swizzle(v, 0x1F | (J << 10));  // bit mode: lane ^ J
    else return __shfl_xor(v, 32);
}

template <int J>
__device__ __forceinline__ void cswap_lane(float &d, uint32_t &id, bool keep_small) {
    const float pd = __int_as_float(xor_lane<J>(__float_as_int(d)));
    const uint32_t pi = (uint32_t)xor_lane<J>((int)id);
    const bool other_less = dless(pd, pi, d, id);
    if (other_less == keep_small) {
        d = pd;
        id = pi;
    }
}

// one stage (k) of the bitonic sort: steps j = k/2 .. 1
template <int K, int J>
__device__ __forceinline__ void bitonic_steps(float &d, uint32_t &id, int lane) {
    if constexpr (J >= 1) {
        cswap_lane<J>(d, id, ((lane & J) == 0) == ((lane & K) == 0 || K == 64));
        bitonic_steps<K, J / 2>(d, id, lane);
    }
}

// ascending sort of one pair per lane (a full bitonic sort network)
__device__ __forceinline__ void sort64(float &cd, uint32_t &ci) {
    const int lane = threadIdx.x & 63;
    bitonic_steps<2, 1>(cd, ci, lane);
    bitonic_steps<4, 2>(cd, ci, lane);
    bitonic_steps<8, 4>(cd, ci, lane);
    bitonic_steps<16, 8>(cd, ci, lane);
    bitonic_steps<32, 16>(cd, ci, lane);
    bitonic_steps<64, 32>(cd, ci, lane);
}

// ascending order of a bitonic sequence of one pair per lane (a bitonic merge)
__device__ __forceinline__ void merge64(float &d, uint32_t &id) {
    const int lane = threadIdx.x & 63;
    cswap_lane<32>(d, id, (lane & 32) == 0);
    cswap_lane<16>(d, id, (lane & 16) == 0);
    cswap_lane<8>(d, id, (lane & 8) == 0);
    cswap_lane<4>(d, id, (lane & 4) == 0);
    cswap_lane<2>(d, id, (lane & 2) == 0);
    cswap_lane<1>(d, id, (lane & 1) == 0);
}

__device__ __forceinline__ void reverse64(float &d, uint32_t &id) {
    const int lane = threadIdx.x & 63;
    d = __shfl(d, 63 - lane);
    id = (uint32_t)__shfl((int)id, 63 - lane);
}

// The pairwise-smaller of (a, ai) and (b, bi) into a, the larger into b.
__device__ __forceinline__ void order_pair(float &a, uint32_t &ai, float &b, uint32_t &bi) {
    if (dless(b, bi, a, ai)) {
        const float t = a;
        const uint32_t ti = ai;
        a = b;
        ai = bi;
        b = t;
        bi = ti;
    }
}

// A sorted list of 64 KR pairs (entry 64 r + lane in d[r] of lane; KR = 1 or 2) takes up to 64
// candidates (one per lane, (INF, 0xFFFFFFFF) for none) and keeps the 64 KR smallest of both,
// sorted.  The candidates are sorted and reversed; against the list's last 64 entries L they form
// a bitonic sequence whose pairwise-smaller half Y holds the 64 smallest of L and the candidates
// (the larger half lies above all of them and above the list's first 64 entries, so it is
// dropped).  SORTED: the candidates arrive sorted.  For KR = 2, Y is sorted, reversed and split against the first 64 entries the same
// way into the new first and last 64 entries.
template <int KR, bool SORTED = false>
__device__ __forceinline__ void merge_into(float (&d)[KR], uint32_t (&id)[KR], float cd, uint32_t ci) {
    if (!SORTED) sort64(cd, ci);
    reverse64(cd, ci);
    order_pair(d[KR - 1], id[KR - 1], cd, ci);
    merge64(d[KR - 1], id[KR - 1]);
    if constexpr (KR == 2) {
        reverse64(d[1], id[1]);
        order_pair(d[0], id[0], d[1], id[1]);
        merge64(d[0], id[0]);
        merge64(d[1], id[1]);
    }
}

// ------------------------------------------------------------------ pair preparation
// One wave per grouped (query, list) pair: the residual r = q_rot - c_rot of every segment,
// quantised to int8 (plus the second int8 term for int8x2) as the pair's row of the MFMA A
// operand, and the pair's constants.  Rows are stored in grouped order, so a work tile's
// queries are consecutive rows.
struct PairPrepArgs {
    BatchedView bv;
    const float *qrot;  // nq x D_pad
    const uint32_t *probes;
    const int32_t *pairs;
    const int32_t *npairs;  // active pairs (device scalar)
    int NP, qs;
    int8_t *qq1, *qq2;  // [pairs][kq]
    float *pc;          // [pairs][nqs][3]: a1 = sqd * s1, a2 = sqd * s2, and the constant term
                        // (-1 + sqd / 2) * sum r + a1 * corr1 + a2 * corr2 (corr: the code shift's
                        // correction 128 * sum q of each term, b >= 8)
    float *pq2;         // [pairs]: |r|^2 over all segments
};

// BF16 form of a segment's residual (precision SAQGPU_PREC_BF16): r / amax rounded to bf16 (an fp32
// scale s1 = amax), stored 2 bytes per dimension in the row; the tile kernel uses the codes unshifted,
// so the constant term is (-1 + sqd / 2) * sum r alone.  r[k] holds the 4 dimensions from
// base + 4 lane + 256 k of the padded vector; those outside the segment are skipped.
__device__ __forceinline__ uint32_t bf16_pair(float x, float y) {
    // round to nearest even (finite inputs)
    auto rne = [](float v) {
        const uint32_t u = __float_as_uint(v);
        return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
    };
    return rne(x) | (rne(y) << 16);
}

template <int NK>
__device__ __forceinline__ void bf16_segment(const PairPrepArgs &a, const BSeg &bs, int64_t g, const float4 (&r)[NK],
                                             int base, float amax, float sr) {
    const int lane = threadIdx.x & 63;
    const float inv = amax > 0.0f ? 1.0f / amax : 0.0f;
    uint16_t *row = reinterpret_cast<uint16_t *>(a.qq1) + g * a.bv.kq + bs.kq_off;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        const int d = base + 4 * lane + 256 * k;  // absolute padded dimension of r[k].x
        if (d < bs.doff || d >= bs.doff + bs.dim) continue;
        uint2 v;
        v.x = bf16_pair(r[k].x * inv, r[k].y * inv);
        v.y = bf16_pair(r[k].z * inv, r[k].w * inv);
        *reinterpret_cast<uint2 *>(row + (d - bs.doff)) = v;
    }
    if (lane == 0) {
        float *pc = a.pc + (g * a.bv.nqs + bs.qs) * 3;
        pc[0] = bs.sqd * amax;
        pc[1] = 0.0f;
        pc[2] = (-1.0f + bs.sqd / 2) * sr;
    }
}

// The pair's constants of one quantised segment: the term scales a1 = sqd * s1, a2 = sqd * s2 and
// the constant term of its inner product, (-1 + sqd / 2) * sum r plus the code shift's correction
// a_t * 128 * sum q_t of each term (codes of 8 and more bits are stored shifted by -128), so that
// the tile kernel forms ip_s = a1 * acc1 + a2 * acc2 + c with fused multiply-adds.
__device__ __forceinline__ void write_pair_consts(const PairPrepArgs &a, const BSeg &bs, int64_t g, float s1, float s2,
                                                  float sr, int sum1, int sum2) {
    const float a1 = bs.sqd * s1, a2 = bs.sqd * s2;
    float c = (-1.0f + bs.sqd / 2) * sr;
    if (bs.bits >= 8) {
        c = fmaf(a1, (float)(128 * sum1), c);
        if (a.qs == 2) c = fmaf(a2, (float)(128 * sum2), c);
    }
    float *pc = a.pc + (g * a.bv.nqs + bs.qs) * 3;
    pc[0] = a1;
    pc[1] = a2;
    pc[2] = c;
}

// One butterfly step (partner lane ^ J) of wave reductions, through xor_lane's DPP / swizzle.
template <int J>
__device__ __forceinline__ float xor_f(float v) {
    return __int_as_float(xor_lane<J>(__float_as_int(v)));
}

// Quantises one segment's residual r (held as float4 per lane: element 4 lane + 256 k) and
// writes the pair's row bytes and constants; NK = float4 registers per lane (dim <= 256 NK).
template <int NK>
__device__ __forceinline__ void pairprep_segment(const PairPrepArgs &a, const BSeg &bs, int64_t g, const float *qr,
                                                 const float *cr, float &q2) {
    const int lane = threadIdx.x & 63;
    const BatchedView &bv = a.bv;
    float4 r[NK];
    float amax = 0.0f, sr = 0.0f, sq = 0.0f;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        const int d = 4 * lane + 256 * k;
        if (d < bs.dim) {
            const float4 x = *reinterpret_cast<const float4 *>(qr + bs.doff + d);
            const float4 y = *reinterpret_cast<const float4 *>(cr + bs.doff + d);
            r[k] = make_float4(x.x - y.x, x.y - y.y, x.z - y.z, x.w - y.w);
        } else {
            r[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        amax = fmaxf(amax, fmaxf(fmaxf(fabsf(r[k].x), fabsf(r[k].y)), fmaxf(fabsf(r[k].z), fabsf(r[k].w))));
        sr += (r[k].x + r[k].y) + (r[k].z + r[k].w);
        sq += (r[k].x * r[k].x + r[k].y * r[k].y) + (r[k].z * r[k].z + r[k].w * r[k].w);
    }
    auto red3 = [&](auto step) {
        sq += step(sq);
        sr += step(sr);
        amax = fmaxf(amax, step(amax));
    };
    red3([](float v) { return xor_f<32>(v); });
    red3([](float v) { return xor_f<16>(v); });
    red3([](float v) { return xor_f<8>(v); });
    red3([](float v) { return xor_f<4>(v); });
    red3([](float v) { return xor_f<2>(v); });
    red3([](float v) { return xor_f<1>(v); });
    q2 += sq;
    if (bs.qs < 0) return;
    if (a.qs == 3) {
        bf16_segment<NK>(a, bs, g, r, bs.doff, amax, sr);
        return;
    }
    const float s1 = amax / 127.0f, inv1 = amax > 0.0f ? 127.0f / amax : 0.0f;
    const float s2 = s1 / 254.0f, inv2 = amax > 0.0f ? 254.0f * 127.0f / amax : 0.0f;
    int sum1 = 0, sum2 = 0;
    int8_t *row1 = a.qq1 + g * bv.kq + bs.kq_off;
    int8_t *row2 = a.qs == 2 ? a.qq2 + g * bv.kq + bs.kq_off : nullptr;
    auto q8 = [](float x) { return (int)fminf(127.0f, fmaxf(-127.0f, rintf(x))); };
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        const int d = 4 * lane + 256 * k;
        if (d >= bs.dim) continue;
        const float e[4] = {r[k].x, r[k].y, r[k].z, r[k].w};
        int v1[4], v2[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            v1[u] = q8(e[u] * inv1);
            sum1 += v1[u];
            v2[u] = q8((e[u] - (float)v1[u] * s1) * inv2);
        }
        *reinterpret_cast<uint32_t *>(row1 + d) =
            (uint32_t)(uint8_t)v1[0] | ((uint32_t)(uint8_t)v1[1] << 8) | ((uint32_t)(uint8_t)v1[2] << 16) |
            ((uint32_t)(uint8_t)v1[3] << 24);
        if (row2) {
            sum2 += (v2[0] + v2[1]) + (v2[2] + v2[3]);
            *reinterpret_cast<uint32_t *>(row2 + d) =
                (uint32_t)(uint8_t)v2[0] | ((uint32_t)(uint8_t)v2[1] << 8) | ((uint32_t)(uint8_t)v2[2] << 16) |
                ((uint32_t)(uint8_t)v2[3] << 24);
        }
    }
    sum1 += xor_lane<32>(sum1), sum2 += xor_lane<32>(sum2);
    sum1 += xor_lane<16>(sum1), sum2 += xor_lane<16>(sum2);
    sum1 += xor_lane<8>(sum1), sum2 += xor_lane<8>(sum2);
    sum1 += xor_lane<4>(sum1), sum2 += xor_lane<4>(sum2);
    sum1 += xor_lane<2>(sum1), sum2 += xor_lane<2>(sum2);
    sum1 += xor_lane<1>(sum1), sum2 += xor_lane<1>(sum2);
    if (lane == 0) write_pair_consts(a, bs, g, s1, s2, sr, sum1, sum2);
}

// The pair's whole residual in NKT float4 registers per lane (element 4 lane + 256 k of the
// padded vector, D_pad <= 256 NKT), loaded in one round; then every segment from the lanes'
// elements inside it (segment offsets and dimensions are multiples of 64, so no float4 straddles
// two segments).  Same quantisation and constants as pairprep_segment.
template <int NKT>
__device__ __forceinline__ void pairprep_whole(const PairPrepArgs &a, const BSeg *segs, int64_t g, const float *qr,
                                               const float *cr) {
    const int lane = threadIdx.x & 63;
    const BatchedView &bv = a.bv;
    float4 r[NKT];
#pragma unroll
    for (int k = 0; k < NKT; ++k) {
        const int d = 4 * lane + 256 * k;
        if (d < bv.D_pad) {
            const float4 x = *reinterpret_cast<const float4 *>(qr + d);
            const float4 y = *reinterpret_cast<const float4 *>(cr + d);
            r[k] = make_float4(x.x - y.x, x.y - y.y, x.z - y.z, x.w - y.w);
        } else {
            r[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    float q2 = 0.0f;
    for (int s = 0; s < bv.nseg; ++s) {
        const BSeg bs = segs[s];
        float amax = 0.0f, sr = 0.0f, sq = 0.0f;
#pragma unroll
        for (int k = 0; k < NKT; ++k) {
            const int d = 4 * lane + 256 * k;
            if (d >= bs.doff && d < bs.doff + bs.dim) {
                amax = fmaxf(amax, fmaxf(fmaxf(fabsf(r[k].x), fabsf(r[k].y)), fmaxf(fabsf(r[k].z), fabsf(r[k].w))));
                sr += (r[k].x + r[k].y) + (r[k].z + r[k].w);
                sq += (r[k].x * r[k].x + r[k].y * r[k].y) + (r[k].z * r[k].z + r[k].w * r[k].w);
            }
        }
        q2 += sq;  // |r|^2 is only needed whole: one reduction after the segments
        if (bs.qs < 0) continue;
        sr += xor_f<32>(sr), amax = fmaxf(amax, xor_f<32>(amax));
        sr += xor_f<16>(sr), amax = fmaxf(amax, xor_f<16>(amax));
        sr += xor_f<8>(sr), amax = fmaxf(amax, xor_f<8>(amax));
        sr += xor_f<4>(sr), amax = fmaxf(amax, xor_f<4>(amax));
        sr += xor_f<2>(sr), amax = fmaxf(amax, xor_f<2>(amax));
        sr += xor_f<1>(sr), amax = fmaxf(amax, xor_f<1>(amax));
        if (a.qs == 3) {
            bf16_segment<NKT>(a, bs, g, r, 0, amax, sr);
            continue;
        }
        const float s1 = amax / 127.0f, inv1 = amax > 0.0f ? 127.0f / amax : 0.0f;
        const float s2 = s1 / 254.0f, inv2 = amax > 0.0f ? 254.0f * 127.0f / amax : 0.0f;
        int sum1 = 0, sum2 = 0;
        int8_t *row1 = a.qq1 + g * bv.kq + bs.kq_off;
        int8_t *row2 = a.qs == 2 ? a.qq2 + g * bv.kq + bs.kq_off : nullptr;
#pragma unroll
        for (int k = 0; k < NKT; ++k) {
            const int d = 4 * lane + 256 * k;
            if (d < bs.doff || d >= bs.doff + bs.dim) continue;
            const float e[4] = {r[k].x, r[k].y, r[k].z, r[k].w};
            int v1[4], v2[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                v1[u] = (int)fminf(127.0f, fmaxf(-127.0f, rintf(e[u] * inv1)));
                sum1 += v1[u];
                v2[u] = (int)fminf(127.0f, fmaxf(-127.0f, rintf((e[u] - (float)v1[u] * s1) * inv2)));
            }
            *reinterpret_cast<uint32_t *>(row1 + d - bs.doff) =
                (uint32_t)(uint8_t)v1[0] | ((uint32_t)(uint8_t)v1[1] << 8) | ((uint32_t)(uint8_t)v1[2] << 16) |
                ((uint32_t)(uint8_t)v1[3] << 24);
            if (row2) {
                sum2 += (v2[0] + v2[1]) + (v2[2] + v2[3]);
                *reinterpret_cast<uint32_t *>(row2 + d - bs.doff) =
                    (uint32_t)(uint8_t)v2[0] | ((uint32_t)(uint8_t)v2[1] << 8) | ((uint32_t)(uint8_t)v2[2] << 16) |
                    ((uint32_t)(uint8_t)v2[3] << 24);
            }
        }
        if (bs.bits >= 8) {  // the code shift's correction: only 8 and more bits are stored shifted
            sum1 += xor_lane<32>(sum1), sum2 += xor_lane<32>(sum2);
            sum1 += xor_lane<16>(sum1), sum2 += xor_lane<16>(sum2);
            sum1 += xor_lane<8>(sum1), sum2 += xor_lane<8>(sum2);
            sum1 += xor_lane<4>(sum1), sum2 += xor_lane<4>(sum2);
            sum1 += xor_lane<2>(sum1), sum2 += xor_lane<2>(sum2);
            sum1 += xor_lane<1>(sum1), sum2 += xor_lane<1>(sum2);
        }
        if (lane == 0) write_pair_consts(a, bs, g, s1, s2, sr, sum1, sum2);
    }
    q2 += xor_f<32>(q2);
    q2 += xor_f<16>(q2);
    q2 += xor_f<8>(q2);
    q2 += xor_f<4>(q2);
    q2 += xor_f<2>(q2);
    q2 += xor_f<1>(q2);
    if (lane == 0) a.pq2[g] = q2;
}

// One wave per grouped pair.  NKT > 0: vectors of up to 256 NKT padded dimensions, loaded whole
// in one round; NKT = 0: segment by segment, segments up to 2048 dimensions held in registers (a
// segment dimension and offset are multiples of 64, rows are 16-byte aligned).
template <int NKT>
__global__ __launch_bounds__(256) void bq_pairprep_kernel(PairPrepArgs a) {
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    // the segment descriptors in LDS: each segment's step then reads them without a global round trip
    __shared__ BSeg seg_s[kMaxSeg];
    for (int s = threadIdx.x; s < a.bv.nseg; s += 256) seg_s[s] = a.bv.segs[s];
    __syncthreads();
    if (g >= *a.npairs) return;
    BatchedView bv = a.bv;
    bv.segs = seg_s;
    const int pair = a.pairs[g];
    const int64_t q = pair / a.NP;
    const int c = (int)a.probes[pair];
    const float *qr = a.qrot + q * bv.D_pad;
    const float *cr = bv.crot + (int64_t)c * bv.D_pad;
    if constexpr (NKT > 0) {
        pairprep_whole<NKT>(a, bv.segs, g, qr, cr);
        return;
    }
    float q2 = 0.0f;
    for (int s = 0; s < bv.nseg; ++s) {
        const BSeg bs = bv.segs[s];
        if (bs.dim <= 256)
            pairprep_segment<1>(a, bs, g, qr, cr, q2);
        else if (bs.dim <= 512)
            pairprep_segment<2>(a, bs, g, qr, cr, q2);
        else if (bs.dim <= 1024)
            pairprep_segment<4>(a, bs, g, qr, cr, q2);
        else
            pairprep_segment<8>(a, bs, g, qr, cr, q2);
    }
    if (lane == 0) a.pq2[g] = q2;
}

// ------------------------------------------------------------------ tile kernel
// Persistent.  A workgroup of W waves takes one work tile at a time (one list, one vector range of
// it and up to W * 16 * QTM of its queries) from a shared counter; wave w owns the tile's query
// rows 16 QTM w .. 16 QTM (w + 1) - 1 and walks the same vector range as its sibling waves, so
// the W waves read every code chunk of the range at about the same time from the same CU: one
// HBM read serves the workgroup's W * 16 * QTM queries, the other waves' reads hitting the caches.
// Per chunk of 16*NVW vectors and per quantised segment, the B operands of 1-4 64-dim chunks for
// all NVW mblocks are loaded together, then the MFMAs run; each segment's int32 sums fold into the
// fp32 distances with three fused multiply-adds per (query, vector); after the segments the
// chunk's distances go through the wave's LDS to the per-query top-k, which a query only touches
// when one of the chunk's candidates beats its k-th pair.
struct BScanArgs {
    BatchedView bv;
    const int8_t *qq1, *qq2;
    const float *pc;
    const float *pq2;
    const int32_t *tile_clu, *tile_first, *tile_cnt, *tile_v0, *tile_cand;
    const int32_t *ntiles, *sv;
    int32_t *work;  // tile counter, zero at launch
    const int32_t *pairs;  // grouped pair -> q * NP + p
    int NP;
    uint32_t *qbound;      // per query: an upper bound of its k-th distance (float_key), shared by its tiles
    int topk;
    float *cand_d;  // (nq * NP) x topk
    uint32_t *cand_id;
};

// A query's distance bound: the k-th distance of any one of its lists' full top-k lists is at
// least its final k-th distance, so a candidate above it can be dropped by every other tile of the
// query.  Kept as float_key (order-preserving) for atomicMin.
__device__ __forceinline__ float key_float(uint32_t k) {
    if (k >= 0xFF800000u) return INFINITY;
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

struct WSmem {
    float *dt;       // [QT][CVW]
    uint32_t *dids;  // [CVW]
    float *ka;       // [nqs][QT] a1 of each row
    float *kb;       // [nqs][QT] a2
    float *kc;       // [nqs][QT] constant term
    float *q2;       // [QT]
    float *kd;       // [QT] k-th distance (the query's bound until the row's list is full)
    uint32_t *ki;    // [QT] k-th id
    float *rb;       // [QT] the query's bound at the tile's start
    int32_t *tkn;    // [QT]
    float *tkd;      // [QT][topk]
    uint32_t *tki;   // [QT][topk]
    float *fac;      // [2][NVW][16 * (2 + nqs)]: factor blocks of the current and the next chunk
    float *cbd;      // [QT][kCap]: candidates of a row not yet merged into its list (k <= 64)
    uint32_t *cbi;   // [QT][kCap]
    int32_t *cbn;    // [QT]
};

// mblocks per wave chunk.  A wave re-reads its query operands (16 rows x sum D_s int8 per row
// tile) from global memory for every chunk, so the longer the chunk the fewer of those reads:
// int8 waves of one row tile take 64 vectors (at 2 waves per SIMD, 32-query tiles 10.4 ms against
// 11.9 ms for the best 32-vector form on 10M x 1024; now 139 VGPRs, 3 waves), two row tiles 32;
// int8x2 (a second term) and BF16 (query operands twice as wide) take 32, or 16 at two row tiles
__host__ __device__ constexpr int nvw_of(int qtm, int qs) {
    return qs >= 2 ? (qtm >= 2 ? 1 : 2) : (qtm == 1 ? 4 : (qtm <= 2 ? 2 : 1));
}

// candidate buffer places per row: 28 for 16-query tiles, fewer for larger tiles (LDS: a wave of
// one int8 row tile at k = 10 and 4 quantised segments then takes 13.2 KB, and 12 waves fit a
// CU's 160 KB)
__host__ __device__ constexpr int cand_cap(int qtm) { return 28 / qtm; }

// LDS of one wave (QTM query tiles of 16 rows): the areas' places from base into s, and the bytes.
// s is a reference, never tested against null: a null test of a stack object's address keeps
// the compiler from holding its fields in registers.
template <int QTM, int QS>
__host__ __device__ inline size_t wave_layout(int nqs, int topk, uint8_t *base, WSmem &s) {
    constexpr int QT = 16 * QTM, CVW = 16 * nvw_of(QTM, QS);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        uint8_t *p = base ? base + off : nullptr;
        off += (bytes + 15) / 16 * 16;
        return p;
    };
    s.dt = (float *)take((size_t)QT * CVW * 4);
    s.dids = (uint32_t *)take(CVW * 4);
    s.ka = (float *)take((size_t)QT * nqs * 4);
    s.kb = (float *)take(QS == 2 ? (size_t)QT * nqs * 4 : 0);
    s.kc = (float *)take((size_t)QT * nqs * 4);
    s.q2 = (float *)take(QT * 4);
    s.kd = (float *)take(QT * 4);
    s.ki = (uint32_t *)take(QT * 4);
    s.rb = (float *)take(QT * 4);
    s.tkn = (int32_t *)take(QT * 4);
    s.tkd = (float *)take((size_t)QT * topk * 4);
    s.tki = (uint32_t *)take((size_t)QT * topk * 4);
    s.fac = (float *)take((size_t)2 * nvw_of(QTM, QS) * 64 * (2 + nqs));
    const int cap = topk <= 128 ? cand_cap(QTM) : 0;
    s.cbd = (float *)take((size_t)QT * cap * 4);
    s.cbi = (uint32_t *)take((size_t)QT * cap * 4);
    s.cbn = (int32_t *)take((size_t)QT * 4);
    return off;
}

template <int QTM, int QS>
__host__ __device__ inline size_t wave_smem(int nqs, int topk) {
    WSmem unused;
    return wave_layout<QTM, QS>(nqs, topk, nullptr, unused);
}

// LDS of a workgroup of W waves: W waves' areas, the tile index it broadcasts and the quantised
// segments' descriptors (read once per chunk and segment: LDS instead of global memory).
template <int W, int QTM, int QS>
__host__ __device__ inline size_t wg_smem(int nqs, int topk) {
    return W * wave_smem<QTM, QS>(nqs, topk) + 16 + (size_t)nqs * sizeof(BSeg);
}

// LDS traffic between the lanes of one wave: keep the compiler from moving accesses across
// this point (a wave's LDS operations complete in order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Entry e of a list held as KR registers per lane (entry 64 r + lane in register r), read in
// every lane; e is wave-uniform.
template <int KR>
__device__ __forceinline__ void list_entry(const float (&d)[KR], const uint32_t (&id)[KR], int e, float &x,
                                           uint32_t &xi) {
    const int r = KR == 1 ? 0 : e >> 6;
    x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r ? d[KR - 1] : d[0]), e & 63));
    xi = (uint32_t)__builtin_amdgcn_readlane((int)(r ? id[KR - 1] : id[0]), e & 63);
}

// Row i's sorted list (topk <= 64 KR places) from LDS into registers; places past topk hold
// (INF, 0xFFFFFFFF).
template <int KR>
__device__ __forceinline__ void load_row(const WSmem &m, int i, int topk, float (&d)[KR], uint32_t (&id)[KR]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int r = 0; r < KR; ++r) {
        const int e = 64 * r + lane;
        d[r] = e < topk ? m.tkd[i * topk + e] : INFINITY;
        id[r] = e < topk ? m.tki[i * topk + e] : 0xFFFFFFFFu;
    }
}

// The row's sorted list back to LDS with its k-th pair; while the list is not full (its k-th
// place still +inf) the threshold stays the query's bound.
template <int KR>
__device__ __forceinline__ void store_row(const WSmem &m, int i, int topk, const float (&d)[KR],
                                          const uint32_t (&id)[KR]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int r = 0; r < KR; ++r) {
        const int e = 64 * r + lane;
        if (e < topk) {
            m.tkd[i * topk + e] = d[r];
            m.tki[i * topk + e] = id[r];
        }
    }
    float kd;
    uint32_t ki;
    list_entry<KR>(d, id, topk - 1, kd, ki);
    if (kd == INFINITY) {
        kd = m.rb[i];
        ki = 0xFFFFFFFFu;
    }
    if (lane == 0) {
        m.kd[i] = kd;
        m.ki[i] = ki;
    }
    wave_sync();
}

// Merges row i's nb buffered candidates into its list and empties the buffer: up to
// kSeqInsert of them one by one (a position ballot and a one-place shift each), more by the
// bitonic merge.
constexpr int kSeqInsert = 8;

template <int KR>
__device__ __forceinline__ void flush_row(const WSmem &m, int i, int topk, int nb, int cap) {
    const int lane = threadIdx.x & 63;
    float d[KR];
    uint32_t id[KR];
    load_row<KR>(m, i, topk, d, id);
    if (nb <= kSeqInsert) {
        float kd = m.kd[i];
        uint32_t ki = m.ki[i];
        for (int e = 0; e < nb; ++e) {
            const float x = m.cbd[i * cap + e];
            const uint32_t xi = m.cbi[i * cap + e];
            if (!dless(x, xi, kd, ki)) continue;
            int pos = 0;
#pragma unroll
            for (int r = 0; r < KR; ++r) pos += __popcll(__ballot(dless(d[r], id[r], x, xi)));
            // entry 64 r + lane takes entry 64 r + lane - 1 above pos
            float carry = 0.0f;
            uint32_t carryi = 0u;
#pragma unroll
            for (int r = 0; r < KR; ++r) {
                float ud = __shfl_up(d[r], 1);
                uint32_t ui = (uint32_t)__shfl_up((int)id[r], 1);
                const float top = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(d[r]), 63));
                const uint32_t topi = (uint32_t)__builtin_amdgcn_readlane((int)id[r], 63);
                if (lane == 0) {
                    ud = carry;
                    ui = carryi;
                }
                carry = top;
                carryi = topi;
                const int ent = 64 * r + lane;
                if (ent > pos) {
                    d[r] = ud;
                    id[r] = ui;
                } else if (ent == pos) {
                    d[r] = x;
                    id[r] = xi;
                }
            }
            list_entry<KR>(d, id, topk - 1, kd, ki);
        }
    } else {
        const float cd = lane < nb ? m.cbd[i * cap + lane] : INFINITY;
        const uint32_t ci = lane < nb ? m.cbi[i * cap + lane] : 0xFFFFFFFFu;
        merge_into<KR>(d, id, cd, ci);
    }
    store_row<KR>(m, i, topk, d, id);
    if (lane == 0) m.cbn[i] = 0;
    wave_sync();
}

__device__ __forceinline__ v4i unpack_nibbles(uint2 x) {
    v4i r;
    r.x = (int)(x.x & 0x0F0F0F0Fu);
    r.y = (int)((x.x >> 4) & 0x0F0F0F0Fu);
    r.z = (int)(x.y & 0x0F0F0F0Fu);
    r.w = (int)((x.y >> 4) & 0x0F0F0F0Fu);
    return r;
}

__device__ __forceinline__ v4i mfma(v4i a, v4i b, v4i c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// Raw buffer loads: the base and size are uniform (scalar registers) and a lane's offset is one
// 32-bit register, so a load costs no vector address arithmetic.  Loads past the size read 0.
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc make_rsrc(const void *base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, bytes, 0x00020000);
}
__device__ __forceinline__ v4i ld16(Rsrc r, int voff, int soff) {
    return __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
__device__ __forceinline__ uint2 ld8(Rsrc r, int voff, int soff) {
    return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// Operands of a chunk's MFMAs: rb covers the chunk's NVW mblock records, mblock j at byte
// j * mb_stride; b16[j] / b8[j] are this lane's offsets of mblock j's 16-byte / 8-byte code
// operands.  The mblock offset is in the per-lane offset, which the buffer's range check covers, so
// mblocks past the list's end read 0.  ra1 / ra2: the wave's quantised query rows, aoff[t] this
// lane's row offset in them.  Segment and chunk offsets go in the uniform offset.
template <int QTM, int NVW>
struct MmaOps {
    Rsrc rb, ra1, ra2;
    int b16[NVW], b8[NVW];
    int aoff[QTM];
};

// 64-dim chunks per load step: 4 for nibble codes (loads half as wide), 2 for one int8 plane, 1
// for two planes (the step's code operands then take 32 registers at 4 mblocks, as the others do)
__host__ __device__ constexpr int chunks_per_step(int kind) { return kind == 0 ? 4 : (kind == 1 ? 2 : 1); }

// One load-then-multiply step of a segment: the code operands of kKB 64-dim chunks for all NVW
// mblocks and the matching query operands, all in flight before the first MFMA waits.  A chunk
// past nk loads the last one again and skips its MFMAs (a uniform branch).  FIRST: the step's
// first MFMAs start from zero instead of the accumulators.  KIND: 0 = 4-bit codes stored as nibbles, 1 = one int8 plane, 2 = a
// shifted low plane and a high plane (9+ bits).
template <int KIND, int QTM, int QS, int NVW, bool FIRST>
__device__ __forceinline__ void seg_step(const BSeg &bs, int kc0, int nk, const MmaOps<QTM, NVW> &o,
                                         v4i (&acc1)[QTM][NVW], v4i (&acc2)[QTM][NVW], v4i (&ach1)[QTM][NVW],
                                         v4i (&ach2)[QTM][NVW], int lane) {
    constexpr int kKB = chunks_per_step(KIND);
    v4i A1[kKB][QTM], A2[kKB][QTM], B[kKB][NVW], BH[kKB][NVW];
    uint2 BN[kKB][NVW];
#pragma unroll
    for (int u = 0; u < kKB; ++u) {
        const int kc = min(kc0 + u, nk - 1);
        const int sa = bs.kq_off + kc * 64;
#pragma unroll
        for (int t = 0; t < QTM; ++t) {
            A1[u][t] = ld16(o.ra1, o.aoff[t], sa);
            if (QS == 2) A2[u][t] = ld16(o.ra2, o.aoff[t], sa);
        }
        const int sb = bs.code_off + kc * bs.chunk_bytes;
#pragma unroll
        for (int j = 0; j < NVW; ++j) {
            if (KIND == 0) {
                BN[u][j] = ld8(o.rb, o.b8[j], sb);
            } else {
                B[u][j] = ld16(o.rb, o.b16[j], sb);
                if (KIND == 2) BH[u][j] = ld16(o.rb, o.b16[j] + 1024, sb);
            }
        }
    }
    // keep the scheduler from interleaving the step's loads with its MFMAs: all of them are in
    // flight before the first MFMA waits
    __builtin_amdgcn_sched_barrier(0);
    const v4i z = {0, 0, 0, 0};
#pragma unroll
    for (int u = 0; u < kKB; ++u) {
        if (u > 0 && kc0 + u >= nk) break;
        const bool zc = FIRST && u == 0;
#pragma unroll
        for (int t = 0; t < QTM; ++t)
#pragma unroll
            for (int j = 0; j < NVW; ++j) {
                const v4i b = KIND == 0 ? unpack_nibbles(BN[u][j]) : B[u][j];
                acc1[t][j] = mfma(A1[u][t], b, zc ? z : acc1[t][j]);
                if (QS == 2) acc2[t][j] = mfma(A2[u][t], b, zc ? z : acc2[t][j]);
                if (KIND == 2) {
                    ach1[t][j] = mfma(A1[u][t], BH[u][j], zc ? z : ach1[t][j]);
                    if (QS == 2) ach2[t][j] = mfma(A2[u][t], BH[u][j], zc ? z : ach2[t][j]);
                }
            }
    }
}

// One quantised segment's contraction for a chunk of NVW mblocks, step by step; the accumulators
// of the segment's plane(s) are set by its first step.
template <int KIND, int QTM, int QS, int NVW>
__device__ __forceinline__ void seg_mma(const BSeg &bs, const MmaOps<QTM, NVW> &o, v4i (&acc1)[QTM][NVW],
                                        v4i (&acc2)[QTM][NVW], v4i (&ach1)[QTM][NVW], v4i (&ach2)[QTM][NVW],
                                        int lane) {
    constexpr int kKB = chunks_per_step(KIND);
    const int nk = bs.dim / 64;
    seg_step<KIND, QTM, QS, NVW, true>(bs, 0, nk, o, acc1, acc2, ach1, ach2, lane);
    for (int kc0 = kKB; kc0 < nk; kc0 += kKB) seg_step<KIND, QTM, QS, NVW, false>(bs, kc0, nk, o, acc1, acc2, ach1, ach2, lane);
}

__device__ __forceinline__ v4f mfma_bf16(v8bf a, v8bf b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// A lane's 16 code bytes (dims 0..15 of its slice, 4 per dword) as two bf16 operands (dims 0..7
// and 8..15), exactly: every code is an integer below 256.  flip: 0x80808080 undoes the -128 shift
// of 8-bit codes and of the low plane of 9+-bit codes, 0 otherwise.
__device__ __forceinline__ void codes_bf16(v4i c, uint32_t flip, v8bf &lo, v8bf &hi) {
    uint32_t o[8];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const uint32_t x = (uint32_t)c[d] ^ flip;
        const float f0 = (float)(x & 0xFFu), f1 = (float)((x >> 8) & 0xFFu);
        const float f2 = (float)((x >> 16) & 0xFFu), f3 = (float)(x >> 24);
        // the high halves of two floats side by side: their bf16 (exact here)
        o[2 * d] = __builtin_amdgcn_perm(__float_as_uint(f1), __float_as_uint(f0), 0x07060302u);
        o[2 * d + 1] = __builtin_amdgcn_perm(__float_as_uint(f3), __float_as_uint(f2), 0x07060302u);
    }
    const uint4 l = make_uint4(o[0], o[1], o[2], o[3]), h = make_uint4(o[4], o[5], o[6], o[7]);
    lo = __builtin_bit_cast(v8bf, l);
    hi = __builtin_bit_cast(v8bf, h);
}

// seg_mma for BF16 queries (SAQGPU_PREC_BF16): the same code loads, each lane's 16 codes of a
// 64-dim chunk split into two v_mfma_f32_16x16x32_bf16 (dims 0..7 and 8..15 of the lane's slice,
// paired with the same dims of the query row, so the contraction runs over the same dimensions),
// fp32 accumulators; 9+-bit codes as a low and a high plane.
template <int KIND, int QTM, int NVW>
__device__ __forceinline__ void seg_mma_bf16(const BSeg &bs, const uint8_t *const (&rec)[NVW],
                                             const uint16_t *const (&arow)[QTM], v4f (&acc)[QTM][NVW],
                                             v4f (&ach)[QTM][NVW], int lane) {
    constexpr int kKB = KIND == 0 ? 4 : 2;
    const int nk = bs.dim / 64;
    const uint32_t flip = (KIND == 2 || bs.bits == 8) ? 0x80808080u : 0u;
    for (int kc0 = 0; kc0 < nk; kc0 += kKB) {
        v8bf Al[kKB][QTM], Ah[kKB][QTM];
        v4i B[kKB][NVW], BH[kKB][NVW];
        uint2 BN[kKB][NVW];
#pragma unroll
        for (int u = 0; u < kKB; ++u) {
            const int kc = min(kc0 + u, nk - 1);
            const bool live = kc0 + u < nk;
#pragma unroll
            for (int t = 0; t < QTM; ++t) {
                const v8bf *ap = reinterpret_cast<const v8bf *>(arow[t] + bs.kq_off + kc * 64);
                const v8bf zero = __builtin_bit_cast(v8bf, make_uint4(0u, 0u, 0u, 0u));
                Al[u][t] = live ? ap[0] : zero;
                Ah[u][t] = live ? ap[1] : zero;
            }
#pragma unroll
            for (int j = 0; j < NVW; ++j) {
                const uint8_t *p = rec[j] + bs.code_off + (int64_t)kc * bs.chunk_bytes;
                if (KIND == 0) {
                    BN[u][j] = *reinterpret_cast<const uint2 *>(p + lane * 8);
                } else {
                    B[u][j] = *reinterpret_cast<const v4i *>(p + lane * 16);
                    if (KIND == 2) BH[u][j] = *reinterpret_cast<const v4i *>(p + 1024 + lane * 16);
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kKB; ++u)
#pragma unroll
            for (int j = 0; j < NVW; ++j) {
                v8bf bl, bh;
                codes_bf16(KIND == 0 ? unpack_nibbles(BN[u][j]) : B[u][j], flip, bl, bh);
                v8bf hl, hh;
                if (KIND == 2) codes_bf16(BH[u][j], 0u, hl, hh);
#pragma unroll
                for (int t = 0; t < QTM; ++t) {
                    acc[t][j] = mfma_bf16(Al[u][t], bl, acc[t][j]);
                    acc[t][j] = mfma_bf16(Ah[u][t], bh, acc[t][j]);
                    if (KIND == 2) {
                        ach[t][j] = mfma_bf16(Al[u][t], hl, ach[t][j]);
                        ach[t][j] = mfma_bf16(Ah[u][t], hh, ach[t][j]);
                    }
                }
            }
    }
}

template <int W, int QTM, int QS, int KR>
__global__ __launch_bounds__(64 * W, (W * QTM >= 8 || (QTM == 4 && QS == 2)) ? 1 : (QS == 1 && QTM == 1 && W >= 2 ? 3 : 2))
void bscan_kernel(BScanArgs a) {
    constexpr int QT = 16 * QTM;          // query rows of a wave
    constexpr int NVW = nvw_of(QTM, QS);  // mblocks per chunk
    constexpr int CVW = 16 * NVW;         // vectors per chunk
    extern __shared__ __align__(16) uint8_t smem_raw[];
    const BatchedView &bv = a.bv;
    // the wave index through readfirstlane: the compiler then keeps the LDS layout's pointers in
    // scalar registers
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const size_t per_wave = wave_smem<QTM, QS>(bv.nqs, a.topk);
    WSmem m;
    wave_layout<QTM, QS>(bv.nqs, a.topk, smem_raw + w * per_wave, m);
    int32_t *tile_s = reinterpret_cast<int32_t *>(smem_raw + W * per_wave);
    BSeg *qseg = reinterpret_cast<BSeg *>(smem_raw + W * per_wave + 16);  // quantised segments by qs
    for (int s = threadIdx.x; s < bv.nseg; s += 64 * W) {
        const BSeg b = bv.segs[s];
        if (b.qs >= 0) qseg[b.qs] = b;
    }
    __syncthreads();
    const int ntiles = *a.ntiles, sv = *a.sv;
    for (;;) {
        int tile;
        // every tile comes from the launch's counter
        if (W == 1) {
            if (lane == 0) tile = atomicAdd(a.work, 1);
            tile = __shfl(tile, 0);
        } else {
            if (threadIdx.x == 0) *tile_s = atomicAdd(a.work, 1);
            __syncthreads();
            tile = *tile_s;
            __syncthreads();  // every wave has read it before the next round writes it
        }
        if (tile >= ntiles) break;
        const int c = a.tile_clu[tile];
        const int first = a.tile_first[tile] + w * QT, cnt = min(QT, a.tile_cnt[tile] - w * QT);
        if (cnt <= 0) continue;  // the tile's queries end before this wave's rows
        // ---- the wave's pair constants to LDS; rows past cnt repeat the last pair
        for (int e = lane; e < QT * bv.nqs; e += 64) {
            const int row = e % QT, s = e / QT;
            const float *pc = a.pc + ((int64_t)(first + min(row, cnt - 1)) * bv.nqs + s) * 3;
            m.ka[e] = pc[0];
            if (QS == 2) m.kb[e] = pc[1];
            m.kc[e] = pc[2];
        }
        for (int row = lane; row < QT; row += 64) {
            const int g = first + min(row, cnt - 1);
            m.q2[row] = a.pq2[g];
            // top-k lists (k <= 128) start from the query's bound; longer ones from +inf
            const float b = KR <= 2 ? key_float(a.qbound[a.pairs[g] / a.NP]) : INFINITY;
            m.kd[row] = b;
            m.rb[row] = b;
            m.ki[row] = 0xFFFFFFFFu;
            m.tkn[row] = 0;
            m.cbn[row] = 0;
        }
        if (KR <= 2)
            for (int e = lane; e < QT * a.topk; e += 64) {
                m.tkd[e] = INFINITY;
                m.tki[e] = 0xFFFFFFFFu;
            }
        wave_sync();
        // A operand rows of this lane: query row t*16 + lane%16 of the wave's rows
        const uint16_t *arowb[QTM];  // BF16 rows (QS == 3)
#pragma unroll
        for (int t = 0; t < QTM; ++t) {
            const int64_t g = first + min(t * 16 + lane % 16, cnt - 1);
            arowb[t] = reinterpret_cast<const uint16_t *>(a.qq1) + g * bv.kq + 16 * (lane / 16);
        }
        // int8 operands through buffer loads; the query rows of this wave: first .. first + QT - 1
        MmaOps<QTM, NVW> ops;
        ops.ra1 = make_rsrc(a.qq1 + (int64_t)first * bv.kq, QT * bv.kq);
        ops.ra2 = make_rsrc(QS == 2 ? a.qq2 + (int64_t)first * bv.kq : a.qq1, QT * bv.kq);
#pragma unroll
        for (int t = 0; t < QTM; ++t) ops.aoff[t] = min(t * 16 + lane % 16, cnt - 1) * bv.kq + 16 * (lane / 16);
#pragma unroll
        for (int j = 0; j < NVW; ++j) {
            ops.b16[j] = (int)(j * bv.mb_stride) + lane * 16;
            ops.b8[j] = (int)(j * bv.mb_stride) + lane * 8;
        }
        // the tile's vector range [v0, v0 + sv) of the list, in mblocks (v0 and sv are multiples of 64)
        const int n = bv.clu_n[c];
        const int v0 = a.tile_v0[tile];
        const int mb_beg = v0 / kMB, mb_end = min((n + kMB - 1) / kMB, (v0 + sv) / kMB);
        const int64_t mb0 = bv.clu_mb[c];
        // factor blocks (|o|^2, -2 x the rescale of every quantised segment, the ids) of a chunk's
        // mblocks move to LDS by DMA one chunk ahead of their use
        const int fac_words = kMB * (2 + bv.nqs);
        auto fetch_fac = [&](int cbn, int slot) {
#pragma unroll
            for (int j = 0; j < NVW; ++j) {
                const uint8_t *src = bv.mb + (mb0 + min(cbn + j, mb_end - 1)) * bv.mb_stride + bv.fac_off;
                float *dst = m.fac + (slot * NVW + j) * fac_words;
                for (int off = 0; off < fac_words * 4; off += 1024)
                    if (off + lane * 16 < fac_words * 4)
                        __builtin_amdgcn_global_load_lds(src + off + lane * 16, reinterpret_cast<uint8_t *>(dst) + off,
                                                         16, 0, 0);
            }
        };
        int slot = 0;
        if (mb_beg < mb_end) fetch_fac(mb_beg, 0);
        for (int cb = mb_beg; cb < mb_end; cb += NVW, slot ^= 1) {
            const uint8_t *rec[NVW];
            float dist[QTM][NVW][4];
            dma_wait();  // this chunk's factor blocks (the previous chunk's reads of the other slot are done)
            const float *fac = m.fac + slot * NVW * fac_words;
#pragma unroll
            for (int t = 0; t < QTM; ++t) {
                const float4 q2 = *reinterpret_cast<const float4 *>(m.q2 + t * 16 + 4 * (lane / 16));
#pragma unroll
                for (int j = 0; j < NVW; ++j) {
                    const float o2 = fac[j * fac_words + lane % kMB];
                    dist[t][j][0] = o2 + q2.x;
                    dist[t][j][1] = o2 + q2.y;
                    dist[t][j][2] = o2 + q2.z;
                    dist[t][j][3] = o2 + q2.w;
                }
            }
#pragma unroll
            for (int j = 0; j < NVW; ++j)
                rec[j] = bv.mb + (mb0 + min(cb + j, mb_end - 1)) * bv.mb_stride;  // clamped duplicates are discarded
            ops.rb = make_rsrc(bv.mb + (mb0 + cb) * bv.mb_stride, (int)(min(NVW, mb_end - cb) * bv.mb_stride));
            if (cb + NVW < mb_end) fetch_fac(cb + NVW, slot ^ 1);
            for (int k = 0; k < bv.nqs; ++k) {
                const BSeg bs = qseg[k];
                // int8 accumulators are set by their segment's first MFMAs
                v4i acc1[QTM][NVW], acc2[QTM][NVW], ach1[QTM][NVW], ach2[QTM][NVW];
                v4f facc[QTM][NVW], fach[QTM][NVW];
                if (QS == 3)
#pragma unroll
                    for (int t = 0; t < QTM; ++t)
#pragma unroll
                        for (int j = 0; j < NVW; ++j) facc[t][j] = fach[t][j] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
                // epilogue of the segment: ip = a1 acc1 (+ 256 a1 acch1) (+ a2 acc2 ...) + c, then
                // dist += (-2 rescale) ip, all fused multiply-adds.  TWO (a high plane) is a
                // template constant: the one-plane form carries none of its instructions.
                auto epilogue = [&](auto two_c) {
                    constexpr bool TWO = decltype(two_c)::value;
#pragma unroll
                    for (int t = 0; t < QTM; ++t) {
                        const int rb = bs.qs * QT + t * 16 + 4 * (lane / 16);
                        const float4 ka = *reinterpret_cast<const float4 *>(m.ka + rb);
                        const float4 kc = *reinterpret_cast<const float4 *>(m.kc + rb);
                        float4 kb = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        if (QS == 2) kb = *reinterpret_cast<const float4 *>(m.kb + rb);
                        const float av[4] = {ka.x, ka.y, ka.z, ka.w}, bv2[4] = {kb.x, kb.y, kb.z, kb.w};
                        const float cv[4] = {kc.x, kc.y, kc.z, kc.w};
#pragma unroll
                        for (int j = 0; j < NVW; ++j) {
                            const float m2r = fac[j * fac_words + kMB * (1 + bs.qs) + lane % kMB];
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const float x1 = QS == 3 ? facc[t][j][r] : (float)acc1[t][j][r];
                                float ip = fmaf(av[r], x1, cv[r]);
                                if (TWO) ip = fmaf(256.0f * av[r], QS == 3 ? fach[t][j][r] : (float)ach1[t][j][r], ip);
                                if (QS == 2) {
                                    ip = fmaf(bv2[r], (float)acc2[t][j][r], ip);
                                    if (TWO) ip = fmaf(256.0f * bv2[r], (float)ach2[t][j][r], ip);
                                }
                                dist[t][j][r] = fmaf(m2r, ip, dist[t][j][r]);
                            }
                        }
                    }
                };
                using One = std::integral_constant<bool, false>;
                using Two = std::integral_constant<bool, true>;
                if (QS == 3) {
                    if (bs.bits <= 4) {
                        seg_mma_bf16<0, QTM, NVW>(bs, rec, arowb, facc, fach, lane);
                        epilogue(One{});
                    } else if (bs.bits <= 8) {
                        seg_mma_bf16<1, QTM, NVW>(bs, rec, arowb, facc, fach, lane);
                        epilogue(One{});
                    } else {
                        seg_mma_bf16<2, QTM, NVW>(bs, rec, arowb, facc, fach, lane);
                        epilogue(Two{});
                    }
                } else if (bs.bits <= 4) {
                    seg_mma<0, QTM, QS, NVW>(bs, ops, acc1, acc2, ach1, ach2, lane);
                    epilogue(One{});
                } else if (bs.bits <= 8) {
                    seg_mma<1, QTM, QS, NVW>(bs, ops, acc1, acc2, ach1, ach2, lane);
                    epilogue(One{});
                } else {
                    seg_mma<2, QTM, QS, NVW>(bs, ops, acc1, acc2, ach1, ach2, lane);
                    epilogue(Two{});
                }
            }
            // ---- rows (queries) with a distance of this chunk at or below their k-th distance: the
            // k-th distance only falls, so no other row can take a candidate of this chunk
            unsigned long long rows = 0;
#pragma unroll
            for (int t = 0; t < QTM; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float kdr = m.kd[t * 16 + 4 * (lane / 16) + r];
                    bool hit = false;
#pragma unroll
                    for (int j = 0; j < NVW; ++j) hit |= cb + j < mb_end && dist[t][j][r] <= kdr;
                    const unsigned long long b = __ballot(hit);
#pragma unroll
                    for (int g = 0; g < 4; ++g)
                        if ((b >> (16 * g)) & 0xFFFFull) rows |= 1ull << (t * 16 + 4 * g + r);
                }
            if (cnt < 64) rows &= (1ull << cnt) - 1;
            if (!rows) continue;
            // ---- the chunk's distances to LDS, then the per-query lists
#pragma unroll
            for (int j = 0; j < NVW; ++j) {
                const int lv = j * kMB + lane % kMB;
                const bool valid = cb + j < mb_end;
#pragma unroll
                for (int t = 0; t < QTM; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        m.dt[(t * 16 + 4 * (lane / 16) + r) * CVW + lv] = valid ? dist[t][j][r] : INFINITY;
                if (lane < kMB)
                    m.dids[lv] = valid ? __float_as_uint(fac[j * fac_words + kMB * (1 + bv.nqs) + lane]) : 0xFFFFFFFFu;
            }
            wave_sync();
            const float x_id_pad = INFINITY;
            while (KR <= 2 && rows) {
                // buffered: candidates that beat the row's k-th pair are appended to its buffer;
                // a full buffer is merged into the list (which lowers the k-th pair)
                const int i = __builtin_ctzll(rows);
                rows &= rows - 1;
                const float x = lane < CVW ? m.dt[i * CVW + lane] : x_id_pad;
                const uint32_t xi = lane < CVW ? m.dids[lane] : 0xFFFFFFFFu;
                bool pass = dless(x, xi, m.kd[i], m.ki[i]);
                unsigned long long b = __ballot(pass);
                if (!b) continue;
                constexpr int kCap = cand_cap(QTM);
                int nb = m.cbn[i];
                if (nb + __popcll(b) > kCap) {
                    if (nb) flush_row<KR>(m, i, a.topk, nb, kCap);
                    nb = 0;
                    pass = dless(x, xi, m.kd[i], m.ki[i]);
                    b = __ballot(pass);
                    if (__popcll(b) > kCap) {  // more than a buffer holds: merge the chunk's row directly
                        float d[KR];
                        uint32_t id[KR];
                        load_row<KR>(m, i, a.topk, d, id);
                        merge_into<KR>(d, id, pass ? x : INFINITY, pass ? xi : 0xFFFFFFFFu);
                        store_row<KR>(m, i, a.topk, d, id);
                        b = 0;
                    }
                }
                if (pass && b) {
                    const int pos = nb + __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0));
                    m.cbd[i * kCap + pos] = x;
                    m.cbi[i * kCap + pos] = xi;
                }
                if (lane == 0) m.cbn[i] = nb + __popcll(b);
                wave_sync();
            }
            while (KR > 2 && rows) {
                const int i = __builtin_ctzll(rows);
                rows &= rows - 1;
                const float x = lane < CVW ? m.dt[i * CVW + lane] : x_id_pad;
                const uint32_t xi = lane < CVW ? m.dids[lane] : 0xFFFFFFFFu;
                if (!__ballot(dless(x, xi, m.kd[i], m.ki[i]))) continue;
                WaveTopK<KR> tk;
                tk.load(m.tkd + i * a.topk, m.tki + i * a.topk, m.tkn[i], a.topk);
                tk.offer(x, xi);
                wave_sync();
                tk.store(m.tkd + i * a.topk, m.tki + i * a.topk);
                if (lane == 0) {
                    m.tkn[i] = tk.n;
                    m.kd[i] = tk.kd;
                    m.ki[i] = tk.kid;
                }
                wave_sync();
            }
            wave_sync();  // the next chunk rewrites dt and dids
        }
        // ---- the wave's per-query lists to global memory
        if (KR <= 2)
            for (int i = 0; i < cnt; ++i) {
                const int nb = m.cbn[i];
                if (nb) flush_row<KR>(m, i, a.topk, nb, cand_cap(QTM));
            }
        for (int i = 0; i < cnt; ++i) {
            const int64_t pr = (int64_t)a.tile_cand[tile] + w * QT + i;
            if (KR <= 2) {  // padded lists: unused places already hold (INF, 0xFFFFFFFF)
                for (int e = lane; e < a.topk; e += 64) {
                    a.cand_d[pr * a.topk + e] = m.tkd[i * a.topk + e];
                    a.cand_id[pr * a.topk + e] = m.tki[i * a.topk + e];
                }
                // a full list's k-th distance bounds the query's for the tiles still to come
                const float kth = m.tkd[i * a.topk + a.topk - 1];
                if (lane == 0 && kth < m.rb[i]) atomicMin(&a.qbound[a.pairs[first + i] / a.NP], float_key(kth));
                continue;
            }
            const int nn = m.tkn[i];
            for (int e = lane; e < a.topk; e += 64) {
                a.cand_d[pr * a.topk + e] = e < nn ? m.tkd[i * a.topk + e] : INFINITY;
                a.cand_id[pr * a.topk + e] = e < nn ? m.tki[i * a.topk + e] : 0xFFFFFFFFu;
            }
        }
        wave_sync();
    }
}

// ------------------------------------------------------------------ merge
// One wave per query: the k smallest (distance, id) pairs over the candidate lists of its
// pairs (one list per vector range of the probed list).  The lanes take 64 probes at a time
// and load the heads of their r-th lists together; only lists whose head beats the current
// k-th pair are then walked (each list is sorted, so a list whose head does not beat it holds
// nothing that does).
template <int KR>
__global__ __launch_bounds__(256) void bq_merge_kernel(const float *cand_d, const uint32_t *cand_id,
                                                      const uint32_t *probes, const int32_t *slot_of,
                                                      const int32_t *counts, const int32_t *candstart,
                                                      const int32_t *nsplit, const int32_t *npq, int64_t nq,
                                                      int NP, int topk, uint32_t *out_ids, float *out_d) {
    const int lane = threadIdx.x & 63;
    const int64_t q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nq) return;
    WaveTopK<KR> tk;
    tk.k = topk;
    tk.n = 0;
#pragma unroll
    for (int j = 0; j < KR; ++j) {
        tk.d[j] = INFINITY;
        tk.id[j] = 0xFFFFFFFFu;
    }
    tk.refresh();
    const int lim = npq ? npq[q] : NP;
    if (KR == 1) {
        // k <= 64: the lists whose head beats the k-th pair, merged g = 64 / k at a time (lane l
        // takes element l % k of the (l / k)-th of them) by one bitonic merge each
        float d = INFINITY;
        uint32_t id = 0xFFFFFFFFu;
        float kd = INFINITY;
        uint32_t ki = 0xFFFFFFFFu;
        const int g = 64 / topk, slot = lane / topk, e = lane % topk;
        for (int p0 = 0; p0 < lim; p0 += 64) {
            const int p = p0 + lane;
            int ns = 0, cn = 0;
            int64_t base = 0;
            if (p < lim) {
                const int64_t i = q * NP + p;
                const int c = (int)probes[i];
                ns = nsplit[c];
                cn = counts[c];
                base = (int64_t)candstart[c] + slot_of[i];
            }
            int rmax = ns;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) rmax = max(rmax, __shfl_xor(rmax, o));
            for (int r = 0; r < rmax; ++r) {
                const int64_t li = base + (int64_t)r * cn;  // this lane's r-th list
                const bool has = r < ns;
                const float hd = has ? cand_d[li * topk] : INFINITY;
                const uint32_t hi = has ? cand_id[li * topk] : 0xFFFFFFFFu;
                unsigned long long m = __ballot(has && dless(hd, hi, kd, ki));
                while (m) {
                    // the slot-th set lane of m holds this lane's list
                    unsigned long long mm = m;
                    for (int j = 0; j < slot && mm; ++j) mm &= mm - 1;
                    const int src = mm && slot < g ? __builtin_ctzll(mm) : -1;
                    for (int j = 0; j < g && m; ++j) m &= m - 1;
                    const int64_t L = __shfl(li, src < 0 ? 0 : src);
                    const float x = src >= 0 ? cand_d[L * topk + e] : INFINITY;
                    const uint32_t xi = src >= 0 ? cand_id[L * topk + e] : 0xFFFFFFFFu;
                    float dd[1] = {d};
                    uint32_t ii[1] = {id};
                    merge_into<1>(dd, ii, x, xi);
                    d = dd[0];
                    id = ii[0];
                    kd = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(d), topk - 1));
                    ki = (uint32_t)__builtin_amdgcn_readlane((int)id, topk - 1);
                    m &= __ballot(has && dless(hd, hi, kd, ki));
                }
            }
        }
        if (lane < topk) {
            const bool ok = id != 0xFFFFFFFFu || d != INFINITY;
            out_ids[q * topk + lane] = ok ? id : 0u;
            if (out_d) out_d[q * topk + lane] = ok ? d : 0.0f;
        }
        return;
    }
    if (KR == 2) {
        // 64 < k <= 128: every list whose head beats the k-th pair is merged 64 entries at a time
        // (a chunk only while its head beats it) by the two-register bitonic merge
        float d[2] = {INFINITY, INFINITY};
        uint32_t id[2] = {0xFFFFFFFFu, 0xFFFFFFFFu};
        float kd = INFINITY;
        uint32_t ki = 0xFFFFFFFFu;
        for (int p0 = 0; p0 < lim; p0 += 64) {
            const int p = p0 + lane;
            int ns = 0, cn = 0;
            int64_t base = 0;
            if (p < lim) {
                const int64_t i = q * NP + p;
                const int c = (int)probes[i];
                ns = nsplit[c];
                cn = counts[c];
                base = (int64_t)candstart[c] + slot_of[i];
            }
            int rmax = ns;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) rmax = max(rmax, __shfl_xor(rmax, o));
            for (int r = 0; r < rmax; ++r) {
                const int64_t li = base + (int64_t)r * cn;  // this lane's r-th list
                const bool has = r < ns;
                const float hd = has ? cand_d[li * topk] : INFINITY;
                const uint32_t hi = has ? cand_id[li * topk] : 0xFFFFFFFFu;
                unsigned long long m = __ballot(has && dless(hd, hi, kd, ki));
                while (m) {
                    const int src = __builtin_ctzll(m);
                    m &= m - 1;
                    const int64_t L = __shfl(li, src);
                    for (int e0 = 0; e0 < topk; e0 += 64) {
                        const int e = e0 + lane;
                        const float x = e < topk ? cand_d[L * topk + e] : INFINITY;
                        const uint32_t xi = e < topk ? cand_id[L * topk + e] : 0xFFFFFFFFu;
                        const float x0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
                        const uint32_t xi0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)xi);
                        if (!dless(x0, xi0, kd, ki)) break;
                        merge_into<2, true>(d, id, x, xi);  // a list chunk is sorted
                        list_entry<2>(d, id, topk - 1, kd, ki);
                    }
                    m &= __ballot(has && dless(hd, hi, kd, ki));
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int e = j * 64 + lane;
            if (e < topk) {
                const bool ok = id[j] != 0xFFFFFFFFu || d[j] != INFINITY;
                out_ids[q * topk + e] = ok ? id[j] : 0u;
                if (out_d) out_d[q * topk + e] = ok ? d[j] : 0.0f;
            }
        }
        return;
    }
    for (int p0 = 0; p0 < lim; p0 += 64) {
        const int p = p0 + lane;
        int ns = 0, cn = 0;
        int64_t base = 0;
        if (p < lim) {
            const int64_t i = q * NP + p;
            const int c = (int)probes[i];
            ns = nsplit[c];
            cn = counts[c];
            base = (int64_t)candstart[c] + slot_of[i];
        }
        int rmax = ns;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) rmax = max(rmax, __shfl_xor(rmax, o));
        for (int r = 0; r < rmax; ++r) {
            const int64_t li = base + (int64_t)r * cn;  // this lane's r-th list
            const bool has = r < ns;
            const float hd = has ? cand_d[li * topk] : INFINITY;
            const uint32_t hi = has ? cand_id[li * topk] : 0xFFFFFFFFu;
            unsigned long long m = __ballot(has && dless(hd, hi, tk.kd, tk.kid));
            while (m) {
                const int src = __builtin_ctzll(m);
                m &= m - 1;
                const int64_t L = __shfl(li, src);
                const float *cd = cand_d + L * topk;
                const uint32_t *ci = cand_id + L * topk;
                for (int e0 = 0; e0 < topk; e0 += 64) {
                    const int e = e0 + lane;
                    const float x = e < topk ? cd[e] : INFINITY;
                    const uint32_t xi = e < topk ? ci[e] : 0xFFFFFFFFu;
                    if (!dless(__shfl(x, 0), __shfl(xi, 0), tk.kd, tk.kid)) break;
                    tk.offer(x, xi);
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < KR; ++j) {
        const int i = j * 64 + lane;
        if (i < topk) {
            out_ids[q * topk + i] = i < tk.n ? tk.id[j] : 0u;
            if (out_d) out_d[q * topk + i] = i < tk.n ? tk.d[j] : 0.0f;
        }
    }
}

// ------------------------------------------------------------------ host-side launchers
// The (waves per workgroup, query tiles of 16 per wave) of a query tile of qt queries.
static void tile_shape(int qt, int qs, int &W, int &qtm) {
    // (one row tile per wave up to 64 queries: with 64-vector chunks the int8 tiles of 16 / 32 / 64
    // queries ran 11.8 / 10.5 / 11.2 ms on 10M x 1024, 2 waves x 2 row tiles of 32-vector chunks
    // 11.9 ms; the wider query forms 21 ms as 4 x 1 against 25-31 ms; 4 row tiles per wave hold 2
    // waves per SIMD and ran 19-21 ms)
    switch (qt) {
    case 16: W = 1, qtm = 1; break;
    case 32: W = 2, qtm = 1; break;
    case 64: W = 4, qtm = 1; break;
    default: W = 4, qtm = 2; break;  // 128
    }
}

hipError_t launch_bq_build(const IndexView &ix, const BatchedView &bv, const int32_t *mb_clu, int64_t total_mb,
                           const int32_t *chunk_seg, const int32_t *chunk_kc, int nchunks, uint8_t *out,
                           hipStream_t st) {
    if (total_mb == 0) return hipSuccess;
    if (nchunks > 0) {
        hipLaunchKernelGGL(bq_build_codes_kernel, dim3((unsigned)total_mb, nchunks), dim3(64), 0, st, ix, bv, mb_clu,
                           chunk_seg, chunk_kc, out);
    }
    hipLaunchKernelGGL(bq_build_factors_kernel, dim3((unsigned)total_mb), dim3(64), 0, st, ix, bv, mb_clu, out);
    return hipGetLastError();
}

hipError_t launch_bq_group(const uint32_t *probes, const int32_t *npq, int64_t nq, int NP, int K, int QT, int qs,
                           const int32_t *order, const int32_t *clu_n, const BqWork &w, hipStream_t st) {
    int W, qtm;
    tile_shape(QT, qs, W, qtm);
    const int workers = kSplitWaves / W;  // resident workgroups of the tile kernel
    hipLaunchKernelGGL(bq_zero_kernel, dim3(16), dim3(256), 0, st, w.counts, K, w.work, w.qbound, nq);
    const int64_t n = nq * NP;
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(bq_count_kernel, dim3(g), dim3(256), 0, st, probes, npq, nq, NP, w.counts, w.slot);
    hipLaunchKernelGGL(bq_scan_kernel, dim3(1), dim3(1024), 0, st, w.counts, order, clu_n, K, QT, workers, w.cstart,
                       w.candstart, w.nsplit, w.ntiles, w.npairs, w.sv, w.tfirst);
    hipLaunchKernelGGL(bq_tiles_kernel, dim3(256), dim3(256), 0, st, order, w.tfirst, K, QT, w.counts, w.cstart,
                       w.candstart, w.nsplit, w.ntiles, w.sv, w.tile_clu, w.tile_first, w.tile_cnt, w.tile_v0,
                       w.tile_cand);
    hipLaunchKernelGGL(bq_scatter_kernel, dim3(g), dim3(256), 0, st, probes, w.slot, n, w.cstart, w.pairs);
    return hipGetLastError();
}

template <int W, int QTM, int QS, int KR>
static hipError_t launch_bscan_t(const BScanArgs &a, int grid, hipStream_t st) {
    const size_t smem = wg_smem<W, QTM, QS>(a.bv.nqs, a.topk);
    auto kern = bscan_kernel<W, QTM, QS, KR>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * W, smem);
    if (e != hipSuccess) return e;
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int g = std::max(1, std::min(grid, std::max(1, per_cu) * ncu));
    hipLaunchKernelGGL(kern, dim3(g), dim3(64 * W), smem, st, a);
    return hipGetLastError();
}

size_t bscan_smem_bytes(const BatchedView &bv, int topk, int qt, int qs) {
    int W, qtm;
    tile_shape(qt, qs, W, qtm);
    if (qs == 3) {
        if (qtm == 1) return (size_t)W * wave_smem<1, 3>(bv.nqs, topk) + 16 + (size_t)bv.nqs * sizeof(BSeg);
        return (size_t)W * wave_smem<2, 3>(bv.nqs, topk) + 16 + (size_t)bv.nqs * sizeof(BSeg);
    }
    if (qs == 2) {
        if (qtm == 1) return (size_t)W * wave_smem<1, 2>(bv.nqs, topk) + 16 + (size_t)bv.nqs * sizeof(BSeg);
        return (size_t)W * wave_smem<2, 2>(bv.nqs, topk) + 16 + (size_t)bv.nqs * sizeof(BSeg);
    }
    if (qtm == 1) return (size_t)W * wave_smem<1, 1>(bv.nqs, topk) + 16 + (size_t)bv.nqs * sizeof(BSeg);
    return (size_t)W * wave_smem<2, 1>(bv.nqs, topk) + 16 + (size_t)bv.nqs * sizeof(BSeg);
}

template <int W, int QTM, int QS>
static hipError_t launch_bscan_q(const BScanArgs &a, int grid, hipStream_t st) {
    if (a.topk <= 64) return launch_bscan_t<W, QTM, QS, 1>(a, grid, st);
    if (a.topk <= 128) return launch_bscan_t<W, QTM, QS, 2>(a, grid, st);
    return launch_bscan_t<W, QTM, QS, 4>(a, grid, st);
}

hipError_t launch_bq_pairprep(const BatchedView &bv, const float *qrot, const uint32_t *probes, const BqWork &w,
                              const BqPairBufs &pb, int NP, int64_t max_pairs, int qs, hipStream_t st) {
    PairPrepArgs a;
    a.bv = bv;
    a.qrot = qrot;
    a.probes = probes;
    a.pairs = w.pairs;
    a.npairs = w.npairs;
    a.NP = NP;
    a.qs = qs;
    a.qq1 = pb.qq1;
    a.qq2 = pb.qq2;
    a.pc = pb.pc;
    a.pq2 = pb.pq2;
    const dim3 grid((unsigned)((max_pairs + 3) / 4));
    if (bv.D_pad <= 512)
        hipLaunchKernelGGL(bq_pairprep_kernel<2>, grid, dim3(256), 0, st, a);
    else if (bv.D_pad <= 1024)
        hipLaunchKernelGGL(bq_pairprep_kernel<4>, grid, dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(bq_pairprep_kernel<0>, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_bscan_tiles(const BatchedView &bv, const BqWork &w, const BqPairBufs &pb, int NP, int topk,
                              float *cand_d, uint32_t *cand_id, int qt, int qs, int grid, hipStream_t st) {
    BScanArgs a;
    a.bv = bv;
    a.qq1 = pb.qq1;
    a.qq2 = pb.qq2;
    a.pc = pb.pc;
    a.pq2 = pb.pq2;
    a.tile_clu = w.tile_clu;
    a.tile_first = w.tile_first;
    a.tile_cnt = w.tile_cnt;
    a.tile_v0 = w.tile_v0;
    a.tile_cand = w.tile_cand;
    a.ntiles = w.ntiles;
    a.sv = w.sv;
    a.work = w.work;
    a.pairs = w.pairs;
    a.NP = NP;
    a.qbound = w.qbound;
    a.topk = topk;
    a.cand_d = cand_d;
    a.cand_id = cand_id;
    if (qs == 3) {
        switch (qt) {
        case 16: return launch_bscan_q<1, 1, 3>(a, grid, st);
        case 32: return launch_bscan_q<2, 1, 3>(a, grid, st);
        case 64: return launch_bscan_q<4, 1, 3>(a, grid, st);
        default: return launch_bscan_q<4, 2, 3>(a, grid, st);
        }
    }
    switch (qt) {
    case 16: return qs == 2 ? launch_bscan_q<1, 1, 2>(a, grid, st) : launch_bscan_q<1, 1, 1>(a, grid, st);
    case 32: return qs == 2 ? launch_bscan_q<2, 1, 2>(a, grid, st) : launch_bscan_q<2, 1, 1>(a, grid, st);
    case 64: return qs == 2 ? launch_bscan_q<4, 1, 2>(a, grid, st) : launch_bscan_q<4, 1, 1>(a, grid, st);
    default: return qs == 2 ? launch_bscan_q<4, 2, 2>(a, grid, st) : launch_bscan_q<4, 2, 1>(a, grid, st);
    }
}

hipError_t launch_bq_merge(const float *cand_d, const uint32_t *cand_id, const uint32_t *probes, const BqWork &w,
                           const int32_t *npq, int64_t nq, int NP, int topk, uint32_t *out_ids, float *out_d,
                           hipStream_t st) {
    const unsigned g = (unsigned)((nq + 3) / 4);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(g), dim3(256), 0, st, cand_d, cand_id, probes, w.slot, w.counts, w.candstart,
                           w.nsplit, npq, nq, NP, topk, out_ids, out_d);
    };
    if (topk <= 64)
        go(bq_merge_kernel<1>);
    else if (topk <= 128)
        go(bq_merge_kernel<2>);
    else
        go(bq_merge_kernel<4>);
    return hipGetLastError();
}

}  // namespace saqgpu
