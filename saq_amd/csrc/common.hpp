// common.hpp -- shared layout and device helpers of the gfx950 SAQ kernels.
//
// Device layout of an index (one allocation per array, clusters concatenated):
//
//   blocks      per 32-vector block, a record of  nseg*32 f32 o_l2norm (segment-major)
//               followed by, for every segment with bits > 0, its short codes as
//               D_s/64 groups x 32 lanes x 8 bytes.  Lane j of group g holds bytes
//               8g..8g+7 of vector j's MSB-first short code (the bytes the reference
//               packs in ClusterPacker::pack_short_codes, cluster_packer.hpp:124-139).
//               One vector = one lane: a thread streams its own vector with 8-byte loads
//               and 32 lanes of a group form one 256-byte coalesced line.
//   long_codes  per vector, every segment's extra-bit code in the reference's
//               CodeHelper<b-1> compacted layout, each slot padded to 16 bytes
//               (the reference's non-compact per-vector layout, cluster_data.hpp:226-229).
//   exfac       per vector per segment {rescale, error} (ExFactor, cluster_data.hpp:27-30).
//   ipcent      per block per segment 32 f32 ip_cent_oa (kept for export and IP).
//   ids         per vector PID.
//   crot        per cluster the rotated centroid, segments concatenated (padded dim).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace saqgpu {

constexpr int kBlock = 32;
// Bytes past the last block record the block array keeps readable: the list scan loads the
// short codes of 8 groups at a time and one step ahead without clamping at the record's end,
// so a load may reach up to 15 groups (3840 bytes) past a record; those values are never summed.
constexpr int kBlockTailPad = 4096;
constexpr int kMaxSeg = 64;

struct SegDesc {
    int32_t dim;         // D_s
    int32_t bits;        // b_s
    int32_t doff;        // first padded dimension of the segment
    int32_t group_off;   // first 8-byte code group of the segment in a block (bits > 0)
    int32_t ngroups;     // D_s / 64 when bits > 0, else 0
    int32_t long_off;    // byte offset of the segment's slot in a vector's long record
    int32_t long_bytes;  // D_s * (b_s - 1) / 8
    float rescale_sqrtD; // (4 / 0.8f) * (float)(1.0 / sqrt((float)D_s)), lut.hpp:47,492
    float sqd;           // (float)(2.0 / (1 << b_s)), caq_estimator.hpp:59
    int32_t pad_;
};

struct IndexView {
    int32_t nseg, num_dim, D_pad, num_cen;
    int32_t ngroups;       // total 8-byte code groups per vector (sum over bits>0 segments)
    int32_t fac_floats;    // nseg * 32
    int64_t block_stride;  // bytes per block record
    int64_t long_stride;   // bytes per vector long record
    const uint8_t *blocks;
    const uint8_t *long_codes;
    const float *exfac;
    const uint32_t *ids;
    const float *crot;
    const float *ipcent;
    const int64_t *clu_blk;  // first block of cluster c
    const int64_t *clu_vec;  // first vector of cluster c
    const int32_t *clu_n;    // vectors in cluster c
    const SegDesc *segs;
};

// ---------------------------------------------------------------- reductions
// The reference's fp32 reduction tree over 16 lanes (Eigen predux<Packet16f> and GCC's
// _mm512_reduce_add_ps): a[i]+a[i+8], then +4, then +2, then +1.  Executed as a xor
// butterfly over 16 consecutive lanes: both partners of every step compute the same
// commutative sum, so lane 0 of each 16-lane group ends with exactly the tree's value.
__device__ __forceinline__ float tree16_xor(float x) {
    x = x + __shfl_xor(x, 8, 16);
    x = x + __shfl_xor(x, 4, 16);
    x = x + __shfl_xor(x, 2, 16);
    x = x + __shfl_xor(x, 1, 16);
    return x;
}

// Eigen's linear-vectorised redux (Redux.h:227-272) over n elements (n a multiple of 16,
// n >= 16) evaluated by the 32 lanes 0..31 of a wavefront: lane t < 16 is packet
// accumulator p0[t], lane t >= 16 is p1[t-16].  elem(i) yields element i.  Every lane of
// the wavefront must call it (shuffles); the result is valid in all lanes.
template <class F>
__device__ __forceinline__ float eigen_redux_wave(int n, F elem) {
    const int lane = threadIdx.x & 63;
    const int t = lane & 31;
    const int aligned2 = n / 32 * 32;
    float v = 0.0f;
    if (n >= 32) {
        v = elem(t);
        // unrolled so that elements read from global memory are loaded eight at a time
#pragma unroll 8
        for (int idx = 32; idx < aligned2; idx += 32) v = v + elem(idx + t);
    } else if (t < 16) {
        v = elem(t);
    }
    const float hi = __shfl_down(v, 16, 32);
    if (n >= 32) {
        v = v + hi;  // p0 + p1 in lanes 0..15
        if (n > aligned2 && t < 16) v = v + elem(aligned2 + t);
    }
    // tree over lanes 0..15 of each 32-lane half (lanes 16..31 compute garbage, unused)
    v = tree16_xor(v);
    return __shfl(v, 0, 64);
}

__device__ __forceinline__ float max0(float v) { return (0.0f > v) ? 0.0f : v; }

// static_cast<uint16_t>(float) as GCC emits it on x86-64 (cvttss2si, low 16 bits).
__device__ __forceinline__ uint32_t cvt_u16(float x) {
    int32_t v;
    if (x != x || !(x > -2147483904.0f && x < 2147483648.0f))
        v = INT32_MIN;
    else
        v = (int32_t)x;
    return (uint32_t)(uint16_t)v;
}

// Order-preserving key of a float (ascending keys <=> ascending floats, -0 < +0).
__device__ __forceinline__ uint32_t float_key(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// LDS DMA of `bytes` (a multiple of 16) from global memory to LDS by a 256-thread workgroup:
// every wavefront moves 1 KB per global_load_lds_dwordx4 (lane l -> dst + 16 l).
__device__ __forceinline__ void dma_to_lds(const uint8_t *src, uint8_t *dst, int bytes, int wave, int lane) {
    for (int off = wave * 1024; off < bytes; off += 4 * 1024) {
        const int i = off + lane * 16;
        if (i < bytes) __builtin_amdgcn_global_load_lds(src + i, dst + off, 16, 0, 0);
    }
}

// Waits for this wavefront's outstanding vector-memory operations (LDS DMA included).
__device__ __forceinline__ void dma_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

}  // namespace saqgpu
