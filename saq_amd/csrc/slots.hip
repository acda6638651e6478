// slots.hip -- IVF::construct's bucketing on the GPU (saqlib/index/ivf.hpp:149-159): the members
// of every list in increasing id order.  A stable radix sort of the (cluster id, id) pairs gives
// each slot of the index (list c occupies slots clu_vec[c] .. clu_vec[c] + clu_n[c] - 1) its
// source row; a fill pass derives the slot's block and lane and checks the lists' sizes.
#include <hipcub/hipcub.hpp>

#include "common.hpp"

namespace saqgpu {

// ids 0..n-1, and bit 0 of *bad for a cluster id outside [0, num_cen)
__global__ void slot_ids_kernel(const uint32_t *cid, int64_t n, uint32_t num_cen, uint32_t *ids, int *bad) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    ids[i] = (uint32_t)i;
    if (cid[i] >= num_cen) atomicOr(bad, 1);
}

// slot t holds sorted pair t: its list c and rank r = t - clu_vec[c] within it (bit 1 of *bad
// where the sorted lists do not start where the index's sizes put them)
__global__ void slot_fill_kernel(const uint32_t *clu, int64_t n, uint32_t num_cen, const int64_t *clu_vec,
                                 const int64_t *clu_blk, const int32_t *clu_n, int64_t *gvec, int64_t *gblk,
                                 int32_t *lane, int *bad) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const uint32_t c = clu[t];
    int64_t r = 0, blk = 0;
    if (c >= num_cen) {
        atomicOr(bad, 1);
    } else {
        r = t - clu_vec[c];
        if (r < 0 || r >= clu_n[c]) {
            atomicOr(bad, 2);
            r = 0;
        }
        blk = clu_blk[c] + r / kBlock;
    }
    gvec[t] = t;
    gblk[t] = blk;
    lane[t] = (int32_t)(r % kBlock);
}

size_t slot_sort_bytes(int64_t n) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                             (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0, 32);
    return bytes;
}

hipError_t launch_make_slots(const uint32_t *cid, int64_t n, int num_cen, const int64_t *clu_vec,
                             const int64_t *clu_blk, const int32_t *clu_n, uint32_t *iota, uint32_t *src,
                             uint32_t *clu, int64_t *gvec, int64_t *gblk, int32_t *lane, void *temp,
                             size_t temp_bytes, int *bad, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(slot_ids_kernel, dim3(g), dim3(256), 0, st, cid, n, (uint32_t)num_cen, iota, bad);
    int end_bit = 1;
    while (end_bit < 32 && (1u << end_bit) < (uint32_t)num_cen) ++end_bit;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, cid, clu, iota, src, (int)n, 0, end_bit, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(slot_fill_kernel, dim3(g), dim3(256), 0, st, clu, n, (uint32_t)num_cen, clu_vec, clu_blk,
                       clu_n, gvec, gblk, lane, bad);
    return hipGetLastError();
}

}  // namespace saqgpu
