// slice_scan.hpp -- arguments and record layouts of the sliced plain scan (slice_scan.hip),
// shared with the host side that sizes its buffers (api.cpp).
#pragma once
#include "common.hpp"

namespace saqgpu {

constexpr uint32_t kNoPair = 0xFFFFFFFFu;  // pair without block info (empty list)

// block info record (dwords): [0] first candidate record, [1] candidates in the block;
// with several segments also [2] stage-1 block minimum, [3] the largest stage-2 block
// minimum, [4 + s] the block minimum after segment s's fast estimate.
__host__ __device__ inline int binfo_stride(int nseg) { return nseg == 1 ? 2 : 4 + nseg; }
// candidate record (dwords): [0] PID, [1] lane in its block, [2] fast estimate (the value
// the masks compare), [3 + s] running accurate sum after segment s.
__host__ __device__ inline int cand_stride(int nseg) { return nseg + 3; }

struct SliceArgs {
    IndexView ix;
    const float *qrot;
    const float *bound;
    const uint32_t *probes;
    int32_t nprobe;          // probe-list stride
    const int32_t *npq;      // lists scanned per query (nullptr: nprobe)
    int32_t p0, p1;          // the slice
    int32_t topk;
    int64_t nq;
    int64_t q0;              // first query of this launch (list-scan launches are split by query)
    int32_t first, last;     // first slice: empty pools; last slice: write the results
    float *pool_d;           // nq x topk raw pool distances (IP: negated, as ResultPool keeps them)
    uint32_t *pool_id;       // nq x topk
    int32_t *pool_size;      // nq
    uint64_t *met;           // nq x 3 running QueryRuntimeMetrics
    uint32_t *pair_blk;      // nq x (p1 - p0): first block-info dword of each pair
    uint32_t *binfo;
    uint32_t *cand;
    uint32_t binfo_cap;      // dwords
    uint32_t cand_cap;       // records
    uint32_t *ctr;           // [0] block-info dwords used, [1] candidate records used, [2] overflow
    int32_t cmax;            // candidates whose long records are staged in LDS at once
    uint32_t *out_ids;
    float *out_dists;
    uint64_t *out_metrics;
};

hipError_t launch_slice_reset(uint32_t *ctr, int all, hipStream_t st);
hipError_t launch_list_scan(const SliceArgs &args, int dist_type, size_t smem, hipStream_t st);
hipError_t launch_replay(const SliceArgs &a, int dist_type, hipStream_t st);
size_t list_scan_smem_bytes(const IndexView &ix, int cmax);

}  // namespace saqgpu
