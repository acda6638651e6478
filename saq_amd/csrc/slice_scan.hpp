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
// what the list scan hands the refine kernel per candidate (dwords): [0] vector in its list,
// [1] fast estimate, [2 + s] segment s's fast-scan inner product (Lut::compFastIP)
__host__ __device__ inline int pre_stride(int nseg) { return nseg + 2; }
// chunk descriptor (dwords): [0] query, [1] list id, [2] first candidate, [3] candidates, then
// per segment the list scan's scalars SC_SUMQCB, SC_QL2, SC_SUMQ, SC_VFAC, SC_IPQC
constexpr int kDescHead = 4, kDescSc = 5;
__host__ __device__ inline int desc_stride(int nseg) { return kDescHead + kDescSc * nseg; }

// Block info, candidate records and descriptors are reserved from kSlots independent counters,
// each owning an equal share of every buffer: atomics on one address serialise.
constexpr int kSlots = 64;
// counters (uint32 entries of SliceArgs::ctr)
enum {
    CTR_OVERFLOW = 0,  // some slice ran out of space: the search is repeated with more
    CTR_MAX_BINFO,     // the most block-info dwords / candidate records / descriptors one slot
    CTR_MAX_CAND,      //   of one slice asked for
    CTR_MAX_DESC,
    CTR_BINFO = 4,                    // [kSlots] block-info dwords used per slot
    CTR_CAND = CTR_BINFO + kSlots,    // [kSlots] 64-bit: descriptors << 32 | candidate records
    CTR_N = CTR_CAND + 2 * kSlots
};

// A query's probe sequence as work items: item 0 is the first chunk (8 blocks, 256 vectors) of
// probed list 0, item 1 the rest of list 0, item j >= 2 all of probed list j - 1.  Slices are
// ranges of items, so the first slice -- refined without a bound -- stays one chunk long.
constexpr int kItemChunk = 8;  // blocks of item 0
__host__ __device__ inline int item_list(int item) { return item == 0 ? 0 : item - 1; }
__host__ __device__ inline int item_first_block(int item, int nb) {
    return item == 1 ? (nb < kItemChunk ? nb : kItemChunk) : 0;
}
__host__ __device__ inline int item_end_block(int item, int nb) {
    return item == 0 ? (nb < kItemChunk ? nb : kItemChunk) : nb;
}

struct SliceArgs {
    IndexView ix;
    const float *qrot;
    const float *bound;
    const uint32_t *probes;
    int32_t nprobe;          // probe-list stride
    const int32_t *npq;      // lists scanned per query (nullptr: nprobe)
    int32_t p0, p1;          // the slice: items [p0, p1)
    int32_t topk;
    int64_t nq;
    int64_t q0;              // first query of this launch (list-scan launches are split by query)
    int32_t first, last;     // first slice: empty pools; last slice: write the results
    float *pool_d;           // nq x topk raw pool distances (IP: negated, as ResultPool keeps them)
    uint32_t *pool_id;       // nq x topk
    int32_t *pool_size;      // nq
    uint64_t *met;           // nq x 3 running QueryRuntimeMetrics
    uint32_t *pair_blk;      // nq x (p1 - p0): first block-info dword of each (query, item)
    uint32_t *binfo;
    uint32_t *cand;          // candidate records
    uint32_t *pre;           // list-scan output per candidate (pre_stride dwords)
    uint32_t *desc;          // chunk descriptors
    uint32_t binfo_cap;      // dwords per slot
    uint32_t cand_cap;       // records per slot
    uint32_t desc_cap;       // descriptors per slot
    uint32_t *ctr;           // CTR_N counters
    int32_t cmax;            // candidates whose long records a refine wavefront stages in LDS at once
    uint32_t *out_ids;
    float *out_dists;
    uint64_t *out_metrics;
};

hipError_t launch_slice_reset(uint32_t *ctr, int all, hipStream_t st);
hipError_t launch_list_scan(const SliceArgs &args, int dist_type, size_t smem, hipStream_t st);
hipError_t launch_refine(const SliceArgs &args, int dist_type, size_t smem, int grid, hipStream_t st);
hipError_t launch_replay(const SliceArgs &a, int dist_type, hipStream_t st);
size_t list_scan_smem_bytes(const IndexView &ix);
size_t refine_smem_bytes(const IndexView &ix, int cmax);

}  // namespace saqgpu
