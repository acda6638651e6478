// batched.hpp -- device layout and launchers of the batched int8-MFMA scan (batched.hip).
#pragma once
#include "common.hpp"

namespace saqgpu {

constexpr int kMB = 16;     // vectors per MFMA column block (one mblock record)
// Vector ranges of a list's work tiles: at least kMinSplit vectors, at most about the launch's
// work / (kSplitWaves / waves per workgroup) (kSplitWaves: the resident wavefronts of a 256-CU
// chip at 3 waves per SIMD, the int8 tile kernels' occupancy).
constexpr int kMinSplit = 512;
constexpr int kSplitWaves = 3072;
// Lists the grouping scan holds in registers: 1024 threads x kScanRounds each (the nlist limit of the
// batched scan).
constexpr int kScanRounds = 16;

struct BSeg {
    int32_t dim, bits, doff;
    int32_t kq_off;       // offset of the segment in a quantised query row (bits > 0)
    int32_t code_off;     // byte offset of the segment's code chunks in an mblock record
    int32_t chunk_bytes;  // bytes of one 64-dim chunk: 512 (b <= 4), 1024 (b <= 8), 2048 (b >= 9)
    float sqd;            // 2 / 2^b
    int32_t qs;           // index among the quantised segments, -1 for b = 0
};

struct BatchedView {
    int32_t nseg, nqs, D_pad, kq;  // segments, quantised segments, padded dim, sum of quantised dims
    int32_t fac_off;               // byte offset of the factors in an mblock record
    int32_t kq_stride;             // LDS row stride of a quantised query (kq + 16)
    int64_t mb_stride;             // bytes per mblock record
    const uint8_t *mb;             // records
    const int64_t *clu_mb;         // first mblock of list c
    const int32_t *clu_n;          // vectors in list c
    const float *crot;             // K x D_pad rotated centroids
    const BSeg *segs;
};

// Grouping workspace: (query, list) pairs grouped by list, and the work tiles over them (a
// work tile: up to QT pairs of one list and one vector range of it, whose per-pair top-k goes
// to candidate lists tile_cand .. + cnt - 1).
struct BqWork {
    int32_t *counts, *slot, *cstart, *candstart, *nsplit, *ntiles, *npairs, *sv, *pairs;
    int32_t *tile_clu, *tile_first, *tile_cnt, *tile_v0, *tile_cand, *work;
    uint32_t *qbound;  // per query: the distance bound its tiles share (nq entries)
    int32_t *tfirst;   // per list in the scan order: its first work tile
};
// Upper bounds of the work tiles and candidate lists of a launch.
inline int64_t bq_max_tiles(int64_t npairs, int K) { return npairs + K + kSplitWaves; }
inline int64_t bq_max_cands(int64_t npairs, int QT) { return npairs + (int64_t)kSplitWaves * QT; }

// Per grouped pair: the quantised residual rows (A operand) and the pair's constants.
struct BqPairBufs {
    int8_t *qq1, *qq2;  // [pairs][kq]
    float *pc;          // [pairs][nqs][3]
    float *pq2;         // [pairs]
};

// LDS of the tile kernel's workgroup for a query tile of qt (16, 32, 64 or 128) queries.
size_t bscan_smem_bytes(const BatchedView &bv, int topk, int qt, int qs);
hipError_t launch_bq_build(const IndexView &ix, const BatchedView &bv, const int32_t *mb_clu, int64_t total_mb,
                           const int32_t *chunk_seg, const int32_t *chunk_kc, int nchunks, uint8_t *out,
                           hipStream_t st);
// order: the lists by decreasing size (the order work tiles are handed out in).
hipError_t launch_bq_group(const uint32_t *probes, const int32_t *npq, int64_t nq, int NP, int K, int QT, int qs,
                           const int32_t *order, const int32_t *clu_n, const BqWork &w, hipStream_t st);
hipError_t launch_bq_merge(const float *cand_d, const uint32_t *cand_id, const uint32_t *probes, const BqWork &w,
                           const int32_t *npq, int64_t nq, int NP, int topk, uint32_t *out_ids, float *out_d,
                           hipStream_t st);

// Quantised residual rows and constants of every grouped pair (max_pairs bounds the grid).
hipError_t launch_bq_pairprep(const BatchedView &bv, const float *qrot, const uint32_t *probes, const BqWork &w,
                              const BqPairBufs &pb, int NP, int64_t max_pairs, int qs, hipStream_t st);

// One launch of the tile kernel: qt = queries per work tile (16: one wave; 32, 64: 2 or 4 waves of
// 16 queries; 128: 4 waves of 32), qs = terms of the quantised query (1 = int8, 2 = int8x2).  grid
// is an upper bound on the workgroups.
hipError_t launch_bscan_tiles(const BatchedView &bv, const BqWork &w, const BqPairBufs &pb, int NP, int topk,
                              float *cand_d, uint32_t *cand_id, int qt, int qs, int grid, hipStream_t st);

}  // namespace saqgpu
