// api.cpp -- host side of the C ABI declared in include/saqgpu.h.
//
// Owns the device-resident index, converts between the reference's SaqCluData layout
// (saqlib/quantization/cluster_data.hpp:179-337) and the device layout documented in
// common.hpp, and dispatches the gfx950 kernels.  No CPU fallback exists: every search
// and encode runs on the GPU or fails with a status.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/saqgpu.h"
#include "batched.hpp"
#include "common.hpp"
#include "slice_scan.hpp"

namespace saqgpu {
hipError_t launch_query_prep(const IndexView &ix, const float *queries, int64_t nq, const float *variance,
                             const float *rot, const int64_t *rot_off, float m, float *qrot, float *bound,
                             int max_dim, hipStream_t st);
hipError_t launch_probe(const float *centroids, int K, int D, const float *queries, int64_t nq, int nprobe,
                        int dist_type, float *dist_ws, int32_t *ties, uint32_t *probes, hipStream_t st);
hipError_t launch_sqnorm_rows(const float *x, int64_t rows, int D, float *out, hipStream_t st);
hipError_t launch_split_bf16(const float *x, int64_t rows, int D, int second, uint16_t *out, hipStream_t st);
size_t estimate_smem_bytes(const IndexView &ix);
hipError_t launch_estimate(const IndexView &ix, const float *qrot, const float *bound, const uint32_t *probes,
                           const int64_t *offs, const int2 *pieces, int npieces, int dist_type, uint32_t *out_ids,
                           float *out_acc, float *out_fast, float *out_vars, hipStream_t st);
hipError_t launch_probe_select_ip(const float *ip, const float *cnorm, int K, int64_t nq, int nprobe, int dist_type,
                                  uint32_t *probes, hipStream_t st);
hipError_t launch_probe_exact(float *ip, const float *cnorm, const float *centroids, int K, int D,
                              const float *queries, int64_t nq, int nprobe, int dist_type, uint32_t *probes,
                              int32_t *ties, hipStream_t st);
hipError_t launch_encode_segment(const IndexView &ix, const float *data, const uint32_t *src, const int64_t *gvec,
                                 const int64_t *gblk, const int32_t *lane, const uint32_t *clu, const float *cent,
                                 const float *rot, int seg, int seg_dim, int64_t nslots, int64_t ld, int rd_lmt,
                                 float eps, int ori_qb, bool shared_cent, float *O, uint16_t *code, int *fail,
                                 hipStream_t st);
size_t slot_sort_bytes(int64_t n);
hipError_t launch_make_slots(const uint32_t *cid, int64_t n, int num_cen, const int64_t *clu_vec,
                             const int64_t *clu_blk, const int32_t *clu_n, uint32_t *iota, uint32_t *src,
                             uint32_t *clu, int64_t *gvec, int64_t *gblk, int32_t *lane, void *temp,
                             size_t temp_bytes, int *bad, hipStream_t st);
hipError_t launch_caq_only(const IndexView &ix, const int64_t *gvec, const int64_t *gblk, const int32_t *lane,
                           const uint32_t *clu, const float *cent, int seg, int64_t nslots, int64_t ld, int rd_lmt,
                           float eps, int ori_qb, float *O, uint16_t *code, int *fail, hipStream_t st);
hipError_t launch_single_quantize_segment(const IndexView &ix, const float *data, const float *rot, int seg,
                                          int seg_dim, int64_t nslots, int64_t ld, int rd_lmt, float eps, int ori_qb,
                                          float *O, uint16_t *code, int *fail, uint8_t *srec, int64_t sstride,
                                          int32_t code_off, int32_t long_off, int32_t exf_off, hipStream_t st);
struct SingleQ;
size_t single_query_state_bytes();
hipError_t launch_single_estimate(const IndexView &ix, const float *qrot, const float *bound, int64_t nq,
                                  SingleQ *qs, uint64_t *planes, const uint8_t *recs, int64_t stride, int64_t nrec,
                                  const int32_t *soff, int32_t exf_off, int32_t dist_type, float *out_vars,
                                  float *out_fast, float *out_acc, hipStream_t st);
}  // namespace saqgpu

using namespace saqgpu;

// The fixed device scratch of saqgpu_probe_noalloc: one copy per GPU, held until the process ends.
// Host code never uses the array's name as an address: hipGetSymbolAddress gives each device's
// copy (scratch_setup, from saqgpu_index_create), and only those pointers reach kernels.
static __device__ __align__(256) float scratch[SAQGPU_NOALLOC_SCRATCH_FLOATS];

namespace {
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(SAQGPU_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kPermLo[16] = {0, 8, 1, 9, 2, 10, 3, 11, 4, 12, 5, 13, 6, 14, 7, 15};  // fastscan.hpp:41-42

inline int64_t rd_up(int64_t x, int64_t y) { return (x + y - 1) / y * y; }

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    bool borrowed = false;  // another index's buffer (a search handle's view): never freed here
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }  // a function's scratch buffers go when it returns
    void borrow(const DevBuf &o) {
        release();
        p = o.p;
        n = o.n;
        borrowed = true;
    }
    int alloc(size_t cnt) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = cnt;
        if (cnt == 0) return 0;
        if (hipMalloc(&p, cnt * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return -1;
        }
        return 0;
    }
    int ensure(size_t cnt) { return cnt <= n && p ? 0 : alloc(cnt); }
    void release() {
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr;
        n = 0;
        borrowed = false;
    }
};

template <class T>
int upload(DevBuf<T> &b, const std::vector<T> &v) {
    if (b.ensure(std::max<size_t>(1, v.size()))) return fail(SAQGPU_ENOMEM, "allocation failed");
    if (!v.empty() && hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return fail(SAQGPU_EDEVICE, "upload failed");
    return SAQGPU_OK;
}
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch[kMaxDevices] = {};

// Resolves the device address of this device's scratch array (the device must be current).
int scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(SAQGPU_EINVAL, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return SAQGPU_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SAQGPU_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<float *>(p);
    return SAQGPU_OK;
}

float *scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}
}  // namespace

// What a captured batched step depends on: its shape, parameters and every buffer it touches.
struct BatchedGraphKey {
    int64_t nq;
    int nprobe, topk, precision, query_tile, exact_probe;
    const void *queries, *probes, *npq, *out_ids, *out_dists, *stream, *ws, *qq, *cand, *cdist, *wprobes;
    bool operator==(const BatchedGraphKey &o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

struct saqgpu_index {
    // a search handle (saqgpu_index_search_handle) reads its parent's device data through
    // borrowed buffers; the parent counts its open handles and refuses changes while any is open
    saqgpu_index *parent = nullptr;
    std::atomic<int> handles{0};
    int device = 0;
    int64_t num_data = 0;
    int num_dim = 0, num_cen = 0, nseg = 0, D_pad = 0, ngroups = 0, max_dim = 0;
    bool compact = false, has_rot = false;
    std::vector<int32_t> seg_dim, seg_bits;
    std::vector<SegDesc> segs;
    std::vector<int64_t> clu_blk, clu_vec, rot_off;
    std::vector<int32_t> clu_n;
    std::vector<uint8_t> clu_set;
    int64_t total_blocks = 0, block_stride = 0, long_stride = 0;
    DevBuf<uint8_t> blocks, long_codes;
    DevBuf<float> exfac, ipcent, crot, centroids, rot, variance;
    DevBuf<uint32_t> ids;
    DevBuf<int64_t> d_clu_blk, d_clu_vec, d_rot_off;
    DevBuf<int32_t> d_clu_n;
    DevBuf<SegDesc> d_segs;
    // search workspace; calls on the index serialise on mu and, on the device, on ws_free
    std::mutex mu;
    hipEvent_t ws_free = nullptr;
    DevBuf<float> ws_q, ws_qrot, ws_bound, ws_dists, ws_cdist;
    DevBuf<int32_t> ws_ties;   // per query: its probe list holds exact distance ties
    DevBuf<int64_t> ws_eoffs;  // IVF::estimate: per probe its first output slot
    DevBuf<int2> ws_epieces;   //   per workgroup its (probe, first block)
    DevBuf<uint32_t> ws_eids;  //   and its outputs
    DevBuf<float> ws_eout;
    DevBuf<uint32_t> ws_probes, ws_ids;
    DevBuf<uint64_t> ws_met;
    // sliced scan (slice_scan.hip): pools between slices, block info and candidate records;
    // the capacities grow to what a search needed and are kept for the next one
    DevBuf<float> sl_pool_d;
    DevBuf<uint32_t> sl_pool_id, sl_pair, sl_binfo, sl_cand, sl_pre, sl_desc, sl_ctr;
    DevBuf<int32_t> sl_pool_size;
    DevBuf<uint64_t> sl_met;
    // query groups of the sliced scan: each runs its slices on its own stream
    struct SliceGroup {
        hipStream_t stream;
        hipEvent_t done;
    };
    std::vector<SliceGroup> sl_groups;
    hipEvent_t sl_fork = nullptr;
    // rocBLAS handle of the coarse quantiser's GEMM (created on first use)
    rocblas_handle blas = nullptr;
    // batched layout (batched.hip): built from the clusters on first use
    bool bq_ready = false;
    int bq_tile = 0;  // autotuned queries per work tile (0 = not tuned)
    int bq_nqs = 0, bq_kq = 0, bq_fac_off = 0;
    int64_t bq_stride = 0, bq_total_mb = 0;
    std::vector<BSeg> bsegs;
    std::vector<int64_t> clu_mb;
    DevBuf<uint8_t> bq_mb;
    DevBuf<int64_t> d_clu_mb;
    DevBuf<int32_t> d_bq_order;  // lists by decreasing size
    DevBuf<BSeg> d_bsegs;
    DevBuf<int32_t> bq_ws, bq_npq;
    DevBuf<int8_t> bq_qq;
    DevBuf<float> bq_pc;
    DevBuf<float> cnorm;  // |c|^2 per centroid (GEMM-based coarse quantisers), computed once
    bool cnorm_ready = false;
    DevBuf<uint16_t> cen_bf3;  // centroid rows as bfloat16 hi | lo | hi (the batched coarse quantiser)
    bool cen_bf3_ready = false;
    DevBuf<uint16_t> ws_qbf3;  // query rows as bfloat16 hi | hi | lo
    DevBuf<float> bq_cand_d;
    DevBuf<uint32_t> bq_cand_id;
    // the captured batched step (SAQGPU_USE_GRAPH) and what it was captured with
    hipGraphExec_t bq_graph = nullptr;
    BatchedGraphKey bq_graph_key{};
    // measurement
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;

    IndexView view() const {
        IndexView v{};
        v.nseg = nseg;
        v.num_dim = num_dim;
        v.D_pad = D_pad;
        v.num_cen = num_cen;
        v.ngroups = ngroups;
        v.fac_floats = nseg * kBlock;
        v.block_stride = block_stride;
        v.long_stride = long_stride;
        v.blocks = blocks.p;
        v.long_codes = long_codes.p;
        v.exfac = exfac.p;
        v.ids = ids.p;
        v.crot = crot.p;
        v.ipcent = ipcent.p;
        v.clu_blk = d_clu_blk.p;
        v.clu_vec = d_clu_vec.p;
        v.clu_n = d_clu_n.p;
        v.segs = d_segs.p;
        return v;
    }

    // reference layout of one cluster (SaqCluData ctor, cluster_data.hpp:206-294)
    struct RefLayout {
        int64_t nb, fac_per_block, code_per_block, long_total, long_stride_ref;
        std::vector<int64_t> code_off, long_off, long_bytes;
        bool single, compact;
    };
    RefLayout ref_layout(int64_t n) const {
        RefLayout L;
        L.nb = (n + kBlock - 1) / kBlock;
        L.single = nseg == 1;
        L.compact = L.single || compact;
        L.code_off.resize(nseg);
        L.long_off.resize(nseg);
        L.long_bytes.resize(nseg);
        int64_t code = 0, lpv = 0, ltot = 0;
        for (int s = 0; s < nseg; ++s) {
            L.code_off[s] = code;
            code += seg_bits[s] ? (int64_t)seg_dim[s] * kBlock / 8 : 0;
            L.long_bytes[s] = seg_bits[s] ? (int64_t)seg_dim[s] * (seg_bits[s] - 1) / 8 : 0;
            if (L.compact) {
                L.long_off[s] = ltot;
                ltot += rd_up(L.long_bytes[s] * n, 16);
            } else {
                L.long_off[s] = lpv;
                lpv += rd_up(L.long_bytes[s], 16);
            }
        }
        if (L.single) {
            L.fac_per_block = 0;
            for (auto &c : L.code_off) c += 2 * kBlock * 4;
            L.code_per_block = 2 * kBlock * 4 + code;
        } else {
            L.fac_per_block = (int64_t)nseg * 2 * kBlock;
            L.code_per_block = code;
        }
        L.long_stride_ref = L.compact ? 0 : lpv;
        L.long_total = L.compact ? ltot : lpv * n;
        return L;
    }
};

namespace {
// One call on an index.  The host mutex serialises the index's host-side state.  On the device,
// the previous call recorded ws_free on its stream once its work was queued; this call's stream
// waits for it before touching the workspace.  Calls with SAQGPU_DEVICE_PTRS return before their
// kernels end, and the next call may come on another stream or thread.
struct IndexCall {
    saqgpu_index *ix;
    hipStream_t st;
    std::lock_guard<std::mutex> lk;
    bool begun = false;
    IndexCall(saqgpu_index *i, hipStream_t s) : ix(i), st(s), lk(i->mu) {}
    int begin() {
        HIPCHK(hipSetDevice(ix->device));
        HIPCHK(hipStreamWaitEvent(st, ix->ws_free, 0));
        begun = true;
        return SAQGPU_OK;
    }
    ~IndexCall() {
        if (begun) (void)hipEventRecord(ix->ws_free, st);
    }
};

// Entry points that change the index: refused on a search handle and while handles are open.
int check_mutable(const saqgpu_index *ix) {
    if (ix->parent) return fail(SAQGPU_ESTATE, "a search handle cannot change the index");
    if (ix->handles.load() > 0) return fail(SAQGPU_ESTATE, "the index has open search handles");
    return SAQGPU_OK;
}
}  // namespace

namespace {
// |c|^2 of every centroid, once per index (the centroids do not change).
int ensure_cnorm(saqgpu_index *ix, hipStream_t st) {
    if (ix->cnorm_ready) return SAQGPU_OK;
    if (ix->cnorm.ensure(ix->num_cen)) return fail(SAQGPU_ENOMEM, "allocation failed");
    HIPCHK(launch_sqnorm_rows(ix->centroids.p, ix->num_cen, ix->num_dim, ix->cnorm.p, st));
    ix->cnorm_ready = true;
    return SAQGPU_OK;
}

// The centroid rows split into bfloat16 parts (hi | lo | hi), once per index.
int ensure_cen_bf3(saqgpu_index *ix, hipStream_t st) {
    if (ix->cen_bf3_ready) return SAQGPU_OK;
    if (ix->cen_bf3.ensure((size_t)ix->num_cen * 3 * ix->num_dim)) return fail(SAQGPU_ENOMEM, "allocation failed");
    HIPCHK(launch_split_bf16(ix->centroids.p, ix->num_cen, ix->num_dim, 0, ix->cen_bf3.p, st));
    ix->cen_bf3_ready = true;
    return SAQGPU_OK;
}

int blas_on(saqgpu_index *ix, hipStream_t st) {
    if (!ix->blas && rocblas_create_handle(&ix->blas) != rocblas_status_success) {
        ix->blas = nullptr;
        return fail(SAQGPU_EDEVICE, "rocblas_create_handle failed");
    }
    if (rocblas_set_stream(ix->blas, st) != rocblas_status_success) return fail(SAQGPU_EDEVICE, "rocblas_set_stream failed");
    return SAQGPU_OK;
}

// q.c for every (query, centroid) into ws_cdist (nq x K) by rocBLAS's fp32 GEMM: in column-major
// terms ip^T (K x nq, ld K) = centroids (D x K, ld D)^T * queries (D x nq, ld D).
int ip_gemm(saqgpu_index *ix, const float *dq, int64_t nq, hipStream_t st) {
    const int K = ix->num_cen, D = ix->num_dim;
    if (int rc = blas_on(ix, st)) return rc;
    const float one = 1.0f, zero = 0.0f;
    if (rocblas_sgemm(ix->blas, rocblas_operation_transpose, rocblas_operation_none, K, (int)nq, D, &one,
                      ix->centroids.p, D, dq, D, &zero, ix->ws_cdist.p, K) != rocblas_status_success)
        return fail(SAQGPU_EDEVICE, "rocblas_sgemm failed");
    return SAQGPU_OK;
}

// FlatInitializer::centroids_distances + partial_sort (initializer.hpp:52-76): the GEMM filter
// (probe_exact_kernel: candidates within a proven fp32 error bound, their distances in Eigen's
// order, an exact selection) when nprobe <= 64 and a query fits in LDS, else the distances of
// every pair in Eigen's order (probe_tile_kernel).  Both give the same lists.  With a larger
// nprobe the window around the nprobe-th key holds too many centroids for the filter to pay
// (at 16384 lists and nprobe 256, often more than its 1024 places).  ws_cdist holds nq x num_cen
// floats.  SAQGPU_PROBE=tile / =gemm forces a form (gemm: any nprobe <= 1024).
int probe_gemm(saqgpu_index *ix, const float *dq, int64_t nq, int nprobe, int dist_type, uint32_t *probes,
               hipStream_t st);

int exact_probe(saqgpu_index *ix, const float *dq, int64_t nq, int nprobe, int dist_type, uint32_t *probes,
                hipStream_t st) {
    if (ix->ws_ties.ensure(std::max<int64_t>(1, nq))) return fail(SAQGPU_ENOMEM, "workspace allocation failed");
    const char *e = getenv("SAQGPU_PROBE");
    const bool tile = e && !strcmp(e, "tile"), gemm = e && !strcmp(e, "gemm");
    if (!tile && nprobe <= (gemm ? 1024 : 64) && (size_t)ix->num_dim * 4 <= 48 * 1024) {
        int rc;
        if ((rc = ensure_cnorm(ix, st)) || (rc = ip_gemm(ix, dq, nq, st))) return rc;
        HIPCHK(launch_probe_exact(ix->ws_cdist.p, ix->cnorm.p, ix->centroids.p, ix->num_cen, ix->num_dim, dq, nq,
                                  nprobe, dist_type, probes, ix->ws_ties.p, st));
        return SAQGPU_OK;
    }
    HIPCHK(launch_probe(ix->centroids.p, ix->num_cen, ix->num_dim, dq, nq, nprobe, dist_type, ix->ws_cdist.p,
                        ix->ws_ties.p, probes, st));
    return SAQGPU_OK;
}

}  // namespace

extern "C" {

const char *saqgpu_last_error(void) { return g_err.c_str(); }

int saqgpu_index_create(const saqgpu_index_desc *d, saqgpu_index **out) {
    if (!d || !out) return fail(SAQGPU_EINVAL, "null argument");
    *out = nullptr;
    if (d->num_seg < 1 || d->num_seg > SAQGPU_MAX_SEG) return fail(SAQGPU_EINVAL, "num_seg out of range");
    if (d->num_dim < 1 || d->num_cen < 1 || d->num_data < 0) return fail(SAQGPU_EINVAL, "bad sizes");
    if (!d->seg_dim || !d->seg_bits) return fail(SAQGPU_EINVAL, "null plan");
    int dsum = 0;
    for (int s = 0; s < d->num_seg; ++s) {
        if (d->seg_dim[s] <= 0 || d->seg_dim[s] % 64 || d->seg_bits[s] < 0 || d->seg_bits[s] > 13)
            return fail(SAQGPU_EINVAL, "segment dims must be multiples of 64 and bits in 0..13");
        dsum += d->seg_dim[s];
    }
    if (dsum != (int)rd_up(d->num_dim, 64)) return fail(SAQGPU_EINVAL, "segment dims must sum to the padded dimension");
    int qdims = 0;
    for (int s = 0; s < d->num_seg; ++s) qdims += d->seg_bits[s] ? d->seg_dim[s] : 0;
    if (qdims > 64 * SAQGPU_MAX_GROUPS)
        return fail(SAQGPU_EUNSUPPORTED, "at most 4096 quantised dimensions (64 code groups of 64) are supported");
    auto *ix = new saqgpu_index();
    ix->device = d->device;
    if (hipSetDevice(d->device) != hipSuccess) {
        delete ix;
        return fail(SAQGPU_EDEVICE, "hipSetDevice failed");
    }
    if (int rc = scratch_setup(d->device)) {
        delete ix;
        return rc;
    }
    if (hipEventCreateWithFlags(&ix->ws_free, hipEventDisableTiming) != hipSuccess) {
        delete ix;
        return fail(SAQGPU_EDEVICE, "hipEventCreate failed");
    }
    ix->num_data = d->num_data;
    ix->num_dim = d->num_dim;
    ix->num_cen = d->num_cen;
    ix->nseg = d->num_seg;
    ix->compact = d->use_compact_layout != 0;
    ix->has_rot = d->rotations != nullptr;
    int doff = 0, goff = 0, loff = 0;
    int64_t roff = 0;
    for (int s = 0; s < d->num_seg; ++s) {
        const int D = d->seg_dim[s], b = d->seg_bits[s];
        if (D <= 0 || D % 64 || b < 0 || b > 13) {
            delete ix;
            return fail(SAQGPU_EINVAL, "segment dims must be multiples of 64 and bits in 0..13");
        }
        SegDesc sd{};
        sd.dim = D;
        sd.bits = b;
        sd.doff = doff;
        sd.group_off = goff;
        sd.ngroups = b ? D / 64 : 0;
        sd.long_off = loff;
        sd.long_bytes = b ? D * (b - 1) / 8 : 0;
        sd.rescale_sqrtD = (4 / 0.8f) * (float)(1.0 / (double)std::sqrt((float)D));
        sd.sqd = (float)(2.0 / (1 << b));
        ix->segs.push_back(sd);
        ix->seg_dim.push_back(D);
        ix->seg_bits.push_back(b);
        ix->rot_off.push_back(roff);
        roff += (int64_t)D * D;
        doff += D;
        goff += sd.ngroups;
        loff += (int)rd_up(sd.long_bytes, 16);
        ix->max_dim = std::max(ix->max_dim, D);
    }
    ix->D_pad = doff;
    ix->ngroups = goff;
    ix->long_stride = loff;
    ix->block_stride = (int64_t)ix->nseg * kBlock * 4 + (int64_t)goff * 256;
    if (ix->D_pad != (int)rd_up(d->num_dim, 64)) {
        delete ix;
        return fail(SAQGPU_EINVAL, "segment dims must sum to the padded dimension");
    }
    int64_t blk = 0, vec = 0;
    for (int c = 0; c < d->num_cen; ++c) {
        const int64_t n = d->cluster_sizes ? d->cluster_sizes[c] : 0;
        ix->clu_blk.push_back(blk);
        ix->clu_vec.push_back(vec);
        ix->clu_n.push_back((int32_t)n);
        blk += (n + kBlock - 1) / kBlock;
        vec += n;
    }
    ix->clu_set.assign(d->num_cen, 0);
    if (vec != d->num_data) {
        delete ix;
        return fail(SAQGPU_EINVAL, "cluster sizes do not sum to num_data");
    }
    ix->total_blocks = blk;
    int rc = 0;
    rc |= ix->blocks.alloc(std::max<int64_t>(1, blk) * ix->block_stride + kBlockTailPad);
    rc |= ix->long_codes.alloc(std::max<int64_t>(1, vec) * std::max<int64_t>(16, ix->long_stride));
    rc |= ix->exfac.alloc(std::max<int64_t>(1, vec) * ix->nseg * 2);
    rc |= ix->ipcent.alloc(std::max<int64_t>(1, blk) * ix->nseg * kBlock);
    rc |= ix->ids.alloc(std::max<int64_t>(1, vec));
    rc |= ix->crot.alloc((size_t)d->num_cen * ix->D_pad);
    rc |= ix->centroids.alloc((size_t)d->num_cen * d->num_dim);
    rc |= ix->rot.alloc(std::max<int64_t>(1, roff));
    rc |= ix->variance.alloc(ix->D_pad);
    rc |= ix->d_clu_blk.alloc(d->num_cen);
    rc |= ix->d_clu_vec.alloc(d->num_cen);
    rc |= ix->d_clu_n.alloc(d->num_cen);
    rc |= ix->d_rot_off.alloc(ix->nseg);
    rc |= ix->d_segs.alloc(ix->nseg);
    if (rc) {
        saqgpu_index_destroy(ix);
        return fail(SAQGPU_ENOMEM, "device allocation failed");
    }
    hipMemset(ix->blocks.p, 0, ix->blocks.n);
    hipMemset(ix->long_codes.p, 0, ix->long_codes.n);
    hipMemset(ix->exfac.p, 0, ix->exfac.n * 4);
    hipMemset(ix->ipcent.p, 0, ix->ipcent.n * 4);
    hipMemset(ix->ids.p, 0, ix->ids.n * 4);
    hipMemset(ix->crot.p, 0, ix->crot.n * 4);
    std::vector<float> var(ix->D_pad, 0.0f);
    if (d->variance) std::copy(d->variance, d->variance + ix->D_pad, var.begin());
    hipMemcpy(ix->variance.p, var.data(), var.size() * 4, hipMemcpyHostToDevice);
    if (d->centroids)
        hipMemcpy(ix->centroids.p, d->centroids, (size_t)d->num_cen * d->num_dim * 4, hipMemcpyHostToDevice);
    if (d->rotations) hipMemcpy(ix->rot.p, d->rotations, roff * 4, hipMemcpyHostToDevice);
    hipMemcpy(ix->d_clu_blk.p, ix->clu_blk.data(), d->num_cen * 8, hipMemcpyHostToDevice);
    hipMemcpy(ix->d_clu_vec.p, ix->clu_vec.data(), d->num_cen * 8, hipMemcpyHostToDevice);
    hipMemcpy(ix->d_clu_n.p, ix->clu_n.data(), d->num_cen * 4, hipMemcpyHostToDevice);
    hipMemcpy(ix->d_rot_off.p, ix->rot_off.data(), ix->nseg * 8, hipMemcpyHostToDevice);
    if (hipMemcpy(ix->d_segs.p, ix->segs.data(), ix->nseg * sizeof(SegDesc), hipMemcpyHostToDevice) != hipSuccess) {
        saqgpu_index_destroy(ix);
        return fail(SAQGPU_EDEVICE, "upload failed");
    }
    *out = ix;
    return SAQGPU_OK;
}

int saqgpu_index_search_handle(saqgpu_index *src, saqgpu_index **out) {
    if (!src || !out) return fail(SAQGPU_EINVAL, "null argument");
    if (src->parent) return fail(SAQGPU_EINVAL, "a search handle cannot be cloned: clone its index");
    std::lock_guard<std::mutex> lk(src->mu);  // the index's state holds still while it is copied
    if (hipSetDevice(src->device) != hipSuccess) return fail(SAQGPU_EDEVICE, "hipSetDevice failed");
    auto *h = new saqgpu_index;
    h->parent = src;
    h->device = src->device;
    h->num_data = src->num_data;
    h->num_dim = src->num_dim;
    h->num_cen = src->num_cen;
    h->nseg = src->nseg;
    h->D_pad = src->D_pad;
    h->ngroups = src->ngroups;
    h->max_dim = src->max_dim;
    h->compact = src->compact;
    h->has_rot = src->has_rot;
    h->seg_dim = src->seg_dim;
    h->seg_bits = src->seg_bits;
    h->segs = src->segs;
    h->clu_blk = src->clu_blk;
    h->clu_vec = src->clu_vec;
    h->rot_off = src->rot_off;
    h->clu_n = src->clu_n;
    h->clu_set = src->clu_set;
    h->total_blocks = src->total_blocks;
    h->block_stride = src->block_stride;
    h->long_stride = src->long_stride;
    h->blocks.borrow(src->blocks);
    h->long_codes.borrow(src->long_codes);
    h->exfac.borrow(src->exfac);
    h->ipcent.borrow(src->ipcent);
    h->crot.borrow(src->crot);
    h->centroids.borrow(src->centroids);
    h->rot.borrow(src->rot);
    h->variance.borrow(src->variance);
    h->ids.borrow(src->ids);
    h->d_clu_blk.borrow(src->d_clu_blk);
    h->d_clu_vec.borrow(src->d_clu_vec);
    h->d_rot_off.borrow(src->d_rot_off);
    h->d_clu_n.borrow(src->d_clu_n);
    h->d_segs.borrow(src->d_segs);
    if (src->bq_ready) {  // the batched layout, where the index has built it
        h->bq_ready = true;
        h->bq_tile = src->bq_tile;
        h->bq_nqs = src->bq_nqs;
        h->bq_kq = src->bq_kq;
        h->bq_fac_off = src->bq_fac_off;
        h->bq_stride = src->bq_stride;
        h->bq_total_mb = src->bq_total_mb;
        h->bsegs = src->bsegs;
        h->clu_mb = src->clu_mb;
        h->bq_mb.borrow(src->bq_mb);
        h->d_clu_mb.borrow(src->d_clu_mb);
        h->d_bq_order.borrow(src->d_bq_order);
        h->d_bsegs.borrow(src->d_bsegs);
    }
    if (hipEventCreateWithFlags(&h->ws_free, hipEventDisableTiming) != hipSuccess) {
        h->parent = nullptr;
        saqgpu_index_destroy(h);
        return fail(SAQGPU_EDEVICE, "event creation failed");
    }
    src->handles++;
    *out = h;
    return SAQGPU_OK;
}

void saqgpu_index_destroy(saqgpu_index *ix) {
    if (!ix) return;
    if (ix->parent) ix->parent->handles--;
    hipSetDevice(ix->device);
    ix->blocks.release();
    ix->long_codes.release();
    ix->exfac.release();
    ix->ipcent.release();
    ix->crot.release();
    ix->centroids.release();
    ix->rot.release();
    ix->variance.release();
    ix->ids.release();
    ix->d_clu_blk.release();
    ix->d_clu_vec.release();
    ix->d_rot_off.release();
    ix->d_clu_n.release();
    ix->d_segs.release();
    ix->ws_q.release();
    ix->ws_qrot.release();
    ix->ws_bound.release();
    ix->ws_dists.release();
    ix->ws_cdist.release();
    ix->ws_probes.release();
    ix->ws_ids.release();
    ix->ws_met.release();
    ix->ws_ties.release();
    ix->ws_eoffs.release();
    ix->ws_epieces.release();
    ix->ws_eids.release();
    ix->ws_eout.release();
    ix->bq_mb.release();
    ix->d_clu_mb.release();
    ix->d_bq_order.release();
    ix->d_bsegs.release();
    ix->bq_ws.release();
    ix->bq_qq.release();
    ix->bq_pc.release();
    ix->cnorm.release();
    ix->cen_bf3.release();
    ix->ws_qbf3.release();
    ix->bq_npq.release();
    ix->bq_cand_d.release();
    ix->bq_cand_id.release();
    ix->sl_pool_d.release();
    ix->sl_pool_id.release();
    ix->sl_pool_size.release();
    ix->sl_met.release();
    ix->sl_pair.release();
    ix->sl_binfo.release();
    ix->sl_cand.release();
    ix->sl_pre.release();
    ix->sl_desc.release();
    ix->sl_ctr.release();
    for (auto &g : ix->sl_groups) {
        (void)hipStreamDestroy(g.stream);
        (void)hipEventDestroy(g.done);
    }
    if (ix->sl_fork) (void)hipEventDestroy(ix->sl_fork);
    if (ix->blas) (void)rocblas_destroy_handle(ix->blas);
    if (ix->bq_graph) (void)hipGraphExecDestroy(ix->bq_graph);
    if (ix->ws_free) (void)hipEventDestroy(ix->ws_free);
    delete ix;
}

int saqgpu_timing(saqgpu_index *ix, int32_t enable, double *scan_ms, int64_t *launches) {
    if (!ix) return fail(SAQGPU_EINVAL, "null index");
    std::lock_guard<std::mutex> lk(ix->mu);
    HIPCHK(hipSetDevice(ix->device));
    double tot = 0;
    for (auto &pr : ix->events) {
        float ms = 0;
        HIPCHK(hipEventSynchronize(pr.second));
        HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
        tot += ms;
    }
    if (scan_ms) *scan_ms = tot;
    if (launches) *launches = (int64_t)ix->events.size();
    if (enable != (int)ix->timing || enable) {
        for (auto &pr : ix->events) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
        ix->events.clear();
    }
    ix->timing = enable != 0;
    return SAQGPU_OK;
}

int64_t saqgpu_index_device_bytes(const saqgpu_index *ix) {
    if (!ix) return 0;
    return (int64_t)(ix->blocks.n + ix->long_codes.n + 4 * (ix->exfac.n + ix->ipcent.n + ix->ids.n + ix->crot.n +
                                                            ix->centroids.n + ix->rot.n + ix->variance.n));
}

int saqgpu_cluster_sizes_of(const saqgpu_index *ix, int32_t cid, saqgpu_cluster_sizes *out) {
    if (!ix || !out || cid < 0 || cid >= ix->num_cen) return fail(SAQGPU_EINVAL, "bad cluster id");
    const int64_t n = ix->clu_n[cid];
    auto L = ix->ref_layout(n);
    out->num_blocks = L.nb;
    out->short_factor_floats = L.single ? 0 : L.fac_per_block * L.nb;
    out->short_code_bytes = L.code_per_block * L.nb;
    out->long_code_bytes = L.long_total;
    out->long_factor_floats = n * ix->nseg * 2;
    out->centroid_floats = ix->D_pad;
    return SAQGPU_OK;
}

int saqgpu_index_set_cluster(saqgpu_index *ix, int32_t cid, const float *short_factors, const uint8_t *short_code,
                             const uint8_t *long_code, const float *long_factors, const uint32_t *ids,
                             const float *centroid_rot) {
    if (!ix || cid < 0 || cid >= ix->num_cen) return fail(SAQGPU_EINVAL, "bad cluster id");
    if (int rc_ = check_mutable(ix)) return rc_;
    IndexCall call(const_cast<saqgpu_index *>(ix), nullptr);
    if (int rc_ = call.begin()) return rc_;
    const int64_t n = ix->clu_n[cid];
    const auto L = ix->ref_layout(n);
    const int S = ix->nseg;
    if (n > 0 && (!short_code || !long_factors || !ids || (!L.single && !short_factors)))
        return fail(SAQGPU_EINVAL, "missing cluster arrays");
    // block records
    std::vector<uint8_t> rec(L.nb * ix->block_stride, 0);
    std::vector<float> ipc(L.nb * S * kBlock, 0.0f);
    for (int64_t b = 0; b < L.nb; ++b) {
        uint8_t *r = &rec[b * ix->block_stride];
        float *fac = reinterpret_cast<float *>(r);
        for (int s = 0; s < S; ++s) {
            const float *src = L.single ? reinterpret_cast<const float *>(short_code + b * L.code_per_block)
                                        : short_factors + b * L.fac_per_block + (int64_t)s * 2 * kBlock;
            std::memcpy(fac + s * kBlock, src, kBlock * 4);
            std::memcpy(&ipc[(b * S + s) * kBlock], src + kBlock, kBlock * 4);
        }
        uint8_t *codes = r + (int64_t)S * kBlock * 4;
        for (int s = 0; s < S; ++s) {
            const SegDesc &sd = ix->segs[s];
            if (!sd.bits) continue;
            const uint8_t *blk = short_code + b * L.code_per_block + L.code_off[s];
            const int bytes = sd.dim / 8;
            for (int v = 0; v < kBlock; ++v) {
                // nibble of codebook m for vector v (pack_codes, fastscan.hpp:75-112)
                int jpos = 0;
                while (kPermLo[jpos] != (v & 15)) ++jpos;
                for (int i = 0; i < bytes; ++i) {
                    const uint8_t b0 = blk[16 * (2 * i) + jpos], b1 = blk[16 * (2 * i + 1) + jpos];
                    const int n0 = v < 16 ? (b0 & 15) : (b0 >> 4);
                    const int n1 = v < 16 ? (b1 & 15) : (b1 >> 4);
                    const int g = sd.group_off + i / 8;
                    codes[(int64_t)g * 256 + v * 8 + (i % 8)] = (uint8_t)((n0 << 4) | n1);
                }
            }
        }
    }
    // long codes into per-vector records
    std::vector<uint8_t> lc(std::max<int64_t>(1, n) * ix->long_stride, 0);
    for (int64_t v = 0; v < n; ++v)
        for (int s = 0; s < S; ++s) {
            const int64_t lb = L.long_bytes[s];
            if (!lb) continue;
            const uint8_t *src = L.compact ? long_code + L.long_off[s] + v * lb
                                           : long_code + L.long_off[s] + v * L.long_stride_ref;
            std::memcpy(&lc[v * ix->long_stride + ix->segs[s].long_off], src, lb);
        }
    const int64_t b0 = ix->clu_blk[cid], v0 = ix->clu_vec[cid];
    if (L.nb) {
        HIPCHK(hipMemcpy(ix->blocks.p + b0 * ix->block_stride, rec.data(), rec.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ix->ipcent.p + b0 * S * kBlock, ipc.data(), ipc.size() * 4, hipMemcpyHostToDevice));
    }
    if (n) {
        HIPCHK(hipMemcpy(ix->long_codes.p + v0 * ix->long_stride, lc.data(), n * ix->long_stride,
                         hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ix->exfac.p + v0 * S * 2, long_factors, n * S * 2 * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ix->ids.p + v0, ids, n * 4, hipMemcpyHostToDevice));
    }
    if (centroid_rot)
        HIPCHK(hipMemcpy(ix->crot.p + (int64_t)cid * ix->D_pad, centroid_rot, ix->D_pad * 4, hipMemcpyHostToDevice));
    ix->clu_set[cid] = 1;
    ix->bq_ready = false;
    return SAQGPU_OK;
}

int saqgpu_index_get_cluster(const saqgpu_index *ix, int32_t cid, float *short_factors, uint8_t *short_code,
                             uint8_t *long_code, float *long_factors, uint32_t *ids, float *centroid_rot) {
    if (!ix || cid < 0 || cid >= ix->num_cen) return fail(SAQGPU_EINVAL, "bad cluster id");
    IndexCall call(const_cast<saqgpu_index *>(ix), nullptr);
    if (int rc_ = call.begin()) return rc_;
    const int64_t n = ix->clu_n[cid];
    const auto L = ix->ref_layout(n);
    const int S = ix->nseg;
    const int64_t b0 = ix->clu_blk[cid], v0 = ix->clu_vec[cid];
    std::vector<uint8_t> rec(std::max<int64_t>(1, L.nb) * ix->block_stride);
    std::vector<float> ipc(std::max<int64_t>(1, L.nb) * S * kBlock);
    std::vector<uint8_t> lc(std::max<int64_t>(1, n) * ix->long_stride);
    if (L.nb) {
        HIPCHK(hipMemcpy(rec.data(), ix->blocks.p + b0 * ix->block_stride, L.nb * ix->block_stride,
                         hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ipc.data(), ix->ipcent.p + b0 * S * kBlock, L.nb * S * kBlock * 4, hipMemcpyDeviceToHost));
    }
    if (n) {
        HIPCHK(hipMemcpy(lc.data(), ix->long_codes.p + v0 * ix->long_stride, n * ix->long_stride,
                         hipMemcpyDeviceToHost));
        if (long_factors)
            HIPCHK(hipMemcpy(long_factors, ix->exfac.p + v0 * S * 2, n * S * 2 * 4, hipMemcpyDeviceToHost));
        if (ids) HIPCHK(hipMemcpy(ids, ix->ids.p + v0, n * 4, hipMemcpyDeviceToHost));
    }
    if (centroid_rot)
        HIPCHK(hipMemcpy(centroid_rot, ix->crot.p + (int64_t)cid * ix->D_pad, ix->D_pad * 4, hipMemcpyDeviceToHost));
    if (short_code) {
        if (!L.single && !short_factors) return fail(SAQGPU_EINVAL, "short_factors required");
        for (int64_t b = 0; b < L.nb; ++b) {
            const uint8_t *r = &rec[b * ix->block_stride];
            const float *fac = reinterpret_cast<const float *>(r);
            for (int s = 0; s < S; ++s) {
                float *dst = L.single ? reinterpret_cast<float *>(short_code + b * L.code_per_block)
                                      : short_factors + b * L.fac_per_block + (int64_t)s * 2 * kBlock;
                std::memcpy(dst, fac + s * kBlock, kBlock * 4);
                std::memcpy(dst + kBlock, &ipc[(b * S + s) * kBlock], kBlock * 4);
            }
            const uint8_t *codes = r + (int64_t)S * kBlock * 4;
            for (int s = 0; s < S; ++s) {
                const SegDesc &sd = ix->segs[s];
                if (!sd.bits) continue;
                uint8_t *blk = short_code + b * L.code_per_block + L.code_off[s];
                const int bytes = sd.dim / 8;
                for (int i = 0; i < bytes; ++i) {
                    const int g = sd.group_off + i / 8;
                    for (int j = 0; j < 16; ++j) {
                        const int v = kPermLo[j];
                        const uint8_t lo = codes[(int64_t)g * 256 + v * 8 + (i % 8)];
                        const uint8_t hi = codes[(int64_t)g * 256 + (v + 16) * 8 + (i % 8)];
                        blk[16 * (2 * i) + j] = (uint8_t)((lo >> 4) | ((hi >> 4) << 4));
                        blk[16 * (2 * i + 1) + j] = (uint8_t)((lo & 15) | ((hi & 15) << 4));
                    }
                }
            }
        }
    }
    if (long_code) {
        std::memset(long_code, 0, L.long_total);
        for (int64_t v = 0; v < n; ++v)
            for (int s = 0; s < S; ++s) {
                const int64_t lb = L.long_bytes[s];
                if (!lb) continue;
                uint8_t *dst = L.compact ? long_code + L.long_off[s] + v * lb
                                         : long_code + L.long_off[s] + v * L.long_stride_ref;
                std::memcpy(dst, &lc[v * ix->long_stride + ix->segs[s].long_off], lb);
            }
    }
    return SAQGPU_OK;
}

int saqgpu_probe(saqgpu_index *ix, const float *queries, int64_t nq, int32_t nprobe, int32_t dist_type,
                 uint32_t *probes, int32_t flags, void *stream) {
    if (!ix || !queries || !probes || nq < 0) return fail(SAQGPU_EINVAL, "null argument");
    if (nprobe < 1 || nprobe > ix->num_cen) return fail(SAQGPU_EINVAL, "nprobe must be in 1..K");
    if (dist_type != 0 && dist_type != 1) return fail(SAQGPU_EINVAL, "dist_type must be 0 (L2) or 1 (IP)");
    if (nq == 0) return SAQGPU_OK;
    IndexCall call(ix, (hipStream_t)stream);
    if (int rc_ = call.begin()) return rc_;
    hipStream_t st = (hipStream_t)stream;
    const bool dev = flags & SAQGPU_DEVICE_PTRS;
    const float *dq = queries;
    uint32_t *dp = probes;
    if (!dev) {
        if (ix->ws_q.ensure(nq * ix->num_dim) || ix->ws_probes.ensure(nq * nprobe))
            return fail(SAQGPU_ENOMEM, "workspace allocation failed");
        HIPCHK(hipMemcpyAsync(ix->ws_q.p, queries, nq * ix->num_dim * 4, hipMemcpyHostToDevice, st));
        dq = ix->ws_q.p;
        dp = ix->ws_probes.p;
    }
    if (ix->ws_cdist.ensure((size_t)nq * ix->num_cen)) return fail(SAQGPU_ENOMEM, "workspace allocation failed");
    const char *form = getenv("SAQGPU_PROBE");
    if (form && !strcmp(form, "bf16x3")) {  // the batched scan's coarse quantiser (not exact), for tests of its select
        if (nprobe > 4096) return fail(SAQGPU_EUNSUPPORTED, "the bf16x3 coarse quantiser selects at most 4096 lists");
        if (ix->ws_qbf3.ensure((size_t)nq * 3 * ix->num_dim)) return fail(SAQGPU_ENOMEM, "workspace allocation failed");
        if (int rc = probe_gemm(ix, dq, nq, nprobe, dist_type, dp, st)) return rc;
    } else if (int rc = exact_probe(ix, dq, nq, nprobe, dist_type, dp, st)) {
        return rc;
    }
    if (!dev) {
        HIPCHK(hipMemcpyAsync(probes, dp, nq * nprobe * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    return SAQGPU_OK;
}

int saqgpu_probe_noalloc(saqgpu_index *ix, const float *queries, int64_t nq, int32_t nprobe, int32_t dist_type,
                         uint32_t *probes, void *stream) {
    if (!ix || !queries || !probes || nq < 0) return fail(SAQGPU_EINVAL, "null argument");
    if (nprobe < 1 || nprobe > ix->num_cen) return fail(SAQGPU_EINVAL, "nprobe must be in 1..K");
    if (nprobe > 4096) return fail(SAQGPU_EUNSUPPORTED, "nprobe above 4096 is not supported");
    if (dist_type != 0 && dist_type != 1) return fail(SAQGPU_EINVAL, "dist_type must be 0 (L2) or 1 (IP)");
    if (nq > SAQGPU_NOALLOC_SCRATCH_FLOATS / (ix->num_cen + 1))
        return fail(SAQGPU_EINVAL, "nq x (num_cen + 1) exceeds the fixed scratch of SAQGPU_NOALLOC_SCRATCH_FLOATS "
                                   "floats: split the queries");
    if (nq == 0) return SAQGPU_OK;
    hipStream_t st = (hipStream_t)stream;
    IndexCall call(ix, st);
    if (int rc = call.begin()) return rc;
    float *ws = scratch_of(ix->device);
    if (!ws) return fail(SAQGPU_ESTATE, "device scratch not set up");
    // the scratch holds nq x num_cen distances, then nq tie flags; every one is written before it is
    // read, so nothing in it needs clearing
    HIPCHK(launch_probe(ix->centroids.p, ix->num_cen, ix->num_dim, queries, nq, nprobe, dist_type, ws,
                        reinterpret_cast<int32_t *>(ws + nq * ix->num_cen), probes, st));
    return SAQGPU_OK;
}

namespace {
// The plain scan as slices of the probe sequence (slice_scan.hip): lists [0,1), [1,4),
// [4,16), [16,64), ... of every query (below 64 items the same count of slices, spaced evenly on a
// log scale); per slice one list-scan launch over all (query, list)
// pairs and one replay launch.  Block-info and candidate space that a slice runs out of is
// flagged on the device; the search is then repeated with the space it asked for (the
// capacities are kept for later searches), so results never depend on the capacities.
// e0 / e1 (optional): events recorded around the kernels of the attempt that succeeds.
//
// The queries are split into up to kGroups contiguous groups (200 queries or more each), each
// with its own share of the workspace; group 0 runs on st, the others on streams of their own
// forked from and joined back into st (four streams in all, the hardware queues a process
// gets by default).  A group's replay, one latency-bound wavefront per query, then overlaps
// the other groups' list scans and refines.  SAQGPU_GROUPS=n caps the group count.
int sliced_scan(saqgpu_index *ix, const IndexView &v, const uint32_t *dp, const saqgpu_search_params *p,
                const int32_t *dnpq, int64_t nq, uint32_t *dids, float *dd, uint64_t *dm, hipStream_t st,
                hipEvent_t e0, hipEvent_t e1) {
    constexpr int kGroups = 4;
    const int K = p->topk, S = ix->nseg;
    // slices of the work items (item 0: list 0's first chunk; 1: its rest; j >= 2: list j - 1)
    const int items = p->nprobe + 1;
    std::vector<int> bounds{0};
    for (int b = 1; b < items; b *= 4) bounds.push_back(b);
    const size_t nb = bounds.size();
    if (nb > 2 && (items <= 64 || items - bounds[nb - 1] < (bounds[nb - 1] - bounds[nb - 2]) / 2)) {
        // few slices, or a short last one: the same number of slices, their ends spaced evenly on a
        // log scale from 1 to items (17 items: [0,1) [1,3) [3,7) [7,17) rather than [0,1) [1,4)
        // [4,16) [16,17)), so no slice is a short tail and the candidate superset of the largest
        // one shrinks
        const int ns = (int)bounds.size();  // slices
        const double r = std::pow((double)items, 1.0 / (ns - 1));
        bounds.assign(1, 0);
        double x = 1.0;
        for (int i = 1; i < ns; ++i, x *= r) {
            const int b = std::max(bounds.back() + 1, (int)std::lround(x));
            if (b >= items) break;
            bounds.push_back(b);
        }
    }
    bounds.push_back(items);
    int Lmax = 1;
    for (size_t i = 0; i + 1 < bounds.size(); ++i) Lmax = std::max(Lmax, bounds[i + 1] - bounds[i]);
    int max_smem = 0, ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&max_smem, hipDeviceAttributeMaxSharedMemoryPerBlock, ix->device));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ix->device));
    // refine wavefronts stage up to cmax long records each (4 wavefronts within 64 KB)
    const size_t ls_smem = list_scan_smem_bytes(v);
    int cmax = 32;
    while (cmax > 4 && refine_smem_bytes(v, cmax) > 64 * 1024) cmax /= 2;
    const size_t rf_smem = refine_smem_bytes(v, cmax);
    if ((int64_t)std::max(ls_smem, rf_smem) > max_smem)
        return fail(SAQGPU_EUNSUPPORTED, "plan needs more LDS than a workgroup has");
    int G = (int)std::max<int64_t>(1, std::min<int64_t>(kGroups, nq / 200));
    if (const char *ge = getenv("SAQGPU_GROUPS")) G = std::max(1, std::min(G, atoi(ge)));
    // the refine kernel is persistent (workgroups loop over the descriptors); with four groups
    // in flight each one's refine takes half the chip's slots
    const int rf_grid = std::max(1, ncu) * (G >= 4 ? 2 : 4);
    // group 0 runs on st itself, groups 1.. on streams of their own
    while ((int)ix->sl_groups.size() < G - 1) {
        saqgpu_index::SliceGroup g{};
        HIPCHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
        ix->sl_groups.push_back(g);
    }
    if (!ix->sl_fork) HIPCHK(hipEventCreateWithFlags(&ix->sl_fork, hipEventDisableTiming));
    int rc = 0;
    rc |= ix->sl_pool_d.ensure(nq * K);
    rc |= ix->sl_pool_id.ensure(nq * K);
    rc |= ix->sl_pool_size.ensure(nq);
    rc |= ix->sl_met.ensure(nq * 3);
    rc |= ix->sl_pair.ensure(nq * Lmax);
    rc |= ix->sl_ctr.ensure((size_t)kGroups * CTR_N);
    if (rc) return fail(SAQGPU_ENOMEM, "workspace allocation failed");
    // first sizes: every list of the widest slice, a quarter of its vectors as candidates
    int64_t nonempty = 0;
    for (int c = 0; c < ix->num_cen; ++c) nonempty += ix->clu_n[c] > 0;
    const double mean_n = nonempty ? (double)ix->num_data / nonempty : 1.0;
    const int64_t kCapMax = ((int64_t)1 << 31) - 1;
    const int64_t unit = (int64_t)kSlots * kGroups;  // buffers split into kGroups x kSlots shares
    auto clamp_cap = [&](double x) { return (int64_t)std::min<double>((double)kCapMax, std::max(x, 64.0 * unit)); };
    // SAQGPU_SLICE_CAP=n (tests): start every capacity at n instead
    const char *cap_env = getenv("SAQGPU_SLICE_CAP");
    const int64_t cap0 = cap_env ? std::max<int64_t>(1, atoll(cap_env)) : 0;
    const int bst = binfo_stride(S), rst = cand_stride(S), pst = pre_stride(S), dst = desc_stride(S);
    if (ix->sl_binfo.n == 0 &&
        ix->sl_binfo.alloc(cap0 ? cap0 : clamp_cap((double)nq * Lmax * (mean_n / kBlock + 1.5) * bst)))
        return fail(SAQGPU_ENOMEM, "workspace allocation failed");
    if (ix->sl_cand.n == 0) {
        const int64_t nc = cap0 ? cap0 : clamp_cap((double)nq * Lmax * mean_n / 4 + 64.0 * nq);
        if (ix->sl_cand.alloc(nc * rst) || ix->sl_pre.alloc(nc * pst))
            return fail(SAQGPU_ENOMEM, "workspace allocation failed");
    }
    if (ix->sl_desc.n == 0 &&
        ix->sl_desc.alloc((cap0 ? cap0 : clamp_cap((double)nq * Lmax * (mean_n / 64 + 1))) * dst))
        return fail(SAQGPU_ENOMEM, "workspace allocation failed");
    for (int attempt = 0;; ++attempt) {
        // per group and slot shares of each buffer (every slot's reservations are independent)
        const int64_t bcap = std::min<int64_t>(kCapMax, (int64_t)ix->sl_binfo.n) / unit;
        const int64_t ccap = std::min<int64_t>(kCapMax, (int64_t)(ix->sl_cand.n / rst)) / unit;
        const int64_t dcap = std::min<int64_t>(kCapMax, (int64_t)(ix->sl_desc.n / dst)) / unit;
        std::vector<SliceArgs> ga(G);
        for (int g = 0; g < G; ++g) {
            const int64_t q0 = nq * g / G, q1 = nq * (g + 1) / G;
            SliceArgs &a = ga[g];
            a = SliceArgs{};
            a.ix = v;
            a.qrot = ix->ws_qrot.p + q0 * ix->D_pad;
            a.bound = ix->ws_bound.p + q0 * S;
            a.probes = dp + q0 * p->nprobe;
            a.nprobe = p->nprobe;
            a.npq = dnpq ? dnpq + q0 : nullptr;
            a.topk = K;
            a.nq = q1 - q0;
            a.pool_d = ix->sl_pool_d.p + q0 * K;
            a.pool_id = ix->sl_pool_id.p + q0 * K;
            a.pool_size = ix->sl_pool_size.p + q0;
            a.met = ix->sl_met.p + q0 * 3;
            a.pair_blk = ix->sl_pair.p + q0 * Lmax;
            a.binfo = ix->sl_binfo.p + g * bcap * kSlots;
            a.cand = ix->sl_cand.p + g * ccap * kSlots * rst;
            a.pre = ix->sl_pre.p + g * ccap * kSlots * pst;
            a.desc = ix->sl_desc.p + g * dcap * kSlots * dst;
            a.binfo_cap = (uint32_t)bcap;
            a.cand_cap = (uint32_t)ccap;
            a.desc_cap = (uint32_t)dcap;
            a.ctr = ix->sl_ctr.p + g * CTR_N;
            a.cmax = cmax;
            a.out_ids = dids + q0 * K;
            a.out_dists = dd ? dd + q0 * K : nullptr;
            a.out_metrics = dm ? dm + q0 * 3 : nullptr;
        }
        if (e0) HIPCHK(hipEventRecord(e0, st));
        auto gstream = [&](int g) { return g == 0 ? st : ix->sl_groups[g - 1].stream; };
        if (G > 1) HIPCHK(hipEventRecord(ix->sl_fork, st));
        for (int g = 0; g < G; ++g) {
            if (g > 0) HIPCHK(hipStreamWaitEvent(gstream(g), ix->sl_fork, 0));
            HIPCHK(launch_slice_reset(ga[g].ctr, 1, gstream(g)));
        }
        for (size_t i = 0; i + 1 < bounds.size(); ++i) {
            for (int g = 0; g < G; ++g) {
                SliceArgs &a = ga[g];
                const hipStream_t gs = gstream(g);
                a.p0 = bounds[i];
                a.p1 = bounds[i + 1];
                a.first = i == 0;
                a.last = i + 2 == bounds.size();
                if (i > 0) HIPCHK(launch_slice_reset(a.ctr, 0, gs));
                HIPCHK(launch_list_scan(a, p->dist_type, ls_smem, gs));
                HIPCHK(launch_refine(a, p->dist_type, rf_smem, rf_grid, gs));
                HIPCHK(launch_replay(a, p->dist_type, gs));
            }
        }
        for (int g = 0; g < G; ++g) {
            HIPCHK(launch_slice_reset(ga[g].ctr, 0, gstream(g)));
            if (g == 0) continue;
            HIPCHK(hipEventRecord(ix->sl_groups[g - 1].done, gstream(g)));
            HIPCHK(hipStreamWaitEvent(st, ix->sl_groups[g - 1].done, 0));
        }
        if (e1) HIPCHK(hipEventRecord(e1, st));
        std::vector<uint32_t> h((size_t)G * CTR_N);
        HIPCHK(hipMemcpyAsync(h.data(), ix->sl_ctr.p, h.size() * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        uint32_t over = 0, mb = 0, mc = 0, md = 0;
        for (int g = 0; g < G; ++g) {
            over |= h[g * CTR_N + CTR_OVERFLOW];
            mb = std::max(mb, h[g * CTR_N + CTR_MAX_BINFO]);
            mc = std::max(mc, h[g * CTR_N + CTR_MAX_CAND]);
            md = std::max(md, h[g * CTR_N + CTR_MAX_DESC]);
        }
        if (!over) return SAQGPU_OK;
        if (attempt >= 3) return fail(SAQGPU_ENOMEM, "scan workspace did not converge");
        // grow to what the slices asked for: every share gets the largest share's need (a slice
        // that ran short only raises later T, so the counts are upper bounds once block info fits)
        const int64_t nb = clamp_cap(std::max<double>((double)mb * unit * 1.25, (double)bcap * unit));
        const int64_t nc = clamp_cap(std::max<double>((double)mc * unit * 1.25, (double)ccap * unit));
        const int64_t nd = clamp_cap(std::max<double>((double)md * unit * 1.25, (double)dcap * unit));
        if (nb > bcap * unit && ix->sl_binfo.alloc(nb)) return fail(SAQGPU_ENOMEM, "workspace allocation failed");
        if (nc > ccap * unit && (ix->sl_cand.alloc(nc * rst) || ix->sl_pre.alloc(nc * pst)))
            return fail(SAQGPU_ENOMEM, "workspace allocation failed");
        if (nd > dcap * unit && ix->sl_desc.alloc(nd * dst)) return fail(SAQGPU_ENOMEM, "workspace allocation failed");
    }
}
}  // namespace

int saqgpu_search(saqgpu_index *ix, const float *queries, int64_t nq, const uint32_t *probes,
                  const saqgpu_search_params *p, uint32_t *out_ids, float *out_dists, saqgpu_metrics *metrics,
                  int32_t flags, void *stream) {
    if (!ix || !queries || !p || !out_ids || nq < 0) return fail(SAQGPU_EINVAL, "null argument");
    if (p->topk < 1 || p->topk > SAQGPU_MAX_TOPK) return fail(SAQGPU_EINVAL, "topk must be in 1..1024");
    if (p->nprobe < 1 || p->nprobe > ix->num_cen) return fail(SAQGPU_EINVAL, "nprobe must be in 1..K");
    if (p->dist_type != 0 && p->dist_type != 1) return fail(SAQGPU_EINVAL, "dist_type must be 0 (L2) or 1 (IP)");
    for (int c = 0; c < ix->num_cen; ++c)
        if (!ix->clu_set[c] && ix->clu_n[c] > 0) return fail(SAQGPU_ESTATE, "index has clusters not yet set");
    if (nq == 0) return SAQGPU_OK;
    const bool dev = flags & SAQGPU_DEVICE_PTRS;
    if (probes && !dev)
        for (int64_t i = 0; i < nq * p->nprobe; ++i)
            if (probes[i] >= (uint32_t)ix->num_cen) return fail(SAQGPU_EINVAL, "probe id out of range");
    if (p->nprobe_per_query && !dev)
        for (int64_t i = 0; i < nq; ++i)
            if (p->nprobe_per_query[i] < 1 || p->nprobe_per_query[i] > p->nprobe)
                return fail(SAQGPU_EINVAL, "nprobe_per_query entries must be in 1..nprobe");
    IndexCall call(ix, (hipStream_t)stream);
    if (int rc_ = call.begin()) return rc_;
    hipStream_t st = (hipStream_t)stream;
    const IndexView v = ix->view();
    int rc = 0;
    rc |= ix->ws_qrot.ensure(nq * ix->D_pad);
    rc |= ix->ws_bound.ensure(nq * ix->nseg);
    if (!dev) {
        rc |= ix->ws_q.ensure(nq * ix->num_dim);
        rc |= ix->ws_ids.ensure(nq * p->topk);
        rc |= ix->ws_dists.ensure(nq * p->topk);
        rc |= ix->ws_met.ensure(nq * 3);
        if (p->nprobe_per_query) rc |= ix->bq_npq.ensure(nq);
    }
    if (!probes || !dev) rc |= ix->ws_probes.ensure(nq * p->nprobe);
    if (rc) return fail(SAQGPU_ENOMEM, "workspace allocation failed");
    const float *dq = queries;
    const uint32_t *dp = probes;
    const int32_t *dnpq = p->nprobe_per_query;
    uint32_t *dids = out_ids;
    float *dd = out_dists;
    uint64_t *dm = reinterpret_cast<uint64_t *>(metrics);
    if (!dev) {
        HIPCHK(hipMemcpyAsync(ix->ws_q.p, queries, nq * ix->num_dim * 4, hipMemcpyHostToDevice, st));
        dq = ix->ws_q.p;
        if (probes) {
            HIPCHK(hipMemcpyAsync(ix->ws_probes.p, probes, nq * p->nprobe * 4, hipMemcpyHostToDevice, st));
            dp = ix->ws_probes.p;
        }
        dids = ix->ws_ids.p;
        dd = out_dists ? ix->ws_dists.p : nullptr;
        dm = metrics ? ix->ws_met.p : nullptr;
        if (dnpq) {
            HIPCHK(hipMemcpyAsync(ix->bq_npq.p, p->nprobe_per_query, nq * 4, hipMemcpyHostToDevice, st));
            dnpq = ix->bq_npq.p;
        }
    }
    if (!probes) {
        if (ix->ws_cdist.ensure((size_t)nq * ix->num_cen)) return fail(SAQGPU_ENOMEM, "workspace allocation failed");
        int rc;
        if ((rc = exact_probe(ix, dq, nq, p->nprobe, p->dist_type, ix->ws_probes.p, st))) return rc;
        dp = ix->ws_probes.p;
    }
    HIPCHK(launch_query_prep(v, dq, nq, ix->variance.p, ix->has_rot ? ix->rot.p : nullptr, ix->d_rot_off.p,
                             p->searcher_vars_bound_m, ix->ws_qrot.p, ix->ws_bound.p, ix->max_dim, st));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ix->timing) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
    }
    const int rc2 = sliced_scan(ix, v, dp, p, dnpq, nq, dids, dd, dm, st, e0, e1);
    if (rc2 != SAQGPU_OK) return rc2;
    if (ix->timing) ix->events.emplace_back(e0, e1);
    if (!dev) {
        HIPCHK(hipMemcpyAsync(out_ids, dids, nq * p->topk * 4, hipMemcpyDeviceToHost, st));
        if (out_dists) HIPCHK(hipMemcpyAsync(out_dists, dd, nq * p->topk * 4, hipMemcpyDeviceToHost, st));
        if (metrics) HIPCHK(hipMemcpyAsync(metrics, dm, nq * 3 * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    return SAQGPU_OK;
}

int saqgpu_estimate(saqgpu_index *ix, const float *query, const uint32_t *probes, int32_t nprobe, float vars_bound_m,
                    int32_t dist_type, uint32_t *out_ids, float *out_acc, float *out_fast, float *out_vars,
                    int64_t *n_out) {
    if (!ix || !query || !out_ids || !out_acc) return fail(SAQGPU_EINVAL, "null argument");
    if (nprobe < 1 || nprobe > ix->num_cen) return fail(SAQGPU_EINVAL, "nprobe must be in 1..K");
    if (dist_type != 0 && dist_type != 1) return fail(SAQGPU_EINVAL, "dist_type must be 0 (L2) or 1 (IP)");
    for (int c = 0; c < ix->num_cen; ++c)
        if (!ix->clu_set[c] && ix->clu_n[c] > 0) return fail(SAQGPU_ESTATE, "index has clusters not yet set");
    if (probes)
        for (int i = 0; i < nprobe; ++i)
            if (probes[i] >= (uint32_t)ix->num_cen) return fail(SAQGPU_EINVAL, "probe id out of range");
    IndexCall call(ix, nullptr);
    if (int rc_ = call.begin()) return rc_;
    const size_t smem = estimate_smem_bytes(ix->view());
    int max_smem = 0;
    HIPCHK(hipDeviceGetAttribute(&max_smem, hipDeviceAttributeMaxSharedMemoryPerBlock, ix->device));
    if ((int64_t)smem > max_smem) return fail(SAQGPU_EUNSUPPORTED, "plan needs more LDS than a workgroup has");
    int rc = 0;
    rc |= ix->ws_q.ensure(ix->num_dim);
    rc |= ix->ws_qrot.ensure(ix->D_pad);
    rc |= ix->ws_bound.ensure(ix->nseg);
    rc |= ix->ws_probes.ensure(nprobe);
    rc |= ix->ws_cdist.ensure(ix->num_cen);
    if (rc) return fail(SAQGPU_ENOMEM, "workspace allocation failed");
    HIPCHK(hipMemcpy(ix->ws_q.p, query, ix->num_dim * 4, hipMemcpyHostToDevice));
    std::vector<uint32_t> pr(nprobe);
    if (probes) {
        std::copy(probes, probes + nprobe, pr.begin());
        HIPCHK(hipMemcpy(ix->ws_probes.p, probes, nprobe * 4, hipMemcpyHostToDevice));
    } else {  // FlatInitializer::centroids_distances, as IVF::estimate does (ivf.hpp:321-322)
        if ((rc = exact_probe(ix, ix->ws_q.p, 1, nprobe, dist_type, ix->ws_probes.p, 0))) return rc;
        HIPCHK(hipMemcpy(pr.data(), ix->ws_probes.p, nprobe * 4, hipMemcpyDeviceToHost));
    }
    std::vector<int64_t> offs(nprobe);
    std::vector<int2> pieces;  // 8-block pieces of the probed lists, one workgroup each
    int64_t tot = 0;
    for (int i = 0; i < nprobe; ++i) {
        offs[i] = tot;
        tot += ix->clu_n[pr[i]];
        const int nb = (ix->clu_n[pr[i]] + kBlock - 1) / kBlock;
        for (int st = 0; st < nb; st += 8) pieces.push_back(make_int2(i, st));
    }
    // the index's estimate workspace, grown to the largest call and kept
    DevBuf<int64_t> &d_offs = ix->ws_eoffs;
    DevBuf<uint32_t> &d_ids = ix->ws_eids;
    DevBuf<float> &d_out = ix->ws_eout;
    if ((rc = upload(d_offs, offs)) || (rc = upload(ix->ws_epieces, pieces))) return rc;
    if (d_ids.ensure(std::max<int64_t>(1, tot)) || d_out.ensure(3 * std::max<int64_t>(1, tot)))
        return fail(SAQGPU_ENOMEM, "allocation failed");
    HIPCHK(launch_query_prep(ix->view(), ix->ws_q.p, 1, ix->variance.p, ix->has_rot ? ix->rot.p : nullptr,
                             ix->d_rot_off.p, vars_bound_m, ix->ws_qrot.p, ix->ws_bound.p, ix->max_dim, 0));
    HIPCHK(launch_estimate(ix->view(), ix->ws_qrot.p, ix->ws_bound.p, ix->ws_probes.p, d_offs.p, ix->ws_epieces.p,
                           (int)pieces.size(), dist_type, d_ids.p, d_out.p, d_out.p + tot, d_out.p + 2 * tot, 0));
    if (tot) {
        HIPCHK(hipMemcpy(out_ids, d_ids.p, tot * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(out_acc, d_out.p, tot * 4, hipMemcpyDeviceToHost));
        if (out_fast) HIPCHK(hipMemcpy(out_fast, d_out.p + tot, tot * 4, hipMemcpyDeviceToHost));
        if (out_vars) HIPCHK(hipMemcpy(out_vars, d_out.p + 2 * tot, tot * 4, hipMemcpyDeviceToHost));
    }
    if (n_out) *n_out = tot;
    return SAQGPU_OK;
}

}  // extern "C"

namespace {
// QuantSingleConfig checks of the encoder (CAQEncoder's constructor, caq_encoder.hpp:163-169: the
// original bit count exceeds every segment's; codes are kept in 16 bits here).
int check_caq_cfg(const saqgpu_index *ix, const saqgpu_caq_cfg *cfg) {
    if (cfg->caq_adj_rd_lmt < 0) return fail(SAQGPU_EINVAL, "caq_adj_rd_lmt must be >= 0");
    if (cfg->caq_ori_qB) {
        int maxb = 0;
        for (int b : ix->seg_bits) maxb = std::max(maxb, b);
        if (cfg->caq_ori_qB <= maxb || cfg->caq_ori_qB > 16)
            return fail(SAQGPU_EINVAL, "caq_ori_qB must exceed every segment's bits and be at most 16");
    }
    return SAQGPU_OK;
}

}  // namespace

extern "C" {

static int encode_impl(saqgpu_index *ix, const float *data, const uint32_t *cluster_ids, int64_t n,
                       const float *one_centroid, const saqgpu_caq_cfg *cfg, int32_t flags, int64_t *n_check_failed,
                       void *stream) {
    if (!ix || !data || !cluster_ids || !cfg) return fail(SAQGPU_EINVAL, "null argument");
    if (n != ix->num_data) return fail(SAQGPU_EINVAL, "n must equal the index's num_data");
    if (int rc_ = check_caq_cfg(ix, cfg)) return rc_;
    if (int rc_ = check_mutable(ix)) return rc_;
    IndexCall call(ix, (hipStream_t)stream);
    if (int rc_ = call.begin()) return rc_;
    hipStream_t st = (hipStream_t)stream;
    const bool dev = flags & SAQGPU_DEVICE_PTRS;
    // the lists' members in increasing id order (IVF::construct's bucketing), sorted on the GPU
    DevBuf<uint32_t> d_cid, d_iota, d_src, d_clu;
    DevBuf<int64_t> d_gvec, d_gblk;
    DevBuf<int32_t> d_lane;
    DevBuf<uint8_t> d_sort;
    DevBuf<int> d_fail;
    const size_t sort_bytes = slot_sort_bytes(n);
    const size_t n1 = std::max<int64_t>(1, n);
    if (d_iota.ensure(n1) || d_src.ensure(n1) || d_clu.ensure(n1) || d_gvec.ensure(n1) || d_gblk.ensure(n1) ||
        d_lane.ensure(n1) || d_sort.ensure(std::max<size_t>(1, sort_bytes)) || d_fail.ensure(2))
        return fail(SAQGPU_ENOMEM, "allocation failed");
    const uint32_t *cid = cluster_ids;
    if (!dev) {
        if (d_cid.ensure(n1)) return fail(SAQGPU_ENOMEM, "allocation failed");
        HIPCHK(hipMemcpyAsync(d_cid.p, cluster_ids, n * 4, hipMemcpyHostToDevice, st));
        cid = d_cid.p;
    }
    HIPCHK(hipMemsetAsync(d_fail.p, 0, 2 * sizeof(int), st));
    HIPCHK(launch_make_slots(cid, n, ix->num_cen, ix->d_clu_vec.p, ix->d_clu_blk.p, ix->d_clu_n.p, d_iota.p, d_src.p,
                             d_clu.p, d_gvec.p, d_gblk.p, d_lane.p, d_sort.p, sort_bytes, d_fail.p + 1, st));
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, d_fail.p + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (bad & 1) return fail(SAQGPU_EINVAL, "cluster id out of range");
    if (bad & 2) return fail(SAQGPU_EINVAL, "cluster ids disagree with the index's cluster sizes");
    DevBuf<float> d_data, d_O, d_bound;
    DevBuf<uint16_t> d_code;
    const float *ddata = data;
    if (!dev) {
        if (d_data.ensure(std::max<int64_t>(1, n) * ix->num_dim)) return fail(SAQGPU_ENOMEM, "allocation failed");
        HIPCHK(hipMemcpy(d_data.p, data, n * ix->num_dim * 4, hipMemcpyHostToDevice));
        ddata = d_data.p;
    }
    const IndexView v = ix->view();
    // rotated centroids c_s * P_s of every list (quantizer.hpp:46), Eigen GEMV order; with one
    // shared centroid, its rotation in every list's row
    if (d_bound.ensure((size_t)ix->num_cen * ix->nseg)) return fail(SAQGPU_ENOMEM, "allocation failed");
    DevBuf<float> d_cen1;
    const float *cent = ix->centroids.p;
    if (one_centroid) {
        if (d_cen1.ensure(ix->num_dim)) return fail(SAQGPU_ENOMEM, "allocation failed");
        HIPCHK(hipMemcpyAsync(d_cen1.p, one_centroid, ix->num_dim * 4, hipMemcpyHostToDevice, st));
        cent = d_cen1.p;
        HIPCHK(launch_query_prep(v, cent, 1, ix->variance.p, ix->has_rot ? ix->rot.p : nullptr, ix->d_rot_off.p, 0.0f,
                                 ix->crot.p, d_bound.p, ix->max_dim, st));
        for (int64_t have = 1; have < ix->num_cen; have *= 2) {
            const int64_t cp = std::min<int64_t>(have, ix->num_cen - have);
            HIPCHK(hipMemcpyAsync(ix->crot.p + have * ix->D_pad, ix->crot.p, cp * ix->D_pad * 4,
                                  hipMemcpyDeviceToDevice, st));
        }
    } else {
        HIPCHK(launch_query_prep(v, cent, ix->num_cen, ix->variance.p, ix->has_rot ? ix->rot.p : nullptr,
                                 ix->d_rot_off.p, 0.0f, ix->crot.p, d_bound.p, ix->max_dim, st));
    }
    const int64_t chunk = std::min<int64_t>(std::max<int64_t>(1, n), 1 << 18);
    if (d_O.ensure(chunk * ix->max_dim) || d_code.ensure(chunk * ix->max_dim))
        return fail(SAQGPU_ENOMEM, "allocation failed");
    for (int64_t off = 0; off < n; off += chunk) {
        const int64_t cnt = std::min(chunk, n - off);
        for (int s = 0; s < ix->nseg; ++s)
            HIPCHK(launch_encode_segment(v, ddata, d_src.p + off, d_gvec.p + off, d_gblk.p + off, d_lane.p + off,
                                         d_clu.p + off, cent,
                                         ix->has_rot ? ix->rot.p + ix->rot_off[s] : nullptr, s, ix->segs[s].dim, cnt,
                                         chunk, cfg->caq_adj_rd_lmt, cfg->caq_adj_eps, cfg->caq_ori_qB,
                                         one_centroid != nullptr, d_O.p, d_code.p, d_fail.p, st));
    }
    int nf = 0;
    HIPCHK(hipMemcpyAsync(&nf, d_fail.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (n_check_failed) *n_check_failed = nf;
    if (n) HIPCHK(hipMemcpy(ix->ids.p, d_src.p, n * 4, hipMemcpyDeviceToDevice));
    std::fill(ix->clu_set.begin(), ix->clu_set.end(), 1);
    ix->bq_ready = false;
    return SAQGPU_OK;
}

int saqgpu_encode(saqgpu_index *ix, const float *data, const uint32_t *cluster_ids, int64_t n,
                  const saqgpu_caq_cfg *cfg, int32_t flags, int64_t *n_check_failed, void *stream) {
    return encode_impl(ix, data, cluster_ids, n, nullptr, cfg, flags, n_check_failed, stream);
}

int saqgpu_encode_one_centroid(saqgpu_index *ix, const float *data, const uint32_t *cluster_ids, int64_t n,
                               const float *centroid, const saqgpu_caq_cfg *cfg, int32_t flags,
                               int64_t *n_check_failed, void *stream) {
    if (!centroid) return fail(SAQGPU_EINVAL, "null centroid");
    return encode_impl(ix, data, cluster_ids, n, centroid, cfg, flags, n_check_failed, stream);
}

int saqgpu_encode_residuals(saqgpu_index *ix, int32_t cid, int32_t seg, const float *o, const float *centroid_seg,
                            const saqgpu_caq_cfg *cfg) {
    if (!ix || !o || !cfg || cid < 0 || cid >= ix->num_cen || seg < 0 || seg >= ix->nseg)
        return fail(SAQGPU_EINVAL, "bad argument");
    if (int rc_ = check_caq_cfg(ix, cfg)) return rc_;
    if (int rc_ = check_mutable(ix)) return rc_;
    IndexCall call(ix, nullptr);
    if (int rc_ = call.begin()) return rc_;
    const int64_t n = ix->clu_n[cid];
    if (n == 0) return SAQGPU_OK;
    const int D = ix->segs[seg].dim;
    std::vector<float> ot((size_t)D * n);
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < D; ++j) ot[(size_t)j * n + i] = o[i * D + j];
    std::vector<float> cen(ix->num_dim, 0.0f);
    const int doff = ix->segs[seg].doff;
    for (int j = 0; j < D && doff + j < ix->num_dim; ++j) cen[doff + j] = centroid_seg ? centroid_seg[j] : 0.0f;
    std::vector<int64_t> gvec(n), gblk(n);
    std::vector<int32_t> lane(n);
    std::vector<uint32_t> clu(n, 0);
    for (int64_t r = 0; r < n; ++r) {
        gvec[r] = ix->clu_vec[cid] + r;
        gblk[r] = ix->clu_blk[cid] + r / kBlock;
        lane[r] = (int32_t)(r % kBlock);
    }
    DevBuf<int64_t> d_gvec, d_gblk;
    DevBuf<int32_t> d_lane;
    DevBuf<uint32_t> d_clu;
    DevBuf<float> d_O, d_cen;
    DevBuf<uint16_t> d_code;
    DevBuf<int> d_fail;
    int rc = 0;
    std::vector<float> otv(ot);
    if ((rc = upload(d_gvec, gvec)) || (rc = upload(d_gblk, gblk)) || (rc = upload(d_lane, lane)) ||
        (rc = upload(d_clu, clu)) || (rc = upload(d_O, otv)) || (rc = upload(d_cen, cen)))
        return rc;
    if (d_code.ensure((size_t)D * n) || d_fail.ensure(1)) return fail(SAQGPU_ENOMEM, "allocation failed");
    HIPCHK(hipMemset(d_fail.p, 0, sizeof(int)));
    HIPCHK(launch_caq_only(ix->view(), d_gvec.p, d_gblk.p, d_lane.p, d_clu.p, d_cen.p, seg, n, n,
                           cfg->caq_adj_rd_lmt, cfg->caq_adj_eps, cfg->caq_ori_qB, d_O.p, d_code.p, d_fail.p, 0));
    HIPCHK(hipDeviceSynchronize());
    ix->bq_ready = false;
    return SAQGPU_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batched scan
namespace {

BatchedView batched_view(const saqgpu_index *ix) {
    BatchedView b{};
    b.nseg = ix->nseg;
    b.nqs = ix->bq_nqs;
    b.D_pad = ix->D_pad;
    b.kq = ix->bq_kq;
    b.fac_off = ix->bq_fac_off;
    b.kq_stride = ix->bq_kq + 16;
    b.mb_stride = ix->bq_stride;
    b.mb = ix->bq_mb.p;
    b.clu_mb = ix->d_clu_mb.p;
    b.clu_n = ix->d_clu_n.p;
    b.crot = ix->crot.p;
    b.segs = ix->d_bsegs.p;
    return b;
}

// Layout of one 16-vector record: the code chunks of every quantised segment, then
// |o|^2[16], rescale[nqs][16], id[16].
int prepare_batched_locked(saqgpu_index *ix) {
    if (ix->bq_ready) return SAQGPU_OK;
    if (ix->parent) return fail(SAQGPU_ESTATE, "build the batched layout on the index before opening the handle");
    for (int c = 0; c < ix->num_cen; ++c)
        if (!ix->clu_set[c] && ix->clu_n[c] > 0) return fail(SAQGPU_ESTATE, "index has clusters not yet set");
    for (int s = 0; s < ix->nseg; ++s)
        if (ix->segs[s].dim > 2048)
            return fail(SAQGPU_EUNSUPPORTED, "the batched scan takes segments of up to 2048 dimensions");
    if (ix->num_cen > 1024 * kScanRounds)
        return fail(SAQGPU_EUNSUPPORTED, "the batched scan takes up to 16384 lists");
    ix->bsegs.assign(ix->nseg, BSeg{});
    int kq = 0, nqs = 0;
    int64_t code = 0;
    std::vector<int32_t> chunk_seg, chunk_kc;
    for (int s = 0; s < ix->nseg; ++s) {
        const SegDesc &sd = ix->segs[s];
        BSeg &b = ix->bsegs[s];
        b.dim = sd.dim;
        b.bits = sd.bits;
        b.doff = sd.doff;
        b.sqd = sd.bits ? (float)(2.0 / (double)(1 << sd.bits)) : 0.0f;
        b.qs = -1;
        if (!sd.bits) continue;
        b.qs = nqs++;
        b.kq_off = kq;
        kq += sd.dim;
        b.chunk_bytes = sd.bits <= 4 ? 512 : (sd.bits <= 8 ? 1024 : 2048);
        b.code_off = (int32_t)code;
        code += (int64_t)b.chunk_bytes * (sd.dim / 64);
        for (int kc = 0; kc < sd.dim / 64; ++kc) {
            chunk_seg.push_back(s);
            chunk_kc.push_back(kc);
        }
    }
    ix->bq_nqs = nqs;
    ix->bq_kq = kq;
    ix->bq_fac_off = (int)code;
    ix->bq_stride = rd_up(code + (int64_t)kMB * 4 * (2 + nqs), 256);
    ix->clu_mb.assign(ix->num_cen, 0);
    int64_t tot = 0;
    std::vector<int32_t> mb_clu;
    for (int c = 0; c < ix->num_cen; ++c) {
        ix->clu_mb[c] = tot;
        const int64_t nmb = (ix->clu_n[c] + kMB - 1) / kMB;
        for (int64_t i = 0; i < nmb; ++i) mb_clu.push_back(c);
        tot += nmb;
    }
    ix->bq_total_mb = tot;
    if (ix->bq_mb.ensure(std::max<int64_t>(1, tot) * ix->bq_stride)) return fail(SAQGPU_ENOMEM, "allocation failed");
    int rc;
    std::vector<int32_t> order(ix->num_cen);
    for (int c = 0; c < ix->num_cen; ++c) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return ix->clu_n[a] > ix->clu_n[b]; });
    if ((rc = upload(ix->d_clu_mb, ix->clu_mb)) || (rc = upload(ix->d_bsegs, ix->bsegs)) ||
        (rc = upload(ix->d_bq_order, order)))
        return rc;
    DevBuf<int32_t> d_mb_clu, d_cs, d_ck;
    if ((rc = upload(d_mb_clu, mb_clu)) || (rc = upload(d_cs, chunk_seg)) || (rc = upload(d_ck, chunk_kc))) return rc;
    HIPCHK(launch_bq_build(ix->view(), batched_view(ix), d_mb_clu.p, tot, d_cs.p, d_ck.p, (int)chunk_seg.size(),
                           ix->bq_mb.p, 0));
    if ((rc = ensure_cnorm(ix, 0))) return rc;
    HIPCHK(hipDeviceSynchronize());
    d_mb_clu.release();
    d_cs.release();
    d_ck.release();
    ix->bq_ready = true;
    return SAQGPU_OK;
}

int check_batched_params(const saqgpu_index *ix, const saqgpu_batched_params *p, int64_t nq, int32_t flags) {
    if (p->topk < 1 || p->topk > 256) return fail(SAQGPU_EINVAL, "batched topk must be in 1..256");
    if (p->nprobe < 1 || p->nprobe > ix->num_cen) return fail(SAQGPU_EINVAL, "nprobe must be in 1..K");
    if (p->precision != SAQGPU_PREC_INT8 && p->precision != SAQGPU_PREC_INT8X2 && p->precision != SAQGPU_PREC_BF16)
        return fail(SAQGPU_EINVAL, "precision must be SAQGPU_PREC_INT8, SAQGPU_PREC_INT8X2 or SAQGPU_PREC_BF16");
    if (p->query_tile != 0 && p->query_tile != 16 && p->query_tile != 32 && p->query_tile != 64 &&
        p->query_tile != 128)
        return fail(SAQGPU_EINVAL, "query_tile must be 0, 16, 32, 64 or 128");
    if (p->nprobe_per_query && !(flags & SAQGPU_DEVICE_PTRS))
        for (int64_t i = 0; i < nq; ++i)
            if (p->nprobe_per_query[i] < 1 || p->nprobe_per_query[i] > p->nprobe)
                return fail(SAQGPU_EINVAL, "nprobe_per_query entries must be in 1..nprobe");
    return SAQGPU_OK;
}

// Queries per work tile when the autotuner has not run: the expected queries per probed list.
int default_query_tile(const saqgpu_index *ix, int64_t nq, int nprobe) {
    const double per_list = (double)nq * nprobe / std::max(1, ix->num_cen);
    return per_list >= 96 ? 128 : (per_list >= 48 ? 64 : (per_list >= 24 ? 32 : 16));
}

// The batched scan's coarse quantiser: q.c for every (query, centroid) on the matrix cores, as
// one rocBLAS bfloat16 GEMM over K = 3 D of the rows split into bfloat16 parts (centroids
// hi | lo | hi, queries hi | hi | lo: the products sum to hi.hi + lo.hi + hi.lo, fp32
// accumulation), into ws_cdist; then the exact top-nprobe selection over |c|^2 - 2 q.c (or q.c
// for IP).  The dropped lo.lo term and the parts' rounding are below 2^-16 of each |q_i c_i|,
// about the accumulation error of an fp32 GEMM.  Not FlatInitializer's reduction order: a probe
// list can differ from the reference's where two centroid distances lie within that error of
// each other.  ws_qbf3 and cen_bf3 are allocated by the caller (cen_bf3 by ensure_cen_bf3).
int probe_gemm(saqgpu_index *ix, const float *dq, int64_t nq, int nprobe, int dist_type, uint32_t *probes,
               hipStream_t st) {
    const int K = ix->num_cen, D = ix->num_dim;
    int rc;
    if ((rc = ensure_cnorm(ix, st)) || (rc = ensure_cen_bf3(ix, st)) || (rc = blas_on(ix, st))) return rc;
    HIPCHK(launch_split_bf16(dq, nq, D, 1, ix->ws_qbf3.p, st));
    const float one = 1.0f, zero = 0.0f;
    // column-major: ip^T (K x nq, ld K) = cen_bf3 (3D x K, ld 3D)^T * ws_qbf3 (3D x nq, ld 3D)
    if (rocblas_gemm_ex(ix->blas, rocblas_operation_transpose, rocblas_operation_none, K, (int)nq, 3 * D, &one,
                        ix->cen_bf3.p, rocblas_datatype_bf16_r, 3 * D, ix->ws_qbf3.p, rocblas_datatype_bf16_r, 3 * D,
                        &zero, ix->ws_cdist.p, rocblas_datatype_f32_r, K, ix->ws_cdist.p, rocblas_datatype_f32_r, K,
                        rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0) != rocblas_status_success)
        return fail(SAQGPU_EDEVICE, "rocblas_gemm_ex failed");
    HIPCHK(launch_probe_select_ip(ix->ws_cdist.p, ix->cnorm.p, K, nq, nprobe, dist_type, probes, st));
    return SAQGPU_OK;
}

int search_batched_locked(saqgpu_index *ix, const float *queries, int64_t nq, const uint32_t *probes,
                          const saqgpu_batched_params *p, int query_tile, uint32_t *out_ids, float *out_dists,
                          int32_t flags, hipStream_t st, bool time_it) {
    int rc = prepare_batched_locked(ix);
    if (rc) return rc;
    const bool dev = flags & SAQGPU_DEVICE_PTRS;
    const bool graph = (flags & SAQGPU_USE_GRAPH) != 0;
    if (graph && !dev) return fail(SAQGPU_EINVAL, "SAQGPU_USE_GRAPH needs SAQGPU_DEVICE_PTRS");
    if (graph && !st) return fail(SAQGPU_EINVAL, "SAQGPU_USE_GRAPH needs a stream other than the null stream");
    const int NP = p->nprobe, K = ix->num_cen, topk = p->topk;
    // query terms: 1 (int8), 2 (int8x2) or BF16 (3: one term, rows of 2 bytes per dimension)
    const int qs = p->precision == SAQGPU_PREC_INT8X2 ? 2 : (p->precision == SAQGPU_PREC_BF16 ? 3 : 1);
    const BatchedView bv = batched_view(ix);
    int max_smem = 0;
    HIPCHK(hipDeviceGetAttribute(&max_smem, hipDeviceAttributeMaxSharedMemoryPerBlock, ix->device));
    // a tile the caller left to the library shrinks until its per-query lists fit in LDS (large k)
    if (p->query_tile == 0)
        while (query_tile > 16 && (int64_t)bscan_smem_bytes(bv, topk, query_tile, qs) > max_smem) query_tile /= 2;
    if ((int64_t)bscan_smem_bytes(bv, topk, query_tile, qs) > max_smem)
        return fail(SAQGPU_EUNSUPPORTED, "query tile needs more LDS than a workgroup has; use a smaller query_tile");
    const int64_t npairs = nq * NP;
    const int64_t max_tiles = bq_max_tiles(npairs, K), max_cands = bq_max_cands(npairs, query_tile);
    if (max_tiles >= INT32_MAX || max_cands * topk >= ((int64_t)1 << 40))
        return fail(SAQGPU_EINVAL, "batch too large: split the queries");
    // every allocation first: none may happen while a graph is being captured
    rc = 0;
    rc |= ix->ws_qrot.ensure(nq * ix->D_pad);
    rc |= ix->ws_bound.ensure(nq * ix->nseg);
    rc |= ix->bq_ws.ensure(5 * (size_t)K + 2 * npairs + 5 * max_tiles + 4 + nq);
    rc |= ix->bq_qq.ensure((size_t)(qs == 1 ? 1 : 2) * npairs * ix->bq_kq);
    rc |= ix->bq_pc.ensure((size_t)npairs * (3 * ix->bq_nqs + 1));
    rc |= ix->bq_cand_d.ensure(max_cands * topk);
    rc |= ix->bq_cand_id.ensure(max_cands * topk);
    if (!dev) {
        rc |= ix->ws_q.ensure(nq * ix->num_dim);
        rc |= ix->ws_ids.ensure(nq * topk);
        rc |= ix->ws_dists.ensure(nq * topk);
        if (p->nprobe_per_query) rc |= ix->bq_npq.ensure(nq);
    }
    if (!probes || !dev) rc |= ix->ws_probes.ensure(npairs);
    if (!probes) {
        rc |= ix->ws_cdist.ensure((size_t)nq * K);
        if (p->exact_probe) rc |= ix->ws_ties.ensure(std::max<int64_t>(1, nq));
        else rc |= ix->ws_qbf3.ensure((size_t)nq * 3 * ix->num_dim);
    }
    if (rc) return fail(SAQGPU_ENOMEM, "workspace allocation failed");
    // the replay of a captured graph: the same shape, parameters and buffers as the capture
    BatchedGraphKey key{nq, NP, topk, p->precision, query_tile, p->exact_probe, queries, probes,
                        p->nprobe_per_query, out_ids, out_dists, st, ix->bq_ws.p, ix->bq_qq.p, ix->bq_cand_d.p,
                        ix->ws_cdist.p, ix->ws_probes.p};
    if (graph && ix->bq_graph && key == ix->bq_graph_key) {
        HIPCHK(hipGraphLaunch(ix->bq_graph, st));
        return SAQGPU_OK;
    }
    if (!probes && (rc = ensure_cnorm(ix, st))) return rc;
    if (!probes && !p->exact_probe && (rc = ensure_cen_bf3(ix, st))) return rc;
    const float *dq = queries;
    const uint32_t *dp = probes;
    const int32_t *dnpq = p->nprobe_per_query;
    uint32_t *dids = out_ids;
    float *dd = out_dists;
    if (!dev) {
        HIPCHK(hipMemcpyAsync(ix->ws_q.p, queries, nq * ix->num_dim * 4, hipMemcpyHostToDevice, st));
        dq = ix->ws_q.p;
        if (probes) {
            HIPCHK(hipMemcpyAsync(ix->ws_probes.p, probes, npairs * 4, hipMemcpyHostToDevice, st));
            dp = ix->ws_probes.p;
        }
        if (dnpq) {
            HIPCHK(hipMemcpyAsync(ix->bq_npq.p, p->nprobe_per_query, nq * 4, hipMemcpyHostToDevice, st));
            dnpq = ix->bq_npq.p;
        }
        dids = ix->ws_ids.p;
        dd = out_dists ? ix->ws_dists.p : nullptr;
    }
    int32_t *wsp = ix->bq_ws.p;
    BqWork w;
    w.counts = wsp;
    w.cstart = wsp + K;
    w.candstart = wsp + 2 * K;
    w.nsplit = wsp + 3 * K;
    w.slot = wsp + 4 * K;
    w.pairs = w.slot + npairs;
    w.tile_clu = w.pairs + npairs;
    w.tile_first = w.tile_clu + max_tiles;
    w.tile_cnt = w.tile_first + max_tiles;
    w.tile_v0 = w.tile_cnt + max_tiles;
    w.tile_cand = w.tile_v0 + max_tiles;
    w.ntiles = w.tile_cand + max_tiles;
    w.npairs = w.ntiles + 1;
    w.work = w.npairs + 1;
    w.sv = w.work + 1;
    w.qbound = reinterpret_cast<uint32_t *>(w.sv + 1);
    w.tfirst = w.sv + 1 + nq;
    BqPairBufs pb;
    pb.qq1 = ix->bq_qq.p;
    pb.qq2 = qs == 2 ? ix->bq_qq.p + npairs * ix->bq_kq : nullptr;
    pb.pc = ix->bq_pc.p;
    pb.pq2 = ix->bq_pc.p + npairs * 3 * ix->bq_nqs;
    const int grid = (int)std::min<int64_t>(max_tiles, 1 << 20);
    // a second stream for the query rotation (the plain scan's first group stream)
    if (ix->sl_groups.empty()) {
        saqgpu_index::SliceGroup g{};
        HIPCHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
        ix->sl_groups.push_back(g);
    }
    if (!ix->sl_fork) HIPCHK(hipEventCreateWithFlags(&ix->sl_fork, hipEventDisableTiming));
    const hipStream_t aux = ix->sl_groups[0].stream;
    const hipEvent_t aux_done = ix->sl_groups[0].done;
    // the step's device work (also what a graph captures): query rotation on the second stream
    // beside the coarse quantiser and the grouping, then pair prep, tile kernel, merge
    auto enqueue = [&](bool timed) -> int {
        int rc2;
        HIPCHK(hipEventRecord(ix->sl_fork, st));
        HIPCHK(hipStreamWaitEvent(aux, ix->sl_fork, 0));
        HIPCHK(launch_query_prep(ix->view(), dq, nq, ix->variance.p, ix->has_rot ? ix->rot.p : nullptr,
                                 ix->d_rot_off.p, 4.0f, ix->ws_qrot.p, ix->ws_bound.p, ix->max_dim, aux));
        HIPCHK(hipEventRecord(aux_done, aux));
        if (!probes) {
            if (p->exact_probe) {
                if ((rc2 = exact_probe(ix, dq, nq, NP, 0, ix->ws_probes.p, st))) return rc2;
            } else if ((rc2 = probe_gemm(ix, dq, nq, NP, 0, ix->ws_probes.p, st))) {
                return rc2;
            }
            dp = ix->ws_probes.p;
        }
        HIPCHK(launch_bq_group(dp, dnpq, nq, NP, K, query_tile, qs, ix->d_bq_order.p, ix->d_clu_n.p, w, st));
        HIPCHK(hipStreamWaitEvent(st, aux_done, 0));
        HIPCHK(launch_bq_pairprep(bv, ix->ws_qrot.p, dp, w, pb, NP, npairs, qs, st));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timed) {
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            HIPCHK(hipEventRecord(e0, st));
        }
        HIPCHK(launch_bscan_tiles(bv, w, pb, NP, topk, ix->bq_cand_d.p, ix->bq_cand_id.p, query_tile, qs, grid,
                                  st));
        if (timed) {
            HIPCHK(hipEventRecord(e1, st));
            ix->events.emplace_back(e0, e1);
        }
        HIPCHK(launch_bq_merge(ix->bq_cand_d.p, ix->bq_cand_id.p, dp, w, dnpq, nq, NP, topk, dids, dd, st));
        return SAQGPU_OK;
    };
    if ((rc = enqueue(time_it && !graph))) return rc;
    if (graph) {
        // this call ran eagerly (which also warms the libraries up); capture the same work for the
        // calls that follow with the same key
        if (ix->bq_graph) {
            (void)hipGraphExecDestroy(ix->bq_graph);
            ix->bq_graph = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int crc = enqueue(false);
        const hipError_t ce = hipStreamEndCapture(st, &g);
        if (crc) {
            if (g) (void)hipGraphDestroy(g);
            return crc;
        }
        if (ce != hipSuccess) return fail(SAQGPU_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        hipGraphExec_t ex = nullptr;
        const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) return fail(SAQGPU_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        ix->bq_graph = ex;
        ix->bq_graph_key = key;
    }
    if (!dev) {
        HIPCHK(hipMemcpyAsync(out_ids, dids, nq * topk * 4, hipMemcpyDeviceToHost, st));
        if (out_dists) HIPCHK(hipMemcpyAsync(out_dists, dd, nq * topk * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    return SAQGPU_OK;
}

}  // namespace

extern "C" {

int saqgpu_prepare_batched(saqgpu_index *ix) {
    if (!ix) return fail(SAQGPU_EINVAL, "null index");
    IndexCall call(ix, nullptr);
    if (int rc_ = call.begin()) return rc_;
    return prepare_batched_locked(ix);
}

int saqgpu_search_batched(saqgpu_index *ix, const float *queries, int64_t nq, const uint32_t *probes,
                          const saqgpu_batched_params *p, uint32_t *out_ids, float *out_dists, int32_t flags,
                          void *stream) {
    if (!ix || !queries || !p || !out_ids || nq < 0) return fail(SAQGPU_EINVAL, "null argument");
    int rc = check_batched_params(ix, p, nq, flags);
    if (rc) return rc;
    if (nq == 0) return SAQGPU_OK;
    if (probes && !(flags & SAQGPU_DEVICE_PTRS))
        for (int64_t i = 0; i < nq * p->nprobe; ++i)
            if (probes[i] >= (uint32_t)ix->num_cen) return fail(SAQGPU_EINVAL, "probe id out of range");
    IndexCall call(ix, (hipStream_t)stream);
    if (int rc_ = call.begin()) return rc_;
    const int qt = p->query_tile ? p->query_tile : (ix->bq_tile ? ix->bq_tile : default_query_tile(ix, nq, p->nprobe));
    return search_batched_locked(ix, queries, nq, probes, p, qt, out_ids, out_dists, flags, (hipStream_t)stream,
                                 ix->timing);
}

int saqgpu_autotune_batched(saqgpu_index *ix, const float *queries, int64_t nq, const uint32_t *probes,
                            const saqgpu_batched_params *p, int32_t flags, void *stream, int32_t *best_query_tile) {
    if (!ix || !queries || !p || nq < 1) return fail(SAQGPU_EINVAL, "null argument");
    int rc = check_batched_params(ix, p, nq, flags);
    if (rc) return rc;
    IndexCall call(ix, (hipStream_t)stream);
    if (int rc_ = call.begin()) return rc_;
    hipStream_t st = (hipStream_t)stream;
    const bool dev = flags & SAQGPU_DEVICE_PTRS;
    DevBuf<uint32_t> ids;
    if (ids.ensure(nq * p->topk)) return fail(SAQGPU_ENOMEM, "allocation failed");
    // the timed calls read device-resident copies so that only the search is measured
    DevBuf<float> q;
    DevBuf<uint32_t> pr;
    DevBuf<int32_t> npq;
    const float *dq = queries;
    const uint32_t *dp = probes;
    saqgpu_batched_params pp = *p;
    if (!dev) {
        if (q.ensure(nq * ix->num_dim)) return fail(SAQGPU_ENOMEM, "allocation failed");
        HIPCHK(hipMemcpy(q.p, queries, nq * ix->num_dim * 4, hipMemcpyHostToDevice));
        dq = q.p;
        if (probes) {
            if (pr.ensure(nq * p->nprobe)) return fail(SAQGPU_ENOMEM, "allocation failed");
            HIPCHK(hipMemcpy(pr.p, probes, nq * p->nprobe * 4, hipMemcpyHostToDevice));
            dp = pr.p;
        }
        if (p->nprobe_per_query) {
            if (npq.ensure(nq)) return fail(SAQGPU_ENOMEM, "allocation failed");
            HIPCHK(hipMemcpy(npq.p, p->nprobe_per_query, nq * 4, hipMemcpyHostToDevice));
            pp.nprobe_per_query = npq.p;
        }
    }
    int max_smem = 0;
    HIPCHK(hipDeviceGetAttribute(&max_smem, hipDeviceAttributeMaxSharedMemoryPerBlock, ix->device));
    if ((rc = prepare_batched_locked(ix))) return rc;
    const BatchedView bv = batched_view(ix);
    const int qs = p->precision == SAQGPU_PREC_INT8X2 ? 2 : (p->precision == SAQGPU_PREC_BF16 ? 3 : 1);
    float best = 1e30f;
    int best_qt = 16;
    for (int qt : {16, 32, 64, 128}) {
        if ((int64_t)bscan_smem_bytes(bv, p->topk, qt, qs) > max_smem) continue;
        hipEvent_t a, b;
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        // one untimed call, then two timed ones
        if ((rc = search_batched_locked(ix, dq, nq, dp, &pp, qt, ids.p, nullptr, SAQGPU_DEVICE_PTRS, st, false))) return rc;
        HIPCHK(hipEventRecord(a, st));
        for (int r = 0; r < 2; ++r)
            if ((rc = search_batched_locked(ix, dq, nq, dp, &pp, qt, ids.p, nullptr, SAQGPU_DEVICE_PTRS, st, false)))
                return rc;
        HIPCHK(hipEventRecord(b, st));
        HIPCHK(hipEventSynchronize(b));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (ms < best) {
            best = ms;
            best_qt = qt;
        }
    }
    ix->bq_tile = best_qt;
    if (best_query_tile) *best_query_tile = best_qt;
    return SAQGPU_OK;
}

}  // extern "C"

namespace {
// SaqSingleDataWrapper's sections (single_data.hpp:180-285): factors, short codes (64-byte
// padded), long codes, ExFactors; per segment its short- and long-code offsets (-1: none).
struct SingleLayout {
    int64_t code_off = 0, long_off = 0, exf_off = 0, bytes = 0;
    std::vector<int64_t> seg_code, seg_long;
};

SingleLayout single_layout(const saqgpu_index *ix) {
    SingleLayout L;
    const auto r64 = [](int64_t x) { return (x + 63) / 64 * 64; };
    int64_t code = 0, lng = 0;
    for (int s = 0; s < ix->nseg; ++s) {
        const int64_t D = ix->seg_dim[s], b = ix->seg_bits[s];
        L.seg_code.push_back(b > 0 ? code : -1);
        L.seg_long.push_back(b > 1 ? lng : -1);
        if (b > 0) code += D / 8;
        if (b > 1) lng += D * (b - 1) / 8;
    }
    L.code_off = r64((int64_t)ix->nseg * 8);
    L.long_off = L.code_off + r64(code);
    L.exf_off = L.long_off + lng;
    L.bytes = L.exf_off + (int64_t)ix->nseg * 8;
    for (int s = 0; s < ix->nseg; ++s) {
        if (L.seg_code[s] >= 0) L.seg_code[s] += L.code_off;
        if (L.seg_long[s] >= 0) L.seg_long[s] += L.long_off;
    }
    return L;
}

int check_stride(const SingleLayout &L, int64_t stride) {
    if (stride < L.bytes || stride % 8) return fail(SAQGPU_EINVAL, "record_stride must be >= the record size and a multiple of 8");
    return SAQGPU_OK;
}
}  // namespace

extern "C" {

int saqgpu_single_layout(const saqgpu_index *ix, int64_t *record_bytes, int64_t *offsets) {
    if (!ix || !record_bytes) return fail(SAQGPU_EINVAL, "null argument");
    const SingleLayout L = single_layout(ix);
    *record_bytes = L.bytes;
    if (offsets) {
        offsets[0] = L.code_off;
        offsets[1] = L.long_off;
        offsets[2] = L.exf_off;
        for (int s = 0; s < ix->nseg; ++s) {
            offsets[3 + 2 * s] = L.seg_code[s];
            offsets[4 + 2 * s] = L.seg_long[s];
        }
    }
    return SAQGPU_OK;
}

int saqgpu_single_quantize(saqgpu_index *ix, const float *data, int64_t n, const saqgpu_caq_cfg *cfg,
                           uint8_t *records, int64_t record_stride, int32_t flags, int64_t *n_check_failed,
                           void *stream) {
    if (!ix || !cfg || n < 0 || (n > 0 && (!data || !records))) return fail(SAQGPU_EINVAL, "bad argument");
    if (int rc_ = check_caq_cfg(ix, cfg)) return rc_;
    const SingleLayout L = single_layout(ix);
    if (int rc = check_stride(L, record_stride)) return rc;
    if (n_check_failed) *n_check_failed = 0;
    if (n == 0) return SAQGPU_OK;
    IndexCall call(ix, (hipStream_t)stream);
    if (int rc_ = call.begin()) return rc_;
    hipStream_t st = (hipStream_t)stream;
    const bool dev = flags & SAQGPU_DEVICE_PTRS;
    if (dev && reinterpret_cast<uintptr_t>(records) % 8) return fail(SAQGPU_EINVAL, "records must be 8-byte aligned");
    DevBuf<float> d_data, d_O;
    DevBuf<uint8_t> d_rec;
    DevBuf<uint16_t> d_code;
    DevBuf<int> d_fail;
    const float *ddata = data;
    uint8_t *drec = records;
    if (!dev) {
        if (d_data.ensure(n * ix->num_dim) || d_rec.ensure(n * record_stride)) return fail(SAQGPU_ENOMEM, "allocation failed");
        HIPCHK(hipMemcpyAsync(d_data.p, data, n * ix->num_dim * 4, hipMemcpyHostToDevice, st));
        ddata = d_data.p;
        drec = d_rec.p;
    }
    // the reference writes only the fields a record uses; unused bytes read as zero here
    HIPCHK(hipMemsetAsync(drec, 0, n * record_stride, st));
    const int64_t chunk = std::min<int64_t>(n, 1 << 18);
    if (d_O.ensure(chunk * ix->max_dim) || d_code.ensure(chunk * ix->max_dim) || d_fail.ensure(1))
        return fail(SAQGPU_ENOMEM, "allocation failed");
    HIPCHK(hipMemsetAsync(d_fail.p, 0, sizeof(int), st));
    const IndexView v = ix->view();
    for (int64_t off = 0; off < n; off += chunk) {
        const int64_t cnt = std::min(chunk, n - off);
        for (int s = 0; s < ix->nseg; ++s)
            HIPCHK(launch_single_quantize_segment(
                v, ddata + off * ix->num_dim, ix->has_rot ? ix->rot.p + ix->rot_off[s] : nullptr, s, ix->segs[s].dim,
                cnt, chunk, cfg->caq_adj_rd_lmt, cfg->caq_adj_eps, cfg->caq_ori_qB, d_O.p, d_code.p, d_fail.p,
                drec + off * record_stride, record_stride, (int32_t)L.seg_code[s], (int32_t)L.seg_long[s],
                (int32_t)L.exf_off, st));
    }
    int nf = 0;
    HIPCHK(hipMemcpyAsync(&nf, d_fail.p, sizeof(int), hipMemcpyDeviceToHost, st));
    if (!dev) HIPCHK(hipMemcpyAsync(records, drec, n * record_stride, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (n_check_failed) *n_check_failed = nf;
    return SAQGPU_OK;
}

int saqgpu_single_estimate(saqgpu_index *ix, const float *queries, int64_t nq, const uint8_t *records,
                           int64_t nrec, int64_t record_stride, float vars_bound_m, int32_t dist_type,
                           float *out_vars, float *out_fast, float *out_acc, int32_t flags, void *stream) {
    if (!ix || nq < 0 || nrec < 0 || (nq > 0 && !queries) || (nrec > 0 && !records) || (nq * nrec > 0 && !out_acc))
        return fail(SAQGPU_EINVAL, "bad argument");
    if (dist_type != 0 && dist_type != 1) return fail(SAQGPU_EINVAL, "dist_type must be 0 (L2) or 1 (IP)");
    const SingleLayout L = single_layout(ix);
    if (int rc = check_stride(L, record_stride)) return rc;
    if (nq == 0 || nrec == 0) return SAQGPU_OK;
    if (nq * ((nrec + 15) / 16) > 0x7fffffffLL) return fail(SAQGPU_EUNSUPPORTED, "nq x nrec too large for one launch");
    IndexCall call(ix, (hipStream_t)stream);
    if (int rc_ = call.begin()) return rc_;
    hipStream_t st = (hipStream_t)stream;
    const bool dev = flags & SAQGPU_DEVICE_PTRS;
    if (dev && reinterpret_cast<uintptr_t>(records) % 8) return fail(SAQGPU_EINVAL, "records must be 8-byte aligned");
    const size_t smem = (size_t)ix->D_pad * 4 + (size_t)ix->D_pad / 64 * 64;
    int max_smem = 0;
    HIPCHK(hipDeviceGetAttribute(&max_smem, hipDeviceAttributeMaxSharedMemoryPerBlock, ix->device));
    if ((int64_t)smem > max_smem) return fail(SAQGPU_EUNSUPPORTED, "dimension needs more LDS than a workgroup has");
    std::vector<int32_t> soff(2 * ix->nseg);
    for (int s = 0; s < ix->nseg; ++s) {
        soff[2 * s] = (int32_t)L.seg_code[s];
        soff[2 * s + 1] = (int32_t)L.seg_long[s];
    }
    DevBuf<int32_t> d_soff;
    DevBuf<float> d_q, d_qrot, d_bound, d_out;
    DevBuf<uint8_t> d_rec, d_qs;
    DevBuf<uint64_t> d_pl;
    int rc = 0;
    if ((rc = upload(d_soff, soff))) return rc;
    if (d_qrot.ensure(nq * ix->D_pad) || d_bound.ensure(nq * ix->nseg) ||
        d_qs.ensure(nq * ix->nseg * single_query_state_bytes()) || d_pl.ensure(nq * (ix->D_pad / 64 * 8)))
        return fail(SAQGPU_ENOMEM, "workspace allocation failed");
    const float *dq = queries;
    const uint8_t *drec = records;
    float *ov = out_vars, *of = out_fast, *oa = out_acc;
    if (!dev) {
        if (d_q.ensure(nq * ix->num_dim) || d_rec.ensure(nrec * record_stride) || d_out.ensure(3 * nq * nrec))
            return fail(SAQGPU_ENOMEM, "allocation failed");
        HIPCHK(hipMemcpyAsync(d_q.p, queries, nq * ix->num_dim * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_rec.p, records, nrec * record_stride, hipMemcpyHostToDevice, st));
        dq = d_q.p;
        drec = d_rec.p;
        oa = d_out.p;
        of = out_fast ? d_out.p + nq * nrec : nullptr;
        ov = out_vars ? d_out.p + 2 * nq * nrec : nullptr;
    }
    // SaqEstimatorBase (saq_estimator.hpp:28-48): the rotated segment queries and their bounds
    HIPCHK(launch_query_prep(ix->view(), dq, nq, ix->variance.p, ix->has_rot ? ix->rot.p : nullptr,
                             ix->d_rot_off.p, vars_bound_m, d_qrot.p, d_bound.p, ix->max_dim, st));
    HIPCHK(launch_single_estimate(ix->view(), d_qrot.p, d_bound.p, nq, reinterpret_cast<SingleQ *>(d_qs.p), d_pl.p,
                                  drec, record_stride, nrec, d_soff.p, (int32_t)L.exf_off, dist_type, ov, of, oa,
                                  st));
    if (!dev) {
        HIPCHK(hipMemcpyAsync(out_acc, oa, nq * nrec * 4, hipMemcpyDeviceToHost, st));
        if (out_fast) HIPCHK(hipMemcpyAsync(out_fast, of, nq * nrec * 4, hipMemcpyDeviceToHost, st));
        if (out_vars) HIPCHK(hipMemcpyAsync(out_vars, ov, nq * nrec * 4, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    return SAQGPU_OK;
}

}  // extern "C"
