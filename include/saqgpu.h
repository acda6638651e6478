/*
 * saqgpu.h -- C ABI of the MI355X (gfx950) SAQ hot path.
 *
 * The reference (howarlii/SAQ) is a header-only C++ library with no FFI: its
 * natural seams are IVF::search / SAQSearcher::searchCluster for the scan and
 * IVF::construct / SAQuantizer::quantize_cluster for the encoder.  Every entry
 * point below replaces one of those seams; the host-side saqlib mirror
 * (saq_amd/host/saqlib_gpu.hpp) binds them exactly as a maintainer would bind
 * them from the reference's own IVF class.
 *
 * Conventions: plain C types, no exceptions cross the boundary, every call
 * returns an saqgpu_status (0 = ok) and saqgpu_last_error() describes the last
 * failure of the calling thread.  Pointers flagged SAQGPU_DEVICE_PTRS are HIP
 * device pointers on the index's device; otherwise host memory.
 */
#ifndef SAQGPU_H
#define SAQGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    SAQGPU_OK = 0,
    SAQGPU_EINVAL = 1,      /* bad argument (the reference CHECK-aborts on these) */
    SAQGPU_EDEVICE = 2,     /* HIP runtime error */
    SAQGPU_ENOMEM = 3,      /* device allocation failed */
    SAQGPU_ESTATE = 4,      /* index not in the right state (e.g. cluster not set) */
    SAQGPU_EUNSUPPORTED = 5 /* configuration outside what the kernels implement */
} saqgpu_status;

#define SAQGPU_MAX_SEG 64
/* Segments with bits > 0 hold at most this many 64-dim code groups in all (4096 quantised
 * dimensions); saqgpu_index_create returns SAQGPU_EUNSUPPORTED beyond it. */
#define SAQGPU_MAX_GROUPS 64
#define SAQGPU_MAX_TOPK 1024
#define SAQGPU_DEVICE_PTRS 1
/* saqgpu_search_batched only, with SAQGPU_DEVICE_PTRS and a non-null stream: the call's device work
 * (coarse quantiser, rotation, grouping, pair prep, tile kernel, merge) is captured into a HIP graph,
 * which later calls with the same shape, parameters and buffers replay as one graph launch (the
 * buffers' contents may change between calls).  The first call of a new shape runs eagerly and
 * captures; the index keeps one graph. */
#define SAQGPU_USE_GRAPH 2

typedef struct saqgpu_index saqgpu_index;

/* Index description: the persistent part of IVF + SaqData
 * (reference: saqlib/index/ivf.hpp:27-121, saqlib/quantization/saq_data.hpp:20-57). */
typedef struct {
    int64_t num_data;            /* IVF::num_data_ */
    int32_t num_dim;             /* IVF::num_dim_ (original dimension) */
    int32_t num_cen;             /* IVF::num_cen_ (K) */
    int32_t num_seg;             /* SaqData::quant_plan.size() */
    int32_t use_compact_layout;  /* QuantizeConfig::use_compact_layout */
    const int32_t *seg_dim;      /* quant_plan[i].first  (multiples of 64) */
    const int32_t *seg_bits;     /* quant_plan[i].second (0..13) */
    const float *rotations;      /* concatenated P_s, row-major D_s x D_s; NULL = no rotation */
    const float *variance;       /* SaqData::data_variance, padded dimension floats */
    const float *centroids;      /* FlatInitializer centroids, K x num_dim */
    const int64_t *cluster_sizes;/* K entries */
    int32_t device;              /* HIP device ordinal */
} saqgpu_index_desc;

/* Sizes of one cluster's arrays in the reference SaqCluData layout
 * (saqlib/quantization/cluster_data.hpp:206-294, 317-336). */
typedef struct {
    int64_t num_blocks;
    int64_t short_factor_floats;
    int64_t short_code_bytes;
    int64_t long_code_bytes;
    int64_t long_factor_floats;
    int64_t centroid_floats;
} saqgpu_cluster_sizes;

/* CAQ encoder configuration (QuantSingleConfig, saqlib/quantization/config.h:12-19). */
typedef struct {
    int32_t caq_adj_rd_lmt; /* adjustment round limit (0 disables adjustment) */
    float caq_adj_eps;      /* adjustment epsilon */
    int32_t caq_ori_qB;     /* 0, or: codes found at this many bits (above every segment's, at most
                               16) and shifted down to the segment's (CAQEncoder::downUpSample,
                               caq_encoder.hpp:144-157, 165-169, 217-219) */
} saqgpu_caq_cfg;

/* Search parameters (IVF::search arguments + SearcherConfig, ivf.hpp:91-94, config.h:75-78). */
typedef struct {
    int32_t topk;
    int32_t nprobe;    /* lists per query (the maximum, and the probe-list stride, when
                          nprobe_per_query is set) */
    float searcher_vars_bound_m;
    int32_t dist_type; /* 0 = L2Sqr, 1 = IP (reproduces the reference's behaviour) */
    const int32_t *nprobe_per_query; /* NULL, or nq entries in 1..nprobe, in the queries' memory
                                        kind: query i scans the first entries of its list, as
                                        IVF::search(query_i, topk, nprobe_per_query[i], ...) */
} saqgpu_search_params;

/* Per-query counters of QueryRuntimeMetrics (saqlib/quantization/caq/caq_estimator.hpp:21-25). */
typedef struct {
    uint64_t fast_bitsum;
    uint64_t acc_bitsum;
    uint64_t total_comp_cnt;
} saqgpu_metrics;

const char *saqgpu_last_error(void);

/* Replaces IVF(n, dim, K, cfg) + SaqDataMaker/prepare_quantizers + IVF::load's header part
 * (ivf.hpp:68-72, 230-255).  Allocates every device array of the index. */
int saqgpu_index_create(const saqgpu_index_desc *desc, saqgpu_index **out);
void saqgpu_index_destroy(saqgpu_index *idx);

/* A search handle for concurrent callers (IVF::search called from many host threads, as
 * src/test_qps.cpp:77-84 does): a handle reads idx's device data but has its own workspace,
 * lock, device-order event and library handles, so calls on different handles -- each on its own
 * stream -- run concurrently, while calls on one handle serialise as on an index.  Every search,
 * probe and estimate entry point accepts a handle in place of the index (the batched ones where
 * idx built its batched layout before the handle was opened); entry points that change the index
 * refuse a handle, and refuse idx while any handle is open.  Destroy handles with
 * saqgpu_index_destroy before idx. */
int saqgpu_index_search_handle(saqgpu_index *idx, saqgpu_index **out);

int saqgpu_cluster_sizes_of(const saqgpu_index *idx, int32_t cid, saqgpu_cluster_sizes *out);

/* Uploads one cluster given in the reference layout (SaqCluData::load, cluster_data.hpp:317-326);
 * the library transposes it into its device layout.  Host pointers. */
int saqgpu_index_set_cluster(saqgpu_index *idx, int32_t cid, const float *short_factors,
                             const uint8_t *short_code, const uint8_t *long_code,
                             const float *long_factors, const uint32_t *ids,
                             const float *centroid_rot);

/* Downloads one cluster back into the reference layout (SaqCluData::save, :327-336). */
int saqgpu_index_get_cluster(const saqgpu_index *idx, int32_t cid, float *short_factors,
                             uint8_t *short_code, uint8_t *long_code, float *long_factors,
                             uint32_t *ids, float *centroid_rot);

/* Replaces IVF::construct's quantisation stage (ivf.hpp:149-183 -> SAQuantizer::quantize_cluster,
 * saq_quantizer.hpp:34-60): encodes data (n x num_dim) with its cluster assignment into the
 * device-resident index.  data/cluster_ids are host pointers unless flags has
 * SAQGPU_DEVICE_PTRS.  Returns the number of vectors whose adjustment loop hit the
 * reference's CHECK_GE(ip, 0) in *n_check_failed (may be NULL). */
int saqgpu_encode(saqgpu_index *idx, const float *data, const uint32_t *cluster_ids, int64_t n,
                  const saqgpu_caq_cfg *cfg, int32_t flags, int64_t *n_check_failed, void *stream);

/* IVF::construct with use_1_centroid (ivf.hpp:163-176): as saqgpu_encode, but every list is
 * quantised against the one centroid `centroid` (num_dim floats, host memory; the reference
 * passes data.colwise().mean(), see saqgpu_column_mean_var): every list's rotated centroid
 * becomes that centroid's.  The lists' own centroids still select the probed lists. */
int saqgpu_encode_one_centroid(saqgpu_index *idx, const float *data, const uint32_t *cluster_ids, int64_t n,
                               const float *centroid, const saqgpu_caq_cfg *cfg, int32_t flags,
                               int64_t *n_check_failed, void *stream);

/* Encodes rotated residuals directly (the per-vector CAQEncoder::encode_and_fac + packing,
 * caq_encoder.hpp:222-247, cluster_packer.hpp:59-79) for one segment of one cluster whose
 * rotated centroid is already set: o is num_vec x D_s (host).  Used to pin the encoder
 * independently of the rotation order. */
int saqgpu_encode_residuals(saqgpu_index *idx, int32_t cid, int32_t seg, const float *o,
                            const float *centroid_seg, const saqgpu_caq_cfg *cfg);

/* Replaces FlatInitializer::centroids_distances (index/initializer.hpp:52-76) for a batch:
 * probes[q * nprobe + i] = i-th closest centroid of query q. */
int saqgpu_probe(saqgpu_index *idx, const float *queries, int64_t nq, int32_t nprobe,
                 int32_t dist_type, uint32_t *probes, int32_t flags, void *stream);

/* No counterpart in the reference: saqgpu_probe for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * allocates nothing: queries (nq x num_dim) and probes (nq x nprobe) are device pointers on the
 * index's device, the work is queued on stream, and the (query, centroid) distances go to one
 * fixed device array per GPU of SAQGPU_NOALLOC_SCRATCH_FLOATS floats, held until the process
 * ends.  nq * (num_cen + 1) beyond it is refused (SAQGPU_EINVAL): split the queries.  The lists equal
 * saqgpu_probe's.  Every call on a GPU shares that GPU's array, so calls that overlap on one GPU
 * (other indexes, other threads or streams) race; calls on one index are ordered as usual. */
#define SAQGPU_NOALLOC_SCRATCH_FLOATS (1 << 24)
int saqgpu_probe_noalloc(saqgpu_index *idx, const float *queries, int64_t nq, int32_t nprobe, int32_t dist_type,
                         uint32_t *probes, void *stream);

/* Replaces IVF::search (ivf.hpp:275-311) for a batch of queries (nq x num_dim).  probes may be
 * NULL (the GPU coarse quantiser picks them) or nq x nprobe cluster ids.  out_ids is
 * nq x topk (unfilled = 0, like ResultPool::copy_results); out_dists and metrics may be NULL.
 * Returns when the results are written (also with SAQGPU_DEVICE_PTRS). */
int saqgpu_search(saqgpu_index *idx, const float *queries, int64_t nq, const uint32_t *probes,
                  const saqgpu_search_params *params, uint32_t *out_ids, float *out_dists,
                  saqgpu_metrics *metrics, int32_t flags, void *stream);

/* Replaces IVF::estimate (ivf.hpp:313-362) for one query (host pointers): accurate, fast and
 * variance estimates of every vector of the probed clusters, in probe order; returns the count
 * through *n_out.  probes may be NULL (the GPU coarse quantiser in the reference's order); the
 * outputs must hold the summed sizes of the probed clusters; out_fast/out_vars may be NULL. */
int saqgpu_estimate(saqgpu_index *idx, const float *query, const uint32_t *probes, int32_t nprobe,
                    float vars_bound_m, int32_t dist_type, uint32_t *out_ids, float *out_acc,
                    float *out_fast, float *out_vars, int64_t *n_out);

/* ---- single-vector records (SaqSingleDataWrapper, quantization/single_data.hpp:156-320) -----
 * One record per vector, independent of the lists: per segment {o_l2norm, ip_cent_oa} padded to
 * 64 bytes, the short codes of segments with bits > 0 (MSB-first, bytes of every 8 reversed)
 * padded to 64 bytes, the CodeHelper<b-1> long codes, one ExFactor per segment.  The index
 * supplies the SaqData (plan, rotations, variance); its lists need not be set.  Records of a
 * batch sit record_stride bytes apart (>= the record size, a multiple of 8) from an 8-byte
 * aligned base. */

/* SaqSingleDataWrapper::calculate_memory_size (single_data.hpp:287-320) for the index's plan.
 * offsets (may be NULL) receives 3 + 2 * nseg int64: short-code section, long-code section,
 * ExFactor section, then per segment its short-code and long-code offset (-1 where none). */
int saqgpu_single_layout(const saqgpu_index *idx, int64_t *record_bytes, int64_t *offsets);

/* Replaces SAQuantizerSingle::quantize (quantization/saq_quantizer.hpp:78-94 ->
 * QuantizerSingle::quantize, quantizer.hpp:105-207) for n rows of data (n x num_dim): each row
 * becomes one record.  Host pointers unless flags has SAQGPU_DEVICE_PTRS.  Returns the number of
 * vectors whose adjustment loop hit CHECK_GE(ip, 0) in *n_check_failed (may be NULL). */
int saqgpu_single_quantize(saqgpu_index *idx, const float *data, int64_t n, const saqgpu_caq_cfg *cfg,
                           uint8_t *records, int64_t record_stride, int32_t flags, int64_t *n_check_failed,
                           void *stream);

/* Replaces SaqSingleEstimator<Any> (quantization/saq_estimator.hpp:248-294 over
 * CaqSingleEstimator, caq/caq_estimator.hpp:220-369, 450-497) for nq queries (nq x num_dim) and
 * nrec records: out_vars / out_fast / out_acc[q * nrec + r] = varsEstDist / compFastDist /
 * compAccurateDist of record r for query q.  out_vars and out_fast may be NULL.  Host pointers
 * unless flags has SAQGPU_DEVICE_PTRS; returns when the results are written. */
int saqgpu_single_estimate(saqgpu_index *idx, const float *queries, int64_t nq, const uint8_t *records,
                           int64_t nrec, int64_t record_stride, float vars_bound_m, int32_t dist_type,
                           float *out_vars, float *out_fast, float *out_acc, int32_t flags, void *stream);

/* SaqDataMaker::analyze_plan (saqlib/quantization/saq_data.hpp:119-227): segment plan from the
 * per-dimension variance (num_dim floats).  dims/bits receive up to SAQGPU_MAX_SEG entries. */
int saqgpu_make_plan(const float *variance, int32_t num_dim, float avg_bits, int32_t enable_segmentation,
                     int32_t seg_eqseg, int32_t *dims, int32_t *bits, int32_t *num_seg);

/* Column means and population variances of an n x d row-major host matrix, in fp32 (host code):
 * the data.colwise().mean() of IVF::construct with use_1_centroid (ivf.hpp:163-167) and the two
 * column-wise means of SaqDataMaker::compute_variance (saq_data.hpp:96-101), summed in Eigen
 * 3.4's order for a row-major matrix.  The means equal the reference's release build bit for
 * bit; its -Ofast build reassociates the variance sums, which here keep Eigen's source order.
 * mean or var may be NULL. */
int saqgpu_column_mean_var(const float *data, int64_t n, int32_t d, float *mean, float *var);

/* ---- batched scan (no single counterpart in the reference) ----------------------------
 * Serves a batch of queries as one grouped job instead of nq independent IVF::search calls:
 * (query, list) pairs are grouped by list and each list's codes are contracted against a tile
 * of its queries on MFMA.  Every vector of every probed list gets the accurate estimate
 * of caq_estimator.hpp:190-217 (no fast-scan pruning), computed from the full codes with the
 * residual query quantised per (query, list, segment) to int8 (one term), int8x2 (two
 * terms) or bf16; the per-query top-k is exact for those estimates, ties broken by id.  IDs therefore
 * differ from IVF::search's only where the estimates' quantisation reorders near-ties. */
#define SAQGPU_PREC_INT8 1
#define SAQGPU_PREC_INT8X2 2
/* BF16: the residual per (query, list, segment) as r / max|r| in bf16 (an fp32 scale), contracted
 * with the codes on v_mfma_f32_16x16x32_bf16 with fp32 accumulation. */
#define SAQGPU_PREC_BF16 3

typedef struct {
    int32_t topk;                     /* 1..256 */
    int32_t nprobe;                   /* lists per query (the maximum when nprobe_per_query is set) */
    const int32_t *nprobe_per_query;  /* NULL, or nq entries in 1..nprobe, in the queries' memory kind */
    int32_t precision;                /* SAQGPU_PREC_INT8, SAQGPU_PREC_INT8X2 or SAQGPU_PREC_BF16 */
    int32_t query_tile;               /* queries per work tile: 16, 32, 64 or 128 (a workgroup of 1, 2, 4
                                         or 4 waves reading each code chunk once for them); 0 = the
                                         autotuned choice, halved while the tile's per-query lists of
                                         topk pairs do not fit in a workgroup's LDS (an explicit tile
                                         that does not fit is SAQGPU_EUNSUPPORTED) */
    int32_t exact_probe;              /* with probes == NULL: 1 = the reference's coarse quantiser order
                                         (as saqgpu_probe), 0 = an fp32 GEMM's |c|^2 - 2 q.c (probe lists
                                         may differ where two centroid distances tie within fp32 rounding) */
} saqgpu_batched_params;

/* Builds the batched code layout (full codes in MFMA operand order plus per-vector factors)
 * from the index's clusters; search_batched calls it when needed.  Re-run after clusters change. */
int saqgpu_prepare_batched(saqgpu_index *idx);

/* Batched search: probes as in saqgpu_search (NULL = GPU coarse quantiser, nprobe lists per
 * query); out_ids nq x topk (unfilled = 0), out_dists optional.  Segments of more than 2048
 * dimensions and indexes of more than 16384 lists are not supported (SAQGPU_EUNSUPPORTED). */
int saqgpu_search_batched(saqgpu_index *idx, const float *queries, int64_t nq, const uint32_t *probes,
                          const saqgpu_batched_params *params, uint32_t *out_ids, float *out_dists,
                          int32_t flags, void *stream);

/* The project's autotuner for the batched scan: times every query tile (16, 32, 64, 128) on this
 * batch and keeps the fastest as the index's choice for query_tile = 0; returns it. */
int saqgpu_autotune_batched(saqgpu_index *idx, const float *queries, int64_t nq, const uint32_t *probes,
                            const saqgpu_batched_params *params, int32_t flags, void *stream,
                            int32_t *best_query_tile);

/* Kernel timing for measurement: enable = 1 starts recording a HIP event pair around every scan
 * kernel launch (on the stream it is launched on), enable = 0 stops.  When scan_ms/launches are
 * non-NULL the recorded pairs are synchronised and their total returned.  No counterpart in the
 * reference (its test_qps times whole queries with wall clocks). */
int saqgpu_timing(saqgpu_index *idx, int32_t enable, double *scan_ms, int64_t *launches);

/* Device memory held by the index (bytes), for capacity planning. */
int64_t saqgpu_index_device_bytes(const saqgpu_index *idx);

#ifdef __cplusplus
}
#endif
#endif
