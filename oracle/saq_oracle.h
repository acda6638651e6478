/*
 * saq_oracle.h -- TEST INFRASTRUCTURE ONLY (the parity checker).
 *
 * CPU restatement of the reference SAQ hot path (howarlii/SAQ, saqlib/): the
 * IVF cluster scan (LUT fast scan -> accurate refine -> per-query top-k) and
 * the CAQ encoder (segment rotation -> scalar quantisation -> code adjustment
 * -> factors -> packing).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker; the
 * product path never calls it.
 *
 * Data structures use the reference's own in-memory layout (SaqCluData,
 * quantization/cluster_data.hpp:179-337) so the oracle reads exactly the bytes
 * the reference would.
 */
#ifndef SAQ_ORACLE_H
#define SAQ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_SEG 64

/* Segment plan (SaqData::quant_plan, quantization/saq_data.hpp:21-27). */
typedef struct {
    int32_t num_seg;
    int32_t num_dim;            /* original dimension (query length) */
    int32_t use_compact_layout; /* SaqCluData use_compact_layout flag */
    int32_t seg_dim[ORC_MAX_SEG];
    int32_t seg_bits[ORC_MAX_SEG];
} orc_plan_t;

/* One IVF list in the reference layout (SaqCluData, cluster_data.hpp:179-337). */
typedef struct {
    int64_t num_vec;
    const float *short_factors; /* multi-segment only: [num_blocks][sum_s 64] */
    const uint8_t *short_code;  /* 1 seg: [nb][256 + 4*D] ; multi: [nb][sum_{b_s>0} 4*D_s] */
    const uint8_t *long_code;   /* per SaqCluData long-code layout */
    const float *long_factors;  /* [num_vec][num_seg][2] = {rescale, error} */
    const uint32_t *ids;        /* [num_vec] */
    const float *centroid_rot;  /* [sum_s D_s] rotated centroid per segment */
} orc_cluster_t;

/* Mutable twin used by the encoder. */
typedef struct {
    int64_t num_vec;
    float *short_factors;
    uint8_t *short_code;
    uint8_t *long_code;
    float *long_factors;
    uint32_t *ids;
    float *centroid_rot;
} orc_cluster_out_t;

/* Sizes of the reference arrays of one cluster. */
typedef struct {
    int64_t num_blocks;
    int64_t short_factor_floats; /* total floats in short_factors */
    int64_t short_code_bytes;    /* total bytes in short_code */
    int64_t long_code_bytes;     /* total bytes in long_code (allocation) */
    int64_t long_stride;         /* non-compact: bytes per vector; compact: 0 */
    int64_t long_factor_floats;
    int64_t centroid_floats;
} orc_cluster_sizes_t;

void orc_cluster_sizes(const orc_plan_t *plan, int64_t num_vec, orc_cluster_sizes_t *out);

/* ---- reductions in the reference's order (Eigen 3.4 redux / GEMV, GCC reduce) ---- */
float orc_eigen_sum(const float *x, int64_t n);
float orc_eigen_sqnorm(const float *x, int64_t n);
float orc_eigen_dot(const float *a, const float *b, int64_t n);
float orc_eigen_diff_sqnorm(const float *a, const float *b, int64_t n);
float orc_eigen_var_q2(const float *var, const float *q, int64_t n);
void orc_gemv_rowvec(const float *q, const float *P, int64_t D, float *out);
float orc_tree16(const float *a);

/* ---- fast scan ---- */
void orc_pack_lut(int64_t D, const float *q, float *lut);
void orc_pack_codes(int64_t D, const uint8_t *short_codes, int64_t num, uint8_t *blocks);
typedef struct {
    float delta, sum_q, sum_vl_lut, q_l2sqr, q_l2norm;
} orc_lut_scalars_t;
void orc_lut_prepare(int64_t D, const float *r, uint16_t *lut16, orc_lut_scalars_t *sc);
void orc_accumulate(const uint8_t *block_codes, const uint16_t *lut16, int64_t D, uint32_t *S);
uint16_t orc_cvt_u16(float x);

/* ---- extra-bit codes ---- */
float orc_compute_ip(int ex_bits, const float *q, const uint8_t *code, int64_t D);
void orc_compact_code(int ex_bits, const uint16_t *raw, int64_t D, uint8_t *out);
int orc_code_at(int ex_bits, const uint8_t *code, int64_t D, int64_t d);

/* ---- encoder ---- */
typedef struct {
    int32_t rd_lmt;   /* caq_adj_rd_lmt */
    float eps;        /* caq_adj_eps */
    int32_t rotate;   /* random_rotation: P supplied */
    int32_t ori_qb;   /* caq_ori_qB: 0, or encode at this many bits and shift down */
} orc_caq_cfg_t;

typedef struct {
    float o_l2norm, ip_cent_oa, fac_rescale, fac_error;
    double ip_o_oa, oa_l2sqr, delta, v_mi;
    int32_t rounds;
    int32_t status; /* 0 ok, 1 CHECK_GE(ip, 0) failed */
} orc_code_info_t;

int orc_caq_encode(const float *o, int64_t D, int bits, const orc_caq_cfg_t *cfg,
                   const float *centroid, uint16_t *code, orc_code_info_t *info);
void orc_rotate_rows(const float *X, int64_t n, int64_t D, const float *P, const float *c_rot,
                     float *O);
int orc_quantize_segment(const orc_plan_t *plan, int seg, const float *X, int64_t n,
                         const float *centroid, const float *P, const orc_caq_cfg_t *cfg,
                         orc_cluster_out_t *clu, float *O_out);

/* ---- plan ---- */
int orc_plan_dp(const float *var, int64_t D_pad, float avg_bits, int32_t *dims, int32_t *bits);
int orc_plan_eqseg(int64_t D_pad, float avg_bits, int num_segs, int32_t *dims, int32_t *bits);

/* ---- search ---- */
typedef struct {
    uint64_t fast_bitsum, acc_bitsum, total_comp_cnt;
} orc_metrics_t;

void orc_probe(const float *centroids, int64_t K, int64_t D, const float *q, int64_t nprobe,
               int dist_type, uint32_t *probe, float *probe_dist);
int orc_search(const orc_plan_t *plan, const float *variance, const float *const *P,
               const orc_cluster_t *clusters, const uint32_t *probe, int64_t nprobe,
               const float *query, int64_t topk, float m, int dist_type, uint32_t *out_ids,
               float *out_dists, orc_metrics_t *metrics);
int orc_search_batch(const orc_plan_t *plan, const float *variance, const float *const *P,
                     const orc_cluster_t *clusters, const uint32_t *probes, int64_t nprobe,
                     const float *queries, int64_t nq, int64_t topk, float m, int dist_type,
                     int nthreads, uint32_t *out_ids, float *out_dists, orc_metrics_t *metrics);
int orc_search_batch_timed(const orc_plan_t *plan, const float *variance, const float *const *P,
                           const orc_cluster_t *clusters, const uint32_t *probes, int64_t nprobe,
                           const int32_t *npq, const float *queries, int64_t nq, int64_t topk, float m,
                           int nthreads, uint32_t *out_ids, float *out_dists, double *lat_ms);
int orc_estimate(const orc_plan_t *plan, const float *variance, const float *const *P,
                 const orc_cluster_t *clusters, const uint32_t *probe, int64_t nprobe,
                 const float *query, float m, int dist_type, uint32_t *out_ids, float *out_acc,
                 float *out_fast, float *out_vars, orc_metrics_t *metrics);

void orc_decode_cluster(const orc_plan_t *plan, const orc_cluster_t *c, int seg, int32_t *codes, float *o_l2norm,
                        float *rescale);
int64_t orc_estimate_lanes(const orc_plan_t *plan, const float *variance, const float *const *P,
                           const orc_cluster_t *clusters, const uint32_t *probe, int64_t nprobe, const float *query,
                           float m, float *out_vars, float *out_fast, float *out_acc);

/* ---- single-vector quantiser / estimator (SaqSingleDataWrapper records) ---- */
/* Record layout of SaqSingleDataWrapper (quantization/single_data.hpp:180-246): 2 short
 * factors per segment (o_l2norm, ip_cent_oa), padded to 64 bytes; the short codes of the
 * segments with bits > 0, padded to 64 bytes; the long codes; one ExFactor per segment.
 * bytes = SaqSingleDataWrapper::calculate_memory_size (:263-296). */
typedef struct {
    int64_t fac_off, code_off, long_off, exfac_off, bytes;
    int64_t seg_code_off[ORC_MAX_SEG]; /* -1 for a 0-bit segment */
    int64_t seg_long_off[ORC_MAX_SEG]; /* -1 below 2 bits */
} orc_single_layout_t;

void orc_single_layout(const orc_plan_t *plan, orc_single_layout_t *out);
/* utils::new_transpose_bin with b_query = 8: tq[8 * blk + p] holds bit p of q[64 blk + k] at
 * bit 63 - k. */
void orc_single_transpose_bin(const uint16_t *q, int64_t D, uint64_t *tq);
float orc_single_warmup_ip(const uint64_t *x, const uint64_t *tq, float delta, float vl, int64_t D);
float orc_single_mask_ip(const float *q, const uint64_t *x, int64_t D);
/* SAQuantizerSingle::quantize of n rows (num_dim floats each) into records of `stride` bytes
 * (zeroed first). Returns the number of vectors whose adjustment hit CHECK_GE(ip, 0). */
int64_t orc_single_quantize(const orc_plan_t *plan, const float *const *P, const float *data, int64_t n,
                            const orc_caq_cfg_t *cfg, uint8_t *records, int64_t stride);
/* SaqSingleEstimator<Any> for one query over nrec records: varsEstDist, compFastDist and
 * compAccurateDist of each record. */
void orc_single_estimate(const orc_plan_t *plan, const float *variance, const float *const *P,
                         const uint8_t *records, int64_t nrec, int64_t stride, const float *query,
                         float m, int dist_type, float *out_vars, float *out_fast, float *out_acc);

#ifdef __cplusplus
}
#endif
#endif
