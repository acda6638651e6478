// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Exposes, through a C ABI, the parts of the reference (howarlii/SAQ) whose
// sources compile on their own: the fast-scan kernels (quantization/fastscan/
// fastscan.hpp, fastscan_highacc.hpp), the extra-bit code helpers
// (utils/code_helper.hpp), the configuration structs that IVF::save writes raw
// (quantization/config.h, with the fmt headers the image ships) and the
// reference's vendored Eigen 3.4 (reductions, GEMV, GEMM).  Built by oracle/Makefile from the sources where they lie under
// the reference tree, with the reference's release flags, into oracle/_ref/.
// The headers that need glog/gflags (lut.hpp, caq_encoder.hpp, the searcher)
// are not built: the image has neither library.  Used only to pin the oracle.
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include <cstddef>

#include "defines.hpp"
#include "quantization/config.h"
#include "quantization/fastscan/fastscan.hpp"
#include "quantization/fastscan/fastscan_highacc.hpp"
#include "utils/code_helper.hpp"
#include "utils/space.hpp"

using namespace saqlib;

namespace {
template <class T>
T *amalloc(size_t n) {
    size_t bytes = (n * sizeof(T) + 63) / 64 * 64;
    if (bytes == 0) bytes = 64;
    T *p = static_cast<T *>(std::aligned_alloc(64, bytes));
    std::memset(p, 0, bytes);
    return p;
}
}  // namespace

extern "C" {

// Byte layout of the configuration structs IVF::save writes raw (saq_data.hpp:30,
// quantizer_data.hpp:29): sizes and field offsets.
void ref_config_layout(int64_t *out) {
    out[0] = sizeof(QuantizeConfig);
    out[1] = offsetof(QuantizeConfig, avg_bits);
    out[2] = offsetof(QuantizeConfig, seg_eqseg);
    out[3] = offsetof(QuantizeConfig, enable_segmentation);
    out[4] = offsetof(QuantizeConfig, use_compact_layout);
    out[5] = offsetof(QuantizeConfig, single);
    out[6] = sizeof(QuantSingleConfig);
    out[7] = offsetof(QuantSingleConfig, quant_type);
    out[8] = offsetof(QuantSingleConfig, random_rotation);
    out[9] = offsetof(QuantSingleConfig, use_fastscan);
    out[10] = offsetof(QuantSingleConfig, caq_adj_rd_lmt);
    out[11] = offsetof(QuantSingleConfig, caq_adj_eps);
    out[12] = offsetof(QuantSingleConfig, caq_ori_qB);
    out[13] = sizeof(BaseQuantType);
    const QuantizeConfig d{};
    out[14] = d.single.use_fastscan ? 1 : 0;
}

void ref_pack_codes(int64_t D, const uint8_t *codes, int64_t num, uint8_t *blocks) {
    fastscan::pack_codes(D, codes, num, blocks);
}

void ref_pack_lut(int64_t D, const float *q, float *lut) { fastscan::pack_lut<float>(D, q, lut); }

void ref_accumulate(const uint8_t *codes, const uint16_t *lut16, int64_t D, int32_t *S) {
    uint8_t *hc = amalloc<uint8_t>(D * 8);
    uint8_t *cd = amalloc<uint8_t>(D * 4);
    std::memcpy(cd, codes, D * 4);
    fastscan::transfer_lut_hacc(lut16, D, hc);
    __m512i res[2];
    fastscan::accumulate_hacc(cd, hc, res, D);
    _mm512_storeu_si512(S, res[0]);
    _mm512_storeu_si512(S + 16, res[1]);
    std::free(hc);
    std::free(cd);
}

float ref_compute_ip(int k, const float *q, const uint8_t *code, int64_t D) {
    float *qa = amalloc<float>(D);
    std::memcpy(qa, q, D * sizeof(float));
    float r = utils::get_IP_FUNC(k)(qa, code, D);
    std::free(qa);
    return r;
}

void ref_compact_code(int k, const uint16_t *raw, int64_t D, uint8_t *out) {
    utils::get_compacted_code16_func(k)(out, raw, D);
}

float ref_eigen_sum(const float *x, int64_t n) {
    FloatVec v = Eigen::Map<const FloatVec>(x, n);
    return v.sum();
}

float ref_eigen_sqnorm(const float *x, int64_t n) {
    FloatVec v = Eigen::Map<const FloatVec>(x, n);
    return v.squaredNorm();
}

float ref_eigen_dot_rowcol(const float *a, const float *b, int64_t n) {
    FloatVec ra = Eigen::Map<const FloatVec>(a, n);
    Eigen::VectorXf cb = Eigen::Map<const Eigen::VectorXf>(b, n);
    return ra.dot(cb);
}

float ref_eigen_diff_sqnorm(const float *rows, int64_t nrows, int64_t row, const float *q, int64_t n) {
    FloatRowMat m = Eigen::Map<const FloatRowMat>(rows, nrows, n);
    FloatVec qv = Eigen::Map<const FloatVec>(q, n);
    return (m.row(row) - qv).squaredNorm();
}

float ref_eigen_var_q2(const float *var, int64_t nvar, int64_t off, const float *q, int64_t n) {
    FloatVec vv = Eigen::Map<const FloatVec>(var, nvar);
    FloatVec qv = Eigen::Map<const FloatVec>(q, n);
    return (vv.segment(off, n).array() * qv.array().square()).sum();
}

// Column means and population variances of an n x d row-major matrix as Eigen 3.4 evaluates
// data.colwise().mean() and ((data.rowwise() - mean).array().square()).colwise().mean()
// (the expressions of SaqDataMaker::compute_variance and of IVF::construct's one-centroid mean).
void ref_eigen_colwise_mean_var(const float *X, int64_t n, int64_t d, float *mean, float *var) {
    FloatRowMat m = Eigen::Map<const FloatRowMat>(X, n, d);
    FloatVec mu = m.colwise().mean();
    FloatVec v = ((m.rowwise() - mu).array().square()).colwise().mean();
    std::memcpy(mean, mu.data(), d * sizeof(float));
    std::memcpy(var, v.data(), d * sizeof(float));
}

void ref_eigen_gemv(const float *q, const float *P, int64_t D, float *out) {
    FloatVec qv = Eigen::Map<const FloatVec>(q, D);
    FloatRowMat Pm = Eigen::Map<const FloatRowMat>(P, D, D);
    FloatVec r = qv * Pm;
    std::memcpy(out, r.data(), D * sizeof(float));
}

void ref_eigen_gemm(const float *X, int64_t n, int64_t D, const float *P, float *out) {
    FloatRowMat Xm = Eigen::Map<const FloatRowMat>(X, n, D);
    FloatRowMat Pm = Eigen::Map<const FloatRowMat>(P, D, D);
    FloatRowMat r = Xm * Pm;
    std::memcpy(out, r.data(), n * D * sizeof(float));
}

// The two fp32 element-wise expressions of the path, evaluated by Eigen under the reference's
// -Ofast flags, so the tests can count where fast-math changes a result: the u16 LUT
// quantisation of Lut::prepare and the floor/min code rounding of CAQEncoder::encode.
void ref_lut_quantize(const float *lut, int64_t T, float vl, float delta, uint16_t *out) {
    FloatVec lf = Eigen::Map<const FloatVec>(lut, T);
    Uint16Vec q(T);
    q = ((lf.array() - vl) / delta).cast<uint16_t>();
    std::memcpy(out, q.data(), T * sizeof(uint16_t));
}

float ref_lut_delta(float vl, float vr) { return (vr - vl) / ((1 << 16) - 0.01); }

void ref_code_floor(const float *o, int64_t D, double v_mi, double delta, int code_max, float *out) {
    FloatVec ov = Eigen::Map<const FloatVec>(o, D);
    Eigen::VectorXf cf = ((ov.array() - v_mi) / delta).floor().array().min(code_max);
    std::memcpy(out, cf.data(), D * sizeof(float));
}

// The single-vector estimator's bit-plane kernels (utils/space.hpp) and its query step
// (CaqEstimatorSingleImpl::prepare, caq/caq_estimator.hpp:300-301, with kNumBits = 8).
void ref_new_transpose_bin(const uint16_t *q, int64_t D, uint64_t *tq) { utils::new_transpose_bin(q, tq, D, 8); }

float ref_warmup_ip_x0_q(const uint64_t *x, const uint64_t *tq, float delta, float vl, int64_t D) {
    return utils::warmup_ip_x0_q(x, tq, delta, vl, D, 8);
}

float ref_mask_ip_x0_q(const float *q, const uint64_t *x, int64_t D) { return utils::mask_ip_x0_q(q, x, D); }

float ref_single_delta(float vl, float vr) { return (vr - vl) / ((1 << 8) - 0.01); }

}  // extern "C"
