// saqlib_gpu.hpp -- the saqlib::IVF surface (saqlib/index/ivf.hpp:27-121) over the gfx950
// C ABI of libsaqgpu.so (include/saqgpu.h).
//
// A host program written against the reference's IVF class switches to the GPU by including
// this header and linking libsaqgpu.so: construction, search, the configuration structs and
// the metrics keep their names and argument meaning.  Errors that the reference turns into
// glog CHECK aborts are thrown as saqlib_gpu::Error carrying the library's message.
//
// Differences a caller can see:
//   * search/construct take plain row-major float pointers where the reference takes Eigen
//     matrices (FloatRowMat / RowVectorXf have exactly that memory layout);
//   * search_batch() searches many queries in one launch -- the GPU's unit of work; search()
//     is search_batch() with one query;
//   * the segment rotations P_s are the caller's (set_rotations) or drawn from a seeded
//     Gaussian + Gram-Schmidt in double; the reference draws them from std::rand through
//     Eigen's HouseholderQR (rotator.hpp), so its matrices are only reproduced by passing them.
#pragma once
#include <cmath>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "saqgpu.h"

namespace saqlib_gpu {

using PID = uint32_t;

enum class DistType { Any = 0, L2Sqr = 1, IP = 2 };  // defines.hpp:38-42

struct QuantSingleConfig {  // quantization/config.h:12-19 (CAQ only)
    bool random_rotation = true;
    int caq_adj_rd_lmt = 6;
    float caq_adj_eps = 1e-8f;
};

struct QuantizeConfig {  // quantization/config.h:21-27
    float avg_bits = 0;
    int seg_eqseg = 0;
    bool enable_segmentation = true;
    bool use_compact_layout = false;
    QuantSingleConfig single;
};

struct SearcherConfig {  // quantization/config.h:75-78
    float searcher_vars_bound_m = 4;
    DistType dist_type = DistType::L2Sqr;
};

struct QueryRuntimeMetrics {  // caq/caq_estimator.hpp:21-25
    uint64_t fast_bitsum = 0;
    uint64_t acc_bitsum = 0;
    uint64_t total_comp_cnt = 0;
};

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

inline void check(int rc, const char *where) {
    if (rc != SAQGPU_OK) throw Error(rc, std::string(where) + ": " + saqgpu_last_error());
}

class IVF {
  public:
    IVF(size_t n, size_t num_dim, size_t k, QuantizeConfig cfg, int device = 0)
        : num_data_(n), num_dim_(num_dim), num_cen_(k), cfg_(cfg), device_(device) {
        num_dim_padded_ = (num_dim + 63) / 64 * 64;
    }
    IVF(const IVF &) = delete;
    IVF &operator=(const IVF &) = delete;
    ~IVF() {
        if (h_) saqgpu_index_destroy(h_);
    }

    size_t num_data() const { return num_data_; }
    size_t num_dim() const { return num_dim_; }
    size_t k() const { return num_cen_; }
    const QuantizeConfig &get_config() const { return cfg_; }
    const std::vector<std::pair<int, int>> &quant_plan() const { return plan_; }
    saqgpu_index *handle() const { return h_; }

    // SaqDataMaker::set_variance (saq_data.hpp:85-94): zero-padded to the padded dimension.
    void set_variance(const float *vars, size_t n) {
        variance_.assign(num_dim_padded_, 0.0f);
        for (size_t i = 0; i < n && i < num_dim_padded_; ++i) variance_[i] = vars[i];
    }

    // Segment rotations, row-major D_s x D_s each, in plan order (call after the plan is known,
    // i.e. after set_variance; see make_plan()).
    void set_rotations(std::vector<std::vector<float>> rots) { rotations_ = std::move(rots); }

    // SaqDataMaker::analyze_plan (saq_data.hpp:119-227).
    const std::vector<std::pair<int, int>> &make_plan() {
        if (variance_.empty()) throw Error(SAQGPU_ESTATE, "make_plan: variance not set");
        int32_t dims[SAQGPU_MAX_SEG], bits[SAQGPU_MAX_SEG], ns = 0;
        check(saqgpu_make_plan(variance_.data(), (int32_t)num_dim_padded_, cfg_.avg_bits,
                               cfg_.enable_segmentation ? 1 : 0, cfg_.seg_eqseg, dims, bits, &ns),
              "make_plan");
        plan_.clear();
        for (int i = 0; i < ns; ++i) plan_.emplace_back(dims[i], bits[i]);
        return plan_;
    }

    // IVF::construct (ivf.hpp:130-184): data n x num_dim, centroids K x num_dim, row-major host
    // arrays.  Variance is computed from the data unless set_variance() was called.
    void construct(const float *data, const float *centroids, const PID *cluster_ids, uint64_t seed = 0) {
        if (variance_.empty()) compute_variance(data);
        if (plan_.empty()) make_plan();
        if (cfg_.single.random_rotation && rotations_.empty()) draw_rotations(seed);
        std::vector<int64_t> sizes(num_cen_, 0);
        for (size_t i = 0; i < num_data_; ++i) {
            if (cluster_ids[i] >= num_cen_) throw Error(SAQGPU_EINVAL, "construct: bad cluster id");
            ++sizes[cluster_ids[i]];
        }
        create(centroids, sizes);
        saqgpu_caq_cfg cc{cfg_.single.caq_adj_rd_lmt, cfg_.single.caq_adj_eps};
        int64_t nfail = 0;
        check(saqgpu_encode(h_, data, cluster_ids, (int64_t)num_data_, &cc, 0, &nfail, nullptr), "construct");
        check_failures_ = nfail;
    }

    // IVF::search (ivf.hpp:275-311) for one query of num_dim floats.
    void search(const float *query, size_t topk, size_t nprobe, SearcherConfig scfg, PID *results,
                QueryRuntimeMetrics *metrics = nullptr) const {
        search_batch(query, 1, topk, nprobe, scfg, results, nullptr, metrics);
    }

    // nq queries (row-major nq x num_dim) in one launch; results nq x topk, dists optional.
    void search_batch(const float *queries, size_t nq, size_t topk, size_t nprobe, SearcherConfig scfg,
                      PID *results, float *dists = nullptr, QueryRuntimeMetrics *metrics = nullptr) const {
        require_index("search");
        saqgpu_search_params p{(int32_t)topk, (int32_t)nprobe, scfg.searcher_vars_bound_m,
                               scfg.dist_type == DistType::IP ? 1 : 0};
        static_assert(sizeof(QueryRuntimeMetrics) == sizeof(saqgpu_metrics), "metrics layout");
        check(saqgpu_search(h_, queries, (int64_t)nq, nullptr, &p, results, dists,
                            reinterpret_cast<saqgpu_metrics *>(metrics), 0, nullptr),
              "search");
    }

    // IVF::estimate (ivf.hpp:313-362): every vector of the nprobe closest lists.
    void estimate(const float *query, size_t nprobe, SearcherConfig scfg,
                  std::vector<std::pair<PID, float>> &dist_list, std::vector<float> *fast_dist_list = nullptr,
                  std::vector<float> *vars_dist_list = nullptr) const {
        require_index("estimate");
        std::vector<uint32_t> probes(nprobe);
        const int dt = scfg.dist_type == DistType::IP ? 1 : 0;
        check(saqgpu_probe(h_, query, 1, (int32_t)nprobe, dt, probes.data(), 0, nullptr), "estimate");
        int64_t total = 0;
        for (uint32_t c : probes) total += sizes_[c];
        std::vector<uint32_t> ids(total);
        std::vector<float> acc(total), fast(total), vars(total);
        int64_t n = 0;
        check(saqgpu_estimate(h_, query, probes.data(), (int32_t)nprobe, scfg.searcher_vars_bound_m, dt, ids.data(),
                              acc.data(), fast.data(), vars.data(), &n),
              "estimate");
        dist_list.clear();
        for (int64_t i = 0; i < n; ++i) dist_list.emplace_back(ids[i], acc[i]);
        if (fast_dist_list) fast_dist_list->assign(fast.begin(), fast.begin() + n);
        if (vars_dist_list) vars_dist_list->assign(vars.begin(), vars.begin() + n);
    }

    // Vectors whose adjustment loop met the reference's CHECK_GE(ip, 0) (caq_encoder.hpp:67).
    int64_t check_failures() const { return check_failures_; }
    int64_t device_bytes() const { return h_ ? saqgpu_index_device_bytes(h_) : 0; }

    // Creates the device index from explicit parts (IVF::load's order of events: header,
    // plan, rotations, centroids, then clusters through set_cluster).
    void create(const float *centroids, const std::vector<int64_t> &cluster_sizes) {
        if (h_) {
            saqgpu_index_destroy(h_);
            h_ = nullptr;
        }
        if (plan_.empty()) throw Error(SAQGPU_ESTATE, "create: no plan");
        std::vector<int32_t> dims, bits;
        for (auto &[d, b] : plan_) {
            dims.push_back(d);
            bits.push_back(b);
        }
        std::vector<float> rot;
        if (!rotations_.empty()) {
            if (rotations_.size() != plan_.size()) throw Error(SAQGPU_EINVAL, "create: one rotation per segment");
            for (size_t s = 0; s < plan_.size(); ++s) {
                if (rotations_[s].size() != (size_t)plan_[s].first * plan_[s].first)
                    throw Error(SAQGPU_EINVAL, "create: rotation size");
                rot.insert(rot.end(), rotations_[s].begin(), rotations_[s].end());
            }
        }
        sizes_ = cluster_sizes;
        saqgpu_index_desc d{};
        d.num_data = (int64_t)num_data_;
        d.num_dim = (int32_t)num_dim_;
        d.num_cen = (int32_t)num_cen_;
        d.num_seg = (int32_t)plan_.size();
        d.use_compact_layout = cfg_.use_compact_layout ? 1 : 0;
        d.seg_dim = dims.data();
        d.seg_bits = bits.data();
        d.rotations = rot.empty() ? nullptr : rot.data();
        d.variance = variance_.data();
        d.centroids = centroids;
        d.cluster_sizes = sizes_.data();
        d.device = device_;
        check(saqgpu_index_create(&d, &h_), "create");
    }

    void set_plan(std::vector<std::pair<int, int>> plan) { plan_ = std::move(plan); }
    const std::vector<float> &variance() const { return variance_; }
    const std::vector<std::vector<float>> &rotations() const { return rotations_; }
    const std::vector<int64_t> &cluster_sizes() const { return sizes_; }

  private:
    void require_index(const char *what) const {
        if (!h_) throw Error(SAQGPU_ESTATE, std::string(what) + ": index not constructed");
    }

    // SaqDataMaker::compute_variance (saq_data.hpp:96-101): per-dimension population variance.
    // Accumulated in double here; the reference's fp32 Eigen column means can differ in the last
    // bit, which moves the plan only at exact DP ties -- pass set_variance() to pin it.
    void compute_variance(const float *data) {
        std::vector<double> mean(num_dim_, 0.0), var(num_dim_, 0.0);
        for (size_t i = 0; i < num_data_; ++i)
            for (size_t j = 0; j < num_dim_; ++j) mean[j] += data[i * num_dim_ + j];
        for (auto &m : mean) m /= (double)num_data_;
        for (size_t i = 0; i < num_data_; ++i)
            for (size_t j = 0; j < num_dim_; ++j) {
                const double t = data[i * num_dim_ + j] - mean[j];
                var[j] += t * t;
            }
        variance_.assign(num_dim_padded_, 0.0f);
        for (size_t j = 0; j < num_dim_; ++j) variance_[j] = (float)(var[j] / (double)num_data_);
    }

    // Random orthonormal D_s x D_s matrices (Gaussian rows, modified Gram-Schmidt in double).
    void draw_rotations(uint64_t seed) {
        rotations_.clear();
        for (size_t s = 0; s < plan_.size(); ++s) {
            const int D = plan_[s].first;
            std::mt19937_64 g(seed * 1000003ull + s);
            std::normal_distribution<double> nd;
            std::vector<double> m((size_t)D * D);
            for (auto &x : m) x = nd(g);
            for (int r = 0; r < D; ++r) {
                double *row = &m[(size_t)r * D];
                for (int p = 0; p < r; ++p) {
                    const double *q = &m[(size_t)p * D];
                    double dot = 0;
                    for (int c = 0; c < D; ++c) dot += row[c] * q[c];
                    for (int c = 0; c < D; ++c) row[c] -= dot * q[c];
                }
                double nrm = 0;
                for (int c = 0; c < D; ++c) nrm += row[c] * row[c];
                nrm = std::sqrt(nrm);
                for (int c = 0; c < D; ++c) row[c] /= nrm;
            }
            rotations_.emplace_back(m.begin(), m.end());
        }
    }

    size_t num_data_, num_dim_, num_cen_, num_dim_padded_;
    QuantizeConfig cfg_;
    int device_;
    saqgpu_index *h_ = nullptr;
    std::vector<float> variance_;
    std::vector<std::pair<int, int>> plan_;
    std::vector<std::vector<float>> rotations_;
    std::vector<int64_t> sizes_;
    int64_t check_failures_ = 0;
};

}  // namespace saqlib_gpu
