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
#include <cstring>
#include <fstream>
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
    int caq_ori_qB = 0;  // codes found at this many bits, then shifted down (caq_encoder.hpp:144-169)
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
    // arrays.  Variance is computed from the data unless set_variance() was called; with
    // use_1_centroid every list is quantised against the data's column mean (ivf.hpp:163-176).
    void construct(const float *data, const float *centroids, const PID *cluster_ids, uint64_t seed = 0,
                   bool use_1_centroid = false) {
        if (variance_.empty()) compute_variance(data);
        if (plan_.empty()) make_plan();
        if (cfg_.single.random_rotation && rotations_.empty()) draw_rotations(seed);
        std::vector<int64_t> sizes(num_cen_, 0);
        for (size_t i = 0; i < num_data_; ++i) {
            if (cluster_ids[i] >= num_cen_) throw Error(SAQGPU_EINVAL, "construct: bad cluster id");
            ++sizes[cluster_ids[i]];
        }
        create(centroids, sizes);
        saqgpu_caq_cfg cc{cfg_.single.caq_adj_rd_lmt, cfg_.single.caq_adj_eps, cfg_.single.caq_ori_qB};
        int64_t nfail = 0;
        if (use_1_centroid) {
            std::vector<float> mean(num_dim_);
            check(saqgpu_column_mean_var(data, (int64_t)num_data_, (int32_t)num_dim_, mean.data(), nullptr),
                  "construct");
            check(saqgpu_encode_one_centroid(h_, data, cluster_ids, (int64_t)num_data_, mean.data(), &cc, 0, &nfail,
                                             nullptr),
                  "construct");
        } else {
            check(saqgpu_encode(h_, data, cluster_ids, (int64_t)num_data_, &cc, 0, &nfail, nullptr), "construct");
        }
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
        centroids_.assign(centroids, centroids + num_cen_ * num_dim_);
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

    // IVF::save (ivf.hpp:197-228): the reference's index file, readable by saqlib::IVF::load.
    void save(const char *path) const {
        require_index("save");
        std::ofstream out(path, std::ios::binary);
        if (!out) throw Error(SAQGPU_EINVAL, std::string("save: cannot open ") + path);
        const uint64_t hdr[3] = {num_data_, num_dim_, num_cen_};
        out.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
        out.write(reinterpret_cast<const char *>(centroids_.data()), centroids_.size() * 4);  // FlatInitializer::save
        // SaqData::save (saq_data.hpp:29-39)
        uint8_t cfgb[kCfgBytes];
        encode_cfg(cfgb);
        out.write(reinterpret_cast<const char *>(cfgb), kCfgBytes);
        put64(out, num_dim_);
        put64(out, variance_.size());
        out.write(reinterpret_cast<const char *>(variance_.data()), variance_.size() * 4);
        put64(out, plan_.size());
        for (size_t s = 0; s < plan_.size(); ++s) {  // BaseQuantizerData::save (quantizer_data.hpp:26-35)
            put64(out, (uint64_t)plan_[s].first);
            put64(out, (uint64_t)plan_[s].second);
            uint8_t scfg[kSingleCfgBytes];
            encode_single_cfg(scfg);
            out.write(reinterpret_cast<const char *>(scfg), kSingleCfgBytes);
            const char flag = rotations_.empty() ? 0 : 1;
            out.write(&flag, 1);
            if (flag) out.write(reinterpret_cast<const char *>(rotations_[s].data()), rotations_[s].size() * 4);
        }
        std::vector<uint64_t> sz(sizes_.begin(), sizes_.end());
        out.write(reinterpret_cast<const char *>(sz.data()), sz.size() * 8);
        for (size_t c = 0; c < num_cen_; ++c) {  // SaqCluData::save (cluster_data.hpp:327-336)
            saqgpu_cluster_sizes cs;
            check(saqgpu_cluster_sizes_of(h_, (int32_t)c, &cs), "save");
            std::vector<float> sf(std::max<int64_t>(1, cs.short_factor_floats)), lf(std::max<int64_t>(1, cs.long_factor_floats));
            std::vector<uint8_t> sc(std::max<int64_t>(1, cs.short_code_bytes)), lc(std::max<int64_t>(1, cs.long_code_bytes));
            std::vector<uint32_t> ids(std::max<int64_t>(1, sizes_[c]));
            std::vector<float> cen(cs.centroid_floats);
            check(saqgpu_index_get_cluster(h_, (int32_t)c, sf.data(), sc.data(), lc.data(), lf.data(), ids.data(),
                                           cen.data()),
                  "save");
            out.write(reinterpret_cast<const char *>(sf.data()), cs.short_factor_floats * 4);
            out.write(reinterpret_cast<const char *>(sc.data()), cs.short_code_bytes);
            out.write(reinterpret_cast<const char *>(lc.data()), file_long_bytes(c));
            out.write(reinterpret_cast<const char *>(lf.data()), cs.long_factor_floats * 4);
            out.write(reinterpret_cast<const char *>(ids.data()), sizes_[c] * 4);
            out.write(reinterpret_cast<const char *>(cen.data()), cen.size() * 4);
        }
        if (!out) throw Error(SAQGPU_EDEVICE, std::string("save: write failed on ") + path);
    }

    // IVF::load (ivf.hpp:230-264): an index file written by saqlib::IVF::save (or save above),
    // uploaded to the device.  Replaces the object's configuration, plan, rotations and data.
    void load(const char *path) {
        std::ifstream in(path, std::ios::binary);
        if (!in) throw Error(SAQGPU_EINVAL, std::string("load: cannot open ") + path);
        uint64_t hdr[3];
        read_exact(in, hdr, sizeof(hdr));
        num_data_ = hdr[0];
        num_dim_ = hdr[1];
        num_cen_ = hdr[2];
        num_dim_padded_ = (num_dim_ + 63) / 64 * 64;
        std::vector<float> cent(num_cen_ * num_dim_);
        read_exact(in, cent.data(), cent.size() * 4);
        uint8_t cfgb[kCfgBytes];
        read_exact(in, cfgb, kCfgBytes);
        decode_cfg(cfgb);
        const uint64_t saq_dim = get64(in);
        if (saq_dim != num_dim_) throw Error(SAQGPU_EINVAL, "load: SaqData dimension differs from the header's");
        const uint64_t nvar = get64(in);
        if (nvar > (1u << 24)) throw Error(SAQGPU_EINVAL, "load: bad variance length");
        std::vector<float> var(nvar);
        read_exact(in, var.data(), nvar * 4);
        variance_.assign(num_dim_padded_, 0.0f);
        std::copy(var.begin(), var.begin() + std::min<size_t>(nvar, num_dim_padded_), variance_.begin());
        const uint64_t nseg = get64(in);
        if (nseg == 0 || nseg > SAQGPU_MAX_SEG) throw Error(SAQGPU_EINVAL, "load: bad segment count");
        plan_.clear();
        rotations_.clear();
        bool any_rot = false, all_rot = true;
        std::vector<std::vector<float>> rots;
        for (uint64_t s = 0; s < nseg; ++s) {
            const uint64_t D = get64(in), b = get64(in);
            if (D == 0 || D % 64 || D > 65536 || b > 16) throw Error(SAQGPU_EINVAL, "load: bad segment");
            uint8_t scfg[kSingleCfgBytes];
            read_exact(in, scfg, kSingleCfgBytes);
            char flag = 0;
            read_exact(in, &flag, 1);
            plan_.emplace_back((int)D, (int)b);
            std::vector<float> P;
            if (flag) {
                P.resize(D * D);
                read_exact(in, P.data(), P.size() * 4);
                any_rot = true;
            } else {
                all_rot = false;
                P.assign(D * D, 0.0f);  // no rotator: the identity
                for (uint64_t i = 0; i < D; ++i) P[i * D + i] = 1.0f;
            }
            rots.push_back(std::move(P));
        }
        if (any_rot) rotations_ = std::move(rots);  // identity blocks stand in for segments without one
        cfg_.single.random_rotation = any_rot && all_rot;
        std::vector<uint64_t> sz(num_cen_);
        read_exact(in, sz.data(), sz.size() * 8);
        std::vector<int64_t> sizes(sz.begin(), sz.end());
        int64_t tot = 0;
        for (auto x : sizes) tot += x;
        if ((uint64_t)tot != num_data_) throw Error(SAQGPU_EINVAL, "load: cluster sizes do not add up to num_data");
        create(cent.data(), sizes);
        for (size_t c = 0; c < num_cen_; ++c) {
            saqgpu_cluster_sizes cs;
            check(saqgpu_cluster_sizes_of(h_, (int32_t)c, &cs), "load");
            std::vector<float> sf(std::max<int64_t>(1, cs.short_factor_floats)), lf(std::max<int64_t>(1, cs.long_factor_floats));
            std::vector<uint8_t> sc(std::max<int64_t>(1, cs.short_code_bytes)), lc(std::max<int64_t>(1, cs.long_code_bytes), 0);
            std::vector<uint32_t> ids(std::max<int64_t>(1, sizes[c]));
            std::vector<float> cen(cs.centroid_floats);
            read_exact(in, sf.data(), cs.short_factor_floats * 4);
            read_exact(in, sc.data(), cs.short_code_bytes);
            read_exact(in, lc.data(), file_long_bytes(c));
            read_exact(in, lf.data(), cs.long_factor_floats * 4);
            read_exact(in, ids.data(), sizes[c] * 4);
            read_exact(in, cen.data(), cen.size() * 4);
            check(saqgpu_index_set_cluster(h_, (int32_t)c, sf.data(), sc.data(), lc.data(), lf.data(), ids.data(),
                                           cen.data()),
                  "load");
        }
    }

    void set_plan(std::vector<std::pair<int, int>> plan) { plan_ = std::move(plan); }
    const std::vector<float> &variance() const { return variance_; }
    const std::vector<std::vector<float>> &rotations() const { return rotations_; }
    const std::vector<int64_t> &cluster_sizes() const { return sizes_; }

  private:
    // Byte layouts of QuantizeConfig / QuantSingleConfig as saqlib writes them raw
    // (quantization/config.h:12-27, x86-64): QuantSingleConfig = {int quant_type @0,
    // bool random_rotation @4, bool use_fastscan @5, int caq_adj_rd_lmt @8, float caq_adj_eps
    // @12, int caq_ori_qB @16} (20 bytes); QuantizeConfig = {float avg_bits @0, int seg_eqseg
    // @4, bool enable_segmentation @8, bool use_compact_layout @9, QuantSingleConfig @12}
    // (32 bytes).
    static constexpr int kCfgBytes = 32, kSingleCfgBytes = 20, kSingleOff = 12;

    void encode_single_cfg(uint8_t *b) const {
        std::memset(b, 0, kSingleCfgBytes);
        const int32_t caq = 0, ori_qb = cfg_.single.caq_ori_qB;
        const int32_t rd = cfg_.single.caq_adj_rd_lmt;
        const float eps = cfg_.single.caq_adj_eps;
        std::memcpy(b + 0, &caq, 4);
        b[4] = cfg_.single.random_rotation ? 1 : 0;
        b[5] = 1;  // use_fastscan
        std::memcpy(b + 8, &rd, 4);
        std::memcpy(b + 12, &eps, 4);
        std::memcpy(b + 16, &ori_qb, 4);
    }
    void encode_cfg(uint8_t *b) const {
        std::memset(b, 0, kCfgBytes);
        const float ab = cfg_.avg_bits;
        const int32_t eq = cfg_.seg_eqseg;
        std::memcpy(b + 0, &ab, 4);
        std::memcpy(b + 4, &eq, 4);
        b[8] = cfg_.enable_segmentation ? 1 : 0;
        b[9] = cfg_.use_compact_layout ? 1 : 0;
        encode_single_cfg(b + kSingleOff);
    }
    void decode_cfg(const uint8_t *b) {
        std::memcpy(&cfg_.avg_bits, b + 0, 4);
        std::memcpy(&cfg_.seg_eqseg, b + 4, 4);
        cfg_.enable_segmentation = b[8] != 0;
        cfg_.use_compact_layout = b[9] != 0;
        cfg_.single.random_rotation = b[kSingleOff + 4] != 0;
        std::memcpy(&cfg_.single.caq_adj_rd_lmt, b + kSingleOff + 8, 4);
        std::memcpy(&cfg_.single.caq_adj_eps, b + kSingleOff + 12, 4);
        std::memcpy(&cfg_.single.caq_ori_qB, b + kSingleOff + 16, 4);
    }
    // SaqCluData::save/load write longb_code_bytes_ * num_vec_ long-code bytes
    // (cluster_data.hpp:226-229, 320, 330): the padded per-vector stride in the per-vector
    // layout, the unpadded sum of the segments' code bytes in the compact one (the default for
    // a single segment).
    int64_t file_long_bytes(size_t c) const {
        const int64_t n = sizes_[c];
        const bool compact = cfg_.use_compact_layout || plan_.size() == 1;
        int64_t per = 0;
        for (auto &[D, b] : plan_) {
            const int64_t lb = b > 1 ? (int64_t)D * (b - 1) / 8 : 0;
            per += compact ? lb : (lb + 15) / 16 * 16;
        }
        return per * n;
    }
    static void put64(std::ofstream &o, uint64_t v) { o.write(reinterpret_cast<const char *>(&v), 8); }
    static uint64_t get64(std::ifstream &i) {
        uint64_t v = 0;
        read_exact(i, &v, 8);
        return v;
    }
    static void read_exact(std::ifstream &i, void *p, int64_t n) {
        if (n <= 0) return;
        i.read(reinterpret_cast<char *>(p), n);
        if (i.gcount() != n) throw Error(SAQGPU_EINVAL, "load: file truncated");
    }

    void require_index(const char *what) const {
        if (!h_) throw Error(SAQGPU_ESTATE, std::string(what) + ": index not constructed");
    }

    // SaqDataMaker::compute_variance (saq_data.hpp:96-101): per-dimension population variance in
    // fp32, summed in Eigen's column-wise order (saqgpu_column_mean_var).
    void compute_variance(const float *data) {
        std::vector<float> var(num_dim_);
        check(saqgpu_column_mean_var(data, (int64_t)num_data_, (int32_t)num_dim_, nullptr, var.data()),
              "compute_variance");
        variance_.assign(num_dim_padded_, 0.0f);
        std::copy(var.begin(), var.end(), variance_.begin());
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
    std::vector<float> centroids_;
    int64_t check_failures_ = 0;
};

}  // namespace saqlib_gpu
