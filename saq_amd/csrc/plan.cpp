// plan.cpp -- SAQ segment plan (SaqDataMaker::analyze_plan / equal_segmentation /
// dynamic_programming, saqlib/quantization/saq_data.hpp:119-227).  Host code, run once
// per index before encoding.
#include <algorithm>
#include <cstdint>
#include <limits>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/saqgpu.h"

namespace {

constexpr int64_t kPad = 64;
constexpr int kMaxBits = 13;
constexpr size_t kFactorBits = 2 * sizeof(float) * 8;  // kNumShortFactors floats per segment

int64_t rd_up(int64_t x, int64_t y) { return (x + y - 1) / y * y; }

// float32 sum of 64 contiguous 64-byte-aligned values as Eigen reduces them: packets 0+2 and
// 1+3 accumulate separately, combine, then the 16 -> 8 -> 4 -> 2 -> 1 tree.
float chunk_sum64(const float *x) {
    float p0[16], p1[16];
    for (int l = 0; l < 16; ++l) {
        p0[l] = x[l] + x[32 + l];
        p1[l] = x[16 + l] + x[48 + l];
    }
    for (int l = 0; l < 16; ++l) p0[l] = p0[l] + p1[l];
    float b[8], c[4];
    for (int i = 0; i < 8; ++i) b[i] = p0[i] + p0[i + 8];
    for (int i = 0; i < 4; ++i) c[i] = b[i] + b[i + 4];
    return (c[0] + c[2]) + (c[1] + c[3]);
}

int equal_plan(int64_t D_pad, float avg_bits, int num_segs, int32_t *dims, int32_t *bits) {
    const int b = (int)(size_t)avg_bits;
    int64_t d = 0;
    int n = 0;
    for (int i = 0; i < num_segs && d < D_pad; ++i) {
        const int64_t t = rd_up((D_pad - d) / (num_segs - i), kPad);
        dims[n] = (int32_t)t;
        bits[n] = b;
        ++n;
        d += t;
    }
    return n;
}

int dp_plan(const float *var, int64_t D_pad, float avg_bits, int32_t *dims, int32_t *bits) {
    const size_t tot = (size_t)(avg_bits * (float)D_pad + (float)kFactorBits);
    const size_t chunks = D_pad / kPad;
    const size_t maxseg = avg_bits < 2 ? chunks : chunks / 2;
    const double lim = std::numeric_limits<double>::max();
    const size_t NB = tot + 1, NI = chunks + 1;
    std::vector<double> f((maxseg + 1) * NI * NB, lim);
    std::vector<uint32_t> from((maxseg + 1) * NI * NB, 0);
    auto at = [&](size_t ns, size_t i, size_t b) { return (ns * NI + i) * NB + b; };
    std::vector<double> cs(chunks);
    for (size_t c = 0; c < chunks; ++c) cs[c] = chunk_sum64(var + c * kPad);
    size_t ans_ns = 0, ans_i = chunks, ans_b = tot;
    f[at(0, 0, 0)] = 0;
    for (size_t ns = 0; ns <= maxseg; ++ns)
        for (size_t i = 0; i <= chunks; ++i)
            for (size_t ub = 0; ub <= tot; ++ub) {
                const double cur = f[at(ns, i, ub)];
                if (!(cur < lim)) continue;
                if (i == chunks) {
                    if (cur * 1.01 < f[at(ans_ns, ans_i, ans_b)]) {
                        ans_ns = ns;
                        ans_i = i;
                        ans_b = ub;
                    }
                    continue;
                }
                if (ns == maxseg) continue;
                double vs = 0;
                for (size_t j = 1; i + j <= chunks; ++j) {
                    vs += cs[i + j - 1];
                    for (int b = 1; b <= kMaxBits; ++b) {
                        const size_t bn = ub + b * j * kPad + kFactorBits;
                        if (bn > tot) break;
                        const double v = vs / (1 << b);
                        const size_t t = at(ns + 1, i + j, bn);
                        if (f[t] > cur + v) {
                            f[t] = cur + v;
                            from[t] = (uint32_t)((i << 4) + b);
                        }
                    }
                }
                const size_t t0 = at(ns + 1, chunks, ub);
                if (cur + vs < f[t0]) {  // the rest of the dims as one 0-bit segment
                    f[t0] = cur + vs;
                    from[t0] = (uint32_t)(i << 4);
                }
            }
    std::vector<std::pair<int32_t, int32_t>> plan;
    size_t ns = ans_ns, i = ans_i, B = ans_b;
    while (i > 0) {
        const uint32_t sec = from[at(ns, i, B)];
        const size_t pi = sec >> 4, cb = sec & 0xF, dl = (i - pi) * kPad;
        plan.emplace_back((int32_t)dl, (int32_t)cb);
        --ns;
        i = pi;
        if (cb) B -= cb * dl + kFactorBits;
    }
    std::reverse(plan.begin(), plan.end());
    for (size_t k = 0; k < plan.size(); ++k) {
        dims[k] = plan[k].first;
        bits[k] = plan[k].second;
    }
    return (int)plan.size();
}

}  // namespace

extern "C" int saqgpu_make_plan(const float *variance, int32_t num_dim, float avg_bits, int32_t enable_segmentation,
                                int32_t seg_eqseg, int32_t *dims, int32_t *bits, int32_t *num_seg) {
    if (!variance || !dims || !bits || !num_seg || num_dim < 1 || avg_bits < 0) return SAQGPU_EINVAL;
    const int64_t D_pad = rd_up(num_dim, kPad);
    std::vector<float> v(D_pad, 0.0f);
    std::copy(variance, variance + num_dim, v.begin());
    int n;
    if (enable_segmentation)
        n = seg_eqseg > 0 ? equal_plan(D_pad, avg_bits, seg_eqseg, dims, bits)
                          : dp_plan(v.data(), D_pad, avg_bits, dims, bits);
    else
        n = equal_plan(D_pad, avg_bits, 1, dims, bits);
    if (n < 1 || n > SAQGPU_MAX_SEG) return SAQGPU_EUNSUPPORTED;
    *num_seg = n;
    return SAQGPU_OK;
}

namespace {

// One column's sum in the order Eigen 3.4 reduces a row-major matrix column-wise: columns inside
// the 16-float packets take packetwise_redux_impl's order (PartialReduxEvaluator.h:100-121: the
// first row, then groups of 4 rows added as (r1 + r2) + (r3 + r4), then the rows left over);
// the columns after the last whole packet take the scalar redux, row after row.
template <class F>
float column_sum(int64_t n, bool packet, F elem) {
    float p = elem(0);
    int64_t i = 1;
    if (packet) {
        const int64_t size4 = (n - 1) & ~int64_t(3);
        for (; i < size4; i += 4) p = p + ((elem(i) + elem(i + 1)) + (elem(i + 2) + elem(i + 3)));
    }
    for (; i < n; ++i) p = p + elem(i);
    return p;
}

}  // namespace

extern "C" int saqgpu_column_mean_var(const float *data, int64_t n, int32_t d, float *mean, float *var) {
    if (!data || n < 1 || d < 1 || (!mean && !var)) return SAQGPU_EINVAL;
    const int32_t packet_cols = d / 16 * 16;
    // the release build's -freciprocal-math multiplies by 1/n where the source divides
    const float rn = 1.0f / (float)n;
    std::vector<float> mu(d);
    auto work = [&](int32_t j0, int32_t j1) {
        for (int32_t j = j0; j < j1; ++j) {
            const bool pk = j < packet_cols;
            const float m = column_sum(n, pk, [&](int64_t i) { return data[i * d + j]; }) * rn;
            mu[j] = m;
            if (mean) mean[j] = m;
            if (var)
                var[j] = column_sum(n, pk, [&](int64_t i) {
                             const float x = data[i * d + j] - m;
                             return x * x;
                         }) * rn;
        }
    };
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(16, d), n * (int64_t)d >> 20));
    std::vector<std::thread> th;
    const int32_t per = (d + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int32_t j0 = t * per, j1 = std::min<int32_t>(d, j0 + per);
        if (j0 < j1) th.emplace_back(work, j0, j1);
    }
    for (auto &t : th) t.join();
    return SAQGPU_OK;
}
