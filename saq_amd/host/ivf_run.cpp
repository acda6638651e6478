// ivf_run.cpp -- builds and searches an index through the saqlib_gpu::IVF host API, the way a
// program written against saqlib::IVF (create_index.cpp / test_qps.cpp) would.
//
//   ivf_run <input.bin> <output.bin> [--save <index file>] [--load <index file>]
//
// --save writes the constructed index in the reference's file format (IVF::save);
// --load skips construction and searches an index read from such a file (IVF::load).
//
// input.bin (little endian): int64 n, d, K, nq, topk, nprobe, nseg, rd_lmt, then nseg pairs
// (int64 dim, int64 bits), float32 eps, float32 avg_bits, then float32 base[n*d],
// centroids[K*d], variance[d], queries[nq*d], int32 has_rotations, rotations[sum dim^2] (only
// when has_rotations), uint32 cluster_ids[n].
// output.bin: uint32 ids[nq*topk], float32 dists[nq*topk], uint64 metrics[nq*3],
// int64 check_failures.
#include <cstdio>
#include <cstring>
#include <vector>

#include "saqlib_gpu.hpp"

namespace {

template <class T>
bool rd(FILE *f, T *p, size_t n) {
    return n == 0 || fread(p, sizeof(T), n, f) == n;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3 && argc != 5) {
        fprintf(stderr, "usage: %s input.bin output.bin [--save|--load index_file]\n", argv[0]);
        return 2;
    }
    const char *save_path = nullptr, *load_path = nullptr;
    if (argc == 5) {
        if (std::strcmp(argv[3], "--save") == 0)
            save_path = argv[4];
        else if (std::strcmp(argv[3], "--load") == 0)
            load_path = argv[4];
        else {
            fprintf(stderr, "unknown option %s\n", argv[3]);
            return 2;
        }
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) {
        perror("input");
        return 2;
    }
    int64_t h[8];
    if (!rd(f, h, 8)) return 3;
    const int64_t n = h[0], d = h[1], K = h[2], nq = h[3], topk = h[4], nprobe = h[5], nseg = h[6];
    std::vector<std::pair<int, int>> plan;
    for (int64_t s = 0; s < nseg; ++s) {
        int64_t p[2];
        if (!rd(f, p, 2)) return 3;
        plan.emplace_back((int)p[0], (int)p[1]);
    }
    float eps_bits[2];
    if (!rd(f, eps_bits, 2)) return 3;
    std::vector<float> base(n * d), cent(K * d), var(d), q(nq * d);
    if (!rd(f, base.data(), base.size()) || !rd(f, cent.data(), cent.size()) || !rd(f, var.data(), var.size()) ||
        !rd(f, q.data(), q.size()))
        return 3;
    int32_t has_rot = 0;
    if (!rd(f, &has_rot, 1)) return 3;
    std::vector<std::vector<float>> rots;
    if (has_rot) {
        for (auto &[dim, b] : plan) {
            std::vector<float> r((size_t)dim * dim);
            if (!rd(f, r.data(), r.size())) return 3;
            rots.push_back(std::move(r));
        }
    }
    std::vector<uint32_t> cids(n);
    if (!rd(f, cids.data(), cids.size())) return 3;
    fclose(f);

    saqlib_gpu::QuantizeConfig cfg;
    cfg.avg_bits = eps_bits[1];
    cfg.single.caq_adj_rd_lmt = (int)h[7];
    cfg.single.caq_adj_eps = eps_bits[0];
    cfg.single.random_rotation = has_rot != 0;
    std::vector<uint32_t> ids(nq * topk);
    std::vector<float> dists(nq * topk);
    std::vector<saqlib_gpu::QueryRuntimeMetrics> met(nq);
    int64_t fails = 0;
    try {
        saqlib_gpu::IVF ivf((size_t)n, (size_t)d, (size_t)K, cfg);
        if (load_path) {
            ivf.load(load_path);
            if (ivf.quant_plan() != plan) {
                fprintf(stderr, "loaded plan differs from the input's\n");
                return 4;
            }
        } else {
            ivf.set_variance(var.data(), var.size());
            const auto &auto_plan = ivf.make_plan();
            if (auto_plan != plan) {
                // the caller's plan wins (a loaded index carries its own); report the difference
                fprintf(stderr, "note: DP plan differs from the given plan; using the given plan\n");
                ivf.set_plan(plan);
            }
            if (has_rot) ivf.set_rotations(rots);
            ivf.construct(base.data(), cent.data(), cids.data());
            fails = ivf.check_failures();
            if (save_path) ivf.save(save_path);
        }
        saqlib_gpu::SearcherConfig scfg;
        // one query through search(), the rest as one batch
        ivf.search(q.data(), (size_t)topk, (size_t)nprobe, scfg, ids.data(), &met[0]);
        if (nq > 1)
            ivf.search_batch(q.data() + d, (size_t)nq - 1, (size_t)topk, (size_t)nprobe, scfg, ids.data() + topk,
                             dists.data() + topk, met.data() + 1);
        // distances of query 0 from a batch of one
        std::vector<uint32_t> i0(topk);
        ivf.search_batch(q.data(), 1, (size_t)topk, (size_t)nprobe, scfg, i0.data(), dists.data(), nullptr);
        if (std::memcmp(i0.data(), ids.data(), topk * 4) != 0) {
            fprintf(stderr, "search() and search_batch() disagree on query 0\n");
            return 4;
        }
    } catch (const saqlib_gpu::Error &e) {
        fprintf(stderr, "saqlib_gpu error %d: %s\n", e.code(), e.what());
        return 5;
    }
    FILE *o = fopen(argv[2], "wb");
    if (!o) {
        perror("output");
        return 2;
    }
    fwrite(ids.data(), 4, ids.size(), o);
    fwrite(dists.data(), 4, dists.size(), o);
    fwrite(met.data(), sizeof(saqlib_gpu::QueryRuntimeMetrics), met.size(), o);
    fwrite(&fails, 8, 1, o);
    fclose(o);
    printf("ivf_run: %lld queries searched\n", (long long)nq);
    return 0;
}
