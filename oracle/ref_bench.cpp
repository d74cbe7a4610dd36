// ref_bench.cpp -- times the reference e8qlib CPU search (built from the
// reference sources by oracle/ref.mk) on an E8AX index file, with the
// reference binding's threading (OpenMP over queries, schedule(dynamic, 1),
// one QueryScratch per thread: e8q_bindings.cpp:119-144) and the reference
// timing protocol (3 warm-ups, then rounds until >= 9 rounds and >= the
// minimum time; median round QPS: experiments/protocol.py:559-597).
//
//   ref_bench INDEX.e8ax|INDEX.me8q QUERIES.f32 NQ K NPROBE THREADS MIN_SECONDS OUT_IDS.u32
// An ME8QIDXX file is searched with Me8qIndex::SearchIVFFastBatch, the
// reference binding's entry for nq > 1 (me8q_bindings.cpp:133-142).
// QUERIES.f32 / OUT_IDS.u32 are raw little-endian arrays [NQ][dim] / [NQ][K].
// Prints one JSON object on stdout.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <cstring>
#include <functional>

#include "e8qlib/index/e8q.hpp"
#include "e8qlib/index/me8q.hpp"

int main(int argc, char** argv) {
    if (argc != 9) {
        std::fprintf(stderr, "usage: ref_bench INDEX QUERIES.f32 NQ K NPROBE THREADS MIN_SECONDS OUT_IDS\n");
        return 2;
    }
    char magic[8] = {};
    {
        FILE* mf = std::fopen(argv[1], "rb");
        if (!mf || std::fread(magic, 1, 8, mf) != 8) {
            std::fprintf(stderr, "cannot read index\n");
            return 1;
        }
        std::fclose(mf);
    }
    const bool me8q = std::memcmp(magic, "ME8QIDXX", 8) == 0;
    e8qlib::E8qIndex idx;
    e8qlib::me8q::Me8qIndex midx;
    const auto tl = std::chrono::steady_clock::now();
    if (me8q) midx.Load(argv[1]);
    else idx.Load(argv[1]);
    const double load_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - tl).count();
    const size_t nq = std::strtoull(argv[3], nullptr, 10);
    const size_t k = std::strtoull(argv[4], nullptr, 10);
    const size_t nprobe = std::strtoull(argv[5], nullptr, 10);
    const int threads = std::atoi(argv[6]);
    const double min_s = std::atof(argv[7]);
    const size_t dim = me8q ? midx.dim() : idx.dim();
    std::vector<float> q(nq * dim);
    FILE* f = std::fopen(argv[2], "rb");
    if (!f || std::fread(q.data(), 4, q.size(), f) != q.size()) {
        std::fprintf(stderr, "cannot read queries\n");
        return 1;
    }
    std::fclose(f);
    std::vector<uint32_t> ids(nq * k);
    std::vector<float> dists(nq * k);
    auto round = [&]() {
        const auto t0 = std::chrono::steady_clock::now();
        if (me8q) {
            midx.SearchIVFFastBatch(q.data(), nq, k, nprobe, ids.data(), nullptr, dists.data(), threads);
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
#pragma omp parallel num_threads(threads)
        {
            e8qlib::QueryScratch scratch;
#pragma omp for schedule(dynamic, 1)
            for (long long i = 0; i < static_cast<long long>(nq); ++i)
                idx.Search(q.data() + static_cast<size_t>(i) * dim, k, nprobe, ids.data() + static_cast<size_t>(i) * k,
                           scratch, nullptr, dists.data() + static_cast<size_t>(i) * k);
        }
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    for (int w = 0; w < 3; ++w) round();
    std::vector<double> qps;
    double total = 0.0;
    while (qps.size() < 9 || total < min_s) {
        const double s = round();
        total += s;
        qps.push_back(static_cast<double>(nq) / s);
    }
    std::sort(qps.begin(), qps.end());
    f = std::fopen(argv[8], "wb");
    if (f) {
        std::fwrite(ids.data(), 4, ids.size(), f);
        std::fclose(f);
    }
    std::printf("{\"qps_median\": %.3f, \"qps_p25\": %.3f, \"qps_p75\": %.3f, \"rounds\": %zu, \"seconds\": %.3f, "
                "\"threads\": %d, \"nq\": %zu, \"load_seconds\": %.3f}\n",
                qps[qps.size() / 2], qps[qps.size() / 4], qps[(3 * qps.size()) / 4], qps.size(), total, threads, nq,
                load_s);
    return 0;
}
