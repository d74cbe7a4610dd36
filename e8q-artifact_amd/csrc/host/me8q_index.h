// me8q_index.h -- e8qlib::me8q::Me8qIndex with a device-resident index.
//
// Same public surface as the reference class (include/e8qlib/index/me8q.hpp:
// 31-136): BuildOptions, SearchStats, Build / Save / Load / SearchIVFFast /
// SearchIVFFastBatch and the shape getters, the same ME8QIDXX file and the
// same exceptions.  The plan (SAQ dynamic program or a manual spec) and the
// rotator signs are computed on the host as the reference does; the encode,
// the FastScan layout and every search run on the GPU through
// include/me8q_gpu.h.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "me8q_gpu.h"

struct e8q_ctx;
struct e8q_m8index;

namespace e8qlib::me8q {

constexpr size_t kSegDim = 64;      // me8q_plan.hpp:12
constexpr int kMaxBits = 13;        // me8q_plan.hpp:13
constexpr size_t kFactorBits = 64;  // me8q_plan.hpp:14
constexpr size_t kMaxDeviceK = 1024;

struct PlanSeg {
    size_t len = 0;
    int bits = 0;
};
using Plan = std::vector<PlanSeg>;

size_t PaddedDim(size_t dim);

struct PlanInput {
    const float* var = nullptr;
    size_t dim = 0;
    float avg_bits = 4.0F;
    int max_bits = kMaxBits;
    double extra_bits = 0.0;
};
// SAQ's DP (me8q_plan.cpp:47-167).
Plan Allocate(const PlanInput& in);
// "L:B,L:B,..." (me8q_common.cpp:37-75).
Plan ParsePlanSpec(const std::string& spec, size_t padded_dim);
// Per-dimension variance of the base rows (me8q_common.cpp:11-35).
void DatasetVariance(const float* base, size_t ld, size_t n, size_t dim, std::vector<float>* var);

struct BuildOptions {
    float avg_bits = 4.0F;
    std::string plan_spec;
    uint64_t rot_seed = 0x5EED1234ULL;
    int max_bits = kMaxBits;
    bool drop_energy = true;
    float prune_alpha = 0.0F;
    double extra_budget_bits = 0.0;
    size_t threads = 1;
};

struct SearchStats {
    size_t candidates = 0;
    size_t clusters = 0;
    double route_us = 0.0;
    double scan_us = 0.0;
    double rotate_us = 0.0;
    double lut_us = 0.0;
    double filter_us = 0.0;
    double lane_us = 0.0;
    double rest_us = 0.0;
    double refine_us = 0.0;
    size_t survivors = 0;
    size_t bits_scored = 0;
    size_t bits_total = 0;
    size_t pruned = 0;
    double device_scan_ms = 0.0;
    double device_total_ms = 0.0;
    size_t batches = 0;   // 32-vector scan records read
    size_t refined = 0;   // candidates that entered the exact cascade
    size_t groups = 0;    // cascade groups (up to four candidates)
    int form = 0;         // the device form that ran (E8Q_MODE_PLAIN / E8Q_MODE_BATCHED)
    size_t overflowed = 0;     // batched form: queries searched again by the plain form
    size_t pairs_scored = 0;   // batched form: (query, vector) pairs of the MFMA contraction
    size_t items = 0;          // batched form: its work items
    // batched form, the exact refine kernel per phase (0: first runs, 1:
    // cascades and the threshold pass): launch time, launches, algorithmic bytes
    double refine_ms[2] = {0.0, 0.0};
    size_t refine_launches[2] = {0, 0};
    size_t refine_bytes[2] = {0, 0};
};

class Me8qIndex {
   public:
    Me8qIndex();
    ~Me8qIndex();
    Me8qIndex(Me8qIndex&&) noexcept;
    Me8qIndex& operator=(Me8qIndex&&) noexcept;
    Me8qIndex(const Me8qIndex&) = delete;
    Me8qIndex& operator=(const Me8qIndex&) = delete;

    void Build(const float* base, size_t ld, const float* centroids, size_t ld_c, const uint32_t* cluster_ids,
               size_t n, size_t dim, size_t num_clusters, const BuildOptions& opt,
               const float* data_variance = nullptr);
    // Same, with the base rows (n x dim, dense) and a copy of cluster_ids in
    // device memory; centroids, cluster_ids and variance on the host.  The
    // plan needs data_variance (or opt.plan_spec): the rows are not read on
    // the host.
    void BuildFromDevice(const float* base_device, const float* centroids, const uint32_t* cluster_ids_host,
                         const uint32_t* cluster_ids_device, size_t n, size_t dim, size_t num_clusters,
                         const BuildOptions& opt, const float* data_variance);
    void Save(const char* path) const;
    void Load(const char* path);

    void SearchIVFFast(const float* query, size_t k, size_t nprobe, uint32_t* out_ids, SearchStats* stats = nullptr,
                       float* out_dists = nullptr) const;
    void SearchIVFFastBatch(const float* queries, size_t nq, size_t k, size_t nprobe, uint32_t* out_ids,
                            SearchStats* stats, float* out_dists, int num_threads) const;
    // The device entry: nprobe_per_query (nullable) overrides nprobe; queries
    // / outputs may be device pointers; per_query_stats (nullable) [nq*5];
    // mode: E8Q_MODE_PLAIN (the reference's schedule) or E8Q_MODE_BATCHED.
    void SearchDevice(const float* queries, size_t nq, size_t k, size_t nprobe, const uint32_t* nprobe_per_query,
                      uint32_t* out_ids, float* out_dists, SearchStats* stats, bool queries_on_device,
                      bool outputs_on_device, uint64_t* per_query_stats = nullptr, int mode = 0) const;

    // Search on a caller-owned device context (a serving worker's): no
    // process-wide lock.  Host queries and outputs, per-query nprobe.
    void SearchOnContext(e8q_ctx* ctx, const float* queries, size_t nq, size_t k, const uint32_t* nprobe_per_query,
                         uint32_t* out_ids, float* out_dists, int mode = 0, e8q_m8_stats* stats = nullptr) const;
    [[nodiscard]] uint64_t SearchWorkspaceBytes(size_t nq, size_t k, size_t max_nprobe, int mode = 0) const;

    // The full-estimate paths (me8q.hpp:85-139): SearchIVF scores every vector
    // of the nprobe nearest lists (me8q_search.cpp:136-236), ExhaustiveTopK
    // every stored vector (:238-266), both keeping PushCandidate's list;
    // Estimate one stored vector (me8q_measure.cpp:41-67; std::out_of_range
    // for an unknown id).  k <= 1024 on the device.
    void SearchIVF(const float* query, size_t k, size_t nprobe, uint32_t* out_ids, SearchStats* stats = nullptr,
                   float* out_dists = nullptr) const;
    void ExhaustiveTopK(const float* query, size_t k, std::vector<std::pair<float, uint32_t>>* out) const;
    [[nodiscard]] float Estimate(const float* query, uint32_t id) const;
    // Batch forms of the three (nq queries in one device call; stats6 per
    // query: candidates, clusters, bits_scored, bits_total, survivors, pruned).
    void SearchIVFBatch(const float* queries, size_t nq, size_t k, size_t nprobe, uint32_t* out_ids,
                        float* out_dists, uint64_t* stats6) const;
    void ExhaustiveTopKBatch(const float* queries, size_t nq, size_t k, uint32_t* out_ids, float* out_dists) const;
    void EstimateBatch(const float* queries, size_t nq, const uint32_t* qidx, const uint32_t* ids, size_t npair,
                       float* out) const;

    [[nodiscard]] size_t n() const { return n_; }
    [[nodiscard]] size_t dim() const { return dim_; }
    [[nodiscard]] size_t padded_dim() const { return padded_; }
    [[nodiscard]] size_t route_dim() const;
    [[nodiscard]] size_t num_clusters() const { return nlist_; }
    [[nodiscard]] const Plan& shared_plan() const { return plan_; }
    [[nodiscard]] double bits_per_dim() const;
    [[nodiscard]] size_t memory_bytes() const;
    [[nodiscard]] size_t device_bytes() const;
    [[nodiscard]] bool built() const { return dev_ != nullptr; }
    [[nodiscard]] float prune_alpha() const { return opt_.prune_alpha; }
    [[nodiscard]] e8q_m8index* device_index() const { return dev_; }

   private:
    void BuildImpl(const float* base, size_t ld, const float* centroids, size_t ld_c, const uint32_t* cluster_ids,
                   const uint32_t* cluster_ids_device, size_t n, size_t dim, size_t num_clusters,
                   const BuildOptions& opt, const float* data_variance);
    void CreateDevice();
    void Reset();
    size_t n_ = 0, dim_ = 0, padded_ = 0, nlist_ = 0, bytes_per_vector_ = 0;
    BuildOptions opt_;
    float ortho_alpha_ = 1.0F;
    Plan plan_;
    std::vector<std::vector<uint8_t>> run_signs_;
    std::vector<uint64_t> list_begin_;
    e8q_ctx* ctx_ = nullptr;
    e8q_m8index* dev_ = nullptr;
};

}  // namespace e8qlib::me8q
