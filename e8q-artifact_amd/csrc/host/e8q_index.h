// e8q_index.h -- e8qlib::E8qIndex with a device-resident index.
//
// Same public surface as the reference class (include/e8qlib/index/e8q.hpp:
// 121-187): Build / Save / Load / Search / ExhaustiveTopK / SetDisablePrune /
// SetPruneAlpha and the shape getters, the same E8AX file, the same
// exceptions (std::invalid_argument for bad arguments, std::runtime_error for
// I/O and device failures).  Build and Search run on the GPU through the C ABI
// of libe8q_gpu.so (include/e8q_gpu.h); SearchBatch is the many-query entry
// the binding uses.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "e8q_gpu.h"

struct e8q_ctx;
struct e8q_dindex;

namespace e8qlib {

constexpr size_t kMaxK = 128;

struct E8qBuildOptions {
    float prune_alpha = 0.0F;   // per-query prune-loss budget (0 = 1.0)
};

struct SearchStats {
    size_t candidates = 0;      // points scanned by the coarse pass
    size_t survivors = 0;       // lanes that entered exact refinement
    double rotate_us = 0.0;
    double prepare_route_us = 0.0;
    double scan_us = 0.0;
    double refine_us = 0.0;
    size_t batches = 0;         // 32-vector scan records of the probed clusters
    size_t scanned = 0;         // of those, records whose codes were read
    double device_scan_ms = 0.0;   // scan kernel, HIP events on the device stream
    double device_total_ms = 0.0;  // whole search on the device stream
};

// Kept for signature compatibility with the reference's per-thread scratch;
// the device path keeps its scratch on the device context.
class QueryScratch {
   public:
    bool disable_prune = false;
};

// kMfma: the unpruned estimator over every probed vector as a split-BF16
// MFMA contraction (bits <= 4); distances within its rounding of the fp32
// refine's.
enum class SearchMode { kPlain = 0, kBatched = 1, kMfma = 2 };

class E8qIndex {
   public:
    E8qIndex();
    ~E8qIndex();
    E8qIndex(E8qIndex&&) noexcept;
    E8qIndex& operator=(E8qIndex&&) noexcept;
    E8qIndex(const E8qIndex&) = delete;
    E8qIndex& operator=(const E8qIndex&) = delete;

    void Build(const float* base, const float* centroids, const uint32_t* cluster_ids, size_t n, size_t dim,
               size_t num_clusters, int bits, uint64_t seed, size_t threads, const E8qBuildOptions& options = {});
    // Same, with base/centroids/cluster_ids already in device memory.
    void BuildFromDevice(const float* base, const float* centroids, const uint32_t* cluster_ids_host,
                         const uint32_t* cluster_ids_device, size_t n, size_t dim, size_t num_clusters, int bits,
                         uint64_t seed, const E8qBuildOptions& options = {});

    void Save(const char* path) const;
    void Load(const char* path);

    void Search(const float* query, size_t k, size_t nprobe, uint32_t* out_ids, QueryScratch& scratch,
                SearchStats* stats = nullptr, float* out_dists = nullptr) const;
    // nq queries in one device call; nprobe_per_query (nullable) overrides
    // nprobe per query.  Outputs nq*k, sorted by (dist, id), sentinel padded.
    void SearchBatch(const float* queries, size_t nq, size_t k, size_t nprobe, const uint32_t* nprobe_per_query,
                     uint32_t* out_ids, float* out_dists, SearchStats* stats, SearchMode mode,
                     bool queries_on_device = false, bool outputs_on_device = false) const;
    // The plain-form search on a caller-owned device context (a serving
    // worker's): no process-wide lock, so calls on different contexts run
    // concurrently.  Host queries and outputs, per-query nprobe.
    void SearchOnContext(e8q_ctx* ctx, const float* queries, size_t nq, size_t k, const uint32_t* nprobe_per_query,
                         uint32_t* out_ids, float* out_dists, e8q_search_stats* stats = nullptr) const;
    // Device scratch such a search carves from its context's workspace.
    [[nodiscard]] uint64_t SearchWorkspaceBytes(size_t nq, size_t k, size_t max_nprobe) const;
    // Top-k of the refine estimator over every stored vector (no pruning),
    // ordered by (distance, id) (e8q_oracle.cpp:14-56).
    void ExhaustiveTopK(const float* query, size_t k, std::vector<std::pair<float, uint32_t>>& out) const;
    // The estimator distance of stored vector `id` (e8q_oracle.cpp:58-101);
    // std::invalid_argument for an id no vector carries.
    [[nodiscard]] float Estimate(const float* query, uint32_t id) const;
    // (d_coarse - d_full) / (r_q Omega) of stored vector `id` from its build
    // row (e8q_oracle.cpp:103-219).
    [[nodiscard]] double ProbeBoundW(const float* query, const float* base_row, uint32_t id) const;
    // The binding's batch form: pair i = (queries[qidx[i]], base_rows[i],
    // ids[i]); NaN for qidx >= nq or id >= n (e8q_bindings.cpp:181-234).
    void ProbeBoundWBatch(const float* queries, size_t nq, const float* base_rows, const uint32_t* qidx,
                          const uint32_t* ids, size_t npair, double* out) const;

    void SetDisablePrune(bool on) { disable_prune_ = on; }
    [[nodiscard]] bool disable_prune() const { return disable_prune_; }
    void SetPruneAlpha(float alpha) { alpha_ = alpha; }
    [[nodiscard]] float prune_alpha() const { return alpha_; }

    [[nodiscard]] size_t n() const { return n_; }
    [[nodiscard]] size_t dim() const { return dim_; }
    [[nodiscard]] size_t padded_dim() const { return padded_; }
    [[nodiscard]] int bits() const { return bits_; }
    [[nodiscard]] size_t num_clusters() const { return nlist_; }
    [[nodiscard]] size_t memory_bytes() const;
    [[nodiscard]] size_t device_bytes() const;
    [[nodiscard]] bool built() const { return dev_ != nullptr; }
    [[nodiscard]] const std::vector<uint64_t>& cluster_begin() const { return cluster_begin_; }

   private:
    void CreateDevice(const std::vector<uint8_t>& signs);
    void Reset();
    size_t n_ = 0, dim_ = 0, padded_ = 0, nlist_ = 0;
    int bits_ = 0;
    float alpha_ = 1.0F;
    bool disable_prune_ = false;
    std::vector<uint8_t> signs_;
    std::vector<uint64_t> cluster_begin_, arena_begin_, ref_arena_begin_;
    e8q_ctx* ctx_ = nullptr;
    e8q_dindex* dev_ = nullptr;
};

// Process-wide device context (one per process; device from E8Q_DEVICE, default 0).
e8q_ctx* DefaultContext();
// Serializes the calls made on that context (one host thread at a time).
std::mutex& DeviceMutex();

// Launch options of that context: graph replay of repeated search shapes,
// and the batched form's scan shape (all zero: the context's autotuner).
void SetSearchGraphs(bool on);
void SetBatchedShape(uint32_t qtile, uint32_t btile, uint32_t grid);
struct ContextInfo {
    uint32_t qtile = 0, btile = 0, grid = 0;   // shape of the last batched search
    bool graphs = false;
    uint64_t tunings = 0, captures = 0, replays = 0;
    // the batched ME8Q form: grids of its last search, autotuner runs
    uint32_t m8_mfma_grid = 0, m8_refine_grid = 0;
    uint64_t m8_tunings = 0;
};
ContextInfo GetContextInfo();
// The batched ME8Q form's persistent grids (MFMA contraction, exact refine;
// zero: the context's autotuner).
void SetMe8qGrids(uint32_t mfma_grid, uint32_t refine_grid);

}  // namespace e8qlib
