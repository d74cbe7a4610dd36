// e8q_kernels.h -- kernel launchers (host-callable) of the E8Q device library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "e8q_device.h"

namespace e8qdev {

// Largest survivor flush size: the reference's staging capacity
// (kSurvCapMax, e8q_internal.hpp:33).
constexpr int kMaxFlush = 1024;

struct PlainArgs {
    uint64_t nq;
    const float* rotq;        // [nq][padded] rotated, centered queries
    const uint8_t* lut;       // [nq][blocks*32] compact u8 LUT
    const float* consts;      // [nq][8]: delta, base, lut_err, score_max, qn2, U (max |S| of any code)
    const uint32_t* probe_ids;  // [nq][maxp]
    const float* probe_d2;      // [nq][maxp]
    uint32_t maxp;
    const uint32_t* nprobe_q;   // optional [nq]
    uint32_t nprobe_all;
    uint64_t k;
    float z;
    int disable_prune;
    int flush_every;
    uint32_t* out_ids;          // [nq][k]
    float* out_dists;
    unsigned long long* stats;  // [4] candidates, survivors, batches, records read
    const uint32_t* qsel;       // optional: search only queries qsel[0..nsel)
    uint64_t nsel;
    uint32_t max_vectors;       // nonzero (a multiple of 32): scan at most this many vectors of each list
};

hipError_t LaunchRotate(hipStream_t st, const float* x, uint64_t rows, uint32_t dim, const IndexView& ix, float* out,
                        bool center);
hipError_t LaunchQueryPrep(hipStream_t st, const float* rotq, uint64_t nq, const IndexView& ix, uint8_t* lut,
                           float* consts);
hipError_t LaunchRoute(hipStream_t st, const float* rotq, const float* consts, uint64_t nq, const IndexView& ix,
                       float* dist);
// Route + top-nprobe of the E8Q index (LaunchRouteSelect over its route arguments).
hipError_t LaunchRouteProbes(hipStream_t st, const float* rotq, const float* consts, uint64_t nq, const IndexView& ix,
                             float* dist, const uint32_t* nprobe_q, uint32_t nprobe_all, uint32_t maxp,
                             uint32_t* probe_ids, float* probe_d2);
// Route distances over the first `dim` coordinates of rows `cstride` apart
// (ivf_route.cpp:9-27; ME8Q routes over its plan's coded prefix).
struct RouteArgs {
    const float* rotq;        // nq rows, qstride floats apart
    uint32_t qstride;
    const float* qn2;         // |q|^2 over the prefix, qn2_stride floats apart
    uint32_t qn2_stride;
    const float* cent;        // rotated centroids
    const float* cn2;         // their squared norms
    uint64_t nlist;
    uint32_t dim, cstride;
    uint64_t nq;
    float* dist;              // [nq][nlist]
};
hipError_t LaunchRouteDist(hipStream_t st, const RouteArgs& ra);
// Route distances and the top-nprobe per query in one step: the two-stage
// selection (approximate distances by an fp32 MFMA contraction, the
// reference-order distances of the candidates its error bound leaves) where
// the shape allows, else LaunchRouteDist + LaunchSelectProbes; the same probe
// lists and distances either way.  ra.dist is scratch [nq][nlist].
hipError_t LaunchRouteSelect(hipStream_t st, const RouteArgs& ra, const uint32_t* nprobe_q, uint32_t nprobe_all,
                             uint32_t maxp, uint32_t* probe_ids, float* probe_d2);
size_t SelectProbesLds(uint32_t nlist, uint32_t maxp);
hipError_t LaunchSelectProbes(hipStream_t st, const float* dist, uint64_t nq, uint32_t nlist,
                              const uint32_t* nprobe_q, uint32_t nprobe_all, uint32_t maxp, uint32_t* probe_ids,
                              float* probe_d2);
// The plain scan reads its parameters from device memory (`dparams`, filled
// by the launcher on the stream) so that rarely used fields are loaded where
// they are used instead of occupying scalar registers for the whole kernel.
struct PlainParams {
    IndexView ix;
    PlainArgs a;
};
hipError_t LaunchPlainSearch(hipStream_t st, const IndexView& ix, const PlainArgs& a, PlainParams* dparams,
                             bool upload = true);
hipError_t LaunchCoarseAccumulate(hipStream_t st, const uint8_t* rowa, uint64_t nbatch, const uint8_t* lut, int blocks,
                                  int32_t* out);

// ---- batched search ---------------------------------------------------------
// One unit of the persistent batched scan: a tile of the (query, probe)
// pairs of one cluster against a range of the cluster's 32-vector batches.
struct BatchItem {
    uint32_t c;    // cluster
    uint32_t p0;   // first pair (index into the pair list)
    uint32_t np;   // pairs in the tile
    uint32_t b0;   // first batch
    uint32_t nb;   // batches
    uint32_t pad[3];
};

struct BatchedArgs {
    uint64_t nq;
    uint32_t maxp;
    const uint32_t* nprobe_q;   // optional [nq]
    uint32_t nprobe_all;
    const uint32_t* probe_ids;  // [nq][maxp]
    const float* probe_d2;      // [nq][maxp]
    const uint8_t* lut;         // [nq][blocks*32]
    const float* consts;        // [nq][8] (see PlainArgs)
    const float* rotq;          // [nq][padded]
    uint32_t k;
    float z;
    int disable_prune;
    const uint32_t* first_ids;  // [nq][k] top-k over the first probe
    const float* first_dists;
    uint32_t* ccount;           // [nlist] pairs per cluster
    uint32_t* cfill;            // [nlist] scatter cursor
    uint32_t* coff;             // [nlist+1] pair offsets
    uint32_t* ioff;             // [nlist+1] item offsets
    uint64_t* pairs;            // (query << 32) | probe position, grouped by cluster
    BatchItem* items;
    uint64_t items_cap;         // capacity of items (an upper bound of the count)
    uint32_t* counters;         // [0] items, [1] next item, [2] overflowed queries
    uint32_t qtile, btile;      // pairs per item, batches per item
    uint64_t* cand;             // [nq][cap] (probe position << 32) | vector offset
    uint32_t* ncand;            // [nq]
    uint32_t cap;
    uint32_t* overflow;         // [nq] queries whose candidates did not fit
    uint32_t* out_ids;          // [nq][k]
    float* out_dists;
    unsigned long long* stats;  // [4] as PlainArgs
    uint32_t p_lo;              // first probe position of the work list (1: the first probe is the threshold pass)
    const uint8_t* qmask;       // optional [nq]: only queries with a nonzero entry get pairs
};

struct BatchedParams {
    IndexView ix;
    BatchedArgs a;
};

constexpr uint32_t kBatchCandCap = 1024;   // candidates kept per query (more: plain search)

// Default item shape of the batched scan: up to 32 batches (~24 KB of
// records) and up to 32 pairs (~16 KB of LUTs) per item.
inline void BatchTiles(const IndexView& ix, uint32_t* qtile, uint32_t* btile) {
    uint32_t bt = (24576u / ix.batch_bytes) & ~1u;
    bt = bt < 2 ? 2 : (bt > 32 ? 32 : bt);
    uint32_t qt = 16384u / (static_cast<uint32_t>(ix.blocks) * 32u);
    qt = qt < 4 ? 4 : (qt > 32 ? 32 : qt);
    *qtile = qt;
    *btile = bt;
}

// The work list of the pairs (q, p), 1 <= p < nprobe(q), of non-empty
// clusters: count, scan, scatter, items.  Fills the candidate and batch
// counts of those pairs into stats.
hipError_t LaunchBatchedWorkList(hipStream_t st, const IndexView& ix, const BatchedArgs& a, BatchedParams* dparams,
                                 bool upload = true);
// The persistent scan over the items; `grid` workgroups.
size_t BatchedScanLds(const IndexView& ix, uint32_t qtile, uint32_t btile);
hipError_t LaunchBatchedScan(hipStream_t st, const IndexView& ix, const BatchedArgs& a, const BatchedParams* dparams,
                             uint32_t grid);
// Refine of every query's candidates merged into its first-probe top-k.
hipError_t LaunchBatchedRefine(hipStream_t st, const IndexView& ix, const BatchedArgs& a,
                               const BatchedParams* dparams);
// ---- MFMA form (e8q_mfma.hip) ----------------------------------------------
// Exact estimator over every probed vector as a split-BF16 MFMA contraction:
// S(q, v) = sum_i q_i L(d_i) with the branch folded by the level table's
// reflection (L1[d] = -L0[Q-1-d]), codes as one-hot {-1, 0, +1} rows of
// K = 16 per dimension, the query side q_i L0[l] split into hi + lo bf16,
// fp32 accumulation.
constexpr uint32_t kMfmaQTile = 32;    // pairs per item (the MFMA's M)
constexpr uint32_t kMfmaBTile = 16;    // batches per item (4 waves x 4 tiles)
constexpr uint32_t kMfmaMaxBits = 4;   // 2^bits folded levels fit one K = 16 slice
constexpr uint32_t kMfmaDynK = 32;     // k up to this: the bound falls during the scan

struct MfmaExtra {
    const uint16_t* qexp;       // [nq][padded][32] bf16 bits: 16 hi, then 16 lo
    uint32_t* thr;              // [nq] candidate bound as a monotone distance key
    uint32_t dyn;               // 1: each tile lowers thr to its k-th distance
    const uint8_t* mcodes;      // [records][blocks][32][8] folded level | branch << 7
    uint64_t* cand;             // [rows][cap] (monotone dist bits << 32) | id
    uint32_t* ncand;            // [nq]
    uint32_t cap;
    const uint32_t* qslot;      // optional [nq]: candidate row of each query
    const uint32_t* qsel;       // optional [nsel]: queries the select kernel handles
    uint64_t nsel;
};

struct MfmaParams {
    BatchedParams bp;   // first, so that the work-list kernels read it as theirs
    MfmaExtra m;
};

// Folded one-hot code layout and per-list max |g1| of an index.
hipError_t LaunchMfmaCodes(hipStream_t st, const IndexView& ix, uint64_t records, uint8_t* mcodes, float* g1max);
// Per query: the hi/lo bf16 query expansion, and (first_dists non-null) the
// candidate bound thr = tau (first-probe k-th distance) + a bound of both
// forms' rounding, as a monotone key.
hipError_t LaunchMfmaQuery(hipStream_t st, const IndexView& ix, const float* rotq, uint64_t nq, uint32_t k,
                           const float* first_dists, const uint32_t* probe_ids, uint32_t maxp, const float* g1max,
                           float lmax, uint16_t* qexp, uint32_t* thr);
hipError_t LaunchMfmaScan(hipStream_t st, const MfmaParams* dp, uint32_t grid);
hipError_t LaunchMfmaSelect(hipStream_t st, const MfmaParams* dp, uint64_t nrun, uint32_t k);

// Fill a device buffer with a 32-bit value (graph-safe, no memset nodes).
hipError_t LaunchFill32(hipStream_t st, uint32_t* p, uint64_t n, uint32_t v);
// The queries by estimated scan work (the sizes begin[c + 1] - begin[c] of
// their probed lists), heaviest first, into order[nq]; key[nq] and
// hist[2 * kOrderBuckets] are scratch.
constexpr uint32_t kOrderBuckets = 132;
hipError_t LaunchQueryOrder(hipStream_t st, const uint64_t* begin, uint64_t nlist, const uint32_t* probe,
                            const uint32_t* nprobe_q, uint32_t nprobe_all, uint32_t maxp, uint64_t nq, uint32_t* key,
                            uint32_t* hist, uint32_t* order);

// Build (encode) path.
struct EncodeArgs {
    const float* base;          // [n][dim] device
    const uint32_t* cluster_ids;  // [n]
    const uint32_t* slot;       // [n] cluster-major slot of each row
    uint64_t n;
    uint8_t* arena;
    uint8_t* ref_arena;
    uint32_t* ids;
    float* ob;                  // [n] by slot: bank inputs base / W / Omega
    float* ow;
    float* op;
};

hipError_t LaunchMean(hipStream_t st, const float* base, uint64_t n, uint32_t dim, double* mu);
hipError_t LaunchMeanFinish(hipStream_t st, const double* mu, uint64_t n, const IndexView& ix, float* mean_rot);
hipError_t LaunchCentroids(hipStream_t st, const float* centroids, const IndexView& ix, float* centroid_rot,
                           float* centroid_n2);
hipError_t LaunchEncodeRows(hipStream_t st, const IndexView& ix, const EncodeView& ev, const EncodeArgs& a);
hipError_t LaunchBanks(hipStream_t st, const IndexView& ix, const float* ob, const float* ow, const float* op,
                       uint8_t* arena);
// Scan records [nrec] between the E8AX rowA layout ([blk][lane]) and the
// device layout ([lane][blk]); src may equal dst.
hipError_t LaunchRecordLayout(hipStream_t st, const IndexView& ix, const uint8_t* src, uint8_t* dst, uint64_t nrec,
                              bool to_device);
// Fills ix.csum from the scan-record headers (after build or upload).
hipError_t LaunchClusterSummary(hipStream_t st, const IndexView& ix, float4* csum);
hipError_t LaunchEncodeUnits(hipStream_t st, const IndexView& ix, const EncodeView& ev, const float* unit,
                             uint64_t rows, double scale, uint16_t* digits, uint8_t* branch);

}  // namespace e8qdev
