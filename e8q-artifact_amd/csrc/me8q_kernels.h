// me8q_kernels.h -- host-callable launchers of the ME8Q kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "e8q_kernels.h"
#include "me8q_device.h"

namespace e8qdev {

// ---- build (me8q_build.hip)
hipError_t LaunchM8Centroids(hipStream_t st, const M8View& v);
size_t M8EncodeLds(const M8View& v);
// rank[row] = position of `row` inside its list (the reference's list_rank,
// me8q_build.cpp:154-163).
hipError_t LaunchM8Encode(hipStream_t st, const M8View& v, const float* base, const uint32_t* cid,
                          const uint32_t* rank, uint64_t n);
// perp: n*nrun float scratch; the plane buffer must be zero beforehand.
hipError_t LaunchM8Fast(hipStream_t st, const M8View& v, float* perp, uint64_t nbatch);
// The batched form's per-vector refine records and per-batch bank scales (rrec, bscale).
hipError_t LaunchM8RefineRecords(hipStream_t st, const M8View& v, uint64_t nbatch);

// ---- search (me8q_search.hip)
// Per-query prep: padded, per-run rotated query (rotq [nq][padded]), the
// quantized class-mean LUT (lut [nq][plane_cols*32]: per plane column 16
// low-nibble then 16 high-nibble bytes), qc [nq][4] = {lut_min, lut_delta,
// |q|^2 over the route prefix, bits per row} and rc [nq][nrun][4] doubles =
// {|q_r|^2, lut_min_x, qerr, score_max}.
hipError_t LaunchM8QueryPrep(hipStream_t st, const M8View& v, const float* queries, uint64_t nq, float* rotq,
                             uint8_t* lut, float* qc, double* rc);
// sqd2 [nq][maxp][nrun] = sqrt(max(|q_r - c_r|^2, 0)) per probed list.
hipError_t LaunchM8RunDist(hipStream_t st, const M8View& v, const float* rotq, const double* rc,
                           const uint32_t* probe, const uint32_t* nprobe_q, uint32_t nprobe_all, uint32_t maxp,
                           uint64_t nq, double* sqd2);

struct M8ScanArgs {
    M8View v;
    const float* rotq;
    const uint8_t* lut;
    const float* qc;
    const double* rc;
    const uint32_t* probe;
    const float* pdist;
    const double* sqd2;
    const uint32_t* nprobe_q;
    uint32_t nprobe_all, maxp;
    uint64_t nq;
    uint32_t k;
    float z;
    uint32_t* out_ids;
    float* out_d;
    uint64_t* qstats;   // [nq][8] candidates, survivors, bits_scored, bits_total, pruned, batches read,
                        //         candidates refined, refine groups
    uint32_t lev_floats;   // total size of the runs' level tables
    uint32_t run_slot;     // bytes of one run's plane bytes + fields, largest run (a multiple of 16)
    // Threshold pass of the batched form (budget > 0): lists in probe order
    // until `budget` 32-vector batches are scanned (the last list possibly in
    // part); cover[q] = (p << 32) | b, the first (probe, batch) not scanned.
    uint32_t budget;
    uint64_t* cover;
    const uint32_t* qsel;   // optional [nsel]: the queries to search (else 0..nq-1)
    uint64_t nsel;
    const uint32_t* nsel_dev;   // optional: with qsel, the number of queries to search, on the device (<= nsel)
};
size_t M8ScanLds(const M8View& v, uint32_t k, uint32_t lev_floats, uint32_t run_slot);
hipError_t LaunchM8Scan(hipStream_t st, const M8ScanArgs& a);


// ---- batched form (me8q_batch.hip) ----------------------------------------
// Every query of a call as one (query, list) work list.  A threshold pass
// (the plain scan over each query's first `budget` batches in probe order)
// gives tau_q, the k-th distance it found; a persistent kernel then computes
// the integer coarse sums of every remaining (query, vector) pair as an int8
// MFMA contraction (one-hot nibble codes against the u8 LUT shifted to
// int8: exact), applies the reference's coarse bound against tau_q and
// appends the passing vectors to the query's candidate list; the candidates'
// exact distances (the reference's arithmetic, with its cascade bound
// against tau_q) and a top-k selection merged with the threshold pass's list
// follow.
constexpr uint32_t kM8BatchRuns = 4;     // coded runs the batched form supports
constexpr uint32_t kM8BatchQTile = 32;   // pairs per work item (the MFMA's M)
constexpr uint32_t kM8BatchBTile = 8;    // batches per work item (4 waves x 2 tiles)
constexpr uint32_t kM8RefineChunk = 64;  // candidates per refine task
constexpr uint32_t kM8RefineGrab = 8;    // refine tasks a wave takes at a time
constexpr uint32_t kM8RefRowWords = 16;  // plane words (8 columns each) the refine loads in one burst
constexpr uint32_t kM8RefineWaves = 4;   // waves of a refine workgroup (one LDS copy of the level tables)
constexpr uint32_t kM8RefLevMax = 12288; // level-table floats (48 KiB) the refine stages in LDS
constexpr uint32_t kM8LevGlobal = 0xFFFFFFFFu;   // a run whose level table the refine reads from global memory
constexpr uint32_t kM8SelMaxK = 512;     // top-k of the selection kernel
constexpr uint32_t kM8TauBins = 64;      // histogram of candidates' coarse estimates over [0, tau0)
constexpr uint32_t kM8SetA = 1024;       // candidates per query refined first (their exact top-k gives tau1)

struct M8BatchArgs {
    M8View v;
    const uint8_t* lut;       // [nq][plane_cols*32]
    const uint8_t* lut8;      // the same as int8 (u8 ^ 0x80 = u8 - 128): the contraction's operand
    const float* qc;          // [nq][4]
    const double* rc;         // [nq][nrun][4]
    const float* rotq;        // [nq][padded]
    const uint32_t* probe;    // [nq][maxp]
    const float* pdist;       // [nq][maxp]
    const double* sqd2;       // [nq][maxp][nrun]
    uint32_t maxp, k;
    uint64_t nq;
    float z;
    int32_t cascade;          // 1: the reference's cascade bound against tau_q in the refine
    const uint32_t* first_ids;   // [nq][k] threshold pass (ascending, UINT32_MAX / +inf padded)
    const float* first_d;
    uint64_t* cover;             // [nq] (p << 32) | b: first (probe, batch) the threshold pass left
    const BatchItem* items;
    const uint64_t* pairs;
    uint64_t items_cap;
    uint32_t* counters;       // [0] items, [1] next item, [3] next refine task
    uint32_t* ncand;          // [nq] candidates at tau0 (count pass)
    uint32_t* nfill;          // [nq] candidates of the refine launch (its count array)
    float* tau;               // [nq] threshold of the refine (<= tau0)
    uint32_t* hist;           // [nq][kM8TauBins] count pass: candidates by coarse estimate, bins over [0, tau0)
    uint32_t* abin;           // [nq] last histogram bin of the query's set A
    uint32_t* na;             // [nq] |A|: candidates with the best coarse estimates, refined first
    uint32_t* fill_a;         // [nq] write cursors of A and of the rest (B)
    uint32_t* fill_b;
    float* pdlo;              // [pool] each candidate's coarse lower bound at tau0 (less a rounding margin)
    uint64_t* coff_b;         // [nq] first slot of B; after the compaction, B's survivors of the bound at tau1
    uint64_t* coff;           // [nq+1] each query's first pool slot
    uint32_t* toff;           // [nq+1] refine task offsets
    uint64_t* pool;           // [coff[nq]] (probe position << 32) | vector position in its list; the
                              // refine overwrites each task's slots with its survivors, (key << 32) | id
    uint32_t* tsurv;          // [toff[nq]] survivors of each refine task
    uint64_t* cont;           // first refine phase: candidates alive after their first run, at coff[q]
    uint32_t* ncont;          // [nq] their count
    uint32_t* out_ids;        // [nq][k]
    float* out_d;
    unsigned long long* stats;   // [8] pairs scored, candidates, refined, survivors, bits scored, items, -, -
    uint32_t* chk;            // optional [nq]: sum over scored pairs of (pos+1) * sum_r (r+1) * coarse_sum_r, mod 2^32
    uint32_t lev_floats;
    uint32_t set_a;           // candidates per query of set A (kM8SetA)
    // Refine: each run's level table offset in LDS (floats; kM8LevGlobal: read
    // from global memory) and the staged total, at most kM8RefLevMax.
    uint32_t ref_lev_off[kM8BatchRuns];
    uint32_t ref_lev_floats;
    // Records of the count pass (null: none written): each passing (pair,
    // vector) as {position in its list, tau0 bin << 26 | its lower bound
    // rounded down to 26 bits}, a pair's records of an item contiguous.
    uint2* rec;
    uint64_t rec_cap;         // records that fit; past it the count goes on and nothing is written
    unsigned long long* rcount;   // [1] records appended (all of them, also past rec_cap)
    uint2* itab;              // [items_cap * kM8BatchQTile] per (item, pair): {first record,
                              // count | smallest tau0 bin of its records << 16 (0xFFFF: none binned)}
};
// write = 0: count each query's candidates (ncand), and with a.rec write
// their records; write = 1: append them to the pool at coff[q] (same
// arithmetic, so the same vectors pass).
hipError_t LaunchM8MfmaScan(hipStream_t st, const M8BatchArgs& a, uint32_t grid, bool write);
// The count pass's records into the pool, from the records instead of a
// second contraction: set_b = 0, set A at coff[q] (count fill_a); set_b = 1,
// after A's refine, the rest whose lower bound is below min(tau0, tau1) at
// coff[q] + na[q] (count fill_b).
hipError_t LaunchM8Place(hipStream_t st, const M8BatchArgs& a, uint32_t grid, bool set_b);
// out[i] = in[i] ^ 0x80 over nq * plane_cols * 32 bytes (the int8 LUT).
hipError_t LaunchM8LutInt8(hipStream_t st, const uint8_t* in, uint8_t* out, uint64_t bytes);
// Pool offsets coff from ncand (with_offsets) and refine tasks toff from
// `count`; coff[nq] = pool entries, toff[nq] = refine tasks.
hipError_t LaunchM8RefineTasks(hipStream_t st, const M8BatchArgs& a, const uint32_t* count, bool with_offsets);
// Per query from the count pass's histogram: set A = the candidates in the
// lowest bins holding at least kM8SetA of them (abin, na); tau = tau0.
hipError_t LaunchM8Split(hipStream_t st, const M8BatchArgs& a);
// After A's exact top-k is in first_ids / first_d: tau = min(tau0, its k-th
// distance), coff_b = coff + na, and each query's B compacted in place to the
// candidates whose lower bound is below tau (count in nfill); placed = 1: B
// was placed from the records with that filter, nfill = fill_b.
hipError_t LaunchM8Compact(hipStream_t st, const M8BatchArgs& a);
// Queries whose pool slots pass pool_cap: no candidates, listed in ovf (count novf).
hipError_t LaunchM8Fit(hipStream_t st, const M8BatchArgs& a, uint64_t pool_cap, uint32_t* ovf, uint32_t* novf);
size_t M8RefineLds(const M8View& v, uint32_t lev_floats);
// first = true (indexes of two or more runs): each candidate of a.pool gets
// the first run of its cascade; those still alive go to a.cont.  first =
// false: the whole cascade, survivors written over the task's slots.
hipError_t LaunchM8Refine(hipStream_t st, const M8BatchArgs& a, uint32_t grid, bool first);
hipError_t LaunchM8Select(hipStream_t st, const M8BatchArgs& a);
// The threshold pass's pool: each query's first `budget` batches in probe
// order at q * budget * 32 (a.pool), their count (a.nfill), offsets (a.coff)
// and cover (a.cover).
hipError_t LaunchM8ThresholdPool(hipStream_t st, const M8BatchArgs& a, const uint32_t* nprobe_q, uint32_t nprobe_all,
                                 uint32_t budget);

}  // namespace e8qdev
