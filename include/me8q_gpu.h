/* me8q_gpu.h -- C ABI of the ME8Q (mixed-rate E8Q) part of libe8q_gpu.so.
 *
 * The reference's ME8Q boundary is the C++ class e8qlib::me8q::Me8qIndex
 * (include/e8qlib/index/me8q.hpp:69-247) and its pybind11 module
 * (python_bindings/me8q_bindings.cpp:22-315); there is no FFI.  These entry
 * points are the device layer those C++ methods now call.  Conventions as in
 * e8q_gpu.h: plain pointers and sizes, an e8q_status per call, the message
 * of the last failure in e8q_last_error(), host pointers unless a field says
 * device, one host thread per context.
 */
#ifndef ME8Q_GPU_H_
#define ME8Q_GPU_H_

#include <stddef.h>
#include <stdint.h>

#include "e8q_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct e8q_m8index e8q_m8index; /* device-resident ME8Q index */

/* Static description: shape, plan, rotators, codec tables, list offsets.
 * Mirrors Me8qIndex's private members (me8q.hpp:204-246); the host computes
 * the plan (Allocate, me8q_plan.cpp:47-167, or ParsePlanSpec), the rotator
 * signs (BuildRunRotators, me8q_build.cpp:29-40) and the level tables
 * (BranchLevels / FixedScale, e8q_codec.cpp:427-509) as Build and Load do. */
typedef struct {
    uint64_t n, dim, padded, nlist;   /* padded = dim rounded up to 64 */
    uint32_t nseg;                    /* plan runs */
    const uint64_t* seg_len;          /* nseg, multiples of 64 summing to padded */
    const int32_t* seg_bits;          /* nseg, 0 (dropped) or 1..13 */
    const uint8_t* signs;             /* the coded runs' rotator sign bytes, 4*len/8 each, in plan order */
    const double* levels;             /* for b = 1..13: branch 0 then branch 1, 2^b each */
    const double* fixed_scale;        /* [14], FixedScale(b) at index b */
    const uint64_t* list_begin;       /* nlist+1 vector offsets (cluster-major ids) */
    int32_t drop_energy;              /* BuildOptions::drop_energy */
} e8q_m8_desc;

/* The arrays of an ME8QIDXX file (me8q_io.cpp:29-64) that are not implied by
 * the desc: ids, the zero-padded (unrotated) centroids, the code rows. */
typedef struct {
    uint32_t* ids;        /* n, cluster-major                     */
    float* centroids;     /* nlist*padded                         */
    uint8_t* codes;       /* n*row_bytes (e8q_m8_row_bytes)       */
} e8q_m8_arrays;

typedef struct {
    const float* queries;             /* nq*dim (device if queries_on_device) */
    uint64_t nq;
    uint64_t k;                       /* 1..1024 on the device (the reference has no cap) */
    uint64_t nprobe;                  /* used when nprobe_per_query is null */
    const uint32_t* nprobe_per_query; /* optional host array [nq] */
    float z;                          /* lambda = OrthoQuantile(ortho_alpha, max(4k, 32)), host computed
                                         (me8q_fast.cpp:444-450) */
    int32_t queries_on_device;
    uint32_t* out_ids;                /* nq*k, ascending distance, UINT32_MAX padded */
    float* out_dists;                 /* nq*k, +inf padded */
    int32_t outputs_on_device;
    int32_t mode;                     /* E8Q_MODE_PLAIN (0): the reference's per-query schedule, bit for
                                         bit.  E8Q_MODE_BATCHED (1): the whole query set as one (query, list)
                                         work list -- a threshold pass over each query's first lists, the
                                         integer coarse sums of every other (query, vector) pair as an exact
                                         int8 MFMA contraction tested with the reference's coarse bound
                                         against that fixed threshold, the exact cascade of the passing
                                         vectors, a top-k merge.  Needs 1..4 coded runs and k <= 512; other
                                         indexes and k run the plain form (stats.form says which ran).
                                         No reference counterpart: the reference searches query by query. */
    uint32_t batched_budget;          /* threshold pass length in 32-vector batches (0: 16) */
    uint32_t batched_set_a;           /* candidates per query refined first for the exact threshold (0: 1024) */
} e8q_m8_search_args;

typedef struct {
    uint64_t candidates, survivors, bits_scored, bits_total, pruned;   /* me8q::SearchStats (me8q.hpp:54-69) */
    double scan_ms;    /* scan kernel, HIP events on the context stream */
    double total_ms;   /* device time of the whole search */
    uint64_t batches;  /* 32-vector scan records read (plane + banks) */
    uint64_t refined;  /* candidates that entered the exact cascade */
    uint64_t groups;   /* cascade groups (up to four candidates each; plain form and threshold pass) */
    int32_t form;      /* the form that ran: E8Q_MODE_PLAIN or E8Q_MODE_BATCHED */
    uint64_t overflowed;   /* batched form: queries whose candidate lists overflowed, searched again
                              by the plain form */
    uint64_t pairs_scored; /* batched form: (query, vector) pairs of the MFMA contraction */
    uint64_t items;        /* batched form: its work items (32 pairs x 8 batches at most) */
    /* batched form, the exact refine kernel per phase ([0]: every candidate's
     * first run; [1]: the cascades of those still alive and the threshold
     * pass): summed time of its launches (HIP events around each launch on the
     * context stream), number of launches, and algorithmic bytes (per
     * candidate: its 8-byte pool entry, its plane row, its 16-byte run
     * headers, its two u8 bank entries (W, Omega) and 1/32 of the 32-byte
     * batch scales per run, the fields of the runs it scores, and what it writes) */
    double refine_ms[2];
    uint64_t refine_launches[2];
    uint64_t refine_bytes[2];
} e8q_m8_stats;

/* Allocates the device index and its static tables (the allocation half of
 * Me8qIndex::Build / Load). */
e8q_status e8q_m8_create(e8q_ctx* ctx, const e8q_m8_desc* desc, e8q_m8index** out);
e8q_status e8q_m8_destroy(e8q_m8index* idx);
uint64_t e8q_m8_device_bytes(const e8q_m8index* idx);
/* Bytes of one code row: nrun (coef, h2) pairs + the runs' codes (me8q_build.cpp:103-135). */
uint64_t e8q_m8_row_bytes(const e8q_m8index* idx);

/* Uploads the arrays of a loaded file and derives the rotated centroids and
 * the FastScan layout on the device (Me8qIndex::Load, me8q_io.cpp:85-176,
 * PrepareCentroidsRotated + BuildFastLayout). */
e8q_status e8q_m8_upload(e8q_ctx* ctx, e8q_m8index* idx, const e8q_m8_arrays* host);
/* Downloads them (feeds Me8qIndex::Save, me8q_io.cpp:29-64). */
e8q_status e8q_m8_download(e8q_ctx* ctx, const e8q_m8index* idx, e8q_m8_arrays* host);

/* The compute of Me8qIndex::Build (me8q_build.cpp:43-311) on the device:
 * rotated centroids, per-row per-run rotation / E8 encode / pack / fits, then
 * BuildFastLayout.  `centroids` = nlist*padded zero-padded host array;
 * rank[i] = position of row i in its list, host array.  base (n*dim) and
 * cluster_ids (n) are host pointers, or device pointers when on_device. */
e8q_status e8q_m8_encode(e8q_ctx* ctx, e8q_m8index* idx, const float* base, const float* centroids,
                         const uint32_t* cluster_ids, const uint32_t* rank, int on_device);

/* Me8qIndex::SearchIVFFast for each query (me8q_fast.cpp:380-1108; the
 * binding's batch entry point SearchIVFFastBatch gives the same results).
 * per_query_stats: optional host [nq*5] = candidates, survivors, bits_scored,
 * bits_total, pruned. */
e8q_status e8q_m8_search(e8q_ctx* ctx, const e8q_m8index* idx, const e8q_m8_search_args* args,
                         e8q_m8_stats* stats, uint64_t* per_query_stats);
/* Device scratch of one search of nq host queries with per-query nprobe <=
 * max_nprobe and host outputs in `mode`, carved from the context workspace.
 * In the batched form it counts the candidate pool and the count pass's
 * records at their default capacity: min(2^31, what the max_nprobe largest
 * lists hold, nq x 65536) entries of 28 bytes (pool, cascade list, lower
 * bound, record), plus the threshold pass's at the default budget.  A
 * context whose workspace was reserved (e8q_ctx_reserve) to at least this
 * size runs such a call without allocating; on a reserved context a call's
 * pool is what the reservation leaves after the call's other buffers (up to
 * what the probed lists can hold), never more, and the queries whose
 * candidates do not fit are searched by the plain form (stats.overflowed). */
uint64_t e8q_m8_search_workspace_bytes(const e8q_m8index* idx, uint64_t nq, uint64_t k, uint64_t max_nprobe,
                                       int32_t mode);

/* Grids of the batched form's two persistent kernels on this context (no
 * reference counterpart: the reference makes no device launches): the int8
 * MFMA contraction and the exact refine, in workgroups.  0 leaves each to
 * the context's autotuner, which times the whole search at 1, 2 and 4
 * workgroups per CU for the contraction, then 4, 8 and 16 per CU for the
 * refine, on the first batched search of at least 64 queries per (index
 * shape, power-of-two buckets of the query count, nprobe and k).  Results do
 * not depend on the grids.  Graph replay (e8q_ctx_set_graphs) covers
 * e8q_m8_search too: the batched form's whole sequence is one graph. */
e8q_status e8q_ctx_set_m8_grids(e8q_ctx* ctx, uint32_t mfma_grid, uint32_t refine_grid);
typedef struct {
    uint32_t mfma_grid, refine_grid;   /* grids of the last batched search */
    uint32_t budget, set_a;            /* its threshold budget and set-A size */
    uint64_t tunings;                  /* autotuner runs of the batched ME8Q form */
    uint64_t pool_per_query;           /* pool entries per query learned from past calls (0: default) */
} e8q_m8_ctx_info;
e8q_status e8q_ctx_get_m8_info(const e8q_ctx* ctx, e8q_m8_ctx_info* out);
/* Makes ctx use the autotuned-grid table of `from` (a context on the same
 * device) from now on, so that a bucket either one tunes serves both.  No
 * reference counterpart: for the worker contexts of a server, which would
 * otherwise each time the same buckets. */
e8q_status e8q_ctx_share_tuning(e8q_ctx* ctx, const e8q_ctx* from);

/* ---- diagnostics (me8q_measure.cpp:41-67, me8q_search.cpp:136-266) -----
 * Host pointers.  Each value is the reference's bit for bit. */
/* Me8qIndex::Estimate (me8q_measure.cpp:41-67; binding `estimate`,
 * me8q_bindings.cpp:209-218) for npair (query, id) pairs: out[i] = the
 * EstimateRow distance of the vector with id ids[i] from queries[qidx[i]]
 * (nq x dim, the PCA domain).  An unknown id fails the call with E8Q_EINVAL
 * ("Me8qIndex::Estimate: unknown id"), as does qidx[i] >= nq. */
e8q_status e8q_m8_estimate(e8q_ctx* ctx, const e8q_m8index* idx, const float* queries, uint64_t nq,
                           const uint32_t* qidx, const uint32_t* ids, uint64_t npair, float* out);
/* Me8qIndex::ExhaustiveTopK (me8q_search.cpp:238-266; me8q.hpp:122-126) for
 * each of nq queries: the list PushCandidate leaves after every stored
 * vector in list order, k (1..1024) wide -- ascending distance, a later
 * vector first among equal distances; rows [nq*k], UINT32_MAX / +inf padded. */
e8q_status e8q_m8_exhaustive_topk(e8q_ctx* ctx, const e8q_m8index* idx, const float* queries, uint64_t nq,
                                  uint64_t k, uint32_t* out_ids, float* out_dists);
/* Me8qIndex::SearchIVF (me8q_search.cpp:136-236; me8q.hpp:85-96), the full
 * estimate of every vector of the nprobe nearest lists, for each query:
 * ids / dists [nq*k] as ExhaustiveTopK's order; stats6 (optional, [nq*6]) =
 * candidates, clusters, bits_scored, bits_total, survivors, pruned (0).  With
 * k == 0, nprobe == 0 or an empty index the outputs are left untouched, as
 * in the reference; k <= 1024. */
e8q_status e8q_m8_search_ivf(e8q_ctx* ctx, const e8q_m8index* idx, const float* queries, uint64_t nq, uint64_t k,
                             uint64_t nprobe, uint32_t* out_ids, float* out_dists, uint64_t* stats6);

/* ---- diagnostics used by tests (reference layouts, host pointers) -------- */
/* The batched form with a checksum of its integer coarse sums: chk[q] =
 * sum over the (q, vector) pairs of q's probed lists of (pos + 1) *
 * sum_r (r + 1) * S_r, mod 2^32, where pos is the vector's position in its
 * list and S_r its coarse sum of run r (AccumulateCoarseHi over the run's
 * plane columns, me8q_fast.cpp:800-818).  chk: host [nq]. */
e8q_status e8q_m8_debug_batch_checksum(e8q_ctx* ctx, const e8q_m8index* idx, const e8q_m8_search_args* args,
                                       uint32_t* chk);
/* One list's FastLayout as the reference holds it (me8q.hpp:153-184): plane
 * [batches][plane_cols*32 + nrun*160] ([col][lane] bytes), refine
 * [cnt][refine_stride], gain / vpc [nrun][cnt], rc_n2 [nrun].  sizes3 =
 * {plane bytes, refine bytes, gain count}; any array pointer may be null. */
e8q_status e8q_m8_debug_fast(e8q_ctx* ctx, const e8q_m8index* idx, uint64_t list, uint64_t* sizes3,
                             uint8_t* plane, uint8_t* refine, float* gain, float* vpc, double* rc_n2);
/* Rotated centroids [nlist*padded] and their squared norms [nlist]. */
e8q_status e8q_m8_debug_centroids(e8q_ctx* ctx, const e8q_m8index* idx, float* cent_rot, float* cn2);
/* Per-query intermediates: rotated query [nq*padded], u8 LUT in the
 * reference's layout [nq*plane_cols*64], {lut_min, lut_delta} [nq*2], probe
 * lists [nq*nprobe] and their distances [nq*nprobe]. */
e8q_status e8q_m8_debug_query(e8q_ctx* ctx, const e8q_m8index* idx, const float* queries, uint64_t nq,
                              uint64_t nprobe, float* rot_q, uint8_t* lut, float* lut_md, uint32_t* probe,
                              float* pdist);

#ifdef __cplusplus
}
#endif

#endif /* ME8Q_GPU_H_ */
