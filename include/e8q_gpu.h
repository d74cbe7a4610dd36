/* e8q_gpu.h -- C ABI of the MI355X (gfx950) E8Q device library, libe8q_gpu.so.
 *
 * The reference e8qlib has no FFI layer: its boundary is the C++ class
 * e8qlib::E8qIndex (include/e8qlib/index/e8q.hpp:121-307) and the pybind11
 * module over it (python_bindings/e8q_bindings.cpp:22-344).  This header is
 * the thin device layer those C++ entry points now call; each function names
 * the reference interface it replaces.  Plain pointers and sizes only, no
 * exceptions across the boundary: every call returns an e8q_status and
 * e8q_last_error() holds a thread-local message for the last failure.
 *
 * Pointers are host pointers unless a field says "device".  Calls on one
 * context are serialized by the caller (one host thread per context); the
 * device work of a context runs in order on that context's HIP stream.
 */
#ifndef E8Q_GPU_H_
#define E8Q_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    E8Q_OK = 0,
    E8Q_EINVAL = 1,   /* bad argument (reference: std::invalid_argument) */
    E8Q_ERUNTIME = 2, /* HIP/runtime failure (reference: std::runtime_error) */
    E8Q_ENOMEM = 3    /* device allocation failed */
} e8q_status;

typedef struct e8q_ctx e8q_ctx;       /* device + stream + scratch pool */
typedef struct e8q_dindex e8q_dindex; /* device-resident E8Q index      */

/* Static, query-independent description of an index: the shape, the
 * rotator, the level tables and the cluster offsets.  Mirrors the private
 * members of e8qlib::E8qIndex (e8q.hpp:201-246); the host computes the
 * tables exactly as E8qIndex::Build / Load do (e8q_build.cpp:87-182,
 * e8q_io.cpp:213-228). */
typedef struct {
    uint64_t n, dim, padded, nlist;
    int32_t bits;
    float alpha;                      /* prune_alpha (>0), e8q_build.cpp:101 */
    double scale;                     /* encoder scale s(b) = FixedScale(bits) */
    const uint8_t* signs;             /* 4*padded/8 rotator sign bytes */
    const float* refine_tab;          /* 2*max(256, 2^bits) level/sqrt(padded) */
    const double* gate_mid;           /* 16 scan class means (e8q_io.cpp:213) */
    const double* levels;             /* 2 * 2^bits: branch 0 then branch 1 */
    const uint64_t* cluster_begin;    /* nlist+1 vector offsets */
    const uint64_t* arena_begin;      /* nlist+1 scan-record byte offsets */
    const uint64_t* ref_arena_begin;  /* nlist+1 refine-record byte offsets */
} e8q_index_desc;

/* The E8AX arrays of one index (e8q_io.cpp:35-84).  Used to upload a loaded
 * file and to download a built index for Save. Sizes follow the desc. */
typedef struct {
    uint32_t* ids;           /* n, cluster-major                  */
    float* mean_rot;         /* padded                            */
    float* centroid_rot;     /* nlist*padded, centered            */
    float* centroid_n2;      /* nlist                             */
    uint8_t* arena;          /* arena_begin[nlist] + 64           */
    uint8_t* ref_arena;      /* ref_arena_begin[nlist] + 64       */
} e8q_index_arrays;

typedef struct {
    uint64_t candidates; /* SearchStats::candidates (e8q.hpp:55-62) */
    uint64_t survivors;  /* SearchStats::survivors (batched form: vectors refined;
                            a query searched again by the plain form after its
                            candidate list overflowed counts its first probe's) */
    uint64_t batches;    /* 32-vector scan records of the probed clusters */
    double scan_ms;      /* scan kernel time, HIP events on the context stream
                            (MFMA form: the contraction kernel; a graph
                            replay: the whole graph)                            */
    double total_ms;     /* device time of the whole search, same events       */
    uint64_t scanned;    /* of those, records whose codes were read (the rest
                            belong to clusters passed over as a whole) */
} e8q_search_stats;

/* Search modes. PLAIN keeps the reference's per-query schedule bit for bit
 * (clusters in probe order, survivors refined every flush_every rows against
 * the live top-k threshold: e8q_search.cpp:215-293); BATCHED runs the whole
 * query set as one (query, cluster) work list.  MFMA scores every vector of
 * every probed list with the refine estimator as a split-BF16 matrix
 * contraction (bits <= 4, FP32 accumulation; no pruning, so its reference
 * result is the unpruned estimator top-k, within the contraction's rounding)
 * and reports those distances; survivors then counts the vectors that passed
 * its candidate bound. */
enum { E8Q_MODE_PLAIN = 0, E8Q_MODE_BATCHED = 1, E8Q_MODE_MFMA = 2 };

typedef struct {
    const float* queries;          /* nq*dim (host, or device if queries_on_device) */
    uint64_t nq;
    uint64_t k;                    /* 1..128 (reference kMaxK, e8q.hpp:44) */
    uint64_t nprobe;               /* used when nprobe_per_query is null   */
    const uint32_t* nprobe_per_query; /* optional host array [nq]          */
    float z;                       /* union-bound quantile lambda (host computed,
                                      e8q_internal.hpp:102-105)            */
    int32_t disable_prune;         /* E8qIndex::SetDisablePrune            */
    int32_t flush_every;           /* survivor flush size 1..1024 (E8Q_FLUSH_EVERY, default 32; 0: 32) */
    int32_t mode;                  /* E8Q_MODE_*                           */
    int32_t queries_on_device;     /* 1: queries is a device pointer       */
    uint32_t* out_ids;             /* nq*k (host, or device if outputs_on_device) */
    float* out_dists;              /* nq*k                                 */
    int32_t outputs_on_device;
} e8q_search_args;

/* ---- context ---------------------------------------------------------- */
e8q_status e8q_ctx_create(int device, e8q_ctx** out);
e8q_status e8q_ctx_destroy(e8q_ctx* ctx);
e8q_status e8q_ctx_synchronize(e8q_ctx* ctx);
/* Grows the context's device workspace to `bytes` now, so that later calls
 * whose scratch fits it allocate nothing (serving admission; no reference
 * counterpart: the reference allocates per-thread host scratch). */
e8q_status e8q_ctx_reserve(e8q_ctx* ctx, uint64_t bytes);
uint64_t e8q_ctx_workspace_bytes(const e8q_ctx* ctx);

/* Graph replay of e8q_search on this context (no reference counterpart: the
 * reference makes no device launches).  With graphs on, the first search of
 * a call shape (sizes, k, nprobe, buffers, workspace) runs eagerly, the
 * second captures its launches from the query rotation to the results as one
 * HIP graph and later ones replay it.  A replayed search reports the graph's
 * time as both scan_ms and total_ms.  Off by default. */
e8q_status e8q_ctx_set_graphs(e8q_ctx* ctx, int enable);
/* Shape of the batched form's persistent scan: qtile (query, list) pairs and
 * btile 32-vector batches per work item, grid workgroups.  qtile = btile = 0:
 * the context's autotuner picks the tile (and, with grid = 0, the grid) per
 * index shape and power-of-two query-count bucket by timing candidates on the
 * first batched search of at least 64 queries.  Results do not depend on the
 * shape. */
e8q_status e8q_ctx_set_batched_shape(e8q_ctx* ctx, uint32_t qtile, uint32_t btile, uint32_t grid);
typedef struct {
    uint32_t qtile, btile, grid; /* shape of the last batched search */
    uint32_t graphs;             /* e8q_ctx_set_graphs state          */
    uint64_t tunings;            /* autotuner runs                    */
    uint64_t captures, replays;  /* graph captures and launches       */
} e8q_ctx_info;
e8q_status e8q_ctx_get_info(const e8q_ctx* ctx, e8q_ctx_info* out);
const char* e8q_last_error(void);

/* ---- index lifetime --------------------------------------------------- */
/* Allocates the device index and uploads the static tables of `desc`
 * (replaces the allocation half of E8qIndex::Build / Load). */
e8q_status e8q_index_create(e8q_ctx* ctx, const e8q_index_desc* desc, e8q_dindex** out);
e8q_status e8q_index_destroy(e8q_dindex* idx);
/* Bytes of device memory held by the index (E8qIndex::memory_bytes, e8q_io.cpp:27-33). */
uint64_t e8q_index_device_bytes(const e8q_dindex* idx);

/* Uploads the arrays of an E8AX file (replaces E8qIndex::Load's reads,
 * e8q_io.cpp:86-211). */
e8q_status e8q_index_upload(e8q_ctx* ctx, e8q_dindex* idx, const e8q_index_arrays* host);
/* Downloads them (feeds E8qIndex::Save, e8q_io.cpp:35-84). */
e8q_status e8q_index_download(e8q_ctx* ctx, const e8q_dindex* idx, e8q_index_arrays* host);

/* ---- build (encode) ----------------------------------------------------- */
/* The whole E8qIndex::Build compute (e8q_build.cpp:184-490) on the device:
 * fp64 base mean, rotated centroids, per-row rotate / residual / E8
 * nearest-lattice-point encode / fp64 fits / pack, then the per-batch u8
 * banks.  `slot[i]` = cluster_begin[cluster_ids[i]] + rank of row i in its
 * list (host computed, e8q_build.cpp:225-231), always a host array.
 * base / centroids / cluster_ids are host pointers, or device pointers when
 * on_device != 0 (then cluster ids are trusted to be < nlist). */
e8q_status e8q_index_encode(e8q_ctx* ctx, e8q_dindex* idx, const float* base, const float* centroids,
                            const uint32_t* cluster_ids, const uint32_t* slot, int on_device);

/* ---- search -------------------------------------------------------------- */
/* E8qIndex::Search over nq queries (e8q_search.cpp:215-318; the binding's
 * OpenMP loop, e8q_bindings.cpp:74-169).  Outputs are sorted ascending by
 * (dist, id) and padded with UINT32_MAX / +inf. */
e8q_status e8q_search(e8q_ctx* ctx, const e8q_dindex* idx, const e8q_search_args* args,
                      e8q_search_stats* stats);
/* Device scratch of one plain-form search of nq host queries with per-query
 * nprobe <= max_nprobe and host outputs (what e8q_search carves from the
 * context workspace); the serving loop admits work by it. */
uint64_t e8q_search_workspace_bytes(const e8q_dindex* idx, uint64_t nq, uint64_t k, uint64_t max_nprobe);

/* Size of the fixed per-GPU scratch array of e8q_search_static. */
#define E8Q_STATIC_SCRATCH_BYTES (32u << 20)

/* The plain form of e8q_search (mode must be E8Q_MODE_PLAIN) with no device
 * allocation: no hipMalloc, no hipFree, no growth of the context workspace.
 * No counterpart in the reference (its search allocates nothing on a
 * device); it is for callers that cannot allocate: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Every
 * per-call buffer (staged queries and outputs, rotated queries, LUTs, route
 * distances, probe lists, counters, the kernel parameter block) lives in one
 * fixed array of E8Q_STATIC_SCRATCH_BYTES per GPU, held by each GPU the
 * library runs on until the process ends.  A call whose buffers do not fit
 * (e8q_search_workspace_bytes with host queries and outputs is an upper
 * bound) fails with E8Q_EINVAL: search fewer queries per call.  Every call
 * on one GPU shares that GPU's array, so calls that overlap on one GPU
 * (several contexts or threads at once) race; serialize them.  The results,
 * stats and error behaviour are e8q_search's. */
e8q_status e8q_search_static(e8q_ctx* ctx, const e8q_dindex* idx, const e8q_search_args* args,
                             e8q_search_stats* stats);

/* ---- diagnostics (e8q_oracle.cpp:14-219) --------------------------------
 * Host pointers.  Each value is the reference's bit for bit. */
/* E8qIndex::Estimate (e8q_oracle.cpp:58-101; binding `estimate`,
 * e8q_bindings.cpp:171-180) for npair (query, id) pairs: out[i] = the
 * estimator distance of the stored vector with global id ids[i] from query
 * queries[qidx[i]] (nq x dim).  An id no stored vector carries fails the call
 * with E8Q_EINVAL ("Estimate: id out of range"), as does qidx[i] >= nq; an
 * empty index fails with E8Q_ERUNTIME. */
e8q_status e8q_estimate(e8q_ctx* ctx, const e8q_dindex* idx, const float* queries, uint64_t nq,
                        const uint32_t* qidx, const uint32_t* ids, uint64_t npair, float* out);
/* E8qIndex::ProbeBoundW (e8q_oracle.cpp:103-219) in the binding's batch form
 * `probe_bound_w(queries, base_rows, qidx, ids)` (e8q_bindings.cpp:181-234):
 * base_rows[i] (dim floats) is the build row of ids[i]; out[i] = (d_coarse -
 * d_full) / (r_q Omega) in fp64, NaN for a pair with qidx[i] >= nq or
 * ids[i] >= n and where the reference returns NaN (zero residual, degenerate
 * code).  An id below n that no stored vector carries fails the call with
 * E8Q_EINVAL; so does an index at b >= 10, where the reference's 8-byte digit
 * staging overflows (undefined behaviour). */
e8q_status e8q_probe_bound_w(e8q_ctx* ctx, const e8q_dindex* idx, const float* queries, uint64_t nq,
                             const float* base_rows, const uint32_t* qidx, const uint32_t* ids, uint64_t npair,
                             double* out);
/* E8qIndex::ExhaustiveTopK (e8q_oracle.cpp:14-56) for each of nq queries:
 * the k (1..1024) best (distance, id) over every stored vector, ordered by
 * (distance, id) as its partial_sort under Better; rows [nq*k], UINT32_MAX /
 * +inf padded. */
e8q_status e8q_exhaustive_topk(e8q_ctx* ctx, const e8q_dindex* idx, const float* queries, uint64_t nq, uint64_t k,
                               uint32_t* out_ids, float* out_dists);

/* ---- diagnostics used by tests ------------------------------------------ */
/* Coarse integer accumulate of one 32-lane batch (AccumulateCoarseHi,
 * e8q.hpp:333-336): rowa = blocks*32 bytes per batch, lut = blocks*64 bytes
 * per query; out = nbatch*32 int32.  Host pointers. */
e8q_status e8q_debug_coarse_accumulate(e8q_ctx* ctx, const uint8_t* rowa, uint64_t nbatch, const uint8_t* lut,
                                       uint64_t blocks, int32_t* out);
/* Per-query intermediates of the query path: rotated centered query
 * [nq*padded], u8 LUT [nq*blocks*64], LUT constants [nq*4] = {delta, base,
 * lut_err, score_max}, probe ids [nq*nprobe] and all centroid distances
 * [nq*nlist] (FhtKac4Rotator::Rotate, PrepareQuery, RouteQuery). */
e8q_status e8q_debug_query_prep(e8q_ctx* ctx, const e8q_dindex* idx, const float* queries, uint64_t nq,
                                uint64_t nprobe, float* rot_q, uint8_t* lut, float* consts, uint32_t* probe,
                                float* cdist);
/* Device rotation of nrows vectors with the index's rotator (no centering). */
e8q_status e8q_debug_rotate(e8q_ctx* ctx, const e8q_dindex* idx, const float* x, uint64_t nrows, float* out);
/* Device E8 encode of unit vectors [nrows*padded] at scale s. */
e8q_status e8q_debug_encode(e8q_ctx* ctx, const e8q_dindex* idx, const float* unit, uint64_t nrows, double s,
                            uint16_t* digits, uint8_t* branch);

#ifdef __cplusplus
}
#endif

#endif /* E8Q_GPU_H_ */
