// me8q_runtime.hip -- the ME8Q half of libe8q_gpu.so's C ABI
// (include/me8q_gpu.h): device index lifetime, upload / download, the
// device build, search, and the test diagnostics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/me8q_gpu.h"
#include "diag_kernels.h"
#include "e8q_kernels.h"
#include "e8q_rt.h"
#include "me8q_device.h"
#include "me8q_kernels.h"

using namespace e8qdev;
using namespace e8qrt;

struct e8q_m8index {
    e8q_ctx* ctx = nullptr;
    M8View view{};
    std::vector<M8Run> runs;
    std::vector<M8Seg> segs;
    std::vector<uint64_t> list_begin, batch_begin;
    std::vector<uint64_t> top_lists;   // [nlist+1]: vectors in the c largest lists
    uint64_t nbatch = 0;
    uint64_t sum_nbt = 0, max_nbt = 0;   // batch tiles of the batched form, summed and largest over the lists
    // refine-row layout of the reference (for the diagnostics)
    std::vector<uint32_t> ref_run_off;
    uint32_t ref_stride = 0;
    DBuf d_runs, d_segs, d_list_begin, d_batch_begin, ids, codes, centroids, cent_rot, cn2, rc_n2, plane, refine,
        signs, lev, rrec, bscale, bcnt, pids;
    std::vector<DBuf> enc;   // encoder tables, 5 per distinct rate
    uint64_t device_bytes() const {
        uint64_t b = d_runs.bytes + d_segs.bytes + d_list_begin.bytes + d_batch_begin.bytes + ids.bytes + codes.bytes +
                     centroids.bytes + cent_rot.bytes + cn2.bytes + rc_n2.bytes + plane.bytes + refine.bytes +
                     signs.bytes + lev.bytes + rrec.bytes + bscale.bytes + bcnt.bytes + pids.bytes;
        for (const auto& e : enc) b += e.bytes;
        return b;
    }
};

namespace {

void RefreshM8View(e8q_m8index* x) {
    M8View& v = x->view;
    v.runs = x->d_runs.as<M8Run>();
    v.segs = x->d_segs.as<M8Seg>();
    v.list_begin = x->d_list_begin.as<uint64_t>();
    v.batch_begin = x->d_batch_begin.as<uint64_t>();
    v.ids = x->ids.as<uint32_t>();
    v.codes = x->codes.as<uint8_t>();
    v.centroids = x->centroids.as<float>();
    v.cent_rot = x->cent_rot.as<float>();
    v.cn2 = x->cn2.as<float>();
    v.rc_n2 = x->rc_n2.as<double>();
    v.plane = x->plane.as<uint8_t>();
    v.refine = x->refine.as<uint8_t>();
    v.rrec = x->rrec.as<uint8_t>();
    v.bscale = x->bscale.as<float4>();
    v.bcnt = x->bcnt.as<uint8_t>();
    v.pids = x->pids.as<uint32_t>();
}

// Rotated centroids + FastScan layout from codes / centroids already on the device.
void BuildDerived(e8q_ctx* ctx, e8q_m8index* x) {
    hipStream_t st = ctx->stream;
    const M8View& v = x->view;
    Check(LaunchM8Centroids(st, v), "centroid kernel");
    Check(hipMemsetAsync(x->plane.p, 0, x->plane.bytes, st), "plane clear");
    if (v.n > 0 && v.nrun > 0) {
        DBuf perp;
        perp.alloc(v.n * v.nrun * sizeof(float));
        Check(LaunchM8Fast(st, v, perp.as<float>(), x->nbatch), "fast-layout kernels");
        if (v.rrec) Check(LaunchM8RefineRecords(st, v, x->nbatch), "refine records");
        Check(hipStreamSynchronize(st), "fast layout");
    } else {
        Check(hipStreamSynchronize(st), "centroids");
    }
}

// Device scratch of one search, carved from the context workspace.
struct M8Bufs {
    float *dq, *rotq, *qc, *dist, *pd, *od;
    double *rc, *sqd2;
    uint8_t* lut;
    uint32_t *probe, *npq, *oid;
    uint64_t* qst;
    uint32_t *qord, *qkey, *qhist;   // the plain scan's query order and its scratch
    // batched form
    uint32_t *fid, *ccount, *cfill, *coff, *ioff, *counters, *ncand, *nfill, *ncont, *toff, *toff2, *chk, *hist;
    uint32_t *abin, *na, *fill_a, *fill_b;
    uint64_t* coff_b;
    float *fd, *tau;
    uint64_t *cover, *pairs, *poff;
    uint8_t* lut8;
    uint64_t *pool, *cont, *tpool;
    float* pdlo;
    uint32_t *tsurv, *ttsurv, *ovf;
    uint2* rec;
    BatchItem* items;
    unsigned long long *wstats, *bstats, *rcount;
    uint2* itab;
    BatchedParams* wparams;
};

// Capacities of the batched form's candidate pool and count-pass records,
// fixed before its first launch, so a search never waits on the host
// between its kernels.  The pool holds
//   min(2^31, the vectors the probed lists can hold, nq x per-query entries),
// per-query entries = 65536 (E8Q_M8_POOL_PER_QUERY overrides it), raised
// after a call that needed more (to 1.25 x its need) unless the context's
// workspace was reserved (e8q_ctx_reserve): a reserved context gives the
// pool all of the reservation the rest of the call leaves (up to what the
// probed lists can hold) and never grows.  The records hold as many as
// the pool (E8Q_M8_RECORD_CAP lowers it; 0 = always the write pass).  Queries
// whose candidates pass the pool's end are searched by the plain form.
constexpr uint64_t kM8PoolMax = uint64_t{1} << 31;
constexpr uint64_t kM8PoolPerQuery = 65536;
constexpr uint32_t kM8DefaultBudget = 16;
constexpr uint32_t kM8MaxBudget = 1024;    // threshold-pass batches per query at most
constexpr uint64_t kM8TuneMinQueries = 64;   // smaller calls run the default grids untuned

uint64_t EnvU64(const char* name, uint64_t dflt) {
    const char* e = std::getenv(name);
    return e ? std::strtoull(e, nullptr, 10) : dflt;
}

// Whether the batched form runs for this index and k.
bool M8BatchedFits(const e8q_m8index* x, uint64_t k) {
    const M8View& v = x->view;
    if (v.nrun < 1 || v.nrun > kM8BatchRuns || k > kM8SelMaxK || (v.plane_cols & 7u) != 0) return false;
    if (v.n > 0 && v.rrec == nullptr) return false;   // more batch slots than 32-bit words name
    uint32_t cols = 0;
    for (const auto& r : x->runs) cols += r.nblk;
    return cols == v.plane_cols;
}

// Work items of the batched form: an upper bound of their count (the list
// count of each pair tile times the batch tiles of its list).
uint64_t M8ItemsCap(const e8q_m8index* x, uint64_t nq, uint64_t pairs) {
    const uint64_t qt = kM8BatchQTile;
    return std::min((nq + qt - 1) / qt * x->sum_nbt, (pairs + qt - 1) / qt * x->max_nbt + x->sum_nbt);
}

struct M8Caps {
    uint64_t pool = 0;     // candidate pool entries
    uint64_t rec = 0;      // count-pass records
    uint64_t tasks = 0;    // refine tasks of one launch at most
    uint64_t tpool = 0;    // threshold-pass entries (budget x 32 per query)
    uint64_t ttasks = 0;   // its tasks
};

// Pool entries a call of nq queries probing up to maxp lists asks for (the
// default, or what the context learned), before any reservation cap.
uint64_t M8PoolWant(const e8q_m8index* x, uint64_t nq, uint32_t maxp, uint64_t per_q_learned) {
    const uint64_t per_q = std::min(kM8PoolMax, std::max(EnvU64("E8Q_M8_POOL_PER_QUERY", kM8PoolPerQuery),
                                                         per_q_learned));
    const uint64_t reach = nq * x->top_lists[std::min<uint64_t>(maxp, x->view.nlist)];
    return std::min({kM8PoolMax, reach, nq * per_q});
}

M8Caps M8CapsFor(uint64_t nq, uint64_t pool, uint32_t budget) {
    M8Caps c;
    c.pool = pool;
    c.rec = std::min(pool, EnvU64("E8Q_M8_RECORD_CAP", pool));
    c.tasks = pool / kM8RefineChunk + nq + 1;
    c.tpool = nq * budget * 32ull;
    c.ttasks = nq * ((budget * 32ull + kM8RefineChunk - 1) / kM8RefineChunk) + 1;
    return c;
}

void M8CarveBatched(e8q_ctx::Carve& cv, M8Bufs& b, const e8q_m8index* x, uint64_t nq, uint64_t k, uint64_t pairs,
                    bool chk, const M8Caps& caps) {
    const M8View& v = x->view;
    cv.add(&b.fid, nq * k);
    cv.add(&b.lut8, nq * v.plane_cols * 32);
    cv.add(&b.fd, nq * k);
    cv.add(&b.cover, nq);
    cv.add(&b.ccount, v.nlist);
    cv.add(&b.cfill, v.nlist);
    cv.add(&b.coff, v.nlist + 1);
    cv.add(&b.ioff, v.nlist + 1);
    cv.add(&b.pairs, std::max<uint64_t>(pairs, 1));
    cv.add(&b.items, std::max<uint64_t>(M8ItemsCap(x, nq, pairs), 1));
    cv.add(&b.counters, 4);
    cv.add(&b.ncand, nq);
    cv.add(&b.nfill, nq);
    cv.add(&b.tau, nq);
    cv.add(&b.hist, nq * kM8TauBins);
    cv.add(&b.ncont, nq);
    cv.add(&b.abin, nq);
    cv.add(&b.na, nq);
    cv.add(&b.fill_a, nq);
    cv.add(&b.fill_b, nq);
    cv.add(&b.coff_b, nq);
    cv.add(&b.poff, nq + 1);
    cv.add(&b.toff, nq + 1);
    cv.add(&b.toff2, nq + 1);
    cv.add(&b.wstats, 4);
    cv.add(&b.bstats, 8);
    cv.add(&b.rcount, 1);
    cv.add(&b.itab, std::max<uint64_t>(M8ItemsCap(x, nq, pairs), 1) * kM8BatchQTile);
    cv.add(&b.wparams, 1);
    cv.add(&b.ovf, nq);
    cv.add(&b.pool, std::max<uint64_t>(caps.pool, 1));
    cv.add(&b.cont, std::max<uint64_t>(caps.pool, 1));
    cv.add(&b.pdlo, std::max<uint64_t>(caps.pool, 1));
    cv.add(&b.tsurv, 4 * caps.tasks);
    if (caps.rec > 0) cv.add(&b.rec, caps.rec);
    cv.add(&b.tpool, std::max<uint64_t>(caps.tpool, 1));
    cv.add(&b.ttsurv, caps.ttasks);
    if (chk) cv.add(&b.chk, nq);
}

void M8Carve(e8q_ctx::Carve& cv, M8Bufs& b, const M8View& v, uint64_t nq, uint64_t k, uint64_t maxp, bool q_host,
             bool out_host, bool per_query_nprobe) {
    const uint64_t R1 = std::max<uint64_t>(v.nrun, 1);
    if (q_host) cv.add(&b.dq, nq * v.dim);
    cv.add(&b.rotq, nq * v.padded);
    cv.add(&b.lut, nq * v.plane_cols * 32 + 16);
    cv.add(&b.qc, nq * 4);
    cv.add(&b.rc, nq * R1 * 4);
    cv.add(&b.dist, nq * v.nlist);
    cv.add(&b.probe, nq * maxp);
    cv.add(&b.pd, nq * maxp);
    cv.add(&b.sqd2, nq * maxp * R1);
    if (per_query_nprobe) cv.add(&b.npq, nq);
    if (out_host) {
        cv.add(&b.oid, nq * k);
        cv.add(&b.od, nq * k);
    }
    cv.add(&b.qst, nq * 8);
    cv.add(&b.qord, nq);
    cv.add(&b.qkey, nq);
    cv.add(&b.qhist, 2 * kOrderBuckets);
}

// Graph-replay key words of an M8View.
static_assert(sizeof(M8View) % 8 == 0, "M8View is compared as 64-bit words");

uint64_t Bucket(uint64_t n) {
    uint64_t b = 0;
    while ((uint64_t{1} << b) < n) ++b;
    return b;
}

uint64_t F32Bits(float f) {
    uint32_t u = 0;
    std::memcpy(&u, &f, 4);
    return u;
}

// One search: plain (the reference's per-query schedule) or batched.  chk
// (host [nq], batched only): the coarse-sum checksum of the diagnostic.
//
// The batched form is one device sequence with no host decision inside:
// its buffers are carved before the first launch, a query whose candidates
// would pass the pool's end is listed on the device and searched by the
// plain form at the end of the same sequence, and whether the count pass's
// records fit (placement from them, else a second contraction) is read by
// the kernels themselves.  So the sequence can be captured as a HIP graph,
// and its persistent grids timed and chosen (results do not depend on them:
// every per-query threshold is fixed before the kernel that applies it, and
// each top-k is sorted by (distance, id)).
void M8Search(e8q_ctx* ctx, const e8q_m8index* x, const e8q_m8_search_args* a, e8q_m8_stats* stats,
              uint64_t* per_query_stats, uint32_t* chk) {
    if (!ctx || !x || !a) Throw(E8Q_EINVAL, "null argument");
    if (a->k == 0 || a->k > kM8MaxK) Throw(E8Q_EINVAL, "k must be in [1, 1024]");
    if (!a->queries || !a->out_ids || !a->out_dists) Throw(E8Q_EINVAL, "null buffer");
    if (a->mode != E8Q_MODE_PLAIN && a->mode != E8Q_MODE_BATCHED) Throw(E8Q_EINVAL, "mode must be plain or batched");
    if (a->batched_budget > kM8MaxBudget) Throw(E8Q_EINVAL, "batched_budget must be <= 1024");
    const M8View& v = x->view;
    const uint64_t nq = a->nq;
    if (stats) {
        *stats = e8q_m8_stats{};
        stats->form = E8Q_MODE_PLAIN;
    }
    if (nq == 0) return;
    uint32_t maxp = 0;
    uint64_t pairs = 0;
    if (a->nprobe_per_query) {
        for (uint64_t i = 0; i < nq; ++i) {
            if (a->nprobe_per_query[i] == 0) Throw(E8Q_EINVAL, "nprobe must be positive");
            maxp = std::max<uint32_t>(maxp, a->nprobe_per_query[i]);
            pairs += std::min<uint64_t>(a->nprobe_per_query[i], v.nlist);
        }
    } else {
        if (a->nprobe == 0) Throw(E8Q_EINVAL, "nprobe must be positive");
        maxp = static_cast<uint32_t>(std::min<uint64_t>(a->nprobe, v.nlist));
        pairs = nq * maxp;
    }
    maxp = static_cast<uint32_t>(std::min<uint64_t>(maxp, v.nlist));
    const bool batched = a->mode == E8Q_MODE_BATCHED && M8BatchedFits(x, a->k);
    ctx->Bind();
    hipStream_t st = ctx->stream;
    const uint64_t P = v.padded;
    const uint32_t cus = static_cast<uint32_t>(std::max(ctx->cus, 1));
    const uint32_t budget = a->batched_budget ? a->batched_budget : kM8DefaultBudget;
    const uint32_t set_a = a->batched_set_a ? a->batched_set_a : kM8SetA;
    M8Bufs b{};
    M8Caps caps;
    {
        e8q_ctx::Carve cv;
        M8Carve(cv, b, v, nq, a->k, maxp, !a->queries_on_device, !a->outputs_on_device, a->nprobe_per_query != nullptr);
        if (batched) {
            // A reserved context offers its whole reservation to the pool (up
            // to what the probed lists can hold): a small call in a
            // workspace sized for a large one may have a query whose
            // candidates alone pass the per-query default.
            uint64_t pool = ctx->reserved
                                ? std::min(kM8PoolMax, nq * x->top_lists[std::min<uint64_t>(maxp, v.nlist)])
                                : M8PoolWant(x, nq, maxp, ctx->m8_pool_per_q);
            auto bytes_at = [&](uint64_t pl) {
                e8q_ctx::Carve c2 = cv;
                M8Bufs b2{};
                M8CarveBatched(c2, b2, x, nq, a->k, pairs, chk != nullptr, M8CapsFor(nq, pl, budget));
                return static_cast<uint64_t>(c2.bytes());
            };
            // a reserved workspace is never grown for the pool: the largest
            // pool that fits it (down to none: then the plain form answers)
            if (ctx->reserved && bytes_at(pool) > ctx->ws.bytes && bytes_at(0) <= ctx->ws.bytes) {
                uint64_t lo = 0, hi = pool;
                while (lo < hi) {
                    const uint64_t mid = lo + (hi - lo + 1) / 2;
                    if (bytes_at(mid) <= ctx->ws.bytes) lo = mid;
                    else hi = mid - 1;
                }
                pool = lo;
            }
            caps = M8CapsFor(nq, pool, budget);
            M8CarveBatched(cv, b, x, nq, a->k, pairs, chk != nullptr, caps);
        }
        ctx->Take(cv);
    }
    float *dq = b.dq, *rotq = b.rotq, *qc = b.qc, *dist = b.dist, *pd = b.pd, *od = b.od;
    double *rc = b.rc, *sqd2 = b.sqd2;
    uint8_t* lut = b.lut;
    uint32_t *probe = b.probe, *npq = b.npq, *oid = b.oid;
    uint64_t* qst = b.qst;
    const float* q = a->queries_on_device ? a->queries : dq;
    const uint32_t nprobe_all = static_cast<uint32_t>(std::min<uint64_t>(a->nprobe, v.nlist));
    RouteArgs ra{};
    ra.rotq = rotq;
    ra.qstride = static_cast<uint32_t>(P);
    ra.qn2 = qc + 2;
    ra.qn2_stride = 4;
    ra.cent = v.cent_rot;
    ra.cn2 = v.cn2;
    ra.nlist = v.nlist;
    ra.dim = v.route_dim;
    ra.cstride = static_cast<uint32_t>(P);
    ra.nq = nq;
    ra.dist = dist;
    uint32_t* out_ids = a->outputs_on_device ? a->out_ids : oid;
    float* out_d = a->outputs_on_device ? a->out_dists : od;
    M8ScanArgs sa{};
    sa.v = v;
    sa.rotq = rotq;
    sa.lut = lut;
    sa.qc = qc;
    sa.rc = rc;
    sa.probe = probe;
    sa.pdist = pd;
    sa.sqd2 = sqd2;
    sa.nprobe_q = npq;
    sa.nprobe_all = nprobe_all;
    sa.maxp = maxp;
    sa.nq = nq;
    sa.k = static_cast<uint32_t>(a->k);
    sa.z = a->z;
    sa.out_ids = out_ids;
    sa.out_d = out_d;
    sa.qstats = qst;
    sa.lev_floats = 0;
    sa.run_slot = 16;
    for (const auto& r : x->runs) {
        if (r.rate <= kM8LdsLevRate) sa.lev_floats += 2u << r.rate;
        sa.run_slot = std::max(sa.run_slot, (r.rate * r.nblk + 15) & ~15u);
    }

    // Arguments of the batched kernels (pointers into the carve).
    IndexView lv{};   // the work-list kernels read the list count and offsets only
    BatchedArgs w{};
    M8BatchArgs ba{};
    if (batched) {
        lv.nlist = v.nlist;
        lv.cluster_begin = v.list_begin;
        w.nq = nq;
        w.maxp = maxp;
        w.nprobe_q = npq;
        w.nprobe_all = nprobe_all;
        w.probe_ids = probe;
        w.k = static_cast<uint32_t>(a->k);
        w.ccount = b.ccount;
        w.cfill = b.cfill;
        w.coff = b.coff;
        w.ioff = b.ioff;
        w.pairs = b.pairs;
        w.items = b.items;
        w.items_cap = M8ItemsCap(x, nq, pairs);
        w.counters = b.counters;
        w.qtile = kM8BatchQTile;
        w.btile = kM8BatchBTile;
        w.stats = b.wstats;
        w.p_lo = 0;
        ba.v = v;
        ba.lut = lut;
        ba.lut8 = b.lut8;
        ba.qc = qc;
        ba.rc = rc;
        ba.rotq = rotq;
        ba.probe = probe;
        ba.pdist = pd;
        ba.sqd2 = sqd2;
        ba.maxp = maxp;
        ba.k = static_cast<uint32_t>(a->k);
        ba.nq = nq;
        ba.z = a->z;
        ba.cascade = 1;
        ba.first_ids = b.fid;
        ba.first_d = b.fd;
        ba.cover = b.cover;
        ba.items = b.items;
        ba.pairs = b.pairs;
        ba.items_cap = w.items_cap;
        ba.counters = b.counters;
        ba.ncand = b.ncand;
        ba.nfill = b.nfill;
        ba.tau = b.tau;
        ba.hist = b.hist;
        ba.abin = b.abin;
        ba.na = b.na;
        ba.fill_a = b.fill_a;
        ba.fill_b = b.fill_b;
        ba.coff_b = b.coff_b;
        ba.coff = b.poff;
        ba.toff = b.toff;
        ba.out_ids = out_ids;
        ba.out_d = out_d;
        ba.stats = b.bstats;
        ba.chk = chk ? b.chk : nullptr;
        ba.lev_floats = sa.lev_floats;
        {   // the level tables the refine stages in LDS: run by run while they fit
            uint32_t o = 0;
            for (uint32_t r = 0; r < kM8BatchRuns; ++r) ba.ref_lev_off[r] = kM8LevGlobal;
            for (uint32_t r = 0; r < v.nrun && r < kM8BatchRuns; ++r) {
                const uint32_t nq2 = 2u << x->runs[r].rate;
                if (o + nq2 <= kM8RefLevMax) {
                    ba.ref_lev_off[r] = o;
                    o += nq2;
                }
            }
            ba.ref_lev_floats = o;
        }
        ba.set_a = set_a;
        ba.rec = caps.rec > 0 ? b.rec : nullptr;
        ba.rec_cap = caps.rec;
        ba.rcount = b.rcount;
        ba.itab = b.itab;
        ba.pool = b.pool;
        ba.cont = b.cont;
        ba.pdlo = b.pdlo;
        ba.ncont = b.ncont;
    }

    // Grids of the persistent batched kernels: fixed, tuned before, or tuned
    // in this call (cands non-empty).
    struct Grids {
        uint32_t mfma, refine;
    };
    Grids grids{ctx->m8_fix_mfma ? ctx->m8_fix_mfma : 2 * cus, ctx->m8_fix_refine ? ctx->m8_fix_refine : 8 * cus};
    const uint64_t tkey[6] = {v.n, v.plane_cols, v.nlist, Bucket(nq), Bucket(maxp), Bucket(a->k)};
    bool tune = false;
    if (batched && (ctx->m8_fix_mfma == 0 || ctx->m8_fix_refine == 0)) {
        bool found = false;
        {
            std::lock_guard<std::mutex> lk(ctx->m8_table->mu);
            for (const auto& t : ctx->m8_table->v)
                if (std::equal(t.shape, t.shape + 6, tkey)) {
                    if (!ctx->m8_fix_mfma) grids.mfma = t.mfma_grid;
                    if (!ctx->m8_fix_refine) grids.refine = t.refine_grid;
                    found = true;
                }
        }
        tune = !found && nq >= kM8TuneMinQueries && chk == nullptr;
    }

    // The refine's launches, each between two of the context's kernel events
    // (eager, timed runs only): kind 0 = first-run phase, 1 = cascades.
    int nkl = 0, kind[e8q_ctx::kKernelEvents / 2] = {};
    // The device sequence from the query rotation to the results.  eager:
    // events around the scan and the refine launches, and the work list's
    // parameter block uploaded on the stream (a capture records kernels
    // only: the block is uploaded before the graph launch).
    auto device_seq = [&](const Grids& g, bool eager, bool timed, bool fallback) {
        if (eager) Check(hipEventRecord(ctx->ev[0], st), "event");
        Check(LaunchM8QueryPrep(st, v, q, nq, rotq, lut, qc, rc), "query prep");
        Check(LaunchRouteSelect(st, ra, npq, nprobe_all, maxp, probe, pd), "route");
        Check(LaunchM8RunDist(st, v, rotq, rc, probe, npq, nprobe_all, maxp, nq, sqd2), "run distances");
        if (!batched) {
            Check(LaunchQueryOrder(st, v.batch_begin, v.nlist, probe, npq, nprobe_all, maxp, nq, b.qkey, b.qhist, b.qord),
                  "query order");
            M8ScanArgs so = sa;   // the queries heaviest first
            so.qsel = b.qord;
            so.nsel = nq;
            if (eager) Check(hipEventRecord(ctx->ev[1], st), "event");
            Check(LaunchM8Scan(st, so), "scan");
            if (eager) Check(hipEventRecord(ctx->ev[2], st), "event");
            return;
        }
        auto refine_launch = [&](const M8BatchArgs& r, bool first, const char* what) {
            const bool t = timed && nkl < e8q_ctx::kKernelEvents / 2;
            if (t) Check(hipEventRecord(ctx->kev[2 * nkl], st), "event");
            Check(LaunchM8Refine(st, r, g.refine, first), what);
            if (t) {
                Check(hipEventRecord(ctx->kev[2 * nkl + 1], st), "event");
                kind[nkl++] = first ? 0 : 1;
            }
        };
        // the (query, list) work list: every probe, grouped by list
        Check(LaunchFill32(st, b.ccount, v.nlist, 0), "clear");
        Check(LaunchFill32(st, b.ncand, nq, 0), "clear");
        Check(LaunchFill32(st, reinterpret_cast<uint32_t*>(b.wstats), 8, 0), "clear");
        Check(LaunchFill32(st, reinterpret_cast<uint32_t*>(b.bstats), 16, 0), "clear");
        if (chk) Check(LaunchFill32(st, b.chk, nq, 0), "clear");
        Check(LaunchBatchedWorkList(st, lv, w, b.wparams, eager), "work list");
        Check(LaunchFill32(st, b.nfill, nq, 0), "clear");
        Check(LaunchFill32(st, b.hist, nq * kM8TauBins, 0), "clear");
        Check(LaunchFill32(st, b.fill_a, nq, 0), "clear");
        Check(LaunchFill32(st, b.fill_b, nq, 0), "clear");
        {   // the threshold pass: exact distances of each query's first `budget`
            // batches (whole cascade, no death test, tau = +inf) and their top-k
            M8BatchArgs ta = ba;
            ta.pool = b.tpool;
            ta.tsurv = b.ttsurv;
            ta.coff = b.coff_b;
            ta.nfill = b.nfill;
            ta.toff = b.toff;
            ta.tau = b.tau;
            ta.cascade = 0;
            ta.first_ids = ta.out_ids = b.fid;
            ta.first_d = ta.out_d = b.fd;
            Check(LaunchFill32(st, b.fid, nq * a->k, 0xFFFFFFFFu), "clear");
            Check(LaunchFill32(st, reinterpret_cast<uint32_t*>(b.fd), nq * a->k, 0x7F800000u), "clear");
            Check(LaunchFill32(st, reinterpret_cast<uint32_t*>(b.tau), nq, 0x7F800000u), "clear");
            Check(LaunchFill32(st, reinterpret_cast<uint32_t*>(qst), nq * 16, 0), "clear");
            Check(LaunchM8ThresholdPool(st, ta, npq, nprobe_all, budget), "threshold pool");
            Check(LaunchM8RefineTasks(st, ta, b.nfill, false), "threshold tasks");
            refine_launch(ta, false, "threshold refine");
            Check(LaunchM8Select(st, ta), "threshold select");
        }
        // 3. the MFMA contraction and the coarse bound: a count pass (with
        // its records when they fit), the pool offsets, the pool's fit
        Check(LaunchFill32(st, reinterpret_cast<uint32_t*>(b.rcount), 2, 0), "clear");
        Check(LaunchM8LutInt8(st, lut, b.lut8, nq * v.plane_cols * 32), "int8 LUT");
        if (eager) Check(hipEventRecord(ctx->ev[1], st), "event");
        Check(LaunchM8MfmaScan(st, ba, g.mfma, false), "mfma count pass");
        if (eager) Check(hipEventRecord(ctx->ev[2], st), "event");
        Check(LaunchM8Split(st, ba), "candidate split");
        Check(LaunchM8RefineTasks(st, ba, b.ncand, true), "pool offsets");
        Check(LaunchM8Fit(st, ba, caps.pool, b.ovf, b.counters + 2), "pool fit");
        // set A placed from the records, or (when they did not fit) every
        // candidate written by a second contraction; each kernel reads which
        Check(LaunchM8Place(st, ba, 8 * cus, false), "placement of set A");
        Check(LaunchFill32(st, b.counters + 1, 1, 0), "clear");
        Check(LaunchM8MfmaScan(st, ba, g.mfma, true), "mfma write pass");
        // 4. exact cascades (every candidate's first run, then the whole
        // cascade of those still alive) and a top-k merge, first of each
        // query's set A (its exact top-k gives tau1), then of the rest whose
        // lower bound is below tau1
        const uint64_t tsn = caps.tasks;
        auto refine = [&](M8BatchArgs r, uint32_t* tsurv1, uint32_t* tsurv2, uint32_t* sel_ids, float* sel_d) {
            r.tsurv = tsurv1;
            r.toff = b.toff;
            Check(LaunchFill32(st, b.ncont, nq, 0), "clear");
            Check(LaunchM8RefineTasks(st, r, r.nfill, false), "refine tasks");
            refine_launch(r, true, "refine, first runs");
            if (v.nrun > 1) {
                r.pool = r.cont;
                r.nfill = b.ncont;
                r.toff = b.toff2;
                r.tsurv = tsurv2;
                Check(LaunchM8RefineTasks(st, r, b.ncont, false), "refine tasks");
                refine_launch(r, false, "refine, cascades");
            }
            r.out_ids = sel_ids;
            r.out_d = sel_d;
            Check(LaunchM8Select(st, r), "select");
        };
        M8BatchArgs rA = ba;   // set A
        rA.nfill = b.na;
        refine(rA, b.tsurv, b.tsurv + tsn, b.fid, b.fd);
        Check(LaunchM8Place(st, ba, 8 * cus, true), "placement of the rest");
        Check(LaunchM8Compact(st, ba), "compaction");
        M8BatchArgs rB = ba;   // the rest
        rB.coff = b.coff_b;
        refine(rB, b.tsurv + 2 * tsn, b.tsurv + 3 * tsn, out_ids, out_d);
        // queries whose candidates passed the pool's end: the plain form
        // (their candidates and batches are counted already)
        if (!fallback) return;
        M8ScanArgs oa = sa;
        oa.qsel = b.ovf;
        oa.nsel = nq;
        oa.nsel_dev = b.counters + 2;
        oa.qstats = nullptr;
        Check(LaunchM8Scan(st, oa), "plain search of overflowed queries");
    };
    // An unreserved context whose pool can still grow answers an overflow
    // by learning the need and running the whole call again (outside a
    // graph, where the host decides after the sequence); the plain
    // fallback is then left out of the first run.
    const uint64_t pool_ceiling = std::min(kM8PoolMax, nq * x->top_lists[std::min<uint64_t>(maxp, v.nlist)]);
    const bool can_grow = batched && !ctx->reserved && chk == nullptr && caps.pool < pool_ceiling;
    auto learn = [&](uint64_t total) {
        ctx->m8_pool_per_q = std::max<uint64_t>(ctx->m8_pool_per_q, (total + total / 4) / nq + 1);
    };

    // Graph replay: the key is everything the captured launches bake in.
    e8q_ctx::GraphEntry* ge = nullptr;
    if (ctx->graphs && !tune && chk == nullptr) {
        std::vector<uint64_t> key{0x4D38u,   // the ME8Q search
                                  reinterpret_cast<uint64_t>(x),
                                  static_cast<uint64_t>(batched ? 1 : 0),
                                  nq,
                                  a->k,
                                  maxp,
                                  nprobe_all,
                                  npq != nullptr ? 1u : 0u,
                                  F32Bits(a->z),
                                  budget,
                                  set_a,
                                  grids.mfma,
                                  grids.refine,
                                  reinterpret_cast<uint64_t>(q),
                                  reinterpret_cast<uint64_t>(out_ids),
                                  reinterpret_cast<uint64_t>(out_d),
                                  reinterpret_cast<uint64_t>(ctx->ws.p),
                                  ctx->ws.bytes,
                                  caps.pool,
                                  caps.rec,
                                  pairs};
        const size_t w0 = key.size();
        key.resize(w0 + sizeof(M8View) / 8);
        std::memcpy(key.data() + w0, &v, sizeof(M8View));
        ge = ctx->FindGraph(key);
        if (!ge) ctx->AddGraph(key);   // first call of this shape runs eagerly
    }
    if (!a->queries_on_device)
        Check(hipMemcpyAsync(dq, a->queries, nq * v.dim * 4, hipMemcpyHostToDevice, st), "upload queries");
    if (npq) Check(hipMemcpyAsync(npq, a->nprobe_per_query, nq * 4, hipMemcpyHostToDevice, st), "upload nprobe");
    bool replayed = false;
    if (ge) {
        if (batched) {
            const BatchedParams hw{lv, w};
            Check(hipMemcpyAsync(b.wparams, &hw, sizeof(hw), hipMemcpyHostToDevice, st), "upload parameters");
        }
        if (!ge->exec) {
            Check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
            try {
                device_seq(grids, false, false, true);
            } catch (...) {
                hipGraph_t gr = nullptr;
                (void)hipStreamEndCapture(st, &gr);
                if (gr) (void)hipGraphDestroy(gr);
                throw;
            }
            hipGraph_t gr = nullptr;
            Check(hipStreamEndCapture(st, &gr), "end capture");
            hipGraphExec_t ex = nullptr;
            const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            Check(e, "graph instantiate");
            ge->exec = ex;
            ++ctx->captures;
        }
        Check(hipEventRecord(ctx->ev[0], st), "event");
        Check(hipEventRecord(ctx->ev[1], st), "event");
        Check(hipGraphLaunch(ge->exec, st), "graph launch");
        Check(hipEventRecord(ctx->ev[2], st), "event");
        ++ctx->replays;
        replayed = true;
    } else {
        if (tune && can_grow) {   // size the pool before timing anything
            device_seq(grids, true, false, false);
            unsigned long long t0 = 0;
            Check(hipMemcpyAsync(&t0, b.poff + nq, 8, hipMemcpyDeviceToHost, st), "download pool size");
            Check(hipStreamSynchronize(st), "sizing run");
            if (t0 > caps.pool) {
                learn(t0);
                M8Search(ctx, x, a, stats, per_query_stats, chk);
                return;
            }
        }
        if (tune) {
            // Times the whole sequence at each candidate grid (the MFMA
            // contraction's at 1, 2 and 4 per CU, then the refine's at 4, 8
            // and 16 per CU at the best of those); each run writes the same
            // results.
            auto time_grids = [&](const Grids& g) {
                Check(hipEventRecord(ctx->kev[0], st), "event");
                device_seq(g, true, false, !can_grow);
                Check(hipEventRecord(ctx->kev[1], st), "event");
                Check(hipEventSynchronize(ctx->kev[1]), "tuning");
                float ms = 0.0F;
                Check(hipEventElapsedTime(&ms, ctx->kev[0], ctx->kev[1]), "event time");
                return ms;
            };
            Grids best = grids;
            float best_ms = time_grids(best);
            if (!ctx->m8_fix_mfma)
                for (uint32_t m : {cus, 4 * cus}) {
                    const Grids c{m, best.refine};
                    const float ms = time_grids(c);
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = c;
                    }
                }
            if (!ctx->m8_fix_refine)
                for (uint32_t r : {4 * cus, 16 * cus}) {
                    const Grids c{best.mfma, r};
                    const float ms = time_grids(c);
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = c;
                    }
                }
            {
                std::lock_guard<std::mutex> lk(ctx->m8_table->mu);
                ctx->m8_table->v.push_back(
                    {{tkey[0], tkey[1], tkey[2], tkey[3], tkey[4], tkey[5]}, best.mfma, best.refine});
            }
            ++ctx->m8_tunings;
            grids = best;
        }
        device_seq(grids, true, stats != nullptr, !can_grow);
    }
    Check(hipEventRecord(ctx->ev[3], st), "event");
    if (batched) {
        ctx->m8_last_mfma = grids.mfma;
        ctx->m8_last_refine = grids.refine;
        ctx->m8_last_budget = budget;
        ctx->m8_last_set_a = set_a;
    }
    std::vector<unsigned long long> bst(8, 0), wst(4, 0);
    uint32_t cnt[4] = {0, 0, 0, 0};
    unsigned long long nrec = 0, total = 0;
    if (batched) {
        Check(hipMemcpyAsync(bst.data(), b.bstats, 8 * 8, hipMemcpyDeviceToHost, st), "download stats");
        Check(hipMemcpyAsync(wst.data(), b.wstats, 4 * 8, hipMemcpyDeviceToHost, st), "download stats");
        Check(hipMemcpyAsync(cnt, b.counters, sizeof(cnt), hipMemcpyDeviceToHost, st), "download counters");
        Check(hipMemcpyAsync(&nrec, b.rcount, 8, hipMemcpyDeviceToHost, st), "download record count");
        Check(hipMemcpyAsync(&total, b.poff + nq, 8, hipMemcpyDeviceToHost, st), "download pool size");
        if (chk) Check(hipMemcpyAsync(chk, b.chk, nq * 4, hipMemcpyDeviceToHost, st), "download checksums");
    }
    if (!a->outputs_on_device) {
        Check(hipMemcpyAsync(a->out_ids, oid, nq * a->k * 4, hipMemcpyDeviceToHost, st), "download ids");
        Check(hipMemcpyAsync(a->out_dists, od, nq * a->k * 4, hipMemcpyDeviceToHost, st), "download dists");
    }
    std::vector<uint64_t> hs(nq * 8);
    Check(hipMemcpyAsync(hs.data(), qst, nq * 8 * 8, hipMemcpyDeviceToHost, st), "download stats");
    Check(hipStreamSynchronize(st), "search");
    uint64_t novf = 0;
    if (batched) {
        if (cnt[0] > w.items_cap) Throw(E8Q_ERUNTIME, "batched search: work items exceed their bound");
        if (cnt[2] > nq) Throw(E8Q_ERUNTIME, "batched search: overflow count exceeds the query count");
        if (ba.rec != nullptr && nrec <= caps.rec && nrec != total)
            Throw(E8Q_ERUNTIME, "batched search: records and candidate counts differ");
        novf = cnt[2];
        // a call whose candidates passed the pool's end raises the next
        // call's pool (a reserved context keeps its size); run without the
        // fallback, it runs again at that size
        if (total > caps.pool && !ctx->reserved) learn(total);
        if (total > caps.pool && can_grow && !replayed) {
            M8Search(ctx, x, a, stats, per_query_stats, chk);
            return;
        }
    }
    if (per_query_stats) {
        if (batched) std::fill(per_query_stats, per_query_stats + nq * 5, 0ull);
        else
            for (uint64_t i = 0; i < nq; ++i) std::memcpy(per_query_stats + i * 5, hs.data() + i * 8, 5 * 8);
    }
    if (stats) {
        uint64_t s8[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // threshold pass (batched) or the whole search (plain)
        for (uint64_t i = 0; i < nq; ++i)
            for (int f = 0; f < 8; ++f) s8[f] += hs[i * 8 + f];
        if (!batched) {
            stats->candidates = s8[0];
            stats->survivors = s8[1];
            stats->bits_scored = s8[2];
            stats->bits_total = s8[3];
            stats->pruned = s8[4];
            stats->batches = s8[5];
            stats->refined = s8[6];
            stats->groups = s8[7];
        } else {
            // candidates: the vectors of every probed list; the threshold
            // pass's refined / survivors / bits plus the batched ones (the
            // plain search of overflowed queries is not counted again)
            uint64_t bits_row = 0;
            for (const auto& r : x->runs) bits_row += static_cast<uint64_t>(r.rate) * r.len;
            stats->form = E8Q_MODE_BATCHED;
            stats->candidates = wst[0];
            stats->batches = wst[2];
            stats->refined = s8[6] + bst[2];
            stats->survivors = s8[1] + bst[3];
            stats->bits_scored = s8[2] + bst[4];
            stats->bits_total = wst[0] * bits_row;
            stats->pruned = stats->candidates > stats->survivors ? stats->candidates - stats->survivors : 0;
            stats->groups = s8[7];
            stats->overflowed = novf;
            stats->pairs_scored = bst[0];
            stats->items = bst[5];
            stats->refine_bytes[0] = bst[6];
            stats->refine_bytes[1] = bst[7];
            for (int i = 0; i < nkl; ++i) {
                float ms = 0.0F;
                Check(hipEventElapsedTime(&ms, ctx->kev[2 * i], ctx->kev[2 * i + 1]), "event time");
                stats->refine_ms[kind[i]] += ms;
                ++stats->refine_launches[kind[i]];
            }
        }
        float ms = 0.0F;
        Check(hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]), "event time");
        stats->scan_ms = ms;
        Check(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]), "event time");
        stats->total_ms = ms;
        (void)replayed;
    }
}

}  // namespace

extern "C" {

uint64_t e8q_m8_search_workspace_bytes(const e8q_m8index* x, uint64_t nq, uint64_t k, uint64_t max_nprobe,
                                        int32_t mode) {
    if (!x) return 0;
    M8Bufs b{};
    e8q_ctx::Carve cv;
    const uint64_t maxp = std::min<uint64_t>(max_nprobe, x->view.nlist);
    M8Carve(cv, b, x->view, nq, k, maxp, true, true, true);
    if (mode == E8Q_MODE_BATCHED && M8BatchedFits(x, k) && nq > 0) {
        const uint32_t mp = static_cast<uint32_t>(maxp);
        M8CarveBatched(cv, b, x, nq, k, nq * maxp, false,
                       M8CapsFor(nq, M8PoolWant(x, nq, mp, 0), kM8DefaultBudget));
    }
    return cv.bytes();
}

e8q_status e8q_m8_create(e8q_ctx* ctx, const e8q_m8_desc* d, e8q_m8index** out) {
    return Guard([&] {
        if (!ctx || !d || !out) Throw(E8Q_EINVAL, "null argument");
        if (d->dim == 0 || d->nlist == 0) Throw(E8Q_EINVAL, "empty index shape");
        if (d->padded != (d->dim + 63) / 64 * 64) Throw(E8Q_EINVAL, "padded must be dim rounded up to 64");
        if (!d->seg_len || !d->seg_bits || !d->levels || !d->fixed_scale || !d->list_begin)
            Throw(E8Q_EINVAL, "null table");
        auto* x = new e8q_m8index();
        x->ctx = ctx;
        try {
            ctx->Bind();
            hipStream_t st = ctx->stream;
            M8View& v = x->view;
            v.n = d->n;
            v.dim = d->dim;
            v.padded = d->padded;
            v.nlist = d->nlist;
            v.drop_energy = d->drop_energy ? 1 : 0;
            // plan -> runs
            uint64_t off = 0, total_signs = 0;
            uint32_t nrun = 0, code = 0, blk = 0;
            for (uint32_t s = 0; s < d->nseg; ++s) {
                const uint64_t len = d->seg_len[s];
                const int bits = d->seg_bits[s];
                if (len == 0 || len % 64 != 0 || bits < 0 || bits > 13) Throw(E8Q_EINVAL, "bad plan segment");
                x->segs.push_back(M8Seg{static_cast<uint32_t>(off), static_cast<uint32_t>(len),
                                        static_cast<uint32_t>(bits), 0});
                if (bits > 0) {
                    M8Run r{};
                    r.rate = static_cast<uint32_t>(bits);
                    r.len = static_cast<uint32_t>(len);
                    r.nblk = static_cast<uint32_t>(len / 8);
                    r.dim_off = static_cast<uint32_t>(off);
                    r.blk0 = blk;
                    uint32_t lg = 0;
                    while ((uint64_t{1} << (lg + 1)) <= len) ++lg;
                    r.trunc = 1u << lg;
                    r.fac = 1.0F / std::sqrt(static_cast<float>(r.trunc));
                    x->runs.push_back(r);
                    blk += r.nblk;
                    total_signs += 4 * len / 8;
                    ++nrun;
                }
                off += len;
            }
            if (off != d->padded) Throw(E8Q_EINVAL, "plan does not cover the padded dimension");
            if (nrun > kM8MaxRuns) Throw(E8Q_EINVAL, "more than 32 coded runs");
            if (nrun > 0 && !d->signs) Throw(E8Q_EINVAL, "null rotator signs");
            v.nrun = nrun;
            v.nseg = d->nseg;
            v.plane_cols = blk;
            uint32_t route = static_cast<uint32_t>(d->padded);
            for (uint32_t s = d->nseg; s-- > 0;) {
                if (d->seg_bits[s] != 0) break;
                route -= static_cast<uint32_t>(d->seg_len[s]);
            }
            v.route_dim = route;
            code = nrun * 8;
            uint32_t ref = (nrun * 16 + 15) & ~15u, refr = nrun * 8;
            for (auto& r : x->runs) {
                r.code_off = code;
                code += r.rate * r.len / 8;
                r.ref_off = ref;
                x->ref_run_off.push_back(refr);
                if (r.rate >= 2) {
                    ref += (r.rate - 1) * r.nblk;
                    refr += (r.rate - 1) * r.nblk;
                }
            }
            v.stride = code;
            v.rstride = (ref + 15) & ~15u;
            v.rstride_fields = ref - nrun * 16;
            v.rrec_shift = (2 * nrun + 7) & ~7u;
            v.rrec_bytes = (nrun * 16 + v.rrec_shift + v.rstride_fields + 127) & ~127u;
            x->ref_stride = refr;
            v.rec = 32u * v.plane_cols + kM8BankBytes * nrun;
            // list / batch offsets
            x->list_begin.assign(d->list_begin, d->list_begin + d->nlist + 1);
            if (x->list_begin[0] != 0 || x->list_begin[d->nlist] != d->n) Throw(E8Q_EINVAL, "list offsets inconsistent");
            x->batch_begin.assign(d->nlist + 1, 0);
            for (uint64_t c = 0; c < d->nlist; ++c) {
                if (x->list_begin[c + 1] < x->list_begin[c]) Throw(E8Q_EINVAL, "list offsets not monotone");
                x->batch_begin[c + 1] = x->batch_begin[c] + (x->list_begin[c + 1] - x->list_begin[c] + 31) / 32;
            }
            x->nbatch = x->batch_begin[d->nlist];
            for (uint64_t c = 0; c < d->nlist; ++c) {
                const uint64_t nbt = (x->batch_begin[c + 1] - x->batch_begin[c] + kM8BatchBTile - 1) / kM8BatchBTile;
                x->sum_nbt += nbt;
                x->max_nbt = std::max(x->max_nbt, nbt);
            }
            {   // the largest lists first: what c probed lists can hold at most
                std::vector<uint64_t> len(d->nlist);
                for (uint64_t c = 0; c < d->nlist; ++c) len[c] = x->list_begin[c + 1] - x->list_begin[c];
                std::sort(len.begin(), len.end(), std::greater<uint64_t>());
                x->top_lists.assign(d->nlist + 1, 0);
                for (uint64_t c = 0; c < d->nlist; ++c) x->top_lists[c + 1] = x->top_lists[c] + len[c];
            }
            // tables: signs, float levels per run, encoder tables per distinct rate
            Upload(x->signs, d->signs, total_signs, st);
            size_t lev_total = 0;
            for (const auto& r : x->runs) lev_total += size_t{2} << r.rate;
            std::vector<float> levf;
            levf.reserve(lev_total);
            std::vector<size_t> lev_off;
            size_t lev_base[14] = {};
            for (int b = 1; b <= 13; ++b) lev_base[b] = (b == 1 ? 0 : lev_base[b - 1] + (size_t{2} << (b - 1)));
            int enc_of_rate[14];
            bool merged_of_rate[14] = {};
            for (int& e : enc_of_rate) e = -1;
            for (auto& r : x->runs) {
                const double* L = d->levels + lev_base[r.rate];
                const uint32_t q = 1u << r.rate;
                lev_off.push_back(levf.size());
                for (uint32_t i = 0; i < 2 * q; ++i) levf.push_back(static_cast<float>(L[i]));
                // class means pbar (ClassModelOf, me8q_codec.cpp:263-280), from the float levels
                for (int cl = 0; cl < 2; ++cl) {
                    double sum = 0.0;
                    for (int br = 0; br < 2; ++br)
                        for (uint32_t dd = cl * (q / 2); dd < (cl + 1) * (q / 2); ++dd)
                            sum += static_cast<double>(static_cast<float>(L[br * q + dd]));
                    const float pb = static_cast<float>(sum / static_cast<double>(q));
                    (cl == 0 ? r.pbar0 : r.pbar1) = pb;
                }
                if (enc_of_rate[r.rate] < 0) {
                    enc_of_rate[r.rate] = static_cast<int>(x->enc.size());
                    const EncTabHost t = MakeEncoderTables(L, static_cast<int>(r.rate));
                    x->enc.resize(x->enc.size() + 5);
                    DBuf* e = &x->enc[x->enc.size() - 5];
                    Upload(e[0], L, 2 * q, st);
                    Upload(e[1], t.mlev.data(), t.mlev.size(), st);
                    Upload(e[2], t.mdig.data(), t.mdig.size(), st);
                    Upload(e[3], t.cl.data(), t.cl.size(), st);
                    Upload(e[4], t.cd.data(), t.cd.size(), st);
                    merged_of_rate[r.rate] = t.merged;
                }
            }
            Upload(x->lev, levf.data(), levf.size(), st);
            uint64_t soff = 0;
            for (size_t i = 0; i < x->runs.size(); ++i) {
                M8Run& r = x->runs[i];
                r.signs = x->signs.as<uint8_t>() + soff;
                soff += 4 * static_cast<uint64_t>(r.len) / 8;
                r.lev = x->lev.as<float>() + lev_off[i];
                const DBuf* e = &x->enc[static_cast<size_t>(enc_of_rate[r.rate])];
                r.ev.bits = static_cast<int>(r.rate);
                r.ev.merged = merged_of_rate[r.rate] ? 1 : 0;
                r.ev.mlen = 2u << r.rate;
                r.ev.scale = d->fixed_scale[r.rate];
                r.ev.levels = e[0].as<double>();
                r.ev.mlev = e[1].as<double>();
                r.ev.mdig = e[2].as<uint16_t>();
                r.ev.cls_lev = e[3].as<double>();
                r.ev.cls_dig = e[4].as<uint16_t>();
            }
            Upload(x->d_runs, x->runs.data(), x->runs.size(), st);
            Upload(x->d_segs, x->segs.data(), x->segs.size(), st);
            Upload(x->d_list_begin, x->list_begin.data(), x->list_begin.size(), st);
            Upload(x->d_batch_begin, x->batch_begin.data(), x->batch_begin.size(), st);
            x->ids.alloc(v.n * sizeof(uint32_t));
            x->codes.alloc(v.n * v.stride + 16);
            x->centroids.alloc(v.nlist * v.padded * sizeof(float));
            x->cent_rot.alloc(v.nlist * v.padded * sizeof(float));
            x->cn2.alloc(v.nlist * sizeof(float));
            x->rc_n2.alloc(v.nlist * std::max<uint32_t>(nrun, 1) * sizeof(double));
            x->plane.alloc(x->nbatch * v.rec + 16);
            x->refine.alloc(v.n * v.rstride + 16);
            // the batched form's refine records, per batch slot, for
            // indexes it runs on (batch slots are named by 32-bit words)
            if (nrun >= 1 && nrun <= kM8BatchRuns && v.n > 0 && x->nbatch * 32 < (uint64_t{1} << 32)) {
                x->rrec.alloc(x->nbatch * 32 * v.rrec_bytes);
                x->bscale.alloc(x->nbatch * nrun * sizeof(float4));
                x->bcnt.alloc(x->nbatch);
                x->pids.alloc(x->nbatch * 32 * sizeof(uint32_t));
            }
            Check(hipMemsetAsync(x->codes.p, 0, x->codes.bytes, st), "clear");
            Check(hipMemsetAsync(x->refine.p, 0, x->refine.bytes, st), "clear");
            Check(hipStreamSynchronize(st), "m8 create");
            RefreshM8View(x);
        } catch (...) {
            delete x;
            throw;
        }
        *out = x;
    });
}

e8q_status e8q_m8_destroy(e8q_m8index* x) {
    return Guard([&] {
        if (!x) return;
        x->ctx->Bind();
        delete x;
    });
}

uint64_t e8q_m8_device_bytes(const e8q_m8index* x) { return x ? x->device_bytes() : 0; }
uint64_t e8q_m8_row_bytes(const e8q_m8index* x) { return x ? x->view.stride : 0; }

e8q_status e8q_m8_upload(e8q_ctx* ctx, e8q_m8index* x, const e8q_m8_arrays* h) {
    return Guard([&] {
        if (!ctx || !x || !h) Throw(E8Q_EINVAL, "null argument");
        const M8View& v = x->view;
        // a plan with no coded run has empty code rows
        if ((v.n && !h->ids) || !h->centroids || (v.n * v.stride && !h->codes)) Throw(E8Q_EINVAL, "null array");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        if (v.n) Check(hipMemcpyAsync(x->ids.p, h->ids, v.n * 4, hipMemcpyHostToDevice, st), "upload ids");
        Check(hipMemcpyAsync(x->centroids.p, h->centroids, v.nlist * v.padded * 4, hipMemcpyHostToDevice, st),
              "upload centroids");
        if (v.n * v.stride)
            Check(hipMemcpyAsync(x->codes.p, h->codes, v.n * v.stride, hipMemcpyHostToDevice, st), "upload codes");
        BuildDerived(ctx, x);
    });
}

e8q_status e8q_m8_download(e8q_ctx* ctx, const e8q_m8index* x, e8q_m8_arrays* h) {
    return Guard([&] {
        if (!ctx || !x || !h) Throw(E8Q_EINVAL, "null argument");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        const M8View& v = x->view;
        if (h->ids && v.n) Check(hipMemcpyAsync(h->ids, x->ids.p, v.n * 4, hipMemcpyDeviceToHost, st), "download ids");
        if (h->centroids)
            Check(hipMemcpyAsync(h->centroids, x->centroids.p, v.nlist * v.padded * 4, hipMemcpyDeviceToHost, st),
                  "download centroids");
        if (h->codes && v.n * v.stride)
            Check(hipMemcpyAsync(h->codes, x->codes.p, v.n * v.stride, hipMemcpyDeviceToHost, st), "download codes");
        Check(hipStreamSynchronize(st), "download");
    });
}

e8q_status e8q_m8_encode(e8q_ctx* ctx, e8q_m8index* x, const float* base, const float* centroids,
                         const uint32_t* cluster_ids, const uint32_t* rank, int on_device) {
    return Guard([&] {
        if (!ctx || !x || !base || !centroids || !cluster_ids || !rank) Throw(E8Q_EINVAL, "null argument");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        const M8View& v = x->view;
        if (!on_device)
            for (uint64_t i = 0; i < v.n; ++i)
                if (cluster_ids[i] >= v.nlist) Throw(E8Q_EINVAL, "cluster id out of range");
        Check(hipMemcpyAsync(x->centroids.p, centroids, v.nlist * v.padded * 4, hipMemcpyHostToDevice, st),
              "upload centroids");
        Check(LaunchM8Centroids(st, v), "centroid kernel");
        float* db = nullptr;
        uint32_t *dc = nullptr, *dr = nullptr;
        e8q_ctx::Carve cv;
        if (!on_device) {
            cv.add(&db, v.n * v.dim);
            cv.add(&dc, v.n);
        }
        cv.add(&dr, v.n);
        ctx->Take(cv);
        if (!on_device) {
            Check(hipMemcpyAsync(db, base, v.n * v.dim * 4, hipMemcpyHostToDevice, st), "upload base");
            Check(hipMemcpyAsync(dc, cluster_ids, v.n * 4, hipMemcpyHostToDevice, st), "upload cluster ids");
        }
        Check(hipMemcpyAsync(dr, rank, v.n * 4, hipMemcpyHostToDevice, st), "upload ranks");
        Check(LaunchM8Encode(st, v, on_device ? base : db, on_device ? cluster_ids : dc, dr, v.n), "encode kernel");
        Check(hipStreamSynchronize(st), "encode");
        BuildDerived(ctx, x);
    });
}

e8q_status e8q_m8_search(e8q_ctx* ctx, const e8q_m8index* x, const e8q_m8_search_args* a, e8q_m8_stats* stats,
                         uint64_t* per_query_stats) {
    return Guard([&] { M8Search(ctx, x, a, stats, per_query_stats, nullptr); });
}

e8q_status e8q_ctx_set_m8_grids(e8q_ctx* ctx, uint32_t mfma_grid, uint32_t refine_grid) {
    return Guard([&] {
        if (!ctx) Throw(E8Q_EINVAL, "ctx is null");
        if (mfma_grid > 65536 || refine_grid > 65536) Throw(E8Q_EINVAL, "grids must be <= 65536");
        ctx->m8_fix_mfma = mfma_grid;
        ctx->m8_fix_refine = refine_grid;
    });
}

e8q_status e8q_ctx_share_tuning(e8q_ctx* ctx, const e8q_ctx* from) {
    return Guard([&] {
        if (!ctx || !from) Throw(E8Q_EINVAL, "null argument");
        if (ctx->device != from->device) Throw(E8Q_EINVAL, "contexts on different devices");
        ctx->m8_table = from->m8_table;
    });
}

e8q_status e8q_ctx_get_m8_info(const e8q_ctx* ctx, e8q_m8_ctx_info* out) {
    return Guard([&] {
        if (!ctx || !out) Throw(E8Q_EINVAL, "null argument");
        *out = e8q_m8_ctx_info{ctx->m8_last_mfma, ctx->m8_last_refine, ctx->m8_last_budget, ctx->m8_last_set_a,
                               ctx->m8_tunings, ctx->m8_pool_per_q};
    });
}

e8q_status e8q_m8_debug_batch_checksum(e8q_ctx* ctx, const e8q_m8index* x, const e8q_m8_search_args* a,
                                       uint32_t* chk) {
    return Guard([&] {
        if (!chk) Throw(E8Q_EINVAL, "null argument");
        if (!a || a->mode != E8Q_MODE_BATCHED) Throw(E8Q_EINVAL, "the checksum needs the batched form");
        if (!x || !M8BatchedFits(x, a->k)) Throw(E8Q_EINVAL, "the batched form does not run for this index and k");
        M8Search(ctx, x, a, nullptr, nullptr, chk);
    });
}

e8q_status e8q_m8_debug_fast(e8q_ctx* ctx, const e8q_m8index* x, uint64_t c, uint64_t* sizes3, uint8_t* plane,
                             uint8_t* refine, float* gain, float* vpc, double* rc_n2) {
    return Guard([&] {
        if (!ctx || !x || !sizes3) Throw(E8Q_EINVAL, "null argument");
        const M8View& v = x->view;
        if (c >= v.nlist) Throw(E8Q_EINVAL, "list out of range");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        const uint64_t R = v.nrun, PC = v.plane_cols;
        const uint64_t cnt = x->list_begin[c + 1] - x->list_begin[c];
        const uint64_t nb = x->batch_begin[c + 1] - x->batch_begin[c];
        const uint64_t rrec = PC * 32 + R * kM8BankBytes;   // reference record bytes
        sizes3[0] = nb * rrec;
        sizes3[1] = cnt * x->ref_stride;
        sizes3[2] = R * cnt;
        std::vector<uint8_t> dp(nb * v.rec), dr(cnt * v.rstride);
        std::vector<double> rn(R);
        if (nb) Check(hipMemcpyAsync(dp.data(), x->plane.as<uint8_t>() + x->batch_begin[c] * v.rec, dp.size(),
                                     hipMemcpyDeviceToHost, st), "download plane");
        if (cnt) Check(hipMemcpyAsync(dr.data(), x->refine.as<uint8_t>() + x->list_begin[c] * v.rstride, dr.size(),
                                      hipMemcpyDeviceToHost, st), "download refine");
        if (R) Check(hipMemcpyAsync(rn.data(), x->rc_n2.as<double>() + c * R, R * 8, hipMemcpyDeviceToHost, st),
                     "download rc_n2");
        Check(hipStreamSynchronize(st), "debug fast");
        if (plane)
            for (uint64_t b = 0; b < nb; ++b) {
                const uint8_t* s = dp.data() + b * v.rec;
                uint8_t* o = plane + b * rrec;
                for (uint64_t col = 0; col < PC; ++col)
                    for (uint64_t l = 0; l < 32; ++l) o[col * 32 + l] = s[l * PC + col];
                std::memcpy(o + PC * 32, s + 32 * PC, R * kM8BankBytes);
            }
        for (uint64_t i = 0; i < cnt; ++i) {
            const uint8_t* s = dr.data() + i * v.rstride;
            if (refine) {
                uint8_t* o = refine + i * x->ref_stride;
                for (uint64_t r = 0; r < R; ++r) std::memcpy(o + r * 8, s + r * 16, 8);
                for (uint64_t r = 0; r < R; ++r) {
                    const M8Run& ru = x->runs[r];
                    if (ru.rate >= 2) std::memcpy(o + x->ref_run_off[r], s + ru.ref_off, (ru.rate - 1) * ru.nblk);
                }
            }
            for (uint64_t r = 0; r < R; ++r) {
                float f4[4];
                std::memcpy(f4, s + r * 16, 16);
                if (gain) gain[r * cnt + i] = f4[2];
                if (vpc) vpc[r * cnt + i] = f4[3];
            }
        }
        if (rc_n2) std::memcpy(rc_n2, rn.data(), R * 8);
    });
}

e8q_status e8q_m8_debug_centroids(e8q_ctx* ctx, const e8q_m8index* x, float* cent_rot, float* cn2) {
    return Guard([&] {
        if (!ctx || !x) Throw(E8Q_EINVAL, "null argument");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        const M8View& v = x->view;
        if (cent_rot) Check(hipMemcpyAsync(cent_rot, x->cent_rot.p, v.nlist * v.padded * 4, hipMemcpyDeviceToHost, st), "d");
        if (cn2) Check(hipMemcpyAsync(cn2, x->cn2.p, v.nlist * 4, hipMemcpyDeviceToHost, st), "d");
        Check(hipStreamSynchronize(st), "debug centroids");
    });
}

e8q_status e8q_m8_debug_query(e8q_ctx* ctx, const e8q_m8index* x, const float* queries, uint64_t nq, uint64_t nprobe,
                              float* rot_q, uint8_t* lut, float* lut_md, uint32_t* probe, float* pdist) {
    return Guard([&] {
        if (!ctx || !x || !queries) Throw(E8Q_EINVAL, "null argument");
        if (nprobe == 0) Throw(E8Q_EINVAL, "nprobe must be positive");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        const M8View& v = x->view;
        const uint64_t P = v.padded, PC = v.plane_cols, R = std::max<uint64_t>(v.nrun, 1);
        const uint32_t maxp = static_cast<uint32_t>(std::min<uint64_t>(nprobe, v.nlist));
        float *dq = nullptr, *rotq = nullptr, *qc = nullptr, *dist = nullptr, *pd = nullptr;
        double* rc = nullptr;
        uint8_t* dl = nullptr;
        uint32_t* pr = nullptr;
        e8q_ctx::Carve cv;
        cv.add(&dq, nq * v.dim);
        cv.add(&rotq, nq * P);
        cv.add(&dl, nq * PC * 32 + 16);
        cv.add(&qc, nq * 4);
        cv.add(&rc, nq * R * 4);
        cv.add(&dist, nq * v.nlist);
        cv.add(&pr, nq * maxp);
        cv.add(&pd, nq * maxp);
        ctx->Take(cv);
        Check(hipMemcpyAsync(dq, queries, nq * v.dim * 4, hipMemcpyHostToDevice, st), "upload");
        Check(LaunchM8QueryPrep(st, v, dq, nq, rotq, dl, qc, rc), "query prep");
        RouteArgs ra{};
        ra.rotq = rotq;
        ra.qstride = static_cast<uint32_t>(P);
        ra.qn2 = qc + 2;
        ra.qn2_stride = 4;
        ra.cent = v.cent_rot;
        ra.cn2 = v.cn2;
        ra.nlist = v.nlist;
        ra.dim = v.route_dim;
        ra.cstride = static_cast<uint32_t>(P);
        ra.nq = nq;
        ra.dist = dist;
        Check(LaunchRouteSelect(st, ra, nullptr, maxp, maxp, pr, pd), "route");
        std::vector<uint8_t> hl(nq * PC * 32);
        std::vector<float> hq(nq * 4);
        if (rot_q) Check(hipMemcpyAsync(rot_q, rotq, nq * P * 4, hipMemcpyDeviceToHost, st), "d");
        if (!hl.empty()) Check(hipMemcpyAsync(hl.data(), dl, hl.size(), hipMemcpyDeviceToHost, st), "d");
        Check(hipMemcpyAsync(hq.data(), qc, nq * 16, hipMemcpyDeviceToHost, st), "d");
        if (probe) Check(hipMemcpyAsync(probe, pr, nq * maxp * 4, hipMemcpyDeviceToHost, st), "d");
        if (pdist) Check(hipMemcpyAsync(pdist, pd, nq * maxp * 4, hipMemcpyDeviceToHost, st), "d");
        Check(hipStreamSynchronize(st), "debug query");
        for (uint64_t i = 0; i < nq; ++i) {
            if (lut_md) {
                lut_md[i * 2] = hq[i * 4];
                lut_md[i * 2 + 1] = hq[i * 4 + 1];
            }
            if (lut)
                for (uint64_t col = 0; col < PC; ++col) {
                    const uint8_t* s = hl.data() + (i * PC + col) * 32;
                    uint8_t* o = lut + (i * PC + col) * 64;
                    for (int e = 0; e < 16; ++e) {
                        o[e] = o[16 + e] = s[e];
                        o[32 + e] = o[48 + e] = s[16 + e];
                    }
                }
        }
    });
}

// ---- diagnostics (me8q_measure.cpp:41-67, me8q_search.cpp:136-266) --------

namespace {

uint64_t M8DiagChunk(uint64_t nq, uint64_t per_query) {
    const uint64_t lim = uint64_t{1} << 28;   // candidates held at once
    const uint64_t c = std::min<uint64_t>({nq, lim / std::max<uint64_t>(per_query, 1), 65535});   // grid y
    return std::max<uint64_t>(1, c);
}

// Queries on the device and their per-run rotations (the query prep's other
// outputs land in scratch).
struct M8DiagQueries {
    float *dq = nullptr, *rotq = nullptr, *qc = nullptr;
    uint8_t* lut = nullptr;
    double* rc = nullptr;
    void Carve(e8q_ctx::Carve& cv, const M8View& v, uint64_t nq) {
        cv.add(&dq, nq * v.dim);
        cv.add(&rotq, nq * v.padded);
        cv.add(&lut, nq * v.plane_cols * 32 + 16);
        cv.add(&qc, nq * 4);
        cv.add(&rc, nq * std::max<uint64_t>(v.nrun, 1) * 4);
    }
    void Prepare(hipStream_t st, const M8View& v, const float* queries, uint64_t nq) {
        Check(hipMemcpyAsync(dq, queries, nq * v.dim * 4, hipMemcpyHostToDevice, st), "upload queries");
        Check(LaunchM8QueryPrep(st, v, dq, nq, rotq, lut, qc, rc), "query prep");
    }
};

}  // namespace

e8q_status e8q_m8_estimate(e8q_ctx* ctx, const e8q_m8index* x, const float* queries, uint64_t nq,
                           const uint32_t* qidx, const uint32_t* ids, uint64_t npair, float* out) {
    return Guard([&] {
        if (!ctx || !x) Throw(E8Q_EINVAL, "null argument");
        if (npair && (!queries || !qidx || !ids || !out)) Throw(E8Q_EINVAL, "null buffer");
        const M8View& v = x->view;
        for (uint64_t i = 0; i < npair; ++i)
            if (qidx[i] >= nq) Throw(E8Q_EINVAL, "Me8qIndex::Estimate: query index out of range");
        if (npair == 0) return;
        if (v.n == 0) Throw(E8Q_EINVAL, "Me8qIndex::Estimate: unknown id");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        M8DiagQueries dqs;
        float* dout = nullptr;
        uint32_t *dqi = nullptr, *dids = nullptr, *inv = nullptr, *bad = nullptr;
        e8q_ctx::Carve cv;
        dqs.Carve(cv, v, nq);
        cv.add(&dqi, npair);
        cv.add(&dids, npair);
        cv.add(&inv, v.n);
        cv.add(&dout, npair);
        cv.add(&bad, 1);
        ctx->Take(cv);
        dqs.Prepare(st, v, queries, nq);
        Check(hipMemcpyAsync(dqi, qidx, npair * 4, hipMemcpyHostToDevice, st), "upload pairs");
        Check(hipMemcpyAsync(dids, ids, npair * 4, hipMemcpyHostToDevice, st), "upload pairs");
        Check(LaunchFill32(st, inv, v.n, 0), "clear");
        Check(LaunchFill32(st, bad, 1, 0), "clear");
        Check(LaunchDiagInverseIds(st, v.ids, v.n, inv), "id map");
        Check(LaunchM8Estimate(st, v, dqs.dq, dqs.rotq, dqi, dids, npair, inv, dout, bad), "estimate");
        uint32_t hbad = 0;
        Check(hipMemcpyAsync(out, dout, npair * 4, hipMemcpyDeviceToHost, st), "download");
        Check(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st), "download");
        Check(hipStreamSynchronize(st), "estimate");
        if (hbad) Throw(E8Q_EINVAL, "Me8qIndex::Estimate: unknown id");
    });
}

e8q_status e8q_m8_exhaustive_topk(e8q_ctx* ctx, const e8q_m8index* x, const float* queries, uint64_t nq, uint64_t k,
                                  uint32_t* out_ids, float* out_dists) {
    return Guard([&] {
        if (!ctx || !x) Throw(E8Q_EINVAL, "null argument");
        if (k == 0 || k > kDiagMaxK) Throw(E8Q_EINVAL, "k must be in [1, 1024]");
        if (nq && (!queries || !out_ids || !out_dists)) Throw(E8Q_EINVAL, "null buffer");
        if (nq == 0) return;
        const M8View& v = x->view;
        ctx->Bind();
        hipStream_t st = ctx->stream;
        const uint64_t chunk = M8DiagChunk(nq, v.n);
        M8DiagQueries dqs;
        double* d2c = nullptr;
        float *dist = nullptr, *od = nullptr;
        uint32_t* oid = nullptr;
        e8q_ctx::Carve cv;
        dqs.Carve(cv, v, nq);
        cv.add(&d2c, nq * v.nlist);
        cv.add(&dist, std::max<uint64_t>(chunk * v.n, 1));
        cv.add(&oid, nq * k);
        cv.add(&od, nq * k);
        ctx->Take(cv);
        dqs.Prepare(st, v, queries, nq);
        Check(LaunchM8ListD2(st, v, dqs.dq, nq, d2c), "list distances");
        for (uint64_t q0 = 0; q0 < nq; q0 += chunk) {
            const uint64_t nqc = std::min(chunk, nq - q0);
            Check(LaunchM8Exhaustive(st, v, dqs.rotq, d2c, q0, nqc, dist), "exhaustive distances");
            Check(LaunchDiagSelect(st, true, dist, v.ids, 0, nullptr, v.n, v.n, q0, nqc, static_cast<uint32_t>(k), oid,
                                   od),
                  "top-k");
        }
        Check(hipMemcpyAsync(out_ids, oid, nq * k * 4, hipMemcpyDeviceToHost, st), "download");
        Check(hipMemcpyAsync(out_dists, od, nq * k * 4, hipMemcpyDeviceToHost, st), "download");
        Check(hipStreamSynchronize(st), "exhaustive top-k");
    });
}

e8q_status e8q_m8_search_ivf(e8q_ctx* ctx, const e8q_m8index* x, const float* queries, uint64_t nq, uint64_t k,
                             uint64_t nprobe, uint32_t* out_ids, float* out_dists, uint64_t* stats6) {
    return Guard([&] {
        if (!ctx || !x) Throw(E8Q_EINVAL, "null argument");
        if (k > kDiagMaxK) Throw(E8Q_EINVAL, "k must be <= 1024");
        const M8View& v = x->view;
        if (nq == 0 || k == 0 || nprobe == 0 || v.n == 0) return;   // the reference returns untouched
        if (!queries || !out_ids || !out_dists) Throw(E8Q_EINVAL, "null buffer");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        const uint32_t take = static_cast<uint32_t>(std::min<uint64_t>(nprobe, v.nlist));
        const uint64_t cap = std::max<uint64_t>(x->top_lists[take], 1);   // candidates of one query at most
        const uint64_t chunk = M8DiagChunk(nq, cap);
        M8DiagQueries dqs;
        float *dist = nullptr, *pd = nullptr, *cd = nullptr, *od = nullptr;
        uint32_t *probe = nullptr, *cid = nullptr, *oid = nullptr;
        uint64_t *off = nullptr, *count = nullptr;
        e8q_ctx::Carve cv;
        dqs.Carve(cv, v, nq);
        cv.add(&dist, nq * v.nlist);
        cv.add(&probe, nq * take);
        cv.add(&pd, nq * take);
        cv.add(&off, nq * (take + 1));
        cv.add(&count, nq);
        cv.add(&cd, chunk * cap);
        cv.add(&cid, chunk * cap);
        cv.add(&oid, nq * k);
        cv.add(&od, nq * k);
        ctx->Take(cv);
        dqs.Prepare(st, v, queries, nq);
        RouteArgs ra{};
        ra.rotq = dqs.rotq;
        ra.qstride = static_cast<uint32_t>(v.padded);
        ra.qn2 = dqs.qc + 2;
        ra.qn2_stride = 4;
        ra.cent = v.cent_rot;
        ra.cn2 = v.cn2;
        ra.nlist = v.nlist;
        ra.dim = v.route_dim;
        ra.cstride = static_cast<uint32_t>(v.padded);
        ra.nq = nq;
        ra.dist = dist;
        Check(LaunchRouteSelect(st, ra, nullptr, take, take, probe, pd), "route");
        Check(LaunchM8IvfOffsets(st, v, probe, take, take, nq, off, count), "candidate offsets");
        for (uint64_t q0 = 0; q0 < nq; q0 += chunk) {
            const uint64_t nqc = std::min(chunk, nq - q0);
            Check(LaunchM8Ivf(st, v, dqs.rotq, probe, pd, take, take, off, q0, nqc, cap, cd, cid), "distances");
            Check(LaunchDiagSelect(st, true, cd, cid, cap, count, 0, cap, q0, nqc, static_cast<uint32_t>(k), oid, od),
                  "top-k");
        }
        std::vector<uint64_t> hc(nq);
        Check(hipMemcpyAsync(out_ids, oid, nq * k * 4, hipMemcpyDeviceToHost, st), "download");
        Check(hipMemcpyAsync(out_dists, od, nq * k * 4, hipMemcpyDeviceToHost, st), "download");
        Check(hipMemcpyAsync(hc.data(), count, nq * 8, hipMemcpyDeviceToHost, st), "download");
        Check(hipStreamSynchronize(st), "search ivf");
        if (stats6) {
            uint64_t bits_row = 0;
            for (const auto& r : x->runs) bits_row += static_cast<uint64_t>(r.rate) * r.len;
            for (uint64_t i = 0; i < nq; ++i) {
                uint64_t* o = stats6 + i * 6;
                o[0] = hc[i];
                o[1] = take;
                o[2] = hc[i] * bits_row;
                o[3] = hc[i] * bits_row;
                o[4] = hc[i];
                o[5] = 0;
            }
        }
    });
}

}  // extern "C"
