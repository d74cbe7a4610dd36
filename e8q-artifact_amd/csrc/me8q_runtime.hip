// me8q_runtime.hip -- the ME8Q half of libe8q_gpu.so's C ABI
// (include/me8q_gpu.h): device index lifetime, upload / download, the
// device build, search, and the test diagnostics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/me8q_gpu.h"
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
    uint64_t nbatch = 0;
    // refine-row layout of the reference (for the diagnostics)
    std::vector<uint32_t> ref_run_off;
    uint32_t ref_stride = 0;
    DBuf d_runs, d_segs, d_list_begin, d_batch_begin, ids, codes, centroids, cent_rot, cn2, rc_n2, plane, refine,
        signs, lev, rrec, bscale;
    std::vector<DBuf> enc;   // encoder tables, 5 per distinct rate
    uint64_t device_bytes() const {
        uint64_t b = d_runs.bytes + d_segs.bytes + d_list_begin.bytes + d_batch_begin.bytes + ids.bytes + codes.bytes +
                     centroids.bytes + cent_rot.bytes + cn2.bytes + rc_n2.bytes + plane.bytes + refine.bytes +
                     signs.bytes + lev.bytes + rrec.bytes + bscale.bytes;
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
    // batched form
    uint32_t *fid, *ccount, *cfill, *coff, *ioff, *counters, *ncand, *nfill, *ncont, *toff, *toff2, *chk, *hist;
    uint32_t *abin, *na, *fill_a, *fill_b;
    uint64_t* coff_b;
    float *fd, *tau;
    uint64_t *cover, *pairs, *poff;
    uint8_t* lut8;
    BatchItem* items;
    unsigned long long *wstats, *bstats, *rcount;
    uint2* itab;
    BatchedParams* wparams;
};

constexpr uint64_t kM8PoolMax = uint64_t{1} << 31;   // candidate pool entries (more: the plain form)
// Records of the count pass (8 B each): the first call of a context reserves
// up to kM8RecFirst of them, a call that needed more raises the next call's
// reserve to 1.25 x its count, up to kM8RecMax (2^31 = 16 GiB); a call whose
// records do not fit places its candidates with the write pass.  The
// environment variable E8Q_M8_RECORD_CAP lowers the limit (0: always the
// write pass).
constexpr uint64_t kM8RecFirst = uint64_t{1} << 28;
constexpr uint64_t kM8RecMax = uint64_t{1} << 31;
constexpr uint32_t kM8DefaultBudget = 16;

// Whether the batched form runs for this index and k.
bool M8BatchedFits(const e8q_m8index* x, uint64_t k) {
    const M8View& v = x->view;
    if (v.nrun < 1 || v.nrun > kM8BatchRuns || k > kM8SelMaxK || (v.plane_cols & 7u) != 0) return false;
    uint32_t cols = 0;
    for (const auto& r : x->runs) cols += r.nblk;
    return cols == v.plane_cols;
}

// Work items of the batched form: an upper bound of their count (the list
// count of each pair tile times the batch tiles of its list).
uint64_t M8ItemsCap(const e8q_m8index* x, uint64_t nq, uint64_t pairs) {
    uint64_t sum_nbt = 0, max_nbt = 0;
    for (uint64_t c = 0; c < x->view.nlist; ++c) {
        const uint64_t nb = (x->list_begin[c + 1] - x->list_begin[c] + 31) / 32;
        const uint64_t nbt = (nb + kM8BatchBTile - 1) / kM8BatchBTile;
        sum_nbt += nbt;
        max_nbt = std::max(max_nbt, nbt);
    }
    const uint64_t qt = kM8BatchQTile;
    return std::min((nq + qt - 1) / qt * sum_nbt, (pairs + qt - 1) / qt * max_nbt + sum_nbt);
}

void M8CarveBatched(e8q_ctx::Carve& cv, M8Bufs& b, const e8q_m8index* x, uint64_t nq, uint64_t k, uint64_t pairs,
                    bool chk) {
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
    if (chk) cv.add(&b.chk, nq);
}

// Records the count pass of this call may write.
uint64_t M8RecordCap(e8q_ctx* ctx, uint64_t items_cap) {
    uint64_t lim = kM8RecMax;
    if (const char* e = std::getenv("E8Q_M8_RECORD_CAP")) lim = std::min<uint64_t>(lim, std::strtoull(e, nullptr, 10));
    const uint64_t bound = items_cap * kM8BatchQTile * kM8BatchBTile * 32;   // every pair of every item
    const uint64_t want = std::max(kM8RecFirst, ctx->rec_want);
    return std::min({lim, bound, want});
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
}

// One search: plain (the reference's per-query schedule) or batched.  chk
// (host [nq], batched only): the coarse-sum checksum of the diagnostic.
void M8Search(e8q_ctx* ctx, const e8q_m8index* x, const e8q_m8_search_args* a, e8q_m8_stats* stats,
              uint64_t* per_query_stats, uint32_t* chk) {
    if (!ctx || !x || !a) Throw(E8Q_EINVAL, "null argument");
    if (a->k == 0 || a->k > kM8MaxK) Throw(E8Q_EINVAL, "k must be in [1, 1024]");
    if (!a->queries || !a->out_ids || !a->out_dists) Throw(E8Q_EINVAL, "null buffer");
    if (a->mode != E8Q_MODE_PLAIN && a->mode != E8Q_MODE_BATCHED) Throw(E8Q_EINVAL, "mode must be plain or batched");
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
    M8Bufs b{};
    e8q_ctx::Carve cv;
    M8Carve(cv, b, v, nq, a->k, maxp, !a->queries_on_device, !a->outputs_on_device, a->nprobe_per_query != nullptr);
    if (batched) M8CarveBatched(cv, b, x, nq, a->k, pairs, chk != nullptr);
    ctx->Take(cv);
    float *dq = b.dq, *rotq = b.rotq, *qc = b.qc, *dist = b.dist, *pd = b.pd, *od = b.od;
    double *rc = b.rc, *sqd2 = b.sqd2;
    uint8_t* lut = b.lut;
    uint32_t *probe = b.probe, *npq = b.npq, *oid = b.oid;
    uint64_t* qst = b.qst;
    const float* q = a->queries;
    if (!a->queries_on_device) {
        Check(hipMemcpyAsync(dq, a->queries, nq * v.dim * 4, hipMemcpyHostToDevice, st), "upload queries");
        q = dq;
    }
    if (npq) Check(hipMemcpyAsync(npq, a->nprobe_per_query, nq * 4, hipMemcpyHostToDevice, st), "upload nprobe");
    const uint32_t nprobe_all = static_cast<uint32_t>(std::min<uint64_t>(a->nprobe, v.nlist));
    Check(hipEventRecord(ctx->ev[0], st), "event");
    Check(LaunchM8QueryPrep(st, v, q, nq, rotq, lut, qc, rc), "query prep");
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
    Check(LaunchRouteSelect(st, ra, npq, nprobe_all, maxp, probe, pd), "route");
    Check(LaunchM8RunDist(st, v, rotq, rc, probe, npq, nprobe_all, maxp, nq, sqd2), "run distances");
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
    for (const auto& r : x->runs)
        if (r.rate <= kM8LdsLevRate) sa.lev_floats += 2u << r.rate;
    std::vector<unsigned long long> bst(8, 0), wst(4, 0);
    uint64_t novf = 0;
    // The refine's launches, each between two of the context's kernel events:
    // kind 0 = first-run phase, 1 = cascades.
    int nkl = 0, kind[e8q_ctx::kKernelEvents / 2] = {};
    auto refine_launch = [&](const M8BatchArgs& r, bool first, const char* what) {
        const bool timed = stats && nkl < e8q_ctx::kKernelEvents / 2;
        if (timed) Check(hipEventRecord(ctx->kev[2 * nkl], st), "event");
        Check(LaunchM8Refine(st, r, 8 * static_cast<uint32_t>(std::max(ctx->cus, 1)), first), what);
        if (timed) {
            Check(hipEventRecord(ctx->kev[2 * nkl + 1], st), "event");
            kind[nkl++] = first ? 0 : 1;
        }
    };
    if (!batched) {
        Check(hipEventRecord(ctx->ev[1], st), "event");
        Check(LaunchM8Scan(st, sa), "scan");
        Check(hipEventRecord(ctx->ev[2], st), "event");
        Check(hipEventRecord(ctx->ev[3], st), "event");
    } else {
        const uint32_t budget = a->batched_budget ? a->batched_budget : kM8DefaultBudget;
        // the (query, list) work list: every probe, grouped by list
        Check(LaunchFill32(st, b.ccount, v.nlist, 0), "clear");
        Check(LaunchFill32(st, b.ncand, nq, 0), "clear");
        Check(LaunchFill32(st, reinterpret_cast<uint32_t*>(b.wstats), 8, 0), "clear");
        Check(LaunchFill32(st, reinterpret_cast<uint32_t*>(b.bstats), 16, 0), "clear");
        if (chk) Check(LaunchFill32(st, b.chk, nq, 0), "clear");
        IndexView lv{};   // the work-list kernels read the list count and offsets only
        lv.nlist = v.nlist;
        lv.cluster_begin = v.list_begin;
        BatchedArgs w{};
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
        Check(LaunchBatchedWorkList(st, lv, w, b.wparams, true), "work list");
        // 3. the MFMA contraction and the coarse bound: a count pass, the pool
        // offsets, a write pass
        Check(LaunchFill32(st, b.nfill, nq, 0), "clear");
        Check(LaunchFill32(st, b.hist, nq * kM8TauBins, 0), "clear");
        Check(LaunchFill32(st, b.fill_a, nq, 0), "clear");
        Check(LaunchFill32(st, b.fill_b, nq, 0), "clear");
        M8BatchArgs ba{};
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
        ba.set_a = a->batched_set_a ? a->batched_set_a : kM8SetA;
        const uint32_t cus = static_cast<uint32_t>(std::max(ctx->cus, 1));
        {   // the threshold pass: exact distances of each query's first `budget`
            // batches (whole cascade, no death test, tau = +inf) and their top-k
            M8BatchArgs ta = ba;
            const uint64_t cap = static_cast<uint64_t>(budget) * 32;
            const uint64_t pool_b = (nq * cap * 8 + 255) & ~uint64_t{255};
            const uint64_t need = pool_b + (nq * ((cap + kM8RefineChunk - 1) / kM8RefineChunk) + 1) * 4 + 256;
            if (ctx->tpool.bytes < need) ctx->tpool.alloc(need);
            ta.pool = ctx->tpool.as<uint64_t>();
            ta.tsurv = reinterpret_cast<uint32_t*>(ctx->tpool.as<char>() + pool_b);
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
        // records of the count pass, when the reserve holds any
        const uint64_t rcap = M8RecordCap(ctx, w.items_cap);
        if (rcap > 0 && ctx->recs.bytes < rcap * sizeof(uint2)) ctx->recs.alloc(rcap * sizeof(uint2));
        ba.rec = rcap > 0 ? ctx->recs.as<uint2>() : nullptr;
        ba.rec_cap = rcap;
        ba.rcount = b.rcount;
        ba.itab = b.itab;
        Check(LaunchFill32(st, reinterpret_cast<uint32_t*>(b.rcount), 2, 0), "clear");
        Check(LaunchM8LutInt8(st, lut, b.lut8, nq * v.plane_cols * 32), "int8 LUT");
        Check(hipEventRecord(ctx->ev[1], st), "event");
        Check(LaunchM8MfmaScan(st, ba, 2 * cus, false), "mfma count pass");
        Check(hipEventRecord(ctx->ev[2], st), "event");
        Check(LaunchM8Split(st, ba), "candidate split");
        Check(LaunchM8RefineTasks(st, ba, b.ncand, true), "pool offsets");
        uint64_t total = 0, nrec = 0;
        uint32_t ntask = 0, cnt[4] = {0, 0, 0, 0};
        Check(hipMemcpyAsync(&total, b.poff + nq, 8, hipMemcpyDeviceToHost, st), "download pool size");
        Check(hipMemcpyAsync(&ntask, b.toff + nq, 4, hipMemcpyDeviceToHost, st), "download task count");
        Check(hipMemcpyAsync(cnt, b.counters, sizeof(cnt), hipMemcpyDeviceToHost, st), "download counters");
        Check(hipMemcpyAsync(&nrec, b.rcount, 8, hipMemcpyDeviceToHost, st), "download record count");
        Check(hipStreamSynchronize(st), "batched count pass");
        if (cnt[0] > w.items_cap) Throw(E8Q_ERUNTIME, "batched search: work items exceed their bound");
        const bool placed = ba.rec != nullptr && nrec <= rcap;
        if (ba.rec != nullptr && nrec != total && nrec <= rcap)
            Throw(E8Q_ERUNTIME, "batched search: records and candidate counts differ");
        if (!placed && nrec > rcap && nrec <= kM8RecMax) ctx->rec_want = nrec + nrec / 4;
        if (total > kM8PoolMax) {   // too many candidates for one pool: the plain form
            novf = nq;
            Check(hipEventRecord(ctx->ev[1], st), "event");
            Check(LaunchM8Scan(st, sa), "plain search");
            Check(hipEventRecord(ctx->ev[2], st), "event");
        } else {
            // pool: the candidates (A then B per query), their lower bounds,
            // the candidates alive after their first run, then the per-task
            // survivor counts of the four refine launches
            const uint64_t pool_b = (total * 8 + 255) & ~uint64_t{255};
            const uint64_t dlo_b = (total * 4 + 255) & ~uint64_t{255};
            const uint64_t task_b = (static_cast<uint64_t>(ntask) * 4 + 255) & ~uint64_t{255};
            const uint64_t need = 2 * pool_b + dlo_b + 4 * task_b + 256;
            if (ctx->pool.bytes < need) ctx->pool.alloc(need + need / 4);
            char* pb = ctx->pool.as<char>();
            ba.pool = reinterpret_cast<uint64_t*>(pb);
            ba.cont = reinterpret_cast<uint64_t*>(pb + pool_b);
            ba.pdlo = reinterpret_cast<float*>(pb + 2 * pool_b);
            ba.ncont = b.ncont;
            uint32_t* ts = reinterpret_cast<uint32_t*>(pb + 2 * pool_b + dlo_b);
            const uint64_t tsn = task_b / 4;
            if (placed) {
                Check(LaunchM8Place(st, ba, 8 * cus, false), "placement of set A");
            } else {
                Check(LaunchFill32(st, b.counters + 1, 1, 0), "clear");
                Check(LaunchM8MfmaScan(st, ba, 2 * cus, true), "mfma write pass");
            }
            // 4. exact cascades (every candidate's first run, then the whole
            // cascade of those still alive) and a top-k merge, first of each
            // query's set A (its exact top-k gives tau1), then of the rest
            // whose lower bound is below tau1
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
            M8BatchArgs ra = ba;   // set A
            ra.nfill = b.na;
            refine(ra, ts, ts + tsn, b.fid, b.fd);
            if (placed) Check(LaunchM8Place(st, ba, 8 * cus, true), "placement of the rest");
            Check(LaunchM8Compact(st, ba, placed), "compaction");
            M8BatchArgs rb = ba;   // the rest
            rb.coff = b.coff_b;
            refine(rb, ts + 2 * tsn, ts + 3 * tsn, out_ids, out_d);
        }
        Check(hipEventRecord(ctx->ev[3], st), "event");
        Check(hipMemcpyAsync(bst.data(), b.bstats, 8 * 8, hipMemcpyDeviceToHost, st), "download stats");
        Check(hipMemcpyAsync(wst.data(), b.wstats, 4 * 8, hipMemcpyDeviceToHost, st), "download stats");
        if (chk) Check(hipMemcpyAsync(chk, b.chk, nq * 4, hipMemcpyDeviceToHost, st), "download checksums");
    }
    if (!a->outputs_on_device) {
        Check(hipMemcpyAsync(a->out_ids, oid, nq * a->k * 4, hipMemcpyDeviceToHost, st), "download ids");
        Check(hipMemcpyAsync(a->out_dists, od, nq * a->k * 4, hipMemcpyDeviceToHost, st), "download dists");
    }
    std::vector<uint64_t> hs(nq * 8);
    Check(hipMemcpyAsync(hs.data(), qst, nq * 8 * 8, hipMemcpyDeviceToHost, st), "download stats");
    Check(hipStreamSynchronize(st), "search");
    const bool ran_batched = batched && novf == 0;   // else the plain form answered every query
    if (per_query_stats) {
        if (ran_batched) std::fill(per_query_stats, per_query_stats + nq * 5, 0ull);
        else
            for (uint64_t i = 0; i < nq; ++i) std::memcpy(per_query_stats + i * 5, hs.data() + i * 8, 5 * 8);
    }
    if (stats) {
        uint64_t s8[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // threshold pass (batched) or the whole search (plain)
        for (uint64_t i = 0; i < nq; ++i)
            for (int f = 0; f < 8; ++f) s8[f] += hs[i * 8 + f];
        if (!ran_batched) {
            stats->overflowed = novf;
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
            // pass's refined / survivors / bits plus the batched ones
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
    if (mode == E8Q_MODE_BATCHED && M8BatchedFits(x, k)) M8CarveBatched(cv, b, x, nq, k, nq * maxp, false);
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
            // the batched form's refine records, for indexes it runs on
            if (nrun >= 1 && nrun <= kM8BatchRuns && v.n > 0) {
                x->rrec.alloc(v.n * v.rrec_bytes);
                x->bscale.alloc(x->nbatch * nrun * sizeof(float4));
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

}  // extern "C"
