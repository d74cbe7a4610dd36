// e8q_runtime.hip -- the C ABI of libe8q_gpu.so (include/e8q_gpu.h):
// device context, device-resident index, workspace, and the host side of
// every kernel launch.  No exception crosses the ABI: failures return an
// e8q_status and leave a thread-local message for e8q_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/e8q_gpu.h"
#include "diag_kernels.h"
#include "e8q_device.h"
#include "e8q_kernels.h"

using namespace e8qdev;

#include "e8q_rt.h"

using namespace e8qrt;

struct e8q_dindex {
    e8q_ctx* ctx = nullptr;
    IndexView view{};
    EncodeView ev{};
    float alpha = 1.0F;
    uint64_t arena_bytes = 0;   // E8AX arena size: records + 64-byte tail
    uint64_t records() const { return view.batch_bytes ? (arena_bytes - 64) / view.batch_bytes : 0; }
    std::vector<uint64_t> cluster_begin, arena_begin, ref_arena_begin;
    DBuf signs, mean_rot, centroid_rot, centroid_n2, d_cluster_begin, d_arena_begin, d_ref_arena_begin, arena,
        ref_arena, ids, refine_tab, gate_mid, levels, mlev, mdig, cls_lev, cls_dig, csum;
    // MFMA form: folded one-hot code layout and per-list max |g1|, built on
    // its first search and dropped when the arrays change.
    struct Mfma {
        DBuf codes, g1max;
        float lmax = 0.0F;
        bool ready = false;
    };
    mutable Mfma mfma;
    uint64_t device_bytes() const {
        return signs.bytes + mean_rot.bytes + centroid_rot.bytes + centroid_n2.bytes + d_cluster_begin.bytes +
               d_arena_begin.bytes + d_ref_arena_begin.bytes + arena.bytes + ref_arena.bytes + ids.bytes +
               refine_tab.bytes + gate_mid.bytes + levels.bytes + mlev.bytes + mdig.bytes + cls_lev.bytes +
               cls_dig.bytes + csum.bytes + mfma.codes.bytes + mfma.g1max.bytes;
    }
};

namespace {

void BuildEncoderTables(e8q_dindex* ix, const double* levels, int bits, hipStream_t st) {
    const EncTabHost t = MakeEncoderTables(levels, bits);
    const uint32_t q = 1u << bits;
    const bool alt = t.merged;
    Upload(ix->levels, levels, 2 * q, st);
    Upload(ix->mlev, t.mlev.data(), t.mlev.size(), st);
    Upload(ix->mdig, t.mdig.data(), t.mdig.size(), st);
    Upload(ix->cls_lev, t.cl.data(), t.cl.size(), st);
    Upload(ix->cls_dig, t.cd.data(), t.cd.size(), st);
    Check(hipStreamSynchronize(st), "encoder tables");
    ix->ev.bits = bits;
    ix->ev.merged = alt ? 1 : 0;
    ix->ev.mlen = 2 * q;
    ix->ev.mlev = ix->mlev.as<double>();
    ix->ev.mdig = ix->mdig.as<uint16_t>();
    ix->ev.levels = ix->levels.as<double>();
    ix->ev.cls_lev = ix->cls_lev.as<double>();
    ix->ev.cls_dig = ix->cls_dig.as<uint16_t>();
}

void RefreshView(e8q_dindex* x) {
    IndexView& v = x->view;
    v.signs = x->signs.as<uint8_t>();
    v.mean_rot = x->mean_rot.as<float>();
    v.centroid_rot = x->centroid_rot.as<float>();
    v.centroid_n2 = x->centroid_n2.as<float>();
    v.cluster_begin = x->d_cluster_begin.as<uint64_t>();
    v.arena_begin = x->d_arena_begin.as<uint64_t>();
    v.ref_arena_begin = x->d_ref_arena_begin.as<uint64_t>();
    v.arena = x->arena.as<uint8_t>();
    v.ref_arena = x->ref_arena.as<uint8_t>();
    v.ids = x->ids.as<uint32_t>();
    v.refine_tab = x->refine_tab.as<float>();
    v.gate_mid = x->gate_mid.as<double>();
    v.csum = x->csum.as<float4>();
}

}  // namespace

// ===========================================================================
namespace {

uint64_t FloatBits(float f) {
    uint32_t u = 0;
    std::memcpy(&u, &f, 4);
    return u;
}

// Power-of-two bucket of a query count (the tuner's cache key).
uint64_t QueryBucket(uint64_t nq) {
    uint64_t b = 0;
    while ((uint64_t{1} << b) < nq) ++b;
    return b;
}

// Upper bound of the batched scan's work items at a tile shape: over the
// lists, ceil(pairs_c / qtile) * ceil(batches_c / btile).
uint64_t ItemsCap(const e8q_dindex* x, uint64_t nq, uint64_t pairs, uint32_t qtile, uint32_t btile) {
    uint64_t sum_nbt = 0, max_nbt = 0;
    for (uint64_t c = 0; c < x->view.nlist; ++c) {
        const uint64_t nb = (x->cluster_begin[c + 1] - x->cluster_begin[c] + 31) / 32;
        const uint64_t nbt = (nb + btile - 1) / btile;
        sum_nbt += nbt;
        max_nbt = std::max(max_nbt, nbt);
    }
    return std::min((nq + qtile - 1) / qtile * sum_nbt, (pairs + qtile - 1) / qtile * max_nbt + sum_nbt);
}

struct ScanShape {
    uint32_t qtile, btile, grid;
};

bool ShapeFits(const IndexView& v, uint32_t qtile, uint32_t btile) {
    return qtile >= 1 && qtile <= 64 && btile >= 2 && btile <= 64 && (btile & 1u) == 0 &&
           BatchedScanLds(v, qtile, btile) <= 160 * 1024;
}

// The tuner's candidates: pairs and batches per item at half, once and
// twice the default shape (default first, so a tie keeps it).
std::vector<ScanShape> ShapeCandidates(const IndexView& v, uint32_t grid) {
    uint32_t q0 = 0, b0 = 0;
    BatchTiles(v, &q0, &b0);
    std::vector<ScanShape> out{{q0, b0, grid}};
    for (uint32_t qm = 0; qm < 3; ++qm)
        for (uint32_t bm = 0; bm < 3; ++bm) {
            const uint32_t qt = std::clamp<uint32_t>(qm == 0 ? q0 / 2 : q0 << (qm - 1), 4, 32);
            const uint32_t bt = std::clamp<uint32_t>((bm == 0 ? b0 / 2 : b0 << (bm - 1)) & ~1u, 2, 32);
            if (!ShapeFits(v, qt, bt)) continue;
            bool dup = false;
            for (const ScanShape& c : out) dup = dup || (c.qtile == qt && c.btile == bt);
            if (!dup) out.push_back({qt, bt, grid});
        }
    return out;
}

constexpr uint64_t kTuneMinQueries = 64;   // smaller calls use the default shape untuned
constexpr uint32_t kMfmaCap = 4096;         // MFMA-form candidates kept per query (more: a second pass)
constexpr uint32_t kMfmaTauVectors = 16384;  // the MFMA form's threshold pass reads this much of the first list

// Builds the MFMA form's code layout of an index on its first MFMA search.
// The fold needs the refine table's reflection symmetry, checked here.
void PrepareMfma(e8q_ctx* ctx, const e8q_dindex* x) {
    if (x->mfma.ready) return;
    const IndexView& v = x->view;
    if (v.bits < 1 || v.bits > static_cast<int>(kMfmaMaxBits)) Throw(E8Q_EINVAL, "the MFMA form needs bits <= 4");
    hipStream_t st = ctx->stream;
    const uint32_t Q = 1u << v.bits;
    std::vector<float> tab(2ull * v.tab_stride);
    Check(hipMemcpyAsync(tab.data(), v.refine_tab, tab.size() * 4, hipMemcpyDeviceToHost, st), "download table");
    Check(hipStreamSynchronize(st), "download table");
    float lmax = 0.0F;
    for (uint32_t d = 0; d < Q; ++d) {
        if (tab[v.tab_stride + d] != -tab[Q - 1 - d])
            Throw(E8Q_EINVAL, "the MFMA form needs a reflection-symmetric level table");
        lmax = std::max(lmax, std::fabs(tab[d]));
    }
    for (uint64_t c = 0; c < v.nlist; ++c)
        if (x->arena_begin[c] % v.batch_bytes != 0) Throw(E8Q_EINVAL, "scan records are not record-aligned");
    const uint64_t records = x->records();
    e8q_dindex::Mfma m;
    m.codes.alloc(records * 32 * v.padded);
    m.g1max.alloc(v.nlist * 4);
    Check(LaunchMfmaCodes(st, v, records, m.codes.as<uint8_t>(), m.g1max.as<float>()), "mfma codes");
    Check(hipStreamSynchronize(st), "mfma codes");
    m.lmax = lmax;
    m.ready = true;
    x->mfma = std::move(m);
}

}  // namespace

extern "C" {

const char* e8q_last_error(void) { return g_err.c_str(); }

e8q_status e8q_ctx_create(int device, e8q_ctx** out) {
    return Guard([&] {
        if (!out) Throw(E8Q_EINVAL, "out is null");
        int n = 0;
        Check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (device < 0 || device >= n) Throw(E8Q_EINVAL, "device index out of range");
        auto* c = new e8q_ctx();
        c->device = device;
        try {
            c->Bind();
            Check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
            Check(hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device), "device attribute");
            for (auto& e : c->ev) Check(hipEventCreate(&e), "hipEventCreate");
            for (auto& e : c->kev) Check(hipEventCreate(&e), "hipEventCreate");
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
    });
}

e8q_status e8q_ctx_destroy(e8q_ctx* ctx) {
    return Guard([&] {
        if (!ctx) return;
        ctx->Bind();
        (void)hipStreamSynchronize(ctx->stream);
        ctx->ClearGraphs();
        ctx->ws.reset();
        for (auto& e : ctx->ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : ctx->kev)
            if (e) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

e8q_status e8q_ctx_reserve(e8q_ctx* ctx, uint64_t bytes) {
    return Guard([&] {
        if (!ctx) Throw(E8Q_EINVAL, "ctx is null");
        ctx->Bind();
        Check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        if (ctx->ws.bytes < bytes) ctx->ws.alloc(bytes);
        ctx->reserved = std::max(ctx->reserved, bytes);
    });
}

uint64_t e8q_ctx_workspace_bytes(const e8q_ctx* ctx) { return ctx ? ctx->ws.bytes : 0; }

e8q_status e8q_ctx_synchronize(e8q_ctx* ctx) {
    return Guard([&] {
        if (!ctx) Throw(E8Q_EINVAL, "ctx is null");
        ctx->Bind();
        Check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

e8q_status e8q_index_create(e8q_ctx* ctx, const e8q_index_desc* d, e8q_dindex** out) {
    return Guard([&] {
        if (!ctx || !d || !out) Throw(E8Q_EINVAL, "null argument");
        if (d->bits < 1 || d->bits > kMaxBits) Throw(E8Q_EINVAL, "E8Q bits must be in [1, 13]");
        if (d->dim == 0 || d->nlist == 0) Throw(E8Q_EINVAL, "dim and nlist must be positive");
        if (d->padded != (d->dim + 63) / 64 * 64) Throw(E8Q_EINVAL, "padded must be dim rounded up to 64");
        if (!d->signs || !d->refine_tab || !d->gate_mid || !d->levels || !d->cluster_begin || !d->arena_begin ||
            !d->ref_arena_begin)
            Throw(E8Q_EINVAL, "missing table");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        auto* x = new e8q_dindex();
        try {
            x->ctx = ctx;
            x->alpha = d->alpha;
            IndexView& v = x->view;
            v.n = d->n;
            v.dim = d->dim;
            v.padded = d->padded;
            v.nlist = d->nlist;
            v.bits = d->bits;
            v.blocks = static_cast<int>(d->padded / 8);
            v.batch_bytes = static_cast<uint32_t>((kHdr + 32 * static_cast<uint64_t>(v.blocks) + 63) & ~63ull);
            v.refine_stride = d->bits >= 2 ? static_cast<uint32_t>(v.blocks * (d->bits - 1)) : 0;
            v.ref_batch_bytes = static_cast<uint32_t>((32ull * v.refine_stride + 256 + 63) & ~63ull);
            v.tab_stride = std::max<uint32_t>(256, 1u << d->bits);
            uint32_t lg = 0;
            while ((uint64_t{1} << (lg + 1)) <= d->dim) ++lg;
            v.trunc = 1u << lg;
            v.fac = 1.0F / std::sqrt(static_cast<float>(v.trunc));
            const uint64_t L = d->nlist;
            x->cluster_begin.assign(d->cluster_begin, d->cluster_begin + L + 1);
            x->arena_begin.assign(d->arena_begin, d->arena_begin + L + 1);
            x->ref_arena_begin.assign(d->ref_arena_begin, d->ref_arena_begin + L + 1);
            if (x->cluster_begin[0] != 0 || x->cluster_begin[L] != d->n) Throw(E8Q_EINVAL, "cluster offsets inconsistent");
            uint64_t ab = 0, rb = 0;
            for (uint64_t c = 0; c < L; ++c) {
                if (x->cluster_begin[c + 1] < x->cluster_begin[c]) Throw(E8Q_EINVAL, "cluster offsets inconsistent");
                const uint64_t nb = (x->cluster_begin[c + 1] - x->cluster_begin[c] + 31) / 32;
                if (x->arena_begin[c] != ab || x->ref_arena_begin[c] != rb) Throw(E8Q_EINVAL, "arena offsets inconsistent");
                ab += nb * v.batch_bytes;
                rb += nb * v.ref_batch_bytes;
            }
            if (x->arena_begin[L] != ab || x->ref_arena_begin[L] != rb) Throw(E8Q_EINVAL, "arena offsets inconsistent");
            Upload(x->signs, d->signs, 4 * d->padded / 8, st);
            Upload(x->refine_tab, d->refine_tab, 2 * static_cast<size_t>(v.tab_stride), st);
            Upload(x->gate_mid, d->gate_mid, 16, st);
            Upload(x->d_cluster_begin, d->cluster_begin, L + 1, st);
            Upload(x->d_arena_begin, d->arena_begin, L + 1, st);
            Upload(x->d_ref_arena_begin, d->ref_arena_begin, L + 1, st);
            x->mean_rot.alloc(d->padded * sizeof(float));
            x->centroid_rot.alloc(L * d->padded * sizeof(float));
            x->centroid_n2.alloc(L * sizeof(float));
            // The E8AX arena carries a 64-byte zero tail (e8q_io.cpp:35-84).
            x->arena_bytes = ab + 64;
            x->arena.alloc(ab + 64);
            x->ref_arena.alloc(rb + 64);
            x->ids.alloc(std::max<uint64_t>(d->n, 1) * sizeof(uint32_t));
            x->csum.alloc(L * sizeof(float4));
            Check(hipMemsetAsync(x->arena.p, 0, x->arena.bytes, st), "memset");
            Check(hipMemsetAsync(x->ref_arena.p, 0, x->ref_arena.bytes, st), "memset");
            BuildEncoderTables(x, d->levels, d->bits, st);
            x->ev.scale = d->scale;
            RefreshView(x);
        } catch (...) {
            delete x;
            throw;
        }
        *out = x;
    });
}

e8q_status e8q_index_destroy(e8q_dindex* idx) {
    return Guard([&] {
        if (!idx) return;
        idx->ctx->Bind();
        (void)hipStreamSynchronize(idx->ctx->stream);
        delete idx;
    });
}

uint64_t e8q_index_device_bytes(const e8q_dindex* idx) { return idx ? idx->device_bytes() : 0; }

e8q_status e8q_index_upload(e8q_ctx* ctx, e8q_dindex* x, const e8q_index_arrays* h) {
    return Guard([&] {
        if (x) x->mfma = e8q_dindex::Mfma{};
        if (!ctx || !x || !h || !h->ids || !h->mean_rot || !h->centroid_rot || !h->centroid_n2 || !h->arena ||
            !h->ref_arena)
            Throw(E8Q_EINVAL, "null argument");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        const IndexView& v = x->view;
        Check(hipMemcpyAsync(x->ids.p, h->ids, v.n * 4, hipMemcpyHostToDevice, st), "upload ids");
        Check(hipMemcpyAsync(x->mean_rot.p, h->mean_rot, v.padded * 4, hipMemcpyHostToDevice, st), "upload mean");
        Check(hipMemcpyAsync(x->centroid_rot.p, h->centroid_rot, v.nlist * v.padded * 4, hipMemcpyHostToDevice, st),
              "upload centroids");
        Check(hipMemcpyAsync(x->centroid_n2.p, h->centroid_n2, v.nlist * 4, hipMemcpyHostToDevice, st), "upload cn2");
        Check(hipMemcpyAsync(x->arena.p, h->arena, x->arena_bytes, hipMemcpyHostToDevice, st), "upload arena");
        Check(LaunchRecordLayout(st, v, x->arena.as<uint8_t>(), x->arena.as<uint8_t>(), x->records(), true),
              "arena layout");
        Check(hipMemcpyAsync(x->ref_arena.p, h->ref_arena, x->ref_arena.bytes, hipMemcpyHostToDevice, st),
              "upload refine arena");
        Check(LaunchClusterSummary(st, v, x->csum.as<float4>()), "cluster summary");
        Check(hipStreamSynchronize(st), "upload");
    });
}

e8q_status e8q_index_download(e8q_ctx* ctx, const e8q_dindex* x, e8q_index_arrays* h) {
    return Guard([&] {
        if (!ctx || !x || !h) Throw(E8Q_EINVAL, "null argument");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        const IndexView& v = x->view;
        auto dl = [&](void* dst, const DBuf& src, size_t bytes, const char* what) {
            if (dst) Check(hipMemcpyAsync(dst, src.p, bytes, hipMemcpyDeviceToHost, st), what);
        };
        dl(h->ids, x->ids, v.n * 4, "download ids");
        dl(h->mean_rot, x->mean_rot, v.padded * 4, "download mean");
        dl(h->centroid_rot, x->centroid_rot, v.nlist * v.padded * 4, "download centroids");
        dl(h->centroid_n2, x->centroid_n2, v.nlist * 4, "download cn2");
        if (h->arena) {   // back to the E8AX rowA layout, through a scratch copy
            DBuf tmp;
            tmp.alloc(x->arena_bytes);
            Check(hipMemcpyAsync(tmp.p, x->arena.p, x->arena_bytes, hipMemcpyDeviceToDevice, st), "copy arena");
            Check(LaunchRecordLayout(st, v, tmp.as<uint8_t>(), tmp.as<uint8_t>(), x->records(), false),
                  "arena layout");
            Check(hipMemcpyAsync(h->arena, tmp.p, x->arena_bytes, hipMemcpyDeviceToHost, st), "download arena");
            Check(hipStreamSynchronize(st), "download");
        }
        dl(h->ref_arena, x->ref_arena, x->ref_arena.bytes, "download refine arena");
        Check(hipStreamSynchronize(st), "download");
    });
}

e8q_status e8q_index_encode(e8q_ctx* ctx, e8q_dindex* x, const float* base, const float* centroids,
                            const uint32_t* cluster_ids, const uint32_t* slot, int on_device) {
    return Guard([&] {
        if (x) x->mfma = e8q_dindex::Mfma{};
        if (!ctx || !x || !base || !centroids || !cluster_ids || !slot) Throw(E8Q_EINVAL, "null argument");
        const IndexView& v = x->view;
        if (v.n == 0) Throw(E8Q_EINVAL, "empty build set");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        DBuf d_base, d_cent, d_cid, d_slot, mu, ob, ow, op;
        const float* pb = base;
        const float* pc = centroids;
        const uint32_t* pid = cluster_ids;
        for (uint64_t i = 0; i < v.n; ++i)
            if (slot[i] >= v.n) Throw(E8Q_EINVAL, "slot out of range");
        if (!on_device) {
            for (uint64_t i = 0; i < v.n; ++i)
                if (cluster_ids[i] >= v.nlist) Throw(E8Q_EINVAL, "cluster id out of range");
            Upload(d_base, base, v.n * v.dim, st);
            Upload(d_cent, centroids, v.nlist * v.dim, st);
            Upload(d_cid, cluster_ids, v.n, st);
            pb = d_base.as<float>();
            pc = d_cent.as<float>();
            pid = d_cid.as<uint32_t>();
        }
        Upload(d_slot, slot, v.n, st);
        const uint32_t* ps = d_slot.as<uint32_t>();
        mu.alloc(v.dim * sizeof(double));
        ob.alloc(v.n * sizeof(float));
        ow.alloc(v.n * sizeof(float));
        op.alloc(v.n * sizeof(float));
        Check(hipMemsetAsync(ob.p, 0, ob.bytes, st), "memset");
        Check(hipMemsetAsync(ow.p, 0, ow.bytes, st), "memset");
        Check(hipMemsetAsync(op.p, 0, op.bytes, st), "memset");
        Check(hipMemsetAsync(x->arena.p, 0, x->arena.bytes, st), "memset");
        Check(hipMemsetAsync(x->ref_arena.p, 0, x->ref_arena.bytes, st), "memset");
        Check(LaunchMean(st, pb, v.n, static_cast<uint32_t>(v.dim), mu.as<double>()), "mean");
        Check(LaunchMeanFinish(st, mu.as<double>(), v.n, v, x->mean_rot.as<float>()), "mean finish");
        Check(LaunchCentroids(st, pc, v, x->centroid_rot.as<float>(), x->centroid_n2.as<float>()), "centroids");
        EncodeArgs a{};
        a.base = pb;
        a.cluster_ids = pid;
        a.slot = ps;
        a.n = v.n;
        a.arena = x->arena.as<uint8_t>();
        a.ref_arena = x->ref_arena.as<uint8_t>();
        a.ids = x->ids.as<uint32_t>();
        a.ob = ob.as<float>();
        a.ow = ow.as<float>();
        a.op = op.as<float>();
        Check(LaunchEncodeRows(st, v, x->ev, a), "encode rows");
        Check(LaunchBanks(st, v, ob.as<float>(), ow.as<float>(), op.as<float>(), x->arena.as<uint8_t>()), "banks");
        Check(LaunchClusterSummary(st, v, x->csum.as<float4>()), "cluster summary");
        Check(hipStreamSynchronize(st), "encode");
    });
}

uint64_t e8q_search_workspace_bytes(const e8q_dindex* x, uint64_t nq, uint64_t k, uint64_t max_nprobe) {
    if (!x) return 0;
    const IndexView& v = x->view;
    const uint64_t maxp = std::min<uint64_t>(max_nprobe, v.nlist);
    const uint64_t P = v.padded, E = static_cast<uint64_t>(v.blocks) * 32;
    float *dq = nullptr, *rotq = nullptr, *consts = nullptr, *dist = nullptr, *pd2 = nullptr, *odist = nullptr;
    uint8_t* lut = nullptr;
    uint32_t *pids = nullptr, *npq = nullptr, *oids = nullptr;
    unsigned long long* dstats = nullptr;
    PlainParams* dparams = nullptr;
    e8q_ctx::Carve cv;   // the plain form's carve in e8q_search, host queries and outputs
    cv.add(&dq, nq * v.dim);
    cv.add(&rotq, nq * P);
    cv.add(&lut, nq * E);
    cv.add(&consts, nq * 8);
    cv.add(&dist, nq * v.nlist);
    cv.add(&pids, nq * maxp);
    cv.add(&pd2, nq * maxp);
    cv.add(&npq, nq);
    cv.add(&oids, nq * k);
    cv.add(&odist, nq * k);
    cv.add(&dstats, 4);
    cv.add(&dparams, 1);
    return cv.bytes();
}


e8q_status e8q_ctx_set_graphs(e8q_ctx* ctx, int enable) {
    return Guard([&] {
        if (!ctx) Throw(E8Q_EINVAL, "ctx is null");
        ctx->Bind();
        Check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        ctx->graphs = enable != 0;
        if (!ctx->graphs) ctx->ClearGraphs();
    });
}

e8q_status e8q_ctx_set_batched_shape(e8q_ctx* ctx, uint32_t qtile, uint32_t btile, uint32_t grid) {
    return Guard([&] {
        if (!ctx) Throw(E8Q_EINVAL, "ctx is null");
        if ((qtile == 0) != (btile == 0)) Throw(E8Q_EINVAL, "qtile and btile are set together");
        if (qtile != 0 && (qtile > 64 || btile < 2 || btile > 64 || (btile & 1u)))
            Throw(E8Q_EINVAL, "qtile must be in [1, 64] and btile even in [2, 64]");
        if (grid > 65536) Throw(E8Q_EINVAL, "grid must be <= 65536");
        ctx->fix_qtile = qtile;
        ctx->fix_btile = btile;
        ctx->fix_grid = grid;
    });
}

e8q_status e8q_ctx_get_info(const e8q_ctx* ctx, e8q_ctx_info* out) {
    return Guard([&] {
        if (!ctx || !out) Throw(E8Q_EINVAL, "null argument");
        *out = e8q_ctx_info{ctx->last_qtile, ctx->last_btile, ctx->last_grid, ctx->graphs ? 1u : 0u,
                            ctx->tunings, ctx->captures, ctx->replays};
    });
}

e8q_status e8q_search(e8q_ctx* ctx, const e8q_dindex* x, const e8q_search_args* a, e8q_search_stats* stats) {
    return Guard([&] {
        if (!ctx || !x || !a) Throw(E8Q_EINVAL, "null argument");
        if (a->k == 0 || a->k > static_cast<uint64_t>(kMaxK))
            Throw(E8Q_EINVAL, "k must be in [1, " + std::to_string(kMaxK) + "]");
        if (!a->queries || !a->out_ids || !a->out_dists) Throw(E8Q_EINVAL, "null buffer");
        const int flush = a->flush_every > 0 ? a->flush_every : 32;
        if (flush > kMaxFlush)
            Throw(E8Q_EINVAL, "flush_every must be <= 1024 (the survivor staging capacity, as in the reference)");
        if (a->mode != E8Q_MODE_PLAIN && a->mode != E8Q_MODE_BATCHED && a->mode != E8Q_MODE_MFMA)
            Throw(E8Q_EINVAL, "unknown search mode");
        const IndexView& v = x->view;
        const uint64_t nq = a->nq;
        if (stats) *stats = e8q_search_stats{0, 0, 0, 0.0, 0.0, 0};
        if (nq == 0) return;
        uint32_t maxp = 0;
        if (a->nprobe_per_query) {
            for (uint64_t i = 0; i < nq; ++i) {
                if (a->nprobe_per_query[i] == 0) Throw(E8Q_EINVAL, "nprobe must be positive");
                maxp = std::max<uint32_t>(maxp, a->nprobe_per_query[i]);
            }
        } else {
            if (a->nprobe == 0) Throw(E8Q_EINVAL, "nprobe must be positive");
            maxp = static_cast<uint32_t>(std::min<uint64_t>(a->nprobe, v.nlist));
        }
        maxp = static_cast<uint32_t>(std::min<uint64_t>(maxp, v.nlist));
        const uint32_t nprobe_all = static_cast<uint32_t>(std::min<uint64_t>(a->nprobe, v.nlist));
        const size_t sel_lds = SelectProbesLds(static_cast<uint32_t>(v.nlist), maxp);
        if (sel_lds > 160 * 1024) Throw(E8Q_EINVAL, "nlist too large for the probe selector");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        const bool batched = a->mode == E8Q_MODE_BATCHED, mfma = a->mode == E8Q_MODE_MFMA;
        const bool listed = batched || mfma;   // forms over the (query, list) work list
        if (mfma) PrepareMfma(ctx, x);
        const uint64_t P = v.padded, E = static_cast<uint64_t>(v.blocks) * 32;
        float *dq = nullptr, *rotq = nullptr, *consts = nullptr, *dist = nullptr, *pd2 = nullptr, *odist = nullptr;
        uint8_t* lut = nullptr;
        uint32_t *pids = nullptr, *npq = nullptr, *oids = nullptr;
        unsigned long long *dstats = nullptr, *tstats = nullptr;
        PlainParams* dparams = nullptr;
        e8q_ctx::Carve cv;
        if (!a->queries_on_device) cv.add(&dq, nq * v.dim);
        cv.add(&rotq, nq * P);
        cv.add(&lut, nq * E);
        cv.add(&consts, nq * 8);
        cv.add(&dist, nq * v.nlist);
        cv.add(&pids, nq * maxp);
        cv.add(&pd2, nq * maxp);
        if (a->nprobe_per_query) cv.add(&npq, nq);
        if (!a->outputs_on_device) {
            cv.add(&oids, nq * a->k);
            cv.add(&odist, nq * a->k);
        }
        cv.add(&dstats, 4);
        cv.add(&dparams, 1);
        // Batched and MFMA forms: the scan shape (batched: fixed, tuned
        // before, or to be tuned now), first-probe lists, work list,
        // candidates.
        BatchedArgs ba{};
        MfmaExtra mx{};
        MfmaParams* mparams = nullptr;
        uint16_t* m_qexp = nullptr;
        uint32_t* m_thr = nullptr;
        uint64_t* m_cand = nullptr;
        BatchedParams* bparams = nullptr;
        uint32_t* first_ids = nullptr;
        float* first_dists = nullptr;
        uint64_t pairs = 0;
        ScanShape shape{0, 0, 0};
        std::vector<ScanShape> cands;   // non-empty: tune in this call
        const uint64_t tkey[4] = {v.batch_bytes, static_cast<uint64_t>(v.blocks), v.nlist, QueryBucket(nq)};
        if (mfma) {
            shape = {kMfmaQTile, kMfmaBTile, static_cast<uint32_t>(std::max(1, ctx->cus) * 2)};
        } else if (batched) {
            const uint32_t grid0 = static_cast<uint32_t>(std::max(1, ctx->cus) * 2);
            if (ctx->fix_qtile != 0) {
                shape = {ctx->fix_qtile, ctx->fix_btile, ctx->fix_grid ? ctx->fix_grid : grid0};
                if (!ShapeFits(v, shape.qtile, shape.btile))
                    Throw(E8Q_EINVAL, "batched shape needs more than 160 KiB of LDS for this index");
            } else {
                bool found = false;
                for (const auto& t : ctx->tuned)
                    if (std::equal(t.shape, t.shape + 4, tkey)) {
                        shape = {t.qtile, t.btile, t.grid};
                        found = true;
                    }
                if (!found) {
                    cands = ShapeCandidates(v, grid0);
                    shape = cands[0];
                    if (nq < kTuneMinQueries) cands.clear();
                }
                if (ctx->fix_grid != 0) shape.grid = ctx->fix_grid;
            }
        }
        if (listed) {
            const uint64_t p_lo = batched ? 1 : 0;   // the MFMA form scores the first probe too
            if (a->nprobe_per_query) {
                for (uint64_t i = 0; i < nq; ++i)
                    pairs += std::min<uint64_t>(a->nprobe_per_query[i], v.nlist) - p_lo;
            } else {
                pairs = nq * (maxp - p_lo);
            }
            uint64_t items_cap = ItemsCap(x, nq, pairs, shape.qtile, shape.btile);
            for (const ScanShape& c : cands) items_cap = std::max(items_cap, ItemsCap(x, nq, pairs, c.qtile, c.btile));
            ba.qtile = shape.qtile;
            ba.btile = shape.btile;
            ba.cap = batched ? kBatchCandCap : kMfmaCap;
            cv.add(&first_ids, nq * a->k);
            cv.add(&first_dists, nq * a->k);
            cv.add(&ba.ccount, v.nlist);
            cv.add(&ba.cfill, v.nlist);
            cv.add(&ba.coff, v.nlist + 1);
            cv.add(&ba.ioff, v.nlist + 1);
            cv.add(&ba.pairs, std::max<uint64_t>(pairs, 1));
            cv.add(&ba.items, std::max<uint64_t>(items_cap, 1));
            ba.items_cap = ItemsCap(x, nq, pairs, shape.qtile, shape.btile);
            cv.add(&ba.counters, 4);
            cv.add(&ba.ncand, nq);
            cv.add(&ba.overflow, nq);
            if (batched) {
                cv.add(&ba.cand, nq * ba.cap);
                cv.add(&bparams, 1);
            } else {
                cv.add(&m_qexp, nq * P * 32);
                cv.add(&m_thr, nq);
                cv.add(&m_cand, nq * kMfmaCap);
                cv.add(&mparams, 1);
            }
            if (!cands.empty() || mfma) cv.add(&tstats, 4);
        }
        ctx->Take(cv);
        if (mfma) {   // carve pointers are final only after Take
            bparams = reinterpret_cast<BatchedParams*>(mparams);   // its first member
            mx.qexp = m_qexp;
            mx.thr = m_thr;
            mx.dyn = a->k <= kMfmaDynK ? 1u : 0u;
            mx.cand = m_cand;
            mx.mcodes = x->mfma.codes.as<uint8_t>();
            mx.ncand = ba.ncand;
            mx.cap = kMfmaCap;
        }
        const float* qsrc = a->queries_on_device ? a->queries : dq;
        if (a->outputs_on_device) {
            oids = a->out_ids;
            odist = a->out_dists;
        }
        PlainArgs pa{};
        pa.nq = nq;
        pa.rotq = rotq;
        pa.lut = lut;
        pa.consts = consts;
        pa.probe_ids = pids;
        pa.probe_d2 = pd2;
        pa.maxp = maxp;
        pa.nprobe_q = npq;
        pa.nprobe_all = maxp;
        pa.k = a->k;
        pa.z = a->z;
        pa.disable_prune = a->disable_prune ? 1 : 0;
        pa.flush_every = flush;
        pa.out_ids = oids;
        pa.out_dists = odist;
        pa.stats = dstats;
        // Threshold pass of the batched form: the plain search over each
        // query's first probe.
        PlainArgs fa = pa;
        fa.nprobe_q = nullptr;
        fa.nprobe_all = 1;
        fa.out_ids = first_ids;
        fa.out_dists = first_dists;
        if (listed) {
            ba.first_ids = first_ids;
            ba.first_dists = first_dists;
            ba.nq = nq;
            ba.maxp = maxp;
            ba.nprobe_q = npq;
            ba.nprobe_all = maxp;
            ba.probe_ids = pids;
            ba.probe_d2 = pd2;
            ba.lut = lut;
            ba.consts = consts;
            ba.rotq = rotq;
            ba.k = static_cast<uint32_t>(a->k);
            ba.z = a->z;
            ba.disable_prune = a->disable_prune ? 1 : 0;
            ba.out_ids = oids;
            ba.out_dists = odist;
            ba.stats = dstats;
            ba.p_lo = mfma ? 0 : 1;
            ba.qmask = nullptr;
        }
        if (mfma) {
            fa.stats = tstats;   // the work list counts the first probe's candidates
            fa.max_vectors = kMfmaTauVectors;
        }
        auto batched_tail = [&](const BatchedArgs& b, uint32_t grid, bool upload) {
            Check(LaunchFill32(st, b.ccount, v.nlist, 0), "fill");
            Check(LaunchFill32(st, b.ncand, nq, 0), "fill");
            Check(LaunchBatchedWorkList(st, v, b, bparams, upload), "work list");
            Check(LaunchBatchedScan(st, v, b, bparams, grid), "batched scan");
            Check(LaunchBatchedRefine(st, v, b, bparams), "batched refine");
        };
        // Times the batched tail at each candidate shape (twice, keeping the
        // faster), then the persistent grid at the best tile; results do not
        // depend on the shape.  Counts go to a scratch stats block.
        auto tune = [&] {
            auto time_shape = [&](const ScanShape& c) {
                BatchedArgs t = ba;
                t.qtile = c.qtile;
                t.btile = c.btile;
                t.items_cap = ItemsCap(x, nq, pairs, c.qtile, c.btile);
                t.stats = tstats;
                float best = 3.0e38F;
                for (int rep = 0; rep < 2; ++rep) {
                    Check(hipEventRecord(ctx->ev[2], st), "event");
                    batched_tail(t, c.grid, true);
                    Check(hipEventRecord(ctx->ev[3], st), "event");
                    Check(hipEventSynchronize(ctx->ev[3]), "tuning");
                    float ms = 0.0F;
                    Check(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]), "event time");
                    best = std::min(best, ms);
                }
                return best;
            };
            ScanShape best = cands[0];
            float best_ms = time_shape(best);
            for (size_t i = 1; i < cands.size(); ++i) {
                const float ms = time_shape(cands[i]);
                if (ms < best_ms) {
                    best_ms = ms;
                    best = cands[i];
                }
            }
            if (ctx->fix_grid == 0) {
                const uint32_t cu = static_cast<uint32_t>(std::max(1, ctx->cus));
                for (uint32_t g : {cu, 4 * cu}) {
                    const ScanShape c{best.qtile, best.btile, g};
                    const float ms = time_shape(c);
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = c;
                    }
                }
            }
            ctx->tuned.push_back({{tkey[0], tkey[1], tkey[2], tkey[3]}, best.qtile, best.btile, best.grid});
            ++ctx->tunings;
            shape = best;
            if (ctx->fix_grid != 0) shape.grid = ctx->fix_grid;
            ba.qtile = shape.qtile;
            ba.btile = shape.btile;
            ba.items_cap = ItemsCap(x, nq, pairs, shape.qtile, shape.btile);
        };
        // The device sequence from the query rotation to the results.  With
        // upload false the launchers' parameter blocks are already on the
        // device (a graph capture records kernel launches only).
        auto device_seq = [&](bool upload, bool eager) {
            Check(LaunchFill32(st, reinterpret_cast<uint32_t*>(dstats), 8, 0), "fill");
            Check(LaunchRotate(st, qsrc, nq, static_cast<uint32_t>(v.dim), v, rotq, true), "rotate");
            Check(LaunchQueryPrep(st, rotq, nq, v, lut, consts), "query prep");
            Check(LaunchRouteProbes(st, rotq, consts, nq, v, dist, npq, nprobe_all, maxp, pids, pd2), "route");
            if (eager && !mfma) Check(hipEventRecord(ctx->ev[1], st), "event");
            if (!listed) {
                Check(LaunchPlainSearch(st, v, pa, dparams, upload), "plain search");
                return;
            }
            if (mfma) {
                // The bound: the threshold pass over (the first 16384 vectors
                // of) the first probe, plus the rounding margin; for k <= 32
                // every scored tile lowers it further.
                Check(LaunchPlainSearch(st, v, fa, dparams, upload), "first-probe search");
                Check(LaunchMfmaQuery(st, v, rotq, nq, static_cast<uint32_t>(a->k), first_dists, pids, maxp,
                                      x->mfma.g1max.as<float>(), x->mfma.lmax, m_qexp, m_thr),
                      "mfma query");
                if (upload) {
                    const MfmaParams hm{{v, ba}, mx};
                    Check(hipMemcpyAsync(mparams, &hm, sizeof(hm), hipMemcpyHostToDevice, st), "upload parameters");
                }
                Check(LaunchFill32(st, ba.ccount, v.nlist, 0), "fill");
                Check(LaunchFill32(st, ba.ncand, nq, 0), "fill");
                Check(LaunchBatchedWorkList(st, v, ba, bparams, false), "work list");
                if (eager) Check(hipEventRecord(ctx->ev[1], st), "event");   // the contraction alone
                Check(LaunchMfmaScan(st, mparams, shape.grid), "mfma scan");
                if (eager) Check(hipEventRecord(ctx->ev[3], st), "event");
                Check(LaunchMfmaSelect(st, mparams, nq, static_cast<uint32_t>(a->k)), "mfma select");
                return;
            }
            Check(LaunchPlainSearch(st, v, fa, dparams, upload), "first-probe search");
            if (eager && !cands.empty()) tune();
            batched_tail(ba, shape.grid, upload);
        };
        // Graph replay: the key is everything the captured launches bake in.
        e8q_ctx::GraphEntry* ge = nullptr;
        if (ctx->graphs && cands.empty()) {
            static_assert(sizeof(IndexView) % 8 == 0, "IndexView is compared as 64-bit words");
            std::vector<uint64_t> key{reinterpret_cast<uint64_t>(x),
                                      static_cast<uint64_t>(a->mode),
                                      nq,
                                      a->k,
                                      maxp,
                                      nprobe_all,
                                      npq != nullptr ? 1u : 0u,
                                      FloatBits(a->z),
                                      static_cast<uint64_t>(pa.disable_prune),
                                      static_cast<uint64_t>(flush),
                                      reinterpret_cast<uint64_t>(qsrc),
                                      reinterpret_cast<uint64_t>(oids),
                                      reinterpret_cast<uint64_t>(odist),
                                      reinterpret_cast<uint64_t>(ctx->ws.p),
                                      ctx->ws.bytes,
                                      shape.qtile,
                                      shape.btile,
                                      shape.grid,
                                      ba.items_cap,
                                      pairs};
            const size_t w0 = key.size();
            key.resize(w0 + sizeof(IndexView) / 8);
            std::memcpy(key.data() + w0, &v, sizeof(IndexView));
            ge = ctx->FindGraph(key);
            if (!ge) ctx->AddGraph(key);   // first call of this shape runs eagerly
        }
        if (!a->queries_on_device)
            Check(hipMemcpyAsync(dq, a->queries, nq * v.dim * 4, hipMemcpyHostToDevice, st), "upload queries");
        if (npq) Check(hipMemcpyAsync(npq, a->nprobe_per_query, nq * 4, hipMemcpyHostToDevice, st), "upload nprobe");
        if (ge) {
            const PlainParams hp{v, listed ? fa : pa};
            Check(hipMemcpyAsync(dparams, &hp, sizeof(hp), hipMemcpyHostToDevice, st), "upload parameters");
            if (batched) {
                const BatchedParams hb{v, ba};
                Check(hipMemcpyAsync(bparams, &hb, sizeof(hb), hipMemcpyHostToDevice, st), "upload parameters");
            } else if (mfma) {
                const MfmaParams hm{{v, ba}, mx};
                Check(hipMemcpyAsync(mparams, &hm, sizeof(hm), hipMemcpyHostToDevice, st), "upload parameters");
            }
            if (!ge->exec) {
                Check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
                try {
                    device_seq(false, false);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(st, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                hipGraph_t g = nullptr;
                Check(hipStreamEndCapture(st, &g), "end capture");
                hipGraphExec_t ex = nullptr;
                const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                Check(e, "graph instantiate");
                ge->exec = ex;
                ++ctx->captures;
            }
            Check(hipEventRecord(ctx->ev[0], st), "event");
            Check(hipEventRecord(ctx->ev[1], st), "event");
            Check(hipGraphLaunch(ge->exec, st), "graph launch");
            ++ctx->replays;
        } else {
            Check(hipEventRecord(ctx->ev[0], st), "event");
            device_seq(true, true);
        }
        if (listed) {
            ctx->last_qtile = shape.qtile;
            ctx->last_btile = shape.btile;
            ctx->last_grid = shape.grid;
        }
        Check(hipEventRecord(ctx->ev[2], st), "event");
        if (batched) {   // queries whose candidate lists overflowed: the plain search
            uint32_t novf = 0;
            Check(hipMemcpyAsync(&novf, ba.counters + 2, 4, hipMemcpyDeviceToHost, st), "download counters");
            Check(hipStreamSynchronize(st), "batched search");
            if (novf > 0) {
                PlainArgs oa = pa;
                oa.qsel = ba.overflow;
                oa.nsel = novf;
                oa.stats = nullptr;   // their candidates and batches are counted already
                Check(LaunchPlainSearch(st, v, oa, dparams), "plain search of overflowed queries");
            }
        }
        DBuf r_cand, r_slot, r_mask, r_sel;   // the MFMA form's second pass (held until the final sync)
        uint64_t extra_survivors = 0;
        if (mfma) {   // queries with more candidates than a list holds: again, with lists of their length
            uint32_t novf = 0;
            Check(hipMemcpyAsync(&novf, ba.counters + 2, 4, hipMemcpyDeviceToHost, st), "download counters");
            Check(hipStreamSynchronize(st), "mfma search");
            if (novf > nq) Throw(E8Q_ERUNTIME, "mfma search: overflow count exceeds the query count");
            if (novf > 0) {
                std::vector<uint32_t> ovf(novf), cnt(nq);
                Check(hipMemcpyAsync(ovf.data(), ba.overflow, novf * 4ull, hipMemcpyDeviceToHost, st), "download");
                Check(hipMemcpyAsync(cnt.data(), ba.ncand, nq * 4, hipMemcpyDeviceToHost, st), "download");
                Check(hipStreamSynchronize(st), "download");
                uint32_t cap2 = 0;
                std::vector<uint32_t> slot(nq, 0xFFFFFFFFu);
                std::vector<uint8_t> mask(nq, 0);
                for (uint32_t j = 0; j < novf; ++j) {
                    if (ovf[j] >= nq) Throw(E8Q_ERUNTIME, "mfma search: bad overflow entry");
                    extra_survivors += cnt[ovf[j]];
                    cap2 = std::max(cap2, cnt[ovf[j]]);
                    slot[ovf[j]] = j;
                    mask[ovf[j]] = 1;
                }
                r_cand.alloc(static_cast<uint64_t>(novf) * cap2 * 8);
                Upload(r_slot, slot.data(), nq, st);
                Upload(r_mask, mask.data(), nq, st);
                Upload(r_sel, ovf.data(), novf, st);
                BatchedArgs b2 = ba;
                b2.qmask = r_mask.as<uint8_t>();
                b2.stats = tstats;   // counted already
                MfmaExtra m2 = mx;
                m2.cand = r_cand.as<uint64_t>();
                m2.cap = cap2;
                m2.qslot = r_slot.as<uint32_t>();
                m2.qsel = r_sel.as<uint32_t>();
                m2.nsel = novf;
                const MfmaParams hm{{v, b2}, m2};
                Check(hipMemcpyAsync(mparams, &hm, sizeof(hm), hipMemcpyHostToDevice, st), "upload parameters");
                Check(LaunchFill32(st, ba.ccount, v.nlist, 0), "fill");
                Check(LaunchFill32(st, ba.ncand, nq, 0), "fill");
                Check(LaunchBatchedWorkList(st, v, b2, bparams, false), "work list");
                Check(LaunchMfmaScan(st, mparams, shape.grid), "mfma scan");
                Check(LaunchMfmaSelect(st, mparams, novf, static_cast<uint32_t>(a->k)), "mfma select");
            }
        }
        if (!a->outputs_on_device) {
            Check(hipMemcpyAsync(a->out_ids, oids, nq * a->k * 4, hipMemcpyDeviceToHost, st), "download ids");
            Check(hipMemcpyAsync(a->out_dists, odist, nq * a->k * 4, hipMemcpyDeviceToHost, st), "download dists");
        }
        unsigned long long hs[4] = {0, 0, 0, 0};
        Check(hipMemcpyAsync(hs, dstats, 32, hipMemcpyDeviceToHost, st), "download stats");
        Check(hipStreamSynchronize(st), "search");
        if (stats) {
            float ms_scan = 0.0F, ms_all = 0.0F;
            const bool contraction = mfma && !ge;   // eager MFMA form: the contraction kernel alone
            Check(hipEventElapsedTime(&ms_scan, ctx->ev[1], ctx->ev[contraction ? 3 : 2]), "event time");
            Check(hipEventElapsedTime(&ms_all, ctx->ev[0], ctx->ev[2]), "event time");
            stats->candidates = hs[0];
            stats->survivors = hs[1] + extra_survivors;
            stats->batches = hs[2];
            stats->scanned = hs[3];
            stats->scan_ms = ms_scan;
            stats->total_ms = ms_all;
        }
    });
}

// ---- the allocation-free plain search --------------------------------------
// Every per-call buffer of the plain form (staged queries, rotated queries,
// LUTs, LUT constants, route distances, probe lists, staged outputs, the
// counters and the kernel parameter block) is carved from `scratch`, one
// fixed array per GPU, instead of the context workspace.  In host code the
// array's name is only a stand-in for the symbol: kernels and copies get the
// device address hipGetSymbolAddress returns, resolved once per device.
static __device__ __align__(256) unsigned char scratch[E8Q_STATIC_SCRATCH_BYTES];

namespace {

constexpr int kStaticDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_dev[kStaticDevices] = {};

// The device address of `scratch` on `device`, which must be current.
char* StaticScratch(int device) {
    if (device < 0 || device >= kStaticDevices) Throw(E8Q_EINVAL, "device id out of range for the static scratch");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* ptr = nullptr;
        Check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) Throw(E8Q_ERUNTIME, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_dev[device] = static_cast<char*>(ptr);
    }
    return g_scratch_dev[device];
}

}  // namespace

e8q_status e8q_search_static(e8q_ctx* ctx, const e8q_dindex* x, const e8q_search_args* a, e8q_search_stats* stats) {
    return Guard([&] {
        if (!ctx || !x || !a) Throw(E8Q_EINVAL, "null argument");
        if (a->k == 0 || a->k > static_cast<uint64_t>(kMaxK))
            Throw(E8Q_EINVAL, "k must be in [1, " + std::to_string(kMaxK) + "]");
        if (!a->queries || !a->out_ids || !a->out_dists) Throw(E8Q_EINVAL, "null buffer");
        if (a->mode != E8Q_MODE_PLAIN) Throw(E8Q_EINVAL, "the allocation-free search runs the plain form only");
        const int flush = a->flush_every > 0 ? a->flush_every : 32;
        if (flush > kMaxFlush)
            Throw(E8Q_EINVAL, "flush_every must be <= 1024 (the survivor staging capacity, as in the reference)");
        const IndexView& v = x->view;
        const uint64_t nq = a->nq;
        if (stats) *stats = e8q_search_stats{0, 0, 0, 0.0, 0.0, 0};
        if (nq == 0) return;
        uint32_t maxp = 0;
        if (a->nprobe_per_query) {
            for (uint64_t i = 0; i < nq; ++i) {
                if (a->nprobe_per_query[i] == 0) Throw(E8Q_EINVAL, "nprobe must be positive");
                maxp = std::max<uint32_t>(maxp, a->nprobe_per_query[i]);
            }
        } else {
            if (a->nprobe == 0) Throw(E8Q_EINVAL, "nprobe must be positive");
            maxp = static_cast<uint32_t>(std::min<uint64_t>(a->nprobe, v.nlist));
        }
        maxp = static_cast<uint32_t>(std::min<uint64_t>(maxp, v.nlist));
        if (SelectProbesLds(static_cast<uint32_t>(v.nlist), maxp) > 160 * 1024)
            Throw(E8Q_EINVAL, "nlist too large for the probe selector");
        const uint64_t P = v.padded, E = static_cast<uint64_t>(v.blocks) * 32;
        float *dq = nullptr, *rotq = nullptr, *consts = nullptr, *dist = nullptr, *pd2 = nullptr, *odist = nullptr;
        uint8_t* lut = nullptr;
        uint32_t *pids = nullptr, *npq = nullptr, *oids = nullptr;
        unsigned long long* dstats = nullptr;
        PlainParams* dparams = nullptr;
        e8q_ctx::Carve cv;   // the same buffers e8q_search carves for the plain form
        if (!a->queries_on_device) cv.add(&dq, nq * v.dim);
        cv.add(&rotq, nq * P);
        cv.add(&lut, nq * E);
        cv.add(&consts, nq * 8);
        cv.add(&dist, nq * v.nlist);
        cv.add(&pids, nq * maxp);
        cv.add(&pd2, nq * maxp);
        if (a->nprobe_per_query) cv.add(&npq, nq);
        if (!a->outputs_on_device) {
            cv.add(&oids, nq * a->k);
            cv.add(&odist, nq * a->k);
        }
        cv.add(&dstats, 4);
        cv.add(&dparams, 1);
        const uint64_t need = cv.bytes();
        if (need > E8Q_STATIC_SCRATCH_BYTES)
            Throw(E8Q_EINVAL, "this search needs " + std::to_string(need) + " bytes of scratch; the allocation-free "
                              "entry point holds " + std::to_string(static_cast<uint64_t>(E8Q_STATIC_SCRATCH_BYTES)) +
                              " (search fewer queries per call)");
        ctx->Bind();
        char* base = StaticScratch(ctx->device);
        size_t off = 0;
        for (auto& w : cv.want) {
            *w.first = base + off;
            off += (w.second + 255) & ~size_t{255};
        }
        hipStream_t st = ctx->stream;
        const float* qsrc = a->queries_on_device ? a->queries : dq;
        if (a->outputs_on_device) {
            oids = a->out_ids;
            odist = a->out_dists;
        }
        PlainArgs pa{};
        pa.nq = nq;
        pa.rotq = rotq;
        pa.lut = lut;
        pa.consts = consts;
        pa.probe_ids = pids;
        pa.probe_d2 = pd2;
        pa.maxp = maxp;
        pa.nprobe_q = npq;
        pa.nprobe_all = maxp;
        pa.k = a->k;
        pa.z = a->z;
        pa.disable_prune = a->disable_prune ? 1 : 0;
        pa.flush_every = flush;
        pa.out_ids = oids;
        pa.out_dists = odist;
        pa.stats = dstats;
        if (!a->queries_on_device)
            Check(hipMemcpyAsync(dq, a->queries, nq * v.dim * 4, hipMemcpyHostToDevice, st), "upload queries");
        if (npq) Check(hipMemcpyAsync(npq, a->nprobe_per_query, nq * 4, hipMemcpyHostToDevice, st), "upload nprobe");
        Check(hipEventRecord(ctx->ev[0], st), "event");
        // The counters start at zero on every call: a fill kernel, since the
        // array keeps the previous call's contents.
        Check(LaunchFill32(st, reinterpret_cast<uint32_t*>(dstats), 8, 0), "fill");
        Check(LaunchRotate(st, qsrc, nq, static_cast<uint32_t>(v.dim), v, rotq, true), "rotate");
        Check(LaunchQueryPrep(st, rotq, nq, v, lut, consts), "query prep");
        Check(LaunchRouteProbes(st, rotq, consts, nq, v, dist, npq, maxp, maxp, pids, pd2), "route");
        Check(hipEventRecord(ctx->ev[1], st), "event");
        Check(LaunchPlainSearch(st, v, pa, dparams, true), "plain search");
        Check(hipEventRecord(ctx->ev[2], st), "event");
        if (!a->outputs_on_device) {
            Check(hipMemcpyAsync(a->out_ids, oids, nq * a->k * 4, hipMemcpyDeviceToHost, st), "download ids");
            Check(hipMemcpyAsync(a->out_dists, odist, nq * a->k * 4, hipMemcpyDeviceToHost, st), "download dists");
        }
        unsigned long long hs[4] = {0, 0, 0, 0};
        Check(hipMemcpyAsync(hs, dstats, 32, hipMemcpyDeviceToHost, st), "download stats");
        Check(hipStreamSynchronize(st), "search");
        if (stats) {
            float ms_scan = 0.0F, ms_all = 0.0F;
            Check(hipEventElapsedTime(&ms_scan, ctx->ev[1], ctx->ev[2]), "event time");
            Check(hipEventElapsedTime(&ms_all, ctx->ev[0], ctx->ev[2]), "event time");
            stats->candidates = hs[0];
            stats->survivors = hs[1];
            stats->batches = hs[2];
            stats->scanned = hs[3];
            stats->scan_ms = ms_scan;
            stats->total_ms = ms_all;
        }
    });
}

e8q_status e8q_debug_coarse_accumulate(e8q_ctx* ctx, const uint8_t* rowa, uint64_t nbatch, const uint8_t* lut,
                                       uint64_t blocks, int32_t* out) {
    return Guard([&] {
        if (!ctx || !rowa || !lut || !out) Throw(E8Q_EINVAL, "null argument");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        DBuf dr, dl, dout;
        Upload(dr, rowa, nbatch * blocks * 32, st);
        Upload(dl, lut, blocks * 64, st);
        dout.alloc(nbatch * 32 * 4);
        Check(LaunchCoarseAccumulate(st, dr.as<uint8_t>(), nbatch, dl.as<uint8_t>(), static_cast<int>(blocks),
                                     dout.as<int32_t>()),
              "coarse accumulate");
        Check(hipMemcpyAsync(out, dout.p, nbatch * 32 * 4, hipMemcpyDeviceToHost, st), "download");
        Check(hipStreamSynchronize(st), "coarse accumulate");
    });
}

e8q_status e8q_debug_query_prep(e8q_ctx* ctx, const e8q_dindex* x, const float* queries, uint64_t nq, uint64_t nprobe,
                                float* rot_q, uint8_t* lut, float* consts, uint32_t* probe, float* cdist) {
    return Guard([&] {
        if (!ctx || !x || !queries) Throw(E8Q_EINVAL, "null argument");
        const IndexView& v = x->view;
        if (nprobe == 0) Throw(E8Q_EINVAL, "nprobe must be positive");
        ctx->Bind();
        hipStream_t st = ctx->stream;
        const uint64_t P = v.padded, E = static_cast<uint64_t>(v.blocks) * 32;
        const uint32_t maxp = static_cast<uint32_t>(std::min<uint64_t>(nprobe, v.nlist));
        DBuf dq, drq, dlut, dc, dd, dp, dp2;
        Upload(dq, queries, nq * v.dim, st);
        drq.alloc(nq * P * 4);
        dlut.alloc(nq * E);
        dc.alloc(nq * 8 * 4);
        dd.alloc(nq * v.nlist * 4);
        dp.alloc(nq * maxp * 4);
        dp2.alloc(nq * maxp * 4);
        Check(LaunchRotate(st, dq.as<float>(), nq, static_cast<uint32_t>(v.dim), v, drq.as<float>(), true), "rotate");
        Check(LaunchQueryPrep(st, drq.as<float>(), nq, v, dlut.as<uint8_t>(), dc.as<float>()), "prep");
        Check(LaunchRoute(st, drq.as<float>(), dc.as<float>(), nq, v, dd.as<float>()), "route");
        Check(LaunchSelectProbes(st, dd.as<float>(), nq, static_cast<uint32_t>(v.nlist), nullptr, maxp, maxp,
                                 dp.as<uint32_t>(), dp2.as<float>()),
              "select");
        std::vector<uint8_t> hl(nq * E);
        std::vector<float> hc(nq * 8);
        if (rot_q) Check(hipMemcpyAsync(rot_q, drq.p, nq * P * 4, hipMemcpyDeviceToHost, st), "dl");
        Check(hipMemcpyAsync(hl.data(), dlut.p, nq * E, hipMemcpyDeviceToHost, st), "dl");
        Check(hipMemcpyAsync(hc.data(), dc.p, nq * 32, hipMemcpyDeviceToHost, st), "dl");
        if (probe) Check(hipMemcpyAsync(probe, dp.p, nq * maxp * 4, hipMemcpyDeviceToHost, st), "dl");
        if (cdist) Check(hipMemcpyAsync(cdist, dd.p, nq * v.nlist * 4, hipMemcpyDeviceToHost, st), "dl");
        Check(hipStreamSynchronize(st), "query prep");
        // Expand the compact LUT to the reference's 64-byte-per-block layout.
        if (lut)
            for (uint64_t q = 0; q < nq; ++q)
                for (int b = 0; b < v.blocks; ++b) {
                    const uint8_t* s = &hl[q * E + b * 32];
                    uint8_t* o = lut + q * v.blocks * 64 + b * 64;
                    std::memcpy(o, s, 16);
                    std::memcpy(o + 16, s, 16);
                    std::memcpy(o + 32, s + 16, 16);
                    std::memcpy(o + 48, s + 16, 16);
                }
        if (consts)
            for (uint64_t q = 0; q < nq; ++q) std::memcpy(consts + q * 4, &hc[q * 8], 16);
    });
}

e8q_status e8q_debug_rotate(e8q_ctx* ctx, const e8q_dindex* x, const float* in, uint64_t nrows, float* out) {
    return Guard([&] {
        if (!ctx || !x || !in || !out) Throw(E8Q_EINVAL, "null argument");
        const IndexView& v = x->view;
        ctx->Bind();
        hipStream_t st = ctx->stream;
        DBuf di, dout;
        Upload(di, in, nrows * v.dim, st);
        dout.alloc(nrows * v.padded * 4);
        Check(LaunchRotate(st, di.as<float>(), nrows, static_cast<uint32_t>(v.dim), v, dout.as<float>(), false), "rotate");
        Check(hipMemcpyAsync(out, dout.p, nrows * v.padded * 4, hipMemcpyDeviceToHost, st), "dl");
        Check(hipStreamSynchronize(st), "rotate");
    });
}

e8q_status e8q_debug_encode(e8q_ctx* ctx, const e8q_dindex* x, const float* unit, uint64_t nrows, double s,
                            uint16_t* digits, uint8_t* branch) {
    return Guard([&] {
        if (!ctx || !x || !unit || !digits || !branch) Throw(E8Q_EINVAL, "null argument");
        const IndexView& v = x->view;
        ctx->Bind();
        hipStream_t st = ctx->stream;
        DBuf du, dd, db;
        Upload(du, unit, nrows * v.padded, st);
        dd.alloc(nrows * v.padded * 2);
        db.alloc(nrows * v.blocks);
        Check(LaunchEncodeUnits(st, v, x->ev, du.as<float>(), nrows, s, dd.as<uint16_t>(), db.as<uint8_t>()), "encode");
        Check(hipMemcpyAsync(digits, dd.p, nrows * v.padded * 2, hipMemcpyDeviceToHost, st), "dl");
        Check(hipMemcpyAsync(branch, db.p, nrows * v.blocks, hipMemcpyDeviceToHost, st), "dl");
        Check(hipStreamSynchronize(st), "encode");
    });
}

// ---- diagnostics (e8q_oracle.cpp:14-219) ------------------------------

namespace {

uint64_t DiagChunk(uint64_t nq, uint64_t per_query) {
    const uint64_t lim = uint64_t{1} << 28;   // candidates held at once (1 GiB of distances)
    const uint64_t c = std::min<uint64_t>({nq, lim / std::max<uint64_t>(per_query, 1), 65535});   // grid y
    return std::max<uint64_t>(1, c);
}

}  // namespace

e8q_status e8q_estimate(e8q_ctx* ctx, const e8q_dindex* x, const float* queries, uint64_t nq, const uint32_t* qidx,
                        const uint32_t* ids, uint64_t npair, float* out) {
    return Guard([&] {
        if (!ctx || !x) Throw(E8Q_EINVAL, "null argument");
        if (npair && (!queries || !qidx || !ids || !out)) Throw(E8Q_EINVAL, "null buffer");
        const IndexView& v = x->view;
        if (v.n == 0) Throw(E8Q_ERUNTIME, "Estimate: empty index");
        for (uint64_t i = 0; i < npair; ++i)
            if (qidx[i] >= nq) Throw(E8Q_EINVAL, "Estimate: query index out of range");
        if (npair == 0) return;
        ctx->Bind();
        hipStream_t st = ctx->stream;
        float *dq = nullptr, *rotq = nullptr, *dout = nullptr;
        uint32_t *dqi = nullptr, *dids = nullptr, *inv = nullptr, *bad = nullptr;
        e8q_ctx::Carve cv;
        cv.add(&dq, nq * v.dim);
        cv.add(&rotq, nq * v.padded);
        cv.add(&dqi, npair);
        cv.add(&dids, npair);
        cv.add(&inv, v.n);
        cv.add(&dout, npair);
        cv.add(&bad, 1);
        ctx->Take(cv);
        Check(hipMemcpyAsync(dq, queries, nq * v.dim * 4, hipMemcpyHostToDevice, st), "upload queries");
        Check(hipMemcpyAsync(dqi, qidx, npair * 4, hipMemcpyHostToDevice, st), "upload pairs");
        Check(hipMemcpyAsync(dids, ids, npair * 4, hipMemcpyHostToDevice, st), "upload pairs");
        Check(LaunchFill32(st, inv, v.n, 0), "clear");
        Check(LaunchFill32(st, bad, 1, 0), "clear");
        Check(LaunchDiagInverseIds(st, v.ids, v.n, inv), "id map");
        Check(LaunchRotate(st, dq, nq, static_cast<uint32_t>(v.dim), v, rotq, true), "rotate");
        Check(LaunchE8Estimate(st, v, rotq, dqi, dids, npair, inv, dout, bad), "estimate");
        uint32_t hbad = 0;
        Check(hipMemcpyAsync(out, dout, npair * 4, hipMemcpyDeviceToHost, st), "download");
        Check(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st), "download");
        Check(hipStreamSynchronize(st), "estimate");
        if (hbad) Throw(E8Q_EINVAL, "Estimate: id out of range");
    });
}

e8q_status e8q_probe_bound_w(e8q_ctx* ctx, const e8q_dindex* x, const float* queries, uint64_t nq,
                             const float* base_rows, const uint32_t* qidx, const uint32_t* ids, uint64_t npair,
                             double* out) {
    return Guard([&] {
        if (!ctx || !x) Throw(E8Q_EINVAL, "null argument");
        if (npair && (!queries || !base_rows || !qidx || !ids || !out)) Throw(E8Q_EINVAL, "null buffer");
        const IndexView& v = x->view;
        if (v.n == 0) Throw(E8Q_ERUNTIME, "ProbeBoundW: empty index");
        if (v.bits >= 10)
            Throw(E8Q_EINVAL, "ProbeBoundW: b >= 10 is undefined in the reference (its 8-byte digit staging overflows)");
        if (npair == 0) return;
        ctx->Bind();
        hipStream_t st = ctx->stream;
        const uint64_t nqa = std::max<uint64_t>(nq, 1);
        float *dq = nullptr, *rotq = nullptr, *db = nullptr, *rotx = nullptr;
        double* dout = nullptr;
        uint32_t *dqi = nullptr, *dids = nullptr, *inv = nullptr, *bad = nullptr;
        e8q_ctx::Carve cv;
        cv.add(&dq, nqa * v.dim);
        cv.add(&rotq, nqa * v.padded);
        cv.add(&db, npair * v.dim);
        cv.add(&rotx, npair * v.padded);
        cv.add(&dqi, npair);
        cv.add(&dids, npair);
        cv.add(&inv, v.n);
        cv.add(&dout, npair);
        cv.add(&bad, 1);
        ctx->Take(cv);
        if (nq) Check(hipMemcpyAsync(dq, queries, nq * v.dim * 4, hipMemcpyHostToDevice, st), "upload queries");
        Check(hipMemcpyAsync(db, base_rows, npair * v.dim * 4, hipMemcpyHostToDevice, st), "upload base rows");
        Check(hipMemcpyAsync(dqi, qidx, npair * 4, hipMemcpyHostToDevice, st), "upload pairs");
        Check(hipMemcpyAsync(dids, ids, npair * 4, hipMemcpyHostToDevice, st), "upload pairs");
        Check(LaunchFill32(st, inv, v.n, 0), "clear");
        Check(LaunchFill32(st, bad, 1, 0), "clear");
        Check(LaunchDiagInverseIds(st, v.ids, v.n, inv), "id map");
        Check(LaunchRotate(st, dq, nq, static_cast<uint32_t>(v.dim), v, rotq, true), "rotate");
        Check(LaunchRotate(st, db, npair, static_cast<uint32_t>(v.dim), v, rotx, true), "rotate");
        Check(LaunchE8ProbeBoundW(st, v, x->levels.as<double>(), rotq, rotx, dqi, dids, nq, npair, inv, dout, bad),
              "probe bound");
        uint32_t hbad = 0;
        Check(hipMemcpyAsync(out, dout, npair * 8, hipMemcpyDeviceToHost, st), "download");
        Check(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st), "download");
        Check(hipStreamSynchronize(st), "probe bound");
        if (hbad) Throw(E8Q_EINVAL, "ProbeBoundW: id out of range");
    });
}

e8q_status e8q_exhaustive_topk(e8q_ctx* ctx, const e8q_dindex* x, const float* queries, uint64_t nq, uint64_t k,
                               uint32_t* out_ids, float* out_dists) {
    return Guard([&] {
        if (!ctx || !x) Throw(E8Q_EINVAL, "null argument");
        if (k == 0 || k > kDiagMaxK) Throw(E8Q_EINVAL, "k must be in [1, 1024]");
        if (nq && (!queries || !out_ids || !out_dists)) Throw(E8Q_EINVAL, "null buffer");
        if (nq == 0) return;
        const IndexView& v = x->view;
        ctx->Bind();
        hipStream_t st = ctx->stream;
        const uint64_t chunk = DiagChunk(nq, v.n);
        float *dq = nullptr, *rotq = nullptr, *d2c = nullptr, *dist = nullptr, *od = nullptr;
        uint32_t* oid = nullptr;
        e8q_ctx::Carve cv;
        cv.add(&dq, nq * v.dim);
        cv.add(&rotq, nq * v.padded);
        cv.add(&d2c, nq * v.nlist);
        cv.add(&dist, std::max<uint64_t>(chunk * v.n, 1));
        cv.add(&oid, nq * k);
        cv.add(&od, nq * k);
        ctx->Take(cv);
        Check(hipMemcpyAsync(dq, queries, nq * v.dim * 4, hipMemcpyHostToDevice, st), "upload queries");
        Check(LaunchRotate(st, dq, nq, static_cast<uint32_t>(v.dim), v, rotq, true), "rotate");
        Check(LaunchE8ListD2(st, v, rotq, nq, d2c), "list distances");
        for (uint64_t q0 = 0; q0 < nq; q0 += chunk) {
            const uint64_t nqc = std::min(chunk, nq - q0);
            Check(LaunchE8Exhaustive(st, v, rotq, d2c, q0, nqc, dist), "exhaustive distances");
            Check(LaunchDiagSelect(st, false, dist, v.ids, 0, nullptr, v.n, v.n, q0, nqc, static_cast<uint32_t>(k),
                                   oid, od),
                  "top-k");
        }
        Check(hipMemcpyAsync(out_ids, oid, nq * k * 4, hipMemcpyDeviceToHost, st), "download");
        Check(hipMemcpyAsync(out_dists, od, nq * k * 4, hipMemcpyDeviceToHost, st), "download");
        Check(hipStreamSynchronize(st), "exhaustive top-k");
    });
}

}  // extern "C"
