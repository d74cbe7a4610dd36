// e8q_index.cpp -- host side of e8qlib::E8qIndex over the device C ABI.
//
// Build (e8q_build.cpp:87-182): shape, rotator signs, list counts and the
// scan / refine arena offsets, level tables, then the device encode.
// Save / Load: the reference's E8AX file byte for byte (e8q_io.cpp:35-211),
// including the EPOQ options trailer and Load's consistency checks.
#include "e8q_index.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>

#include "e8q_gpu.h"
#include "me8q_gpu.h"
#include "e8q_host_math.h"

namespace e8qlib {
namespace {

constexpr uint32_t kMagic = 0x58413845;     // "E8AX"
constexpr uint32_t kOptMagic = 0x514F5045;  // "EPOQ"

void Check(e8q_status st) {
    if (st == E8Q_OK) return;
    const std::string msg = e8q_last_error();
    if (st == E8Q_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error("e8q device: " + msg);
}

size_t BatchBytes(size_t blocks) { return (128 + 32 * blocks + 63) & ~size_t{63}; }
size_t RefBatchBytes(size_t blocks, int bits) {
    const size_t stride = bits >= 2 ? blocks * static_cast<size_t>(bits - 1) : 0;
    return (32 * stride + 256 + 63) & ~size_t{63};
}

}  // namespace

std::mutex& DeviceMutex() {
    static std::mutex m;
    return m;
}

e8q_ctx* DefaultContext() {
    static e8q_ctx* ctx = [] {
        const char* e = std::getenv("E8Q_DEVICE");
        e8q_ctx* c = nullptr;
        Check(e8q_ctx_create(e != nullptr ? std::atoi(e) : 0, &c));
        return c;
    }();
    return ctx;
}

void SetSearchGraphs(bool on) {
    std::lock_guard<std::mutex> lk(DeviceMutex());
    Check(e8q_ctx_set_graphs(DefaultContext(), on ? 1 : 0));
}

void SetBatchedShape(uint32_t qtile, uint32_t btile, uint32_t grid) {
    std::lock_guard<std::mutex> lk(DeviceMutex());
    Check(e8q_ctx_set_batched_shape(DefaultContext(), qtile, btile, grid));
}

ContextInfo GetContextInfo() {
    std::lock_guard<std::mutex> lk(DeviceMutex());
    e8q_ctx_info i{};
    Check(e8q_ctx_get_info(DefaultContext(), &i));
    ContextInfo o;
    o.qtile = i.qtile;
    o.btile = i.btile;
    o.grid = i.grid;
    o.graphs = i.graphs != 0;
    o.tunings = i.tunings;
    o.captures = i.captures;
    o.replays = i.replays;
    e8q_m8_ctx_info m{};
    Check(e8q_ctx_get_m8_info(DefaultContext(), &m));
    o.m8_mfma_grid = m.mfma_grid;
    o.m8_refine_grid = m.refine_grid;
    o.m8_tunings = m.tunings;
    return o;
}

void SetMe8qGrids(uint32_t mfma_grid, uint32_t refine_grid) {
    std::lock_guard<std::mutex> lk(DeviceMutex());
    Check(e8q_ctx_set_m8_grids(DefaultContext(), mfma_grid, refine_grid));
}

E8qIndex::E8qIndex() = default;

E8qIndex::~E8qIndex() { Reset(); }

E8qIndex::E8qIndex(E8qIndex&& o) noexcept { *this = std::move(o); }

E8qIndex& E8qIndex::operator=(E8qIndex&& o) noexcept {
    if (this != &o) {
        Reset();
        n_ = o.n_; dim_ = o.dim_; padded_ = o.padded_; nlist_ = o.nlist_; bits_ = o.bits_;
        alpha_ = o.alpha_; disable_prune_ = o.disable_prune_;
        signs_ = std::move(o.signs_);
        cluster_begin_ = std::move(o.cluster_begin_);
        arena_begin_ = std::move(o.arena_begin_);
        ref_arena_begin_ = std::move(o.ref_arena_begin_);
        ctx_ = o.ctx_;
        dev_ = o.dev_;
        o.dev_ = nullptr;
    }
    return *this;
}

void E8qIndex::Reset() {
    if (dev_ != nullptr) {
        std::lock_guard<std::mutex> lk(DeviceMutex());
        (void)e8q_index_destroy(dev_);
        dev_ = nullptr;
    }
}

void E8qIndex::CreateDevice(const std::vector<uint8_t>& signs) {
    const std::vector<float> tab = hostmath::RefineTables(bits_, padded_);
    double gate[16];
    hostmath::GateMid(bits_, gate);
    std::vector<double> levels = hostmath::BranchLevels(bits_, 0);
    const std::vector<double> l1 = hostmath::BranchLevels(bits_, 1);
    levels.insert(levels.end(), l1.begin(), l1.end());
    e8q_index_desc d{};
    d.n = n_;
    d.dim = dim_;
    d.padded = padded_;
    d.nlist = nlist_;
    d.bits = bits_;
    d.alpha = alpha_;
    d.scale = hostmath::FixedScale(bits_);
    d.signs = signs.data();
    d.refine_tab = tab.data();
    d.gate_mid = gate;
    d.levels = levels.data();
    d.cluster_begin = cluster_begin_.data();
    d.arena_begin = arena_begin_.data();
    d.ref_arena_begin = ref_arena_begin_.data();
    ctx_ = DefaultContext();
    std::lock_guard<std::mutex> lk(DeviceMutex());
    Check(e8q_index_create(ctx_, &d, &dev_));
}

namespace {

// Shape and offsets shared by both Build forms; returns per-row slots.
std::vector<uint32_t> Layout(const uint32_t* cluster_ids, size_t n, size_t nlist, size_t blocks, int bits,
                             std::vector<uint64_t>& cb, std::vector<uint64_t>& ab, std::vector<uint64_t>& rb) {
    std::vector<uint64_t> counts(nlist, 0);
    for (size_t i = 0; i < n; ++i) {
        if (cluster_ids[i] >= nlist) throw std::invalid_argument("cluster id out of range");
        ++counts[cluster_ids[i]];
    }
    cb.assign(nlist + 1, 0);
    ab.assign(nlist + 1, 0);
    rb.assign(nlist + 1, 0);
    const size_t bb = BatchBytes(blocks), rbb = RefBatchBytes(blocks, bits);
    for (size_t c = 0; c < nlist; ++c) {
        cb[c + 1] = cb[c] + counts[c];
        const uint64_t nb = (counts[c] + 31) / 32;
        ab[c + 1] = ab[c] + nb * bb;
        rb[c + 1] = rb[c] + nb * rbb;
    }
    std::vector<uint32_t> slot(n);
    std::vector<uint64_t> cur(cb.begin(), cb.end() - 1);
    for (size_t i = 0; i < n; ++i) slot[i] = static_cast<uint32_t>(cur[cluster_ids[i]]++);
    return slot;
}

}  // namespace

void E8qIndex::Build(const float* base, const float* centroids, const uint32_t* cluster_ids, size_t n, size_t dim,
                     size_t num_clusters, int bits, uint64_t seed, size_t threads, const E8qBuildOptions& options) {
    (void)threads;   // the device sizes its own parallelism
    if (bits < 1 || bits > hostmath::kMaxQuantizeBits) throw std::invalid_argument("E8Q bits must be in [1, 13]");
    if (n == 0 || dim == 0 || num_clusters == 0) throw std::invalid_argument("empty build set");
    if (n > std::numeric_limits<uint32_t>::max()) throw std::invalid_argument("n exceeds 2^32 ids");
    Reset();
    alpha_ = options.prune_alpha > 0.0F ? options.prune_alpha : 1.0F;
    n_ = n;
    dim_ = dim;
    nlist_ = num_clusters;
    bits_ = bits;
    padded_ = hostmath::PaddedDim(dim);
    const std::vector<uint32_t> slot =
        Layout(cluster_ids, n, num_clusters, padded_ / 8, bits, cluster_begin_, arena_begin_, ref_arena_begin_);
    signs_ = hostmath::RotatorSigns(dim, seed);
    CreateDevice(signs_);
    std::lock_guard<std::mutex> lk(DeviceMutex());
    Check(e8q_index_encode(ctx_, dev_, base, centroids, cluster_ids, slot.data(), 0));
}

void E8qIndex::BuildFromDevice(const float* base, const float* centroids, const uint32_t* cluster_ids_host,
                               const uint32_t* cluster_ids_device, size_t n, size_t dim, size_t num_clusters,
                               int bits, uint64_t seed, const E8qBuildOptions& options) {
    if (bits < 1 || bits > hostmath::kMaxQuantizeBits) throw std::invalid_argument("E8Q bits must be in [1, 13]");
    if (n == 0 || dim == 0 || num_clusters == 0) throw std::invalid_argument("empty build set");
    Reset();
    alpha_ = options.prune_alpha > 0.0F ? options.prune_alpha : 1.0F;
    n_ = n;
    dim_ = dim;
    nlist_ = num_clusters;
    bits_ = bits;
    padded_ = hostmath::PaddedDim(dim);
    const std::vector<uint32_t> slot =
        Layout(cluster_ids_host, n, num_clusters, padded_ / 8, bits, cluster_begin_, arena_begin_, ref_arena_begin_);
    signs_ = hostmath::RotatorSigns(dim, seed);
    CreateDevice(signs_);
    // The slots go up through the host path of the ABI; base, centroids and
    // ids stay in device memory.
    std::lock_guard<std::mutex> lk(DeviceMutex());
    Check(e8q_index_encode(ctx_, dev_, base, centroids, cluster_ids_device, slot.data(), 1));
}

size_t E8qIndex::memory_bytes() const {
    if (dev_ == nullptr) return 0;
    const size_t stride = std::max<size_t>(256, size_t{1} << bits_);
    return (arena_begin_.back() + 64) + (ref_arena_begin_.back() + 64) + n_ * 4 + nlist_ * padded_ * 4 +
           nlist_ * 4 + 3 * (nlist_ + 1) * 8 + padded_ * 4 + 2 * stride * 4;
}

size_t E8qIndex::device_bytes() const { return dev_ != nullptr ? e8q_index_device_bytes(dev_) : 0; }

void E8qIndex::Save(const char* path) const {
    if (dev_ == nullptr) throw std::runtime_error("E8qIndex must be built or loaded before save");
    std::vector<uint32_t> ids(n_);
    std::vector<float> mean(padded_), crot(nlist_ * padded_), cn2(nlist_);
    std::vector<uint8_t> arena(arena_begin_.back() + 64), ref(ref_arena_begin_.back() + 64);
    e8q_index_arrays h{ids.data(), mean.data(), crot.data(), cn2.data(), arena.data(), ref.data()};
    {
        std::lock_guard<std::mutex> lk(DeviceMutex());
        Check(e8q_index_download(ctx_, dev_, &h));
    }
    std::FILE* f = std::fopen(path, "wb");
    if (f == nullptr) throw std::runtime_error("cannot open index file for writing");
    bool ok = true;
    auto wr = [&](const void* p, size_t bytes) {
        if (bytes != 0 && std::fwrite(p, 1, bytes, f) != bytes) ok = false;
    };
    const uint32_t magic = kMagic;
    const uint64_t n = n_, dim = dim_, nl = nlist_, rot_bytes = signs_.size();
    const int32_t bits = bits_;
    wr(&magic, 4);
    wr(&n, 8);
    wr(&dim, 8);
    wr(&nl, 8);
    wr(&bits, 4);
    wr(&rot_bytes, 8);
    wr(ids.data(), ids.size() * 4);
    wr(signs_.data(), signs_.size());
    wr(mean.data(), mean.size() * 4);
    wr(crot.data(), crot.size() * 4);
    wr(cn2.data(), cn2.size() * 4);
    wr(cluster_begin_.data(), cluster_begin_.size() * 8);
    wr(arena_begin_.data(), arena_begin_.size() * 8);
    wr(arena.data(), arena.size());
    wr(ref_arena_begin_.data(), ref_arena_begin_.size() * 8);
    wr(ref.data(), ref.size());
    uint32_t lam = 0;
    std::memcpy(&lam, &alpha_, 4);
    const uint32_t om = kOptMagic;
    wr(&om, 4);
    wr(&lam, 4);
    wr(&om, 4);
    std::fclose(f);
    if (!ok) throw std::runtime_error("short write while saving index");
}

void E8qIndex::Load(const char* path) {
    std::FILE* f = std::fopen(path, "rb");
    if (f == nullptr) throw std::runtime_error("cannot open index file for reading");
    struct Closer {
        std::FILE* f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    auto rd = [&](void* p, size_t bytes) {
        if (bytes != 0 && std::fread(p, 1, bytes, f) != bytes) throw std::runtime_error("short read while loading index");
    };
    uint32_t magic = 0;
    rd(&magic, 4);
    if (magic != kMagic) throw std::runtime_error("not an E8Q index file");
    uint64_t n = 0, dim = 0, nl = 0, rot_bytes = 0;
    uint32_t bits = 0;
    rd(&n, 8);
    rd(&dim, 8);
    rd(&nl, 8);
    rd(&bits, 4);
    rd(&rot_bytes, 8);
    if (bits < 1 || bits > static_cast<uint32_t>(hostmath::kMaxQuantizeBits) || nl == 0 || dim == 0)
        throw std::runtime_error("corrupt E8Q index header");
    const size_t padded = hostmath::PaddedDim(dim);
    if (rot_bytes != 4 * padded / 8) throw std::runtime_error("rotator payload size mismatch");
    // Header sanity against the file size before large allocations.
    std::fseek(f, 0, SEEK_END);
    const long fsize = std::ftell(f);
    std::fseek(f, 40, SEEK_SET);
    const uint64_t fixed = n * 4 + rot_bytes + padded * 4 + nl * padded * 4 + nl * 4 + 3 * (nl + 1) * 8;
    if (fsize < 0 || fixed > static_cast<uint64_t>(fsize)) throw std::runtime_error("short read while loading index");
    std::vector<uint32_t> ids(n);
    std::vector<uint8_t> signs(rot_bytes);
    std::vector<float> mean(padded), crot(nl * padded), cn2(nl);
    std::vector<uint64_t> cb(nl + 1), ab(nl + 1), rb(nl + 1);
    rd(ids.data(), n * 4);
    rd(signs.data(), rot_bytes);
    rd(mean.data(), padded * 4);
    rd(crot.data(), crot.size() * 4);
    rd(cn2.data(), nl * 4);
    rd(cb.data(), (nl + 1) * 8);
    rd(ab.data(), (nl + 1) * 8);
    if (cb.back() != n || cb[0] != 0) throw std::runtime_error("cluster offsets inconsistent");
    const size_t blocks = padded / 8, bb = BatchBytes(blocks), rbb = RefBatchBytes(blocks, static_cast<int>(bits));
    uint64_t abytes = 0, rbytes = 0;
    for (size_t c = 0; c < nl; ++c) {
        if (cb[c + 1] < cb[c]) throw std::runtime_error("cluster offsets inconsistent");
        if (ab[c] != abytes) throw std::runtime_error("arena offsets inconsistent");
        abytes += (cb[c + 1] - cb[c] + 31) / 32 * bb;
    }
    if (ab[nl] != abytes) throw std::runtime_error("arena offsets inconsistent");
    std::vector<uint8_t> arena(abytes + 64);
    rd(arena.data(), arena.size());
    rd(rb.data(), (nl + 1) * 8);
    for (size_t c = 0; c < nl; ++c) {
        if (rb[c] != rbytes) throw std::runtime_error("refine arena offsets inconsistent");
        rbytes += (cb[c + 1] - cb[c] + 31) / 32 * rbb;
    }
    if (rb[nl] != rbytes) throw std::runtime_error("refine arena offsets inconsistent");
    std::vector<uint8_t> ref(rbytes + 64);
    rd(ref.data(), ref.size());
    float alpha = 1.0F;
    uint32_t lead = 0;
    if (std::fread(&lead, 4, 1, f) == 1 && lead == kOptMagic) {
        uint32_t lam = 0, tail = 0;
        if (std::fread(&lam, 4, 1, f) != 1 || std::fread(&tail, 4, 1, f) != 1 || tail != kOptMagic)
            throw std::runtime_error("build-options trailer corrupt");
        float v = 0.0F;
        std::memcpy(&v, &lam, 4);
        alpha = (v > 0.0F && v == v) ? v : 1.0F;
    }
    Reset();
    n_ = n;
    dim_ = dim;
    nlist_ = nl;
    bits_ = static_cast<int>(bits);
    padded_ = padded;
    alpha_ = alpha;
    signs_ = std::move(signs);
    cluster_begin_ = std::move(cb);
    arena_begin_ = std::move(ab);
    ref_arena_begin_ = std::move(rb);
    CreateDevice(signs_);
    e8q_index_arrays h{ids.data(), mean.data(), crot.data(), cn2.data(), arena.data(), ref.data()};
    std::lock_guard<std::mutex> lk(DeviceMutex());
    Check(e8q_index_upload(ctx_, dev_, &h));
}

void E8qIndex::SearchBatch(const float* queries, size_t nq, size_t k, size_t nprobe,
                           const uint32_t* nprobe_per_query, uint32_t* out_ids, float* out_dists, SearchStats* stats,
                           SearchMode mode, bool queries_on_device, bool outputs_on_device) const {
    if (dev_ == nullptr) throw std::runtime_error("E8qIndex must be built or loaded before search");
    if (k == 0) throw std::invalid_argument("k must be positive");
    if (k > kMaxK) throw std::invalid_argument("k must be <= " + std::to_string(kMaxK));
    if (nprobe == 0 && nprobe_per_query == nullptr) throw std::invalid_argument("nprobe must be positive");
    e8q_search_args a{};
    a.queries = queries;
    a.nq = nq;
    a.k = k;
    a.nprobe = nprobe;
    a.nprobe_per_query = nprobe_per_query;
    a.z = hostmath::OrthoQuantile(alpha_, std::max<size_t>(4 * k, 32));
    a.disable_prune = disable_prune_ ? 1 : 0;
    a.flush_every = hostmath::FlushEvery();
    a.mode = mode == SearchMode::kBatched ? E8Q_MODE_BATCHED
             : mode == SearchMode::kMfma  ? E8Q_MODE_MFMA
                                          : E8Q_MODE_PLAIN;
    a.queries_on_device = queries_on_device ? 1 : 0;
    a.out_ids = out_ids;
    a.out_dists = out_dists;
    a.outputs_on_device = outputs_on_device ? 1 : 0;
    e8q_search_stats st{};
    {
        std::lock_guard<std::mutex> lk(DeviceMutex());
        Check(e8q_search(ctx_, dev_, &a, &st));
    }
    if (stats != nullptr) {
        stats->candidates += st.candidates;
        stats->survivors += st.survivors;
        stats->batches += st.batches;
        stats->scanned += st.scanned;
        stats->device_scan_ms += st.scan_ms;
        stats->device_total_ms += st.total_ms;
    }
}

void E8qIndex::SearchOnContext(e8q_ctx* ctx, const float* queries, size_t nq, size_t k,
                               const uint32_t* nprobe_per_query, uint32_t* out_ids, float* out_dists,
                               e8q_search_stats* stats) const {
    if (dev_ == nullptr) throw std::runtime_error("E8qIndex must be built or loaded before search");
    if (k == 0 || k > kMaxK) throw std::invalid_argument("k must be in [1, " + std::to_string(kMaxK) + "]");
    e8q_search_args a{};
    a.queries = queries;
    a.nq = nq;
    a.k = k;
    a.nprobe = 1;
    a.nprobe_per_query = nprobe_per_query;
    a.z = hostmath::OrthoQuantile(alpha_, std::max<size_t>(4 * k, 32));
    a.disable_prune = disable_prune_ ? 1 : 0;
    a.flush_every = hostmath::FlushEvery();
    a.mode = E8Q_MODE_PLAIN;
    a.out_ids = out_ids;
    a.out_dists = out_dists;
    Check(e8q_search(ctx, dev_, &a, stats));
}

uint64_t E8qIndex::SearchWorkspaceBytes(size_t nq, size_t k, size_t max_nprobe) const {
    return dev_ ? e8q_search_workspace_bytes(dev_, nq, k, max_nprobe) : 0;
}

void E8qIndex::Search(const float* query, size_t k, size_t nprobe, uint32_t* out_ids, QueryScratch& scratch,
                      SearchStats* stats, float* out_dists) const {
    (void)scratch;
    std::vector<float> d(k);
    SearchBatch(query, 1, k, nprobe, nullptr, out_ids, out_dists != nullptr ? out_dists : d.data(), stats,
                SearchMode::kPlain);
}

void E8qIndex::ExhaustiveTopK(const float* query, size_t k, std::vector<std::pair<float, uint32_t>>& out) const {
    if (dev_ == nullptr) throw std::runtime_error("E8qIndex must be built or loaded before search");
    out.clear();
    const size_t kk = std::min(k, n_);
    if (kk == 0) return;
    if (kk > 1024) throw std::invalid_argument("ExhaustiveTopK: k above the device limit of 1024");
    std::vector<uint32_t> ids(kk);
    std::vector<float> d(kk);
    {
        std::lock_guard<std::mutex> lk(DeviceMutex());
        Check(e8q_exhaustive_topk(ctx_, dev_, query, 1, kk, ids.data(), d.data()));
    }
    for (size_t i = 0; i < kk && ids[i] != std::numeric_limits<uint32_t>::max(); ++i) out.emplace_back(d[i], ids[i]);
}

float E8qIndex::Estimate(const float* query, uint32_t id) const {
    if (dev_ == nullptr || n_ == 0) throw std::runtime_error("Estimate: empty index");
    float out = 0.0F;
    const uint32_t q0 = 0;
    std::lock_guard<std::mutex> lk(DeviceMutex());
    const e8q_status st = e8q_estimate(ctx_, dev_, query, 1, &q0, &id, 1, &out);
    if (st == E8Q_EINVAL) throw std::invalid_argument(e8q_last_error());
    Check(st);
    return out;
}

double E8qIndex::ProbeBoundW(const float* query, const float* base_row, uint32_t id) const {
    double out = 0.0;
    const uint32_t q0 = 0;
    ProbeBoundWBatch(query, 1, base_row, &q0, &id, 1, &out);
    return out;
}

void E8qIndex::ProbeBoundWBatch(const float* queries, size_t nq, const float* base_rows, const uint32_t* qidx,
                                const uint32_t* ids, size_t npair, double* out) const {
    if (dev_ == nullptr || n_ == 0) throw std::runtime_error("ProbeBoundW: empty index");
    std::lock_guard<std::mutex> lk(DeviceMutex());
    const e8q_status st = e8q_probe_bound_w(ctx_, dev_, queries, nq, base_rows, qidx, ids, npair, out);
    if (st == E8Q_EINVAL) throw std::invalid_argument(e8q_last_error());
    Check(st);
}

}  // namespace e8qlib
