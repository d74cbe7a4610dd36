// me8q_index.cpp -- host side of e8qlib::me8q::Me8qIndex over the device C ABI.
//
// Build (me8q_build.cpp:43-311): plan (SAQ DP or manual spec), code-row
// layout, list offsets, per-run rotator signs, level tables, then the device
// encode + FastScan layout.  Save / Load: the ME8QIDXX file byte for byte
// (me8q_io.cpp:29-176).  Search: lambda on the host, everything else on the
// device.
#include "me8q_index.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>

#include "e8q_gpu.h"
#include "e8q_host_math.h"
#include "e8q_index.h"
#include "me8q_gpu.h"

namespace e8qlib::me8q {
namespace {

constexpr char kMagic[8] = {'M', 'E', '8', 'Q', 'I', 'D', 'X', 'X'};

void Check(e8q_status st) {
    if (st == E8Q_OK) return;
    const std::string msg = e8q_last_error();
    if (st == E8Q_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error("e8q device: " + msg);
}

void Put(std::ofstream& o, const void* p, size_t n) { o.write(static_cast<const char*>(p), static_cast<std::streamsize>(n)); }
void Get(std::ifstream& i, void* p, size_t n, const char* what) {
    i.read(static_cast<char*>(p), static_cast<std::streamsize>(n));
    if (!i) throw std::runtime_error(std::string("Me8qIndex::Load: truncated ") + what);
}

}  // namespace

size_t PaddedDim(size_t dim) { return (dim + kSegDim - 1) / kSegDim * kSegDim; }

// Minimum total distortion sum_runs var/2^b over plans of at most `max_segs`
// runs whose payload + 64-bit factors fit the budget; a run may also drop
// every remaining dimension at the cost of its full variance.  The budget is
// tracked in 64-bit units (every run spends b*j*64 + 64 bits).  Ties: a plan
// replaces the best only when 1% cheaper, scanning shorter plans first.
Plan Allocate(const PlanInput& in) {
    if (in.var == nullptr || in.dim == 0 || in.dim % kSegDim != 0)
        throw std::invalid_argument("Allocate: dim must be a positive multiple of 64");
    if (in.max_bits < 1 || in.max_bits > kMaxBits) throw std::invalid_argument("Allocate: max_bits must be in [1, 13]");
    const size_t nb = in.dim / kSegDim;
    std::vector<double> cum(nb + 1, 0.0);
    for (size_t b = 0; b < nb; ++b) {
        double s = 0.0;
        for (size_t k = 0; k < kSegDim; ++k) s += static_cast<double>(in.var[b * kSegDim + k]);
        cum[b + 1] = cum[b] + s;
    }
    const double budget = static_cast<double>(in.avg_bits) * static_cast<double>(in.dim) + in.extra_bits +
                          static_cast<double>(kFactorBits);
    const size_t units = (budget > 0.0 ? static_cast<size_t>(budget) : 0) / kSegDim;
    const size_t max_segs = in.avg_bits < 2.0F ? nb : nb / 2;
    const double kInf = std::numeric_limits<double>::max();
    const size_t span_i = nb + 1, span_t = units + 1;
    std::vector<double> cost((max_segs + 1) * span_i * span_t, kInf);
    std::vector<uint32_t> from(cost.size(), 0);
    auto cell = [&](size_t ns, size_t i, size_t t) { return (ns * span_i + i) * span_t + t; };
    cost[cell(0, 0, 0)] = 0.0;
    double best = kInf;
    size_t best_ns = 0, best_t = 0;
    for (size_t ns = 0; ns <= max_segs; ++ns) {
        for (size_t i = 0; i <= nb; ++i) {
            for (size_t t = 0; t <= units; ++t) {
                const double cur = cost[cell(ns, i, t)];
                if (cur >= kInf) continue;
                if (i == nb) {
                    if (cur * 1.01 < best) {
                        best = cur;
                        best_ns = ns;
                        best_t = t;
                    }
                    continue;
                }
                if (ns == max_segs) continue;
                for (size_t j = 1; i + j <= nb; ++j) {
                    const double span_var = cum[i + j] - cum[i];
                    for (int b = 1; b <= in.max_bits; ++b) {
                        const size_t t2 = t + static_cast<size_t>(b) * j + 1;
                        if (t2 > units) break;
                        const double c2 = cur + span_var / static_cast<double>(int64_t{1} << b);
                        const size_t to = cell(ns + 1, i + j, t2);
                        if (c2 < cost[to]) {
                            cost[to] = c2;
                            from[to] = static_cast<uint32_t>((i << 4) | static_cast<size_t>(b));
                        }
                    }
                }
                const double c_drop = cur + (cum[nb] - cum[i]);
                const size_t to = cell(ns + 1, nb, t);
                if (c_drop < cost[to]) {
                    cost[to] = c_drop;
                    from[to] = static_cast<uint32_t>(i << 4);
                }
            }
        }
    }
    if (best >= kInf) {
        std::fprintf(stderr,
                     "e8qlib::me8q::Allocate: SAQ cap admits no plan (dim=%zu, avg_bits=%g); "
                     "emitting the all-dropped run as the reference does\n",
                     in.dim, static_cast<double>(in.avg_bits));
        return Plan{{in.dim, 0}};
    }
    Plan plan;
    size_t ns = best_ns, i = nb, t = best_t;
    while (i > 0) {
        const uint32_t f = from[cell(ns, i, t)];
        const size_t i0 = f >> 4;
        const int bits = static_cast<int>(f & 0xF);
        plan.push_back({(i - i0) * kSegDim, bits});
        --ns;
        if (bits != 0) t -= static_cast<size_t>(bits) * (i - i0) + 1;
        i = i0;
    }
    std::reverse(plan.begin(), plan.end());
    return plan;
}

Plan ParsePlanSpec(const std::string& spec, size_t padded_dim) {
    Plan plan;
    size_t pos = 0, total = 0;
    while (pos <= spec.size()) {
        const size_t comma = spec.find(',', pos);
        const std::string tok = spec.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos);
        if (!tok.empty()) {
            const size_t colon = tok.find(':');
            if (colon == std::string::npos) throw std::invalid_argument("plan_spec: expected L:B, got " + tok);
            const long len = std::atol(tok.substr(0, colon).c_str());
            const long bits = std::atol(tok.substr(colon + 1).c_str());
            if (len <= 0 || len % static_cast<long>(kSegDim) != 0 || bits < 0 || bits > kMaxBits)
                throw std::invalid_argument("plan_spec: bad run " + tok);
            plan.push_back({static_cast<size_t>(len), static_cast<int>(bits)});
            total += static_cast<size_t>(len);
        }
        if (comma == std::string::npos) break;
        pos = comma + 1;
    }
    if (total != padded_dim)
        throw std::invalid_argument("plan_spec: runs cover " + std::to_string(total) + " of " +
                                    std::to_string(padded_dim) + " padded dims");
    return plan;
}

// fp32 column means, then fp64 squared deviations; the reference build fuses
// each `acc += d*d` into one multiply-add.
void DatasetVariance(const float* base, size_t ld, size_t n, size_t dim, std::vector<float>* var) {
    std::vector<float> mean(dim, 0.0F);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < dim; ++j) mean[j] += base[i * ld + j];
    for (size_t j = 0; j < dim; ++j) mean[j] /= static_cast<float>(n);
    std::vector<double> sq(dim, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < dim; ++j) {
            const double d = static_cast<double>(base[i * ld + j]) - static_cast<double>(mean[j]);
            sq[j] = std::fma(d, d, sq[j]);
        }
    for (size_t j = 0; j < dim; ++j) (*var)[j] = static_cast<float>(sq[j] / static_cast<double>(n));
}

Me8qIndex::Me8qIndex() = default;
Me8qIndex::~Me8qIndex() { Reset(); }

Me8qIndex::Me8qIndex(Me8qIndex&& o) noexcept { *this = std::move(o); }

Me8qIndex& Me8qIndex::operator=(Me8qIndex&& o) noexcept {
    if (this != &o) {
        Reset();
        n_ = o.n_;
        dim_ = o.dim_;
        padded_ = o.padded_;
        nlist_ = o.nlist_;
        bytes_per_vector_ = o.bytes_per_vector_;
        opt_ = std::move(o.opt_);
        ortho_alpha_ = o.ortho_alpha_;
        plan_ = std::move(o.plan_);
        run_signs_ = std::move(o.run_signs_);
        list_begin_ = std::move(o.list_begin_);
        ctx_ = o.ctx_;
        dev_ = o.dev_;
        o.dev_ = nullptr;
    }
    return *this;
}

void Me8qIndex::Reset() {
    if (dev_) {
        std::lock_guard<std::mutex> lk(DeviceMutex());
        (void)e8q_m8_destroy(dev_);
        dev_ = nullptr;
    }
}

void Me8qIndex::CreateDevice() {
    ctx_ = DefaultContext();
    std::vector<uint64_t> seg_len;
    std::vector<int32_t> seg_bits;
    std::vector<uint8_t> signs;
    for (const PlanSeg& s : plan_) {
        seg_len.push_back(s.len);
        seg_bits.push_back(s.bits);
    }
    for (const auto& s : run_signs_) signs.insert(signs.end(), s.begin(), s.end());
    std::vector<double> levels;
    for (int b = 1; b <= kMaxBits; ++b)
        for (int br = 0; br < 2; ++br) {
            const auto l = hostmath::BranchLevels(b, br);
            levels.insert(levels.end(), l.begin(), l.end());
        }
    double fixed[kMaxBits + 1] = {};
    for (int b = 1; b <= kMaxBits; ++b) fixed[b] = hostmath::FixedScale(b);
    e8q_m8_desc d{};
    d.n = n_;
    d.dim = dim_;
    d.padded = padded_;
    d.nlist = nlist_;
    d.nseg = static_cast<uint32_t>(plan_.size());
    d.seg_len = seg_len.data();
    d.seg_bits = seg_bits.data();
    d.signs = signs.empty() ? nullptr : signs.data();
    d.levels = levels.data();
    d.fixed_scale = fixed;
    d.list_begin = list_begin_.data();
    d.drop_energy = opt_.drop_energy ? 1 : 0;
    std::lock_guard<std::mutex> lk(DeviceMutex());
    Check(e8q_m8_create(ctx_, &d, &dev_));
}

void Me8qIndex::Build(const float* base, size_t ld, const float* centroids, size_t ld_c, const uint32_t* cluster_ids,
                      size_t n, size_t dim, size_t num_clusters, const BuildOptions& opt, const float* data_variance) {
    BuildImpl(base, ld, centroids, ld_c, cluster_ids, nullptr, n, dim, num_clusters, opt, data_variance);
}

void Me8qIndex::BuildFromDevice(const float* base_device, const float* centroids, const uint32_t* cluster_ids_host,
                                const uint32_t* cluster_ids_device, size_t n, size_t dim, size_t num_clusters,
                                const BuildOptions& opt, const float* data_variance) {
    if (cluster_ids_device == nullptr) throw std::invalid_argument("Me8qIndex::BuildFromDevice: null device cluster ids");
    if (data_variance == nullptr && opt.plan_spec.empty())
        throw std::invalid_argument("Me8qIndex::BuildFromDevice: the plan needs a variance or a plan_spec");
    BuildImpl(base_device, dim, centroids, dim, cluster_ids_host, cluster_ids_device, n, dim, num_clusters, opt,
              data_variance);
}

void Me8qIndex::BuildImpl(const float* base, size_t ld, const float* centroids, size_t ld_c,
                          const uint32_t* cluster_ids, const uint32_t* cluster_ids_device, size_t n, size_t dim,
                          size_t num_clusters, const BuildOptions& opt, const float* data_variance) {
    if (base == nullptr || centroids == nullptr || cluster_ids == nullptr || n == 0 || dim == 0 || num_clusters == 0)
        throw std::invalid_argument("Me8qIndex::Build: empty input");
    if (opt.max_bits < 1 || opt.max_bits > kMaxBits)
        throw std::invalid_argument("Me8qIndex::Build: max_bits must be in [1, 13]");
    if (ld < dim || ld_c < dim) throw std::invalid_argument("Me8qIndex::Build: leading dimension below dim");
    Reset();
    opt_ = opt;
    ortho_alpha_ = opt.prune_alpha > 0.0F ? opt.prune_alpha : 1.0F;
    n_ = n;
    dim_ = dim;
    padded_ = PaddedDim(dim);
    nlist_ = num_clusters;
    std::vector<float> cent(nlist_ * padded_, 0.0F);
    for (size_t c = 0; c < nlist_; ++c) std::memcpy(&cent[c * padded_], centroids + c * ld_c, dim * sizeof(float));
    list_begin_.assign(nlist_ + 1, 0);
    for (size_t i = 0; i < n; ++i) {
        if (cluster_ids[i] >= nlist_) throw std::invalid_argument("cluster id out of range");
        ++list_begin_[cluster_ids[i] + 1];
    }
    for (size_t c = 0; c < nlist_; ++c) list_begin_[c + 1] += list_begin_[c];
    std::vector<uint32_t> rank(n);
    {
        std::vector<uint64_t> cur(list_begin_.begin(), list_begin_.end() - 1);
        for (size_t i = 0; i < n; ++i) rank[i] = static_cast<uint32_t>(cur[cluster_ids[i]]++ - list_begin_[cluster_ids[i]]);
    }
    if (!opt_.plan_spec.empty()) {
        plan_ = ParsePlanSpec(opt_.plan_spec, padded_);
    } else {
        std::vector<float> var(padded_, 0.0F);
        if (data_variance != nullptr) std::memcpy(var.data(), data_variance, dim_ * sizeof(float));
        else DatasetVariance(base, ld, n_, dim_, &var);
        PlanInput pin;
        pin.var = var.data();
        pin.dim = padded_;
        pin.avg_bits = opt_.avg_bits;
        pin.max_bits = opt_.max_bits;
        pin.extra_bits = opt_.extra_budget_bits;
        plan_ = Allocate(pin);
        std::string s2;
        for (const PlanSeg& sg : plan_) s2 += std::to_string(sg.len) + ":" + std::to_string(sg.bits) + " ";
        std::fprintf(stderr, "me8q: plan=%s\n", s2.c_str());
    }
    // per-run rotators: seed advanced by the golden-ratio step per coded run
    run_signs_.clear();
    uint64_t s = opt_.rot_seed;
    size_t nseg = 0, stride = 0;
    for (const PlanSeg& sg : plan_) {
        if (sg.bits == 0) continue;
        s += 0x9E3779B97F4A7C15ULL;
        run_signs_.push_back(hostmath::RotatorSigns(sg.len, s));
        ++nseg;
        stride += static_cast<size_t>(sg.bits) * sg.len / 8;
    }
    bytes_per_vector_ = nseg * 8 + stride;
    CreateDevice();
    std::vector<float> packed;
    const float* rows = base;
    if (ld != dim) {
        packed.resize(n * dim);
        for (size_t i = 0; i < n; ++i) std::memcpy(&packed[i * dim], base + i * ld, dim * sizeof(float));
        rows = packed.data();
    }
    std::lock_guard<std::mutex> lk(DeviceMutex());
    if (cluster_ids_device != nullptr)
        Check(e8q_m8_encode(ctx_, dev_, rows, cent.data(), cluster_ids_device, rank.data(), 1));
    else
        Check(e8q_m8_encode(ctx_, dev_, rows, cent.data(), cluster_ids, rank.data(), 0));
}

void Me8qIndex::Save(const char* path) const {
    if (!dev_) throw std::runtime_error("Me8qIndex::Save: index not built");
    std::vector<uint32_t> ids(n_);
    std::vector<float> cent(nlist_ * padded_);
    std::vector<uint8_t> codes(n_ * bytes_per_vector_);
    {
        e8q_m8_arrays a{ids.data(), cent.data(), codes.data()};
        std::lock_guard<std::mutex> lk(DeviceMutex());
        Check(e8q_m8_download(ctx_, dev_, &a));
    }
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error(std::string("Me8qIndex::Save: cannot write ") + path);
    Put(out, kMagic, 8);
    const uint64_t hdr[5] = {n_, dim_, padded_, nlist_, bytes_per_vector_};
    Put(out, hdr, sizeof(hdr));
    Put(out, &opt_.avg_bits, 4);
    int32_t flags[4] = {0, opt_.max_bits, opt_.drop_energy ? 1 : 0, 0};
    std::memcpy(&flags[3], &opt_.prune_alpha, 4);
    Put(out, flags, sizeof(flags));
    const uint64_t np = plan_.size();
    Put(out, &np, 8);
    for (const PlanSeg& sg : plan_) {
        const uint64_t len = sg.len;
        const int32_t bits = sg.bits;
        Put(out, &len, 8);
        Put(out, &bits, 4);
    }
    Put(out, cent.data(), cent.size() * 4);
    Put(out, ids.data(), ids.size() * 4);
    Put(out, list_begin_.data(), list_begin_.size() * 8);
    std::vector<uint64_t> byte_begin(nlist_ + 1);
    for (size_t c = 0; c <= nlist_; ++c) byte_begin[c] = list_begin_[c] * bytes_per_vector_;
    Put(out, byte_begin.data(), byte_begin.size() * 8);
    Put(out, codes.data(), codes.size());
    Put(out, &opt_.rot_seed, 8);
    const uint64_t nrot = run_signs_.size();
    Put(out, &nrot, 8);
    for (const auto& sg : run_signs_) {
        const uint64_t nb = sg.size();
        Put(out, &nb, 8);
        Put(out, sg.data(), nb);
    }
    if (!out) throw std::runtime_error(std::string("Me8qIndex::Save: write failed ") + path);
}

void Me8qIndex::Load(const char* path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error(std::string("Me8qIndex::Load: cannot read ") + path);
    char magic[8] = {};
    Get(in, magic, 8, "magic");
    if (std::memcmp(magic, kMagic, 8) != 0) throw std::runtime_error(std::string("Me8qIndex::Load: bad magic ") + path);
    Reset();
    uint64_t hdr[5] = {};
    Get(in, hdr, sizeof(hdr), "header");
    n_ = hdr[0];
    dim_ = hdr[1];
    padded_ = hdr[2];
    nlist_ = hdr[3];
    bytes_per_vector_ = hdr[4];
    float avg = 0.0F;
    int32_t flags[4] = {};
    Get(in, &avg, 4, "avg_bits");
    Get(in, flags, sizeof(flags), "flags");
    opt_ = BuildOptions{};
    opt_.avg_bits = avg;
    opt_.max_bits = flags[1];
    opt_.drop_energy = flags[2] != 0;
    std::memcpy(&opt_.prune_alpha, &flags[3], 4);
    ortho_alpha_ = (opt_.prune_alpha > 0.0F && opt_.prune_alpha == opt_.prune_alpha) ? opt_.prune_alpha : 1.0F;
    uint64_t np = 0;
    Get(in, &np, 8, "plan size");
    plan_.clear();
    for (uint64_t i = 0; i < np; ++i) {
        uint64_t len = 0;
        int32_t bits = 0;
        Get(in, &len, 8, "plan segment length");
        Get(in, &bits, 4, "plan segment bits");
        plan_.push_back({static_cast<size_t>(len), static_cast<int>(bits)});
    }
    if (padded_ != PaddedDim(dim_)) throw std::runtime_error("Me8qIndex::Load: padded dim mismatch");
    std::vector<float> cent(nlist_ * padded_);
    Get(in, cent.data(), cent.size() * 4, "centroids");
    std::vector<uint32_t> ids(n_);
    Get(in, ids.data(), ids.size() * 4, "ids");
    list_begin_.assign(nlist_ + 1, 0);
    Get(in, list_begin_.data(), list_begin_.size() * 8, "list_begin");
    std::vector<uint64_t> byte_begin(nlist_ + 1);
    Get(in, byte_begin.data(), byte_begin.size() * 8, "byte_begin");
    std::vector<uint8_t> codes(byte_begin[nlist_]);
    Get(in, codes.data(), codes.size(), "codes");
    Get(in, &opt_.rot_seed, 8, "rot_seed");
    uint64_t nrot = 0;
    Get(in, &nrot, 8, "rotator count");
    size_t ncoded = 0, stride = 0;
    for (const PlanSeg& sg : plan_) {
        if (sg.bits == 0) continue;
        ++ncoded;
        stride += static_cast<size_t>(sg.bits) * sg.len / 8;
    }
    if (nrot != ncoded) throw std::runtime_error(std::string("Me8qIndex::Load: rotator count mismatch ") + path);
    run_signs_.clear();
    for (const PlanSeg& sg : plan_) {
        if (sg.bits == 0) continue;
        uint64_t nb = 0;
        Get(in, &nb, 8, "rot_bytes");
        if (nb != 4 * sg.len / 8) throw std::runtime_error(std::string("Me8qIndex::Load: rotator payload size mismatch ") + path);
        std::vector<uint8_t> sgn(nb);
        Get(in, sgn.data(), nb, "rotator");
        run_signs_.push_back(std::move(sgn));
    }
    if (ncoded * 8 + stride != bytes_per_vector_ || codes.size() != n_ * bytes_per_vector_ ||
        list_begin_[nlist_] != n_)
        throw std::runtime_error(std::string("Me8qIndex::Load: layout mismatch ") + path);
    CreateDevice();
    e8q_m8_arrays a{ids.data(), cent.data(), codes.data()};
    std::lock_guard<std::mutex> lk(DeviceMutex());
    Check(e8q_m8_upload(ctx_, dev_, &a));
}

size_t Me8qIndex::route_dim() const {
    size_t r = padded_;
    for (auto it = plan_.rbegin(); it != plan_.rend(); ++it) {
        if (it->bits != 0) break;
        r -= it->len;
    }
    return r;
}

double Me8qIndex::bits_per_dim() const {
    return dim_ == 0 ? 0.0 : 8.0 * static_cast<double>(bytes_per_vector_) / static_cast<double>(dim_);
}

// The reference's accounting (me8q_io.cpp:182-196): codes, ids, centroids,
// offsets, per-list segment slots, rotator signs, rotated centroids.
size_t Me8qIndex::memory_bytes() const {
    size_t coded = 0, rot = 0;
    for (const PlanSeg& sg : plan_) coded += sg.bits > 0 ? 1 : 0;
    for (const auto& s : run_signs_) rot += s.size();
    return n_ * bytes_per_vector_ + n_ * 4 + nlist_ * padded_ * 4 + 2 * (nlist_ + 1) * 8 + nlist_ * coded * 16 + rot +
           nlist_ * padded_ * 4;
}

size_t Me8qIndex::device_bytes() const { return dev_ ? e8q_m8_device_bytes(dev_) : 0; }

namespace {
// Threshold-pass length of the batched form in 32-vector batches
// (E8Q_M8_BATCH_BUDGET; 0 = the library default).
uint32_t BatchedBudget() {
    const char* e = std::getenv("E8Q_M8_BATCH_BUDGET");
    return e ? static_cast<uint32_t>(std::strtoul(e, nullptr, 10)) : 0u;
}
// Candidates per query the batched form refines first (E8Q_M8_SET_A; 0 =
// the library default).
uint32_t BatchedSetA() {
    const char* e = std::getenv("E8Q_M8_SET_A");
    return e ? static_cast<uint32_t>(std::strtoul(e, nullptr, 10)) : 0u;
}
}  // namespace

void Me8qIndex::SearchDevice(const float* queries, size_t nq, size_t k, size_t nprobe, const uint32_t* nprobe_per_query,
                             uint32_t* out_ids, float* out_dists, SearchStats* stats, bool queries_on_device,
                             bool outputs_on_device, uint64_t* per_query_stats, int mode) const {
    if (!dev_) throw std::runtime_error("Me8qIndex must be built or loaded before search");
    if (k > kMaxDeviceK) throw std::invalid_argument("k above the device limit of 1024");
    e8q_m8_search_args a{};
    a.queries = queries;
    a.nq = nq;
    a.k = k;
    a.nprobe = nprobe;
    a.nprobe_per_query = nprobe_per_query;
    a.z = hostmath::OrthoQuantile(static_cast<double>(ortho_alpha_), std::max<size_t>(4 * k, 32));
    a.queries_on_device = queries_on_device ? 1 : 0;
    a.out_ids = out_ids;
    a.out_dists = out_dists;
    a.outputs_on_device = outputs_on_device ? 1 : 0;
    a.mode = mode;
    a.batched_budget = BatchedBudget();
    a.batched_set_a = BatchedSetA();
    e8q_m8_stats st{};
    {
        std::lock_guard<std::mutex> lk(DeviceMutex());
        Check(e8q_m8_search(ctx_, dev_, &a, &st, per_query_stats));
    }
    if (stats) {
        stats->candidates += st.candidates;
        stats->survivors += st.survivors;
        stats->bits_scored += st.bits_scored;
        stats->bits_total += st.bits_total;
        stats->pruned += st.pruned;
        stats->clusters += nq * std::min(nprobe, nlist_);
        stats->device_scan_ms += st.scan_ms;
        stats->device_total_ms += st.total_ms;
        stats->batches += st.batches;
        stats->refined += st.refined;
        stats->groups += st.groups;
        stats->form = st.form;
        stats->overflowed += st.overflowed;
        stats->pairs_scored += st.pairs_scored;
        stats->items += st.items;
        for (int i = 0; i < 2; ++i) {
            stats->refine_ms[i] += st.refine_ms[i];
            stats->refine_launches[i] += st.refine_launches[i];
            stats->refine_bytes[i] += st.refine_bytes[i];
        }
    }
}

void Me8qIndex::SearchOnContext(e8q_ctx* ctx, const float* queries, size_t nq, size_t k,
                                const uint32_t* nprobe_per_query, uint32_t* out_ids, float* out_dists,
                                int mode, e8q_m8_stats* stats) const {
    if (!dev_) throw std::runtime_error("Me8qIndex must be built or loaded before search");
    if (k == 0 || k > kMaxDeviceK) throw std::invalid_argument("k must be in [1, 1024] on the device");
    e8q_m8_search_args a{};
    a.queries = queries;
    a.nq = nq;
    a.k = k;
    a.nprobe = 1;
    a.nprobe_per_query = nprobe_per_query;
    a.z = hostmath::OrthoQuantile(static_cast<double>(ortho_alpha_), std::max<size_t>(4 * k, 32));
    a.out_ids = out_ids;
    a.out_dists = out_dists;
    a.mode = mode;
    a.batched_budget = BatchedBudget();
    a.batched_set_a = BatchedSetA();
    Check(e8q_m8_search(ctx, dev_, &a, stats, nullptr));
}

uint64_t Me8qIndex::SearchWorkspaceBytes(size_t nq, size_t k, size_t max_nprobe, int mode) const {
    return dev_ ? e8q_m8_search_workspace_bytes(dev_, nq, k, max_nprobe, mode) : 0;
}

void Me8qIndex::SearchIVFBatch(const float* queries, size_t nq, size_t k, size_t nprobe, uint32_t* out_ids,
                               float* out_dists, uint64_t* stats6) const {
    if (!dev_) throw std::runtime_error("Me8qIndex must be built or loaded before search");
    std::lock_guard<std::mutex> lk(DeviceMutex());
    Check(e8q_m8_search_ivf(ctx_, dev_, queries, nq, k, nprobe, out_ids, out_dists, stats6));
}

void Me8qIndex::SearchIVF(const float* query, size_t k, size_t nprobe, uint32_t* out_ids, SearchStats* stats,
                          float* out_dists) const {
    if (k == 0 || nprobe == 0 || n_ == 0) return;
    std::vector<float> d(k);
    uint64_t st[6] = {0, 0, 0, 0, 0, 0};
    SearchIVFBatch(query, 1, k, nprobe, out_ids, d.data(), st);
    if (out_dists) std::copy(d.begin(), d.end(), out_dists);
    if (stats) {
        stats->candidates = st[0];
        stats->clusters = st[1];
        stats->bits_scored = st[2];
        stats->bits_total = st[3];
        stats->survivors = st[4];
        stats->pruned = st[5];
    }
}

void Me8qIndex::ExhaustiveTopKBatch(const float* queries, size_t nq, size_t k, uint32_t* out_ids,
                                    float* out_dists) const {
    if (!dev_) throw std::runtime_error("Me8qIndex must be built or loaded before search");
    std::lock_guard<std::mutex> lk(DeviceMutex());
    Check(e8q_m8_exhaustive_topk(ctx_, dev_, queries, nq, k, out_ids, out_dists));
}

void Me8qIndex::ExhaustiveTopK(const float* query, size_t k, std::vector<std::pair<float, uint32_t>>* out) const {
    out->clear();
    const size_t kk = std::min(k, n_);
    if (kk == 0) return;
    std::vector<uint32_t> ids(kk);
    std::vector<float> d(kk);
    ExhaustiveTopKBatch(query, 1, kk, ids.data(), d.data());
    for (size_t i = 0; i < kk && ids[i] != std::numeric_limits<uint32_t>::max(); ++i) out->emplace_back(d[i], ids[i]);
}

void Me8qIndex::EstimateBatch(const float* queries, size_t nq, const uint32_t* qidx, const uint32_t* ids,
                              size_t npair, float* out) const {
    if (!dev_) throw std::runtime_error("Me8qIndex must be built or loaded before estimate");
    std::lock_guard<std::mutex> lk(DeviceMutex());
    const e8q_status st = e8q_m8_estimate(ctx_, dev_, queries, nq, qidx, ids, npair, out);
    if (st == E8Q_EINVAL) throw std::out_of_range(e8q_last_error());
    Check(st);
}

float Me8qIndex::Estimate(const float* query, uint32_t id) const {
    float out = 0.0F;
    const uint32_t q0 = 0;
    EstimateBatch(query, 1, &q0, &id, 1, &out);
    return out;
}

void Me8qIndex::SearchIVFFastBatch(const float* queries, size_t nq, size_t k, size_t nprobe, uint32_t* out_ids,
                                   SearchStats* stats, float* out_dists, int /*num_threads*/) const {
    if (nq == 0) return;
    if (k == 0 || nprobe == 0 || n_ == 0 || !dev_) {
        for (size_t i = 0; i < nq * k; ++i) {
            out_ids[i] = std::numeric_limits<uint32_t>::max();
            if (out_dists) out_dists[i] = std::numeric_limits<float>::infinity();
        }
        return;
    }
    std::vector<float> dists;
    float* d = out_dists;
    if (!d) {
        dists.resize(nq * k);
        d = dists.data();
    }
    SearchDevice(queries, nq, k, nprobe, nullptr, out_ids, d, stats, false, false);
}

void Me8qIndex::SearchIVFFast(const float* query, size_t k, size_t nprobe, uint32_t* out_ids, SearchStats* stats,
                              float* out_dists) const {
    SearchIVFFastBatch(query, 1, k, nprobe, out_ids, stats, out_dists, 1);
}

}  // namespace e8qlib::me8q
