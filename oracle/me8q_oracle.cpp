// me8q_oracle.cpp -- CPU restatement of e8qlib's ME8Q (mixed-rate E8Q) build
// and FastScan search.  TEST INFRASTRUCTURE ONLY: the checker the HIP path is
// compared against (tests/, bench.py's cpu_baseline leg).  Nothing in the
// product links or calls it.
//
// Numerics follow the reference as its own build compiles it (GCC, C++17,
// -O3 with FMA, E8Q_ISA=avx2): every multiply-add GCC contracts in the
// reference's scalar code and every explicit _mm256_fmadd/_fnmadd of its
// AVX2 kernels is written as std::fma here; the rest is evaluated with
// contraction off (this file is compiled with -ffp-contract=off).
//
// Parity pinning: tests/test_oracle_me8q.py compares the plan, the ME8QIDXX
// image, the fast layout (plane and refine bytes, u8 banks, gain/vpc), the
// rotated queries, the probe lists and the search results (ids, distance
// bits, candidate/survivor/bit counters) with fixtures written by the
// reference library itself (tests/golden/gen_golden.cpp).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "orc_common.h"

namespace orc::m8 {

constexpr size_t kSegDim = 64;         // me8q_plan.hpp:12
constexpr size_t kFactorBits = 64;     // me8q_plan.hpp:14
constexpr size_t kBankRunBytes = 160;  // me8q.hpp:183

struct Seg {
    size_t len = 0;
    int bits = 0;
};
using Plan = std::vector<Seg>;

size_t PaddedDim(size_t dim) { return (dim + kSegDim - 1) / kSegDim * kSegDim; }

// SAQ dynamic program (me8q_plan.cpp:45-167): cost of a run = its variance
// / 2^b, a 64-bit factor per coded run, segment cap nblocks/2 (or nblocks
// below 2 bits), a tail-drop action priced at the remaining variance, and a
// plan accepted only when 1% better than the best so far.
Plan Allocate(const float* var, size_t dim, float avg_bits, int max_bits, double extra_bits) {
    const size_t nb = dim / kSegDim;
    std::vector<double> p1(nb + 1, 0.0);
    for (size_t b = 0; b < nb; ++b) {
        double s = 0.0;
        for (size_t k = 0; k < kSegDim; ++k) s += static_cast<double>(var[b * kSegDim + k]);
        p1[b + 1] = p1[b] + s;
    }
    const double budget = static_cast<double>(avg_bits) * static_cast<double>(dim) + extra_bits +
                          static_cast<double>(kFactorBits);
    const size_t tot_bits = budget > 0.0 ? static_cast<size_t>(budget) : 0;
    const size_t units = tot_bits / kSegDim;
    const size_t max_segs = avg_bits < 2.0F ? nb : nb / 2;
    const double kInf = std::numeric_limits<double>::max();
    const size_t si = nb + 1, st = units + 1;
    std::vector<double> f((max_segs + 1) * si * st, kInf);
    std::vector<uint32_t> pred(f.size(), 0);
    auto at = [&](size_t ns, size_t i, size_t t) { return (ns * si + i) * st + t; };
    f[at(0, 0, 0)] = 0.0;
    double best = kInf;
    size_t ans_ns = 0, ans_t = 0;
    for (size_t ns = 0; ns <= max_segs; ++ns)
        for (size_t i = 0; i <= nb; ++i)
            for (size_t t = 0; t <= units; ++t) {
                const double cur = f[at(ns, i, t)];
                if (cur >= kInf) continue;
                if (i == nb) {
                    if (cur * 1.01 < best) {
                        best = cur;
                        ans_ns = ns;
                        ans_t = t;
                    }
                    continue;
                }
                if (ns == max_segs) continue;
                for (size_t j = 1; i + j <= nb; ++j) {
                    const double vs = p1[i + j] - p1[i];
                    for (int b = 1; b <= max_bits; ++b) {
                        const size_t t2 = t + static_cast<size_t>(b) * j + 1;
                        if (t2 > units) break;
                        const double v = cur + vs / static_cast<double>(int64_t{1} << b);
                        const size_t to = at(ns + 1, i + j, t2);
                        if (v < f[to]) {
                            f[to] = v;
                            pred[to] = static_cast<uint32_t>((i << 4) | static_cast<size_t>(b));
                        }
                    }
                }
                const double rest = p1[nb] - p1[i];
                const size_t to = at(ns + 1, nb, t);
                if (cur + rest < f[to]) {
                    f[to] = cur + rest;
                    pred[to] = static_cast<uint32_t>(i << 4);
                }
            }
    if (best >= kInf) return Plan{{dim, 0}};
    Plan plan;
    size_t ns = ans_ns, i = nb, t = ans_t;
    while (i > 0) {
        const uint32_t p = pred[at(ns, i, t)];
        const size_t pi = p >> 4;
        const int bits = static_cast<int>(p & 0xF);
        plan.push_back({(i - pi) * kSegDim, bits});
        --ns;
        if (bits != 0) t -= static_cast<size_t>(bits) * (i - pi) + 1;
        i = pi;
    }
    std::reverse(plan.begin(), plan.end());
    return plan;
}

// DatasetVariance (me8q_common.cpp:11-35): fp32 mean, fp64 squared
// deviations (the multiply-add is contracted).
void DatasetVariance(const float* base, size_t ld, size_t n, size_t dim, float* var) {
    std::vector<float> mean(dim, 0.0F);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < dim; ++j) mean[j] += base[i * ld + j];
    for (size_t j = 0; j < dim; ++j) mean[j] /= static_cast<float>(n);
    std::vector<double> acc(dim, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < dim; ++j) {
            const double d = static_cast<double>(base[i * ld + j]) - mean[j];
            acc[j] = std::fma(d, d, acc[j]);
        }
    for (size_t j = 0; j < dim; ++j) var[j] = static_cast<float>(acc[j] / static_cast<double>(n));
}

Plan ParsePlanSpec(const std::string& spec, size_t padded) {
    Plan plan;
    size_t pos = 0, total = 0;
    while (pos <= spec.size()) {
        const size_t comma = spec.find(',', pos);
        const std::string tok = spec.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos);
        if (!tok.empty()) {
            const size_t colon = tok.find(':');
            if (colon == std::string::npos) throw std::invalid_argument("plan_spec");
            const long len = std::atol(tok.substr(0, colon).c_str());
            const long bits = std::atol(tok.substr(colon + 1).c_str());
            if (len <= 0 || len % 64 != 0 || bits < 0 || bits > kMaxBits) throw std::invalid_argument("plan_spec");
            plan.push_back({static_cast<size_t>(len), static_cast<int>(bits)});
            total += static_cast<size_t>(len);
        }
        if (comma == std::string::npos) break;
        pos = comma + 1;
    }
    if (total != padded) throw std::invalid_argument("plan_spec coverage");
    return plan;
}

// Float level tables per rate and the branch-averaged top-bit class model
// (me8q_codec.cpp:13-44, 263-291).
struct Codec {
    std::vector<float> tab[2][kMaxBits + 1];
    float pbar[kMaxBits + 1][2] = {};
    Codec() {
        for (int b = 1; b <= kMaxBits; ++b) {
            for (int br = 0; br < 2; ++br) {
                const auto lev = BranchLevels(b, br);
                tab[br][b].assign(lev.size(), 0.0F);
                for (size_t i = 0; i < lev.size(); ++i) tab[br][b][i] = static_cast<float>(lev[i]);
            }
            const int half = 1 << (b - 1);
            for (int c = 0; c < 2; ++c) {
                double s = 0.0;
                for (int br = 0; br < 2; ++br)
                    for (int d = c * half; d < (c + 1) * half; ++d) s += static_cast<double>(tab[br][b][d]);
                pbar[b][c] = static_cast<float>(s / static_cast<double>(2 * half));
            }
        }
    }
};
const Codec& GetCodec() {
    static const Codec c;
    return c;
}

// Bit-tight block pack (me8q_codec.cpp:117-152): digits 0..6 at `bits` each,
// digit 7 without its parity bit at bits-1, the branch in the field's top bit.
void PackBlock(const uint16_t* d, int branch, int bits, uint8_t* out) {
    unsigned __int128 x = 0;
    const uint64_t m = bits >= 64 ? ~uint64_t{0} : (uint64_t{1} << bits) - 1;
    const uint64_t m7 = (uint64_t{1} << (bits - 1)) - 1;
    for (int i = 0; i < 7; ++i) x |= static_cast<unsigned __int128>(d[i] & m) << (bits * i);
    x |= static_cast<unsigned __int128>((static_cast<uint64_t>(d[7]) >> 1) & m7) << (bits * 7);
    x |= static_cast<unsigned __int128>(branch & 1) << (8 * bits - 1);
    std::memcpy(out, &x, static_cast<size_t>(bits));
}

void UnpackBlock(const uint8_t* in, int bits, uint16_t* d, uint8_t* branch) {
    unsigned __int128 x = 0;
    std::memcpy(&x, in, static_cast<size_t>(bits));
    const uint64_t m = (uint64_t{1} << bits) - 1;
    const uint64_t m7 = (uint64_t{1} << (bits - 1)) - 1;
    int par = 0;
    for (int i = 0; i < 7; ++i) {
        const uint64_t f = static_cast<uint64_t>(x >> (bits * i)) & m;
        d[i] = static_cast<uint16_t>(f);
        par ^= static_cast<int>(f & 1);
    }
    const uint64_t f7 = static_cast<uint64_t>(x >> (bits * 7)) & m7;
    d[7] = static_cast<uint16_t>((f7 << 1) | static_cast<uint64_t>(par));
    *branch = static_cast<uint8_t>(static_cast<uint64_t>(x >> (8 * bits - 1)) & 1);
}

struct Index {
    size_t n = 0, dim = 0, padded = 0, nlist = 0, bytes_per_vector = 0, route_dim = 0;
    float avg_bits = 4.0F, prune_alpha = 0.0F, ortho_alpha = 1.0F;
    int max_bits = kMaxBits;
    bool drop_energy = true;
    uint64_t rot_seed = 0x5EED1234ULL;
    Plan plan;
    std::vector<Rotator> rot;
    std::vector<float> centroids, cent_rot, centroid_n2;
    std::vector<uint32_t> ids;
    std::vector<uint64_t> list_begin, byte_begin;
    std::vector<uint8_t> codes;
    // Shared layout (one plan for every list): coded runs only.
    size_t nrun = 0, plane_cols = 0, refine_stride = 0, stride = 0;
    std::vector<uint32_t> run_rate, run_len, run_dim_off, run_blk0, run_nblk, run_code_off, refine_run_off;
    // Fast layout per list.
    struct Fast {
        size_t cnt = 0, batches = 0;
        std::vector<uint8_t> plane, refine;
        std::vector<float> gain, vpc;
        std::vector<double> rc_n2;
    };
    std::vector<Fast> fast;
    size_t rec_bytes() const { return plane_cols * 32 + nrun * kBankRunBytes; }

    void Layout() {
        nrun = 0;
        run_rate.clear(); run_len.clear(); run_dim_off.clear(); run_blk0.clear(); run_nblk.clear();
        run_code_off.clear(); refine_run_off.clear();
        size_t nseg = 0;
        for (const Seg& s : plan) nseg += s.bits > 0 ? 1 : 0;
        size_t code = nseg * 8, dimo = 0, blk = 0;
        for (const Seg& s : plan) {
            if (s.bits > 0) {
                run_rate.push_back(static_cast<uint32_t>(s.bits));
                run_len.push_back(static_cast<uint32_t>(s.len));
                run_dim_off.push_back(static_cast<uint32_t>(dimo));
                run_blk0.push_back(static_cast<uint32_t>(blk));
                run_nblk.push_back(static_cast<uint32_t>(s.len / 8));
                run_code_off.push_back(static_cast<uint32_t>(code));
                code += static_cast<size_t>(s.bits) * s.len / 8;
                blk += s.len / 8;
                ++nrun;
            }
            dimo += s.len;
        }
        stride = code;
        plane_cols = blk;
        size_t off = nrun * 8;
        for (size_t r = 0; r < nrun; ++r) {
            refine_run_off.push_back(static_cast<uint32_t>(off));
            if (run_rate[r] >= 2) off += static_cast<size_t>(run_rate[r] - 1) * run_len[r] / 8;
        }
        refine_stride = off;
    }

    void BuildRotators() {
        rot.clear();
        uint64_t s = rot_seed;
        for (const Seg& g : plan) {
            if (g.bits == 0) continue;
            s += 0x9E3779B97F4A7C15ULL;
            rot.emplace_back(g.len, s);
        }
    }

    // PrepareCentroidsRotated (me8q_search.cpp:45-72) and CentroidNorm2
    // (ivf_common.hpp:43-54, contracted).
    void RotateCentroids() {
        cent_rot.assign(nlist * padded, 0.0F);
        for (size_t c = 0; c < nlist; ++c)
            for (size_t r = 0; r < nrun; ++r)
                rot[r].Rotate(&centroids[c * padded + run_dim_off[r]], &cent_rot[c * padded + run_dim_off[r]]);
        centroid_n2.assign(nlist, 0.0F);
        for (size_t c = 0; c < nlist; ++c) {
            double s = 0.0;
            for (size_t d = 0; d < padded; ++d) {
                const double v = cent_rot[c * padded + d];
                s = std::fma(v, v, s);
            }
            centroid_n2[c] = static_cast<float>(s);
        }
        route_dim = padded;
        for (auto it = plan.rbegin(); it != plan.rend(); ++it) {
            if (it->bits != 0) break;
            route_dim -= it->len;
        }
    }

    void BuildFast();
};

// Build (me8q_build.cpp:43-311): per-run rotation of the residual, per-run
// encode at FixedScale(rate), bit-tight pack, fits coef = |r|^2/<r,lev> and
// h2 = |r|^2 + 2 coef <c',lev> (+ the dropped tail's energy on the last run).
void Build(Index& ix, const float* base, const float* cent, const uint32_t* cid, size_t n, size_t dim, size_t nlist,
           float avg_bits, uint64_t seed, float prune_alpha, const std::string& plan_spec, const float* variance,
           bool drop_energy) {
    ix = Index();
    ix.n = n;
    ix.dim = dim;
    ix.nlist = nlist;
    ix.padded = PaddedDim(dim);
    ix.avg_bits = avg_bits;
    ix.prune_alpha = prune_alpha;
    ix.ortho_alpha = prune_alpha > 0.0F ? prune_alpha : 1.0F;
    ix.drop_energy = drop_energy;
    if (seed != 0) ix.rot_seed = seed;
    const size_t P = ix.padded;
    ix.centroids.assign(nlist * P, 0.0F);
    for (size_t c = 0; c < nlist; ++c) std::memcpy(&ix.centroids[c * P], cent + c * dim, dim * 4);
    ix.list_begin.assign(nlist + 1, 0);
    for (size_t i = 0; i < n; ++i) ++ix.list_begin[cid[i] + 1];
    for (size_t c = 0; c < nlist; ++c) ix.list_begin[c + 1] += ix.list_begin[c];
    ix.ids.resize(n);
    std::vector<uint64_t> cur(ix.list_begin.begin(), ix.list_begin.end() - 1);
    for (size_t i = 0; i < n; ++i) ix.ids[cur[cid[i]]++] = static_cast<uint32_t>(i);
    if (!plan_spec.empty()) {
        ix.plan = ParsePlanSpec(plan_spec, P);
    } else {
        std::vector<float> var(P, 0.0F);
        if (variance) std::memcpy(var.data(), variance, dim * 4);
        else DatasetVariance(base, dim, n, dim, var.data());
        ix.plan = Allocate(var.data(), P, avg_bits, ix.max_bits, 0.0);
    }
    ix.Layout();
    ix.bytes_per_vector = ix.stride;
    ix.byte_begin.assign(nlist + 1, 0);
    for (size_t c = 0; c < nlist; ++c)
        ix.byte_begin[c + 1] = ix.byte_begin[c] + (ix.list_begin[c + 1] - ix.list_begin[c]) * ix.stride;
    ix.codes.assign(ix.byte_begin[nlist], 0);
    ix.BuildRotators();
    ix.RotateCentroids();
    std::vector<uint32_t> rank(n, 0);
    for (size_t c = 0; c < nlist; ++c)
        for (uint64_t k = ix.list_begin[c]; k < ix.list_begin[c + 1]; ++k)
            rank[ix.ids[k]] = static_cast<uint32_t>(k - ix.list_begin[c]);
    const Codec& cd = GetCodec();
    std::vector<float> res(P), rr(P), unit;
    std::vector<uint16_t> dig;
    std::vector<uint8_t> brs;
    for (size_t i = 0; i < n; ++i) {
        const size_t c = cid[i];
        std::fill(res.begin(), res.end(), 0.0F);
        for (size_t j = 0; j < dim; ++j) res[j] = base[i * dim + j] - ix.centroids[c * P + j];
        for (size_t r = 0; r < ix.nrun; ++r) ix.rot[r].Rotate(&res[ix.run_dim_off[r]], &rr[ix.run_dim_off[r]]);
        uint8_t* row = &ix.codes[ix.byte_begin[c] + rank[i] * ix.stride];
        double drop = 0.0;
        if (drop_energy) {
            size_t off = 0;
            for (const Seg& g : ix.plan) {
                if (g.bits == 0)
                    for (size_t j = 0; j < g.len; ++j) {
                        const double v = res[off + j];
                        drop = std::fma(v, v, drop);
                    }
                off += g.len;
            }
        }
        for (size_t r = 0; r < ix.nrun; ++r) {
            const int len = static_cast<int>(ix.run_len[r]), bits = static_cast<int>(ix.run_rate[r]);
            const float* x = &rr[ix.run_dim_off[r]];
            double n2 = 0.0;
            for (int j = 0; j < len; ++j) n2 = std::fma(static_cast<double>(x[j]), static_cast<double>(x[j]), n2);
            const double nrm = std::sqrt(n2);
            dig.assign(static_cast<size_t>(len), 0);
            brs.assign(static_cast<size_t>(len / 8), 0);
            if (nrm > 0.0) {
                unit.resize(static_cast<size_t>(len));
                const double inv = 1.0 / nrm;
                for (int j = 0; j < len; ++j) unit[j] = static_cast<float>(static_cast<double>(x[j]) * inv);
                EncodeVector(unit.data(), len, bits, FixedScale(bits), dig.data(), brs.data());
            }
            uint8_t* dst = row + ix.run_code_off[r];
            for (int b = 0; b < len / 8; ++b) PackBlock(&dig[b * 8], brs[b], bits, dst + b * bits);
            const float* rc = &ix.cent_rot[c * P + ix.run_dim_off[r]];
            double dr = 0.0, dc = 0.0;
            for (int b = 0; b < len / 8; ++b) {
                const float* lev = cd.tab[brs[b]][bits].data();
                for (int t = 0; t < 8; ++t) {
                    const double lv = lev[dig[b * 8 + t]];
                    dr = std::fma(static_cast<double>(x[b * 8 + t]), lv, dr);
                    dc = std::fma(static_cast<double>(rc[b * 8 + t]), lv, dc);
                }
            }
            double coef = dr > 1e-12 ? n2 / dr : 0.0;
            if (!std::isfinite(coef)) coef = 0.0;
            double h2 = std::fma(2.0 * coef, dc, n2);
            if (r + 1 == ix.nrun) h2 += drop;
            float fac[2] = {static_cast<float>(coef), static_cast<float>(std::isfinite(h2) ? h2 : 0.0)};
            std::memcpy(row + r * 8, fac, 8);
        }
    }
    ix.BuildFast();
}

// BuildFastLayout (me8q_fast.cpp:37-309): plane bytes (digit top bits,
// vector-in-batch interleaved), (rate-1)-bit refine fields, ZPR fit per
// (vector, run) from the stored code, u8 banks per batch and run.
void Index::BuildFast() {
    const Codec& cd = GetCodec();
    fast.assign(nlist, Fast{});
    std::vector<uint16_t> dig;
    std::vector<uint8_t> brs;
    for (size_t c = 0; c < nlist; ++c) {
        Fast& f = fast[c];
        f.cnt = list_begin[c + 1] - list_begin[c];
        f.batches = (f.cnt + 31) / 32;
        f.rc_n2.assign(nrun, 0.0);
        for (size_t r = 0; r < nrun; ++r) {
            double a = 0.0;
            for (uint32_t t = 0; t < run_len[r]; ++t) {
                const double v = cent_rot[c * padded + run_dim_off[r] + t];
                a = std::fma(v, v, a);
            }
            f.rc_n2[r] = a;
        }
        std::vector<uint8_t> prow(f.cnt * plane_cols, 0);
        f.refine.assign(f.cnt * refine_stride, 0);
        std::vector<float> coef_v(nrun * f.cnt), h2_v(nrun * f.cnt), perp(nrun * f.cnt, 0.0F);
        f.gain.assign(nrun * f.cnt, 0.0F);
        f.vpc.assign(nrun * f.cnt, 0.0F);
        for (size_t i = 0; i < f.cnt; ++i) {
            const uint8_t* row = &codes[byte_begin[c] + i * stride];
            uint8_t* rrow = &f.refine[i * refine_stride];
            for (size_t r = 0; r < nrun; ++r) {
                std::memcpy(rrow + r * 8, row + r * 8, 8);
                float fac[2];
                std::memcpy(fac, row + r * 8, 8);
                coef_v[r * f.cnt + i] = fac[0];
                h2_v[r * f.cnt + i] = fac[1];
            }
            for (size_t r = 0; r < nrun; ++r) {
                const int bits = static_cast<int>(run_rate[r]), len = static_cast<int>(run_len[r]), nblk = len / 8;
                dig.assign(static_cast<size_t>(len), 0);
                brs.assign(static_cast<size_t>(nblk), 0);
                for (int b = 0; b < nblk; ++b) UnpackBlock(row + run_code_off[r] + b * bits, bits, &dig[b * 8], &brs[b]);
                if (bits >= 2)
                    for (int b = 0; b < nblk; ++b)
                        PackBlock(&dig[b * 8], brs[b], bits - 1, rrow + refine_run_off[r] + b * (bits - 1));
                uint8_t* pb = &prow[i * plane_cols + run_blk0[r]];
                for (int b = 0; b < nblk; ++b) {
                    uint32_t byte = 0;
                    for (int t = 0; t < 8; ++t) {
                        const int top = bits == 1 && t == 7 ? (brs[b] & 1) : ((dig[b * 8 + t] >> (bits - 1)) & 1);
                        byte |= static_cast<uint32_t>(top) << (7 - t);
                    }
                    pb[b] = static_cast<uint8_t>(byte);
                }
                double dew = 0.0, n2w = 0.0, n2e = 0.0, dce = 0.0, dcw = 0.0;
                const float* rc = &cent_rot[c * padded + run_dim_off[r]];
                for (int b = 0; b < nblk; ++b) {
                    const float* lev = cd.tab[brs[b]][bits].data();
                    for (int t = 0; t < 8; ++t) {
                        const int d = dig[b * 8 + t];
                        const double ve = lev[d];
                        const double wc = bits == 1 && t == 7 ? 0.0 : static_cast<double>(cd.pbar[bits][d >> (bits - 1)]);
                        const double cc = rc[b * 8 + t];
                        dew = std::fma(ve, wc, dew);
                        n2w = std::fma(wc, wc, n2w);
                        n2e = std::fma(ve, ve, n2e);
                        dce = std::fma(cc, ve, dce);
                        dcw = std::fma(cc, wc, dcw);
                    }
                }
                const double coef = coef_v[r * f.cnt + i], h2 = h2_v[r * f.cnt + i];
                const double rn2 = std::max(std::fma(-(2.0 * coef), dce, h2), 0.0);
                const double rn = std::sqrt(rn2);
                const double slm1 = std::sqrt(static_cast<double>(std::max<uint32_t>(run_len[r] > 0 ? run_len[r] - 1 : 1, 1)));
                const double ac = std::fabs(coef);
                const size_t o = r * f.cnt + i;
                if (n2e > 1e-30 && ac > 1e-30) {
                    const double ven = std::sqrt(n2e);
                    const double n2ws = std::max(n2w, 1e-30);
                    const double duw = dew / ven;
                    const double eps = 1e-12 * std::sqrt(n2ws);
                    double wx, nrx;
                    if (std::abs(duw) > eps) {
                        wx = 2.0 * rn / duw;
                        const double c4 = 4.0 * coef;
                        const double rx2 = std::fma(-(c4 * wx), dew, std::fma(c4 * coef, n2e, (wx * wx) * n2w));
                        nrx = std::sqrt(std::max(0.0, rx2));
                    } else {
                        const double g = dew / n2ws;
                        wx = 2.0 * coef * g;
                        const double vp2 = std::max(std::fma(-(g * g), n2w, n2e), 0.0);
                        nrx = 2.0 * ac * std::sqrt(vp2);
                    }
                    f.gain[o] = static_cast<float>(wx / (2.0 * ac));
                    f.vpc[o] = static_cast<float>(std::fma(-wx, dcw, h2 - rn2) / (2.0 * ac));
                    perp[o] = static_cast<float>((nrx / slm1) / (2.0 * ac));
                } else if (n2w > 1e-30) {
                    const double g = dew / n2w;
                    const double vp2 = std::max(std::fma(-(g * g), n2w, n2e), 0.0);
                    f.gain[o] = static_cast<float>(g);
                    f.vpc[o] = static_cast<float>(std::fma(-g, dcw, dce));
                    perp[o] = static_cast<float>(std::sqrt(vp2) / slm1);
                }
            }
        }
        const size_t rec = rec_bytes();
        f.plane.assign(f.batches * rec, 0);
        for (size_t b = 0; b < f.batches; ++b) {
            uint8_t* rb = &f.plane[b * rec];
            for (size_t col = 0; col < plane_cols; ++col)
                for (size_t l = 0; l < 32; ++l) {
                    const size_t row = b * 32 + l;
                    rb[col * 32 + l] = row < f.cnt ? prow[row * plane_cols + col] : 0;
                }
            const size_t hi = std::min(f.cnt, b * 32 + 32);
            for (size_t r = 0; r < nrun; ++r) {
                uint8_t* bp = rb + plane_cols * 32 + r * kBankRunBytes;
                float vals[32][3], vmin[3] = {0, 0, 0}, vmax[3] = {0, 0, 0};
                for (size_t row = b * 32; row < hi; ++row) {
                    const size_t l = row - b * 32, o = r * f.cnt + row;
                    const float cf = coef_v[o], h2 = h2_v[o];
                    vals[l][0] = std::fma(-(2.0F * cf), f.vpc[o], h2);
                    vals[l][1] = 2.0F * cf * f.gain[o];
                    vals[l][2] = perp[o] * 2.0F * std::fabs(cf);
                    for (int k = 0; k < 3; ++k) {
                        if (l == 0) {
                            vmin[k] = vmax[k] = vals[l][k];
                        } else {
                            vmin[k] = std::min(vmin[k], vals[l][k]);
                            vmax[k] = std::max(vmax[k], vals[l][k]);
                        }
                    }
                }
                float sc[3];
                for (int k = 0; k < 3; ++k) sc[k] = vmax[k] > vmin[k] ? (vmax[k] - vmin[k]) / 255.0F : 1.0F;
                for (size_t row = b * 32; row < hi; ++row) {
                    const size_t l = row - b * 32;
                    bp[l] = static_cast<uint8_t>(std::clamp(static_cast<int>((vals[l][0] - vmin[0]) / sc[0]), 0, 255));
                    bp[32 + l] = static_cast<uint8_t>(
                        std::clamp(static_cast<int>(std::lround((vals[l][1] - vmin[1]) / sc[1])), 0, 255));
                    bp[64 + l] = static_cast<uint8_t>(
                        std::clamp(static_cast<int>(std::ceil((vals[l][2] - vmin[2]) / sc[2])), 0, 255));
                }
                const float scl[8] = {vmin[0], sc[0], vmin[1], sc[1], vmin[2], sc[2], 0.0F, 0.0F};
                std::memcpy(bp + 128, scl, 32);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Search (me8q_fast.cpp:380-1108, SearchIVFFast / ScanPreRouted)

struct Stats {
    uint64_t candidates = 0, survivors = 0, bits_scored = 0, bits_total = 0, pruned = 0;
};

// Exact run score (me8q_score_avx2.cpp:292-425): level[br][digit] over the
// run's blocks, component i of even blocks in acc[i], of odd blocks in
// acc2[i] (FMA chains), then the AVX horizontal sum of acc + acc2.
static float ScoreRun(const Index& ix, size_t r, const uint8_t* rrow, const uint8_t* pbatch, size_t lane,
                      const float* q) {
    const Codec& cd = GetCodec();
    const int bits = static_cast<int>(ix.run_rate[r]), nblk = static_cast<int>(ix.run_nblk[r]);
    const int T = bits - 1;
    float acc[8] = {}, acc2[8] = {};
    for (int b = 0; b < nblk; ++b) {
        const uint8_t plane = pbatch[(ix.run_blk0[r] + static_cast<size_t>(b)) * 32 + lane];
        uint16_t d[8];
        int br;
        if (bits == 1) {
            int par = 0;
            for (int i = 0; i < 7; ++i) {
                d[i] = static_cast<uint16_t>((plane >> (7 - i)) & 1);
                par ^= d[i];
            }
            d[7] = static_cast<uint16_t>(par);
            br = plane & 1;
        } else {
            // The refine field is the block packed at rate-1 bits (digit 7's
            // LSB is the parity shell, the branch sits in the top bit).
            uint8_t b8;
            UnpackBlock(rrow + ix.refine_run_off[r] + static_cast<size_t>(b) * T, T, d, &b8);
            br = b8;
            for (int i = 0; i < 8; ++i) d[i] = static_cast<uint16_t>(d[i] | (((plane >> (7 - i)) & 1) << T));
        }
        const float* lev = cd.tab[br][bits].data();
        const float* qb = q + static_cast<size_t>(b) * 8;
        for (int i = 0; i < 8; ++i) {
            if (b & 1) acc2[i] = std::fma(lev[d[i]], qb[i], acc2[i]);
            else acc[i] = std::fma(lev[d[i]], qb[i], acc[i]);
        }
    }
    float s[8];
    for (int i = 0; i < 8; ++i) s[i] = acc[i] + acc2[i];
    return Hsum8(s);
}

// PushCandidate (me8q_search.cpp:121-134): sorted by distance only, a new
// item goes before equal ones, and one equal to a full list's worst is
// rejected.
static void Push(std::vector<std::pair<float, uint32_t>>& h, size_t k, float d, uint32_t id) {
    if (h.size() >= k && !(d < h.back().first)) return;
    if (h.size() >= k) h.pop_back();
    auto it = std::lower_bound(h.begin(), h.end(), d,
                               [](const std::pair<float, uint32_t>& a, float v) { return a.first < v; });
    h.insert(it, {d, id});
}

struct QueryPrep {
    std::vector<float> q;            // padded, per-run rotated
    std::vector<float> lut_f;        // plane_cols * 64
    std::vector<uint8_t> lut;        // plane_cols * 64
    float lut_min = 0.0F, lut_delta = 1.0F;
    std::vector<double> run_qn2, lut_min_x, qerr, score_max;
    uint64_t bits_row = 0;
};

void PrepareQuery(const Index& ix, const float* query, QueryPrep& p) {
    const Codec& cd = GetCodec();
    std::vector<float> pad(ix.padded, 0.0F);
    std::memcpy(pad.data(), query, ix.dim * 4);
    p.q.assign(ix.padded, 0.0F);
    for (size_t r = 0; r < ix.nrun; ++r) ix.rot[r].Rotate(&pad[ix.run_dim_off[r]], &p.q[ix.run_dim_off[r]]);
    // Class-mean LUT (me8q_fast.cpp:627-667): per block, dims 4..7 under the
    // low nibble at [0,32), dims 0..3 under the high nibble at [32,64), each
    // 16-entry quarter twice; a rate-1 block prices dim 7 (the branch) at 0.
    p.lut_f.assign(ix.plane_cols * 64, 0.0F);
    for (size_t r = 0; r < ix.nrun; ++r) {
        const int rate = static_cast<int>(ix.run_rate[r]);
        const float p0 = cd.pbar[rate][0], p1 = cd.pbar[rate][1];
        const float p70 = rate == 1 ? 0.0F : p0, p71 = rate == 1 ? 0.0F : p1;
        for (uint32_t b = 0; b < ix.run_nblk[r]; ++b) {
            const float* qb = &p.q[ix.run_dim_off[r] + b * 8];
            float* gl = &p.lut_f[(ix.run_blk0[r] + b) * 64];
            float* gh = gl + 32;
            for (int v = 0; v < 16; ++v) {
                const float a4 = (v >> 3) & 1 ? p1 : p0, a5 = (v >> 2) & 1 ? p1 : p0, a6 = (v >> 1) & 1 ? p1 : p0;
                const float a7 = v & 1 ? p71 : p70;
                // `qb4*a4 + qb5*a5 + ...`: GCC fuses the first product of
                // each sum into the add, the second stays a plain product.
                gl[v] = std::fma(qb[7], a7, std::fma(qb[6], a6, std::fma(qb[4], a4, qb[5] * a5)));
                const float a0 = (v >> 3) & 1 ? p1 : p0, a1 = (v >> 2) & 1 ? p1 : p0, a2 = (v >> 1) & 1 ? p1 : p0;
                const float a3 = v & 1 ? p1 : p0;
                gh[v] = std::fma(qb[3], a3, std::fma(qb[2], a2, std::fma(qb[0], a0, qb[1] * a1)));
            }
            for (int v = 0; v < 16; ++v) {
                gl[16 + v] = gl[v];
                gh[16 + v] = gh[v];
            }
        }
    }
    // QuantizeLut (fastscan_avx2.cpp:8-51): min/max, delta = range/255,
    // u8 = clamp(trunc((x - lo)/delta + 0.5)).
    const size_t len = p.lut_f.size();
    p.lut.assign(len, 0);
    if (len > 0) {
        float lo = p.lut_f[0], hi = p.lut_f[0];
        for (size_t i = 0; i < len; ++i) {
            lo = std::min(lo, p.lut_f[i]);
            hi = std::max(hi, p.lut_f[i]);
        }
        const float delta = hi > lo ? (hi - lo) / 255.0F : 1.0F;
        for (size_t i = 0; i < len; ++i) {
            const float t = (p.lut_f[i] - lo) / delta + 0.5F;
            const int v = static_cast<int>(t);
            p.lut[i] = static_cast<uint8_t>(std::clamp(v, 0, 255));
        }
        p.lut_min = lo;
        p.lut_delta = delta;
    }
    p.run_qn2.assign(ix.nrun, 0.0);
    p.lut_min_x.assign(ix.nrun, 0.0);
    p.qerr.assign(ix.nrun, 0.0);
    p.score_max.assign(ix.nrun, 0.0);
    p.bits_row = 0;
    for (size_t r = 0; r < ix.nrun; ++r) {
        double a = 0.0;
        for (uint32_t t = 0; t < ix.run_len[r]; ++t) {
            const double v = p.q[ix.run_dim_off[r] + t];
            a = std::fma(v, v, a);
        }
        p.run_qn2[r] = a;
        p.bits_row += static_cast<uint64_t>(ix.run_rate[r]) * ix.run_len[r];
        const double nb2 = 2.0 * static_cast<double>(ix.run_nblk[r]);
        p.lut_min_x[r] = static_cast<double>(p.lut_min) * nb2;
        const float sq = static_cast<float>(std::sqrt(static_cast<double>(std::max<uint32_t>(ix.run_nblk[r], 1))));
        p.qerr[r] = static_cast<double>(2.0F * sq * p.lut_delta);
        const double hiv = std::fma(2.0 * static_cast<double>(ix.run_nblk[r]) * 255.0,
                                    static_cast<double>(p.lut_delta), p.lut_min_x[r]);
        p.score_max[r] = std::max(std::abs(p.lut_min_x[r]), std::abs(hiv));
    }
}

// RouteL2 over the coded prefix (ivf_route.cpp:9-27), AVX2 dot order.
void Route(const Index& ix, const float* q, size_t nprobe, std::vector<float>& pd, std::vector<uint32_t>& probe) {
    const size_t L = ix.nlist, D = ix.route_dim;
    const float qn2 = QueryNorm2(q, D);
    std::vector<uint64_t> keys(L);
    std::vector<float> dist(L);
    for (size_t c = 0; c < L; ++c) {
        const float dot = RouteDot(q, &ix.cent_rot[c * ix.padded], D);
        float d2 = std::fma(-dot, 2.0F, ix.centroid_n2[c] + qn2);
        if (d2 < 0.0F) d2 = 0.0F;
        dist[c] = d2;
        uint32_t bv;
        std::memcpy(&bv, &d2, 4);
        keys[c] = (static_cast<uint64_t>(bv) << 32) | c;
    }
    const size_t take = std::min(nprobe, L);
    std::partial_sort(keys.begin(), keys.begin() + static_cast<long>(take), keys.end());
    probe.resize(take);
    pd.resize(take);
    for (size_t i = 0; i < take; ++i) {
        probe[i] = static_cast<uint32_t>(keys[i]);
        pd[i] = dist[probe[i]];
    }
}

// lambda_override: the union-bound quantile to use instead of the index's
// (OrthoQuantile of its prune_alpha); NaN keeps the index's.  +inf prunes
// nothing: every probed vector is scored and the result is the exact
// estimator top-k over the probed lists.
void Search(const Index& ix, const float* query, size_t k, size_t nprobe, uint32_t* out_ids, float* out_d,
            Stats* st, float lambda_override = std::numeric_limits<float>::quiet_NaN()) {
    const float inf = std::numeric_limits<float>::infinity();
    for (size_t i = 0; i < k; ++i) {
        out_ids[i] = UINT32_MAX;
        out_d[i] = inf;
    }
    if (k == 0 || nprobe == 0 || ix.n == 0) return;
    QueryPrep p;
    PrepareQuery(ix, query, p);
    std::vector<float> pd;
    std::vector<uint32_t> probe;
    Route(ix, p.q.data(), nprobe, pd, probe);
    const float lambda = std::isnan(lambda_override)
                             ? OrthoQuantile(static_cast<double>(ix.ortho_alpha), std::max<size_t>(4 * k, 32))
                             : lambda_override;
    const double dl = lambda;
    std::vector<std::pair<float, uint32_t>> heap;
    Stats s;
    const size_t R = ix.nrun;
    std::vector<double> sqd2(R);
    std::vector<int32_t> acc(32 * (R > 0 ? R : 1));
    std::vector<float> l_dc(32), l_pool(32), l_qa(32), l_sc(32 * (R > 0 ? R : 1)), l_tr(32 * (R > 0 ? R : 1));
    auto distk_f = [&]() { return heap.size() < k ? inf : heap.back().first; };
    for (size_t pi = 0; pi < probe.size(); ++pi) {
        const size_t c = probe[pi];
        const Index::Fast& f = ix.fast[c];
        const double d2 = pd[pi];
        if (R == 0) {
            for (size_t i = 0; i < f.cnt; ++i) Push(heap, k, static_cast<float>(d2), ix.ids[ix.list_begin[c] + i]);
            s.candidates += f.cnt;
            s.survivors += f.cnt;
            continue;
        }
        if (f.cnt == 0) continue;
        for (size_t r = 0; r < R; ++r) {
            double dot = 0.0;
            for (uint32_t t = 0; t < ix.run_len[r]; ++t)
                dot = std::fma(static_cast<double>(p.q[ix.run_dim_off[r] + t]),
                               static_cast<double>(ix.cent_rot[c * ix.padded + ix.run_dim_off[r] + t]), dot);
            const double d2r = std::fma(-2.0, dot, p.run_qn2[r] + f.rc_n2[r]);
            sqd2[r] = std::sqrt(std::max(d2r, 0.0));
        }
        const size_t rec = ix.rec_bytes();
        for (size_t b = 0; b < f.batches; ++b) {
            const uint8_t* pb = &f.plane[b * rec];
            const size_t cntb = std::min<size_t>(32, f.cnt - b * 32);
            for (size_t r = 0; r < R; ++r)
                CoarseAccumulate(pb + ix.run_blk0[r] * 32, &p.lut[ix.run_blk0[r] * 64], ix.run_nblk[r], &acc[r * 32]);
            // AVX2 coarse prune over whole groups of 8 lanes
            // (me8q_prune_avx2.cpp:9-98, FillPruneRunScalars me8q_prune.hpp:40-62).
            std::vector<uint32_t> keep;
            size_t covered = 0;
            {
                const float distk = distk_f();
                if (cntb >= 8) {
                    const float sqk = distk < inf ? std::sqrt(std::max(0.0F, distk)) : inf;
                    std::vector<float> rh(R), qe(R), wp(R), lm(R);
                    for (size_t r = 0; r < R; ++r) {
                        const float a = static_cast<float>(sqd2[r]);
                        rh[r] = sqk < a ? sqk : a;
                        qe[r] = static_cast<float>(p.qerr[r]);
                        lm[r] = static_cast<float>(p.lut_min_x[r]);
                        float sc[8];
                        std::memcpy(sc, pb + ix.plane_cols * 32 + r * kBankRunBytes + 128, 32);
                        wp[r] = static_cast<float>(0.5 * static_cast<double>(sc[3]) * (p.score_max[r] + p.qerr[r]));
                    }
                    covered = cntb / 8 * 8;
                    for (size_t l = 0; l < covered; ++l) {
                        float dc = static_cast<float>(d2), pool = 0.0F, qa = 0.0F, wpad = 0.0F;
                        for (size_t r = 0; r < R; ++r) {
                            const uint8_t* bp = pb + ix.plane_cols * 32 + r * kBankRunBytes;
                            float sc[8];
                            std::memcpy(sc, bp + 128, 32);
                            const float s_c = std::fma(static_cast<float>(acc[r * 32 + l]), p.lut_delta, lm[r]);
                            const float bq = std::fma(static_cast<float>(bp[l]), sc[1], sc[0]);
                            const float wq = std::fma(static_cast<float>(bp[32 + l]), sc[3], sc[2]);
                            const float pq = std::fma(static_cast<float>(bp[64 + l]), sc[5], sc[4]);
                            dc = std::fma(-wq, s_c, dc);
                            dc = dc + bq;
                            const float t = pq * rh[r];
                            l_sc[r * 32 + l] = s_c;
                            l_tr[r * 32 + l] = t;
                            pool = std::fma(t, t, pool);
                            qa = std::fma(std::fabs(wq), qe[r], qa);
                            wpad = wpad + wp[r];
                        }
                        const float qw = qa + wpad;
                        const float slack = std::fma(lambda, std::sqrt(pool), qw);
                        l_dc[l] = dc;
                        l_pool[l] = pool;
                        l_qa[l] = qw;
                        if (dc - slack >= distk) {
                            s.candidates += 1;
                            s.bits_total += p.bits_row;
                        } else {
                            keep.push_back(static_cast<uint32_t>(l));
                        }
                    }
                }
                for (size_t l = covered; l < cntb; ++l) keep.push_back(static_cast<uint32_t>(l));
            }
            // Lane pass in order, refined four at a time with a live heap.
            struct Staged {
                size_t lane, row;
                uint32_t id;
                double pool, qadd;
                std::vector<double> sc, tr;
            };
            std::vector<Staged> stg;
            auto flush4 = [&]() {
                const size_t n = stg.size();
                if (n == 0) return;
                std::vector<size_t> order(R);
                for (size_t r = 0; r < R; ++r) order[r] = r;
                if (R > 1) {
                    std::vector<double> mean(R, 0.0);
                    for (size_t j = 0; j < n; ++j)
                        for (size_t r = 0; r < R; ++r) mean[r] = std::fma(stg[j].tr[r], stg[j].tr[r], mean[r]);
                    for (size_t r = 0; r < R; ++r) mean[r] /= static_cast<double>(std::max<uint32_t>(1, ix.run_nblk[r]));
                    // std::sort of <= 16 elements is GCC's insertion sort:
                    // stable for equal means.
                    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b2) { return mean[a] > mean[b2]; });
                }
                std::vector<double> dist(n), rem(n);
                std::vector<char> died(n, 0);
                std::vector<size_t> live;
                for (size_t j = 0; j < n; ++j) {
                    live.push_back(j);
                    rem[j] = stg[j].pool;
                    const uint8_t* rrow = &f.refine[stg[j].row * ix.refine_stride];
                    if (R == 1) {
                        float fac[2];
                        std::memcpy(fac, rrow, 8);
                        dist[j] = d2 + static_cast<double>(fac[1]);
                    } else {
                        double dc = d2;
                        for (size_t r = 0; r < R; ++r) {
                            float fac[2];
                            std::memcpy(fac, rrow + r * 8, 8);
                            const double coef = fac[0];
                            const double g = f.gain[r * f.cnt + stg[j].row];
                            const double v = f.vpc[r * f.cnt + stg[j].row];
                            dc += std::fma(-(2.0 * coef), std::fma(g, stg[j].sc[r], v), static_cast<double>(fac[1]));
                        }
                        dist[j] = dc;
                    }
                }
                const double distk = heap.size() < k ? std::numeric_limits<double>::infinity()
                                                     : static_cast<double>(heap.back().first);
                for (size_t oi = 0; oi < R && !live.empty(); ++oi) {
                    const size_t r = order[oi];
                    std::vector<size_t> nl;
                    for (size_t j : live) {
                        const uint8_t* rrow = &f.refine[stg[j].row * ix.refine_stride];
                        const double se = ScoreRun(ix, r, rrow, pb, stg[j].lane, &p.q[ix.run_dim_off[r]]);
                        float fac[2];
                        std::memcpy(fac, rrow + r * 8, 8);
                        const double coef = fac[0];
                        if (R == 1) {
                            dist[j] = std::fma(-(2.0 * coef), se, dist[j]);
                        } else {
                            const double g = f.gain[r * f.cnt + stg[j].row];
                            const double v = f.vpc[r * f.cnt + stg[j].row];
                            dist[j] = std::fma(2.0 * coef, std::fma(g, stg[j].sc[r], v) - se, dist[j]);
                        }
                        const double t = stg[j].tr[r];
                        rem[j] = std::fma(-t, t, rem[j]);
                        s.bits_scored += static_cast<uint64_t>(ix.run_rate[r]) * ix.run_len[r];
                        if (std::fma(-dl, std::sqrt(std::max(rem[j], 0.0)), dist[j]) - stg[j].qadd >= distk) died[j] = 1;
                        else nl.push_back(j);
                    }
                    live.swap(nl);
                }
                for (size_t j = 0; j < n; ++j) {
                    if (died[j]) continue;
                    ++s.survivors;
                    Push(heap, k, static_cast<float>(dist[j]), stg[j].id);
                }
                stg.clear();
            };
            for (uint32_t lane : keep) {
                const size_t row = b * 32 + lane;
                const double distk = heap.size() < k ? std::numeric_limits<double>::infinity()
                                                     : static_cast<double>(heap.back().first);
                s.candidates += 1;
                s.bits_total += p.bits_row;
                Staged sg;
                sg.lane = lane;
                sg.row = row;
                sg.id = ix.ids[ix.list_begin[c] + row];
                sg.sc.assign(R, 0.0);
                sg.tr.assign(R, 0.0);
                double dcq = d2, pool2 = 0.0, qadd = 0.0;
                if (lane < covered) {
                    dcq = l_dc[lane];
                    pool2 = l_pool[lane];
                    qadd = l_qa[lane];
                    for (size_t r = 0; r < R; ++r) {
                        sg.sc[r] = l_sc[r * 32 + lane];
                        sg.tr[r] = l_tr[r * 32 + lane];
                    }
                } else {
                    // Scalar fp64 twin of the filter (me8q_fast.cpp:1012-1043).
                    for (size_t r = 0; r < R; ++r) {
                        const uint8_t* bp = pb + ix.plane_cols * 32 + r * kBankRunBytes;
                        float scl[8];
                        std::memcpy(scl, bp + 128, 32);
                        const double scr = std::fma(static_cast<double>(p.lut_delta),
                                                    static_cast<double>(acc[r * 32 + lane]), p.lut_min_x[r]);
                        sg.sc[r] = scr;
                        const double bq = std::fma(static_cast<double>(bp[lane]), static_cast<double>(scl[1]),
                                                   static_cast<double>(scl[0]));
                        const double wq = std::fma(static_cast<double>(bp[32 + lane]), static_cast<double>(scl[3]),
                                                   static_cast<double>(scl[2]));
                        const double pq = std::fma(static_cast<double>(bp[64 + lane]), static_cast<double>(scl[5]),
                                                   static_cast<double>(scl[4]));
                        dcq += std::fma(-wq, scr, bq);
                        const double rh = std::min(sqd2[r], distk < std::numeric_limits<double>::infinity()
                                                                ? std::sqrt(std::max(distk, 0.0))
                                                                : std::numeric_limits<double>::infinity());
                        sg.tr[r] = pq * rh;
                        pool2 = std::fma(sg.tr[r], sg.tr[r], pool2);
                        qadd = std::fma(std::abs(wq), p.qerr[r], qadd);
                        qadd = std::fma(0.5 * static_cast<double>(scl[3]), p.score_max[r] + p.qerr[r], qadd);
                    }
                }
                if (dcq - std::fma(dl, std::sqrt(pool2), qadd) >= distk) continue;
                sg.pool = pool2;
                sg.qadd = qadd;
                stg.push_back(std::move(sg));
                if (stg.size() == 4) flush4();
            }
            flush4();
        }
    }
    for (size_t i = 0; i < k && i < heap.size(); ++i) {
        out_ids[i] = heap[i].second;
        out_d[i] = heap[i].first;
    }
    s.pruned = s.candidates - s.survivors;
    if (st) {
        st->candidates += s.candidates;
        st->survivors += s.survivors;
        st->bits_scored += s.bits_scored;
        st->bits_total += s.bits_total;
        st->pruned += s.pruned;
    }
}

// ME8QIDXX image (me8q_io.cpp:29-64).
std::vector<uint8_t> Image(const Index& ix) {
    std::vector<uint8_t> out;
    auto put = [&](const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        out.insert(out.end(), b, b + n);
    };
    put("ME8QIDXX", 8);
    const uint64_t hdr[5] = {ix.n, ix.dim, ix.padded, ix.nlist, ix.bytes_per_vector};
    put(hdr, 40);
    put(&ix.avg_bits, 4);
    int32_t flags[4] = {0, ix.max_bits, ix.drop_energy ? 1 : 0, 0};
    std::memcpy(&flags[3], &ix.prune_alpha, 4);
    put(flags, 16);
    const uint64_t np = ix.plan.size();
    put(&np, 8);
    for (const Seg& g : ix.plan) {
        const uint64_t len = g.len;
        const int32_t bits = g.bits;
        put(&len, 8);
        put(&bits, 4);
    }
    put(ix.centroids.data(), ix.centroids.size() * 4);
    put(ix.ids.data(), ix.ids.size() * 4);
    put(ix.list_begin.data(), ix.list_begin.size() * 8);
    put(ix.byte_begin.data(), ix.byte_begin.size() * 8);
    put(ix.codes.data(), ix.codes.size());
    put(&ix.rot_seed, 8);
    const uint64_t nr = ix.rot.size();
    put(&nr, 8);
    for (const Rotator& r : ix.rot) {
        const uint64_t nb = r.signs.size();
        put(&nb, 8);
        put(r.signs.data(), nb);
    }
    return out;
}

// ---------------------------------------------------------------------------
// Diagnostics: Estimate (me8q_measure.cpp:41-67), ExhaustiveTopK
// (me8q_search.cpp:238-266), SearchIVF (me8q_search.cpp:136-236).

// ScoreFast (me8q_score_avx2.cpp:104-129) over a run's bit-tight code:
// component i of even blocks in acc[i], of odd blocks in acc2[i] (one chain
// at rate 1), FMA chains, then the AVX horizontal sum of acc + acc2.
static float ScoreFast(const uint8_t* code, int len, int bits, const float* q) {
    const Codec& cd = GetCodec();
    float acc[8] = {}, acc2[8] = {};
    for (int b = 0; b < len / 8; ++b) {
        uint16_t d[8];
        uint8_t br;
        UnpackBlock(code + static_cast<size_t>(b) * bits, bits, d, &br);
        const float* lev = cd.tab[br][bits].data();
        const float* qb = q + static_cast<size_t>(b) * 8;
        for (int i = 0; i < 8; ++i) {
            if (bits >= 2 && (b & 1)) acc2[i] = std::fma(lev[d[i]], qb[i], acc2[i]);
            else acc[i] = std::fma(lev[d[i]], qb[i], acc[i]);
        }
    }
    if (bits >= 2)
        for (int i = 0; i < 8; ++i) acc[i] = acc[i] + acc2[i];
    return Hsum8(acc);
}

// EstimateRow (me8q_search.cpp:77-99): d2 + sum over the coded runs of
// h2 - 2 coef S (the product fused into the subtraction; a run with coef 0
// adds h2).
static double EstimateRow(const Index& ix, const uint8_t* row, double d2, const float* q) {
    double acc = d2;
    for (size_t r = 0; r < ix.nrun; ++r) {
        float fac[2];
        std::memcpy(fac, row + r * 8, 8);
        const double coef = fac[0];
        if (coef == 0.0) {
            acc += fac[1];
            continue;
        }
        const double sc = ScoreFast(row + ix.run_code_off[r], static_cast<int>(ix.run_len[r]),
                                    static_cast<int>(ix.run_rate[r]), q + ix.run_dim_off[r]);
        acc += std::fma(-(2.0 * coef), sc, static_cast<double>(fac[1]));
    }
    return acc;
}

// The padded query, its per-run rotation (zero tail) and the fp64 squared
// distance to list c's unrotated centroid (dimension order, contracted).
static void PadRotate(const Index& ix, const float* query, std::vector<float>& pad, std::vector<float>& rot) {
    pad.assign(ix.padded, 0.0F);
    std::memcpy(pad.data(), query, ix.dim * 4);
    rot.assign(ix.padded, 0.0F);
    for (size_t r = 0; r < ix.nrun; ++r) ix.rot[r].Rotate(&pad[ix.run_dim_off[r]], &rot[ix.run_dim_off[r]]);
}

static double CentroidD2(const Index& ix, const float* pad, size_t c) {
    double d2 = 0.0;
    for (size_t j = 0; j < ix.padded; ++j) {
        const double d = static_cast<double>(pad[j]) - static_cast<double>(ix.centroids[c * ix.padded + j]);
        d2 = std::fma(d, d, d2);
    }
    return d2;
}

static const uint8_t* RowOf(const Index& ix, size_t c, size_t i) {
    return ix.codes.data() + ix.byte_begin[c] + (i - ix.list_begin[c]) * ix.stride;
}

// Estimate; false for an unknown id (the last stored vector carrying it
// answers, as EnsureIdMap's overwrite order leaves it).
bool Estimate(const Index& ix, const float* query, uint32_t id, float* out) {
    size_t pos = static_cast<size_t>(-1);
    if (id < ix.n)
        for (size_t i = 0; i < ix.n; ++i)
            if (ix.ids[i] == id) pos = i;
    if (pos == static_cast<size_t>(-1)) return false;
    const auto it = std::upper_bound(ix.list_begin.begin() + 1, ix.list_begin.end(), pos);
    const size_t c = static_cast<size_t>(it - ix.list_begin.begin() - 1);
    std::vector<float> pad, rot;
    PadRotate(ix, query, pad, rot);
    *out = static_cast<float>(EstimateRow(ix, RowOf(ix, c, pos), CentroidD2(ix, pad.data(), c), rot.data()));
    return true;
}

void ExhaustiveTopK(const Index& ix, const float* query, size_t k, std::vector<std::pair<float, uint32_t>>& out) {
    out.clear();
    std::vector<float> pad, rot;
    PadRotate(ix, query, pad, rot);
    for (size_t c = 0; c < ix.nlist; ++c) {
        const double d2 = CentroidD2(ix, pad.data(), c);
        for (size_t i = ix.list_begin[c]; i < ix.list_begin[c + 1]; ++i)
            Push(out, k, static_cast<float>(EstimateRow(ix, RowOf(ix, c, i), d2, rot.data())), ix.ids[i]);
    }
}

// SearchIVF: the route's probe distances, every row of every probed list in
// probe order.  stats6: candidates, clusters, bits_scored, bits_total,
// survivors, pruned (0).
void SearchIVF(const Index& ix, const float* query, size_t k, size_t nprobe, uint32_t* out_ids, float* out_d,
               uint64_t* stats6) {
    if (k == 0 || nprobe == 0 || ix.n == 0) return;
    std::vector<float> pad, rot;
    PadRotate(ix, query, pad, rot);
    std::vector<float> pd;
    std::vector<uint32_t> probe;
    Route(ix, rot.data(), nprobe, pd, probe);
    uint64_t bits_row = 0;
    for (size_t r = 0; r < ix.nrun; ++r) bits_row += static_cast<uint64_t>(ix.run_rate[r]) * ix.run_len[r];
    std::vector<std::pair<float, uint32_t>> h;
    uint64_t cand = 0;
    for (size_t p = 0; p < probe.size(); ++p) {
        const size_t c = probe[p];
        for (size_t i = ix.list_begin[c]; i < ix.list_begin[c + 1]; ++i) {
            Push(h, k, static_cast<float>(EstimateRow(ix, RowOf(ix, c, i), static_cast<double>(pd[p]), rot.data())),
                 ix.ids[i]);
            ++cand;
        }
    }
    for (size_t i = 0; i < k; ++i) {
        out_ids[i] = i < h.size() ? h[i].second : std::numeric_limits<uint32_t>::max();
        out_d[i] = i < h.size() ? h[i].first : std::numeric_limits<float>::infinity();
    }
    if (stats6) {
        stats6[0] = cand;
        stats6[1] = probe.size();
        stats6[2] = cand * bits_row;
        stats6[3] = cand * bits_row;
        stats6[4] = cand;
        stats6[5] = 0;
    }
}

}  // namespace orc::m8

// ---------------------------------------------------------------------------
// C interface for the tests (ctypes).
extern "C" {

void* orc_m8_build(const float* base, const float* cent, const uint32_t* cid, size_t n, size_t dim, size_t nlist,
                   float avg_bits, uint64_t seed, float prune_alpha, const char* plan_spec, const float* variance,
                   int drop_energy) {
    auto* ix = new orc::m8::Index();
    try {
        orc::m8::Build(*ix, base, cent, cid, n, dim, nlist, avg_bits, seed, prune_alpha, plan_spec ? plan_spec : "",
                       variance, drop_energy != 0);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "orc_m8_build: %s\n", e.what());
        delete ix;
        return nullptr;
    }
    return ix;
}

void orc_m8_free(void* h) { delete static_cast<orc::m8::Index*>(h); }

// [n, dim, padded, nlist, nrun, plane_cols, refine_stride, stride, route_dim, image_bytes, plan_len]
void orc_m8_info(void* h, uint64_t* out11) {
    const auto* ix = static_cast<orc::m8::Index*>(h);
    const uint64_t v[11] = {ix->n, ix->dim, ix->padded, ix->nlist, ix->nrun, ix->plane_cols, ix->refine_stride,
                            ix->stride, ix->route_dim, orc::m8::Image(*ix).size(), ix->plan.size()};
    std::memcpy(out11, v, sizeof(v));
}

void orc_m8_image(void* h, uint8_t* out) {
    const auto img = orc::m8::Image(*static_cast<orc::m8::Index*>(h));
    std::memcpy(out, img.data(), img.size());
}

void orc_m8_plan(void* h, int64_t* out) {
    const auto* ix = static_cast<orc::m8::Index*>(h);
    for (size_t i = 0; i < ix->plan.size(); ++i) {
        out[2 * i] = static_cast<int64_t>(ix->plan[i].len);
        out[2 * i + 1] = ix->plan[i].bits;
    }
}

// Sizes of one list's fast layout: [plane bytes, refine bytes, gain floats]
void orc_m8_fast_sizes(void* h, size_t c, uint64_t* out3) {
    const auto& f = static_cast<orc::m8::Index*>(h)->fast[c];
    out3[0] = f.plane.size();
    out3[1] = f.refine.size();
    out3[2] = f.gain.size();
}

void orc_m8_fast(void* h, size_t c, uint8_t* plane, uint8_t* refine, float* gain, float* vpc, double* rc_n2) {
    const auto* ix = static_cast<orc::m8::Index*>(h);
    const auto& f = ix->fast[c];
    if (plane) std::memcpy(plane, f.plane.data(), f.plane.size());
    if (refine) std::memcpy(refine, f.refine.data(), f.refine.size());
    if (gain) std::memcpy(gain, f.gain.data(), f.gain.size() * 4);
    if (vpc) std::memcpy(vpc, f.vpc.data(), f.vpc.size() * 4);
    if (rc_n2) std::memcpy(rc_n2, f.rc_n2.data(), f.rc_n2.size() * 8);
}

void orc_m8_centroids(void* h, float* cent_rot, float* cn2) {
    const auto* ix = static_cast<orc::m8::Index*>(h);
    if (cent_rot) std::memcpy(cent_rot, ix->cent_rot.data(), ix->cent_rot.size() * 4);
    if (cn2) std::memcpy(cn2, ix->centroid_n2.data(), ix->centroid_n2.size() * 4);
}

// Query-side intermediates: rotated query, u8 LUT (plane_cols*64), probe list.
void orc_m8_query(void* h, const float* query, size_t nprobe, float* rot_q, uint8_t* lut, float* lut_md,
                  uint32_t* probe, float* pdist) {
    const auto* ix = static_cast<orc::m8::Index*>(h);
    orc::m8::QueryPrep p;
    orc::m8::PrepareQuery(*ix, query, p);
    std::vector<float> pd;
    std::vector<uint32_t> pr;
    orc::m8::Route(*ix, p.q.data(), nprobe, pd, pr);
    if (rot_q) std::memcpy(rot_q, p.q.data(), p.q.size() * 4);
    if (lut) std::memcpy(lut, p.lut.data(), p.lut.size());
    if (lut_md) {
        lut_md[0] = p.lut_min;
        lut_md[1] = p.lut_delta;
    }
    if (probe) std::memcpy(probe, pr.data(), pr.size() * 4);
    if (pdist) std::memcpy(pdist, pd.data(), pd.size() * 4);
}

// stats5 per query: candidates, survivors, bits_scored, bits_total, pruned.
// lambda: as Search's lambda_override (NaN: the index's own).
int orc_m8_search_z(void* h, const float* queries, size_t nq, size_t k, size_t nprobe, int threads, float lambda,
                    uint32_t* ids, float* dists, uint64_t* stats5) {
    const auto* ix = static_cast<orc::m8::Index*>(h);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (long long i = 0; i < static_cast<long long>(nq); ++i) {
        orc::m8::Stats st;
        orc::m8::Search(*ix, queries + static_cast<size_t>(i) * ix->dim, k, nprobe, ids + static_cast<size_t>(i) * k,
                        dists + static_cast<size_t>(i) * k, &st, lambda);
        if (stats5) {
            uint64_t* o = stats5 + static_cast<size_t>(i) * 5;
            o[0] = st.candidates;
            o[1] = st.survivors;
            o[2] = st.bits_scored;
            o[3] = st.bits_total;
            o[4] = st.pruned;
        }
    }
    return 0;
}


int orc_m8_search(void* h, const float* queries, size_t nq, size_t k, size_t nprobe, int threads, uint32_t* ids,
                  float* dists, uint64_t* stats5) {
    return orc_m8_search_z(h, queries, nq, k, nprobe, threads, std::numeric_limits<float>::quiet_NaN(), ids, dists,
                           stats5);
}

// Diagnostics.  orc_m8_estimate: status 1 = an unknown id.
int orc_m8_estimate(void* h, const float* queries, const uint32_t* qidx, const uint32_t* ids, size_t npair,
                    float* out) {
    const auto& ix = *static_cast<orc::m8::Index*>(h);
    for (size_t i = 0; i < npair; ++i)
        if (!orc::m8::Estimate(ix, queries + static_cast<size_t>(qidx[i]) * ix.dim, ids[i], out + i)) return 1;
    return 0;
}

// k results per query, UINT32_MAX / +inf padded.
void orc_m8_exhaustive_topk(void* h, const float* queries, size_t nq, size_t k, uint32_t* ids, float* dists) {
    const auto& ix = *static_cast<orc::m8::Index*>(h);
    for (size_t i = 0; i < nq; ++i) {
        std::vector<std::pair<float, uint32_t>> out;
        orc::m8::ExhaustiveTopK(ix, queries + i * ix.dim, k, out);
        for (size_t j = 0; j < k; ++j) {
            ids[i * k + j] = j < out.size() ? out[j].second : std::numeric_limits<uint32_t>::max();
            dists[i * k + j] = j < out.size() ? out[j].first : std::numeric_limits<float>::infinity();
        }
    }
}

void orc_m8_search_ivf(void* h, const float* queries, size_t nq, size_t k, size_t nprobe, uint32_t* ids,
                       float* dists, uint64_t* stats6) {
    const auto& ix = *static_cast<orc::m8::Index*>(h);
    for (size_t i = 0; i < nq; ++i)
        orc::m8::SearchIVF(ix, queries + i * ix.dim, k, nprobe, ids + i * k, dists + i * k,
                           stats6 ? stats6 + i * 6 : nullptr);
}

}  // extern "C"
