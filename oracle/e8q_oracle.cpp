// e8q_oracle.cpp -- CPU restatement of e8qlib's E8Q build (encode) and
// search path.  TEST INFRASTRUCTURE ONLY: it is the checker the HIP path is
// compared against (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline
// leg).  Nothing in the product links or calls it.
//
// Numerics follow the reference as compiled by its own flags (GCC, C++17,
// -O3, x86-64-v3 with FMA, E8Q_ISA=avx2): every place where GCC contracts a
// multiply-add in the reference's generic translation units, and every
// explicit _mm256_fmadd/_fnmadd in its AVX2 kernels, is written here as an
// explicit std::fma; everything else is evaluated with contraction off
// (this file is compiled with -ffp-contract=off).
//
// Parity pinning (see tests/test_oracle_golden.py): level tables, rotator
// signs and outputs, encoder digits, whole E8AX index images, query LUT
// bytes, route lists and search results are compared bit-for-bit with
// fixtures written by the reference library (tests/golden/gen_golden.cpp).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "orc_common.h"

namespace orc {

constexpr double kSigmaC = 1.1180339887498949;  // sqrt(5/4), e8q_codec.hpp:24
constexpr size_t kMaxK = 128;                   // e8q.hpp:44
constexpr size_t kHdr = 128;                    // scan-record header, e8q.hpp:193-197

// ---------------------------------------------------------------------------
// Level tables (e8q_codec.cpp:17-44, 427-524)

// Acklam's rational inverse normal CDF.  The reference build contracts each
// Horner step `x*r + c` into one fused multiply-add (its BranchLevels holds
// 30 FMAs); the final `* q` and the divide are separate roundings.
static double AcklamInvNormal(double p) {
    static const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02,
                                -2.759285104469687e+02, 1.383577518672690e+02,
                                -3.066479806614716e+01, 2.506628277459239e+00};
    static const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02,
                                -1.556989798598866e+02, 6.680131188771972e+01,
                                -1.328068155288572e+01};
    static const double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01,
                                -2.400758277161838e+00, -2.549732539343734e+00,
                                4.374664141464968e+00, 2.938163982698783e+00};
    static const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01,
                                2.445134137142996e+00, 3.754408661907416e+00};
    const double plow = 0.02425;
    auto tail = [&](double q) {
        double num = std::fma(c[0], q, c[1]);
        num = std::fma(num, q, c[2]);
        num = std::fma(num, q, c[3]);
        num = std::fma(num, q, c[4]);
        num = std::fma(num, q, c[5]);
        double den = std::fma(d[0], q, d[1]);
        den = std::fma(den, q, d[2]);
        den = std::fma(den, q, d[3]);
        den = std::fma(den, q, 1.0);
        return num / den;
    };
    if (p < plow) {
        return tail(std::sqrt(-2 * std::log(std::max(p, 1e-300))));
    }
    if (p <= 1 - plow) {
        const double q = p - 0.5;
        const double r = q * q;
        double num = std::fma(a[0], r, a[1]);
        num = std::fma(num, r, a[2]);
        num = std::fma(num, r, a[3]);
        num = std::fma(num, r, a[4]);
        num = std::fma(num, r, a[5]);
        double den = std::fma(b[0], r, b[1]);
        den = std::fma(den, r, b[2]);
        den = std::fma(den, r, b[3]);
        den = std::fma(den, r, b[4]);
        den = std::fma(den, r, 1.0);
        return num * q / den;
    }
    return -tail(std::sqrt(-2 * std::log(std::max(1 - p, 1e-300))));
}

double FixedScale(int bits) {
    static const std::array<double, kMaxBits + 1> tab = [] {
        std::array<double, kMaxBits + 1> t = {0.0,   1.118034, 0.895, 0.825, 0.785,
                                              0.755, 0.735,    0.735, 0.735, 0.725,
                                              0.715, 0.705,    0.710, 0.705};
        // E8Q_SCALE_SCHEDULE="s1,s2,..." (e8q_codec.cpp:464-486).
        if (const char* env = std::getenv("E8Q_SCALE_SCHEDULE")) {
            const char* p = env;
            for (int i = 1; i <= kMaxBits && *p != '\0'; ++i) {
                char* end = nullptr;
                const double v = std::strtod(p, &end);
                if (end == p) break;
                t[static_cast<size_t>(i)] = v;
                p = end;
                while (*p == ',' || *p == ' ') ++p;
            }
        }
        return t;
    }();
    return tab[static_cast<size_t>(std::min(std::max(bits, 1), kMaxBits))];
}

std::vector<double> BranchLevels(int bits, int branch) {
    const int q = 1 << bits;
    // eta = 0.25 + 0.5*branch and ceil(-0.5*q - eta) are both contracted in
    // the reference build.
    const double eta = std::fma(0.5, static_cast<double>(branch), 0.25);
    const int start = static_cast<int>(std::ceil(std::fma(-0.5, static_cast<double>(q), -eta)));
    std::vector<double> lev(static_cast<size_t>(q));
    for (int i = 0; i < q; ++i) {
        const long t = start + i;
        const int digit = static_cast<int>(((t % q) + q) % q);
        const double p = 0.5 + (static_cast<double>(t) + eta) / static_cast<double>(q);
        lev[static_cast<size_t>(digit)] =
            kSigmaC * AcklamInvNormal(std::min(1.0 - 1e-12, std::max(1e-12, p)));
    }
    return lev;
}

void ScanClassMeans(int bits, double w[2]) {
    const auto l0 = BranchLevels(bits, 0);
    const auto l1 = BranchLevels(bits, 1);
    const int cv = 1 << (bits - 1);
    for (int t = 0; t < 2; ++t) {
        double s0 = 0.0, s1 = 0.0;
        for (int lo = 0; lo < cv; ++lo) {
            s0 += l0[static_cast<size_t>((t << (bits - 1)) | lo)];
            s1 += l1[static_cast<size_t>((t << (bits - 1)) | lo)];
        }
        w[t] = 0.5 * (s0 / cv + s1 / cv);
    }
}

std::vector<float> RefineTable(int bits, size_t padded) {
    const size_t q = size_t{1} << bits;
    const size_t stride = std::max<size_t>(256, q);
    std::vector<float> tab(2 * stride, 0.0F);
    const double rd = std::sqrt(static_cast<double>(padded));
    for (int br = 0; br < 2; ++br) {
        const auto lev = BranchLevels(bits, br);
        for (size_t d = 0; d < q; ++d) tab[br * stride + d] = static_cast<float>(lev[d] / rd);
    }
    return tab;
}

// ---------------------------------------------------------------------------
// FHT-Kac-4 rotator (rotator.cpp:13-80, kernels_generic.cpp:49-78)


// ---------------------------------------------------------------------------
// E8 nearest-lattice-point encoder (e8q_codec.cpp:142-414, 549-726)

struct ClassTab {       // one (branch, digit-parity) class, sorted by (level, digit)
    std::vector<double> lev;
    std::vector<uint16_t> dig;
};

struct EncTables {
    int bits = 0;
    std::vector<double> l[2];
    ClassTab cls[2][2];           // [branch][parity]
    bool merged = false;          // classes interleave with period 4 (b >= 2)
    std::vector<double> mlev;     // merged sorted levels (2^(b+1))
    std::vector<uint16_t> mdig;
};

static const EncTables& GetEncTables(int bits) {
    static std::vector<EncTables> cache = [] {
        std::vector<EncTables> v(kMaxBits + 1);
        for (int b = 1; b <= kMaxBits; ++b) {
            EncTables& t = v[static_cast<size_t>(b)];
            t.bits = b;
            t.l[0] = BranchLevels(b, 0);
            t.l[1] = BranchLevels(b, 1);
            const int q = 1 << b;
            for (int br = 0; br < 2; ++br)
                for (int par = 0; par < 2; ++par) {
                    std::vector<std::pair<double, uint16_t>> e;
                    for (int d = par; d < q; d += 2)
                        e.push_back({t.l[br][static_cast<size_t>(d)], static_cast<uint16_t>(d)});
                    std::sort(e.begin(), e.end());
                    for (auto& x : e) {
                        t.cls[br][par].lev.push_back(x.first);
                        t.cls[br][par].dig.push_back(x.second);
                    }
                }
            // Merged order: level(d,0) < level(d,1) < level(d+1,0) < ...; the
            // class of position i is (branch i&1, parity (i>>1)&1) when the
            // interleave holds.  Verified, as the reference does; b == 1
            // fails it (digit order 1,0) and keeps the per-class search.
            struct E { double v; uint16_t d; int br; };
            std::vector<E> m;
            for (int d = 0; d < q; ++d) {
                m.push_back({t.l[0][static_cast<size_t>(d)], static_cast<uint16_t>(d), 0});
                m.push_back({t.l[1][static_cast<size_t>(d)], static_cast<uint16_t>(d), 1});
            }
            std::stable_sort(m.begin(), m.end(), [](const E& a, const E& c) { return a.v < c.v; });
            bool alt = true;
            for (size_t i = 0; i < m.size(); ++i) {
                if (i > 0 && !(m[i].v > m[i - 1].v)) alt = false;
                if (m[i].br != static_cast<int>(i & 1) || (m[i].d & 1) != ((i >> 1) & 1)) alt = false;
            }
            t.merged = alt;
            for (auto& x : m) {
                t.mlev.push_back(x.v);
                t.mdig.push_back(x.d);
            }
        }
        return v;
    }();
    return cache[static_cast<size_t>(bits)];
}

// Nearest level of one class table: first index with lev >= v, then the
// lower neighbour wins on a smaller |distance|, or an equal one with a
// smaller digit (e8q_codec.cpp:233-256).
static void NearestClass(double v, const ClassTab& t, uint16_t* dig, double* cost) {
    const size_t n = t.lev.size();
    const size_t pos = static_cast<size_t>(std::lower_bound(t.lev.begin(), t.lev.end(), v) - t.lev.begin());
    size_t best = pos < n ? pos : n - 1;
    if (pos > 0) {
        const double dc = std::abs(v - t.lev[pos - 1]);
        const double db = std::abs(v - t.lev[best]);
        if ((dc < db) || (dc == db && t.dig[pos - 1] < t.dig[best])) best = pos - 1;
    }
    *dig = t.dig[best];
    const double dv = v - t.lev[best];
    *cost = dv * dv;
}

// Merged search (e8q_codec.cpp:387-414): for class r the neighbours are the
// nearest positions of residue r below / at-or-above the insertion point;
// squared distances compared, ties to the smaller digit.
static void NearestMergedClass(double v, const EncTables& t, uint16_t dig[4], double cost[4]) {
    const long n = static_cast<long>(t.mlev.size());
    const long p = static_cast<long>(std::lower_bound(t.mlev.begin(), t.mlev.end(), v) - t.mlev.begin());
    const long p4 = p & 3;
    for (long r = 0; r < 4; ++r) {
        const long ib = p - 1 - ((p4 + 3 - r) & 3);
        const long ia = p + ((r + 4 - p4) & 3);
        const double lb = ib >= 0 ? t.mlev[static_cast<size_t>(ib)] : -std::numeric_limits<double>::infinity();
        const double la = ia < n ? t.mlev[static_cast<size_t>(ia)] : std::numeric_limits<double>::infinity();
        const uint16_t db = ib >= 0 ? t.mdig[static_cast<size_t>(ib)] : 0;
        const uint16_t da = ia < n ? t.mdig[static_cast<size_t>(ia)] : 0;
        const double cb = (v - lb) * (v - lb);
        const double ca = (la - v) * (la - v);
        const bool take_b = (cb < ca) || (cb == ca && db < da);
        dig[r] = take_b ? db : da;
        cost[r] = take_b ? cb : ca;
    }
}

struct BlockSel {
    uint16_t d[8];
    double cost;
};

// Per-branch even/odd choice plus the single-flip parity repair
// (e8q_codec.cpp:549-599).
static BlockSel Assemble(const uint16_t de[8], const double ce[8], const uint16_t dO[8], const double co[8]) {
    BlockSel s{};
    s.cost = 0.0;
    int par = 0;
    for (int i = 0; i < 8; ++i) {
        const bool odd = co[i] < ce[i];
        s.d[i] = odd ? dO[i] : de[i];
        s.cost += odd ? co[i] : ce[i];
        par ^= odd ? 1 : 0;
    }
    if (par != 0) {
        int sel = 0;
        double best = std::numeric_limits<double>::max();
        for (int i = 0; i < 8; ++i) {
            const bool was_odd = s.d[i] == dO[i];
            const double extra = was_odd ? ce[i] - co[i] : co[i] - ce[i];
            if (extra < best) {
                best = extra;
                sel = i;
            }
        }
        s.d[sel] = (s.d[sel] == dO[sel]) ? de[sel] : dO[sel];
        s.cost += std::max(best, 0.0);
    }
    return s;
}

// Whole-vector encode at a fixed scale (EncodeVectorAtScaleInto).
void EncodeVector(const float* unit, int padded, int bits, double s, uint16_t* digits, uint8_t* branch) {
    const EncTables& t = GetEncTables(bits);
    const int blocks = padded / 8;
    const double rd = std::sqrt(static_cast<double>(padded));
    const double sr = s * rd;
    for (int blk = 0; blk < blocks; ++blk) {
        uint16_t de[2][8], dO[2][8];
        double ce[2][8], co[2][8];
        for (int i = 0; i < 8; ++i) {
            const double v = static_cast<double>(unit[blk * 8 + i]) * sr;
            if (t.merged) {
                uint16_t dg[4];
                double cs[4];
                NearestMergedClass(v, t, dg, cs);
                de[0][i] = dg[0]; ce[0][i] = cs[0];
                de[1][i] = dg[1]; ce[1][i] = cs[1];
                dO[0][i] = dg[2]; co[0][i] = cs[2];
                dO[1][i] = dg[3]; co[1][i] = cs[3];
            } else {
                for (int br = 0; br < 2; ++br) {
                    NearestClass(v, t.cls[br][0], &de[br][i], &ce[br][i]);
                    NearestClass(v, t.cls[br][1], &dO[br][i], &co[br][i]);
                }
            }
        }
        const BlockSel e0 = Assemble(de[0], ce[0], dO[0], co[0]);
        const BlockSel e1 = Assemble(de[1], ce[1], dO[1], co[1]);
        const bool one = e1.cost < e0.cost;
        const BlockSel& e = one ? e1 : e0;
        for (int i = 0; i < 8; ++i) digits[blk * 8 + i] = e.d[i];
        branch[blk] = one ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// Index image (E8AX layout, e8q_io.cpp:35-84 and e8q.hpp:189-266)

struct Index {
    size_t n = 0, dim = 0, padded = 0, nlist = 0;
    int bits = 0, blocks = 0;
    Rotator rot;
    std::vector<uint32_t> ids;
    std::vector<float> mean_rot, centroid_rot, centroid_n2, refine_tab;
    std::vector<uint64_t> cluster_begin, arena_begin, ref_arena_begin;
    std::vector<uint8_t> arena, ref_arena;
    size_t batch_bytes = 0, ref_batch_bytes = 0;
    double gate_mid[16] = {};
    float alpha = 1.0F;
    bool disable_prune = false;

    size_t RefineStride() const { return bits >= 2 ? static_cast<size_t>(blocks) * (bits - 1) : 0; }
    size_t RefG1Off() const { return RefineStride() * 32; }
    size_t RefH3Off() const { return RefG1Off() + 128; }

    void InitTables() {
        double w[2];
        ScanClassMeans(bits, w);
        for (int d = 0; d < 8; ++d) {
            gate_mid[2 * d] = (bits == 1 && d == 7) ? 0.0 : w[0];
            gate_mid[2 * d + 1] = (bits == 1 && d == 7) ? 0.0 : w[1];
        }
        refine_tab = RefineTable(bits, padded);
        batch_bytes = (kHdr + static_cast<size_t>(blocks) * 32 + 63) & ~size_t{63};
        ref_batch_bytes = (RefH3Off() + 128 + 63) & ~size_t{63};
    }
};

// rowA byte (e8q_scan.cpp:193-205) and refine field (e8q_codec.cpp:813-838).
static uint8_t RowAByte(const uint16_t d[8], int br, int bits) {
    uint8_t b = 0;
    for (int i = 0; i < 8; ++i) {
        const int top = (bits == 1 && i == 7) ? (br & 1) : ((d[i] >> (bits - 1)) & 1);
        if (top) b = static_cast<uint8_t>(b | (1u << (7 - i)));
    }
    return b;
}

static void RefineField(const uint16_t d[8], int br, int bits, uint8_t* out) {
    const int c = bits - 1;
    unsigned __int128 x = 0;
    const unsigned __int128 mask = (static_cast<unsigned __int128>(1) << c) - 1;
    for (int i = 0; i < 7; ++i) x |= (static_cast<unsigned __int128>(d[i]) & mask) << (c * i);
    x |= (static_cast<unsigned __int128>(d[7] & ~1u) & mask) << (c * 7);
    x |= static_cast<unsigned __int128>(br & 1) << (c * 7);
    for (int i = 0; i < c; ++i) out[i] = static_cast<uint8_t>(x >> (8 * i));
}

// Per-batch u8 banks (e8q_build.cpp:33-80).
static void WriteBanks(const float* ob, const float* ow, const float* op, size_t cnt, uint8_t* bp) {
    float bmin = 0, bmax = 0, wmin = 0, wmax = 0, pmin = 0, pmax = 0;
    if (cnt > 0) {
        bmin = bmax = ob[0];
        wmin = wmax = ow[0];
        pmin = pmax = op[0];
        for (size_t i = 1; i < cnt; ++i) {
            bmin = std::min(bmin, ob[i]); bmax = std::max(bmax, ob[i]);
            wmin = std::min(wmin, ow[i]); wmax = std::max(wmax, ow[i]);
            pmin = std::min(pmin, op[i]); pmax = std::max(pmax, op[i]);
        }
    }
    const float wamax = std::max(std::abs(wmin), std::abs(wmax));
    const float bsc = bmax > bmin ? (bmax - bmin) / 255.0F : 1.0F;
    const float wsc = wmax > wmin ? (wmax - wmin) / 255.0F : 1.0F;
    const float psc = pmax > pmin ? (pmax - pmin) / 255.0F : 1.0F;
    std::memset(bp, 0, kHdr);
    for (size_t i = 0; i < cnt; ++i) {
        bp[i] = static_cast<uint8_t>(std::clamp(static_cast<int>((ob[i] - bmin) / bsc), 0, 255));
        bp[32 + i] = static_cast<uint8_t>(std::clamp(static_cast<int>(std::lround((ow[i] - wmin) / wsc)), 0, 255));
        bp[64 + i] = static_cast<uint8_t>(std::clamp(static_cast<int>(std::ceil((op[i] - pmin) / psc)), 0, 255));
    }
    float sc[8] = {bmin, bsc, wmin, wsc, pmin, psc, wamax, 0.0F};
    std::memcpy(bp + 96, sc, 32);
}

// Build (e8q_build.cpp:87-503), sequential, rows in file order.
void Build(Index& ix, const float* base, const float* cent, const uint32_t* cid, size_t n, size_t dim,
           size_t nlist, int bits, uint64_t seed, float prune_alpha) {
    if (bits < 1 || bits > kMaxBits) throw std::invalid_argument("E8Q bits must be in [1, 13]");
    const bool hilbert = [] {
        const char* p = std::getenv("E8Q_PROJECTION");
        return p != nullptr && std::string(p) == "hilbert";
    }();
    ix.alpha = prune_alpha > 0.0F ? prune_alpha : 1.0F;
    ix.n = n; ix.dim = dim; ix.nlist = nlist; ix.bits = bits;
    ix.rot = Rotator(dim, seed);
    ix.padded = ix.rot.padded;
    ix.blocks = static_cast<int>(ix.padded / 8);
    const size_t P = ix.padded;
    std::vector<uint64_t> counts(nlist, 0);
    for (size_t i = 0; i < n; ++i) {
        if (cid[i] >= nlist) throw std::invalid_argument("cluster id out of range");
        ++counts[cid[i]];
    }
    ix.InitTables();
    ix.cluster_begin.assign(nlist + 1, 0);
    ix.arena_begin.assign(nlist + 1, 0);
    ix.ref_arena_begin.assign(nlist + 1, 0);
    for (size_t c = 0; c < nlist; ++c) {
        ix.cluster_begin[c + 1] = ix.cluster_begin[c] + counts[c];
        const uint64_t nb = (counts[c] + 31) / 32;
        ix.arena_begin[c + 1] = ix.arena_begin[c] + nb * ix.batch_bytes;
        ix.ref_arena_begin[c + 1] = ix.ref_arena_begin[c] + nb * ix.ref_batch_bytes;
    }
    ix.arena.assign(ix.arena_begin.back() + 64, 0);
    ix.ref_arena.assign(ix.ref_arena_begin.back() + 64, 0);
    ix.ids.assign(n, 0);
    ix.centroid_rot.assign(nlist * P, 0.0F);
    ix.centroid_n2.assign(nlist, 0.0F);
    std::vector<uint32_t> rank(n);
    {
        std::vector<uint64_t> cur(nlist, 0);
        for (size_t i = 0; i < n; ++i) rank[i] = static_cast<uint32_t>(cur[cid[i]]++);
    }
    const auto lev0 = BranchLevels(bits, 0), lev1 = BranchLevels(bits, 1);
    const std::vector<double>* lev[2] = {&lev0, &lev1};
    const double s_vec = FixedScale(bits);
    const double inv_ls = 1.0 / std::sqrt(static_cast<double>(P));
    double w_cls[2];
    ScanClassMeans(bits, w_cls);

    // Base mean: fp64 sums in row order, cast, then rotated.
    {
        std::vector<double> mu(dim, 0.0);
        for (size_t i = 0; i < n; ++i)
            for (size_t d = 0; d < dim; ++d) mu[d] += static_cast<double>(base[i * dim + d]);
        std::vector<float> mp(P, 0.0F);
        for (size_t d = 0; d < dim; ++d) mp[d] = static_cast<float>(mu[d] / static_cast<double>(n));
        ix.mean_rot.assign(P, 0.0F);
        ix.rot.Rotate(mp.data(), ix.mean_rot.data());
    }
    for (size_t c = 0; c < nlist; ++c) {
        float* rc = &ix.centroid_rot[c * P];
        ix.rot.Rotate(cent + c * dim, rc);
        double n2 = 0.0;
        for (size_t d = 0; d < P; ++d) {
            rc[d] -= ix.mean_rot[d];
            const double v = static_cast<double>(rc[d]);
            n2 = std::fma(v, v, n2);
        }
        ix.centroid_n2[c] = static_cast<float>(n2);
    }
    std::vector<float> ob(n, 0.0F), ow(n, 0.0F), op(n, 0.0F);  // by cluster-major slot
    std::vector<float> rx(P), unit(P);
    std::vector<uint16_t> dg(P);
    std::vector<uint8_t> brs(ix.blocks);
    const int cbits = bits - 1;
    const size_t run = ix.RefineStride();
    for (size_t r = 0; r < n; ++r) {
        const size_t c = cid[r];
        const size_t rk = rank[r], lane = rk % 32;
        const size_t slot = ix.cluster_begin[c] + rk;
        uint8_t* rec = ix.arena.data() + ix.arena_begin[c] + (rk / 32) * ix.batch_bytes;
        uint8_t* rrec = ix.ref_arena.data() + ix.ref_arena_begin[c] + (rk / 32) * ix.ref_batch_bytes;
        const float* rc = &ix.centroid_rot[c * P];
        ix.ids[slot] = static_cast<uint32_t>(r);
        ix.rot.Rotate(base + r * dim, rx.data());
        for (size_t d = 0; d < P; ++d) rx[d] -= ix.mean_rot[d];
        double rn2 = 0.0;
        for (size_t d = 0; d < P; ++d) {
            const double rv = static_cast<double>(rx[d]) - static_cast<double>(rc[d]);
            rn2 = std::fma(rv, rv, rn2);
        }
        const double rn = std::sqrt(rn2);
        if (rn < 1e-20) continue;
        const double inv_rn = 1.0 / rn;
        for (size_t d = 0; d < P; ++d)
            unit[d] = static_cast<float>((static_cast<double>(rx[d]) - static_cast<double>(rc[d])) * inv_rn);
        EncodeVector(unit.data(), static_cast<int>(P), bits, s_vec, dg.data(), brs.data());
        // Fit dots.  The reference build contracts every accumulation into an
        // FMA except ||w||^2 at the 8th coordinate of a block, which GCC
        // hoists as a separately rounded square.
        double due = 0, n2e = 0, dce = 0, dew = 0, duw = 0, n2w = 0, dcw = 0;
        for (int blk = 0; blk < ix.blocks; ++blk) {
            const int br = brs[static_cast<size_t>(blk)];
            for (int k = 0; k < 8; ++k) {
                const size_t i = static_cast<size_t>(blk) * 8 + k;
                const int digit = dg[i];
                const double ve = (*lev[br])[static_cast<size_t>(digit)] * inv_ls;
                const double wc = (bits == 1 && k == 7) ? 0.0 : w_cls[digit >> (bits - 1)];
                const double u = static_cast<double>(unit[i]);
                const double cc = static_cast<double>(rc[i]);
                due = std::fma(u, ve, due);
                dce = std::fma(ve, cc, dce);
                n2e = std::fma(ve, ve, n2e);
                dew = std::fma(wc, ve, dew);
                duw = std::fma(wc, u, duw);
                if (k == 7) n2w = n2w + wc * wc;
                else n2w = std::fma(wc, wc, n2w);
                dcw = std::fma(wc, cc, dcw);
            }
            const uint16_t* bd = &dg[static_cast<size_t>(blk) * 8];
            rec[kHdr + static_cast<size_t>(blk) * 32 + lane] = RowAByte(bd, br, bits);
            if (bits >= 2) RefineField(bd, br, bits, rrec + lane * run + static_cast<size_t>(blk) * cbits);
        }
        double g1 = 0.0, h3 = 0.0;
        if (n2e > 1e-30 && due > 1e-12) {
            g1 = rn / due;
            h3 = std::fma(2.0 * g1, dce, rn2);
        }
        float* g1b = reinterpret_cast<float*>(rrec + ix.RefG1Off());
        float* h3b = reinterpret_cast<float*>(rrec + ix.RefH3Off());
        g1b[lane] = static_cast<float>(g1);
        h3b[lane] = static_cast<float>(h3);
        const double g1v = static_cast<double>(g1b[lane]);
        const double sqrt_dm1 = std::sqrt(static_cast<double>(std::max<size_t>(dim - 1, 1)));
        const double n2w_safe = std::max(n2w, 1e-30);
        const double uw_eps = 1e-12 * std::sqrt(n2w_safe);
        double w_use, perp;
        if (!hilbert && std::abs(duw) > uw_eps) {
            const double wx = 2.0 * rn / duw;
            const double g4 = g1v * 4.0;
            const double rx2 = std::fma(-(wx * g4), dew, std::fma(g1v * g4, n2e, n2w * (wx * wx)));
            w_use = wx;
            perp = std::sqrt(std::max(0.0, rx2)) / sqrt_dm1;
        } else {
            const double gamma = dew / n2w_safe;
            const double vp2 = std::max(std::fma(-n2w, gamma * gamma, n2e), 0.0);
            w_use = 2.0 * g1v * gamma;
            perp = 2.0 * g1v * std::sqrt(vp2) / sqrt_dm1;
        }
        ow[slot] = static_cast<float>(w_use);
        ob[slot] = static_cast<float>(std::fma(w_use, dcw, rn2));
        op[slot] = static_cast<float>(perp);
    }
    for (size_t c = 0; c < nlist; ++c) {
        const size_t cnt = counts[c], nb = (cnt + 31) / 32, vb = ix.cluster_begin[c];
        for (size_t b = 0; b < nb; ++b) {
            const size_t lo = b * 32, hi = std::min(cnt, lo + 32);
            WriteBanks(&ob[vb + lo], &ow[vb + lo], &op[vb + lo], hi - lo,
                       ix.arena.data() + ix.arena_begin[c] + b * ix.batch_bytes);
        }
    }
}

// ---------------------------------------------------------------------------
// Query side


// PhiInv / OrthoQuantile (e8q_internal.hpp:77-105); `Phi - p` is contracted.
static double PhiInv(double p) {
    if (p <= 0.0) return -8.0;
    if (p >= 1.0) return 8.0;
    const double t = std::min(p, 1.0 - p);
    double x = std::sqrt(-2.0 * std::log(t));
    if (p < 0.5) x = -x;
    for (int it = 0; it < 12; ++it) {
        const double phi = std::exp(x * (x * -0.5)) * 0.3989422804014327;
        const double e = std::erf(x * 0.7071067811865476);
        x -= std::fma(e + 1.0, 0.5, -p) / phi;
    }
    return x;
}

float OrthoQuantile(double alpha, size_t k_eff) {
    const double tail = std::min(alpha / static_cast<double>(std::max<size_t>(k_eff, 1)), 0.49);
    return static_cast<float>(std::max(PhiInv(1.0 - tail), 1.0));
}

struct Prep {
    std::vector<float> q;        // rotated, centered
    std::vector<float> cf;       // blocks*64 float LUT
    std::vector<uint8_t> lut;    // blocks*64 u8
    float delta = 0, base = 0, lut_err = 0, score_max = 0;
};

void RotateQuery(const Index& ix, const float* query, float* out) {
    ix.rot.Rotate(query, out);
    for (size_t d = 0; d < ix.padded; ++d) out[d] -= ix.mean_rot[d];
}

// PrepareQuery (e8q_search.cpp:20-73) with the AVX2 quantizer (a fused
// x*qinv + qoff, round-to-nearest-even, clamp).
void PrepareQuery(const Index& ix, const float* q, Prep& p) {
    const int B = ix.blocks;
    p.cf.assign(static_cast<size_t>(B) * 64, 0.0F);
    for (int blk = 0; blk < B; ++blk) {
        const float* qb = q + static_cast<size_t>(blk) * 8;
        float* A = p.cf.data() + static_cast<size_t>(blk) * 64;
        float* Bt = A + 32;
        for (int v = 0; v < 16; ++v) {
            double av = 0.0, bv = 0.0;
            for (int h = 0; h < 4; ++h) {
                const int bit = (v >> (3 - h)) & 1;
                av = std::fma(static_cast<double>(qb[4 + h]), ix.gate_mid[2 * (4 + h) + bit], av);
                bv = std::fma(static_cast<double>(qb[h]), ix.gate_mid[2 * h + bit], bv);
            }
            A[v] = A[16 + v] = static_cast<float>(av);
            Bt[v] = Bt[16 + v] = static_cast<float>(bv);
        }
    }
    float hc = 0.0F;
    for (float v : p.cf) hc = std::max(hc, std::abs(v));
    p.delta = hc > 0.0F ? 2.0F * hc / 255.0F : 1.0F;
    const float qinv = 1.0F / p.delta, qoff = hc / p.delta;
    p.lut.resize(p.cf.size());
    for (size_t i = 0; i < p.cf.size(); ++i) {
        const float x = std::fma(p.cf[i], qinv, qoff);
        float r = std::nearbyint(x);
        int v = (r >= -2147483648.0F && r < 2147483648.0F) ? static_cast<int>(r) : INT32_MIN;
        p.lut[i] = static_cast<uint8_t>(std::clamp(v, 0, 255));
    }
    p.base = -hc * static_cast<float>(2 * B);
    const float sb = std::sqrt(static_cast<float>(B));
    p.lut_err = p.delta * (sb * 2.0F);
    p.score_max = static_cast<float>(B) * std::fma(p.delta, 255.0F, hc * 2.0F);
}

// Query norm and route distances in the AVX2 reduction order
// (kernels_avx2.cpp:8-103): dot = Hsum8(((a0+a1)+(a2+a3))) with a_j the
// FMA chains over 32-float strides; dist = (qn2 + cn2) - 2*dot fused.
float QueryNorm2(const float* q, size_t P) {
    float a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t d = 0; d < P; d += 8)
        for (int l = 0; l < 8; ++l) a[l] = std::fma(q[d + l], q[d + l], a[l]);
    return Hsum8(a);
}

float RouteDot(const float* q, const float* c, size_t P) {
    float a[4][8] = {};
    size_t d = 0;
    for (; d + 32 <= P; d += 32)
        for (int j = 0; j < 4; ++j)
            for (int l = 0; l < 8; ++l) a[j][l] = std::fma(q[d + 8 * j + l], c[d + 8 * j + l], a[j][l]);
    float s[8];
    for (int l = 0; l < 8; ++l) s[l] = (a[0][l] + a[1][l]) + (a[2][l] + a[3][l]);
    float dot = Hsum8(s);
    for (; d + 8 <= P; d += 8) {
        float m[8];
        for (int l = 0; l < 8; ++l) m[l] = q[d + l] * c[d + l];
        dot += Hsum8(m);
    }
    for (; d < P; ++d) dot = std::fma(q[d], c[d], dot);
    return dot;
}

void Route(const Index& ix, const float* q, size_t nprobe, float* dist, std::vector<uint32_t>& probe) {
    const size_t P = ix.padded, L = ix.nlist;
    const float qn2 = QueryNorm2(q, P);
    std::vector<uint64_t> keys(L);
    for (size_t c = 0; c < L; ++c) {
        const float dot = RouteDot(q, &ix.centroid_rot[c * P], P);
        float d2 = std::fma(-dot, 2.0F, ix.centroid_n2[c] + qn2);
        if (d2 < 0.0F) d2 = 0.0F;
        dist[c] = d2;
        uint32_t bitsv;
        std::memcpy(&bitsv, &d2, 4);
        keys[c] = (static_cast<uint64_t>(bitsv) << 32) | c;
    }
    const size_t take = std::min(nprobe, L);
    std::partial_sort(keys.begin(), keys.begin() + static_cast<long>(take), keys.end());
    probe.resize(take);
    for (size_t i = 0; i < take; ++i) probe[i] = static_cast<uint32_t>(keys[i]);
}

// Integer coarse accumulate (scan_kernels_generic.cpp:14-35).
void CoarseAccumulate(const uint8_t* rowa, const uint8_t* lut, size_t blocks, int32_t out[32]) {
    for (int t = 0; t < 32; ++t) out[t] = 0;
    for (size_t b = 0; b < blocks; ++b) {
        const uint8_t* lc = lut + b * 64;
        for (int t = 0; t < 32; ++t) {
            const uint8_t h = rowa[b * 32 + t];
            const size_t half = static_cast<size_t>(t >> 4) * 16;
            out[t] += lc[half + (h & 15)] + lc[32 + half + (h >> 4)];
        }
    }
}

static inline bool Better(float d1, uint32_t i1, float d2, uint32_t i2) {
    return d1 < d2 || (d1 == d2 && i1 < i2);
}

struct TopK {   // keeps the k best (dist, id) under Better
    size_t k = 0;
    std::vector<std::pair<float, uint32_t>> h;  // max-heap on Better
    static bool Cmp(const std::pair<float, uint32_t>& a, const std::pair<float, uint32_t>& b) {
        return Better(a.first, a.second, b.first, b.second);
    }
    float Limit() const { return h.size() == k ? h.front().first : std::numeric_limits<float>::infinity(); }
    void Insert(float d, uint32_t id) {
        if (h.size() < k) {
            h.push_back({d, id});
            std::push_heap(h.begin(), h.end(), Cmp);
            return;
        }
        if (!Better(d, id, h.front().first, h.front().second)) return;
        std::pop_heap(h.begin(), h.end(), Cmp);
        h.back() = {d, id};
        std::push_heap(h.begin(), h.end(), Cmp);
    }
};

struct Row {
    const uint8_t* refine;
    const uint8_t* plane;   // rowA base of the batch
    uint32_t lane;
    float g1, h3;           // h3 with the cluster d2 folded in
    uint32_t id;
};

// Digits and branch of block `blk` of a row (DecodeSplitBlockLane,
// e8q_scan.cpp:34-90): d_i = rowA bit (7-i) << (b-1) | field_i, digit 7's
// field LSB the parity of digits 0..6's, its stored LSB the branch.
void DecodeBlock(const Index& ix, const Row& r, int blk, int d[8], int* br) {
    const int bits = ix.bits, c = bits - 1;
    const uint8_t rb = r.plane[static_cast<size_t>(blk) * 32 + r.lane];
    if (bits == 1) {
        int par = 0;
        for (int i = 0; i < 7; ++i) {
            d[i] = (rb >> (7 - i)) & 1;
            par ^= d[i];
        }
        d[7] = par;
        *br = rb & 1;
        return;
    }
    unsigned __int128 x = 0;
    for (int i = 0; i < c; ++i)
        x |= static_cast<unsigned __int128>(r.refine[static_cast<size_t>(blk) * c + i]) << (8 * i);
    const unsigned __int128 m = (static_cast<unsigned __int128>(1) << c) - 1;
    int par = 0;
    for (int i = 0; i < 7; ++i) {
        const int f = static_cast<int>((x >> (c * i)) & m);
        par ^= f & 1;
        d[i] = f | (((rb >> (7 - i)) & 1) << c);
    }
    const int f7 = static_cast<int>((x >> (c * 7)) & m);
    *br = f7 & 1;
    d[7] = (f7 & ~1) | par | ((rb & 1) << c);
}

// Exact per-row score in the AVX2 refine order (e8q_search_avx2.cpp:82-403):
// component i of even blocks in acc[i], of odd blocks in acc2[i], FMA
// chains, then Hsum8(acc + acc2).
float RefineScore(const Index& ix, const Row& r, const float* q) {
    const size_t stride = ix.bits >= 9 ? (size_t{1} << ix.bits) : 256;
    float acc[8] = {}, acc2[8] = {};
    for (int blk = 0; blk < ix.blocks; ++blk) {
        int d[8];
        int br;
        DecodeBlock(ix, r, blk, d, &br);
        const float* tab = ix.refine_tab.data() + static_cast<size_t>(br) * stride;
        float* dst = (blk & 1) ? acc2 : acc;
        for (int i = 0; i < 8; ++i) dst[i] = std::fma(tab[d[i]], q[blk * 8 + i], dst[i]);
    }
    float s[8];
    for (int i = 0; i < 8; ++i) s[i] = acc[i] + acc2[i];
    return Hsum8(s);
}

// The same score in fp64: the float table entries and query components,
// exact products, one fp64 sum in dimension order.
double RefineScore64(const Index& ix, const Row& r, const float* q) {
    const size_t stride = ix.bits >= 9 ? (size_t{1} << ix.bits) : 256;
    double s = 0.0;
    for (int blk = 0; blk < ix.blocks; ++blk) {
        int d[8];
        int br;
        DecodeBlock(ix, r, blk, d, &br);
        const float* tab = ix.refine_tab.data() + static_cast<size_t>(br) * stride;
        for (int i = 0; i < 8; ++i) s += static_cast<double>(tab[d[i]]) * static_cast<double>(q[blk * 8 + i]);
    }
    return s;
}

struct Stats {
    uint64_t candidates = 0, survivors = 0;
};

size_t FlushEvery() {
    const char* e = std::getenv("E8Q_FLUSH_EVERY");
    const long v = e != nullptr ? std::strtol(e, nullptr, 10) : 0;
    return v > 0 ? static_cast<size_t>(v) : 32;
}

// Search (e8q_search.cpp:215-318 with ScanClusterAvx2, e8q_search_avx2.cpp:407-598).
void Search(const Index& ix, const float* query, size_t k, size_t nprobe, bool disable_prune, float alpha,
            uint32_t* out_ids, float* out_dists, Stats* st, Prep* prep_out, std::vector<uint32_t>* probe_out) {
    if (k > kMaxK) throw std::invalid_argument("k must be <= 128");
    Prep p;
    p.q.assign(ix.padded, 0.0F);
    RotateQuery(ix, query, p.q.data());
    PrepareQuery(ix, p.q.data(), p);
    std::vector<float> cdist(ix.nlist);
    std::vector<uint32_t> probe;
    Route(ix, p.q.data(), nprobe, cdist.data(), probe);
    const float z = OrthoQuantile(alpha, std::max<size_t>(4 * k, 32));
    TopK heap;
    heap.k = k;
    const size_t flush_every = FlushEvery();
    std::vector<Row> staged;
    auto flush = [&]() {
        for (const Row& r : staged) {
            const float S = RefineScore(ix, r, p.q.data());
            heap.Insert(std::fma(-(2.0F * r.g1), S, r.h3), r.id);
        }
        staged.clear();
    };
    for (uint32_t cidx : probe) {
        const float d2 = cdist[cidx];
        const uint64_t vb = ix.cluster_begin[cidx];
        const size_t cnt = ix.cluster_begin[cidx + 1] - vb;
        if (cnt == 0) continue;
        const float sqrt_d2 = std::sqrt(std::max(0.0F, d2));
        for (size_t off = 0; off < cnt; off += 32) {
            const size_t nl = std::min<size_t>(32, cnt - off);
            const uint8_t* bp = ix.arena.data() + ix.arena_begin[cidx] + (off / 32) * ix.batch_bytes;
            const uint8_t* rr = ix.ref_arena.data() + ix.ref_arena_begin[cidx] + (off / 32) * ix.ref_batch_bytes;
            int32_t acc[32];
            CoarseAccumulate(bp + kHdr, p.lut.data(), static_cast<size_t>(ix.blocks), acc);
            if (st) st->candidates += nl;
            const float distk0 = heap.Limit();
            const float rh = distk0 < d2 ? std::sqrt(std::max(0.0F, distk0)) : sqrt_d2;
            float sc[8];
            std::memcpy(sc, bp + 96, 32);
            const float pad_w_round = (0.5F * sc[3]) * (p.score_max + p.lut_err);
            if (!disable_prune && distk0 < std::numeric_limits<float>::infinity()) {
                const float pad_add = std::fma(p.lut_err, sc[6], pad_w_round);
                const float perp_amax = std::fma(sc[5], 255.0F, sc[4]);
                int32_t mx = acc[0], mn = acc[0];
                for (int t = 1; t < 32; ++t) { mx = std::max(mx, acc[t]); mn = std::min(mn, acc[t]); }
                const float s_hi = std::fma(static_cast<float>(mx), p.delta, p.base);
                const float s_lo = std::fma(static_cast<float>(mn), p.delta, p.base);
                const float sa = std::abs(s_hi) < std::abs(s_lo) ? std::abs(s_lo) : std::abs(s_hi);
                const float lb = std::fma(-perp_amax, rh * z,
                                          std::fma(-sc[6], sa, (sc[0] + d2) - pad_add));
                if (lb >= distk0) continue;
            }
            const float vbase = (sc[0] + d2) - pad_w_round;
            const float slack = z * rh;
            float dlo[32];
            for (int t = 0; t < 32; ++t) {
                const float scv = std::fma(static_cast<float>(acc[t]), p.delta, p.base);
                const float bt = std::fma(static_cast<float>(bp[t]), sc[1], vbase);
                const float w = std::fma(static_cast<float>(bp[32 + t]), sc[3], sc[2]);
                const float pp = std::fma(static_cast<float>(bp[64 + t]), sc[5], sc[4]);
                const float dc = std::fma(-w, scv, bt);
                dlo[t] = std::fma(-slack, pp, std::fma(-std::abs(w), p.lut_err, dc));
            }
            uint32_t mask = 0;
            for (int t = 0; t < 32; ++t)
                if (disable_prune || dlo[t] < distk0) mask |= 1u << t;
            if (nl < 32) mask &= (1u << nl) - 1u;
            const float* g1b = reinterpret_cast<const float*>(rr + ix.RefG1Off());
            const float* h3b = reinterpret_cast<const float*>(rr + ix.RefH3Off());
            while (mask != 0) {
                const int t = __builtin_ctz(mask);
                mask &= mask - 1;
                if (!disable_prune && !(dlo[t] < heap.Limit())) continue;
                if (staged.size() >= flush_every) flush();
                if (st) ++st->survivors;
                staged.push_back(Row{ix.bits >= 2 ? rr + static_cast<size_t>(t) * ix.RefineStride() : nullptr,
                                     bp + kHdr, static_cast<uint32_t>(t), g1b[t], d2 + h3b[t],
                                     ix.ids[vb + off + static_cast<size_t>(t)]});
            }
        }
        flush();
    }
    flush();
    std::vector<std::pair<float, uint32_t>> out = heap.h;
    std::sort(out.begin(), out.end(), TopK::Cmp);
    for (size_t i = 0; i < k; ++i) {
        out_ids[i] = i < out.size() ? out[i].second : std::numeric_limits<uint32_t>::max();
        if (out_dists) out_dists[i] = i < out.size() ? out[i].first : std::numeric_limits<float>::infinity();
    }
    if (prep_out) *prep_out = p;
    if (probe_out) *probe_out = probe;
}

// Estimator over every stored vector of the given clusters, no pruning:
// the top-k the pruned search converges to.  d2 comes from Route's distances.
void ScanAllTopK(const Index& ix, const float* query, size_t k, size_t nprobe, uint32_t* out_ids, float* out_dists) {
    std::vector<float> q(ix.padded);
    RotateQuery(ix, query, q.data());
    std::vector<float> cdist(ix.nlist);
    std::vector<uint32_t> probe;
    Route(ix, q.data(), nprobe, cdist.data(), probe);
    std::vector<std::pair<float, uint32_t>> all;
    for (uint32_t c : probe) {
        const uint64_t vb = ix.cluster_begin[c];
        const size_t cnt = ix.cluster_begin[c + 1] - vb;
        for (size_t i = 0; i < cnt; ++i) {
            const uint8_t* bp = ix.arena.data() + ix.arena_begin[c] + (i / 32) * ix.batch_bytes;
            const uint8_t* rr = ix.ref_arena.data() + ix.ref_arena_begin[c] + (i / 32) * ix.ref_batch_bytes;
            const size_t t = i % 32;
            const float g1 = reinterpret_cast<const float*>(rr + ix.RefG1Off())[t];
            const float h3 = reinterpret_cast<const float*>(rr + ix.RefH3Off())[t];
            Row r{ix.bits >= 2 ? rr + t * ix.RefineStride() : nullptr, bp + kHdr, static_cast<uint32_t>(t), g1,
                  cdist[c] + h3, ix.ids[vb + i]};
            all.push_back({std::fma(-(2.0F * g1), RefineScore(ix, r, q.data()), r.h3), r.id});
        }
    }
    const size_t take = std::min(k, all.size());
    std::partial_sort(all.begin(), all.begin() + static_cast<long>(take), all.end(), TopK::Cmp);
    for (size_t i = 0; i < k; ++i) {
        out_ids[i] = i < take ? all[i].second : std::numeric_limits<uint32_t>::max();
        out_dists[i] = i < take ? all[i].first : std::numeric_limits<float>::infinity();
    }
}

// The estimator of ScanAllTopK in fp64 for given ids of the probed lists:
// ((double) d2_c + h3) - 2 g1 S64 (NaN for an id outside them).
void Dist64(const Index& ix, const float* query, size_t nprobe, const uint32_t* ids, size_t n, double* out) {
    std::vector<float> q(ix.padded);
    RotateQuery(ix, query, q.data());
    std::vector<float> cdist(ix.nlist);
    std::vector<uint32_t> probe;
    Route(ix, q.data(), nprobe, cdist.data(), probe);
    std::unordered_map<uint32_t, size_t> want;
    for (size_t j = 0; j < n; ++j) {
        want[ids[j]] = j;
        out[j] = std::numeric_limits<double>::quiet_NaN();
    }
    for (uint32_t c : probe) {
        const uint64_t vb = ix.cluster_begin[c];
        const size_t cnt = ix.cluster_begin[c + 1] - vb;
        for (size_t i = 0; i < cnt; ++i) {
            const auto it = want.find(ix.ids[vb + i]);
            if (it == want.end()) continue;
            const uint8_t* bp = ix.arena.data() + ix.arena_begin[c] + (i / 32) * ix.batch_bytes;
            const uint8_t* rr = ix.ref_arena.data() + ix.ref_arena_begin[c] + (i / 32) * ix.ref_batch_bytes;
            const size_t t = i % 32;
            const float g1 = reinterpret_cast<const float*>(rr + ix.RefG1Off())[t];
            const float h3 = reinterpret_cast<const float*>(rr + ix.RefH3Off())[t];
            Row r{ix.bits >= 2 ? rr + t * ix.RefineStride() : nullptr, bp + kHdr, static_cast<uint32_t>(t), g1, h3,
                  ix.ids[vb + i]};
            const double S = RefineScore64(ix, r, q.data());
            out[it->second] = (static_cast<double>(cdist[c]) + static_cast<double>(h3)) - 2.0 * static_cast<double>(g1) * S;
        }
    }
}

// ---------------------------------------------------------------------------
// Diagnostics: ExhaustiveTopK, Estimate, ProbeBoundW (e8q_oracle.cpp:14-219).

// ExactScore (e8q_search.cpp:87-126).  b = 1: the AVX2 form
// (e8q_search_avx2.cpp:53-80), one FMA chain per component over the blocks
// then the horizontal sum; b >= 2: one fp64 chain in dimension order (GCC
// contracts `acc += q * tab`).
float ExactScore(const Index& ix, const Row& r, const float* q) {
    const size_t stride = ix.bits >= 9 ? (size_t{1} << ix.bits) : 256;
    if (ix.bits == 1) {
        float acc[8] = {};
        for (int blk = 0; blk < ix.blocks; ++blk) {
            int d[8];
            int br;
            DecodeBlock(ix, r, blk, d, &br);
            for (int i = 0; i < 8; ++i)
                acc[i] = std::fma(ix.refine_tab[static_cast<size_t>(br) * stride + d[i]], q[blk * 8 + i], acc[i]);
        }
        return Hsum8(acc);
    }
    double acc = 0.0;
    for (int blk = 0; blk < ix.blocks; ++blk) {
        int d[8];
        int br;
        DecodeBlock(ix, r, blk, d, &br);
        const float* tab = ix.refine_tab.data() + static_cast<size_t>(br) * stride;
        for (int i = 0; i < 8; ++i)
            acc = std::fma(static_cast<double>(q[blk * 8 + i]), static_cast<double>(tab[d[i]]), acc);
    }
    return static_cast<float>(acc);
}

// (query, cluster) distance of ExhaustiveTopK / Estimate: |q|^2 and <q, c>
// summed in dimension order, (qn2 + cn2) - 2 dot, clamped at 0.  GCC
// vectorizes the two sums as in-order reductions: the products of each run
// of 8 (then 4) dimensions are rounded on their own and added one by one;
// only the last P % 4 dimensions are fused multiply-adds.
float ClusterD2Seq(const Index& ix, const float* q, size_t c) {
    const float* cc = &ix.centroid_rot[c * ix.padded];
    const size_t P = ix.padded, body = P & ~size_t{3};
    float qn2 = 0.0F, dot = 0.0F;
    for (size_t d = 0; d < body; ++d) {
        const float pq = q[d] * q[d], pc = q[d] * cc[d];
        qn2 = qn2 + pq;
        dot = dot + pc;
    }
    for (size_t d = body; d < P; ++d) {
        qn2 = std::fma(q[d], q[d], qn2);
        dot = std::fma(q[d], cc[d], dot);
    }
    float d2 = std::fma(-2.0F, dot, qn2 + ix.centroid_n2[c]);
    return d2 < 0.0F ? 0.0F : d2;
}

// RowsAt (e8q.hpp:289-306): vector i of cluster c.
static Row RowAt(const Index& ix, size_t c, size_t i) {
    const uint8_t* bp = ix.arena.data() + ix.arena_begin[c] + (i / 32) * ix.batch_bytes;
    const uint8_t* rr = ix.ref_arena.data() + ix.ref_arena_begin[c] + (i / 32) * ix.ref_batch_bytes;
    const size_t t = i % 32;
    return Row{ix.bits >= 2 ? rr + t * ix.RefineStride() : nullptr, bp + kHdr, static_cast<uint32_t>(t),
               reinterpret_cast<const float*>(rr + ix.RefG1Off())[t],
               reinterpret_cast<const float*>(rr + ix.RefH3Off())[t], ix.ids[ix.cluster_begin[c] + i]};
}

// ExhaustiveTopK (e8q_oracle.cpp:14-56): every vector, d2 + h3 - 2 g1 S
// (fused), the k best under Better (distance, then id).
void ExhaustiveTopK(const Index& ix, const float* query, size_t k, uint32_t* out_ids, float* out_dists) {
    std::vector<float> q(ix.padded);
    RotateQuery(ix, query, q.data());
    std::vector<std::pair<float, uint32_t>> all;
    for (size_t c = 0; c < ix.nlist; ++c) {
        const float d2 = ClusterD2Seq(ix, q.data(), c);
        const size_t cnt = ix.cluster_begin[c + 1] - ix.cluster_begin[c];
        for (size_t i = 0; i < cnt; ++i) {
            const Row r = RowAt(ix, c, i);
            all.push_back({std::fma(-(2.0F * r.g1), ExactScore(ix, r, q.data()), d2 + r.h3), r.id});
        }
    }
    const size_t take = std::min(k, all.size());
    std::partial_sort(all.begin(), all.begin() + static_cast<long>(take), all.end(), TopK::Cmp);
    for (size_t i = 0; i < k; ++i) {
        out_ids[i] = i < take ? all[i].second : std::numeric_limits<uint32_t>::max();
        out_dists[i] = i < take ? all[i].first : std::numeric_limits<float>::infinity();
    }
}

// Position of global id `id`: the last stored vector carrying it (the
// reference's inv_ids_ map, e8q_oracle.cpp:63-74); npos when none.
static size_t PositionOf(const Index& ix, uint32_t id) {
    size_t pos = static_cast<size_t>(-1);
    if (id >= ix.n) return pos;
    for (size_t i = 0; i < ix.n; ++i)
        if (ix.ids[i] == id) pos = i;
    return pos;
}

static size_t ClusterOf(const Index& ix, size_t pos) {
    const auto it = std::upper_bound(ix.cluster_begin.begin() + 1, ix.cluster_begin.end(), pos);
    return static_cast<size_t>(it - ix.cluster_begin.begin() - 1);
}

// Estimate (e8q_oracle.cpp:58-101); returns false for an unknown id.
bool Estimate(const Index& ix, const float* query, uint32_t id, float* out) {
    const size_t pos = PositionOf(ix, id);
    if (pos == static_cast<size_t>(-1)) return false;
    const size_t c = ClusterOf(ix, pos), local = pos - ix.cluster_begin[c];
    std::vector<float> q(ix.padded);
    RotateQuery(ix, query, q.data());
    const float d2 = ClusterD2Seq(ix, q.data(), c);
    const Row r = RowAt(ix, c, local);
    *out = std::fma(-(2.0F * r.g1), ExactScore(ix, r, q.data()), d2 + r.h3);
    return true;
}

// DecodeSplitBlockLane (e8q_scan.cpp:34-61) as ProbeBoundW calls it: the
// 8-bit-digit form for every rate.  At b = 1 the (empty) field makes digit 7
// the rowA byte's bit 0 and the branch 0; at b = 9 the digits keep their low
// 8 bits.  (At b >= 10 the reference's 8-byte staging is overrun: undefined.)
static void DecodeLiteral(int bits, uint8_t rb, const uint8_t* field, uint8_t d[8], int* br) {
    const int c = bits - 1;
    uint64_t x = 0;
    for (int i = 0; i < c; ++i) x |= static_cast<uint64_t>(field[i]) << (8 * i);
    const uint64_t m = (uint64_t{1} << c) - 1;
    int par = 0;
    for (int i = 0; i < 7; ++i) {
        const uint64_t f = (x >> (c * i)) & m;
        d[i] = static_cast<uint8_t>(f | (static_cast<uint64_t>((rb >> (7 - i)) & 1) << c));
        par ^= static_cast<int>(f & 1u);
    }
    const uint64_t f7 = (x >> (c * 7)) & m;
    d[7] = static_cast<uint8_t>(((f7 >> 1) << 1) | static_cast<uint64_t>(par) | (static_cast<uint64_t>(rb & 1) << c));
    *br = static_cast<int>(f7 & 1u);
}

// ProbeBoundW (e8q_oracle.cpp:103-219): the normalized gap between the
// coarse (class-mean) and the exact estimate of one stored vector from its
// build row, (d_coarse - d_full) / (r_q Omega); NaN on a zero residual or a
// degenerate code.  Returns false for an unknown id (or b >= 10).
bool ProbeBoundW(const Index& ix, const float* query, const float* base_row, uint32_t id, double* out) {
    const size_t pos = PositionOf(ix, id);
    if (pos == static_cast<size_t>(-1) || ix.bits >= 10) return false;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    const size_t c = ClusterOf(ix, pos), local = pos - ix.cluster_begin[c];
    const size_t P = ix.padded;
    const float* cc = &ix.centroid_rot[c * P];
    std::vector<float> rq(P), rx(P);
    RotateQuery(ix, query, rq.data());
    RotateQuery(ix, base_row, rx.data());
    double d2 = 0.0, rn2 = 0.0;
    std::vector<double> unit(P);
    for (size_t d = 0; d < P; ++d) {
        const double qr = static_cast<double>(rq[d]) - static_cast<double>(cc[d]);
        const double xr = static_cast<double>(rx[d]) - static_cast<double>(cc[d]);
        d2 = std::fma(qr, qr, d2);
        rn2 = std::fma(xr, xr, rn2);
        unit[d] = xr;
    }
    const double rn = std::sqrt(rn2);
    if (!(rn > 0.0)) {
        *out = nan;
        return true;
    }
    for (size_t d = 0; d < P; ++d) unit[d] /= rn;
    const std::vector<double> lev[2] = {BranchLevels(ix.bits, 0), BranchLevels(ix.bits, 1)};
    const double rd = std::sqrt(static_cast<double>(P));
    const Row r = RowAt(ix, c, local);
    double dot_ue = 0, n2_e = 0, dot_ce = 0, dot_ew = 0, dot_uw = 0, n2_w = 0, dot_cw = 0, s_exact = 0, s_coarse = 0;
    const int cb = ix.bits - 1;
    for (int blk = 0; blk < ix.blocks; ++blk) {
        uint8_t dg[8];
        int br;
        DecodeLiteral(ix.bits, r.plane[static_cast<size_t>(blk) * 32 + r.lane],
                      ix.bits >= 2 ? r.refine + static_cast<size_t>(blk) * cb : nullptr, dg, &br);
        for (int k = 0; k < 8; ++k) {
            const size_t i = static_cast<size_t>(blk) * 8 + static_cast<size_t>(k);
            const double ve = lev[br][dg[k]] / rd;
            const double wc = ix.gate_mid[2 * k + (dg[k] >> cb)];
            dot_ue = std::fma(unit[i], ve, dot_ue);
            n2_e = std::fma(ve, ve, n2_e);
            dot_ce = std::fma(static_cast<double>(cc[i]), ve, dot_ce);
            dot_ew = std::fma(ve, wc, dot_ew);
            dot_uw = std::fma(unit[i], wc, dot_uw);
            n2_w = std::fma(wc, wc, n2_w);
            dot_cw = std::fma(static_cast<double>(cc[i]), wc, dot_cw);
            s_exact = std::fma(static_cast<double>(rq[i]), ve, s_exact);
            s_coarse = std::fma(static_cast<double>(rq[i]), wc, s_coarse);
        }
    }
    if (!(std::abs(dot_ue) > 0.0) || !(std::abs(dot_uw) > 0.0)) {
        *out = nan;
        return true;
    }
    const double g1 = rn / dot_ue;
    const double h3 = std::fma(2.0 * g1, dot_ce, rn2);
    const double d_full = std::fma(-(2.0 * g1), s_exact, d2 + h3);
    const double w_x = 2.0 * rn / dot_uw;
    const double d_coarse = std::fma(-w_x, s_coarse, std::fma(w_x, dot_cw, d2 + rn2));
    const double rx2 = std::fma(-((4.0 * g1) * w_x), dot_ew, std::fma((4.0 * g1) * g1, n2_e, (w_x * w_x) * n2_w));
    const double sqrt_dm1 = std::sqrt(static_cast<double>(std::max<size_t>(ix.dim - 1, 1)));
    const double omega = std::sqrt(std::max(0.0, rx2)) / sqrt_dm1;
    const double r_q = std::sqrt(std::max(0.0, d2));
    if (!(omega > 0.0) || !(r_q > 0.0)) {
        *out = nan;
        return true;
    }
    *out = (d_coarse - d_full) / (r_q * omega);
    return true;
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C API for the Python tests (ctypes).  Handles are heap Index objects.

extern "C" {

double orc_fixed_scale(int bits) { return orc::FixedScale(bits); }

void orc_branch_levels(int bits, int branch, double* out) {
    const auto l = orc::BranchLevels(bits, branch);
    std::memcpy(out, l.data(), l.size() * sizeof(double));
}

void orc_scan_class_means(int bits, double* out2) { orc::ScanClassMeans(bits, out2); }

void orc_refine_table(int bits, size_t padded, float* out) {
    const auto t = orc::RefineTable(bits, padded);
    std::memcpy(out, t.data(), t.size() * sizeof(float));
}

size_t orc_rotator_signs(size_t dim, uint64_t seed, uint8_t* out) {
    orc::Rotator r(dim, seed);
    if (out) std::memcpy(out, r.signs.data(), r.signs.size());
    return r.signs.size();
}

void orc_rotate(size_t dim, uint64_t seed, const float* src, size_t nrows, float* dst) {
    orc::Rotator r(dim, seed);
    for (size_t i = 0; i < nrows; ++i) r.Rotate(src + i * dim, dst + i * r.padded);
}

void orc_encode(const float* unit, size_t nrows, int padded, int bits, double s, uint16_t* digits, uint8_t* branch) {
    for (size_t i = 0; i < nrows; ++i)
        orc::EncodeVector(unit + i * padded, padded, bits, s, digits + i * padded, branch + i * (padded / 8));
}

void orc_coarse_accumulate(const uint8_t* rowa, const uint8_t* lut, size_t blocks, int32_t* out) {
    orc::CoarseAccumulate(rowa, lut, blocks, out);
}

float orc_ortho_quantile(double alpha, size_t k_eff) { return orc::OrthoQuantile(alpha, k_eff); }

void* orc_build(const float* base, const float* cent, const uint32_t* cid, size_t n, size_t dim, size_t nlist,
                int bits, uint64_t seed, float prune_alpha) {
    auto* ix = new orc::Index();
    try {
        orc::Build(*ix, base, cent, cid, n, dim, nlist, bits, seed, prune_alpha);
    } catch (...) {
        delete ix;
        return nullptr;
    }
    return ix;
}

void orc_free(void* h) { delete static_cast<orc::Index*>(h); }

// Sizes: [n, dim, padded, nlist, bits, blocks, arena_bytes, ref_arena_bytes,
//         batch_bytes, ref_batch_bytes]
void orc_index_info(void* h, uint64_t* out10) {
    const auto* ix = static_cast<orc::Index*>(h);
    const uint64_t v[10] = {ix->n, ix->dim, ix->padded, ix->nlist, static_cast<uint64_t>(ix->bits),
                            static_cast<uint64_t>(ix->blocks), ix->arena.size(), ix->ref_arena.size(),
                            ix->batch_bytes, ix->ref_batch_bytes};
    std::memcpy(out10, v, sizeof(v));
}

// Copies the E8AX arrays out (any pointer may be null).
void orc_index_arrays(void* h, uint32_t* ids, uint8_t* signs, float* mean_rot, float* centroid_rot,
                      float* centroid_n2, uint64_t* cluster_begin, uint64_t* arena_begin, uint8_t* arena,
                      uint64_t* ref_arena_begin, uint8_t* ref_arena) {
    const auto* ix = static_cast<orc::Index*>(h);
    auto cp = [](void* dst, const void* src, size_t bytes) {
        if (dst) std::memcpy(dst, src, bytes);
    };
    cp(ids, ix->ids.data(), ix->ids.size() * 4);
    cp(signs, ix->rot.signs.data(), ix->rot.signs.size());
    cp(mean_rot, ix->mean_rot.data(), ix->mean_rot.size() * 4);
    cp(centroid_rot, ix->centroid_rot.data(), ix->centroid_rot.size() * 4);
    cp(centroid_n2, ix->centroid_n2.data(), ix->centroid_n2.size() * 4);
    cp(cluster_begin, ix->cluster_begin.data(), ix->cluster_begin.size() * 8);
    cp(arena_begin, ix->arena_begin.data(), ix->arena_begin.size() * 8);
    cp(arena, ix->arena.data(), ix->arena.size());
    cp(ref_arena_begin, ix->ref_arena_begin.data(), ix->ref_arena_begin.size() * 8);
    cp(ref_arena, ix->ref_arena.data(), ix->ref_arena.size());
}

// Builds an oracle index over externally supplied E8AX arrays (a loaded file).
void* orc_from_arrays(uint64_t n, uint64_t dim, uint64_t nlist, int bits, const uint32_t* ids, const uint8_t* signs,
                      const float* mean_rot, const float* centroid_rot, const float* centroid_n2,
                      const uint64_t* cluster_begin, const uint64_t* arena_begin, const uint8_t* arena,
                      const uint64_t* ref_arena_begin, const uint8_t* ref_arena, float alpha) {
    auto* ix = new orc::Index();
    ix->n = n; ix->dim = dim; ix->nlist = nlist; ix->bits = bits;
    ix->rot = orc::Rotator(dim, 0);
    std::memcpy(ix->rot.signs.data(), signs, ix->rot.signs.size());
    ix->padded = ix->rot.padded;
    ix->blocks = static_cast<int>(ix->padded / 8);
    ix->InitTables();
    ix->ids.assign(ids, ids + n);
    ix->mean_rot.assign(mean_rot, mean_rot + ix->padded);
    ix->centroid_rot.assign(centroid_rot, centroid_rot + nlist * ix->padded);
    ix->centroid_n2.assign(centroid_n2, centroid_n2 + nlist);
    ix->cluster_begin.assign(cluster_begin, cluster_begin + nlist + 1);
    ix->arena_begin.assign(arena_begin, arena_begin + nlist + 1);
    ix->ref_arena_begin.assign(ref_arena_begin, ref_arena_begin + nlist + 1);
    ix->arena.assign(arena, arena + arena_begin[nlist] + 64);
    ix->ref_arena.assign(ref_arena, ref_arena + ref_arena_begin[nlist] + 64);
    ix->alpha = alpha;
    return ix;
}

void orc_set_alpha(void* h, float alpha) { static_cast<orc::Index*>(h)->alpha = alpha; }

// Sequential reference search for nq queries (OpenMP over queries when
// threads > 1, one query per iteration as the reference binding does).
int orc_search(void* h, const float* queries, size_t nq, size_t k, size_t nprobe, int disable_prune, int threads,
               uint32_t* ids, float* dists, uint64_t* stats2) {
    const auto* ix = static_cast<orc::Index*>(h);
    if (k == 0 || k > orc::kMaxK || nprobe == 0) return -1;
    uint64_t cand = 0, surv = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1) reduction(+ : cand, surv)
    for (long long i = 0; i < static_cast<long long>(nq); ++i) {
        orc::Stats st;
        orc::Search(*ix, queries + static_cast<size_t>(i) * ix->dim, k, nprobe, disable_prune != 0, ix->alpha,
                    ids + static_cast<size_t>(i) * k, dists + static_cast<size_t>(i) * k, &st, nullptr, nullptr);
        cand += st.candidates;
        surv += st.survivors;
    }
    if (stats2) {
        stats2[0] = cand;
        stats2[1] = surv;
    }
    return 0;
}

// Per-query intermediates of one search: rotated query, u8 LUT, LUT
// constants [delta, base, lut_err, score_max, z], probe list.
void orc_query_prep(void* h, const float* query, size_t k, size_t nprobe, float* rot_q, uint8_t* lut,
                    float* consts5, uint32_t* probe, float* cdist) {
    const auto* ix = static_cast<orc::Index*>(h);
    orc::Prep p;
    p.q.assign(ix->padded, 0.0F);
    orc::RotateQuery(*ix, query, p.q.data());
    orc::PrepareQuery(*ix, p.q.data(), p);
    std::vector<float> cd(ix->nlist);
    std::vector<uint32_t> pr;
    orc::Route(*ix, p.q.data(), nprobe, cd.data(), pr);
    if (rot_q) std::memcpy(rot_q, p.q.data(), ix->padded * 4);
    if (lut) std::memcpy(lut, p.lut.data(), p.lut.size());
    if (consts5) {
        consts5[0] = p.delta; consts5[1] = p.base; consts5[2] = p.lut_err; consts5[3] = p.score_max;
        consts5[4] = orc::OrthoQuantile(ix->alpha, std::max<size_t>(4 * k, 32));
    }
    if (probe) std::memcpy(probe, pr.data(), pr.size() * 4);
    if (cdist) std::memcpy(cdist, cd.data(), cd.size() * 4);
}

// fp64 distances of `k` given ids per query (NaN: not in the probed lists).
void orc_dist64(void* h, const float* queries, size_t nq, size_t nprobe, const uint32_t* ids, size_t k, double* out) {
    const auto* ix = static_cast<orc::Index*>(h);
#pragma omp parallel for schedule(dynamic, 1)
    for (long long i = 0; i < static_cast<long long>(nq); ++i)
        orc::Dist64(*ix, queries + static_cast<size_t>(i) * ix->dim, nprobe, ids + static_cast<size_t>(i) * k, k,
                    out + static_cast<size_t>(i) * k);
}

void orc_scan_all(void* h, const float* queries, size_t nq, size_t k, size_t nprobe, uint32_t* ids, float* dists) {
    const auto* ix = static_cast<orc::Index*>(h);
#pragma omp parallel for schedule(dynamic, 1)
    for (long long i = 0; i < static_cast<long long>(nq); ++i)
        orc::ScanAllTopK(*ix, queries + static_cast<size_t>(i) * ix->dim, k, nprobe, ids + static_cast<size_t>(i) * k,
                         dists + static_cast<size_t>(i) * k);
}


// Diagnostics over (query, id) pairs; status 1 = an unknown id (or b >= 10
// for ProbeBoundW).
int orc_estimate(void* h, const float* queries, const uint32_t* qidx, const uint32_t* ids, size_t npair,
                 float* out) {
    const auto& ix = *static_cast<orc::Index*>(h);
    for (size_t i = 0; i < npair; ++i)
        if (!orc::Estimate(ix, queries + static_cast<size_t>(qidx[i]) * ix.dim, ids[i], out + i)) return 1;
    return 0;
}

int orc_probe_bound_w(void* h, const float* queries, const float* base_rows, const uint32_t* qidx,
                      const uint32_t* ids, size_t npair, double* out) {
    const auto& ix = *static_cast<orc::Index*>(h);
    for (size_t i = 0; i < npair; ++i)
        if (!orc::ProbeBoundW(ix, queries + static_cast<size_t>(qidx[i]) * ix.dim, base_rows + i * ix.dim, ids[i],
                              out + i))
            return 1;
    return 0;
}

void orc_exhaustive_topk(void* h, const float* queries, size_t nq, size_t k, uint32_t* ids, float* dists) {
    const auto& ix = *static_cast<orc::Index*>(h);
    for (size_t i = 0; i < nq; ++i) orc::ExhaustiveTopK(ix, queries + i * ix.dim, k, ids + i * k, dists + i * k);
}

}  // extern "C"
