// gen_golden.cpp -- writes golden vectors from the reference e8qlib library
// (built from /root/reference by oracle/ref.mk into oracle/_ref/).  Driven by
// tests/golden/make_golden.py; runs only in the build container, where the
// reference sources exist.  Its outputs (small .npy / .e8ax files) are the
// committed fixtures under tests/golden/.
//
// Subcommands (all paths are files):
//   levels   OUT_LEVELS.npy OUT_SCALE.npy OUT_WCLS.npy
//   rotate   DIM SEED IN.npy OUT_SIGNS.npy OUT_ROT.npy
//   encode   BITS PADDED IN_UNIT.npy OUT_DIGITS.npy OUT_BRANCH.npy
//   accum    PAYLOAD.npy LUT.npy OUT_SINGLE.npy OUT_PAIR.npy
//   route    Q.npy CENT.npy NPROBE OUT_DIST.npy OUT_IDS.npy OUT_CN2.npy
//   build    DATA.npy CENT.npy IDS.npy BITS SEED ALPHA OUT.e8ax
//   search   INDEX.e8ax QUERIES.npy K NPROBE DISABLE_PRUNE OUT_PREFIX
//   mbuild   DATA.npy CENT.npy IDS.npy AVG_BITS SEED ALPHA PLAN_SPEC|- VAR.npy|- DROP_ENERGY
//            OUT.me8q OUT_PREFIX
//   msearch  INDEX.me8q QUERIES.npy K NPROBE OUT_PREFIX
//   diag     INDEX.e8ax QUERIES.npy BASE.npy PAIR_Q.npy PAIR_ID.npy K OUT_PREFIX
//   mdiag    INDEX.me8q QUERIES.npy PAIR_Q.npy PAIR_ID.npy K NPROBE OUT_PREFIX
//
// The ME8Q commands read the index's fast layout (private members) to dump
// it: the class's `private` is mapped to `public` around its header only.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include <array>
#include <mutex>
#include <utility>

#include "e8qlib/index/e8q.hpp"
#define private public
#include "e8qlib/index/me8q.hpp"
#undef private
#include "e8qlib/index/ivf_common.hpp"
#include "e8qlib/quantization/e8q_codec.hpp"
#include "e8qlib/utils/rotator.hpp"

namespace {

struct Npy {
    std::string dtype;            // numpy descr, e.g. "<f4"
    std::vector<size_t> shape;
    std::vector<char> data;
    size_t count() const {
        size_t c = 1;
        for (size_t s : shape) c *= s;
        return c;
    }
};

Npy ReadNpy(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    char magic[8];
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "\x93NUMPY", 6) != 0)
        throw std::runtime_error("not npy: " + path);
    uint32_t hlen = 0;
    if (magic[6] == 1) {
        uint16_t h16 = 0;
        if (std::fread(&h16, 2, 1, f) != 1) throw std::runtime_error("short npy");
        hlen = h16;
    } else {
        if (std::fread(&hlen, 4, 1, f) != 1) throw std::runtime_error("short npy");
    }
    std::string hdr(hlen, ' ');
    if (std::fread(&hdr[0], 1, hlen, f) != hlen) throw std::runtime_error("short npy header");
    Npy a;
    const size_t dp = hdr.find("'descr': '") + 10;
    a.dtype = hdr.substr(dp, hdr.find('\'', dp) - dp);
    if (hdr.find("'fortran_order': False") == std::string::npos) throw std::runtime_error("fortran npy");
    size_t sp = hdr.find('(', hdr.find("'shape'")) + 1;
    const size_t se = hdr.find(')', sp);
    while (sp < se) {
        while (sp < se && (hdr[sp] == ' ' || hdr[sp] == ',')) ++sp;
        if (sp >= se) break;
        a.shape.push_back(std::strtoull(hdr.c_str() + sp, nullptr, 10));
        while (sp < se && hdr[sp] != ',') ++sp;
    }
    size_t esz = static_cast<size_t>(std::atoi(a.dtype.c_str() + 2));
    a.data.resize(a.count() * esz);
    if (!a.data.empty() && std::fread(a.data.data(), 1, a.data.size(), f) != a.data.size())
        throw std::runtime_error("short npy data");
    std::fclose(f);
    return a;
}

void WriteNpy(const std::string& path, const char* descr, const std::vector<size_t>& shape, const void* p,
              size_t bytes) {
    std::string sh = "(";
    for (size_t i = 0; i < shape.size(); ++i) sh += std::to_string(shape[i]) + (shape.size() == 1 ? "," : (i + 1 < shape.size() ? ", " : ""));
    sh += ")";
    std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + sh + ", }";
    while ((10 + hdr.size() + 1) % 64 != 0) hdr += ' ';
    hdr += '\n';
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    const char magic[8] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 1, 0};
    std::fwrite(magic, 1, 8, f);
    const uint16_t hl = static_cast<uint16_t>(hdr.size());
    std::fwrite(&hl, 2, 1, f);
    std::fwrite(hdr.data(), 1, hdr.size(), f);
    if (bytes) std::fwrite(p, 1, bytes, f);
    std::fclose(f);
}

template <class T>
const T* As(const Npy& a, const char* want) {
    if (a.dtype != want) throw std::runtime_error(std::string("dtype ") + a.dtype + " != " + want);
    return reinterpret_cast<const T*>(a.data.data());
}

int CmdLevels(char** av) {
    const size_t W = 8192;
    std::vector<double> lev(13 * 2 * W, std::numeric_limits<double>::quiet_NaN());
    std::vector<double> sc(14, 0.0), wc(14 * 2, 0.0);
    for (int b = 1; b <= 13; ++b) {
        for (int br = 0; br < 2; ++br) {
            const auto l = e8qlib::BranchLevels(b, br);
            std::memcpy(&lev[(static_cast<size_t>(b - 1) * 2 + br) * W], l.data(), l.size() * 8);
        }
        sc[static_cast<size_t>(b)] = e8qlib::FixedScale(b);
        e8qlib::ScanClassMeans(b, &wc[static_cast<size_t>(b) * 2]);
    }
    WriteNpy(av[0], "<f8", {13, 2, W}, lev.data(), lev.size() * 8);
    WriteNpy(av[1], "<f8", {14}, sc.data(), sc.size() * 8);
    WriteNpy(av[2], "<f8", {14, 2}, wc.data(), wc.size() * 8);
    return 0;
}

int CmdRotate(char** av) {
    const size_t dim = std::strtoull(av[0], nullptr, 10);
    const uint64_t seed = std::strtoull(av[1], nullptr, 0);
    const Npy in = ReadNpy(av[2]);
    const float* x = As<float>(in, "<f4");
    e8qlib::FhtKac4Rotator rot(dim, seed);
    std::vector<char> signs(rot.dump_bytes());
    rot.Save(signs.data());
    const size_t rows = in.shape[0], P = rot.padded_dim();
    std::vector<float> out(rows * P);
    for (size_t i = 0; i < rows; ++i) rot.Rotate(x + i * dim, out.data() + i * P);
    WriteNpy(av[3], "|u1", {signs.size()}, signs.data(), signs.size());
    WriteNpy(av[4], "<f4", {rows, P}, out.data(), out.size() * 4);
    return 0;
}

int CmdEncode(char** av) {
    const int bits = std::atoi(av[0]);
    const int padded = std::atoi(av[1]);
    const Npy in = ReadNpy(av[2]);
    const float* u = As<float>(in, "<f4");
    const size_t rows = in.shape[0], blocks = static_cast<size_t>(padded) / 8;
    std::vector<uint16_t> dg(rows * padded);
    std::vector<uint8_t> br(rows * blocks);
    for (size_t i = 0; i < rows; ++i) {
        const e8qlib::VectorCode vc = e8qlib::EncodeVectorAtScale(u + i * padded, padded, bits, e8qlib::FixedScale(bits));
        for (size_t b = 0; b < blocks; ++b) {
            for (int k = 0; k < 8; ++k) dg[i * padded + b * 8 + k] = vc.digits[b][static_cast<size_t>(k)];
            br[i * blocks + b] = static_cast<uint8_t>(vc.branch[b]);
        }
    }
    WriteNpy(av[3], "<u2", {rows, static_cast<size_t>(padded)}, dg.data(), dg.size() * 2);
    WriteNpy(av[4], "|u1", {rows, blocks}, br.data(), br.size());
    return 0;
}

int CmdAccum(char** av) {
    const Npy pay = ReadNpy(av[0]);   // [2, blocks, 32] u8
    const Npy lut = ReadNpy(av[1]);   // [blocks, 64] u8
    const size_t blocks = lut.shape[0];
    const uint8_t* p = As<uint8_t>(pay, "|u1");
    const uint8_t* l = As<uint8_t>(lut, "|u1");
    std::vector<uint8_t> pair(blocks * 64);
    e8qlib::FillPairLut(l, blocks, pair.data());
    alignas(64) int32_t single[32], o1[32], o2[32];
    e8qlib::AccumulateCoarseHi(p, l, pair.data(), blocks, single);
    e8qlib::AccumulateCoarseHiPair(p, p + blocks * 32, l, pair.data(), blocks, o1, o2);
    std::vector<int32_t> po(64);
    std::memcpy(po.data(), o1, 128);
    std::memcpy(po.data() + 32, o2, 128);
    WriteNpy(av[2], "<i4", {32}, single, 128);
    WriteNpy(av[3], "<i4", {2, 32}, po.data(), 256);
    return 0;
}

int CmdRoute(char** av) {
    const Npy q = ReadNpy(av[0]);
    const Npy c = ReadNpy(av[1]);
    const size_t nprobe = std::strtoull(av[2], nullptr, 10);
    const size_t nq = q.shape[0], dim = q.shape[1], nl = c.shape[0];
    std::vector<float> cn2(nl);
    const float* cp = As<float>(c, "<f4");
    // The E8Q index stores cn2 as the fp64 sum of squares (e8q_build.cpp:264-269).
    for (size_t i = 0; i < nl; ++i) {
        double s = 0.0;
        for (size_t d = 0; d < dim; ++d) s += static_cast<double>(cp[i * dim + d]) * cp[i * dim + d];
        cn2[i] = static_cast<float>(s);
    }
    std::vector<float> dist(nq * nl);
    std::vector<uint32_t> ids(nq * nprobe);
    std::vector<uint64_t> keys(nl);
    for (size_t i = 0; i < nq; ++i)
        e8qlib::ivf::RouteL2(As<float>(q, "<f4") + i * dim, cp, cn2.data(), nl, dim, nprobe, dist.data() + i * nl,
                             ids.data() + i * nprobe, keys.data());
    WriteNpy(av[3], "<f4", {nq, nl}, dist.data(), dist.size() * 4);
    WriteNpy(av[4], "<u4", {nq, nprobe}, ids.data(), ids.size() * 4);
    WriteNpy(av[5], "<f4", {nl}, cn2.data(), cn2.size() * 4);
    return 0;
}

int CmdBuild(char** av) {
    const Npy data = ReadNpy(av[0]);
    const Npy cent = ReadNpy(av[1]);
    const Npy ids = ReadNpy(av[2]);
    const int bits = std::atoi(av[3]);
    const uint64_t seed = std::strtoull(av[4], nullptr, 0);
    e8qlib::E8qBuildOptions opt;
    opt.prune_alpha = std::strtof(av[5], nullptr);
    e8qlib::E8qIndex idx;
    idx.Build(As<float>(data, "<f4"), As<float>(cent, "<f4"), As<uint32_t>(ids, "<u4"), data.shape[0],
              data.shape[1], cent.shape[0], bits, seed, 4, opt);
    idx.Save(av[6]);
    return 0;
}

int CmdSearch(char** av) {
    e8qlib::E8qIndex idx;
    idx.Load(av[0]);
    const Npy q = ReadNpy(av[1]);
    const size_t k = std::strtoull(av[2], nullptr, 10);
    const size_t nprobe = std::strtoull(av[3], nullptr, 10);
    idx.SetDisablePrune(std::atoi(av[4]) != 0);
    const std::string pre = av[5];
    const size_t nq = q.shape[0], dim = q.shape[1], P = idx.padded_dim();
    const size_t take = std::min(nprobe, idx.num_clusters());
    std::vector<uint32_t> ids(nq * k);
    std::vector<float> dists(nq * k), rotq(nq * P), consts(nq * 5), pdist(nq * idx.num_clusters());
    std::vector<uint8_t> lut(nq * (P / 8) * 64);
    std::vector<uint32_t> probe(nq * take);
    std::vector<uint64_t> st(nq * 2);
    for (size_t i = 0; i < nq; ++i) {
        e8qlib::QueryScratch s;
        e8qlib::SearchStats stats;
        idx.Search(As<float>(q, "<f4") + i * dim, k, nprobe, ids.data() + i * k, s, &stats, dists.data() + i * k);
        std::memcpy(rotq.data() + i * P, s.rot_q.data(), P * 4);
        std::memcpy(lut.data() + i * (P / 8) * 64, s.lut_c.data(), (P / 8) * 64);
        consts[i * 5 + 0] = s.prep_delta_c;
        consts[i * 5 + 1] = s.prep_base_c;
        consts[i * 5 + 2] = s.ortho_lut_err;
        consts[i * 5 + 3] = s.ortho_score_max;
        consts[i * 5 + 4] = s.ortho_z;
        std::memcpy(probe.data() + i * take, s.probe_id.data(), take * 4);
        std::memcpy(pdist.data() + i * idx.num_clusters(), s.probe_dist.data(), idx.num_clusters() * 4);
        st[i * 2] = stats.candidates;
        st[i * 2 + 1] = stats.survivors;
    }
    WriteNpy(pre + "_ids.npy", "<u4", {nq, k}, ids.data(), ids.size() * 4);
    WriteNpy(pre + "_dists.npy", "<f4", {nq, k}, dists.data(), dists.size() * 4);
    WriteNpy(pre + "_stats.npy", "<u8", {nq, 2}, st.data(), st.size() * 8);
    WriteNpy(pre + "_rotq.npy", "<f4", {nq, P}, rotq.data(), rotq.size() * 4);
    WriteNpy(pre + "_lut.npy", "|u1", {nq, (P / 8) * 64}, lut.data(), lut.size());
    WriteNpy(pre + "_consts.npy", "<f4", {nq, 5}, consts.data(), consts.size() * 4);
    WriteNpy(pre + "_probe.npy", "<u4", {nq, take}, probe.data(), probe.size() * 4);
    WriteNpy(pre + "_pdist.npy", "<f4", {nq, idx.num_clusters()}, pdist.data(), pdist.size() * 4);
    return 0;
}

int CmdMBuild(char** av) {
    const Npy data = ReadNpy(av[0]);
    const Npy cent = ReadNpy(av[1]);
    const Npy ids = ReadNpy(av[2]);
    e8qlib::me8q::BuildOptions opt;
    opt.avg_bits = std::strtof(av[3], nullptr);
    const uint64_t seed = std::strtoull(av[4], nullptr, 0);
    if (seed != 0) opt.rot_seed = seed;
    opt.prune_alpha = std::strtof(av[5], nullptr);
    if (std::string(av[6]) != "-") opt.plan_spec = av[6];
    Npy var;
    const float* varp = nullptr;
    if (std::string(av[7]) != "-") {
        var = ReadNpy(av[7]);
        varp = As<float>(var, "<f4");
    }
    opt.drop_energy = std::atoi(av[8]) != 0;
    opt.threads = 4;
    const size_t n = data.shape[0], dim = data.shape[1], nl = cent.shape[0];
    e8qlib::me8q::Me8qIndex idx;
    idx.Build(As<float>(data, "<f4"), dim, As<float>(cent, "<f4"), dim, As<uint32_t>(ids, "<u4"), n, dim, nl, opt,
              varp);
    idx.Save(av[9]);
    // Fast layout, clusters concatenated.
    const std::string pre = av[10];
    std::vector<uint8_t> plane, refine;
    std::vector<float> gain, vpc;
    std::vector<double> rcn2;
    std::vector<uint64_t> off(nl * 4 + 4, 0);
    for (size_t c = 0; c < nl; ++c) {
        const auto& f = idx.fast_[c];
        off[c * 4 + 0] = plane.size();
        off[c * 4 + 1] = refine.size();
        off[c * 4 + 2] = gain.size();
        off[c * 4 + 3] = rcn2.size();
        plane.insert(plane.end(), f.plane.begin(), f.plane.end());
        refine.insert(refine.end(), f.refine.begin(), f.refine.end());
        gain.insert(gain.end(), f.ortho_gain.begin(), f.ortho_gain.end());
        vpc.insert(vpc.end(), f.ortho_vpc.begin(), f.ortho_vpc.end());
        rcn2.insert(rcn2.end(), f.rc_n2.begin(), f.rc_n2.end());
    }
    off[nl * 4 + 0] = plane.size();
    off[nl * 4 + 1] = refine.size();
    off[nl * 4 + 2] = gain.size();
    off[nl * 4 + 3] = rcn2.size();
    std::vector<float> cn2(idx.centroid_n2_.begin(), idx.centroid_n2_.end());
    std::vector<float> crot(idx.cent_rot_.begin(), idx.cent_rot_.end());
    WriteNpy(pre + "_plane.npy", "|u1", {plane.size()}, plane.data(), plane.size());
    WriteNpy(pre + "_refine.npy", "|u1", {refine.size()}, refine.data(), refine.size());
    WriteNpy(pre + "_gain.npy", "<f4", {gain.size()}, gain.data(), gain.size() * 4);
    WriteNpy(pre + "_vpc.npy", "<f4", {vpc.size()}, vpc.data(), vpc.size() * 4);
    WriteNpy(pre + "_rcn2.npy", "<f8", {rcn2.size()}, rcn2.data(), rcn2.size() * 8);
    WriteNpy(pre + "_off.npy", "<u8", {nl + 1, 4}, off.data(), off.size() * 8);
    WriteNpy(pre + "_cn2.npy", "<f4", {cn2.size()}, cn2.data(), cn2.size() * 4);
    WriteNpy(pre + "_crot.npy", "<f4", {nl, idx.padded_dim()}, crot.data(), crot.size() * 4);
    std::vector<int64_t> plan;
    for (const auto& sg : idx.shared_plan()) {
        plan.push_back(static_cast<int64_t>(sg.len));
        plan.push_back(sg.bits);
    }
    WriteNpy(pre + "_plan.npy", "<i8", {plan.size() / 2, 2}, plan.data(), plan.size() * 8);
    const uint64_t rd = idx.route_dim();
    WriteNpy(pre + "_routedim.npy", "<u8", {1}, &rd, 8);
    return 0;
}

int CmdMSearch(char** av) {
    e8qlib::me8q::Me8qIndex idx;
    idx.Load(av[0]);
    const Npy q = ReadNpy(av[1]);
    const size_t k = std::strtoull(av[2], nullptr, 10);
    const size_t nprobe = std::strtoull(av[3], nullptr, 10);
    const std::string pre = av[4];
    const size_t nq = q.shape[0], dim = q.shape[1], P = idx.padded_dim();
    const size_t take = std::min(nprobe, idx.num_clusters());
    std::vector<uint32_t> ids(nq * k), probe(nq * take);
    std::vector<float> dists(nq * k), rotq(nq * P), pdist(nq * take);
    std::vector<uint64_t> st(nq * 5);
    for (size_t i = 0; i < nq; ++i) {
        e8qlib::me8q::SearchStats stats;
        idx.SearchIVFFast(As<float>(q, "<f4") + i * dim, k, nprobe, ids.data() + i * k, &stats,
                          dists.data() + i * k);
        std::vector<float> qpad;
        idx.PadQuery(As<float>(q, "<f4") + i * dim, &qpad);
        std::vector<float> rq(P, 0.0F);
        idx.PrepareRotatedQuery(qpad.data(), rq.data());
        std::memcpy(rotq.data() + i * P, rq.data(), P * 4);
        std::vector<std::pair<float, uint32_t>> cd;
        idx.RouteQuery(rq.data(), nprobe, &cd);
        for (size_t p = 0; p < take; ++p) {
            probe[i * take + p] = cd[p].second;
            pdist[i * take + p] = cd[p].first;
        }
        st[i * 5 + 0] = stats.candidates;
        st[i * 5 + 1] = stats.survivors;
        st[i * 5 + 2] = stats.bits_scored;
        st[i * 5 + 3] = stats.bits_total;
        st[i * 5 + 4] = stats.pruned;
    }
    // The batched entry point must give the same results.
    std::vector<uint32_t> bids(nq * k);
    std::vector<float> bd(nq * k);
    idx.SearchIVFFastBatch(As<float>(q, "<f4"), nq, k, nprobe, bids.data(), nullptr, bd.data(), 3);
    if (bids != ids || std::memcmp(bd.data(), dists.data(), bd.size() * 4) != 0)
        throw std::runtime_error("SearchIVFFastBatch differs from SearchIVFFast");
    WriteNpy(pre + "_ids.npy", "<u4", {nq, k}, ids.data(), ids.size() * 4);
    WriteNpy(pre + "_dists.npy", "<f4", {nq, k}, dists.data(), dists.size() * 4);
    WriteNpy(pre + "_stats.npy", "<u8", {nq, 5}, st.data(), st.size() * 8);
    WriteNpy(pre + "_rotq.npy", "<f4", {nq, P}, rotq.data(), rotq.size() * 4);
    WriteNpy(pre + "_probe.npy", "<u4", {nq, take}, probe.data(), probe.size() * 4);
    WriteNpy(pre + "_pdist.npy", "<f4", {nq, take}, pdist.data(), pdist.size() * 4);
    return 0;
}

// E8Q diagnostics: Estimate and ProbeBoundW for (query, id) pairs (the base
// row of id is row id of BASE), ExhaustiveTopK of every query at k.
int CmdDiag(char** av) {
    e8qlib::E8qIndex idx;
    idx.Load(av[0]);
    const Npy q = ReadNpy(av[1]);
    const Npy base = ReadNpy(av[2]);
    const Npy pq = ReadNpy(av[3]);
    const Npy pid = ReadNpy(av[4]);
    const size_t k = std::strtoull(av[5], nullptr, 10);
    const std::string pre = av[6];
    const size_t nq = q.shape[0], dim = q.shape[1], np_ = pq.shape[0];
    const float* qp = As<float>(q, "<f4");
    const float* bp = As<float>(base, "<f4");
    const uint32_t* pqp = As<uint32_t>(pq, "<u4");
    const uint32_t* pip = As<uint32_t>(pid, "<u4");
    std::vector<float> est(np_);
    std::vector<double> pbw(np_);
    for (size_t i = 0; i < np_; ++i) {
        est[i] = idx.Estimate(qp + pqp[i] * dim, pip[i]);
        pbw[i] = idx.ProbeBoundW(qp + pqp[i] * dim, bp + static_cast<size_t>(pip[i]) * dim, pip[i]);
    }
    std::vector<uint32_t> ids(nq * k, std::numeric_limits<uint32_t>::max());
    std::vector<float> d(nq * k, std::numeric_limits<float>::infinity());
    for (size_t i = 0; i < nq; ++i) {
        std::vector<std::pair<float, uint32_t>> out;
        idx.ExhaustiveTopK(qp + i * dim, k, out);
        for (size_t j = 0; j < out.size(); ++j) {
            d[i * k + j] = out[j].first;
            ids[i * k + j] = out[j].second;
        }
    }
    WriteNpy(pre + "_est.npy", "<f4", {np_}, est.data(), est.size() * 4);
    WriteNpy(pre + "_pbw.npy", "<f8", {np_}, pbw.data(), pbw.size() * 8);
    WriteNpy(pre + "_exids.npy", "<u4", {nq, k}, ids.data(), ids.size() * 4);
    WriteNpy(pre + "_exd.npy", "<f4", {nq, k}, d.data(), d.size() * 4);
    return 0;
}

// ME8Q diagnostics: Estimate for (query, id) pairs, ExhaustiveTopK at k and
// SearchIVF at (k, nprobe) of every query; the SearchIVF stats per query are
// candidates, clusters, bits_scored, bits_total, survivors, pruned.
int CmdMDiag(char** av) {
    e8qlib::me8q::Me8qIndex idx;
    idx.Load(av[0]);
    const Npy q = ReadNpy(av[1]);
    const Npy pq = ReadNpy(av[2]);
    const Npy pid = ReadNpy(av[3]);
    const size_t k = std::strtoull(av[4], nullptr, 10);
    const size_t nprobe = std::strtoull(av[5], nullptr, 10);
    const std::string pre = av[6];
    const size_t nq = q.shape[0], dim = q.shape[1], np_ = pq.shape[0];
    const float* qp = As<float>(q, "<f4");
    const uint32_t* pqp = As<uint32_t>(pq, "<u4");
    const uint32_t* pip = As<uint32_t>(pid, "<u4");
    std::vector<float> est(np_);
    for (size_t i = 0; i < np_; ++i) est[i] = idx.Estimate(qp + pqp[i] * dim, pip[i]);
    std::vector<uint32_t> ex(nq * k, std::numeric_limits<uint32_t>::max()), iv(nq * k);
    std::vector<float> exd(nq * k, std::numeric_limits<float>::infinity()), ivd(nq * k);
    std::vector<uint64_t> st(nq * 6);
    for (size_t i = 0; i < nq; ++i) {
        std::vector<std::pair<float, uint32_t>> out;
        idx.ExhaustiveTopK(qp + i * dim, k, &out);
        for (size_t j = 0; j < out.size(); ++j) {
            exd[i * k + j] = out[j].first;
            ex[i * k + j] = out[j].second;
        }
        e8qlib::me8q::SearchStats s;
        idx.SearchIVF(qp + i * dim, k, nprobe, iv.data() + i * k, &s, ivd.data() + i * k);
        st[i * 6 + 0] = s.candidates;
        st[i * 6 + 1] = s.clusters;
        st[i * 6 + 2] = s.bits_scored;
        st[i * 6 + 3] = s.bits_total;
        st[i * 6 + 4] = s.survivors;
        st[i * 6 + 5] = s.pruned;
    }
    WriteNpy(pre + "_est.npy", "<f4", {np_}, est.data(), est.size() * 4);
    WriteNpy(pre + "_exids.npy", "<u4", {nq, k}, ex.data(), ex.size() * 4);
    WriteNpy(pre + "_exd.npy", "<f4", {nq, k}, exd.data(), exd.size() * 4);
    WriteNpy(pre + "_ivfids.npy", "<u4", {nq, k}, iv.data(), iv.size() * 4);
    WriteNpy(pre + "_ivfd.npy", "<f4", {nq, k}, ivd.data(), ivd.size() * 4);
    WriteNpy(pre + "_ivfstats.npy", "<u8", {nq, 6}, st.data(), st.size() * 8);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: gen_golden CMD ...\n");
        return 2;
    }
    const std::string cmd = argv[1];
    try {
        if (cmd == "levels" && argc == 5) return CmdLevels(argv + 2);
        if (cmd == "rotate" && argc == 7) return CmdRotate(argv + 2);
        if (cmd == "encode" && argc == 7) return CmdEncode(argv + 2);
        if (cmd == "accum" && argc == 6) return CmdAccum(argv + 2);
        if (cmd == "route" && argc == 8) return CmdRoute(argv + 2);
        if (cmd == "build" && argc == 9) return CmdBuild(argv + 2);
        if (cmd == "search" && argc == 8) return CmdSearch(argv + 2);
        if (cmd == "mbuild" && argc == 13) return CmdMBuild(argv + 2);
        if (cmd == "msearch" && argc == 7) return CmdMSearch(argv + 2);
        if (cmd == "diag" && argc == 9) return CmdDiag(argv + 2);
        if (cmd == "mdiag" && argc == 9) return CmdMDiag(argv + 2);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "gen_golden: %s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "gen_golden: bad command line\n");
    return 2;
}
