// e8q_scan.hip -- the plain search kernel: one wave per query, the
// reference's per-query schedule reproduced bit for bit.
//
// Reference: E8qIndex::Search (e8q_search.cpp:215-293) with ScanClusterAvx2
// (e8q_search_avx2.cpp:407-598), CollectLanes (e8q_scan_common.hpp:87-113),
// FlushSurvivors (e8q_search.cpp:182-207), RefineCoreAvx2 (e8q_search_avx2.cpp:
// 82-403), InsertHeap/ExtractTopK (e8q_search.cpp:135-180, 295-318).
//
// Work per query: the probed clusters in probe order; each cluster's
// 32-vector batches two at a time (one per half-wave).  A step is: integer
// coarse sums from the LDS query LUT over this lane's rowA bytes (held in
// registers: the device scan record stores rowA vector-major, so a lane's
// bytes are one contiguous load), the threshold-free part of the ZPR bound,
// then the batch skip, lane mask and survivor staging against the live top-k
// threshold.  The next step's records are loaded into a second register set
// while a step is scored.  Staged rows are refined whenever flush_every are
// waiting and at the end of each cluster; the top-k is a sorted (dist, id)
// list in LDS whose k-th entry is the reference heap's root, the prune
// threshold.
#include <cstdlib>

#include "e8q_scan_dev.h"

namespace e8qdev {

// The query LUT first, at LDS address 0, so lookups fold the block offset
// into the instruction.
// LDS bytes of the refine's row buffer (kGroup rows of rowA bytes and
// fields) of the plain kernel, 16-byte aligned; + 16: the refine reads each
// field through an 8-byte window.
__host__ __device__ inline uint32_t PlainRowBufBytes(uint32_t blocks, int bits) {
    return (kGroup * blocks * static_cast<uint32_t>(bits) + 16 + 15) & ~15u;
}

// The specialized kernels (rates 2..5) hold only the levels they read.
template <int NQW, bool GEN>
struct PlainLds {
    uint8_t lut[NQW > 0 ? NQW * 256 : 4];
    PlainSharedT<GEN ? 512 : 64> s;
};

// Per-query inputs of the whole-cluster skip test.
struct SkipConsts {
    float U;          // upper bound of |S| over every possible code of this query
    float lut_err;
    float sml;        // score_max + lut_err
    float z;
    bool prune;
};

// Probe metadata of 64 consecutive probes, one per lane.
struct ProbeChunk {
    uint32_t base;
    uint64_t nonempty;   // lanes whose cluster has vectors
    uint64_t active;     // ... and that are not skipped as a whole
    uint32_t cnt, vb_lo, vb_hi, ao_lo, ao_hi, ro_lo, ro_hi;
    float d2, sqd2;
    float lbpc;          // cluster bound before the slack term, with max|S| = U
    float pamx;          // cluster max of Omega_max
};

// Whole-cluster skip.  The reference skips a batch when its bound
//   lb = fma(-Om, z*r_h, fma(-Wmax, max|S|, (sc0 + d2) - fma(lut_err, Wmax, (0.5*sc3)*sml)))
// reaches the threshold.  Every operation is monotone in its inputs, so the
// same expression over the cluster's extreme header values (min sc0, max
// sc3, Wmax, Om) and max|S| := U is a lower bound of every batch's lb; when
// it reaches the threshold, every batch of the cluster would be skipped and
// the cluster is passed over without reading its records.  Skips decided
// with a threshold stay valid as it decreases (lb grows as r_h shrinks).
__device__ __forceinline__ void Refresh(ProbeChunk& c, float L, const SkipConsts& k) {
    bool s = false;
    if (k.prune && L < __builtin_inff()) {
        const float zr = k.z * (L < c.d2 ? sqrtf(fmaxf(0.0F, L)) : c.sqd2);
        s = fmaf(-c.pamx, zr, c.lbpc) >= L;
    }
    c.active = c.nonempty & ~__ballot(s);
}

// lane_cand / lane_nb: this lane's share of the candidate and batch counts
// (every vector and record of every probed cluster, skipped or not).
__device__ __forceinline__ void LoadChunk(ProbeChunk& c, uint32_t base, const IndexView& ix, const PlainArgs& a,
                                          uint64_t qi, uint32_t np, float L, const SkipConsts& k,
                                          uint32_t& lane_cand, uint32_t& lane_nb) {
    const uint32_t p = base + threadIdx.x;
    c.base = base;
    c.cnt = c.vb_lo = c.vb_hi = c.ao_lo = c.ao_hi = c.ro_lo = c.ro_hi = 0;
    c.d2 = c.sqd2 = c.lbpc = c.pamx = 0.0F;
    if (p < np) {
        const uint32_t cid = G(a.probe_ids)[qi * a.maxp + p];
        const uint64_t vb = G(ix.cluster_begin)[cid];
        const uint64_t ve = G(ix.cluster_begin)[cid + 1];
        const uint64_t ao = G(ix.arena_begin)[cid];
        const uint64_t ro = G(ix.ref_arena_begin)[cid];
        c.cnt = static_cast<uint32_t>(ve - vb);
        if (a.max_vectors != 0 && c.cnt > a.max_vectors) c.cnt = a.max_vectors;
        c.vb_lo = static_cast<uint32_t>(vb);
        c.vb_hi = static_cast<uint32_t>(vb >> 32);
        c.ao_lo = static_cast<uint32_t>(ao);
        c.ao_hi = static_cast<uint32_t>(ao >> 32);
        c.ro_lo = static_cast<uint32_t>(ro);
        c.ro_hi = static_cast<uint32_t>(ro >> 32);
        c.d2 = G(a.probe_d2)[qi * a.maxp + p];
        const f32x4 cs = *reinterpret_cast<const E8Q_GLOBAL f32x4*>(G(ix.csum) + cid);
        c.sqd2 = sqrtf(fmaxf(0.0F, c.d2));
        const float pad_w_round = (0.5F * cs.y) * k.sml;
        const float pad_add = fmaf(k.lut_err, cs.z, pad_w_round);
        c.lbpc = fmaf(-cs.z, k.U, (cs.x + c.d2) - pad_add);
        c.pamx = cs.w;
        lane_cand += c.cnt;
        lane_nb += (c.cnt + 31) / 32;
    }
    c.nonempty = __ballot(c.cnt > 0);
    asm volatile("" ::"v"(c.vb_lo), "v"(c.vb_hi), "v"(c.ao_lo), "v"(c.ao_hi), "v"(c.ro_lo), "v"(c.ro_hi), "v"(c.d2),
                 "v"(c.sqd2), "v"(c.lbpc), "v"(c.pamx));
    Refresh(c, L, k);
}

// First active probe >= from inside the loaded chunk; base + 64 (or np)
// when there is none there.
__device__ __forceinline__ uint32_t NextInChunk(const ProbeChunk& c, uint32_t from, uint32_t np) {
    const uint32_t off = from - c.base;
    const uint64_t m = off >= 64 ? 0 : (c.active >> off) << off;
    const uint32_t r = m ? c.base + static_cast<uint32_t>(__ffsll(static_cast<long long>(m))) - 1 : c.base + 64;
    return min(r, np);
}

// First active probe >= from, loading chunks as needed; np when none.
__device__ __forceinline__ uint32_t Settle(ProbeChunk& c, uint32_t from, const IndexView& ix, const PlainArgs& a,
                                           uint64_t qi, uint32_t np, float L, const SkipConsts& k,
                                           uint32_t& lane_cand, uint32_t& lane_nb) {
    uint32_t p = from;
    while (p < np && p - c.base >= 64) {
        LoadChunk(c, p, ix, a, qi, np, L, k, lane_cand, lane_nb);
        p = NextInChunk(c, p, np);
    }
    return p;
}

// This lane's share of one step: its rowA bytes (NQW 8-byte words; NQW == 0:
// any size, read in the coarse loop) and the batch header fields it uses.
template <int NQW>
struct Rec {
    u32x2 w[NQW > 0 ? NQW : 1];
    const uint8_t* rp;           // NQW == 0: this lane's rowA
    uint32_t hb, hw, hp;         // base / W / Omega bank bytes
    f32x4 s0;                    // sc[0..3]
    f32x4 s1;                    // sc[4..7]
};

// This lane's byte offsets inside its half's batch record: its rowA bytes
// and its bank bytes (the batch header's scales sit at fixed offsets).
struct LaneOffs {
    uint32_t row, hdr;
};

// The step's records: batch record `rec_p` (wave-uniform) for lanes 0..31
// and the one `dj` bytes further for lanes 32..63 (dj: this lane's half's
// displacement, zero in the lower half), `nb` bytes from rec_p in all, read
// through a buffer descriptor.  A step of one batch points the upper half
// past nb: its loads return zeros, and those lanes do not use them.
template <int NQW>
__device__ __forceinline__ void LoadRec(Rec<NQW>& r, const uint8_t* rec_p, uint32_t nb, uint32_t dj,
                                        const LaneOffs& o) {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(UniformPtr(rec_p)), 0,
                                                                        static_cast<int>(nb), kBufferDword3);
    const uint32_t row = o.row + dj, hdr = o.hdr + dj;
    if constexpr (NQW > 0) {
        if constexpr (NQW % 2 == 0) {
#pragma unroll
            for (int w = 0; w < NQW; w += 2) {
                const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, row, 8 * w, 0);
                r.w[w] = u32x2{v.x, v.y};
                r.w[w + 1] = u32x2{v.z, v.w};
            }
        } else {
#pragma unroll
            for (int w = 0; w < NQW; ++w) r.w[w] = __builtin_amdgcn_raw_buffer_load_b64(rs, row, 8 * w, 0);
        }
    } else {
        r.rp = rec_p + row;
    }
    r.hb = __builtin_amdgcn_raw_buffer_load_b8(rs, hdr, 0, 0);
    r.hw = __builtin_amdgcn_raw_buffer_load_b8(rs, hdr, 32, 0);
    r.hp = __builtin_amdgcn_raw_buffer_load_b8(rs, hdr, 64, 0);
    const u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(rs, dj, 96, 0);
    const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rs, dj, 112, 0);
    r.s0 = f32x4{__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z), __uint_as_float(a.w)};
    r.s1 = f32x4{__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z), __uint_as_float(b.w)};
}

// The batch header scales the U test reads (sc[0], sc[3..6]) of the batch
// this lane stands for in a 64-batch window.
struct WinScales {
    float sc0;
    f32x4 s36;   // sc[3..6]
};

// Loads the scales of batches w0 + lane of a cluster's records at `rec_c`
// with `nbat` batches (lanes past them read zeros through the descriptor's
// range check); `woff` is lane x batch bytes.
__device__ __forceinline__ void LoadWin(WinScales& s, const uint8_t* rec_c, uint32_t w0, uint32_t nbat, uint32_t bb,
                                        uint32_t woff) {
    const uint32_t n = min(nbat - w0, 64u);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint8_t*>(UniformPtr(rec_c + static_cast<uint64_t>(w0) * bb)), 0, static_cast<int>(n * bb),
        kBufferDword3);
    s.sc0 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, woff, 96, 0));
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, woff, 108, 0);
    s.s36 = f32x4{__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w)};
}

template <int NQW>
__device__ __forceinline__ int32_t Coarse(const Rec<NQW>& r, const uint8_t* lut, int blocks) {
    int32_t acc = 0;
    if constexpr (NQW > 0) {
#pragma unroll
        for (int w = 0; w < NQW; ++w) acc += Coarse8(r.w[w], lut + w * 256);
    } else {
        const E8Q_GLOBAL u32x2* rw = reinterpret_cast<const E8Q_GLOBAL u32x2*>(G(r.rp));
        const int nw = blocks / 8;
        for (int w = 0; w < nw; ++w) acc += Coarse8(rw[w], lut + w * 256);
    }
    return acc;
}

// One wave per query.  Each step covers two 32-vector batches (lanes 0..31
// and 32..63): coarse integer sums, the threshold-free part of the ZPR
// bound, and a speculative evaluation of both batches against the threshold
// at the step's start.  When the step's survivors fit the staging area
// without a flush, those decisions are the reference's decisions (the
// threshold cannot move before the next flush) and they are applied at
// once; otherwise the step is replayed batch by batch, lane by lane, exactly
// as CollectLanes does.
// Waves per SIMD the register budget is set for: four for short rows (the
// cold refine and merge paths take the few spills), three for long rows and
// the general decode.
template <int NQW, bool GEN>
constexpr int PlainWaves() {
    return (!GEN && NQW >= 0 && NQW <= 2) ? 4 : 3;
}

template <int NQW, bool GEN>
__global__ __launch_bounds__(kWave, (PlainWaves<NQW, GEN>())) void PlainSearchKernel(
    const PlainParams* __restrict__ prm) {
    const IndexView& ix = prm->ix;
    const PlainArgs& a = prm->a;
    __shared__ PlainLds<NQW, GEN> sh;
    auto& S = sh.s;
    extern __shared__ __align__(16) uint8_t dyn[];
    uint8_t* lutc = NQW > 0 ? sh.lut : dyn;
    float* qv = reinterpret_cast<float*>(dyn + (NQW > 0 ? 0 : static_cast<size_t>(ix.blocks) * 32));  // padded
    uint8_t* rowbuf = reinterpret_cast<uint8_t*>(qv + ix.padded);                                     // kGroup rows
    // staged rows (vector offsets inside the cluster), flush_every of them
    uint32_t* stg = reinterpret_cast<uint32_t*>(rowbuf + PlainRowBufBytes(ix.blocks, ix.bits));
    const int lane = threadIdx.x;
    LoadRefineTab(S, ix, lane, kWave);
    const uint32_t j = lane >> 5, t = lane & 31;
    // All ones in the upper half-wave (kept in a vector register).
    uint32_t jmask = 0u - j;
    asm volatile("" : "+v"(jmask));
    const LaneOffs lo{kHdr + t * static_cast<uint32_t>(ix.blocks), t};
    // byte offset of batch `lane` of a window of records
    const uint32_t woff = static_cast<uint32_t>(lane) * Uniform(ix.batch_bytes);
    const int k = static_cast<int>(a.k);
    const int fe = a.flush_every;
    const float inf = __builtin_inff();
    const uint32_t P = static_cast<uint32_t>(ix.padded);
    const uint32_t E = static_cast<uint32_t>(ix.blocks) * 32;
    // Values the scan loop uses, read once (scalar registers).
    const uint32_t BB = Uniform(ix.batch_bytes);
    const int blocks = static_cast<int>(Uniform(static_cast<uint32_t>(ix.blocks)));
    const uint8_t* const arena = UniformPtr(ix.arena);
    const float z = a.z;
    const bool prune = !a.disable_prune;
    const uint64_t nrun = a.qsel ? a.nsel : a.nq;
    for (uint64_t ii = blockIdx.x; ii < nrun; ii += gridDim.x) {
        const uint64_t qi = a.qsel ? G(a.qsel)[ii] : ii;
        for (uint32_t i = lane; i < P; i += kWave) qv[i] = G(a.rotq)[qi * P + i];
        for (uint32_t i = lane; 4 * i < E; i += kWave)
            reinterpret_cast<uint32_t*>(lutc)[i] = reinterpret_cast<const E8Q_GLOBAL uint32_t*>(G(a.lut) + qi * E)[i];
        const E8Q_GLOBAL float* cst = G(a.consts) + qi * 8;
        const float delta = UniformF(cst[0]), base = UniformF(cst[1]);
        const float lut_err = UniformF(cst[2]);
        const float sml = UniformF(cst[3]) + lut_err;   // score_max + lut_err
        const uint32_t np = min(static_cast<uint32_t>(ix.nlist), a.nprobe_q ? Uniform(G(a.nprobe_q)[qi]) : a.nprobe_all);
        __syncthreads();
        SkipConsts sk;
        sk.U = UniformF(cst[5]);
        sk.lut_err = lut_err;
        sk.sml = sml;
        sk.z = z;
        sk.prune = prune;
        int cnt = 0, staged = 0;
        float limit = inf;
        uint32_t n_surv = 0, n_scan = 0, lane_cand = 0, lane_nb = 0;
        ProbeChunk ch;
        LoadChunk(ch, 0, ix, a, qi, np, limit, sk, lane_cand, lane_nb);
        // The open cluster (the one whose batches are being scored).
        uint32_t cp = np, ccnt = 0, cnb = 0;
        float d2 = 0.0F, sqrt_d2 = 0.0F, zrh = 0.0F;
        const uint8_t* arena_c = nullptr;
        auto open = [&](uint32_t p) {
            const uint32_t l = p - ch.base;
            ccnt = ReadLane(ch.cnt, l);
            cnb = (ccnt + 31) / 32;
            d2 = __uint_as_float(ReadLane(__float_as_uint(ch.d2), l));
            sqrt_d2 = sqrtf(fmaxf(0.0F, d2));
            arena_c = arena + ReadLane64(ch.ao_lo, ch.ao_hi, l);
        };
        // z * r_horizon, r_horizon = min(sqrt(d2), sqrt(limit)) once the heap is live.
        auto horizon = [&]() { zrh = z * (limit < d2 ? sqrtf(fmaxf(0.0F, limit)) : sqrt_d2); };
        auto cluster_ref = [&]() {
            const uint32_t l = cp - ch.base;
            return ClusterRows{arena_c, ix.ref_arena + ReadLane64(ch.ro_lo, ch.ro_hi, l),
                               ReadLane64(ch.vb_lo, ch.vb_hi, l), d2, stg};
        };
        // The batch cursor walks the active probes' batches one 64-batch
        // window at a time.  A window's mask keeps the batches that the U
        // test (the step's first test, on the batch header alone) does not
        // rule out at the threshold of the moment: a batch it rules out is
        // skipped by the reference at that threshold and at every lower one,
        // so its records are never read.  Steps take the mask's batches two
        // at a time, in order; the scales of the window after the cursor's
        // are loaded a window ahead.
        uint32_t kp = np, kw0 = 0, kcnb = 0;   // the cursor's window: probe, first batch, the probe's batches
        uint64_t kwm = 0;                      // its batches not handed out yet
        uint32_t pfp = np, pfw0 = 0;           // the prefetched window (pfp == np: none)
        WinScales pf;
        auto nbat = [&](uint32_t p) { return (ReadLane(ch.cnt, p - ch.base) + 31) / 32; };
        auto rec_of = [&](uint32_t p) { return arena + ReadLane64(ch.ao_lo, ch.ao_hi, p - ch.base); };
        auto window_mask = [&](const WinScales& s, uint32_t p, uint32_t w0, uint32_t nb) -> uint64_t {
            bool need = LaneHere() < nb - w0;
            if (prune && limit < inf) {
                const float d2w = __uint_as_float(ReadLane(__float_as_uint(ch.d2), p - ch.base));
                const float zr = z * (limit < d2w ? sqrtf(fmaxf(0.0F, limit)) : sqrtf(fmaxf(0.0F, d2w)));
                const float sc3 = s.s36.x, sc4 = s.s36.y, sc5 = s.s36.z, sc6 = s.s36.w;
                const float pad_w_round = (0.5F * sc3) * sml;
                const float pad_add = fmaf(lut_err, sc6, pad_w_round);
                const float lbp = (s.sc0 + d2w) - pad_add;
                const float pam = fmaf(sc5, 255.0F, sc4);
                need = need && !(fmaf(-pam, zr, fmaf(-sc6, sk.U, lbp)) >= limit);
            }
            return __ballot(need);
        };
        // The window after (p, w0): the probe's next 64 batches, else the
        // first batches of the next active probe (np: none; past the chunk:
        // base + 64).
        auto after = [&](uint32_t p, uint32_t w0, uint32_t nb, uint32_t& ap, uint32_t& aw0) {
            if (w0 + 64 < nb) {
                ap = p;
                aw0 = w0 + 64;
            } else {
                ap = NextInChunk(ch, p + 1, np);
                aw0 = 0;
            }
        };
        // The cursor enters window (p, w0) of the loaded chunk: its mask
        // (from the prefetched scales when they are this window's), then the
        // next window's scales are loaded.
        auto enter = [&](uint32_t p, uint32_t w0) {
            kp = Uniform(p);
            kw0 = Uniform(w0);
            kcnb = Uniform(nbat(kp));
            if (kp == pfp && kw0 == pfw0) {
                kwm = window_mask(pf, kp, kw0, kcnb);
            } else {
                WinScales s;
                LoadWin(s, rec_of(kp), kw0, kcnb, BB, woff);
                kwm = window_mask(s, kp, kw0, kcnb);
            }
            kwm = Uniform64(kwm);
            uint32_t bp, bw0;
            after(kp, kw0, kcnb, bp, bw0);
            pfp = np;
            if (bp < np && bp - ch.base < 64) {
                LoadWin(pf, rec_of(bp), bw0, nbat(bp), BB, woff);
                pfp = bp;
                pfw0 = bw0;
            }
        };
        // The next step: probe xp, batches xa and xb (xjn of them).  0: a
        // step; 1: no batch is left; 2: the next window lies past the
        // loaded chunk (the cursor waits until the open cluster is done).
        auto cursor = [&](uint32_t& xp, uint32_t& xa, uint32_t& xb, uint32_t& xjn) -> int {
            while (kwm == 0) {
                if (kp >= np) return 1;
                uint32_t ap, aw0;
                after(kp, kw0, kcnb, ap, aw0);
                if (ap >= np) return 1;
                if (ap - ch.base >= 64) return 2;
                enter(ap, aw0);
            }
            xp = kp;
            xa = kw0 + static_cast<uint32_t>(__builtin_ctzll(kwm));
            kwm &= kwm - 1;
            if (kwm != 0) {
                xb = kw0 + static_cast<uint32_t>(__builtin_ctzll(kwm));
                kwm &= kwm - 1;
                xjn = 2;
            } else {
                xb = xa;
                xjn = 1;
            }
            return 0;
        };
        // The same, loading chunks when the cursor has passed the loaded one
        // (only when no cluster of it is open).
        auto cursor_settled = [&](uint32_t& xp, uint32_t& xa, uint32_t& xb, uint32_t& xjn) -> bool {
            for (;;) {
                const int r = cursor(xp, xa, xb, xjn);
                if (r == 0) return true;
                if (r == 1) return false;
                const uint32_t p = Settle(ch, ch.base + 64, ix, a, qi, np, limit, sk, lane_cand, lane_nb);
                ch.base = Uniform(ch.base);
                ch.active = Uniform64(ch.active);
                ch.nonempty = Uniform64(ch.nonempty);
                if (p >= np) {
                    kp = np;
                    kwm = 0;
                    return false;
                }
                pfp = np;
                enter(p, 0);
            }
        };
        // Records of step (xp, xa, xb, xjn) into r.
        auto load_step = [&](Rec<NQW>& r, uint32_t xp, uint32_t xa, uint32_t xb, uint32_t xjn) {
            const uint32_t span = xjn == 2 ? xb - xa : 1u;
            LoadRec<NQW>(r, rec_of(xp) + static_cast<uint64_t>(xa) * BB, (xjn == 2 ? span + 1 : 1u) * BB,
                         jmask & (span * BB), lo);
        };
        // The step being scored: probe cp, batches sa and sb (sjn of them).
        uint32_t sa = 0, sb = 0, sjn = 0;
        // Two record sets: a step scores one while the next step's records
        // load into the other (the loop body runs twice per iteration with
        // the roles swapped, so no set is ever copied).
        Rec<NQW> ra, rb;
        {
            const uint32_t p0 = Settle(ch, NextInChunk(ch, 0, np), ix, a, qi, np, limit, sk, lane_cand, lane_nb);
            ch.base = Uniform(ch.base);
            ch.active = Uniform64(ch.active);
            ch.nonempty = Uniform64(ch.nonempty);
            if (p0 < np) {
                enter(p0, 0);
                uint32_t xp = np;
                if (cursor_settled(xp, sa, sb, sjn)) {
                    cp = xp;
                    open(cp);
                    horizon();
                    load_step(ra, cp, sa, sb, sjn);
                }
            }
        }
        // One step on rc (its records loaded), prefetching into rn; false
        // when the query's probes are done.
        auto step = [&](const Rec<NQW>& rc, Rec<NQW>& rn) -> bool {
            // The scan state is wave-uniform; saying so keeps it in scalar
            // registers and the loop on scalar branches (the compiler's
            // uniformity analysis gives up after the per-lane merge loops).
            cp = Uniform(cp);
            if (cp >= np) return false;
            sa = Uniform(sa);
            sb = Uniform(sb);
            sjn = Uniform(sjn);
            cnb = Uniform(cnb);
            ccnt = Uniform(ccnt);
            staged = static_cast<int>(Uniform(static_cast<uint32_t>(staged)));
            cnt = static_cast<int>(Uniform(static_cast<uint32_t>(cnt)));
            limit = UniformF(limit);
            d2 = UniformF(d2);
            sqrt_d2 = UniformF(sqrt_d2);
            zrh = UniformF(zrh);
            arena_c = UniformPtr(arena_c);
            ch.base = Uniform(ch.base);
            ch.active = Uniform64(ch.active);
            ch.nonempty = Uniform64(ch.nonempty);
            kp = Uniform(kp);
            kw0 = Uniform(kw0);
            kcnb = Uniform(kcnb);
            kwm = Uniform64(kwm);
            pfp = Uniform(pfp);
            pfw0 = Uniform(pfw0);
            n_surv = Uniform(n_surv);
            n_scan = Uniform(n_scan);
            // The next step; its records are loaded now when the cursor has
            // it in the loaded chunk (otherwise after this step, see below).
            // Every lane loads, so that the number of loads in flight is the
            // same on every path and a step waits only for its own records:
            // when the next records come after the step this step's are
            // loaded in their place.
            uint32_t xp = np, xa = 0, xb = 0, xjn = 0;
            const bool x_in = cursor(xp, xa, xb, xjn) == 0;
            load_step(rn, x_in ? xp : cp, x_in ? xa : sa, x_in ? xb : sb, x_in ? xjn : sjn);

            // The record's unused header word (sc[7]) stays live until here,
            // so that no temporary takes its register while the next
            // records' loads are in flight (that would wait for them).
            asm volatile("" ::"v"(rc.s1.w));
            const uint32_t jn = sjn;
            n_scan += jn;
            const uint32_t bj = sa + (jmask & (sb - sa));
            const bool act = j < jn;
            const float L = limit;
            // Header terms of the ZPR bound.  Batch bound:
            // (sc0 + d2 - pad_add) - |W|max * max|S| - Omega_max * z * r_h.
            float lbp = 0.0F, pam = 0.0F, vbase = 0.0F;
            bool uskip = !act;
            if (act) {
                const float sc0 = rc.s0.x, sc3 = rc.s0.w, sc4 = rc.s1.x, sc5 = rc.s1.y, sc6 = rc.s1.z;
                const float pad_w_round = (0.5F * sc3) * sml;
                vbase = (sc0 + d2) - pad_w_round;
                const float pad_add = fmaf(lut_err, sc6, pad_w_round);
                lbp = (sc0 + d2) - pad_add;
                pam = fmaf(sc5, 255.0F, sc4);
                // No code has |S| above U and the bound only grows as max|S|
                // falls, so a batch whose bound with max|S| = U reaches the
                // threshold is one the reference skips: its coarse sums are
                // not needed.
                if (prune && L < inf) uskip = fmaf(-pam, zrh, fmaf(-sc6, sk.U, lbp)) >= L;
            }
            if (__ballot(uskip) != ~uint64_t{0}) {   // some batch of the step needs its sums
            float t1 = inf, pp = 0.0F;
            uint32_t sabs = 0;
            if (act) {
                const int32_t acc = Coarse<NQW>(rc, lutc, blocks);
                const float sc1 = rc.s0.y, sc2 = rc.s0.z, sc3 = rc.s0.w;
                const float sc4 = rc.s1.x, sc5 = rc.s1.y;
                const float scv = fmaf(static_cast<float>(acc), delta, base);
                const float bt = fmaf(static_cast<float>(rc.hb), sc1, vbase);
                const float w = fmaf(static_cast<float>(rc.hw), sc3, sc2);
                pp = fmaf(static_cast<float>(rc.hp), sc5, sc4);
                const float dc = fmaf(-w, scv, bt);
                t1 = fmaf(-fabsf(w), lut_err, dc);
                sabs = __float_as_uint(fabsf(scv));
            }
            // Batch max |S| over its 32 lanes: S is monotone in the integer
            // sum, so this is max(|S(max sum)|, |S(min sum)|), the
            // reference's batch bound input (non-negative floats order as
            // their bit patterns).
            const uint32_t hm = HalfMax(sabs);
            const uint32_t h31 = ReadLane(hm, 31);
            const float sa_ = __uint_as_float(h31 ^ ((h31 ^ ReadLane(hm, 63)) & jmask));
            if (act) lbp = fmaf(-rc.s1.z, sa_, lbp);
            bool skip = false;
            if (prune && L < inf) skip = fmaf(-pam, zrh, lbp) >= L;
            const float dlo = fmaf(-zrh, pp, t1);
            // (a batch's lanes past the cluster's vectors hold no vector)
            const bool pass = act & !skip & (bj * 32 + t < ccnt) & (!prune | (dlo < L));
            const uint64_t bal = __ballot(pass);
            const int total = __popcll(bal);
            if (staged + total <= fe) {
                // No flush inside the step: apply the speculative decisions
                // (staged in lane order: batch by batch, lane by lane).
                if (pass) {
                    const int rank = staged + static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                                                  static_cast<uint32_t>(bal >> 32),
                                                  __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(bal), 0u)));
                    stg[rank] = bj * 32 + t;
                }
                staged += total;
                n_surv += total;
            } else {
                // Exact replay, batch by batch (a flush lands inside this step).
                S.t1[lane] = t1;
                S.pp[lane] = pp;
                if ((LaneHere() & 31) == 0) {
                    S.lbp[j] = lbp;
                    S.pam[j] = pam;
                }
                __syncthreads();
                const ClusterRows cr = cluster_ref();
                for (uint32_t jj = 0; jj < jn; ++jj) {
                    const uint32_t off = (jj == 0 ? sa : sb) * 32;
                    const uint32_t nl = min(32u, ccnt - off);
                    const float limit0 = limit;
                    const float zrh0 = z * (limit0 < d2 ? sqrtf(fmaxf(0.0F, limit0)) : sqrt_d2);
                    if (prune && limit0 < inf) {
                        if (fmaf(-S.pam[jj], zrh0, S.lbp[jj]) >= limit0) continue;
                    }
                    const bool in32 = LaneHere() < 32;
                    const float dl = in32 ? fmaf(-zrh0, S.pp[jj * 32 + lane], S.t1[jj * 32 + lane]) : inf;
                    const uint32_t lmask = nl >= 32 ? 0xFFFFFFFFu : ((1u << nl) - 1u);
                    uint32_t mask = prune ? (static_cast<uint32_t>(__ballot(in32 && dl < limit0)) & lmask) : lmask;
                    while (mask != 0) {
                        const float Lc = limit;
                        const uint32_t ps = prune ? (mask & static_cast<uint32_t>(__ballot(in32 && dl < Lc))) : mask;
                        if (ps == 0) break;
                        const int room = fe - staged;
                        const int pc = __popc(ps);
                        const int tk = pc <= room ? pc : room;
                        if (in32 && ((ps >> lane) & 1u)) {
                            const int rank = __popc(ps & ((1u << lane) - 1u));
                            if (rank < tk) stg[staged + rank] = off + lane;
                        }
                        staged += tk;
                        n_surv += tk;
                        if (pc <= room) break;
                        // The next passing lane meets a full staging area: flush,
                        // then stage it without a re-test (CollectLanes' order).
                        uint32_t rest = ps;
                        for (int q = 0; q < tk; ++q) rest &= rest - 1;
                        const int pl = __ffs(rest) - 1;
                        FlushRows<GEN>(S, staged, cnt, limit, &prm->ix, k, qv, cr, rowbuf);
                        Refresh(ch, limit, sk);
                        if (LaneHere() == 0) stg[0] = off + static_cast<uint32_t>(pl);
                        staged = 1;
                        n_surv += 1;
                        mask &= (pl >= 31) ? 0u : ~((2u << pl) - 1u);
                    }
                }
                __syncthreads();
                staged = static_cast<int>(Uniform(static_cast<uint32_t>(staged)));
                n_surv = Uniform(n_surv);
                cnt = static_cast<int>(Uniform(static_cast<uint32_t>(cnt)));
                limit = UniformF(limit);
                ch.active = Uniform64(ch.active);
                horizon();
            }
            }
            // The open cluster ends with this step when the next one lies in
            // another cluster (or there is none): its staged rows are refined.
            if ((!x_in || xp != cp) && staged > 0) {
                __syncthreads();
                FlushRows<GEN>(S, staged, cnt, limit, &prm->ix, k, qv, cluster_ref(), rowbuf);
                Refresh(ch, limit, sk);
                staged = 0;
                horizon();
            }
            // The next step's records are loaded here when the cursor had to
            // wait for the next chunk, or when a flush has just made the next
            // step's whole cluster skippable (the cursor then passes over the
            // rest of that cluster).
            bool reload = !x_in;
            if (x_in && xp != cp && !((ch.active >> (xp - ch.base)) & 1u)) {
                kp = xp;
                kcnb = nbat(xp);
                kw0 = kcnb;   // no further window of this probe
                kwm = 0;
                reload = true;
            }
            if (reload) {
                if (!cursor_settled(xp, xa, xb, xjn)) {
                    cp = np;
                    return false;
                }
            }
            if (xp != cp) {
                cp = xp;
                open(cp);
                horizon();
            }
            if (reload) load_step(rn, cp, xa, xb, xjn);
            cp = Uniform(cp);
            sa = Uniform(xa);
            sb = Uniform(xb);
            sjn = Uniform(xjn);
            cnb = Uniform(cnb);
            ccnt = Uniform(ccnt);
            staged = static_cast<int>(Uniform(static_cast<uint32_t>(staged)));
            cnt = static_cast<int>(Uniform(static_cast<uint32_t>(cnt)));
            limit = UniformF(limit);
            d2 = UniformF(d2);
            sqrt_d2 = UniformF(sqrt_d2);
            zrh = UniformF(zrh);
            arena_c = UniformPtr(arena_c);
            ch.base = Uniform(ch.base);
            ch.active = Uniform64(ch.active);
            ch.nonempty = Uniform64(ch.nonempty);
            kp = Uniform(kp);
            kw0 = Uniform(kw0);
            kcnb = Uniform(kcnb);
            kwm = Uniform64(kwm);
            pfp = Uniform(pfp);
            pfw0 = Uniform(pfw0);
            n_surv = Uniform(n_surv);
            n_scan = Uniform(n_scan);
            return true;
        };
        while (step(ra, rb) && step(rb, ra)) {
        }
        for (int i = lane; i < k; i += kWave) {
            Gw(a.out_ids)[qi * a.k + i] = i < cnt ? S.ti[i] : 0xFFFFFFFFu;
            Gw(a.out_dists)[qi * a.k + i] = i < cnt ? S.td[i] : inf;
        }
        for (int o = 32; o > 0; o >>= 1) {
            lane_cand += __shfl_xor(lane_cand, o);
            lane_nb += __shfl_xor(lane_nb, o);
        }
        if (lane == 0 && a.stats != nullptr) {
            atomicAdd(&a.stats[0], static_cast<unsigned long long>(lane_cand));
            atomicAdd(&a.stats[1], static_cast<unsigned long long>(n_surv));
            atomicAdd(&a.stats[2], static_cast<unsigned long long>(lane_nb));
            atomicAdd(&a.stats[3], static_cast<unsigned long long>(n_scan));
        }
        __syncthreads();
    }
}

size_t PlainSearchLds(const IndexView& ix, int flush_every) {
    const size_t nqw = static_cast<size_t>(ix.blocks) / 8;
    const bool lut_static = (nqw == 1 || nqw == 2 || nqw == 12 || nqw == 16) && ix.bits >= 2 && ix.bits <= 5;
    return (lut_static ? 0 : static_cast<size_t>(ix.blocks) * 32) + ix.padded * sizeof(float) +
           PlainRowBufBytes(ix.blocks, ix.bits) + static_cast<size_t>(flush_every) * sizeof(uint32_t);
}

template <int NQW, bool GEN>
static hipError_t LaunchPlainT(hipStream_t st, const PlainParams* dp, uint32_t grid, size_t dyn) {
    if (dyn > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&PlainSearchKernel<NQW, GEN>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(dyn));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((PlainSearchKernel<NQW, GEN>), dim3(grid), dim3(kWave), dyn, st, dp);
    return hipGetLastError();
}

hipError_t LaunchPlainSearch(hipStream_t st, const IndexView& ix, const PlainArgs& a, PlainParams* dparams,
                             bool upload) {
    const uint64_t nrun = a.qsel ? a.nsel : a.nq;
    if (nrun == 0) return hipSuccess;
    if (a.flush_every < 1 || a.flush_every > kMaxFlush) return hipErrorInvalidValue;
    if (ix.blocks % 8 != 0) return hipErrorInvalidValue;
    if (upload) {
        const PlainParams hp{ix, a};
        const hipError_t e = hipMemcpyAsync(dparams, &hp, sizeof(hp), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return e;
    }
    const uint32_t grid = static_cast<uint32_t>(nrun < 262144 ? nrun : 262144);
    size_t dyn = PlainSearchLds(ix, a.flush_every);
    if (const char* e = std::getenv("E8Q_PLAIN_LDS_PAD")) dyn += std::strtoul(e, nullptr, 10);
    // Rates 2..5 (refine fields of 1..4 bytes per block) take the specialized
    // kernels; other rates and row sizes the general one.
    if (ix.bits < 2 || ix.bits > 5) return LaunchPlainT<0, true>(st, dparams, grid, dyn);
    switch (ix.blocks / 8) {
        case 1: return LaunchPlainT<1, false>(st, dparams, grid, dyn);
        case 2: return LaunchPlainT<2, false>(st, dparams, grid, dyn);
        case 12: return LaunchPlainT<12, false>(st, dparams, grid, dyn);
        case 16: return LaunchPlainT<16, false>(st, dparams, grid, dyn);
        default: return LaunchPlainT<0, false>(st, dparams, grid, dyn);
    }
}

// Coarse accumulate on caller-provided batches and LUTs (the integer kernel
// on its own, for the exactness tests).  rowa is the reference layout
// [blk][lane]; lut the full 64-byte-per-block reference layout.
__global__ void CoarseAccumulateKernel(const uint8_t* __restrict__ rowa, uint64_t nbatch,
                                       const uint8_t* __restrict__ lut, int blocks, int32_t* __restrict__ out) {
    const uint64_t b = blockIdx.x;
    const int t = threadIdx.x;
    if (b >= nbatch || t >= 32) return;
    const int half = (t >> 4) * 16;
    int32_t acc = 0;
    for (int blk = 0; blk < blocks; ++blk) {
        const uint32_t h = rowa[b * blocks * 32 + blk * 32 + t];
        acc += lut[blk * 64 + half + (h & 15)] + lut[blk * 64 + 32 + half + (h >> 4)];
    }
    out[b * 32 + t] = acc;
}

hipError_t LaunchCoarseAccumulate(hipStream_t st, const uint8_t* rowa, uint64_t nbatch, const uint8_t* lut, int blocks,
                                  int32_t* out) {
    if (nbatch == 0) return hipSuccess;
    hipLaunchKernelGGL(CoarseAccumulateKernel, dim3(static_cast<uint32_t>(nbatch)), dim3(64), 0, st, rowa, nbatch, lut,
                       blocks, out);
    return hipGetLastError();
}

}  // namespace e8qdev
