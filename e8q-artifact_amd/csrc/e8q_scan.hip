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

template <int NQW>
struct PlainLds {
    uint8_t lut[NQW > 0 ? NQW * 256 : 4];
    PlainShared s;
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

// This lane's byte offsets inside a step's records: its rowA bytes, its
// bank bytes and its batch header's scales (lane half j reads batch j of
// the step).  The loads add small constants as their scalar offsets.
struct LaneOffs {
    uint32_t row, hdr, hs96;
};

// The step's records: `nrec` records of `bb` bytes from `rec_p`, the first
// batch's record (wave-uniform), read through a buffer descriptor with these
// per-lane offsets, so a step spends no vector instructions on addresses.
// The second batch may lie past the cluster (then nrec is 1): such loads
// return zeros, and those lanes do not use them.
template <int NQW>
__device__ __forceinline__ void LoadRec(Rec<NQW>& r, const uint8_t* rec_p, uint32_t nrec, uint32_t bb,
                                        const LaneOffs& o) {
    const uint32_t nb = __umulhi(nrec, bb) != 0 ? 0xFFFFFFFFu : nrec * bb;   // byte count, saturated
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(UniformPtr(rec_p)), 0,
                                                                        static_cast<int>(nb), kBufferDword3);
    if constexpr (NQW > 0) {
        if constexpr (NQW % 2 == 0) {
#pragma unroll
            for (int w = 0; w < NQW; w += 2) {
                const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, o.row, 8 * w, 0);
                r.w[w] = u32x2{v.x, v.y};
                r.w[w + 1] = u32x2{v.z, v.w};
            }
        } else {
#pragma unroll
            for (int w = 0; w < NQW; ++w) r.w[w] = __builtin_amdgcn_raw_buffer_load_b64(rs, o.row, 8 * w, 0);
        }
    } else {
        r.rp = rec_p + o.row;
    }
    r.hb = __builtin_amdgcn_raw_buffer_load_b8(rs, o.hdr, 0, 0);
    r.hw = __builtin_amdgcn_raw_buffer_load_b8(rs, o.hdr, 32, 0);
    r.hp = __builtin_amdgcn_raw_buffer_load_b8(rs, o.hdr, 64, 0);
    const u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(rs, o.hs96, 0, 0);
    const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rs, o.hs96, 16, 0);
    r.s0 = f32x4{__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z), __uint_as_float(a.w)};
    r.s1 = f32x4{__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z), __uint_as_float(b.w)};
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
    __shared__ PlainLds<NQW> sh;
    PlainShared& S = sh.s;
    extern __shared__ __align__(16) uint8_t dyn[];
    uint8_t* lutc = NQW > 0 ? sh.lut : dyn;
    float* qv = reinterpret_cast<float*>(dyn + (NQW > 0 ? 0 : static_cast<size_t>(ix.blocks) * 32));  // padded
    uint8_t* rowbuf = reinterpret_cast<uint8_t*>(qv + ix.padded);                                     // kGroup rows
    // staged rows (vector offsets inside the cluster), flush_every of them
    uint32_t* stg = reinterpret_cast<uint32_t*>(rowbuf + PlainRowBufBytes(ix.blocks, ix.bits));
    const int lane = threadIdx.x;
    if (ix.bits <= 8)
        for (int i = lane; i < 512; i += kWave) S.tab[i] = G(ix.refine_tab)[i];
    const uint32_t j = lane >> 5, t = lane & 31;
    // All ones in the upper half-wave (kept in a vector register).
    uint32_t jmask = 0u - j;
    asm volatile("" : "+v"(jmask));
    const LaneOffs lo{j * Uniform(ix.batch_bytes) + kHdr + t * static_cast<uint32_t>(ix.blocks),
                      j * Uniform(ix.batch_bytes) + t, j * Uniform(ix.batch_bytes) + 96};
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
        uint32_t cp = Settle(ch, NextInChunk(ch, 0, np), ix, a, qi, np, limit, sk, lane_cand, lane_nb);
        // The open cluster.
        uint32_t ccnt = 0, cnb = 0, cb0 = 0;
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
        // Two record sets: a step scores one while the next step's records
        // load into the other (the loop body runs twice per iteration with
        // the roles swapped, so no set is ever copied).
        Rec<NQW> ra, rb;
        cp = Uniform(cp);
        ch.base = Uniform(ch.base);
        ch.active = Uniform64(ch.active);
        ch.nonempty = Uniform64(ch.nonempty);
        if (cp < np) {
            open(cp);
            horizon();
            LoadRec<NQW>(ra, arena_c, cnb, BB, lo);
        }
        // One step on rc (its records loaded), prefetching into rn; false
        // when the query's probes are done.
        auto step = [&](const Rec<NQW>& rc, Rec<NQW>& rn) -> bool {
            // The scan state is wave-uniform; saying so keeps it in scalar
            // registers and the loop on scalar branches (the compiler's
            // uniformity analysis gives up after the per-lane merge loops).
            cp = Uniform(cp);
            if (cp >= np) return false;
            cb0 = Uniform(cb0);
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
            n_surv = Uniform(n_surv);
            n_scan = Uniform(n_scan);
            // The next step; its records are loaded now when its probe is in
            // the loaded chunk (otherwise after this step, see below).  Every
            // lane loads, so that the number of loads in flight is the same
            // on every path and a step waits only for its own records: when
            // the next records come after the step this step's are loaded
            // in their place.
            uint32_t xp = cp, xb0 = cb0 + 2;
            if (xb0 >= cnb) {
                xp = NextInChunk(ch, cp + 1, np);
                xb0 = 0;
            }
            const bool x_in = xp < np && xp - ch.base < 64;
            {
                const uint8_t* xa = arena_c;
                uint32_t xnb = cnb, xb = cb0;
                if (x_in) {
                    xb = xb0;
                    if (xp != cp) {
                        const uint32_t l = xp - ch.base;
                        xa = arena + ReadLane64(ch.ao_lo, ch.ao_hi, l);
                        xnb = (ReadLane(ch.cnt, l) + 31) / 32;
                    }
                }
                LoadRec<NQW>(rn, xa + static_cast<uint64_t>(xb) * BB, xnb - xb, BB, lo);
            }

            // The record's unused header word (sc[7]) stays live until here,
            // so that no temporary takes its register while the next
            // records' loads are in flight (that would wait for them).
            asm volatile("" ::"v"(rc.s1.w));
            const uint32_t jn = min(2u, cnb - cb0);
            n_scan += jn;
            const uint32_t bj = cb0 + j;
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
            const float sa = __uint_as_float(h31 ^ ((h31 ^ ReadLane(hm, 63)) & jmask));
            if (act) lbp = fmaf(-rc.s1.z, sa, lbp);
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
                    const uint32_t off = (cb0 + jj) * 32;
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
            if (cb0 + 2 >= cnb && staged > 0) {   // end of the cluster
                __syncthreads();
                FlushRows<GEN>(S, staged, cnt, limit, &prm->ix, k, qv, cluster_ref(), rowbuf);
                Refresh(ch, limit, sk);
                staged = 0;
                horizon();
            }
            // The records of the next step are reloaded here when its probe
            // lies beyond the loaded chunk, or when a flush has just made its
            // whole cluster skippable (then the next active probe is taken).
            bool reload = !x_in;
            if (x_in && xp != cp && !((ch.active >> (xp - ch.base)) & 1u)) {
                xp = NextInChunk(ch, xp + 1, np);
                reload = true;
            }
            if (xp >= np) return false;
            if (xp - ch.base >= 64) {
                xp = Settle(ch, xp, ix, a, qi, np, limit, sk, lane_cand, lane_nb);
                if (xp >= np) return false;
            }
            if (xp != cp) {
                open(xp);
                horizon();
            }
            if (reload) LoadRec<NQW>(rn, arena_c, cnb, BB, lo);
            cp = Uniform(xp);
            cb0 = Uniform(xb0);
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
