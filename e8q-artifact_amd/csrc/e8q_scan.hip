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
#include "e8q_device.h"
#include "e8q_kernels.h"

namespace e8qdev {

constexpr int kWave = 64;       // one wave per query
constexpr int kMaxFlush = 256;
constexpr int kGroup = 32;      // staged rows refined and merged together

__device__ __forceinline__ bool Better(float d1, uint32_t i1, float d2, uint32_t i2) {
    return d1 < d2 || (d1 == d2 && i1 < i2);
}

// Global-memory views of pointers read from the parameter block: the
// compiler cannot infer their address space, and generic (flat) loads would
// also count against the LDS wait counter.
#define E8Q_GLOBAL __attribute__((address_space(1)))
// Plain clang vectors (no constructors) for loads through such pointers.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <class T>
__device__ __forceinline__ const E8Q_GLOBAL T* G(const T* p) {
    return (const E8Q_GLOBAL T*)(p);
}
template <class T>
__device__ __forceinline__ E8Q_GLOBAL T* Gw(T* p) {
    return (E8Q_GLOBAL T*)(p);
}

// Values made opaque to the optimizer where they are used, so that it does
// not hoist loop-invariant lane masks and parameter loads of the rarely taken
// paths to the kernel entry, where they would hold scalar registers for the
// whole kernel.
__device__ __forceinline__ uint32_t LaneHere() {
    uint32_t v = threadIdx.x;
    asm volatile("" : "+v"(v));
    return v;
}
template <class T>
__device__ __forceinline__ const T* Here(const T* p) {
    asm volatile("" : "+s"(p));
    return p;
}

__device__ __forceinline__ uint32_t Uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float UniformF(float v) {
    return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v)));
}
__device__ __forceinline__ uint64_t Uniform64(uint64_t v) {
    return (static_cast<uint64_t>(Uniform(static_cast<uint32_t>(v >> 32))) << 32) | Uniform(static_cast<uint32_t>(v));
}
template <class T>
__device__ __forceinline__ const T* UniformPtr(const T* p) {
    return reinterpret_cast<const T*>(Uniform64(reinterpret_cast<uint64_t>(p)));
}
__device__ __forceinline__ uint32_t ReadLane(uint32_t v, uint32_t l) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(l)));
}
__device__ __forceinline__ uint64_t ReadLane64(uint32_t lo, uint32_t hi, uint32_t l) {
    return (static_cast<uint64_t>(ReadLane(hi, l)) << 32) | ReadLane(lo, l);
}

template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ uint32_t Dpp(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), CTRL, ROWMASK, 0xF, true));
}

// Unsigned max over each 32-lane half of the wave; valid in lanes 31 and 63
// (the other lanes of rows 0 and 2 end with unspecified values).
__device__ __forceinline__ uint32_t HalfMax(uint32_t v) {
    v = max(v, Dpp<0xB1>(v));        // quad_perm [1,0,3,2]
    v = max(v, Dpp<0x4E>(v));        // quad_perm [2,3,0,1]
    v = max(v, Dpp<0x124>(v));       // row_ror:4
    v = max(v, Dpp<0x128>(v));       // row_ror:8
    v = max(v, Dpp<0x142, 0xA>(v));  // row_bcast:15 into rows 1 and 3
    return v;
}

// Digit `i` (0..7) and the branch of one (lane, block) from the rowA byte and
// the (bits-1)-bit refine field (e8q_scan.cpp:34-90).
__device__ __forceinline__ uint32_t FieldBits(uint64_t lo, uint64_t hi, int pos, int n) {
    const uint64_t m = (uint64_t{1} << n) - 1;
    if (pos + n <= 64) return static_cast<uint32_t>((lo >> pos) & m);
    if (pos >= 64) return static_cast<uint32_t>((hi >> (pos - 64)) & m);
    return static_cast<uint32_t>(((lo >> pos) | (hi << (64 - pos))) & m);
}

__device__ __forceinline__ void DecodeDigit(int bits, uint8_t rb, const uint8_t* field, int i, uint32_t* digit,
                                            int* branch) {
    if (bits == 1) {
        const uint32_t par = __popc(static_cast<uint32_t>(rb & 0xFE)) & 1u;
        *digit = i < 7 ? ((rb >> (7 - i)) & 1u) : par;
        *branch = rb & 1;
        return;
    }
    const int c = bits - 1;
    uint64_t lo = 0, hi = 0;
    for (int b = 0; b < c; ++b) {
        const uint64_t v = field[b];
        if (b < 8) lo |= v << (8 * b);
        else hi |= v << (8 * (b - 8));
    }
    const uint32_t f7 = FieldBits(lo, hi, c * 7, c);
    *branch = static_cast<int>(f7 & 1u);
    if (i < 7) {
        *digit = FieldBits(lo, hi, c * i, c) | (((rb >> (7 - i)) & 1u) << c);
    } else {
        uint32_t par = 0;
        for (int j = 0; j < 7; ++j) par ^= FieldBits(lo, hi, c * j, 1);
        *digit = (f7 & ~1u) | par | ((rb & 1u) << c);
    }
}

// Fixed-size LDS state of one query.
struct PlainShared {
    float tab[512];              // refine level table (bits <= 8)
    float t1[kWave];             // exact replay: ZPR bound minus the slack term
    float pp[kWave];             //               perp (Omega) per lane
    float lbp[2], pam[2];        //               batch bound parts
    float td[kMaxK], nd[kMaxK];  // sorted top-k (dist, id) and merge target
    uint32_t ti[kMaxK], ni[kMaxK];
    float cd[kGroup];            // refined rows of the group
    uint32_t ci[kGroup];
    float ad[kGroup];            // ... those that can enter the top-k
    uint32_t ai[kGroup];
    float rg1[kGroup], rh3[kGroup];
    uint32_t rid[kGroup];
    uint32_t staged[kMaxFlush];  // staged rows: vector offset inside the cluster
};

// The query LUT first, at LDS address 0, so lookups fold the block offset
// into the instruction.
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

template <int NQW>
__device__ __forceinline__ void LoadRec(Rec<NQW>& r, const uint8_t* rec_p, uint32_t t, int blocks) {
    const E8Q_GLOBAL uint8_t* rec = G(rec_p);
    const uint32_t row = kHdr + t * static_cast<uint32_t>(blocks);
    if constexpr (NQW > 0) {
        const E8Q_GLOBAL u32x2* rw = reinterpret_cast<const E8Q_GLOBAL u32x2*>(rec + row);
#pragma unroll
        for (int w = 0; w < NQW; ++w) r.w[w] = rw[w];
    } else {
        r.rp = rec_p + row;
    }
    r.hb = rec[t];
    r.hw = rec[32 + t];
    r.hp = rec[64 + t];
    r.s0 = *reinterpret_cast<const E8Q_GLOBAL f32x4*>(rec + 96);
    r.s1 = *reinterpret_cast<const E8Q_GLOBAL f32x4*>(rec + 112);
}

// Eight blocks of coarse sum: bytes of x.x are blocks 0..3, of x.y 4..7; the
// LUT holds per block 16 low-nibble then 16 high-nibble entries.
__device__ __forceinline__ int32_t Coarse8(u32x2 x, const uint8_t* l) {
    int32_t s = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b)
        s += l[b * 32 + __builtin_amdgcn_ubfe(x.x, 8 * b, 4)] + l[b * 32 + 16 + __builtin_amdgcn_ubfe(x.x, 8 * b + 4, 4)];
#pragma unroll
    for (int b = 0; b < 4; ++b)
        s += l[128 + b * 32 + __builtin_amdgcn_ubfe(x.y, 8 * b, 4)] +
             l[128 + b * 32 + 16 + __builtin_amdgcn_ubfe(x.y, 8 * b + 4, 4)];
    return s;
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

// The open cluster, for the refine.
struct ClusterRef {
    const uint8_t* arena_c;
    const uint8_t* ref_c;
    uint64_t vb;
    float d2;
};

// Refines the staged rows, kGroup at a time, and merges them into the
// sorted top-k list.  The rows' code bytes (rowA bytes and the
// (bits-1)-byte refine fields) are gathered into LDS with independent loads;
// then 8 rows x 8 components per pass decode and accumulate in the AVX2
// order (even blocks in acc, odd in acc2, then the 8-lane horizontal sum).
// Only rows that beat the current k-th entry take part in the merge.
template <bool GEN>
__device__ void FlushRows(PlainShared& S, int ns, int& cnt, float& limit, const PlainParams* prm_,
                          const float* qv, const ClusterRef& cr, uint8_t* rowbuf) {
    const PlainParams* prm = Here(prm_);
    const IndexView& ix = prm->ix;
    const int k = static_cast<int>(prm->a.k);
    const bool tab_lds = ix.bits <= 8;
    const int lane = static_cast<int>(LaneHere());
    const int B = ix.blocks;
    const int c = ix.bits - 1;
    const int fbytes = ix.bits >= 2 ? B * c : 0;
    const int rstride = B + fbytes;        // multiple of 8
    const int wpr = rstride / 8, bw = B / 8;
    for (int rb = 0; rb < ns; rb += kGroup) {
        const int nn = min(kGroup, ns - rb);
        __syncthreads();
        // Word e of the group: row e % 32, word e / 32 (rowA words, then
        // refine-field words); four loads in flight per lane.
        const int nw = kGroup * wpr;
        for (int e0 = lane; e0 < nw; e0 += 4 * kWave) {
            u32x2 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + u * kWave;
                const int r = e & (kGroup - 1), w = e >> 5;
                v[u] = u32x2{0u, 0u};
                if (e < nw && r < nn) {
                    const uint32_t off = S.staged[rb + r];
                    const uint8_t* src =
                        w < bw ? cr.arena_c + static_cast<uint64_t>(off >> 5) * ix.batch_bytes + kHdr + (off & 31) * B +
                                     8 * w
                               : cr.ref_c + static_cast<uint64_t>(off >> 5) * ix.ref_batch_bytes +
                                     (off & 31) * ix.refine_stride + 8 * (w - bw);
                    v[u] = *reinterpret_cast<const E8Q_GLOBAL u32x2*>(G(src));
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + u * kWave;
                const int r = e & (kGroup - 1), w = e >> 5;
                if (e < nw && r < nn) *reinterpret_cast<u32x2*>(rowbuf + r * rstride + 8 * w) = v[u];
            }
        }
        if (lane < nn) {
            const uint32_t off = S.staged[rb + lane];
            const E8Q_GLOBAL float* g1b = reinterpret_cast<const E8Q_GLOBAL float*>(
                G(cr.ref_c) + static_cast<uint64_t>(off >> 5) * ix.ref_batch_bytes + ix.refine_stride * 32);
            S.rg1[lane] = g1b[off & 31];
            S.rh3[lane] = cr.d2 + g1b[32 + (off & 31)];
            S.rid[lane] = G(ix.ids)[cr.vb + off];
        }
        __syncthreads();
        for (int sub = 0; sub < nn; sub += 8) {
            const int r = sub + (lane >> 3);
            const int i = lane & 7;
            if (r < nn) {
                const uint8_t* plane = rowbuf + r * rstride;
                const uint8_t* field = plane + B;
                float acc = 0.0F, acc2 = 0.0F;
                if constexpr (!GEN) {   // bits 2..5: the field fits 32 bits
                    // The block's field lies in the 8 bytes from its 4-byte
                    // aligned start (the field area of a row is 8-aligned).
                    const uint32_t fmask = c >= 4 ? ~0u : (1u << (8 * c)) - 1u;
                    const uint32_t cm = (1u << c) - 1u;
                    uint32_t pm = 0;                                // LSB positions of digits 0..6
                    for (int q = 0; q < 7; ++q) pm |= 1u << (c * q);
                    for (int blk = 0; blk < B; ++blk) {
                        const uint32_t bo = static_cast<uint32_t>(blk * c);
                        const uint32_t* wp = reinterpret_cast<const uint32_t*>(field + (bo & ~3u));
                        const uint64_t win = (static_cast<uint64_t>(wp[1]) << 32) | wp[0];
                        const uint32_t f = static_cast<uint32_t>(win >> (8 * (bo & 3u))) & fmask;
                        const uint32_t rbv = plane[blk];
                        const uint32_t f7 = (f >> (7 * c)) & cm;
                        const uint32_t br = f7 & 1u;
                        const uint32_t lo = i < 7 ? ((f >> (c * i)) & cm) : ((f7 & ~1u) | (__popc(f & pm) & 1u));
                        const uint32_t digit = lo | (((rbv >> (7 - i)) & 1u) << c);
                        const uint32_t ti = br * ix.tab_stride + digit;
                        const float lv = tab_lds ? S.tab[ti] : G(ix.refine_tab)[ti];
                        if (blk & 1) acc2 = fmaf(lv, qv[blk * 8 + i], acc2);
                        else acc = fmaf(lv, qv[blk * 8 + i], acc);
                    }
                } else {
                    for (int blk = 0; blk < B; ++blk) {
                        uint32_t digit;
                        int br;
                        DecodeDigit(ix.bits, plane[blk], field + blk * c, i, &digit, &br);
                        const uint32_t ti = static_cast<uint32_t>(br) * ix.tab_stride + digit;
                        const float lv = tab_lds ? S.tab[ti] : G(ix.refine_tab)[ti];
                        if (blk & 1) acc2 = fmaf(lv, qv[blk * 8 + i], acc2);
                        else acc = fmaf(lv, qv[blk * 8 + i], acc);
                    }
                }
                const float sc = Hsum8Lanes(acc + acc2);
                if (i == 0) {
                    S.cd[r] = fmaf(-(2.0F * S.rg1[r]), sc, S.rh3[r]);
                    S.ci[r] = S.rid[r];
                }
            }
        }
        __syncthreads();
        // Rows that cannot enter the top-k (not better than a full list's
        // k-th entry) are dropped; the rest are ranked against the list and
        // each other ((dist, id) is a strict total order: ids are unique).
        bool take = false;
        float d = 0.0F;
        uint32_t id = 0;
        if (lane < nn) {
            d = S.cd[lane];
            id = S.ci[lane];
            take = cnt < k || Better(d, id, S.td[k - 1], S.ti[k - 1]);
        }
        const uint64_t am = __ballot(take);
        const int m = __popcll(am);
        if (m > 0) {
            if (take) {
                const int pos = __popcll(am & ((uint64_t{1} << lane) - 1));
                S.ad[pos] = d;
                S.ai[pos] = id;
            }
            __syncthreads();
            for (int e = lane; e < cnt; e += kWave) {
                const float ed = S.td[e];
                const uint32_t ei = S.ti[e];
                int rank = e;
                for (int q = 0; q < m; ++q) rank += Better(S.ad[q], S.ai[q], ed, ei) ? 1 : 0;
                if (rank < k) {
                    S.nd[rank] = ed;
                    S.ni[rank] = ei;
                }
            }
            if (lane < m) {
                const float nd0 = S.ad[lane];
                const uint32_t ni0 = S.ai[lane];
                int lo = 0, hi = cnt;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (Better(S.td[mid], S.ti[mid], nd0, ni0)) lo = mid + 1;
                    else hi = mid;
                }
                int rank = lo;
                for (int q = 0; q < m; ++q) rank += Better(S.ad[q], S.ai[q], nd0, ni0) ? 1 : 0;
                if (rank < k) {
                    S.nd[rank] = nd0;
                    S.ni[rank] = ni0;
                }
            }
            __syncthreads();
            cnt = min(k, cnt + m);
            for (int e = lane; e < cnt; e += kWave) {
                S.td[e] = S.nd[e];
                S.ti[e] = S.ni[e];
            }
            __syncthreads();
            limit = cnt == k ? UniformF(S.td[k - 1]) : __builtin_inff();
        }
    }
    __syncthreads();
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
    const int lane = threadIdx.x;
    if (ix.bits <= 8)
        for (int i = lane; i < 512; i += kWave) S.tab[i] = G(ix.refine_tab)[i];
    const uint32_t j = lane >> 5, t = lane & 31;
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
    for (uint64_t qi = blockIdx.x; qi < a.nq; qi += gridDim.x) {
        for (uint32_t i = lane; i < P; i += kWave) qv[i] = G(a.rotq)[qi * P + i];
        for (uint32_t i = lane; 4 * i < E; i += kWave)
            reinterpret_cast<uint32_t*>(lutc)[i] = reinterpret_cast<const E8Q_GLOBAL uint32_t*>(G(a.lut) + qi * E)[i];
        const E8Q_GLOBAL float* cst = G(a.consts) + qi * 8;
        const float delta = UniformF(cst[0]), base = UniformF(cst[1]);
        const float lut_err = UniformF(cst[2]);
        const float sml = UniformF(cst[3]) + lut_err;   // score_max + lut_err
        const uint32_t np = min(static_cast<uint32_t>(ix.nlist), a.nprobe_q ? Uniform(G(a.nprobe_q)[qi]) : a.nprobe_all);
        __syncthreads();
        // U: the integer sum of any code lies between the sums of the
        // per-table minima and maxima, and S is monotone in it.
        int32_t tmin = 0, tmax = 0;
        for (uint32_t tb = lane; tb < 2u * static_cast<uint32_t>(ix.blocks); tb += kWave) {
            const uint8_t* tp = lutc + (tb >> 1) * 32 + (tb & 1) * 16;
            int32_t mn = 255, mx = 0;
            for (int e = 0; e < 16; ++e) {
                mn = min(mn, static_cast<int32_t>(tp[e]));
                mx = max(mx, static_cast<int32_t>(tp[e]));
            }
            tmin += mn;
            tmax += mx;
        }
        for (int o = 32; o > 0; o >>= 1) {
            tmin += __shfl_xor(tmin, o);
            tmax += __shfl_xor(tmax, o);
        }
        SkipConsts sk;
        sk.U = UniformF(fmaxf(fabsf(fmaf(static_cast<float>(tmin), delta, base)),
                              fabsf(fmaf(static_cast<float>(tmax), delta, base))));
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
            return ClusterRef{arena_c, ix.ref_arena + ReadLane64(ch.ro_lo, ch.ro_hi, l), ReadLane64(ch.vb_lo, ch.vb_hi, l),
                              d2};
        };
        Rec<NQW> rn;
        cp = Uniform(cp);
        ch.base = Uniform(ch.base);
        ch.active = Uniform64(ch.active);
        ch.nonempty = Uniform64(ch.nonempty);
        if (cp < np) {
            open(cp);
            horizon();
            if (j < min(2u, cnb)) LoadRec<NQW>(rn, arena_c + static_cast<uint64_t>(j) * BB, t, blocks);
        }
        for (;;) {
            // The scan state is wave-uniform; saying so keeps it in scalar
            // registers and the loop on scalar branches (the compiler's
            // uniformity analysis gives up after the per-lane merge loops).
            cp = Uniform(cp);
            if (cp >= np) break;
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
            // the loaded chunk (otherwise after this step, see below).
            uint32_t xp = cp, xb0 = cb0 + 2;
            if (xb0 >= cnb) {
                xp = NextInChunk(ch, cp + 1, np);
                xb0 = 0;
            }
            const bool x_in = xp < np && xp - ch.base < 64;
            const Rec<NQW> rc = rn;
            if (x_in) {
                const uint8_t* xa = arena_c;
                uint32_t xnb = cnb;
                if (xp != cp) {
                    const uint32_t l = xp - ch.base;
                    xa = arena + ReadLane64(ch.ao_lo, ch.ao_hi, l);
                    xnb = (ReadLane(ch.cnt, l) + 31) / 32;
                }
                if (j < min(2u, xnb - xb0)) LoadRec<NQW>(rn, xa + static_cast<uint64_t>(xb0 + j) * BB, t, blocks);
            }

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
            const float sa = __uint_as_float(j ? ReadLane(hm, 63) : ReadLane(hm, 31));
            if (act) lbp = fmaf(-rc.s1.z, sa, lbp);
            bool skip = false;
            if (prune && L < inf) skip = fmaf(-pam, zrh, lbp) >= L;
            const uint32_t n_logic = act ? min(32u, ccnt - bj * 32) : 0u;
            const float dlo = fmaf(-zrh, pp, t1);
            const bool pass = act && !skip && t < n_logic && (!prune || dlo < L);
            const uint64_t bal = __ballot(pass);
            const uint32_t pm0 = static_cast<uint32_t>(bal), pm1 = static_cast<uint32_t>(bal >> 32);
            const int total = __popc(pm0) + __popc(pm1);
            if (staged + total <= fe) {
                // No flush inside the step: apply the speculative decisions.
                if (pass) {
                    const uint32_t m = j ? pm1 : pm0;
                    const int rank = staged + (j ? __popc(pm0) : 0) + __popc(m & ((1u << t) - 1u));
                    S.staged[rank] = bj * 32 + t;
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
                const ClusterRef cr = cluster_ref();
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
                            if (rank < tk) S.staged[staged + rank] = off + lane;
                        }
                        staged += tk;
                        n_surv += tk;
                        if (pc <= room) break;
                        // The next passing lane meets a full staging area: flush,
                        // then stage it without a re-test (CollectLanes' order).
                        uint32_t rest = ps;
                        for (int q = 0; q < tk; ++q) rest &= rest - 1;
                        const int pl = __ffs(rest) - 1;
                        FlushRows<GEN>(S, staged, cnt, limit, prm, qv, cr, rowbuf);
                        Refresh(ch, limit, sk);
                        if (LaneHere() == 0) S.staged[0] = off + static_cast<uint32_t>(pl);
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
                FlushRows<GEN>(S, staged, cnt, limit, prm, qv, cluster_ref(), rowbuf);
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
            if (xp >= np) break;
            if (xp - ch.base >= 64) {
                xp = Settle(ch, xp, ix, a, qi, np, limit, sk, lane_cand, lane_nb);
                if (xp >= np) break;
            }
            if (xp != cp) {
                open(xp);
                horizon();
            }
            if (reload && j < min(2u, cnb)) LoadRec<NQW>(rn, arena_c + static_cast<uint64_t>(j) * BB, t, blocks);
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

size_t PlainSearchLds(const IndexView& ix, int /*flush_every*/) {
    const size_t nqw = static_cast<size_t>(ix.blocks) / 8;
    const bool lut_static = (nqw == 1 || nqw == 2 || nqw == 12 || nqw == 16) && ix.bits >= 2 && ix.bits <= 5;
    const size_t rstride = static_cast<size_t>(ix.blocks) * static_cast<size_t>(ix.bits);
    // + 16: the refine reads each field through an 8-byte window.
    return (lut_static ? 0 : static_cast<size_t>(ix.blocks) * 32) + ix.padded * sizeof(float) + kGroup * rstride + 16;
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

hipError_t LaunchPlainSearch(hipStream_t st, const IndexView& ix, const PlainArgs& a, PlainParams* dparams) {
    if (a.nq == 0) return hipSuccess;
    if (a.flush_every < 1 || a.flush_every > kMaxFlush) return hipErrorInvalidValue;
    if (ix.blocks % 8 != 0) return hipErrorInvalidValue;
    const PlainParams hp{ix, a};
    const hipError_t e = hipMemcpyAsync(dparams, &hp, sizeof(hp), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    const uint32_t grid = static_cast<uint32_t>(a.nq < 262144 ? a.nq : 262144);
    const size_t dyn = PlainSearchLds(ix, a.flush_every);
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
