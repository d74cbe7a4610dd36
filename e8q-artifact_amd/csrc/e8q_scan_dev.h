// e8q_scan_dev.h -- device code shared by the plain and the batched search
// kernels: uniformity and address-space helpers, the refine-field decode,
// the per-query LDS state, the coarse LUT sums, and the refine + top-k merge
// of staged rows (FlushRows).
#pragma once

#include "e8q_device.h"
#include "e8q_kernels.h"

namespace e8qdev {

constexpr int kWave = 64;       // one wave per query
constexpr int kGroup = 32;      // staged rows refined and merged together

__device__ __forceinline__ bool Better(float d1, uint32_t i1, float d2, uint32_t i2) {
    return d1 < d2 || (d1 == d2 && i1 < i2);
}

// Global-memory views of pointers read from the parameter block: the
// compiler cannot infer their address space, and generic (flat) loads would
// also count against the LDS wait counter.
#define E8Q_GLOBAL __attribute__((address_space(1)))
// LDS views of generic pointers into __shared__ arrays (ds_read, not flat loads).
#define E8Q_LDS __attribute__((address_space(3)))
// Plain clang vectors (no constructors) for loads through such pointers.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// Word 3 of a raw buffer descriptor on gfx9 parts (32-bit data format, no
// swizzle, no per-lane index): byte-addressed loads range-checked against
// the descriptor's byte count.
constexpr int kBufferDword3 = 0x00020000;
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

// Fixed-size LDS state of one query.  TAB: floats of the refine level
// table held in LDS, both branches: 512 (bits <= 8, the device table's
// 256-entry branch stride), or 64 (bits <= 5: each branch's 2^bits levels,
// branch stride 2^bits).
template <int TAB>
struct PlainSharedT {
    float tab[TAB];              // refine level table
    float t1[kWave];             // exact replay: ZPR bound minus the slack term
    float pp[kWave];             //               perp (Omega) per lane
    float lbp[2], pam[2];        //               batch bound parts
    float td[kMaxK], nd[kMaxK];  // sorted top-k (dist, id) and merge target
    uint32_t ti[kMaxK], ni[kMaxK];
    float cd[kGroup];            // refined rows of the group
    uint32_t ci[kGroup];
    float ad[kGroup];            // ... those that can enter the top-k
    uint32_t ai[kGroup];
    float rg1[kGroup], rh3[kGroup], rd2[kGroup];
    uint32_t rid[kGroup];
    const uint8_t* pa[kGroup];   // rows' rowA bytes
    const uint8_t* pf[kGroup];   // rows' refine fields
    const float* pg[kGroup];     // rows' g1 (h3 follows 32 floats later)
    const uint32_t* pid[kGroup]; // rows' ids
};
using PlainShared = PlainSharedT<512>;

// Copies the refine level table into S.tab (the whole workgroup).
template <int TAB>
__device__ __forceinline__ void LoadRefineTab(PlainSharedT<TAB>& S, const IndexView& ix, int tid, int nthreads) {
    if (ix.bits > 8) return;
    if constexpr (TAB == 512) {
        for (int i = tid; i < 512; i += nthreads) S.tab[i] = G(ix.refine_tab)[i];
    } else {
        const int d = 1 << ix.bits;   // bits <= 5
        for (int i = tid; i < 2 * d; i += nthreads) S.tab[i] = G(ix.refine_tab)[(i / d) * ix.tab_stride + i % d];
    }
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

// Where a staged row lives: its cluster's scan and refine records, the
// vector offset inside the cluster, the cluster's first vector and the
// query-cluster distance.
struct RowLoc {
    const uint8_t* arena_c;
    const uint8_t* ref_c;
    uint32_t off;
    uint64_t vb;
    float d2;
};

// Rows of one cluster staged by vector offset (the plain search).
struct ClusterRows {
    const uint8_t* arena_c;
    const uint8_t* ref_c;
    uint64_t vb;
    float d2;
    const uint32_t* staged;   // LDS
    __device__ RowLoc operator()(int i) const { return RowLoc{arena_c, ref_c, staged[i], vb, d2}; }
};

// Refines rows 0..ns-1 of `rows`, kGroup at a time, and merges them into the
// sorted top-k list.  The rows' code bytes (rowA bytes and the
// (bits-1)-byte refine fields) are gathered into LDS with independent loads;
// then 8 rows x 8 components per pass decode and accumulate in the AVX2
// order (even blocks in acc, odd in acc2, then the 8-lane horizontal sum).
// Only rows that beat the current k-th entry take part in the merge.
template <bool GEN, int TAB, class Rows>
__device__ void FlushRows(PlainSharedT<TAB>& S, int ns, int& cnt, float& limit, const IndexView* ixp, int k,
                          const float* qv, const Rows& rows, uint8_t* rowbuf) {
    const IndexView& ix = *Here(ixp);
    const bool tab_lds = ix.bits <= 8;
    // branch stride of the LDS table
    const uint32_t tstride = TAB == 512 ? ix.tab_stride : (1u << ix.bits);
    const int lane = static_cast<int>(LaneHere());
    const int B = ix.blocks;
    const int c = ix.bits - 1;
    const int fbytes = ix.bits >= 2 ? B * c : 0;
    const int rstride = B + fbytes;        // multiple of 8
    const int wpr = rstride / 8, bw = B / 8;
    for (int rb = 0; rb < ns; rb += kGroup) {
        const int nn = min(kGroup, ns - rb);
        __syncthreads();
        if (lane < nn) {   // the row's addresses
            const RowLoc rl = rows(rb + lane);
            const uint64_t bo = rl.off >> 5;
            const uint32_t t = rl.off & 31;
            const uint8_t* rr = rl.ref_c + bo * ix.ref_batch_bytes;
            S.pa[lane] = rl.arena_c + bo * ix.batch_bytes + kHdr + t * B;
            S.pf[lane] = rr + t * ix.refine_stride;
            S.pg[lane] = reinterpret_cast<const float*>(rr + ix.refine_stride * 32) + t;
            S.pid[lane] = ix.ids + rl.vb + rl.off;
            S.rd2[lane] = rl.d2;
        }
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
                    const uint8_t* src = w < bw ? S.pa[r] + 8 * w : S.pf[r] + 8 * (w - bw);
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
            const E8Q_GLOBAL float* g1p = G(S.pg[lane]);
            S.rg1[lane] = g1p[0];
            S.rh3[lane] = S.rd2[lane] + g1p[32];
            S.rid[lane] = *G(S.pid[lane]);
        }
        __syncthreads();
        for (int sub = 0; sub < nn; sub += 8) {
            const int r = sub + (lane >> 3);
            const int i = lane & 7;
            if (r < nn) {
                const uint8_t* plane = rowbuf + r * rstride;
                const uint8_t* field = plane + B;
                float acc = 0.0F, acc2 = 0.0F;
                if constexpr (!GEN) {   // bits 2..5: the field fits 32 bits, the levels sit in LDS
                    // The block's field lies in the 8 bytes from its 4-byte
                    // aligned start (the field area of a row is 8-aligned);
                    // the 32 bits from its first byte hold it.  Two blocks per
                    // iteration: the even one into acc, the odd one into acc2.
                    const uint32_t cm = (1u << c) - 1u;
                    uint32_t pm = 0;                                // LSB positions of digits 0..6
                    for (int q = 0; q < 7; ++q) pm |= 1u << (c * q);
                    const uint32_t sh = static_cast<uint32_t>(c * i), rsh = static_cast<uint32_t>(7 - i);
                    // all ones in the lanes of component 7 (its digit is the parity form)
                    uint32_t lastm = i == 7 ? ~0u : 0u;
                    asm volatile("" : "+v"(lastm));
                    uint32_t bo = 0;
                    for (int blk = 0; blk < B; blk += 2) {
#pragma unroll
                        for (int u = 0; u < 2; ++u) {
                            const uint32_t* wp = reinterpret_cast<const uint32_t*>(field + (bo & ~3u));
                            const uint32_t f = __builtin_amdgcn_alignbit(wp[1], wp[0], 8 * (bo & 3u));
                            const uint32_t f7 = (f >> (7 * c)) & cm;
                            const uint32_t par = static_cast<uint32_t>(__popc(f & pm)) & 1u;
                            const uint32_t lo = (((f7 & ~1u) | par) & lastm) | (((f >> sh) & cm) & ~lastm);
                            const uint32_t rbv = plane[blk + u];
                            const uint32_t digit = lo | (((rbv >> rsh) & 1u) << c);
                            const float lv = S.tab[(f7 & 1u) * tstride + digit];
                            const float qq = qv[(blk + u) * 8 + i];
                            if (u == 0) acc = fmaf(lv, qq, acc);
                            else acc2 = fmaf(lv, qq, acc2);
                            bo += static_cast<uint32_t>(c);
                        }
                    }
                } else {
                    for (int blk = 0; blk < B; ++blk) {
                        uint32_t digit;
                        int br;
                        DecodeDigit(ix.bits, plane[blk], field + blk * c, i, &digit, &br);
                        const float lv = tab_lds ? S.tab[static_cast<uint32_t>(br) * tstride + digit]
                                                 : G(ix.refine_tab)[static_cast<uint32_t>(br) * ix.tab_stride + digit];
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

}  // namespace e8qdev
