// me8q_device.h -- device view of an ME8Q (mixed-rate E8Q) index and the
// helpers its build and search kernels share.
//
// HBM layout (one device index):
//   codes    n rows x `stride` bytes, the reference's row format
//            (me8q_build.cpp:103-135): nrun (coef, h2) float pairs, then each
//            coded run's bit-tight code (rate bytes per 8-dim block); kept as
//            is so Save() writes it back unchanged.
//   plane    per 32-vector batch, `rec` bytes: the digit top-bit plane stored
//            vector-major ([lane][plane_cols], one byte per 8-dim block, so a
//            lane reads its own bytes 8 at a time), then nrun x 160 bytes of
//            u8 banks + scales in the reference's format (me8q_fast.cpp:
//            223-301).  The reference keeps the plane [col][lane].
//   refine   per vector, `rstride` bytes: nrun x float4 {coef, h2, gain, vpc}
//            (the reference keeps gain/vpc in separate per-list arrays), then
//            each run's (rate-1)-bit refine fields (me8q_fast.cpp:97-118).
//   rrec     (indexes of up to four coded runs: the batched form's refine)
//            per batch slot (32 per 32-vector batch, the batch tail's slots
//            zero), `rrec_bytes` (whole 128-byte lines): the refine
//            row's nrun float4 header, then the vector's W and Omega bank
//            bytes per run (padded to `rrec_shift` bytes), then the refine
//            row's fields -- everything the exact cascade reads for one
//            candidate besides its plane row, on as few lines as it fits.
//   bscale   per 32-vector batch, nrun x float4 {W_min, W_step, Omega_min,
//            Omega_step}: the bank scales of the batch record.
//   bcnt     per batch, its vector count; pids per batch slot, the id.
//   The batched form's pool entries name a vector by its batch slot
//   (global batch * 32 + lane), so its plane row, refine record, batch
//   scales and count are addressed without the list's offsets.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "e8q_device.h"

namespace e8qdev {

constexpr uint32_t kM8BankBytes = 160;   // per run per batch (me8q.hpp:183)
constexpr uint32_t kM8MaxRuns = 32;      // the reference's flush4 bound (me8q_fast.cpp:865)
constexpr uint32_t kM8MaxK = 1024;       // top-k list held in LDS
constexpr uint32_t kM8LdsLevRate = 7;    // level tables up to this rate are staged in LDS

// One coded run of the plan.
struct M8Run {
    uint32_t rate, len, nblk, dim_off;   // bits/dim, dims, blocks, first dim
    uint32_t blk0;                       // first plane column
    uint32_t code_off;                   // byte offset of its code in a code row
    uint32_t ref_off;                    // byte offset of its fields in a device refine row
    uint32_t trunc;                      // rotator FHT window
    float fac;                           // 1/sqrt(trunc)
    float pbar0, pbar1;                  // top-bit class means (ClassModelOf)
    uint32_t pad;
    const uint8_t* signs;                // rotator, 4*len/8 bytes
    const float* lev;                    // float levels [2][2^rate]
    EncodeView ev;                       // encoder tables at this rate
};

// One plan segment (coded or dropped), for the dropped-energy sum.
struct M8Seg {
    uint32_t off, len, bits, pad;
};

struct M8View {
    uint64_t n, dim, padded, nlist;
    uint32_t nrun, nseg, plane_cols, route_dim;
    uint32_t stride;      // code row bytes
    uint32_t rstride;     // device refine row bytes
    uint32_t rec;         // device plane record bytes
    int32_t drop_energy;
    const M8Run* runs;
    const M8Seg* segs;
    const uint64_t* list_begin;    // [nlist+1] vectors
    const uint64_t* batch_begin;   // [nlist+1] 32-vector batches
    uint32_t* ids;
    uint8_t* codes;
    float* centroids;     // [nlist][padded] zero padded, unrotated
    float* cent_rot;      // [nlist][padded] per-run rotated (dropped dims 0)
    float* cn2;           // [nlist] fp64 sum of squares, stored float
    double* rc_n2;        // [nlist][nrun] per-run centroid energy
    uint8_t* plane;
    uint8_t* refine;
    uint8_t* rrec;        // null when the index has more than four coded runs
    float4* bscale;
    uint8_t* bcnt;        // [nbatch] vectors in each batch (with rrec)
    uint32_t* pids;       // [nbatch * 32] the id at each batch slot, UINT32_MAX past the list (with rrec)
    uint32_t rrec_bytes, rrec_shift;
    uint32_t rstride_fields;   // bytes of a refine row's fields (after its 16 x nrun header)
    uint32_t pad_;
};

// `n` (<= 25) bits at bit `pos` of a little-endian byte string.
__device__ __forceinline__ uint32_t M8Bits(const uint8_t* p, uint32_t pos, uint32_t n) {
    const uint8_t* b = p + (pos >> 3);
    const uint32_t w = static_cast<uint32_t>(b[0]) | (static_cast<uint32_t>(b[1]) << 8) |
                       (static_cast<uint32_t>(b[2]) << 16) | (static_cast<uint32_t>(b[3]) << 24);
    return n == 0 ? 0u : (w >> (pos & 7)) & ((1u << n) - 1u);
}

// Unpacks one block packed at `bits` bits per digit (me8q_codec.cpp:186-213):
// digits 0..6 at bits*i, digit 7's high bits-1 bits at 7*bits with its LSB
// the parity of digits 0..6's LSBs, the branch at bit 8*bits-1.
__device__ __forceinline__ void M8Unpack(const uint8_t* in, uint32_t bits, uint32_t d[8], uint32_t* branch) {
    uint32_t par = 0;
    for (uint32_t i = 0; i < 7; ++i) {
        d[i] = M8Bits(in, bits * i, bits);
        par ^= d[i] & 1u;
    }
    d[7] = (M8Bits(in, bits * 7, bits - 1) << 1) | par;
    *branch = M8Bits(in, 8 * bits - 1, 1);
}

// Packs one block at `bits` bits per digit into `bits` bytes (PackBlock,
// me8q_codec.cpp:117-133).  Byte-wise, so any alignment.
__device__ __forceinline__ void M8Pack(const uint32_t d[8], uint32_t branch, uint32_t bits, uint8_t* out) {
    for (uint32_t k = 0; k < bits; ++k) {
        uint32_t byte = 0;
        for (uint32_t bit = 0; bit < 8; ++bit) {
            const uint32_t pos = 8 * k + bit;
            uint32_t v;
            if (pos == 8 * bits - 1) {
                v = branch & 1u;
            } else if (pos < 7 * bits) {
                v = (d[pos / bits] >> (pos % bits)) & 1u;
            } else {
                v = ((d[7] >> 1) >> (pos - 7 * bits)) & 1u;
            }
            byte |= v << bit;
        }
        out[k] = static_cast<uint8_t>(byte);
    }
}

__device__ __forceinline__ uint32_t M8ListOf(const uint64_t* begin, uint64_t nlist, uint64_t i) {
    uint64_t lo = 0, hi = nlist;   // last c with begin[c] <= i
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (begin[mid] <= i) lo = mid;
        else hi = mid;
    }
    return static_cast<uint32_t>(lo);
}

// One 8-lane group's half of a run score: component `ci` of the blocks of
// parity `half`, as one FMA chain in block order (the reference keeps even
// blocks in acc and odd blocks in acc2, me8q_score_avx2.cpp:404-413).  The
// digit of (block, ci) = its refine field bits (T = rate-1 per digit, digit
// 7's LSB the parity of digits 0..6's, the branch in the field's top bit)
// under the plane's top bit; at rate 1 the plane byte is the whole code.
template <int T, class BP = const uint8_t*, class LP = const float*, class QP = const float*>
__device__ __forceinline__ float M8HalfChainT(BP pb, BP fr, LP lev, QP qb, uint32_t nblk, uint32_t half, uint32_t ci) {
    constexpr uint32_t rate = T + 1;
    typedef const __attribute__((address_space(3))) uint8_t* LdsB;
    if constexpr (T >= 1 && T <= 4 && std::is_same_v<BP, LdsB>) {
        // Fields in LDS at a 4-byte aligned start, T <= 4: a block's 8T field
        // bits fit 32, so one funnel-shifted window of two LDS words holds
        // all this lane needs of a block -- its digit's T bits (component
        // ci < 7: the window starts at them), or, for component 7, the
        // block's whole field (the parity form and f7), and the branch bit.
        typedef const __attribute__((address_space(3))) uint32_t* LdsW;
        const LdsW fw = reinterpret_cast<LdsW>(fr);
        const uint32_t boff = ci < 7 ? T * ci : 0u;   // the window's start in the block's field
        const uint32_t sbr = 8 * T - 1 - boff;       // the branch bit in the window
        uint32_t pm = 0;                             // LSBs of digits 0..6
#pragma unroll
        for (uint32_t jj = 0; jj < 7; ++jj) pm |= 1u << (jj * T);
        // all ones in the lanes of component 7
        uint32_t lastm = ci == 7 ? ~0u : 0u;
        asm volatile("" : "+v"(lastm));
        const uint32_t rsh = 7 - ci;
        float acc = 0.0F;
#pragma unroll 8
        for (uint32_t blk = half; blk < nblk; blk += 2) {
            const uint32_t p = blk * 8 * T + boff;
            const uint32_t di = p >> 5;
            const uint32_t win = __builtin_amdgcn_alignbit(fw[di + 1], fw[di], p & 31u);
            const uint32_t f7 = T > 1 ? (win >> (7 * T)) & ((1u << (T - 1)) - 1u) : 0u;
            const uint32_t par = static_cast<uint32_t>(__popc(win & pm)) & 1u;
            const uint32_t lo = (((f7 << 1) | par) & lastm) | ((win & ((1u << T) - 1u)) & ~lastm);
            const uint32_t br = (win >> sbr) & 1u;
            const uint32_t top = (static_cast<uint32_t>(pb[blk]) >> rsh) & 1u;
            acc = fmaf(lev[(br << rate) + (lo | (top << T))], qb[blk * 8], acc);
        }
        return acc;
    }
    if constexpr (T >= 5 && T <= 12 && std::is_same_v<BP, LdsB>) {
        // Fields in LDS at a 4-byte aligned start, 5 <= T <= 12: a block's 8T
        // field bits fit 96, three funnel-shifted words from the block's
        // start (f0, f1, f2).  Lane ci < 7 takes its digit at T ci, lane 7
        // the T-1 bits of f7 at 7T (the same position formula) and the parity
        // of digits 0..6's LSBs; every lane the branch at 8T-1.  The words
        // past the last block's field that the window may read lie in the
        // next slot or LDS region; their bits are never used.
        typedef const __attribute__((address_space(3))) uint32_t* LdsW;
        const LdsW fw = reinterpret_cast<LdsW>(fr);
        constexpr uint32_t bpos = 8 * T - 1, bk = bpos >> 5, bsh = bpos & 31;
        uint32_t pm0 = 0, pm1 = 0, pm2 = 0;   // LSBs of digits 0..6 in f0, f1, f2
#pragma unroll
        for (uint32_t jj = 0; jj < 7; ++jj) {
            const uint32_t b = jj * T;
            if (b < 32) pm0 |= 1u << b;
            else if (b < 64) pm1 |= 1u << (b - 32);
            else pm2 |= 1u << (b - 64);
        }
        const uint32_t dpos = T * ci, dk = dpos >> 5, dsh = dpos & 31;
        const uint32_t dmask = ci < 7 ? (1u << T) - 1u : (1u << (T - 1)) - 1u;
        uint32_t lastm = ci == 7 ? ~0u : 0u;   // all ones in the lanes of component 7
        asm volatile("" : "+v"(lastm));
        const uint32_t rsh = 7 - ci;
        float acc = 0.0F;
#pragma unroll 4
        for (uint32_t blk = half; blk < nblk; blk += 2) {
            const uint32_t p = blk * 8 * T;
            const uint32_t di = p >> 5, sh = p & 31u;
            const uint32_t w0 = fw[di], w1 = fw[di + 1], w2 = fw[di + 2], w3 = fw[di + 3];
            const uint32_t f0 = __builtin_amdgcn_alignbit(w1, w0, sh);
            const uint32_t f1 = __builtin_amdgcn_alignbit(w2, w1, sh);
            const uint32_t f2 = __builtin_amdgcn_alignbit(w3, w2, sh);
            const uint32_t fa = dk == 0 ? f0 : dk == 1 ? f1 : f2;
            const uint32_t fb = dk == 0 ? f1 : dk == 1 ? f2 : 0u;
            const uint32_t d = __builtin_amdgcn_alignbit(fb, fa, dsh) & dmask;
            const uint32_t par =
                static_cast<uint32_t>(__popc(f0 & pm0) + __popc(f1 & pm1) + __popc(f2 & pm2)) & 1u;
            const uint32_t lo = (((d << 1) | par) & lastm) | (d & ~lastm);
            const uint32_t fbk = bk == 0 ? f0 : bk == 1 ? f1 : f2;
            const uint32_t br = (fbk >> bsh) & 1u;
            const uint32_t top = (static_cast<uint32_t>(pb[blk]) >> rsh) & 1u;
            acc = fmaf(lev[(br << rate) + (lo | (top << T))], qb[blk * 8], acc);
        }
        return acc;
    }
    uint64_t pml = 0;
    uint32_t pmh = 0;
#pragma unroll
    for (uint32_t jj = 0; jj < 7; ++jj) {
        if (jj * T < 64) pml |= uint64_t{1} << (jj * T);
        else pmh |= 1u << (jj * T - 64);
    }
    float acc = 0.0F;
#pragma unroll 2
    for (uint32_t blk = half; blk < nblk; blk += 2) {
        const uint32_t pbyte = pb[blk];
        const uint32_t top = (pbyte >> (7 - ci)) & 1u;
        uint32_t dgt, br;
        if constexpr (T == 0) {
            dgt = ci < 7 ? top : (__popc(pbyte >> 1) & 1u);
            br = pbyte & 1u;
        } else {
            const auto fb = fr + blk * T;
            uint64_t flo = 0;
            uint32_t fhi = 0;
#pragma unroll
            for (uint32_t b = 0; b < static_cast<uint32_t>(T); ++b) {
                const uint32_t byte = fb[b];
                if (b < 8) flo |= static_cast<uint64_t>(byte) << (8 * b);
                else fhi |= byte << (8 * (b - 8));
            }
            auto field = [&](uint32_t pos, uint32_t nb) -> uint32_t {
                const uint64_t m = (uint64_t{1} << nb) - 1;
                if (pos + nb <= 64) return static_cast<uint32_t>((flo >> pos) & m);
                if (pos >= 64) return static_cast<uint32_t>((fhi >> (pos - 64)) & m);
                return static_cast<uint32_t>(((flo >> pos) | (static_cast<uint64_t>(fhi) << (64 - pos))) & m);
            };
            uint32_t lo;
            if (ci < 7) {
                lo = field(T * ci, T);
            } else {
                const uint32_t par = (__popcll(flo & pml) + __popc(fhi & pmh)) & 1u;
                lo = (field(7 * T, T - 1) << 1) | par;
            }
            br = field(8 * T - 1, 1);
            dgt = lo | (top << T);
        }
        acc = fmaf(lev[(br << rate) + dgt], qb[blk * 8], acc);
    }
    return acc;
}

// The chain at a run-time rate.  Level tables up to kM8LdsLevRate come from
// LDS (`lev_lds`), wider ones from global memory (`lev_glob`); with every
// operand's address space known, the chain's LDS reads wait on the LDS
// counter alone and never on global loads in flight.
template <class BP, class LL, class LG, class QP>
__device__ inline float M8HalfChain(uint32_t rate, BP pb, BP fr, LL lev_lds, LG lev_glob, QP qb, uint32_t nblk,
                                    uint32_t half, uint32_t ci) {
    static_assert(kM8LdsLevRate == 7, "rates 1..7 read the LDS tables");
    switch (rate) {
        case 1: return M8HalfChainT<0>(pb, fr, lev_lds, qb, nblk, half, ci);
        case 2: return M8HalfChainT<1>(pb, fr, lev_lds, qb, nblk, half, ci);
        case 3: return M8HalfChainT<2>(pb, fr, lev_lds, qb, nblk, half, ci);
        case 4: return M8HalfChainT<3>(pb, fr, lev_lds, qb, nblk, half, ci);
        case 5: return M8HalfChainT<4>(pb, fr, lev_lds, qb, nblk, half, ci);
        case 6: return M8HalfChainT<5>(pb, fr, lev_lds, qb, nblk, half, ci);
        case 7: return M8HalfChainT<6>(pb, fr, lev_lds, qb, nblk, half, ci);
        case 8: return M8HalfChainT<7>(pb, fr, lev_glob, qb, nblk, half, ci);
        case 9: return M8HalfChainT<8>(pb, fr, lev_glob, qb, nblk, half, ci);
        case 10: return M8HalfChainT<9>(pb, fr, lev_glob, qb, nblk, half, ci);
        case 11: return M8HalfChainT<10>(pb, fr, lev_glob, qb, nblk, half, ci);
        case 12: return M8HalfChainT<11>(pb, fr, lev_glob, qb, nblk, half, ci);
        default: return M8HalfChainT<12>(pb, fr, lev_glob, qb, nblk, half, ci);
    }
}


}  // namespace e8qdev
