// e8q_encoder.h -- the E8 nearest-lattice-point encoder of one 8-dim block
// (e8q_codec.cpp:142-414, 549-726), shared by the E8Q and ME8Q build
// kernels.
#pragma once

#include "e8q_device.h"

namespace e8qdev {

// ---------------------------------------------------------------------------
// E8 encoder, one 8-dim block.

// Nearest level per (branch, parity) class from one insertion point into the
// merged sorted table (e8q_codec.cpp:387-414).  r: 0 = br0 even, 1 = br1
// even, 2 = br0 odd, 3 = br1 odd.
__device__ __forceinline__ void NearestMerged(double v, const double* lev, const uint16_t* mdig, long n,
                                              uint16_t dig[4], double cost[4]) {
    long lo = 0, len = n;   // lower_bound: first index with lev >= v
    while (len > 0) {
        const long half = len >> 1;
        if (lev[lo + half] < v) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    const long p = lo, p4 = p & 3;
#pragma unroll
    for (long r = 0; r < 4; ++r) {
        const long ib = p - 1 - ((p4 + 3 - r) & 3);
        const long ia = p + ((r + 4 - p4) & 3);
        const double lb = ib >= 0 ? lev[ib] : -__builtin_inf();
        const double la = ia < n ? lev[ia] : __builtin_inf();
        const uint16_t db = ib >= 0 ? mdig[ib] : 0;
        const uint16_t da = ia < n ? mdig[ia] : 0;
        const double cb = (v - lb) * (v - lb);
        const double ca = (la - v) * (la - v);
        const bool take_b = (cb < ca) || (cb == ca && db < da);
        dig[r] = take_b ? db : da;
        cost[r] = take_b ? cb : ca;
    }
}

// Single-class search (the reference's four-table path, used when the merged
// interleave does not hold, i.e. b == 1): e8q_codec.cpp:233-256.
__device__ __forceinline__ void NearestClass(double v, const double* lev, const uint16_t* dg, int n, uint16_t* dig,
                                             double* cost) {
    int lo = 0, len = n;
    while (len > 0) {
        const int half = len >> 1;
        if (lev[lo + half] < v) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    const int pos = lo;
    int best = pos < n ? pos : n - 1;
    if (pos > 0) {
        const double dc = fabs(v - lev[pos - 1]);
        const double dbv = fabs(v - lev[best]);
        if ((dc < dbv) || (dc == dbv && dg[pos - 1] < dg[best])) best = pos - 1;
    }
    *dig = dg[best];
    const double dv = v - lev[best];
    *cost = dv * dv;
}

// One branch of the encode, assembled coordinate by coordinate
// (AssembleBlock, e8q_codec.cpp:549-599): per coordinate the cheaper of the
// even and odd classes (odd only if strictly cheaper), costs summed in
// coordinate order, the digit parity, and the coordinate whose flip to the
// other class costs least (strict <, so the first minimum); a block of odd
// parity flips that coordinate and adds max(extra, 0).  Digits are packed
// two per word.
struct BranchAcc {
    uint32_t d[4];
    double cost, best;
    uint32_t par, sel, alt;
    __device__ __forceinline__ void init() {
        d[0] = d[1] = d[2] = d[3] = 0;
        cost = 0.0;
        best = 1.7976931348623157e308;
        par = 0;
        sel = 0;
        alt = 0;
    }
    __device__ __forceinline__ void add(int i, uint32_t de, double ce, uint32_t dodd, double co) {
        const bool odd = co < ce;
        const uint32_t dig = odd ? dodd : de;
        d[i >> 1] |= dig << (16 * (i & 1));
        cost += odd ? co : ce;
        par ^= odd ? 1u : 0u;
        const double extra = odd ? ce - co : co - ce;
        if (extra < best) {
            best = extra;
            sel = static_cast<uint32_t>(i);
            alt = odd ? de : dodd;
        }
    }
    __device__ __forceinline__ void finish() {
        if (par != 0) {
            const uint32_t sh = 16 * (sel & 1);
#pragma unroll
            for (uint32_t w = 0; w < 4; ++w) {   // branch-free, constant indexes: d stays in registers
                const uint32_t m = (sel >> 1) == w ? 0xFFFFu << sh : 0u;
                d[w] = (d[w] & ~m) | ((alt << sh) & m);
            }
            cost += fmax(best, 0.0);
        }
    }
};

// Encodes one block of 8 unit-vector coordinates (EncodeVectorAtScaleInto,
// e8q_codec.cpp:671-726): both branches assembled in one pass over the
// coordinates; branch 1 only if strictly cheaper.  sr = s * sqrt(padded).
// mlev / mdig: the merged table (ev's, or a copy of it in LDS).
__device__ inline void EncodeBlock(const float* u8, double sr, const EncodeView& ev, const double* mlev,
                                   const uint16_t* mdig, uint16_t out[8], int* branch) {
    BranchAcc a0, a1;
    a0.init();
    a1.init();
    const int half = 1 << (ev.bits - 1);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const double v = static_cast<double>(u8[i]) * sr;
        if (ev.merged) {
            uint16_t dg[4];
            double cs[4];
            NearestMerged(v, mlev, mdig, static_cast<long>(ev.mlen), dg, cs);
            a0.add(i, dg[0], cs[0], dg[2], cs[2]);
            a1.add(i, dg[1], cs[1], dg[3], cs[3]);
        } else {
            uint16_t de0, do0, de1, do1;
            double ce0, co0, ce1, co1;
            NearestClass(v, ev.cls_lev, ev.cls_dig, half, &de0, &ce0);
            NearestClass(v, ev.cls_lev + half, ev.cls_dig + half, half, &do0, &co0);
            NearestClass(v, ev.cls_lev + 2 * half, ev.cls_dig + 2 * half, half, &de1, &ce1);
            NearestClass(v, ev.cls_lev + 3 * half, ev.cls_dig + 3 * half, half, &do1, &co1);
            a0.add(i, de0, ce0, do0, co0);
            a1.add(i, de1, ce1, do1, co1);
        }
    }
    a0.finish();
    a1.finish();
    const bool one = a1.cost < a0.cost;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t w = one ? a1.d[i >> 1] : a0.d[i >> 1];
        out[i] = static_cast<uint16_t>(w >> (16 * (i & 1)));
    }
    *branch = one ? 1 : 0;
}

__device__ inline void EncodeBlock(const float* u8, double sr, const EncodeView& ev, uint16_t out[8], int* branch) {
    EncodeBlock(u8, sr, ev, ev.mlev, ev.mdig, out, branch);
}

}  // namespace e8qdev
