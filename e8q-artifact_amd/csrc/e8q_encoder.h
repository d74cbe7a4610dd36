// e8q_encoder.h -- the E8 nearest-lattice-point encoder of one 8-dim block
// (e8q_codec.cpp:142-414, 549-726), shared by the E8Q and ME8Q build
// kernels.
#pragma once

#include "e8q_device.h"

namespace e8qdev {

// ---------------------------------------------------------------------------
// E8 encoder, one 8-dim block.

struct Sel {
    uint16_t d[8];
    double cost;
};

// Nearest level per (branch, parity) class from one insertion point into the
// merged sorted table (e8q_codec.cpp:387-414).  r: 0 = br0 even, 1 = br1
// even, 2 = br0 odd, 3 = br1 odd.
__device__ __forceinline__ void NearestMerged(double v, const EncodeView& ev, uint16_t dig[4], double cost[4]) {
    const double* lev = ev.mlev;
    const long n = ev.mlen;
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
    for (long r = 0; r < 4; ++r) {
        const long ib = p - 1 - ((p4 + 3 - r) & 3);
        const long ia = p + ((r + 4 - p4) & 3);
        const double lb = ib >= 0 ? lev[ib] : -__builtin_inf();
        const double la = ia < n ? lev[ia] : __builtin_inf();
        const uint16_t db = ib >= 0 ? ev.mdig[ib] : 0;
        const uint16_t da = ia < n ? ev.mdig[ia] : 0;
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

__device__ __forceinline__ Sel Assemble(const uint16_t de[8], const double ce[8], const uint16_t dO[8],
                                        const double co[8]) {
    Sel s;
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
        double best = 1.7976931348623157e308;
        for (int i = 0; i < 8; ++i) {
            const bool was_odd = s.d[i] == dO[i];
            const double extra = was_odd ? ce[i] - co[i] : co[i] - ce[i];
            if (extra < best) {
                best = extra;
                sel = i;
            }
        }
        s.d[sel] = (s.d[sel] == dO[sel]) ? de[sel] : dO[sel];
        s.cost += fmax(best, 0.0);
    }
    return s;
}

// Encodes one block of 8 unit-vector coordinates (EncodeVectorAtScaleInto,
// e8q_codec.cpp:671-726).  sr = s * sqrt(padded).
__device__ inline void EncodeBlock(const float* u8, double sr, const EncodeView& ev, uint16_t out[8], int* branch) {
    uint16_t de[2][8], dO[2][8];
    double ce[2][8], co[2][8];
    const int half = 1 << (ev.bits - 1);
    for (int i = 0; i < 8; ++i) {
        const double v = static_cast<double>(u8[i]) * sr;
        if (ev.merged) {
            uint16_t dg[4];
            double cs[4];
            NearestMerged(v, ev, dg, cs);
            de[0][i] = dg[0]; ce[0][i] = cs[0];
            de[1][i] = dg[1]; ce[1][i] = cs[1];
            dO[0][i] = dg[2]; co[0][i] = cs[2];
            dO[1][i] = dg[3]; co[1][i] = cs[3];
        } else {
            for (int br = 0; br < 2; ++br) {
                NearestClass(v, ev.cls_lev + (br * 2 + 0) * half, ev.cls_dig + (br * 2 + 0) * half, half, &de[br][i],
                             &ce[br][i]);
                NearestClass(v, ev.cls_lev + (br * 2 + 1) * half, ev.cls_dig + (br * 2 + 1) * half, half, &dO[br][i],
                             &co[br][i]);
            }
        }
    }
    const Sel e0 = Assemble(de[0], ce[0], dO[0], co[0]);
    const Sel e1 = Assemble(de[1], ce[1], dO[1], co[1]);
    const bool one = e1.cost < e0.cost;
    for (int i = 0; i < 8; ++i) out[i] = one ? e1.d[i] : e0.d[i];
    *branch = one ? 1 : 0;
}

}  // namespace e8qdev
