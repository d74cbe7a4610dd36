// diag.hip -- the diagnostics of both index kinds on the device:
// E8qIndex::Estimate / ProbeBoundW / ExhaustiveTopK (e8q_oracle.cpp:14-219)
// and Me8qIndex::Estimate / ExhaustiveTopK / SearchIVF (me8q_measure.cpp:
// 41-67, me8q_search.cpp:136-266).
//
// These score stored vectors one at a time in the reference's own order of
// operations: one lane per (query, vector) pair runs the dimension-order
// chains the reference runs (fp32 sums of separately rounded products for the
// E8Q list distance, an fp64 fused chain for the E8Q exact score at b >= 2,
// the AVX2 even/odd component chains of ScoreFast for ME8Q), so each value is
// bit-identical to the reference's.  Each query's top-k is then selected by
// one workgroup: a radix select over the candidates' keys and a bitonic sort
// in LDS, with the reference's tie rules -- (distance, id) for E8Q
// (partial_sort under Better), PushCandidate's insertion order for ME8Q.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "diag_kernels.h"
#include "e8q_device.h"
#include "me8q_device.h"

namespace e8qdev {
namespace {

constexpr uint32_t kSelThreads = 1024;

__device__ __forceinline__ float HsumAvx(const float v[8]) {
    return ((v[0] + v[4]) + (v[2] + v[6])) + ((v[1] + v[5]) + (v[3] + v[7]));
}

// Last list c in [0, nlist) with begin[c] <= pos (upper_bound(begin + 1,
// begin + nlist + 1, pos) - begin - 1: empty lists are passed over).
__device__ __forceinline__ uint64_t ListOf(const uint64_t* begin, uint64_t nlist, uint64_t pos) {
    uint64_t lo = 0, hi = nlist - 1;
    while (lo < hi) {
        const uint64_t mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= pos) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Monotone unsigned image of a float (and back).
__device__ __forceinline__ uint32_t OrdOf(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float FloatOf(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

__global__ void InverseIdsKernel(const uint32_t* __restrict__ ids, uint64_t n, uint32_t* __restrict__ inv) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) {
        const uint32_t id = ids[i];
        if (id < n) atomicMax(&inv[id], static_cast<uint32_t>(i + 1));
    }
}

// ---------------------------------------------------------------------------
// E8Q

struct E8Row {
    const uint8_t* rowa;    // the vector's rowA bytes (device layout: one per block)
    const uint8_t* field;   // its (bits-1)-byte refine fields
    float g1, h3;
};

__device__ E8Row E8RowAt(const IndexView& ix, uint64_t c, uint64_t i) {
    const uint64_t bo = i >> 5;
    const uint32_t t = static_cast<uint32_t>(i & 31);
    const uint8_t* rec = ix.arena + ix.arena_begin[c] + bo * ix.batch_bytes;
    const uint8_t* rr = ix.ref_arena + ix.ref_arena_begin[c] + bo * ix.ref_batch_bytes;
    const float* g = reinterpret_cast<const float*>(rr + static_cast<uint64_t>(ix.refine_stride) * 32);
    return E8Row{rec + kHdr + static_cast<uint64_t>(t) * ix.blocks, rr + static_cast<uint64_t>(t) * ix.refine_stride,
                 g[t], g[32 + t]};
}

// Digits and branch of one block (DecodeSplitBlockLane, e8q_scan.cpp:34-90;
// b = 1: the sign bits, digit 7 their parity, the branch in bit 0).
__device__ void E8Block(int bits, uint8_t rb, const uint8_t* field, uint32_t d[8], uint32_t* br) {
    if (bits == 1) {
        uint32_t par = 0;
        for (int i = 0; i < 7; ++i) {
            d[i] = (rb >> (7 - i)) & 1u;
            par ^= d[i];
        }
        d[7] = par;
        *br = rb & 1u;
        return;
    }
    const int c = bits - 1;
    uint64_t lo = 0, hi = 0;
    for (int b = 0; b < c; ++b) {
        if (b < 8) lo |= static_cast<uint64_t>(field[b]) << (8 * b);
        else hi |= static_cast<uint64_t>(field[b]) << (8 * (b - 8));
    }
    auto bitsat = [&](int pos, int n) -> uint32_t {
        const uint64_t m = (uint64_t{1} << n) - 1;
        if (pos + n <= 64) return static_cast<uint32_t>((lo >> pos) & m);
        if (pos >= 64) return static_cast<uint32_t>((hi >> (pos - 64)) & m);
        return static_cast<uint32_t>(((lo >> pos) | (hi << (64 - pos))) & m);
    };
    uint32_t par = 0;
    for (int i = 0; i < 7; ++i) {
        const uint32_t f = bitsat(c * i, c);
        par ^= f & 1u;
        d[i] = f | (((rb >> (7 - i)) & 1u) << c);
    }
    const uint32_t f7 = bitsat(c * 7, c);
    *br = f7 & 1u;
    d[7] = (f7 & ~1u) | par | ((rb & 1u) << c);
}

// ExactScore (e8q_search.cpp:87-126): b = 1 the AVX2 form (one FMA chain
// per component over the blocks, then the horizontal sum); b >= 2 one fp64
// fused chain in dimension order.
__device__ float E8ExactScore(const IndexView& ix, const E8Row& r, const float* q) {
    const uint32_t ts = ix.tab_stride;
    const int bits = ix.bits, c = bits - 1;
    if (bits == 1) {
        float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int blk = 0; blk < ix.blocks; ++blk) {
            uint32_t d[8], br;
            E8Block(1, r.rowa[blk], nullptr, d, &br);
            for (int i = 0; i < 8; ++i) acc[i] = fmaf(ix.refine_tab[br * ts + d[i]], q[blk * 8 + i], acc[i]);
        }
        return HsumAvx(acc);
    }
    double acc = 0.0;
    for (int blk = 0; blk < ix.blocks; ++blk) {
        uint32_t d[8], br;
        E8Block(bits, r.rowa[blk], r.field + static_cast<uint64_t>(blk) * c, d, &br);
        const float* tab = ix.refine_tab + br * ts;
        for (int i = 0; i < 8; ++i)
            acc = fma(static_cast<double>(q[blk * 8 + i]), static_cast<double>(tab[d[i]]), acc);
    }
    return static_cast<float>(acc);
}

// List distance of ExhaustiveTopK / Estimate: |q|^2 and <q, c> as sums in
// dimension order of separately rounded products (the reference's compiled
// in-order reductions; the last P % 4 dimensions fused), then
// (qn2 + cn2) - 2 dot, clamped at 0.
__device__ float E8ListD2(const IndexView& ix, const float* q, uint64_t c) {
    const float* cc = ix.centroid_rot + c * ix.padded;
    const uint64_t P = ix.padded, body = P & ~uint64_t{3};
    float qn2 = 0.0F, dot = 0.0F;
    for (uint64_t d = 0; d < body; ++d) {
        const float pq = q[d] * q[d], pc = q[d] * cc[d];
        qn2 = qn2 + pq;
        dot = dot + pc;
    }
    for (uint64_t d = body; d < P; ++d) {
        qn2 = fmaf(q[d], q[d], qn2);
        dot = fmaf(q[d], cc[d], dot);
    }
    const float d2 = (qn2 + ix.centroid_n2[c]) - (dot + dot);
    return d2 < 0.0F ? 0.0F : d2;
}

__global__ void E8EstimateKernel(IndexView ix, const float* __restrict__ rotq, const uint32_t* __restrict__ qidx,
                                 const uint32_t* __restrict__ ids, uint64_t npair, const uint32_t* __restrict__ inv,
                                 float* __restrict__ out, uint32_t* __restrict__ bad) {
    const uint64_t i = blockIdx.x * 64ull + threadIdx.x;
    if (i >= npair) return;
    const uint32_t id = ids[i];
    const uint32_t p1 = id < ix.n ? inv[id] : 0u;
    if (p1 == 0) {
        out[i] = __builtin_nanf("");
        atomicOr(bad, 1u);
        return;
    }
    const uint64_t pos = p1 - 1;
    const uint64_t c = ListOf(ix.cluster_begin, ix.nlist, pos);
    const float* q = rotq + static_cast<uint64_t>(qidx[i]) * ix.padded;
    const float d2 = E8ListD2(ix, q, c);
    const E8Row r = E8RowAt(ix, c, pos - ix.cluster_begin[c]);
    out[i] = fmaf(-(r.g1 + r.g1), E8ExactScore(ix, r, q), d2 + r.h3);
}

__global__ void E8ListD2Kernel(IndexView ix, const float* __restrict__ rotq, uint64_t nq, float* __restrict__ d2c) {
    const uint64_t t = blockIdx.x * 64ull + threadIdx.x;
    if (t >= nq * ix.nlist) return;
    const uint64_t q = t / ix.nlist, c = t % ix.nlist;
    d2c[t] = E8ListD2(ix, rotq + q * ix.padded, c);
}

// Distance of every stored vector (position order) for queries q0..q0+nqc.
__global__ void E8ExhaustiveKernel(IndexView ix, const float* __restrict__ rotq, const float* __restrict__ d2c,
                                   uint64_t q0, float* __restrict__ dist) {
    const uint64_t qj = blockIdx.y;
    const uint64_t q = q0 + qj;
    const float* qv = rotq + q * ix.padded;
    for (uint64_t i = blockIdx.x * 64ull + threadIdx.x; i < ix.n; i += gridDim.x * 64ull) {
        const uint64_t c = ListOf(ix.cluster_begin, ix.nlist, i);
        const E8Row r = E8RowAt(ix, c, i - ix.cluster_begin[c]);
        dist[qj * ix.n + i] = fmaf(-(r.g1 + r.g1), E8ExactScore(ix, r, qv), d2c[q * ix.nlist + c] + r.h3);
    }
}

// DecodeSplitBlockLane's 8-bit-digit form as ProbeBoundW calls it for every
// rate (e8q_scan.cpp:34-61): at b = 1 the empty field leaves digit 7 = the
// rowA byte's bit 0 and the branch 0; at b = 9 the digits keep their low 8
// bits.  Only b <= 9 (the host rejects more).
__device__ void E8BlockLiteral(int bits, uint8_t rb, const uint8_t* field, uint32_t d[8], uint32_t* br) {
    const int c = bits - 1;
    uint64_t x = 0;
    for (int b = 0; b < c; ++b) x |= static_cast<uint64_t>(field[b]) << (8 * b);
    const uint64_t m = (uint64_t{1} << c) - 1;
    uint32_t par = 0;
    for (int i = 0; i < 7; ++i) {
        const uint64_t f = (x >> (c * i)) & m;
        d[i] = static_cast<uint32_t>((f | (static_cast<uint64_t>((rb >> (7 - i)) & 1u) << c)) & 0xFFu);
        par ^= static_cast<uint32_t>(f & 1u);
    }
    const uint64_t f7 = (x >> (c * 7)) & m;
    d[7] = static_cast<uint32_t>((((f7 >> 1) << 1) | par | (static_cast<uint64_t>(rb & 1u) << c)) & 0xFFu);
    *br = static_cast<uint32_t>(f7 & 1u);
}

// ProbeBoundW (e8q_oracle.cpp:103-219) for pair i: rotq / rotx are the
// centered rotations of its query and its base row.
__global__ void E8ProbeKernel(IndexView ix, const double* __restrict__ levels, const float* __restrict__ rotq,
                              const float* __restrict__ rotx, const uint32_t* __restrict__ qidx,
                              const uint32_t* __restrict__ ids, uint64_t nq, uint64_t npair,
                              const uint32_t* __restrict__ inv, double* __restrict__ out,
                              uint32_t* __restrict__ bad) {
    const uint64_t i = blockIdx.x * 64ull + threadIdx.x;
    if (i >= npair) return;
    const double nan = __builtin_nan("");
    const uint32_t id = ids[i];
    if (qidx[i] >= nq || id >= ix.n) {   // the binding's out-of-range pair
        out[i] = nan;
        return;
    }
    const uint32_t p1 = inv[id];
    if (p1 == 0) {
        out[i] = nan;
        atomicOr(bad, 1u);
        return;
    }
    const uint64_t pos = p1 - 1, P = ix.padded;
    const uint64_t c = ListOf(ix.cluster_begin, ix.nlist, pos);
    const float* cc = ix.centroid_rot + c * P;
    const float* rq = rotq + static_cast<uint64_t>(qidx[i]) * P;
    const float* rx = rotx + i * P;
    double d2 = 0.0, rn2 = 0.0;
    for (uint64_t d = 0; d < P; ++d) {
        const double qr = static_cast<double>(rq[d]) - static_cast<double>(cc[d]);
        const double xr = static_cast<double>(rx[d]) - static_cast<double>(cc[d]);
        d2 = fma(qr, qr, d2);
        rn2 = fma(xr, xr, rn2);
    }
    const double rn = sqrt(rn2);
    if (!(rn > 0.0)) {
        out[i] = nan;
        return;
    }
    const int bits = ix.bits, cb = bits - 1;
    const uint32_t q2 = 1u << bits;
    const double rd = sqrt(static_cast<double>(P));
    const E8Row r = E8RowAt(ix, c, pos - ix.cluster_begin[c]);
    double dot_ue = 0, n2_e = 0, dot_ce = 0, dot_ew = 0, dot_uw = 0, n2_w = 0, dot_cw = 0, s_exact = 0, s_coarse = 0;
    for (int blk = 0; blk < ix.blocks; ++blk) {
        uint32_t dg[8], br;
        E8BlockLiteral(bits, r.rowa[blk], r.field + static_cast<uint64_t>(blk) * cb, dg, &br);
        for (int k = 0; k < 8; ++k) {
            const uint64_t j = static_cast<uint64_t>(blk) * 8 + k;
            const double unit = (static_cast<double>(rx[j]) - static_cast<double>(cc[j])) / rn;
            const double ve = levels[br * q2 + dg[k]] / rd;
            const double wc = ix.gate_mid[2 * k + (dg[k] >> cb)];
            dot_ue = fma(unit, ve, dot_ue);
            n2_e = fma(ve, ve, n2_e);
            dot_ce = fma(static_cast<double>(cc[j]), ve, dot_ce);
            dot_ew = fma(ve, wc, dot_ew);
            dot_uw = fma(unit, wc, dot_uw);
            n2_w = fma(wc, wc, n2_w);
            dot_cw = fma(static_cast<double>(cc[j]), wc, dot_cw);
            s_exact = fma(static_cast<double>(rq[j]), ve, s_exact);
            s_coarse = fma(static_cast<double>(rq[j]), wc, s_coarse);
        }
    }
    if (!(fabs(dot_ue) > 0.0) || !(fabs(dot_uw) > 0.0)) {
        out[i] = nan;
        return;
    }
    const double g1 = rn / dot_ue;
    const double h3 = fma(2.0 * g1, dot_ce, rn2);
    const double d_full = fma(-(2.0 * g1), s_exact, d2 + h3);
    const double w_x = (2.0 * rn) / dot_uw;
    const double d_coarse = fma(-w_x, s_coarse, fma(w_x, dot_cw, d2 + rn2));
    const double rx2 = fma(-((4.0 * g1) * w_x), dot_ew, fma((4.0 * g1) * g1, n2_e, (w_x * w_x) * n2_w));
    const double sqrt_dm1 = sqrt(static_cast<double>(ix.dim > 1 ? ix.dim - 1 : 1));
    const double omega = sqrt(rx2 > 0.0 ? rx2 : 0.0) / sqrt_dm1;
    const double r_q = sqrt(d2 > 0.0 ? d2 : 0.0);
    if (!(omega > 0.0) || !(r_q > 0.0)) {
        out[i] = nan;
        return;
    }
    out[i] = (d_coarse - d_full) / (r_q * omega);
}

// ---------------------------------------------------------------------------
// ME8Q

// ScoreFast (me8q_score_avx2.cpp:104-129) over one run's bit-tight code:
// component i of even blocks in acc[i], of odd blocks in acc2[i] (one chain
// at rate 1), FMA chains, then the horizontal sum of acc + acc2.
__device__ float M8ScoreFast(const M8Run& ru, const uint8_t* code, const float* q) {
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0}, acc2[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t bits = ru.rate;
    for (uint32_t b = 0; b < ru.nblk; ++b) {
        uint32_t d[8], br;
        M8Unpack(code + b * bits, bits, d, &br);
        const float* lev = ru.lev + (br << bits);
        const float* qb = q + b * 8;
        if (bits >= 2 && (b & 1u)) {
            for (int i = 0; i < 8; ++i) acc2[i] = fmaf(lev[d[i]], qb[i], acc2[i]);
        } else {
            for (int i = 0; i < 8; ++i) acc[i] = fmaf(lev[d[i]], qb[i], acc[i]);
        }
    }
    if (bits >= 2)
        for (int i = 0; i < 8; ++i) acc[i] = acc[i] + acc2[i];
    return HsumAvx(acc);
}

// EstimateRow (me8q_search.cpp:77-99): d2 + over the coded runs h2 - 2 coef S
// (the product fused into the subtraction; coef 0 adds h2).
__device__ double M8EstimateRow(const M8View& v, const uint8_t* row, double d2, const float* rq) {
    double acc = d2;
    for (uint32_t r = 0; r < v.nrun; ++r) {
        const float* fac = reinterpret_cast<const float*>(row + 8 * r);
        const double coef = fac[0];
        if (coef == 0.0) {
            acc += static_cast<double>(fac[1]);
            continue;
        }
        const M8Run& ru = v.runs[r];
        const double sc = M8ScoreFast(ru, row + ru.code_off, rq + ru.dim_off);
        acc += fma(-(2.0 * coef), sc, static_cast<double>(fac[1]));
    }
    return acc;
}

// fp64 squared distance of the zero-padded raw query to list c's unrotated
// centroid, dimension order, fused.
__device__ double M8ListD2(const M8View& v, const float* q, uint64_t c) {
    const float* cen = v.centroids + c * v.padded;
    double d2 = 0.0;
    for (uint64_t j = 0; j < v.padded; ++j) {
        const double d = static_cast<double>(j < v.dim ? q[j] : 0.0F) - static_cast<double>(cen[j]);
        d2 = fma(d, d, d2);
    }
    return d2;
}

__global__ void M8EstimateKernel(M8View v, const float* __restrict__ queries, const float* __restrict__ rotq,
                                 const uint32_t* __restrict__ qidx, const uint32_t* __restrict__ ids, uint64_t npair,
                                 const uint32_t* __restrict__ inv, float* __restrict__ out,
                                 uint32_t* __restrict__ bad) {
    const uint64_t i = blockIdx.x * 64ull + threadIdx.x;
    if (i >= npair) return;
    const uint32_t id = ids[i];
    const uint32_t p1 = id < v.n ? inv[id] : 0u;
    if (p1 == 0) {
        out[i] = __builtin_nanf("");
        atomicOr(bad, 1u);
        return;
    }
    const uint64_t pos = p1 - 1, q = qidx[i];
    const uint64_t c = ListOf(v.list_begin, v.nlist, pos);
    const double d2 = M8ListD2(v, queries + q * v.dim, c);
    out[i] = static_cast<float>(M8EstimateRow(v, v.codes + pos * v.stride, d2, rotq + q * v.padded));
}

__global__ void M8ListD2Kernel(M8View v, const float* __restrict__ queries, uint64_t nq, double* __restrict__ d2c) {
    const uint64_t t = blockIdx.x * 64ull + threadIdx.x;
    if (t >= nq * v.nlist) return;
    const uint64_t q = t / v.nlist, c = t % v.nlist;
    d2c[t] = M8ListD2(v, queries + q * v.dim, c);
}

__global__ void M8ExhaustiveKernel(M8View v, const float* __restrict__ rotq, const double* __restrict__ d2c,
                                   uint64_t q0, float* __restrict__ dist) {
    const uint64_t qj = blockIdx.y, q = q0 + qj;
    const float* rq = rotq + q * v.padded;
    for (uint64_t i = blockIdx.x * 64ull + threadIdx.x; i < v.n; i += gridDim.x * 64ull) {
        const uint64_t c = ListOf(v.list_begin, v.nlist, i);
        dist[qj * v.n + i] = static_cast<float>(M8EstimateRow(v, v.codes + i * v.stride, d2c[q * v.nlist + c], rq));
    }
}

// SearchIVF's candidates: query q's probed lists in probe order, each in
// row order; off[q * (take + 1) + p] = first candidate of probe p, count[q]
// = all of them.
__global__ void M8IvfOffsetsKernel(M8View v, const uint32_t* __restrict__ probe, uint32_t maxp, uint32_t take,
                                   uint64_t nq, uint64_t* __restrict__ off, uint64_t* __restrict__ count) {
    const uint64_t q = blockIdx.x * 64ull + threadIdx.x;
    if (q >= nq) return;
    uint64_t s = 0;
    for (uint32_t p = 0; p < take; ++p) {
        off[q * (take + 1) + p] = s;
        const uint32_t c = probe[q * maxp + p];
        s += v.list_begin[c + 1] - v.list_begin[c];
    }
    off[q * (take + 1) + take] = s;
    count[q] = s;
}

// One workgroup per (probe, query): the distances of the probed list's rows
// with the route's distance as d2, and their ids.
__global__ __launch_bounds__(256) void M8IvfKernel(M8View v, const float* __restrict__ rotq,
                                                   const uint32_t* __restrict__ probe, const float* __restrict__ pd,
                                                   uint32_t maxp, uint32_t take, const uint64_t* __restrict__ off,
                                                   uint64_t q0, uint64_t cap, float* __restrict__ dist,
                                                   uint32_t* __restrict__ cid) {
    const uint32_t p = blockIdx.x;
    const uint64_t qj = blockIdx.y, q = q0 + qj;
    const uint32_t c = probe[q * maxp + p];
    const double d2 = static_cast<double>(pd[q * maxp + p]);
    const uint64_t b0 = v.list_begin[c], cnt = v.list_begin[c + 1] - b0;
    const uint64_t o = off[q * (take + 1) + p];
    const float* rq = rotq + q * v.padded;
    for (uint64_t r = threadIdx.x; r < cnt; r += 256) {
        dist[qj * cap + o + r] = static_cast<float>(M8EstimateRow(v, v.codes + (b0 + r) * v.stride, d2, rq));
        cid[qj * cap + o + r] = v.ids[b0 + r];
    }
}

// ---------------------------------------------------------------------------
// Top-k of each query, one workgroup of 1024 threads.

struct SelShared {
    uint32_t hist[256];
    uint32_t wsum[16];
    uint32_t bin, before, nl, nt, run;
    unsigned long long keys[2048];
};

// Inclusive prefix sum over the workgroup of one value per thread.
__device__ uint32_t BlockScan(uint32_t v, uint32_t* wsum) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t j = 0; j < w; ++j) add += wsum[j];
    __syncthreads();
    return x + add;
}

// The rank-th smallest (1-based) of the m keys key(j) restricted to `width`
// top bits (64 or 32 of a 64-bit key), eight bits per pass; returns the key
// and sets *less to the count of keys below it.
template <class KeyF>
__device__ unsigned long long RadixSelect(const KeyF& key, uint64_t m, uint32_t rank, int width, SelShared& S,
                                          uint32_t* less) {
    unsigned long long prefix = 0, mask = 0;
    const uint32_t k = rank;
    for (int shift = 64 - 8; shift >= 64 - width; shift -= 8) {
        if (threadIdx.x < 256) S.hist[threadIdx.x] = 0;
        __syncthreads();
        for (uint64_t j = threadIdx.x; j < m; j += kSelThreads) {
            const unsigned long long kv = key(j);
            if ((kv & mask) == prefix) atomicAdd(&S.hist[(kv >> shift) & 255u], 1u);
        }
        __syncthreads();
        const uint32_t h = threadIdx.x < 256 ? S.hist[threadIdx.x] : 0u;
        const uint32_t incl = BlockScan(h, S.wsum);
        if (threadIdx.x < 256 && incl >= rank && incl - h < rank) {
            S.bin = threadIdx.x;
            S.before = incl - h;
        }
        __syncthreads();
        prefix |= static_cast<unsigned long long>(S.bin) << shift;
        mask |= 255ull << shift;
        rank -= S.before;
        __syncthreads();
    }
    *less = k - rank;
    return prefix;
}

// Ascending bitonic sort of S.keys[0, n2), n2 a power of two <= 2048.
__device__ void BitonicSort(SelShared& S, uint32_t n2) {
    for (uint32_t size = 2; size <= n2; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = threadIdx.x; t < n2 / 2; t += kSelThreads) {
                const uint32_t lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
                const bool up = (lo & size) == 0;
                const unsigned long long a = S.keys[lo], b = S.keys[hi];
                if ((a > b) == up) {
                    S.keys[lo] = b;
                    S.keys[hi] = a;
                }
            }
            __syncthreads();
        }
}

__device__ __forceinline__ uint32_t Pow2AtLeast(uint32_t n) {
    uint32_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

// kPush = false: the k best (distance, id) by (distance, id) -- E8Q's
// partial_sort under Better.  kPush = true: the list PushCandidate leaves
// after the candidates in index order (me8q_search.cpp:121-134): the
// candidates below the k-th distance d_k by distance, a later one first among
// equals; then, of the candidates at d_k among the first k (index order)
// at or below d_k, the last k - #below ones, latest first.
template <bool kPush>
__global__ __launch_bounds__(1024) void SelectKernel(const float* __restrict__ dist, const uint32_t* __restrict__ idv,
                                                     uint64_t id_stride, const uint64_t* __restrict__ count,
                                                     uint64_t m_fixed, uint64_t stride, uint64_t q0, uint32_t k,
                                                     uint32_t* __restrict__ out_ids, float* __restrict__ out_d) {
    __shared__ SelShared S;
    const uint64_t qj = blockIdx.x, q = q0 + qj;
    const uint64_t m = count ? count[q] : m_fixed;
    const float* dq = dist + qj * stride;
    const uint32_t* iq = idv + qj * id_stride;
    uint32_t nout = 0;
    if (m <= k) {   // every candidate
        for (uint64_t j = threadIdx.x; j < 2048; j += kSelThreads) S.keys[j] = ~0ull;
        __syncthreads();
        for (uint64_t j = threadIdx.x; j < m; j += kSelThreads)
            S.keys[j] = (static_cast<unsigned long long>(OrdOf(dq[j])) << 32) |
                        (kPush ? (0xFFFFFFFFu - static_cast<uint32_t>(j)) : iq[j]);
        __syncthreads();
        BitonicSort(S, Pow2AtLeast(static_cast<uint32_t>(m > 1 ? m : 1)));
        nout = static_cast<uint32_t>(m);
    } else if (!kPush) {
        auto key = [&](uint64_t j) {
            return (static_cast<unsigned long long>(OrdOf(dq[j])) << 32) | iq[j];
        };
        uint32_t less = 0;
        const unsigned long long kth = RadixSelect(key, m, k, 64, S, &less);
        if (threadIdx.x == 0) S.nl = 0;
        for (uint32_t j = threadIdx.x; j < 2048; j += kSelThreads) S.keys[j] = ~0ull;
        __syncthreads();
        for (uint64_t j = threadIdx.x; j < m; j += kSelThreads) {
            const unsigned long long kv = key(j);
            if (kv < kth) S.keys[atomicAdd(&S.nl, 1u)] = kv;
        }
        __syncthreads();
        for (uint32_t j = less + threadIdx.x; j < k; j += kSelThreads) S.keys[j] = kth;   // the ties at the k-th key
        __syncthreads();
        BitonicSort(S, Pow2AtLeast(k));
        nout = k;
    } else {
        auto key = [&](uint64_t j) { return static_cast<unsigned long long>(OrdOf(dq[j])) << 32; };
        uint32_t less = 0;
        const uint32_t uk = static_cast<uint32_t>(RadixSelect(key, m, k, 32, S, &less) >> 32);
        // keys[0, 1024): the candidates below d_k as (ord, ~index); keys[1024,
        // 2048): the candidates at d_k among the first k at or below it, by index
        if (threadIdx.x == 0) {
            S.nl = 0;
            S.nt = 0;
            S.run = 0;
        }
        for (uint32_t j = threadIdx.x; j < 2048; j += kSelThreads) S.keys[j] = ~0ull;
        __syncthreads();
        for (uint64_t base = 0; base < m; base += kSelThreads) {
            const uint64_t j = base + threadIdx.x;
            const uint32_t o = j < m ? OrdOf(dq[j]) : 0xFFFFFFFFu;
            const uint32_t le = (j < m && o <= uk) ? 1u : 0u;
            const uint32_t run = S.run;
            const uint32_t incl = BlockScan(le, S.wsum);
            if (le) {
                const uint32_t rank = run + incl - 1;
                if (o < uk) S.keys[atomicAdd(&S.nl, 1u)] = (static_cast<unsigned long long>(o) << 32) |
                                                           (0xFFFFFFFFu - static_cast<uint32_t>(j));
                else if (rank < k) S.keys[1024 + atomicAdd(&S.nt, 1u)] = j;
            }
            __syncthreads();
            if (threadIdx.x == kSelThreads - 1) S.run = run + incl;
            __syncthreads();
        }
        // sort both halves (the tie half by index), then move the last
        // k - less ties, latest first, behind the candidates below d_k
        BitonicSort(S, 2048);   // the tie entries (index < 2^32) sort below the (ord, ~index) keys
        const uint32_t nt = S.nt, nl = S.nl, want = k - less;
        __syncthreads();
        unsigned long long tie = 0, low = 0;
        const uint32_t t = threadIdx.x;
        if (t < want && t < nt) tie = S.keys[nt - 1 - t];   // the t-th latest tie (nt >= want)
        if (t < nl) low = S.keys[nt + t];
        __syncthreads();
        if (t < nl) S.keys[t] = low;
        if (t < want) S.keys[nl + t] = (static_cast<unsigned long long>(uk) << 32) | (0xFFFFFFFFu - static_cast<uint32_t>(tie));
        __syncthreads();
        nout = k;
    }
    for (uint32_t t = threadIdx.x; t < k; t += kSelThreads) {
        uint32_t id = 0xFFFFFFFFu;
        float d = __builtin_inff();
        if (t < nout) {
            const unsigned long long kv = S.keys[t];
            d = FloatOf(static_cast<uint32_t>(kv >> 32));
            const uint32_t lo = static_cast<uint32_t>(kv);
            id = kPush ? iq[0xFFFFFFFFu - lo] : lo;
        }
        out_ids[q * k + t] = id;
        out_d[q * k + t] = d;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers

hipError_t LaunchDiagInverseIds(hipStream_t st, const uint32_t* ids, uint64_t n, uint32_t* inv) {
    if (n == 0) return hipSuccess;
    const uint32_t g = static_cast<uint32_t>(std::min<uint64_t>((n + 255) / 256, 16384));
    hipLaunchKernelGGL(InverseIdsKernel, dim3(g), dim3(256), 0, st, ids, n, inv);
    return hipGetLastError();
}

hipError_t LaunchE8Estimate(hipStream_t st, const IndexView& ix, const float* rotq, const uint32_t* qidx,
                            const uint32_t* ids, uint64_t npair, const uint32_t* inv, float* out, uint32_t* bad) {
    if (npair == 0) return hipSuccess;
    hipLaunchKernelGGL(E8EstimateKernel, dim3(static_cast<uint32_t>((npair + 63) / 64)), dim3(64), 0, st, ix, rotq,
                       qidx, ids, npair, inv, out, bad);
    return hipGetLastError();
}

hipError_t LaunchE8ProbeBoundW(hipStream_t st, const IndexView& ix, const double* levels, const float* rotq,
                               const float* rotx, const uint32_t* qidx, const uint32_t* ids, uint64_t nq,
                               uint64_t npair, const uint32_t* inv, double* out, uint32_t* bad) {
    if (npair == 0) return hipSuccess;
    hipLaunchKernelGGL(E8ProbeKernel, dim3(static_cast<uint32_t>((npair + 63) / 64)), dim3(64), 0, st, ix, levels,
                       rotq, rotx, qidx, ids, nq, npair, inv, out, bad);
    return hipGetLastError();
}

hipError_t LaunchE8ListD2(hipStream_t st, const IndexView& ix, const float* rotq, uint64_t nq, float* d2c) {
    const uint64_t t = nq * ix.nlist;
    if (t == 0) return hipSuccess;
    hipLaunchKernelGGL(E8ListD2Kernel, dim3(static_cast<uint32_t>((t + 63) / 64)), dim3(64), 0, st, ix, rotq, nq, d2c);
    return hipGetLastError();
}

hipError_t LaunchE8Exhaustive(hipStream_t st, const IndexView& ix, const float* rotq, const float* d2c, uint64_t q0,
                              uint64_t nqc, float* dist) {
    if (nqc == 0 || ix.n == 0) return hipSuccess;
    const uint32_t gx = static_cast<uint32_t>(std::min<uint64_t>((ix.n + 63) / 64, 4096));
    hipLaunchKernelGGL(E8ExhaustiveKernel, dim3(gx, static_cast<uint32_t>(nqc)), dim3(64), 0, st, ix, rotq, d2c, q0,
                       dist);
    return hipGetLastError();
}

hipError_t LaunchM8Estimate(hipStream_t st, const M8View& v, const float* queries, const float* rotq,
                            const uint32_t* qidx, const uint32_t* ids, uint64_t npair, const uint32_t* inv, float* out,
                            uint32_t* bad) {
    if (npair == 0) return hipSuccess;
    hipLaunchKernelGGL(M8EstimateKernel, dim3(static_cast<uint32_t>((npair + 63) / 64)), dim3(64), 0, st, v, queries,
                       rotq, qidx, ids, npair, inv, out, bad);
    return hipGetLastError();
}

hipError_t LaunchM8ListD2(hipStream_t st, const M8View& v, const float* queries, uint64_t nq, double* d2c) {
    const uint64_t t = nq * v.nlist;
    if (t == 0) return hipSuccess;
    hipLaunchKernelGGL(M8ListD2Kernel, dim3(static_cast<uint32_t>((t + 63) / 64)), dim3(64), 0, st, v, queries, nq,
                       d2c);
    return hipGetLastError();
}

hipError_t LaunchM8Exhaustive(hipStream_t st, const M8View& v, const float* rotq, const double* d2c, uint64_t q0,
                              uint64_t nqc, float* dist) {
    if (nqc == 0 || v.n == 0) return hipSuccess;
    const uint32_t gx = static_cast<uint32_t>(std::min<uint64_t>((v.n + 63) / 64, 4096));
    hipLaunchKernelGGL(M8ExhaustiveKernel, dim3(gx, static_cast<uint32_t>(nqc)), dim3(64), 0, st, v, rotq, d2c, q0,
                       dist);
    return hipGetLastError();
}

hipError_t LaunchM8IvfOffsets(hipStream_t st, const M8View& v, const uint32_t* probe, uint32_t maxp, uint32_t take,
                              uint64_t nq, uint64_t* off, uint64_t* count) {
    if (nq == 0) return hipSuccess;
    hipLaunchKernelGGL(M8IvfOffsetsKernel, dim3(static_cast<uint32_t>((nq + 63) / 64)), dim3(64), 0, st, v, probe, maxp,
                       take, nq, off, count);
    return hipGetLastError();
}

hipError_t LaunchM8Ivf(hipStream_t st, const M8View& v, const float* rotq, const uint32_t* probe, const float* pd,
                       uint32_t maxp, uint32_t take, const uint64_t* off, uint64_t q0, uint64_t nqc, uint64_t cap,
                       float* dist, uint32_t* cid) {
    if (nqc == 0 || take == 0) return hipSuccess;
    hipLaunchKernelGGL(M8IvfKernel, dim3(take, static_cast<uint32_t>(nqc)), dim3(256), 0, st, v, rotq, probe, pd, maxp,
                       take, off, q0, cap, dist, cid);
    return hipGetLastError();
}

hipError_t LaunchDiagSelect(hipStream_t st, bool push, const float* dist, const uint32_t* idv, uint64_t id_stride,
                            const uint64_t* count, uint64_t m_fixed, uint64_t stride, uint64_t q0, uint64_t nqc,
                            uint32_t k, uint32_t* out_ids, float* out_d) {
    if (nqc == 0) return hipSuccess;
    if (k == 0 || k > kDiagMaxK) return hipErrorInvalidValue;
    if (push)
        hipLaunchKernelGGL(SelectKernel<true>, dim3(static_cast<uint32_t>(nqc)), dim3(kSelThreads), 0, st, dist, idv,
                           id_stride, count, m_fixed, stride, q0, k, out_ids, out_d);
    else
        hipLaunchKernelGGL(SelectKernel<false>, dim3(static_cast<uint32_t>(nqc)), dim3(kSelThreads), 0, st, dist, idv,
                           id_stride, count, m_fixed, stride, q0, k, out_ids, out_d);
    return hipGetLastError();
}

}  // namespace e8qdev
