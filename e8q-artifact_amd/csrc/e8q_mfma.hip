// e8q_mfma.hip -- the MFMA form of the E8Q search: the refine estimator of
// every vector of every probed list as a split-BF16 matrix contraction.
//
// The reference scores a survivor with S = sum_i q_i * level[br][d_i] / sqrt(P)
// (RefineCoreAvx2, e8q_search_avx2.cpp:82-403) and dist = h3' - 2 g1 S.  The
// level table is not linear in the digit, but it is reflection-symmetric,
// L1[d] = -L0[Q-1-d] (e8q_test.cpp:35-47, e8q_search_avx512.cpp:180-188), so
// with  l_i = br ? Q-1-d_i : d_i  and  s_i = br ? -1 : +1
//     S = sum_i sum_l  (q_i L0[l]) * (s_i [l == l_i])
// -- a contraction over K = 16 P of the query expansion Qexp[q, 16 i + l] =
// q_i L0[l] against one-hot code rows with entries in {-1, 0, +1}, exact in
// BF16.  Qexp is split into hi + lo BF16 (hi = bf16(x), lo = bf16(x - hi)),
// both contracted against the same code tile into one FP32 accumulator:
// v_mfma_f32_32x32x16_bf16, one instruction per dimension (its K = 16 is one
// dimension's folded levels) per 32 pairs x 32 vectors.  FP64 is not used at
// search time; the build-time fits (g1, h3) are the reference's.
//
// Candidates: the vectors whose MFMA distance is within the query's bound
// thr_q.  For k <= 32 the bound starts at +inf and every scored tile lowers
// it to the tile's k-th smallest distance for that query (atomicMin): any k
// scored vectors bound the final k-th distance from above, so every vector
// of the final top-k passes whenever it is scored.  For larger k it is
// tau_q + margin_q, tau_q being the k-th fp32 refine distance of the query's
// first probe (the threshold pass) and margin_q a bound of the two forms'
// rounding difference for that list's vectors, so the k vectors behind tau_q
// lie within it in MFMA distance.  One wave per query then selects the top-k.
#include <algorithm>

#include "e8q_scan_dev.h"

namespace e8qdev {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Index side: folded codes, one byte per dimension (level index in bits 0-3,
// branch in bit 4: the row of the one-hot table below), [record][block][lane][8]
// so that a lane reads a block's eight dimensions with one 8-byte load.  Decode as DecodeSplitBlockLane
// (e8q_scan.cpp:34-90): d_i = rowA bit (7-i) << (b-1) | field_i; digit 7's
// field LSB is the parity of digits 0..6's and its stored LSB the branch.

__device__ __forceinline__ uint32_t ListOfRecord(const IndexView& ix, uint64_t rec) {
    const uint64_t byte = rec * ix.batch_bytes;
    uint64_t lo = 0, hi = ix.nlist;   // last c with arena_begin[c] <= byte
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (ix.arena_begin[mid] <= byte) lo = mid;
        else hi = mid;
    }
    return static_cast<uint32_t>(lo);
}

__global__ __launch_bounds__(256) void MfmaCodesKernel(IndexView ix, uint64_t records, uint8_t* __restrict__ mcodes,
                                                        float* __restrict__ g1max) {
    const uint64_t tid = blockIdx.x * 256ull + threadIdx.x;
    if (tid >= records * 32) return;
    const uint64_t rec = tid >> 5;
    const uint32_t t = static_cast<uint32_t>(tid & 31);
    const uint32_t c = ListOfRecord(ix, rec);
    const uint64_t b = (rec * ix.batch_bytes - ix.arena_begin[c]) / ix.batch_bytes;   // batch within the list
    const uint64_t cnt = ix.cluster_begin[c + 1] - ix.cluster_begin[c];
    const int B = ix.blocks, cb = ix.bits - 1;
    const uint32_t Q = 1u << ix.bits;
    const uint8_t* rowa = ix.arena + rec * ix.batch_bytes + kHdr + static_cast<uint64_t>(t) * B;
    const uint8_t* rr = ix.ref_arena + ix.ref_arena_begin[c] + b * ix.ref_batch_bytes;
    const uint8_t* field = rr + static_cast<uint64_t>(t) * ix.refine_stride;
    uint8_t* out = mcodes + rec * 32ull * ix.padded + static_cast<uint64_t>(t) * 8;
    for (int blk = 0; blk < B; ++blk) {
        const uint32_t rb = rowa[blk];
        uint32_t d[8], br;
        if (ix.bits == 1) {
            uint32_t par = 0;
            for (int i = 0; i < 7; ++i) {
                d[i] = (rb >> (7 - i)) & 1u;
                par ^= d[i];
            }
            d[7] = par;
            br = rb & 1u;
        } else {
            uint64_t f = 0;   // 8 fields of cb <= 3 bits
            for (int i = 0; i < cb; ++i) f |= static_cast<uint64_t>(field[blk * cb + i]) << (8 * i);
            const uint32_t m = (1u << cb) - 1u;
            uint32_t par = 0;
            for (int i = 0; i < 7; ++i) {
                const uint32_t fi = static_cast<uint32_t>(f >> (cb * i)) & m;
                par ^= fi & 1u;
                d[i] = fi | (((rb >> (7 - i)) & 1u) << cb);
            }
            const uint32_t f7 = static_cast<uint32_t>(f >> (cb * 7)) & m;
            br = f7 & 1u;
            d[7] = (f7 & ~1u) | par | ((rb & 1u) << cb);
        }
        uint64_t w = 0;
        for (int i = 0; i < 8; ++i) {
            const uint32_t l = br ? Q - 1u - d[i] : d[i];
            w |= static_cast<uint64_t>(l | (br << 4)) << (8 * i);
        }
        *reinterpret_cast<uint64_t*>(out + static_cast<uint64_t>(blk) * 256) = w;
    }
    if (b * 32 + t < cnt) {   // |g1| of stored vectors; non-negative floats order as their bits
        const float g1 = reinterpret_cast<const float*>(rr + 32ull * ix.refine_stride)[t];
        atomicMax(reinterpret_cast<uint32_t*>(g1max) + c, __float_as_uint(fabsf(g1)));
    }
}

hipError_t LaunchMfmaCodes(hipStream_t st, const IndexView& ix, uint64_t records, uint8_t* mcodes, float* g1max) {
    if (records == 0) return hipSuccess;
    if (ix.bits < 1 || ix.bits > static_cast<int>(kMfmaMaxBits)) return hipErrorInvalidValue;
    const hipError_t e = LaunchFill32(st, reinterpret_cast<uint32_t*>(g1max), ix.nlist, 0);
    if (e != hipSuccess) return e;
    const uint64_t n = records * 32;
    hipLaunchKernelGGL(MfmaCodesKernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, st, ix, records,
                       mcodes, g1max);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Query side.

// Order-preserving map of a float distance to 32 bits (negative values
// reversed below the positive ones).
__device__ __forceinline__ uint32_t DistKey32(float d) {
    const uint32_t u = __float_as_uint(d);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Round to nearest even (no NaN reaches here).
__device__ __forceinline__ uint32_t Bf16Bits(float x) {
    const uint32_t u = __float_as_uint(x);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

// Qexp rows: thread per (query, dimension), 32 bf16 = 64 bytes.
__global__ __launch_bounds__(256) void MfmaQexpKernel(const float* __restrict__ rotq, uint64_t nq, uint32_t P,
                                                       const float* __restrict__ tab, uint32_t Q,
                                                       uint16_t* __restrict__ qexp) {
    const uint64_t tid = blockIdx.x * 256ull + threadIdx.x;
    if (tid >= nq * P) return;
    const float qi = rotq[tid];
    uint32_t hi[8], lo[8];
#pragma unroll
    for (int l2 = 0; l2 < 8; ++l2) {
        uint32_t h2 = 0, o2 = 0;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const uint32_t l = 2 * l2 + u;
            const float x = l < Q ? qi * tab[l] : 0.0F;
            const uint32_t hb = Bf16Bits(x);
            const uint32_t lb = Bf16Bits(x - __uint_as_float(hb << 16));
            h2 |= hb << (16 * u);
            o2 |= lb << (16 * u);
        }
        hi[l2] = h2;
        lo[l2] = o2;
    }
    u32x4* o = reinterpret_cast<u32x4*>(qexp + tid * 32);
    o[0] = u32x4{hi[0], hi[1], hi[2], hi[3]};
    o[1] = u32x4{hi[4], hi[5], hi[6], hi[7]};
    o[2] = u32x4{lo[0], lo[1], lo[2], lo[3]};
    o[3] = u32x4{lo[4], lo[5], lo[6], lo[7]};
}

// thr_q = tau + margin, one wave per query.  Both the MFMA score and the
// fp32 refine chain are within (2^-15 + (3P + 32) 2^-24) * sum_i |q_i| Lmax of
// the exact S (split residual 2^-17 per term, fp32 rounding of the products
// and of at most 2P + 16 accumulations, generously rounded up), and the
// distance moves by 2 |g1| times that; |g1| is bounded by the first list's
// max.  2^-20 |tau| covers the final fused multiply-add.
__global__ __launch_bounds__(kWave) void MfmaThrKernel(const float* __restrict__ rotq, uint32_t P, uint32_t k,
                                                        const float* __restrict__ first_dists,
                                                        const uint32_t* __restrict__ probe_ids, uint32_t maxp,
                                                        const float* __restrict__ g1max, float lmax,
                                                        uint32_t* __restrict__ thr) {
    const uint64_t q = blockIdx.x;
    float s = 0.0F;
    for (uint32_t i = threadIdx.x; i < P; i += kWave) s += fabsf(rotq[q * P + i]);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (threadIdx.x == 0) {
        const float tau = first_dists[q * k + k - 1];
        const float eps = 3.0517578e-05F + static_cast<float>(3 * P + 32) * 5.9604645e-08F;
        const float g1 = g1max[probe_ids[q * maxp]];
        const float margin = 2.0F * g1 * (s * 1.001F) * lmax * eps + fabsf(tau) * 9.5367432e-07F;
        thr[q] = DistKey32(tau + margin);   // +inf when the first list holds fewer than k vectors
    }
}

hipError_t LaunchMfmaQuery(hipStream_t st, const IndexView& ix, const float* rotq, uint64_t nq, uint32_t k,
                           const float* first_dists, const uint32_t* probe_ids, uint32_t maxp, const float* g1max,
                           float lmax, uint16_t* qexp, uint32_t* thr) {
    if (nq == 0) return hipSuccess;
    const uint32_t P = static_cast<uint32_t>(ix.padded);
    const uint64_t n = nq * P;
    hipLaunchKernelGGL(MfmaQexpKernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, st, rotq, nq, P,
                       ix.refine_tab, 1u << ix.bits, qexp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || first_dists == nullptr) return e;
    hipLaunchKernelGGL(MfmaThrKernel, dim3(static_cast<uint32_t>(nq)), dim3(kWave), 0, st, rotq, P, k, first_dists,
                       probe_ids, maxp, g1max, lmax, thr);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The contraction.  Items of the batched work list with 32 pairs x 16
// batches; four waves per workgroup, each scoring the 32 pairs against four
// 32-vector tiles (batches wave, wave + 4, wave + 8, wave + 12), so that one
// load of the A operand feeds eight MFMAs.  Per dimension a lane supplies 8
// bf16 of the A operand (query row lane % 32, levels 8 (lane / 32) .. +7, hi
// and lo) and 8 of each B operand (vector column lane % 32, the same
// levels); accumulator register r holds row 8 (r / 4) + 4 (lane / 32) + r % 4,
// column lane % 32.

constexpr int kMfmaTiles = 4;

// The one-hot B slices: entry 32 h + code holds the 8 bf16 of levels
// 8 h .. 8 h + 7 for a folded code (level l, branch s): -1 or +1 at l, else 0.
__device__ __forceinline__ void FillOneHotTable(u32x4* tab) {
    for (uint32_t e = threadIdx.x; e < 64; e += blockDim.x) {
        const uint32_t h = e >> 5, code = e & 31, l = code & 15, s = code >> 4;
        uint32_t dw[4] = {0u, 0u, 0u, 0u};
        if (l >= 8 * h && l < 8 * h + 8) {
            const uint32_t j = l - 8 * h;
            dw[j >> 1] = (s ? 0xBF80u : 0x3F80u) << (16 * (j & 1));
        }
        tab[e] = u32x4{dw[0], dw[1], dw[2], dw[3]};
    }
}

__device__ __forceinline__ f32x16 Mfma(u32x4 a, u32x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

__device__ __forceinline__ float KeyDist(uint64_t key) {
    const uint32_t m = static_cast<uint32_t>(key >> 32);
    return __uint_as_float((m & 0x80000000u) ? (m & 0x7FFFFFFFu) : ~m);
}

// The k-th smallest (1 <= k <= 32) of the 32 values held by the lanes of
// this lane's half of the wave: a bitonic sort across the half.
__device__ __forceinline__ uint32_t HalfKth(uint32_t v, uint32_t n, uint32_t lane, uint32_t k) {
#pragma unroll
    for (uint32_t size = 2; size <= 32; size <<= 1)
#pragma unroll
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            const uint32_t o = static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), static_cast<int>(stride)));
            const bool keep_min = ((n & size) == 0) == ((n & stride) == 0);
            v = keep_min ? min(v, o) : max(v, o);
        }
    return static_cast<uint32_t>(__shfl(static_cast<int>(v), static_cast<int>((lane & 32u) | (k - 1))));
}

constexpr int kMfmaThreads = 256;

__global__ __launch_bounds__(kMfmaThreads) void MfmaScanKernel(const MfmaParams* __restrict__ prm) {
    const IndexView& ix = prm->bp.ix;
    const BatchedArgs& a = prm->bp.a;
    const MfmaExtra& m = prm->m;
    __shared__ uint32_t s_it, s_q[kMfmaQTile], s_thr[kMfmaQTile];
    __shared__ float s_d2[kMfmaQTile];
    __shared__ u32x4 s_onehot[64];
    // The query operand of one 8-dimension block for the item's 32 pairs, two
    // buffers: [buffer][dimension][pair][4 x 16 bytes: hi 0-7, hi 8-15, lo 0-7,
    // lo 8-15] with a 16-byte pad per row (rows 80 bytes apart spread a
    // half-wave's reads over the banks).
    __shared__ u32x4 s_a[2 * 8 * kMfmaQTile * 5];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t h = lane >> 5, n = lane & 31;
    FillOneHotTable(s_onehot);
    const u32x4* onehot = s_onehot + 32 * h;   // this half's levels
    const uint32_t P = static_cast<uint32_t>(ix.padded);
    const int B = ix.blocks;
    const uint32_t k = a.k;
    const bool dyn = m.dyn != 0;
    const uint32_t nitems = Uniform(static_cast<uint32_t>(min<uint64_t>(*G(a.counters), a.items_cap)));
    unsigned long long staged = 0;
    for (;;) {
        if (threadIdx.x == 0) s_it = atomicAdd(&a.counters[1], 1u);
        __syncthreads();
        const uint32_t it = Uniform(s_it);
        if (it >= nitems) break;
        const E8Q_GLOBAL BatchItem* item = G(a.items) + it;
        const uint32_t c = Uniform(item->c), np = Uniform(item->np), nb = Uniform(item->nb), b0 = Uniform(item->b0);
        const uint32_t p0 = Uniform(item->p0);
        if (threadIdx.x < np) {
            const uint64_t pr = G(a.pairs)[p0 + threadIdx.x];
            const uint32_t q = static_cast<uint32_t>(pr >> 32), p = static_cast<uint32_t>(pr);
            s_q[threadIdx.x] = q;
            s_d2[threadIdx.x] = G(a.probe_d2)[static_cast<uint64_t>(q) * a.maxp + p];
            s_thr[threadIdx.x] = m.thr[q];
        }
        __syncthreads();
        // Every wave stages the query operand; waves past nb then skip the
        // contraction (wave-uniform).  This wave's tiles: batches wave + 4 t.
        const bool active = wave < nb;
        const uint32_t nt = active ? min(static_cast<uint32_t>(kMfmaTiles), (nb - wave + 3) / 4) : 0u;
        const uint64_t ccnt = G(ix.cluster_begin)[c + 1] - G(ix.cluster_begin)[c];
        const uint64_t rec0 = G(ix.arena_begin)[c] / ix.batch_bytes + b0 + (active ? wave : 0u);
        const E8Q_GLOBAL uint8_t* code[kMfmaTiles];
#pragma unroll
        for (int t = 0; t < kMfmaTiles; ++t)   // a tile past nt reads the first one (unused)
            code[t] = G(m.mcodes) + (rec0 + 4 * (t < static_cast<int>(nt) ? t : 0)) * 32ull * P + n * 8;
        // Staging: thread t carries pair t / 8, dimension t % 8 of each block
        // (pairs past np carry the first pair's row: computed, never read).
        const uint32_t mq = threadIdx.x >> 3, eq = threadIdx.x & 7;
        const E8Q_GLOBAL u32x4* asrc = reinterpret_cast<const E8Q_GLOBAL u32x4*>(G(m.qexp)) +
                                       (static_cast<uint64_t>(s_q[mq < np ? mq : 0]) * P + eq) * 4;
        u32x4* adst = s_a + (eq * kMfmaQTile + mq) * 5;   // + buffer * 8 * 32 * 5
        u32x4 pf[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) pf[i] = asrc[i];
#pragma unroll
        for (int i = 0; i < 4; ++i) adst[i] = pf[i];
        __syncthreads();
        f32x16 acc[kMfmaTiles] = {};
        for (int blk = 0; blk < B; ++blk) {
            const uint32_t cur = static_cast<uint32_t>(blk) & 1u;
            const bool more = blk + 1 < B;
            if (more) {   // the next block's operand, in flight during this block's MFMAs
#pragma unroll
                for (int i = 0; i < 4; ++i) pf[i] = asrc[static_cast<uint64_t>(blk + 1) * 32 + i];
            }
            if (active) {
                uint64_t w[kMfmaTiles];
#pragma unroll
                for (int t = 0; t < kMfmaTiles; ++t) w[t] = *reinterpret_cast<const E8Q_GLOBAL uint64_t*>(code[t] + blk * 256);
                const u32x4* arow = s_a + (cur * 8 * kMfmaQTile + n) * 5;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const u32x4 xh = arow[e * kMfmaQTile * 5 + h], xl = arow[e * kMfmaQTile * 5 + 2 + h];
                    u32x4 bv[kMfmaTiles];
#pragma unroll
                    for (int t = 0; t < kMfmaTiles; ++t) bv[t] = onehot[static_cast<uint32_t>(w[t] >> (8 * e)) & 31u];
                    // hi for every tile, then lo: no MFMA waits on the previous one's result
#pragma unroll
                    for (int t = 0; t < kMfmaTiles; ++t)
                        if (t < static_cast<int>(nt)) acc[t] = Mfma(xh, bv[t], acc[t]);
#pragma unroll
                    for (int t = 0; t < kMfmaTiles; ++t)
                        if (t < static_cast<int>(nt)) acc[t] = Mfma(xl, bv[t], acc[t]);
                }
            }
            if (more) {   // the other buffer was last read before the previous barrier
                u32x4* d = adst + (cur ^ 1u) * 8 * kMfmaQTile * 5;
#pragma unroll
                for (int i = 0; i < 4; ++i) d[i] = pf[i];
            }
            __syncthreads();
        }
        if (active) {
            // Epilogue: distances, the bound, the candidate test, appends.
            uint32_t tk[16];   // this lane's rows' bounds
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const uint32_t row = 8 * (r / 4) + 4 * h + (r % 4);
                tk[r] = row < np ? s_thr[row] : 0u;
            }
#pragma unroll
            for (int t = 0; t < kMfmaTiles; ++t) {
                if (t >= static_cast<int>(nt)) break;
                const uint32_t bj = b0 + wave + 4 * t;   // batch within the list
                const uint64_t vo = static_cast<uint64_t>(bj) * 32 + n;
                const bool vv = vo < ccnt;
                float g1 = 0.0F, h3 = 0.0F;
                uint32_t id = 0;
                if (vv) {
                    const E8Q_GLOBAL uint8_t* rr =
                        G(ix.ref_arena) + G(ix.ref_arena_begin)[c] + static_cast<uint64_t>(bj) * ix.ref_batch_bytes;
                    const E8Q_GLOBAL float* g1p =
                        reinterpret_cast<const E8Q_GLOBAL float*>(rr + 32ull * ix.refine_stride);
                    g1 = g1p[n];
                    h3 = g1p[32 + n];
                    id = G(ix.ids)[G(ix.cluster_begin)[c] + vo];
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const uint32_t row = 8 * (r / 4) + 4 * h + (r % 4);
                    const bool ok = vv && row < np;
                    const float d = fmaf(-(2.0F * g1), acc[t][r], s_d2[row < np ? row : 0] + h3) + 0.0F;   // -0 -> +0
                    const uint32_t key = ok ? DistKey32(d) : 0xFFFFFFFFu;
                    if (dyn) tk[r] = min(tk[r], HalfKth(key, n, lane, k));
                    if (!ok || key > tk[r]) continue;
                    const uint32_t q = s_q[row];
                    const uint32_t slot = atomicAdd(&m.ncand[q], 1u);
                    const uint64_t crow = m.qslot ? G(m.qslot)[q] : q;
                    if (slot < m.cap) Gw(m.cand)[crow * m.cap + slot] = (static_cast<uint64_t>(key) << 32) | id;
                }
            }
            if (dyn && n == 0) {   // one lane per half publishes its rows' bounds
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const uint32_t row = 8 * (r / 4) + 4 * h + (r % 4);
                    if (row < np && tk[r] < s_thr[row]) atomicMin(&m.thr[s_q[row]], tk[r]);
                }
            }
        }
        if (threadIdx.x == 0) staged += nb;
        __syncthreads();
    }
    if (threadIdx.x == 0 && a.stats) atomicAdd(&a.stats[3], staged);
}

hipError_t LaunchMfmaScan(hipStream_t st, const MfmaParams* dp, uint32_t grid) {
    hipLaunchKernelGGL(MfmaScanKernel, dim3(grid), dim3(kMfmaThreads), 0, st, dp);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Top-k of each query's candidates, one wave per query: a bitonic sort in
// LDS of the best k so far plus the next chunk, repeated over the list.

constexpr uint32_t kSelChunk = 1024;

__global__ __launch_bounds__(kWave) void MfmaSelectKernel(const MfmaParams* __restrict__ prm) {
    const BatchedArgs& a = prm->bp.a;
    const MfmaExtra& m = prm->m;
    __shared__ uint64_t buf[kSelChunk];
    const uint32_t lane = threadIdx.x;
    const uint32_t k = a.k;
    const uint64_t nrun = m.qsel ? m.nsel : a.nq;
    for (uint64_t j = blockIdx.x; j < nrun; j += gridDim.x) {
        const uint32_t q = m.qsel ? G(m.qsel)[j] : static_cast<uint32_t>(j);
        const uint32_t n = Uniform(*G(m.ncand + q));
        if (n > m.cap) {   // searched again with a longer list
            if (lane == 0) a.overflow[atomicAdd(&a.counters[2], 1u)] = q;
            continue;
        }
        const uint64_t crow = m.qslot ? G(m.qslot)[q] : q;
        const E8Q_GLOBAL uint64_t* cl = G(m.cand) + crow * m.cap;
        // chunk size: a power of two >= k + min(n, room), at least 64
        uint32_t ch = 64;
        while (ch < kSelChunk && ch < k + n) ch <<= 1;
        const uint32_t room = ch - k;
        for (uint32_t i = lane; i < k; i += kWave) buf[i] = ~0ull;
        uint32_t s = 0;
        do {
            const uint32_t take = min(room, n - s);
            for (uint32_t i = lane; i < room; i += kWave) buf[k + i] = i < take ? cl[s + i] : ~0ull;
            __syncthreads();
            for (uint32_t size = 2; size <= ch; size <<= 1)
                for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                    for (uint32_t i = lane; i < ch / 2; i += kWave) {
                        const uint32_t lo = 2 * i - (i & (stride - 1));
                        const uint32_t hi = lo + stride;
                        const bool up = (lo & size) == 0;
                        const uint64_t x = buf[lo], y = buf[hi];
                        if ((x > y) == up) {
                            buf[lo] = y;
                            buf[hi] = x;
                        }
                    }
                    __syncthreads();
                }
            s += take;
        } while (s < n);
        for (uint32_t i = lane; i < k; i += kWave) {
            const uint64_t key = buf[i];
            const bool ok = key != ~0ull;
            Gw(a.out_ids)[static_cast<uint64_t>(q) * k + i] = ok ? static_cast<uint32_t>(key) : 0xFFFFFFFFu;
            Gw(a.out_dists)[static_cast<uint64_t>(q) * k + i] = ok ? KeyDist(key) : __builtin_inff();
        }
        if (lane == 0 && a.stats) atomicAdd(&a.stats[1], static_cast<unsigned long long>(n));
        __syncthreads();
    }
}

hipError_t LaunchMfmaSelect(hipStream_t st, const MfmaParams* dp, uint64_t nrun, uint32_t k) {
    if (nrun == 0) return hipSuccess;
    if (k == 0 || k > 512) return hipErrorInvalidValue;
    const uint32_t grid = static_cast<uint32_t>(nrun < 262144 ? nrun : 262144);
    hipLaunchKernelGGL(MfmaSelectKernel, dim3(grid), dim3(kWave), 0, st, dp);
    return hipGetLastError();
}

}  // namespace e8qdev
