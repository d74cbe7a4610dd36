// me8q_batch.hip -- the batched form of the ME8Q search: every query of a
// call as one (query, list) work list, the coarse integer sums of all its
// (query, vector) pairs as an exact int8 MFMA contraction.
//
// The reference scans one query at a time (SearchIVFFast, me8q_fast.cpp:
// 380-1108): per list, the integer coarse sums of each run
// (AccumulateCoarseHi, a u8 LUT entry per nibble of the top-bit plane), the
// ZPR coarse bound against the live k-th distance (me8q_prune_avx2.cpp:
// 9-98), then the exact cascade of the lanes that pass (flush4, me8q_fast.cpp:
// 856-979).  The live threshold makes each query sequential, and under Zipf
// list sizes one query's head list holds a large share of the index.  Here:
//
//   1. threshold pass: the plain kernel over each query's first `budget`
//      32-vector batches in probe order; its top-k is the query's starting
//      list and tau_q, its k-th distance, the threshold;
//   2. the (query, list) pairs are grouped by list and cut into items of 32
//      pairs x 8 batches; a persistent kernel takes items from an atomic
//      counter and computes the coarse sums of all 32 x 256 pairs of an item
//      with v_mfma_i32_32x32x32_i8: the A operand is the query's LUT row of a
//      plane column shifted to int8 (u8 ^ 0x80 = u8 - 128; 16 low-nibble
//      entries in one half of K, 16 high-nibble entries in the other), the B
//      operand the one-hot int8 vector of the vector's nibble (from a
//      16-entry LDS table), so one instruction sums one plane column for 32
//      queries x 32 vectors and the integer sums are exact (minus 256 per
//      column, added back).  Per run the kernel then evaluates the
//      reference's coarse bound terms, and after the last run tests
//      dc - (z sqrt(pool) + qa) against tau_q; passing vectors are appended
//      to the query's candidate list;
//   3. refine: per candidate the coarse terms again in the reference's own
//      arithmetic (fp32 for lanes its AVX2 filter covers, fp64 for the rest),
//      the run order of the reference's cascade for a group of one
//      candidate (decreasing t^2 / nblk), the exact run scores (the plain
//      kernel's even/odd block chains and AVX horizontal sums) and the
//      cascade's death test against tau_q; the survivors' exact distances
//      below tau_q go to a short per-query list;
//   4. select: one wave per query sorts its threshold-pass list and its
//      survivors by (distance, id) and writes the top-k.
//
// With z = +inf nothing is pruned in either form and the result is the
// exact estimator top-k over the probed lists, as the plain form's.  With a
// finite z the bounds are the reference's with tau_q in place of the live
// threshold, which is at least the reference's threshold after the same
// vectors.  Queries whose candidate or survivor list overflows are searched
// again by the plain kernel.
#include <hip/hip_runtime.h>

#include "e8q_scan_dev.h"
#include "me8q_device.h"
#include "me8q_kernels.h"

namespace e8qdev {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

// Order-preserving 32-bit key of a float distance (negatives below positives).
__device__ __forceinline__ uint32_t M8Key(float d) {
    const uint32_t u = __float_as_uint(d);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float M8KeyDist(uint32_t m) {
    return __uint_as_float((m & 0x80000000u) ? (m & 0x7FFFFFFFu) : ~m);
}

// ---------------------------------------------------------------------------
// The contraction.  256 threads = four waves; wave w scores the item's 32
// pairs against batches b0 + w and b0 + w + 4.  Accumulator register i of
// lane l holds row (pair) 8 (i / 4) + 4 (l / 32) + i % 4, column (vector)
// l % 32.

constexpr int kM8Tiles = 2;
constexpr uint32_t kM8ARow = 48;   // bytes per staged (column, pair) row: 32 + 16 pad

template <bool kChk>
__global__ __launch_bounds__(256, 2) void M8MfmaScanKernel(M8BatchArgs a) {
    const M8View& v = a.v;
    __shared__ uint32_t s_it;
    __shared__ uint32_t s_q[kM8BatchQTile], s_p[kM8BatchQTile], s_ps[kM8BatchQTile], s_bs[kM8BatchQTile];
    __shared__ float s_tau[kM8BatchQTile], s_sqk[kM8BatchQTile], s_d2[kM8BatchQTile], s_dl[kM8BatchQTile];
    __shared__ float s_lmx[kM8BatchRuns][kM8BatchQTile], s_qe[kM8BatchRuns][kM8BatchQTile];
    __shared__ float s_ra[kM8BatchRuns][kM8BatchQTile], s_wf[kM8BatchRuns][kM8BatchQTile];
    __shared__ __align__(16) uint32_t s_oh[16 * 4];
    // The query operand of eight plane columns for the item's 32 pairs, two
    // buffers: [buffer][column][pair][32 bytes + pad], already int8.
    __shared__ __align__(16) uint8_t s_a[2 * 8 * kM8BatchQTile * kM8ARow];
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint32_t h = lane >> 5, n = lane & 31;
    if (tid < 16)
        for (uint32_t j = 0; j < 4; ++j) s_oh[tid * 4 + j] = (tid >> 2) == j ? (1u << (8 * (tid & 3))) : 0u;
    const uint32_t PC = v.plane_cols, R = v.nrun, k = a.k, maxp = a.maxp;
    const float finf = __builtin_inff();
    const uint32_t nitems = Uniform(static_cast<uint32_t>(min<uint64_t>(*G(a.counters), a.items_cap)));
    unsigned long long n_pairs = 0, n_cand = 0, n_items = 0;
    for (;;) {
        if (tid == 0) s_it = atomicAdd(&a.counters[1], 1u);
        __syncthreads();
        const uint32_t it = Uniform(s_it);
        if (it >= nitems) break;
        const E8Q_GLOBAL BatchItem* item = G(a.items) + it;
        const uint32_t c = Uniform(item->c), np = Uniform(item->np), nb = Uniform(item->nb), b0 = Uniform(item->b0);
        const uint32_t p0 = Uniform(item->p0);
        if (tid < kM8BatchQTile) {   // rows past np carry the first pair (computed, never appended)
            const uint64_t pr = G(a.pairs)[p0 + (tid < np ? tid : 0u)];
            const uint32_t q = static_cast<uint32_t>(pr >> 32), p = static_cast<uint32_t>(pr);
            const float tau = G(a.first_d)[static_cast<uint64_t>(q) * k + k - 1];
            const uint64_t cv = G(a.cover)[q];
            s_q[tid] = q;
            s_p[tid] = p;
            s_ps[tid] = static_cast<uint32_t>(cv >> 32);
            s_bs[tid] = static_cast<uint32_t>(cv);
            s_tau[tid] = tau;
            s_sqk[tid] = tau < finf ? sqrtf(fmaxf(0.0F, tau)) : finf;
            s_d2[tid] = G(a.pdist)[static_cast<uint64_t>(q) * maxp + p];
            s_dl[tid] = G(a.qc)[static_cast<uint64_t>(q) * 4 + 1];
            for (uint32_t r = 0; r < R; ++r) {
                const E8Q_GLOBAL double* o = G(a.rc) + (static_cast<uint64_t>(q) * R + r) * 4;
                s_lmx[r][tid] = static_cast<float>(o[1]);
                s_qe[r][tid] = static_cast<float>(o[2]);
                s_wf[r][tid] = static_cast<float>(0.5 * (o[3] + o[2]));
                s_ra[r][tid] = static_cast<float>(G(a.sqd2)[(static_cast<uint64_t>(q) * maxp + p) * R + r]);
            }
        }
        __syncthreads();
        const uint32_t nt = (wave < nb ? 1u : 0u) + (wave + 4 < nb ? 1u : 0u);
        const uint64_t lb = G(v.list_begin)[c], cnt = G(v.list_begin)[c + 1] - lb;
        const uint64_t bat0 = G(v.batch_begin)[c];
        const E8Q_GLOBAL uint8_t* prow[kM8Tiles];
#pragma unroll
        for (int t = 0; t < kM8Tiles; ++t) {   // a tile past nt reads the first one (unused)
            const uint32_t bj = b0 + wave + 4 * (t < static_cast<int>(nt) ? t : 0);
            prow[t] = G(v.plane) + (bat0 + bj) * v.rec + static_cast<uint64_t>(n) * PC;
        }
        // Staging: thread t carries pair t / 8, column t % 8 of each 8-column block.
        const uint32_t mq = tid >> 3, eq = tid & 7;
        const E8Q_GLOBAL u32x4* asrc =
            reinterpret_cast<const E8Q_GLOBAL u32x4*>(G(a.lut) + static_cast<uint64_t>(s_q[mq]) * PC * 32 + eq * 32);
        uint8_t* adst = s_a + (eq * kM8BatchQTile + mq) * kM8ARow;
        const uint32_t kFlip = 0x80808080u;
        u32x4 pf0 = asrc[0], pf1 = asrc[1];
        *reinterpret_cast<u32x4*>(adst) = pf0 ^ kFlip;
        *reinterpret_cast<u32x4*>(adst + 16) = pf1 ^ kFlip;
        __syncthreads();
        i32x16 acc[kM8Tiles];
        float dc[kM8Tiles][16], pool[kM8Tiles][16], qw[kM8Tiles][16];
        uint32_t ck[kM8Tiles][16];
#pragma unroll
        for (int t = 0; t < kM8Tiles; ++t) {
            acc[t] = i32x16{};
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                dc[t][i] = s_d2[8 * (i / 4) + 4 * h + (i % 4)];
                pool[t][i] = 0.0F;
                qw[t][i] = 0.0F;
                ck[t][i] = 0u;
            }
        }
        uint32_t r = 0, rend = R > 0 ? Uniform(v.runs[0].blk0 + v.runs[0].nblk) : PC;
        const uint32_t ncb = PC / 8;
        for (uint32_t cb = 0; cb < ncb; ++cb) {
            const uint32_t cur = cb & 1u;
            const bool more = cb + 1 < ncb;
            if (more) {   // the next block's operand, in flight during this block's MFMAs
                pf0 = asrc[(cb + 1) * 16];
                pf1 = asrc[(cb + 1) * 16 + 1];
            }
            if (nt > 0) {
                uint64_t w[kM8Tiles];
#pragma unroll
                for (int t = 0; t < kM8Tiles; ++t) w[t] = *reinterpret_cast<const E8Q_GLOBAL uint64_t*>(prow[t] + cb * 8);
                const uint8_t* arow = s_a + cur * 8 * kM8BatchQTile * kM8ARow + n * kM8ARow + 16 * h;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const i32x4 av = *reinterpret_cast<const i32x4*>(arow + e * kM8BatchQTile * kM8ARow);
#pragma unroll
                    for (int t = 0; t < kM8Tiles; ++t) {
                        if (t >= static_cast<int>(nt)) break;
                        const uint32_t nib = static_cast<uint32_t>(w[t] >> (8 * e + 4 * h)) & 15u;
                        const i32x4 bv = *reinterpret_cast<const i32x4*>(s_oh + nib * 4);
                        acc[t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bv, acc[t], 0, 0, 0);
                    }
                }
            }
            if (more) {   // the other buffer was last read before the previous barrier
                uint8_t* d = adst + (cur ^ 1u) * 8 * kM8BatchQTile * kM8ARow;
                *reinterpret_cast<u32x4*>(d) = pf0 ^ kFlip;
                *reinterpret_cast<u32x4*>(d + 16) = pf1 ^ kFlip;
            }
            __syncthreads();
            if ((cb + 1) * 8 == rend) {   // run r ends here: its terms of the coarse bound
                if (nt > 0) {
                    const int32_t shift = 256 * static_cast<int32_t>(Uniform(v.runs[r].nblk));
#pragma unroll
                    for (int t = 0; t < kM8Tiles; ++t) {
                        if (t >= static_cast<int>(nt)) break;
                        const uint32_t bj = b0 + wave + 4 * t;
                        const E8Q_GLOBAL uint8_t* bp = G(v.plane) + (bat0 + bj) * v.rec + 32ull * PC + r * kM8BankBytes;
                        const E8Q_GLOBAL float* sc = reinterpret_cast<const E8Q_GLOBAL float*>(bp + 128);
                        const float bq = fmaf(static_cast<float>(bp[n]), sc[1], sc[0]);
                        const float wq = fmaf(static_cast<float>(bp[32 + n]), sc[3], sc[2]);
                        const float pq = fmaf(static_cast<float>(bp[64 + n]), sc[5], sc[4]);
                        const float aw = fabsf(wq), s3 = sc[3];
#pragma unroll
                        for (int i = 0; i < 16; ++i) {
                            const uint32_t row = 8 * (i / 4) + 4 * h + (i % 4);
                            const int32_t iv = acc[t][i] + shift;
                            const float s_c = fmaf(static_cast<float>(iv), s_dl[row], s_lmx[r][row]);
                            dc[t][i] = fmaf(-wq, s_c, dc[t][i]) + bq;
                            const float ra = s_ra[r][row], sk = s_sqk[row];
                            const float tt = pq * (sk < ra ? sk : ra);
                            pool[t][i] = fmaf(tt, tt, pool[t][i]);
                            qw[t][i] = fmaf(aw, s_qe[r][row], qw[t][i]) + s3 * s_wf[r][row];
                            if (kChk) ck[t][i] += (r + 1) * static_cast<uint32_t>(iv);
                        }
                        acc[t] = i32x16{};
                    }
                }
                ++r;
                rend = r < R ? Uniform(v.runs[r].blk0 + v.runs[r].nblk) : PC;
            }
        }
        if (nt > 0) {   // the bound test and the appends
            const float z = a.z;
#pragma unroll
            for (int t = 0; t < kM8Tiles; ++t) {
                if (t >= static_cast<int>(nt)) break;
                const uint32_t bj = b0 + wave + 4 * t;
                const uint64_t pos = static_cast<uint64_t>(bj) * 32 + n;
                const bool vv = pos < cnt;
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const uint32_t row = 8 * (i / 4) + 4 * h + (i % 4);
                    const uint32_t q = s_q[row], p = s_p[row];
                    const bool scored = vv && row < np;
                    const bool fresh = p > s_ps[row] || (p == s_ps[row] && bj >= s_bs[row]);
                    const float slack = fmaf(z, sqrtf(pool[t][i]), qw[t][i]);
                    const float tau = s_tau[row];
                    const float margin = 2.4414062e-4F * (fabsf(dc[t][i]) + fabsf(slack) + fabsf(tau));
                    const bool kept = scored && fresh && !(dc[t][i] - slack >= tau + margin);
                    const uint64_t m = __ballot(kept);
                    const uint32_t hm = static_cast<uint32_t>(m >> (32 * h));
                    uint32_t base = 0;
                    if (n == 0 && hm != 0) base = atomicAdd(&a.ncand[q], static_cast<uint32_t>(__popc(hm)));
                    base = static_cast<uint32_t>(__shfl(static_cast<int>(base), static_cast<int>(lane & 32u)));
                    if (kept) {
                        const uint32_t slot = base + __popc(hm & ((1u << n) - 1u));
                        if (slot < a.cap) Gw(a.cand)[static_cast<uint64_t>(q) * a.cap + slot] = (static_cast<uint64_t>(p) << 32) | pos;
                    }
                    if (lane == 0) n_cand += __popcll(m);
                    if (kChk) {
                        uint32_t x = scored ? ck[t][i] * static_cast<uint32_t>(pos + 1) : 0u;
                        for (int o = 16; o > 0; o >>= 1) x += static_cast<uint32_t>(__shfl_xor(static_cast<int>(x), o));
                        if (n == 0 && row < np) atomicAdd(&a.chk[q], x);
                    }
                }
                if (lane == 0) {   // vectors of this tile within the list x pairs
                    const uint64_t at = static_cast<uint64_t>(bj) * 32;
                    n_pairs += (cnt > at ? min<uint64_t>(32, cnt - at) : 0) * np;
                }
            }
        }
        if (tid == 0) ++n_items;
        __syncthreads();
    }
    if (lane == 0 && a.stats) {
        atomicAdd(&a.stats[0], n_pairs);
        atomicAdd(&a.stats[1], n_cand);
        if (tid == 0) atomicAdd(&a.stats[5], n_items);
    }
}

hipError_t LaunchM8MfmaScan(hipStream_t st, const M8BatchArgs& a, uint32_t grid) {
    if (a.v.nrun < 1 || a.v.nrun > kM8BatchRuns || (a.v.plane_cols & 7u) != 0) return hipErrorInvalidValue;
    if (a.chk) hipLaunchKernelGGL(M8MfmaScanKernel<true>, dim3(grid), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(M8MfmaScanKernel<false>, dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Refine tasks: ceil(candidates / kM8RefineChunk) per query (none for a query
// whose list overflowed), exclusive offsets in toff; one workgroup of 1024
// threads, each owning a contiguous run of queries.

__global__ __launch_bounds__(1024) void M8TaskScanKernel(M8BatchArgs a) {
    __shared__ uint32_t sp[1024];
    const uint64_t nq = a.nq;
    const uint64_t per = (nq + 1023) / 1024;
    const uint64_t q0 = min<uint64_t>(nq, threadIdx.x * per), q1 = min<uint64_t>(nq, q0 + per);
    auto tasks = [&](uint64_t q) -> uint32_t {
        const uint32_t nc = a.ncand[q];
        return nc > a.cap ? 0u : (nc + kM8RefineChunk - 1) / kM8RefineChunk;
    };
    uint32_t tot = 0;
    for (uint64_t q = q0; q < q1; ++q) tot += tasks(q);
    sp[threadIdx.x] = tot;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint32_t x = threadIdx.x >= o ? sp[threadIdx.x - o] : 0u;
        __syncthreads();
        sp[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t off = sp[threadIdx.x] - tot;
    for (uint64_t q = q0; q < q1; ++q) {
        a.toff[q] = off;
        off += tasks(q);
    }
    if (threadIdx.x == 1023) {
        a.toff[nq] = sp[1023];
        a.counters[3] = 0;
    }
}

hipError_t LaunchM8RefineTasks(hipStream_t st, const M8BatchArgs& a) {
    hipLaunchKernelGGL(M8TaskScanKernel, dim3(1), dim3(1024), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Refine.  A workgroup of 256 threads takes one task (up to 64 candidates of
// one query); 16 groups of 16 lanes each take four of them, one at a time.
// LDS: the query's rotated vector, its u8 LUT, the runs' level tables.

struct M8RefLayout {
    uint32_t q, lut, lev, total;
    __host__ __device__ M8RefLayout(uint32_t P, uint32_t PC, uint32_t LV) {
        q = 0;
        lut = P * 4;
        lev = (lut + PC * 32 + 15) & ~15u;
        total = (lev + LV * 4 + 15) & ~15u;
    }
};

__global__ __launch_bounds__(256) void M8RefineKernel(M8BatchArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const M8View& v = a.v;
    const uint32_t P = static_cast<uint32_t>(v.padded), PC = v.plane_cols, R = v.nrun, k = a.k, maxp = a.maxp;
    const M8RefLayout L(P, PC, a.lev_floats);
    float* sq = reinterpret_cast<float*>(smem + L.q);
    uint8_t* slut = smem + L.lut;
    float* slev = reinterpret_cast<float*>(smem + L.lev);
    __shared__ uint32_t s_task, s_q;
    const uint32_t tid = threadIdx.x, g = tid >> 4, j = tid & 15;
    {   // level tables of the runs at rates <= kM8LdsLevRate, run after run
        uint32_t o = 0;
        for (uint32_t r = 0; r < R; ++r) {
            if (v.runs[r].rate > kM8LdsLevRate) continue;
            const uint32_t nq2 = 2u << v.runs[r].rate;
            for (uint32_t e = tid; e < nq2; e += 256) slev[o + e] = v.runs[r].lev[e];
            o += nq2;
        }
    }
    const float finf = __builtin_inff();
    const double dinf = __builtin_inf();
    const double dz = a.z;
    const uint32_t ntask = Uniform(a.toff[a.nq]);
    uint32_t qcur = 0xFFFFFFFFu;
    unsigned long long n_ref = 0, n_surv = 0, n_bits = 0;
    for (;;) {
        if (tid == 0) {
            const uint32_t t = atomicAdd(&a.counters[3], 1u);
            s_task = t;
            if (t < ntask) {   // its query: the last q with toff[q] <= t
                uint64_t lo = 0, hi = a.nq;
                while (hi - lo > 1) {
                    const uint64_t mid = (lo + hi) >> 1;
                    if (a.toff[mid] <= t) lo = mid;
                    else hi = mid;
                }
                s_q = static_cast<uint32_t>(lo);
            }
        }
        __syncthreads();
        const uint32_t task = Uniform(s_task);
        if (task >= ntask) break;
        const uint32_t q = Uniform(s_q);
        if (q != qcur) {
            for (uint32_t d = tid; d < P; d += 256) sq[d] = a.rotq[static_cast<uint64_t>(q) * P + d];
            for (uint32_t e = tid; e < PC * 8; e += 256)
                reinterpret_cast<uint32_t*>(slut)[e] = reinterpret_cast<const uint32_t*>(a.lut + static_cast<uint64_t>(q) * PC * 32)[e];
            qcur = q;
        }
        __syncthreads();
        const uint32_t nc = a.ncand[q];
        const uint32_t c0 = (task - a.toff[q]) * kM8RefineChunk;
        const uint32_t c1 = min(nc, c0 + kM8RefineChunk);
        const float tau = a.first_d[static_cast<uint64_t>(q) * k + k - 1];
        const double taud = tau < finf ? static_cast<double>(tau) : dinf;
        const float sqk = tau < finf ? sqrtf(fmaxf(0.0F, tau)) : finf;
        const double sqdk = taud < dinf ? sqrt(fmax(taud, 0.0)) : dinf;
        const float lut_delta = a.qc[static_cast<uint64_t>(q) * 4 + 1];
        const double* rcq = a.rc + static_cast<uint64_t>(q) * R * 4;
        for (uint32_t ci = c0 + g; ci < c0 + 4 * 16 && ci - g < c1; ci += 16) {
            const bool has = ci < c1;
            const uint64_t e = has ? a.cand[static_cast<uint64_t>(q) * a.cap + ci] : 0ull;
            const uint32_t p = static_cast<uint32_t>(e >> 32), pos = static_cast<uint32_t>(e);
            const uint32_t c = has ? a.probe[static_cast<uint64_t>(q) * maxp + p] : 0u;
            const uint64_t lb = v.list_begin[c], cnt = v.list_begin[c + 1] - lb;
            const uint32_t bj = pos >> 5, l = pos & 31u;
            const uint32_t cntb = static_cast<uint32_t>(min<uint64_t>(32, cnt - static_cast<uint64_t>(bj) * 32));
            const uint32_t covered = cntb >= 8 ? (cntb & ~7u) : 0u;
            const uint8_t* rec = v.plane + (v.batch_begin[c] + bj) * v.rec;
            const uint8_t* prow = rec + static_cast<uint64_t>(l) * PC;
            const uint8_t* banks = rec + 32ull * PC;
            // integer coarse sums of every run: lane j sums 8-column chunks j, j + 16, ...
            int32_t part[kM8BatchRuns] = {0, 0, 0, 0};
            if (has) {
                for (uint32_t ch = j; ch < PC / 8; ch += 16) {
                    const uint2 w = *reinterpret_cast<const uint2*>(prow + ch * 8);
                    int32_t s = 0;
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        const uint32_t byte = ((b < 4 ? w.x : w.y) >> (8 * (b & 3))) & 0xFFu;
                        const uint8_t* lc = slut + (ch * 8 + b) * 32;
                        s += lc[byte & 15u] + lc[16 + (byte >> 4)];
                    }
#pragma unroll
                    for (uint32_t r = 0; r < kM8BatchRuns; ++r)
                        if (r < R && ch * 8 >= v.runs[r].blk0 && ch * 8 < v.runs[r].blk0 + v.runs[r].nblk) part[r] += s;
                }
            }
#pragma unroll
            for (uint32_t r = 0; r < kM8BatchRuns; ++r)
                for (int o = 8; o > 0; o >>= 1) part[r] += __shfl_xor(part[r], o, 16);
            // the coarse terms in the reference's arithmetic, against tau
            double ssc[kM8BatchRuns], str[kM8BatchRuns], pool = 0.0, qadd = 0.0;
            const double d2 = has ? static_cast<double>(a.pdist[static_cast<uint64_t>(q) * maxp + p]) : 0.0;
            const double* sqd2 = a.sqd2 + (static_cast<uint64_t>(q) * maxp + p) * R;
            if (has) {
                if (l < covered) {   // the AVX2 filter's fp32 values
                    float pl = 0.0F, qa = 0.0F, wpad = 0.0F;
                    for (uint32_t r = 0; r < R; ++r) {
                        const uint8_t* bp = banks + r * kM8BankBytes;
                        const float* sc = reinterpret_cast<const float*>(bp + 128);
                        const float s_c = fmaf(static_cast<float>(part[r]), lut_delta, static_cast<float>(rcq[r * 4 + 1]));
                        const float wq = fmaf(static_cast<float>(bp[32 + l]), sc[3], sc[2]);
                        const float pq = fmaf(static_cast<float>(bp[64 + l]), sc[5], sc[4]);
                        const float ra = static_cast<float>(sqd2[r]);
                        const float rh = sqk < ra ? sqk : ra;
                        const float t = pq * rh;
                        ssc[r] = s_c;
                        str[r] = t;
                        pl = fmaf(t, t, pl);
                        qa = fmaf(fabsf(wq), static_cast<float>(rcq[r * 4 + 2]), qa);
                        wpad = wpad + static_cast<float>(0.5 * static_cast<double>(sc[3]) * (rcq[r * 4 + 3] + rcq[r * 4 + 2]));
                    }
                    pool = pl;
                    qadd = qa + wpad;
                } else {   // the fp64 twin for the lanes the filter does not cover
                    for (uint32_t r = 0; r < R; ++r) {
                        const uint8_t* bp = banks + r * kM8BankBytes;
                        const float* scl = reinterpret_cast<const float*>(bp + 128);
                        const double scr = fma(static_cast<double>(lut_delta), static_cast<double>(part[r]), rcq[r * 4 + 1]);
                        const double wq = fma(static_cast<double>(bp[32 + l]), static_cast<double>(scl[3]), static_cast<double>(scl[2]));
                        const double pq = fma(static_cast<double>(bp[64 + l]), static_cast<double>(scl[5]), static_cast<double>(scl[4]));
                        const double sa = sqd2[r];
                        const double rh = sqdk < sa ? sqdk : sa;
                        const double tr = pq * rh;
                        ssc[r] = scr;
                        str[r] = tr;
                        pool = fma(tr, tr, pool);
                        qadd = fma(fabs(wq), rcq[r * 4 + 2], qadd);
                        qadd = fma(0.5 * static_cast<double>(scl[3]), rcq[r * 4 + 3] + rcq[r * 4 + 2], qadd);
                    }
                }
            }
            // the cascade's run order for a group of one: decreasing t^2 / nblk, stable
            uint32_t ord[kM8BatchRuns] = {0, 1, 2, 3};
            if (R > 1) {
                double mean[kM8BatchRuns];
                for (uint32_t r = 0; r < R; ++r)
                    mean[r] = fma(str[r], str[r], 0.0) / static_cast<double>(v.runs[r].nblk > 1 ? v.runs[r].nblk : 1);
                for (uint32_t i = 1; i < R; ++i) {
                    const uint32_t x = ord[i];
                    uint32_t m = i;
                    while (m > 0 && mean[x] > mean[ord[m - 1]]) {
                        ord[m] = ord[m - 1];
                        --m;
                    }
                    ord[m] = x;
                }
            }
            const float* f4 = reinterpret_cast<const float*>(v.refine + (lb + pos) * v.rstride);
            double dist = 0.0;
            if (has) {
                if (R == 1) {
                    dist = d2 + static_cast<double>(f4[1]);
                } else {
                    double dcq = d2;
                    for (uint32_t r = 0; r < R; ++r) {
                        const double coef = f4[r * 4], h2 = f4[r * 4 + 1];
                        const double gn = f4[r * 4 + 2], vp = f4[r * 4 + 3];
                        dcq += fma(-(2.0 * coef), fma(gn, ssc[r], vp), h2);
                    }
                    dist = dcq;
                }
            }
            double rem2 = pool;
            bool live = has;
            uint64_t bits = 0;
            const uint32_t half = j >> 3, cmp = j & 7u;
            for (uint32_t oi = 0; oi < R; ++oi) {
                if (!live) break;
                const uint32_t r = ord[oi];
                const M8Run& ru = v.runs[r];
                uint32_t lo_off = 0;
                for (uint32_t r2 = 0; r2 < r; ++r2)
                    if (v.runs[r2].rate <= kM8LdsLevRate) lo_off += 2u << v.runs[r2].rate;
                const float* lev = ru.rate <= kM8LdsLevRate ? slev + lo_off : ru.lev;
                const uint8_t* fr = v.refine + (lb + pos) * v.rstride + ru.ref_off;
                const float acc = M8HalfChain(ru.rate, prow + ru.blk0, fr, lev, sq + ru.dim_off + cmp, ru.nblk, half, cmp);
                const float other = __shfl_xor(acc, 8, 16);
                const double se = Hsum8Lanes(acc + other);
                const double coef = f4[r * 4];
                if (R == 1) {
                    dist = fma(-(2.0 * coef), se, dist);
                } else {
                    const double gn = f4[r * 4 + 2], vp = f4[r * 4 + 3];
                    dist = fma(2.0 * coef, fma(gn, ssc[r], vp) - se, dist);
                }
                const double t = str[r];
                rem2 = fma(-t, t, rem2);
                bits += static_cast<uint64_t>(ru.rate) * ru.len;
                if (a.cascade && fma(-dz, sqrt(fmax(rem2, 0.0)), dist) - qadd >= taud) live = false;
            }
            if (has && j == 0) {
                ++n_ref;
                n_bits += bits;
                const float d = static_cast<float>(dist);
                if (live && (d < tau || !(tau < finf))) {
                    ++n_surv;
                    const uint32_t slot = atomicAdd(&a.nsurv[q], 1u);
                    if (slot < a.scap)
                        a.surv[static_cast<uint64_t>(q) * a.scap + slot] =
                            (static_cast<uint64_t>(M8Key(d)) << 32) | v.ids[lb + pos];
                }
            }
        }
        __syncthreads();
    }
    if (j == 0 && a.stats && (n_ref | n_surv)) {
        atomicAdd(&a.stats[2], n_ref);
        atomicAdd(&a.stats[3], n_surv);
        atomicAdd(&a.stats[4], n_bits);
    }
}

size_t M8RefineLds(const M8View& v, uint32_t lev_floats) {
    return M8RefLayout(static_cast<uint32_t>(v.padded), v.plane_cols, lev_floats).total;
}

hipError_t LaunchM8Refine(hipStream_t st, const M8BatchArgs& a, uint32_t grid) {
    const size_t lds = M8RefineLds(a.v, a.lev_floats);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&M8RefineKernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(M8RefineKernel, dim3(grid), dim3(256), lds, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Top-k of each query: its threshold-pass list and its survivors, sorted by
// (distance, id) with a bitonic sort in LDS of the best k so far plus the
// next chunk.  One wave per query.

constexpr uint32_t kM8SelChunk = 1024;

__global__ __launch_bounds__(64) void M8SelectKernel(M8BatchArgs a) {
    __shared__ uint64_t buf[kM8SelChunk];
    const uint32_t lane = threadIdx.x, k = a.k;
    const float finf = __builtin_inff();
    for (uint64_t q = blockIdx.x; q < a.nq; q += gridDim.x) {
        const uint32_t nc = Uniform(a.ncand[q]), ns = Uniform(a.nsurv[q]);
        if (nc > a.cap || ns > a.scap) {   // searched again by the plain kernel
            if (lane == 0) a.overflow[atomicAdd(&a.counters[2], 1u)] = static_cast<uint32_t>(q);
            continue;
        }
        const uint64_t* sl = a.surv + q * a.scap;
        // the threshold-pass list first (it is already sorted), then the survivors
        const uint32_t n = k + ns;
        uint32_t ch = 64;
        while (ch < kM8SelChunk && ch < n) ch <<= 1;
        const uint32_t room = ch - k;
        for (uint32_t i = lane; i < k; i += 64) {
            const float d = a.first_d[q * k + i];
            buf[i] = d < finf ? (static_cast<uint64_t>(M8Key(d)) << 32) | a.first_ids[q * k + i] : ~0ull;
        }
        uint32_t s = 0;
        do {
            const uint32_t take = min(room, ns - s);
            for (uint32_t i = lane; i < room; i += 64) buf[k + i] = i < take ? sl[s + i] : ~0ull;
            __syncthreads();
            for (uint32_t size = 2; size <= ch; size <<= 1)
                for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                    for (uint32_t i = lane; i < ch / 2; i += 64) {
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
        } while (s < ns);
        for (uint32_t i = lane; i < k; i += 64) {
            const uint64_t key = buf[i];
            const bool ok = key != ~0ull;
            a.out_ids[q * k + i] = ok ? static_cast<uint32_t>(key) : 0xFFFFFFFFu;
            a.out_d[q * k + i] = ok ? M8KeyDist(static_cast<uint32_t>(key >> 32)) : finf;
        }
        __syncthreads();
    }
}

hipError_t LaunchM8Select(hipStream_t st, const M8BatchArgs& a) {
    if (a.nq == 0) return hipSuccess;
    if (a.k == 0 || a.k > kM8SelMaxK) return hipErrorInvalidValue;
    const uint32_t grid = static_cast<uint32_t>(a.nq < 262144 ? a.nq : 262144);
    hipLaunchKernelGGL(M8SelectKernel, dim3(grid), dim3(64), 0, st, a);
    return hipGetLastError();
}

}  // namespace e8qdev
