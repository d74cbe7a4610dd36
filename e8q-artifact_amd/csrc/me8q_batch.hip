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
//   1. threshold pass: the exact distances of each query's first `budget`
//      32-vector batches in probe order (no pruning); their top-k is the
//      query's starting list and tau_q, its k-th distance, the threshold;
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
//      dc - (z sqrt(pool) + qa) against tau_q.  The kernel runs twice: the
//      first pass counts each query's passing vectors, a scan turns the
//      counts into offsets in one candidate pool, the second pass (the same
//      arithmetic, so the same vectors pass) writes them there;
//   3. refine: per candidate the coarse terms again in the reference's own
//      arithmetic (fp32 for lanes its AVX2 filter covers, fp64 for the rest),
//      the run order of the reference's cascade for a group of one
//      candidate (decreasing t^2 / nblk), the exact run scores (the plain
//      kernel's even/odd block chains and AVX horizontal sums) and the
//      cascade's death test against tau_q; the survivors' exact distances
//      below tau_q are written over the task's own pool slots;
//   4. select: one wave per query sorts its threshold-pass list and its
//      survivors by (distance, id) and writes the top-k.
//
// With z = +inf nothing is pruned in either form and the result is the
// exact estimator top-k over the probed lists, as the plain form's.  With a
// finite z the bounds are the reference's with tau_q in place of the live
// threshold, which is at least the reference's threshold after the same
// vectors.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "e8q_scan_dev.h"
#include "me8q_device.h"
#include "me8q_kernels.h"

namespace e8qdev {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Order-preserving 32-bit key of a float distance (negatives below positives).
__device__ __forceinline__ uint32_t M8Key(float d) {
    const uint32_t u = __float_as_uint(d);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float M8KeyDist(uint32_t m) {
    return __uint_as_float((m & 0x80000000u) ? (m & 0x7FFFFFFFu) : ~m);
}

// A record's second word: the tau0 bin (< 64) in bits 26..31, the lower bound
// in bits 0..25 as its float's sign and top 25 magnitude bits, the magnitude
// rounded toward -inf (down for a positive bound, up for a negative one), so
// the value read back is never above the bound.
__device__ __forceinline__ uint32_t M8RecPack(float dm, uint32_t bin) {
    const uint32_t u = __float_as_uint(dm), m = u & 0x7FFFFFFFu, s = u >> 31;
    return (bin << 26) | (s << 25) | min((m + (63u & (0u - s))) >> 6, 0x1FFFFFFu);
}
__device__ __forceinline__ float M8RecBound(uint32_t w) {
    return __uint_as_float(((w >> 25) & 1u) << 31 | (w & 0x1FFFFFFu) << 6);
}

// ---------------------------------------------------------------------------
// The contraction.  256 threads = four waves; wave w scores the item's 32
// pairs against batches b0 + w and b0 + w + 4.  Accumulator register i of
// lane l holds row (pair) 8 (i / 4) + 4 (l / 32) + i % 4, column (vector)
// l % 32.

constexpr int kM8Tiles = 2;
constexpr uint32_t kM8StageBytes = 8 * 1024;   // one stage of the query operand: 8 columns x 2 x 32 x 16 B

// s_waitcnt vmcnt(n) for a wave-uniform n <= 4: this wave's vector memory
// operations, all but the n most recently issued, have completed.
__device__ __forceinline__ void WaitVmcnt(uint32_t n) {
    switch (n) {
        case 0: asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); break;
        case 1: asm volatile("s_waitcnt vmcnt(1)" ::: "memory"); break;
        case 2: asm volatile("s_waitcnt vmcnt(2)" ::: "memory"); break;
        case 3: asm volatile("s_waitcnt vmcnt(3)" ::: "memory"); break;
        default: asm volatile("s_waitcnt vmcnt(4)" ::: "memory"); break;
    }
}

// Rounding margin of the coarse bound's regrouped evaluation: 2^-18 of the
// sum of the magnitudes of its terms (d2, the bq, each |wq| times a bound on
// its run's coarse score, the slack), on top of 2^-12 (|dc| + |slack|).  This
// kernel's fp32 evaluation and the reference's per-run one (fp32 or fp64)
// each stay within a few tens of 2^-24 of that sum, so every vector the
// reference's bound keeps at a threshold is kept here.
constexpr float kM8MarginAbs = 3.814697265625e-06F;   // 2^-18
constexpr float kM8MarginRel = 2.4414062e-4F;         // 2^-12
static_assert(kM8BatchRuns == 4, "the per-pair run tables are float4");

// The coarse bound per (pair, vector): dc = d2 + sum_r (bq_r - wq_r s_r) with
// s_r = S_r delta + lut_min_x_r the run's coarse score, slack = z sqrt(sum_r
// pq_r^2 rh_r^2) + sum_r (|wq_r| qerr_r + s3_r m_r) (me8q_prune_avx2.cpp:
// 9-98).  With S_r = acc_r + 256 nblk_r (acc_r the int8 contraction's sum),
// s_r = acc_r delta + x_r, so sum_r wq_r s_r = delta T + sum_r wq_r x_r where
// T = sum_r wq_r acc_r: T is the only per-(pair, vector) quantity the column
// loop carries; the other sums over runs of a per-pair factor times a
// per-vector factor are themselves small contractions (K = runs), done at the
// end of the item with v_mfma_f32_32x32x2_f32 (two runs per instruction, lane
// l supplying run 2j + l / 32 of pair l % 32 and of vector l % 32), whose
// 32 x 32 result has the int8 MFMA's accumulator layout.
// Whether the count pass's records all fit their reserve: the candidates are
// then placed from the records, else written by a second contraction.
__device__ __forceinline__ bool M8RecordsPlaced(const M8BatchArgs& a) {
    return a.rec != nullptr && *G(a.rcount) <= a.rec_cap;
}

template <bool kChk, bool kWrite>
__global__ __launch_bounds__(256, 2) void M8MfmaScanKernel(M8BatchArgs a) {
    const M8View& v = a.v;
    __shared__ uint32_t s_it;
    __shared__ uint32_t s_q[kM8BatchQTile], s_p[kM8BatchQTile];
    __shared__ float s_d2[kM8BatchQTile];
    // per pair: {delta, tau0 + 2^-12 |tau0|, 64 / tau0 (0 when tau0 is +inf), as bits:
    // the first batch of its probe that the threshold pass has not scored (~0 past np)}
    __shared__ float4 s_rowf[kM8BatchQTile];
    // per (run, pair): {x = lut_min_x + 256 nblk delta, qerr, (score_max + qerr) / 2,
    // min(sqrt(tau0), sqrt(d2_r))^2}; zero for runs past the index's
    __shared__ float4 s_rv[kM8BatchRuns][kM8BatchQTile];
    // per pair, per run: |lut_min_x| + 512 nblk |delta| >= |s_r| (either form)
    __shared__ float4 s_rm[kM8BatchQTile];
    __shared__ __align__(16) uint32_t s_oh[16 * 4];
    __shared__ uint32_t s_hist[kChk || kWrite ? 1 : kM8BatchQTile * kM8TauBins];   // count pass
    __shared__ uint32_t s_abin[kM8BatchQTile], s_na[kM8BatchQTile];                   // write pass
    __shared__ uint64_t s_bal[4][kM8Tiles][16][2];       // per wave, tile, accumulator row: kept / set-A ballots
    __shared__ uint32_t s_off[4][kM8Tiles][kM8BatchQTile][2];   // a tile's first append of each pair (A, B)
    __shared__ uint32_t s_base[kM8BatchQTile][2];        // each pair's first slot of the item (A, B)
    __shared__ uint32_t s_rs[kM8BatchQTile];             // count pass: each pair's first record of the item
    __shared__ uint32_t s_minbin[kM8BatchQTile];         // count pass: each pair's smallest tau0 bin
    __shared__ unsigned long long s_rb;                  // count pass: the item's first record
    // count pass: each (wave, tile, accumulator row, lane)'s record word
    __shared__ uint32_t s_rw[kWrite ? 1 : 4][kM8Tiles][16][64];
    // The query operand, eight plane columns per stage, a ring of three
    // stages: [stage][column][half][pair][16 bytes] of the int8 LUT.
    __shared__ __align__(16) uint8_t s_a[3 * kM8StageBytes];
    // The write pass runs only when the count pass's records did not fit
    // their reserve (else they are placed instead); the choice is made here,
    // on the device.
    if (kWrite && M8RecordsPlaced(a)) return;
    const uint32_t tid = threadIdx.x, wave = Uniform(tid >> 6), lane = tid & 63;
    const uint32_t h = lane >> 5, n = lane & 31;
    if (tid < 16)
        for (uint32_t j = 0; j < 4; ++j) s_oh[tid * 4 + j] = (tid >> 2) == j ? (1u << (8 * (tid & 3))) : 0u;
    const uint8_t* s_ohb = reinterpret_cast<const uint8_t*>(s_oh);
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
            const float tau = G(a.first_d)[static_cast<uint64_t>(q) * k + k - 1];   // tau0
            const uint64_t cv = G(a.cover)[q];
            const uint32_t ps = static_cast<uint32_t>(cv >> 32), bs = static_cast<uint32_t>(cv);
            s_q[tid] = q;
            s_p[tid] = p;
            if (kWrite) {
                s_abin[tid] = G(a.abin)[q];
                s_na[tid] = G(a.na)[q];
            }
            s_d2[tid] = G(a.pdist)[static_cast<uint64_t>(q) * maxp + p];
            const float dl = G(a.qc)[static_cast<uint64_t>(q) * 4 + 1];
            const float sk = tau < finf ? sqrtf(fmaxf(0.0F, tau)) : finf;
            // a vector is new to the query unless the threshold pass scored it
            // (probe positions before ps, and ps's batches before bs)
            const uint32_t first = tid >= np ? ~0u : p > ps ? 0u : p == ps ? bs : ~0u;
            s_rowf[tid] = make_float4(dl, tau + kM8MarginRel * fabsf(tau), tau < finf ? 64.0F / tau : 0.0F,
                                      __uint_as_float(first));
            float rm[kM8BatchRuns];
#pragma unroll
            for (uint32_t r = 0; r < kM8BatchRuns; ++r) {
                float4 rv = make_float4(0.0F, 0.0F, 0.0F, 0.0F);
                rm[r] = 0.0F;
                if (r < R) {
                    const E8Q_GLOBAL double* o = G(a.rc) + (static_cast<uint64_t>(q) * R + r) * 4;
                    const float ra = static_cast<float>(G(a.sqd2)[(static_cast<uint64_t>(q) * maxp + p) * R + r]);
                    const float rh = sk < ra ? sk : ra;
                    const float nbf = static_cast<float>(v.runs[r].nblk), lmx = static_cast<float>(o[1]);
                    rv = make_float4(fmaf(256.0F * nbf, dl, lmx), static_cast<float>(o[2]),
                                     static_cast<float>(0.5 * (o[3] + o[2])), rh * rh);
                    rm[r] = fmaf(512.0F * nbf, fabsf(dl), fabsf(lmx));
                }
                s_rv[r][tid] = rv;
            }
            s_rm[tid] = make_float4(rm[0], rm[1], rm[2], rm[3]);
        }
        if (!kChk && !kWrite)
            for (uint32_t e = tid; e < kM8BatchQTile * kM8TauBins; e += 256) s_hist[e] = 0u;
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
        // The per-vector factors of the bound (each tile's batch, this lane's
        // vector), from the scan banks (me8q.hpp:175-184), computed before the
        // column loop; zero for runs past the index's and tiles past nt.
        float wq[kM8Tiles][kM8BatchRuns], pq2[kM8Tiles][kM8BatchRuns], s3[kM8Tiles][kM8BatchRuns];
        float bqs[kM8Tiles], bqa[kM8Tiles];
#pragma unroll
        for (int t = 0; t < kM8Tiles; ++t) {
            bqs[t] = 0.0F;
            bqa[t] = 0.0F;
            const uint32_t bj = b0 + wave + 4 * t;
#pragma unroll
            for (uint32_t r = 0; r < kM8BatchRuns; ++r) {
                wq[t][r] = 0.0F;
                pq2[t][r] = 0.0F;
                s3[t][r] = 0.0F;
                if (t < static_cast<int>(nt) && r < R) {
                    const E8Q_GLOBAL uint8_t* bp = G(v.plane) + (bat0 + bj) * v.rec + 32ull * PC + r * kM8BankBytes;
                    const E8Q_GLOBAL float* sc = reinterpret_cast<const E8Q_GLOBAL float*>(bp + 128);
                    const float bq = fmaf(static_cast<float>(bp[n]), sc[1], sc[0]);
                    const float pq = fmaf(static_cast<float>(bp[64 + n]), sc[5], sc[4]);
                    wq[t][r] = fmaf(static_cast<float>(bp[32 + n]), sc[3], sc[2]);
                    pq2[t][r] = pq * pq;
                    s3[t][r] = sc[3];
                    bqs[t] += bq;
                    bqa[t] += fabsf(bq);
                }
            }
        }
        // The query operand goes straight to LDS (global_load_lds_dwordx4, 16
        // bytes per lane, no registers), two stages ahead: wave w copies columns
        // 2w and 2w + 1 of a stage, lane l pair l & 31, half l >> 5, so a lane
        // (pair n, half h) of the MFMA reads its 16 bytes at (h * 32 + n) * 16 of
        // a column (consecutive lanes, consecutive banks).
        const E8Q_GLOBAL uint8_t* lsrc =
            G(a.lut8) + static_cast<uint64_t>(s_q[lane & 31]) * PC * 32 + (lane >> 5) * 16;
        auto stage = [&](uint32_t sb) {
            uint8_t* dst = s_a + (sb % 3) * kM8StageBytes + (2 * wave) * 1024;
            __builtin_amdgcn_global_load_lds(lsrc + (sb * 8 + 2 * wave) * 32, dst, 16, 0, 0);
            __builtin_amdgcn_global_load_lds(lsrc + (sb * 8 + 2 * wave + 1) * 32, dst + 1024, 16, 0, 0);
        };
        i32x16 acc[kM8Tiles];
        float T[kM8Tiles][16];
        uint32_t ck[kM8Tiles][16];
#pragma unroll
        for (int t = 0; t < kM8Tiles; ++t) {
            acc[t] = i32x16{};
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                T[t][i] = 0.0F;
                ck[t][i] = 0u;
            }
        }
        // Eight plane columns of NT tiles: per column one A read (shared by the
        // tiles) and per tile the one-hot B row of the vector's nibble, read
        // one column ahead of its MFMA.  A plane byte's nibble for half h,
        // times 16 (the B row's LDS offset), is byte e of the word's masked
        // nibbles shifted up by four.
        auto block = [&](auto ntc, uint32_t cb, const uint64_t* w) {
            constexpr int NT = decltype(ntc)::value;
            const uint8_t* arow = s_a + (cb % 3) * kM8StageBytes + (h * 32 + n) * 16;
            uint32_t lo[NT], hi[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const uint64_t x = ((w[t] >> (4 * h)) & 0x0F0F0F0F0F0F0F0Full) << 4;
                lo[t] = static_cast<uint32_t>(x);
                hi[t] = static_cast<uint32_t>(x >> 32);
            }
            auto brow = [&](int t, int e) -> i32x4 {   // byte e & 3 of the word, zero-extended (v_perm_b32)
                const uint32_t off = __builtin_amdgcn_perm(0u, e < 4 ? lo[t] : hi[t], 0x0C0C0C00u | (e & 3));
                return *reinterpret_cast<const i32x4*>(s_ohb + off);
            };
            i32x4 av = *reinterpret_cast<const i32x4*>(arow);
            i32x4 bv[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) bv[t] = brow(t, 0);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                i32x4 avn = av, bvn[NT];
#pragma unroll
                for (int t = 0; t < NT; ++t) bvn[t] = bv[t];
                if (e < 7) {
                    avn = *reinterpret_cast<const i32x4*>(arow + (e + 1) * 1024);
#pragma unroll
                    for (int t = 0; t < NT; ++t) bvn[t] = brow(t, e + 1);
                }
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bv[t], acc[t], 0, 0, 0);
                av = avn;
#pragma unroll
                for (int t = 0; t < NT; ++t) bv[t] = bvn[t];
            }
            // issue order: column 0's reads, then per column the next column's
            // offsets and reads before this column's MFMAs, so an MFMA waits
            // only for reads issued one column earlier
            __builtin_amdgcn_sched_group_barrier(0x100, NT + 1, 0);
#pragma unroll
            for (int e = 0; e < 7; ++e) {
                __builtin_amdgcn_sched_group_barrier(0x002, 2 * NT, 0);
                __builtin_amdgcn_sched_group_barrier(0x100, NT + 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, NT, 0);
            }
            __builtin_amdgcn_sched_group_barrier(0x008, NT, 0);
        };
        uint32_t r = 0, rend = R > 0 ? Uniform(v.runs[0].blk0 + v.runs[0].nblk) : PC;
        const uint32_t ncb = PC / 8;
        // prologue: stages 0 and 1, then this wave's plane words of block 0
        stage(0);
        if (ncb > 1) stage(1);
        uint64_t wcur[kM8Tiles] = {0ull, 0ull}, wnext[kM8Tiles] = {0ull, 0ull};
#pragma unroll
        for (int t = 0; t < kM8Tiles; ++t)
            if (t < static_cast<int>(nt)) wcur[t] = *reinterpret_cast<const E8Q_GLOBAL uint64_t*>(prow[t]);
        for (uint32_t cb = 0; cb < ncb; ++cb) {
            // Issued after stage cb's copies: this wave's plane words of block cb
            // (nt loads) and stage cb + 1's copies (2, when there is one); all
            // older operations are done once no more than those are in flight.
            WaitVmcnt(cb + 1 < ncb ? nt + 2 : 0);
            __syncthreads();   // every wave's copies of stage cb are in LDS
            if (nt > 0 && cb + 1 < ncb) {
#pragma unroll
                for (int t = 0; t < kM8Tiles; ++t)
                    if (t < static_cast<int>(nt))
                        wnext[t] = *reinterpret_cast<const E8Q_GLOBAL uint64_t*>(prow[t] + (cb + 1) * 8);
            }
            if (cb + 2 < ncb) stage(cb + 2);   // its ring slot was last read before this barrier
            if (nt == 2) block(std::integral_constant<int, 2>{}, cb, wcur);
            else if (nt == 1) block(std::integral_constant<int, 1>{}, cb, wcur);
#pragma unroll
            for (int t = 0; t < kM8Tiles; ++t) wcur[t] = wnext[t];
            if ((cb + 1) * 8 == rend) {   // run r ends here: T += wq_r acc_r
                const int32_t shift = 256 * static_cast<int32_t>(Uniform(v.runs[r].nblk));
#pragma unroll
                for (int t = 0; t < kM8Tiles; ++t) {
                    const float wr = r == 0 ? wq[t][0] : r == 1 ? wq[t][1] : r == 2 ? wq[t][2] : wq[t][3];
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        T[t][i] = fmaf(wr, static_cast<float>(acc[t][i]), T[t][i]);
                        if (kChk) ck[t][i] += (r + 1) * static_cast<uint32_t>(acc[t][i] + shift);
                    }
                    acc[t] = i32x16{};
                }
                ++r;
                rend = r < R ? Uniform(v.runs[r].blk0 + v.runs[r].nblk) : PC;
            }
        }
        // The bound test of every (pair, vector) of the wave's tiles, ballots
        // to LDS; then one thread per pair adds up its appends over the item's
        // tiles (one atomic per query and item), and the write pass places each
        // vector at its query's base + the tile's offset + its rank in the ballot.
        const float z = a.z;
        // MFMA operands of the sums: this lane's pair row n and vector n, runs h
        // and 2 + h (a pair's factors are the same for both tiles)
        const float4 ra0 = s_rv[h][n], ra1 = s_rv[2 + h][n];
        const float4 rmv = s_rm[n];
        const float rm0 = h ? rmv.y : rmv.x, rm1 = h ? rmv.w : rmv.z, d2n = s_d2[n];
        struct Sums {
            f32x16 e, m, p, q;
        };
        // e = d2 + sum_r (bq_r - wq_r x_r); m = |d2| + sum_r (|bq_r| + |wq_r| rm_r);
        // p = sum_r pq_r^2 rh_r^2; q = sum_r (|wq_r| qerr_r + s3_r m_r)
        auto sums = [&](int t) -> Sums {
            const f32x16 zero = {};
            const float w0 = h ? wq[t][1] : wq[t][0], w1 = h ? wq[t][3] : wq[t][2];
            const float p0 = h ? pq2[t][1] : pq2[t][0], p1 = h ? pq2[t][3] : pq2[t][2];
            const float t0 = h ? s3[t][1] : s3[t][0], t1 = h ? s3[t][3] : s3[t][2];
            Sums o;
            o.e = __builtin_amdgcn_mfma_f32_32x32x2f32(h ? 1.0F : d2n, h ? bqs[t] : 1.0F, zero, 0, 0, 0);
            o.e = __builtin_amdgcn_mfma_f32_32x32x2f32(-ra0.x, w0, o.e, 0, 0, 0);
            o.e = __builtin_amdgcn_mfma_f32_32x32x2f32(-ra1.x, w1, o.e, 0, 0, 0);
            o.m = __builtin_amdgcn_mfma_f32_32x32x2f32(h ? 1.0F : fabsf(d2n), h ? bqa[t] : 1.0F, zero, 0, 0, 0);
            o.m = __builtin_amdgcn_mfma_f32_32x32x2f32(rm0, fabsf(w0), o.m, 0, 0, 0);
            o.m = __builtin_amdgcn_mfma_f32_32x32x2f32(rm1, fabsf(w1), o.m, 0, 0, 0);
            o.p = __builtin_amdgcn_mfma_f32_32x32x2f32(ra0.w, p0, zero, 0, 0, 0);
            o.p = __builtin_amdgcn_mfma_f32_32x32x2f32(ra1.w, p1, o.p, 0, 0, 0);
            o.q = __builtin_amdgcn_mfma_f32_32x32x2f32(ra0.y, fabsf(w0), zero, 0, 0, 0);
            o.q = __builtin_amdgcn_mfma_f32_32x32x2f32(ra0.z, t0, o.q, 0, 0, 0);
            o.q = __builtin_amdgcn_mfma_f32_32x32x2f32(ra1.y, fabsf(w1), o.q, 0, 0, 0);
            o.q = __builtin_amdgcn_mfma_f32_32x32x2f32(ra1.z, t1, o.q, 0, 0, 0);
            return o;
        };
        // The test of accumulator element i of tile t: dc = e - delta T, the
        // slack, the lower bound dm (less the margin) against tau0, the tau0 bin
        // of dc; scored (in the list, a pair of the item) and not scored by the
        // threshold pass.
        auto test = [&](int t, int i, const Sums& sm, float& dm, uint32_t& bin) -> bool {
            const uint32_t row = 8 * (i / 4) + 4 * h + (i % 4);
            const uint32_t bj = b0 + wave + 4 * t;
            const uint64_t pos = static_cast<uint64_t>(bj) * 32 + n;
            const float4 rf = s_rowf[row];
            const float dc = fmaf(-rf.x, T[t][i], sm.e[i]);
            // v_sqrt_f32 (within 1 ulp): the margin's 2^-18 |slack| covers it
            const float slack = fmaf(z, __builtin_amdgcn_sqrtf(sm.p[i]), sm.q[i]);
            dm = dc - slack - (kM8MarginRel * (fabsf(dc) + fabsf(slack)) + kM8MarginAbs * (sm.m[i] + fabsf(slack)));
            bin = static_cast<uint32_t>(fminf(fmaxf(dc, 0.0F) * rf.z, static_cast<float>(kM8TauBins - 1)));
            return bj >= __float_as_uint(rf.w) && pos < cnt && !(dm >= rf.y);
        };
        if (nt > 0) {
#pragma unroll
            for (int t = 0; t < kM8Tiles; ++t) {
                if (t >= static_cast<int>(nt)) break;
                const uint32_t bj = b0 + wave + 4 * t;
                const uint64_t pos = static_cast<uint64_t>(bj) * 32 + n;
                const Sums sm = sums(t);
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const uint32_t row = 8 * (i / 4) + 4 * h + (i % 4);
                    float dm;
                    uint32_t bin;
                    const bool kept = test(t, i, sm, dm, bin);
                    if (!kWrite) {
                        if (!kChk && kept) atomicAdd(&s_hist[row * kM8TauBins + bin], 1u);
                        s_rw[kWrite ? 0 : wave][t][i][lane] = M8RecPack(dm, bin);
                    }
                    const uint64_t mk = __ballot(kept);
                    const uint64_t ma = kWrite ? __ballot(kept && bin <= s_abin[row]) : 0ull;
                    if (lane == 0) {
                        s_bal[wave][t][i][0] = mk;
                        if (kWrite) s_bal[wave][t][i][1] = ma;
                        n_cand += __popcll(mk);
                    }
                    if (kChk) {
                        uint32_t x = pos < cnt && row < np ? ck[t][i] * static_cast<uint32_t>(pos + 1) : 0u;
                        for (int o = 16; o > 0; o >>= 1) x += static_cast<uint32_t>(__shfl_xor(static_cast<int>(x), o));
                        if (n == 0 && row < np) atomicAdd(&a.chk[s_q[row]], x);
                    }
                }
                if (lane == 0) {   // vectors of this tile within the list x pairs
                    const uint64_t at = static_cast<uint64_t>(bj) * 32;
                    n_pairs += (cnt > at ? min<uint64_t>(32, cnt - at) : 0) * np;
                }
            }
        }
        __syncthreads();
        uint32_t ta = 0, tb = 0;
        if (tid < np) {   // pair tid: its row's appends over the item's tiles
            const uint32_t m = tid, hh = (m / 4) & 1u;   // accumulator row 8 (i / 4) + 4 h + i % 4
            const uint32_t i = (m / 8) * 4 + (m % 4);
            for (uint32_t w = 0; w < 4; ++w)
                for (uint32_t t = 0; t < static_cast<uint32_t>(kM8Tiles); ++t) {
                    if (w + 4 * t >= nb) continue;
                    const uint32_t mk = static_cast<uint32_t>(s_bal[w][t][i][0] >> (32 * hh));
                    const uint32_t ma = kWrite ? static_cast<uint32_t>(s_bal[w][t][i][1] >> (32 * hh)) : 0u;
                    s_off[w][t][m][0] = ta;
                    s_off[w][t][m][1] = tb;
                    ta += __popc(mk & ma);
                    tb += __popc(mk & ~ma);
                }
            const uint32_t q = s_q[m];
            if (!kWrite) {
                if (ta + tb) atomicAdd(&a.ncand[q], ta + tb);
            } else {
                s_base[m][0] = ta ? atomicAdd(&a.fill_a[q], ta) : 0u;
                s_base[m][1] = tb ? atomicAdd(&a.fill_b[q], tb) : 0u;
            }
        }
        if (!kWrite && a.rec != nullptr) {
            // Records: the item's block of the record array (one atomic), each
            // pair's records contiguous in it, a pair's tiles in (wave, tile)
            // order, the words staged in LDS by the test above.
            if (wave == 0) {   // each pair's smallest tau0 bin: its lowest non-empty histogram bin
                uint32_t mb = ~0u;
                if (!kChk)
                    for (uint32_t r = 0; r < kM8BatchQTile; ++r) {
                        const uint64_t nz = __ballot(s_hist[r * kM8TauBins + lane] != 0u);   // lane = bin
                        if (lane == r) mb = nz ? static_cast<uint32_t>(__builtin_ctzll(nz)) : ~0u;
                    }
                else
                    mb = 0u;   // no histogram in the checksum pass: no unit is skipped
                if (lane < kM8BatchQTile) s_minbin[lane] = mb;
            }
            if (tid < kM8BatchQTile) {   // wave 0, lanes 0..31 = the pairs (count 0 past np)
                const uint32_t tot = ta + tb;
                uint32_t x = tot;
                for (uint32_t o = 1; o < 32; o <<= 1) {
                    const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(x), o, 32));
                    if (n >= o) x += y;
                }
                const uint32_t all = static_cast<uint32_t>(__shfl(static_cast<int>(x), 31, 32));
                unsigned long long rb = 0;
                if (n == 0) {
                    rb = all ? atomicAdd(a.rcount, static_cast<unsigned long long>(all)) : 0ull;
                    s_rb = rb;
                }
                const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(rb)), 0, 32));
                s_rs[n] = x - tot;
                a.itab[static_cast<uint64_t>(it) * kM8BatchQTile + n] = make_uint2(lo + x - tot, tot | (s_minbin[n] << 16));
            }
            __syncthreads();
            const unsigned long long rb = s_rb;
            if (nt > 0) {
#pragma unroll
                for (int t = 0; t < kM8Tiles; ++t) {
                    if (t >= static_cast<int>(nt)) break;
                    const uint32_t bj = b0 + wave + 4 * t;
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        const uint32_t row = 8 * (i / 4) + 4 * h + (i % 4);
                        const uint32_t mk = static_cast<uint32_t>(s_bal[wave][t][i][0] >> (32 * h));
                        if (!((mk >> n) & 1u)) continue;
                        const unsigned long long at =
                            rb + s_rs[row] + s_off[wave][t][row][1] + __popc(mk & ((1u << n) - 1u));
                        if (at < a.rec_cap)
                            a.rec[at] = make_uint2(static_cast<uint32_t>((bat0 + bj) * 32 + n),
                                                   s_rw[kWrite ? 0 : wave][t][i][lane]);
                    }
                }
            }
        }
        if (kWrite) {
            __syncthreads();
            if (nt > 0) {
#pragma unroll
                for (int t = 0; t < kM8Tiles; ++t) {
                    if (t >= static_cast<int>(nt)) break;
                    const uint32_t bj = b0 + wave + 4 * t;
                    const uint64_t pos = static_cast<uint64_t>(bj) * 32 + n;
                    const Sums sm = sums(t);
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        const uint32_t row = 8 * (i / 4) + 4 * h + (i % 4);
                        float dm;
                        uint32_t bin;
                        if (!test(t, i, sm, dm, bin)) continue;
                        const uint32_t q = s_q[row];
                        const bool in_a = bin <= s_abin[row];
                        const uint32_t below = (1u << n) - 1u;
                        const uint32_t mk = static_cast<uint32_t>(s_bal[wave][t][i][0] >> (32 * h));
                        const uint32_t ma = static_cast<uint32_t>(s_bal[wave][t][i][1] >> (32 * h)) & mk;
                        const uint32_t slot = in_a ? s_base[row][0] + s_off[wave][t][row][0] + __popc(ma & below)
                                                   : s_na[row] + s_base[row][1] + s_off[wave][t][row][1] +
                                                         __popc(mk & ~ma & below);
                        if (in_a ? slot < s_na[row] : slot < G(a.ncand)[q]) {   // the count pass saw the same vectors pass
                            const uint64_t at = G(a.coff)[q] + slot;
                            Gw(a.pool)[at] = (static_cast<uint64_t>(s_p[row]) << 32) | ((bat0 + bj) * 32 + n);
                            Gw(a.pdlo)[at] = dm;
                        }
                    }
                }
            }
        }
        if (!kChk && !kWrite) {   // the item's histogram counts into the queries' histograms
            __syncthreads();
            for (uint32_t e = tid; e < kM8BatchQTile * kM8TauBins; e += 256) {
                const uint32_t row = e / kM8TauBins, x = s_hist[e];
                if (x != 0 && row < np) atomicAdd(&a.hist[static_cast<uint64_t>(s_q[row]) * kM8TauBins + e % kM8TauBins], x);
            }
        }
        if (tid == 0) ++n_items;
        __syncthreads();
    }
    if (!kWrite && lane == 0 && a.stats) {
        atomicAdd(&a.stats[0], n_pairs);
        atomicAdd(&a.stats[1], n_cand);
        if (tid == 0) atomicAdd(&a.stats[5], n_items);
    }
}

__global__ __launch_bounds__(256) void M8LutInt8Kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                       uint64_t words) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < words; i += gridDim.x * 256ull) out[i] = in[i] ^ 0x80808080u;
}

hipError_t LaunchM8LutInt8(hipStream_t st, const uint8_t* in, uint8_t* out, uint64_t bytes) {
    const uint64_t words = bytes / 4;   // plane_cols * 32 bytes per query: whole words
    if (words == 0) return hipSuccess;
    const uint64_t g = (words + 255) / 256;
    hipLaunchKernelGGL(M8LutInt8Kernel, dim3(static_cast<uint32_t>(g < 8192 ? g : 8192)), dim3(256), 0, st,
                       reinterpret_cast<const uint32_t*>(in), reinterpret_cast<uint32_t*>(out), words);
    return hipGetLastError();
}

hipError_t LaunchM8MfmaScan(hipStream_t st, const M8BatchArgs& a, uint32_t grid, bool write) {
    if (a.v.nrun < 1 || a.v.nrun > kM8BatchRuns || (a.v.plane_cols & 7u) != 0) return hipErrorInvalidValue;
    if (write) hipLaunchKernelGGL((M8MfmaScanKernel<false, true>), dim3(grid), dim3(256), 0, st, a);
    else if (a.chk) hipLaunchKernelGGL((M8MfmaScanKernel<true, false>), dim3(grid), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((M8MfmaScanKernel<false, false>), dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Placement: 16 lanes per (item, pair) take the pair's records of the item,
// keep those of the set they place (set A: tau0 bin <= abin[q]; B: the
// rest, with a lower bound below min(tau0, tau1) -- the compaction's test),
// take their slots with one atomic and write the pool entries there.

__global__ __launch_bounds__(256) void M8PlaceKernel(M8BatchArgs a, int set_b) {
    // Each 16-lane quarter of a wave takes its own unit, so a wave has four
    // units' load chains in flight; a quarter counts its unit's kept records
    // (ballots), takes their slots with one atomic, then writes them.  Runs
    // only when the records fit their reserve (else the write pass places).
    if (!M8RecordsPlaced(a)) return;
    const uint32_t lane = threadIdx.x & 63, g = lane >> 4, gl = lane & 15;
    const uint64_t nitems = min<uint64_t>(*G(a.counters), a.items_cap);
    const uint64_t units = nitems * kM8BatchQTile;
    const uint64_t below = (1ull << gl) - 1ull;
    const uint64_t wstride = ((static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6) * 4;
    uint32_t* fill = set_b ? a.fill_b : a.fill_a;
    for (uint64_t u0 = Uniform64(((static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6) * 4); u0 < units;
         u0 += wstride) {
        const uint64_t u = u0 + g;
        uint32_t start = 0, cnt = 0, q = 0, p = 0, abin = 0, room = 0;
        uint64_t base = 0;
        float lim = 0.0F;
        if (u < units) {
            const uint32_t it = static_cast<uint32_t>(u / kM8BatchQTile), m = static_cast<uint32_t>(u % kM8BatchQTile);
            const E8Q_GLOBAL BatchItem* item = G(a.items) + it;
            const E8Q_GLOBAL uint2* ep = G(a.itab) + u;
            const uint32_t np = item->np, ey = ep->y;
            start = ep->x;
            cnt = m < np ? (ey & 0xFFFFu) : 0u;
            if (cnt > 0) {
                const uint64_t pr = G(a.pairs)[item->p0 + m];
                q = static_cast<uint32_t>(pr >> 32);
                p = static_cast<uint32_t>(pr);
                abin = G(a.abin)[q];
                const uint32_t na = G(a.na)[q];
                if (!set_b && (ey >> 16) > abin) cnt = 0;   // none of its records is in set A
                room = set_b ? G(a.ncand)[q] - na : na;
                base = G(a.coff)[q] + (set_b ? na : 0u);
                if (set_b) {
                    const float tau0 = G(a.tau)[q], t1 = G(a.first_d)[static_cast<uint64_t>(q) * a.k + a.k - 1];
                    const float tau = t1 < tau0 ? t1 : tau0;
                    lim = tau + 2.4414062e-4F * fabsf(tau);
                }
            }
        }
        uint32_t most = cnt;   // the quarters' largest count
        for (int o = 16; o < 64; o <<= 1) most = max(most, static_cast<uint32_t>(__shfl_xor(static_cast<int>(most), o)));
        auto keep_of = [&](uint32_t j, uint2& r) -> bool {
            if (j + gl >= cnt) return false;
            const E8Q_GLOBAL uint2* rp = G(a.rec) + static_cast<uint64_t>(start) + j + gl;
            r = make_uint2(rp->x, rp->y);
            const bool in_a = (r.y >> 26) <= abin;
            return set_b ? !in_a && !(M8RecBound(r.y) >= lim) : in_a;
        };
        uint32_t kept = 0;
        for (uint32_t j = 0; j < most; j += 16) {
            uint2 r;
            kept += static_cast<uint32_t>(__popcll((__ballot(keep_of(j, r)) >> (16 * g)) & 0xFFFFull));
        }
        uint32_t s0 = 0;
        if (gl == 0 && kept != 0) s0 = atomicAdd(&fill[q], kept);
        s0 = static_cast<uint32_t>(__shfl(static_cast<int>(s0), static_cast<int>(16 * g)));
        if (kept == 0) continue;
        for (uint32_t j = 0; j < most; j += 16) {
            uint2 r = make_uint2(0u, 0u);
            const bool keep = keep_of(j, r);
            const uint64_t mk = (__ballot(keep) >> (16 * g)) & 0xFFFFull;
            const uint32_t slot = s0 + static_cast<uint32_t>(__popcll(mk & below));
            if (keep && slot < room) Gw(a.pool)[base + slot] = (static_cast<uint64_t>(p) << 32) | r.x;
            s0 += static_cast<uint32_t>(__popcll(mk));
        }
    }
}

hipError_t LaunchM8Place(hipStream_t st, const M8BatchArgs& a, uint32_t grid, bool set_b) {
    if (a.rec == nullptr || a.itab == nullptr) return hipSuccess;   // no records: the write pass places
    hipLaunchKernelGGL(M8PlaceKernel, dim3(grid), dim3(256), 0, st, a, set_b ? 1 : 0);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Threshold pass: each query's first `budget` 32-vector batches in probe
// order (empty lists skipped, the last list possibly cut at a batch), as pool
// entries at q * budget * 32, their count, and where the pass stopped
// (cover[q] = probe << 32 | batch: from that probe's batch on, and in every
// later probe, the vectors are the contraction's).  The refine's whole cascade
// without the death test against tau = +inf then gives their exact distances,
// and the select their top-k: the starting list and tau0.  One wave per query.

__global__ __launch_bounds__(64) void M8ThresholdPoolKernel(M8BatchArgs a, const uint32_t* __restrict__ nprobe_q,
                                                            uint32_t nprobe_all, uint32_t budget) {
    const uint32_t lane = threadIdx.x;
    const uint64_t cap = static_cast<uint64_t>(budget) * 32;
    for (uint64_t q = blockIdx.x; q < a.nq; q += gridDim.x) {
        const uint64_t np = min<uint64_t>(nprobe_q ? nprobe_q[q] : nprobe_all, a.v.nlist);
        uint64_t used = 0, p_stop = np, b_stop = 0, n = 0;
        const uint64_t base = q * cap;
        for (uint64_t p = 0; p < np; ++p) {
            const uint32_t c = a.probe[q * a.maxp + p];
            const uint64_t cntc = a.v.list_begin[c + 1] - a.v.list_begin[c];
            if (cntc == 0) continue;
            const uint64_t nbl = (cntc + 31) >> 5, nb = min<uint64_t>(nbl, budget - used);
            const uint64_t npos = min<uint64_t>(nb * 32, cntc);
            const uint64_t s0 = a.v.batch_begin[c] * 32;   // the list's first batch slot
            for (uint64_t i = lane; i < npos; i += 64) a.pool[base + n + i] = (p << 32) | (s0 + i);
            n += npos;
            used += nb;
            if (used >= budget) {
                p_stop = nb == nbl ? p + 1 : p;
                b_stop = nb == nbl ? 0 : nb;
                break;
            }
        }
        if (lane == 0) {
            a.cover[q] = (p_stop << 32) | b_stop;
            a.nfill[q] = static_cast<uint32_t>(n);
            a.coff[q] = base;
        }
    }
}

hipError_t LaunchM8ThresholdPool(hipStream_t st, const M8BatchArgs& a, const uint32_t* nprobe_q, uint32_t nprobe_all,
                                 uint32_t budget) {
    if (a.nq == 0) return hipSuccess;
    if (budget == 0) return hipErrorInvalidValue;
    const uint32_t grid = static_cast<uint32_t>(a.nq < 65536 ? a.nq : 65536);
    hipLaunchKernelGGL(M8ThresholdPoolKernel, dim3(grid), dim3(64), 0, st, a, nprobe_q, nprobe_all, budget);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pool offsets (coff, from the count pass's ncand) and refine tasks (toff,
// ceil(count / kM8RefineChunk) per query); one workgroup of 1024 threads,
// each owning a contiguous run of queries.

__global__ __launch_bounds__(1024) void M8TaskScanKernel(M8BatchArgs a, const uint32_t* __restrict__ count,
                                                         int with_offsets) {
    __shared__ unsigned long long sc[1024];
    __shared__ uint32_t st[1024];
    const uint64_t nq = a.nq;
    const uint64_t per = (nq + 1023) / 1024;
    const uint64_t q0 = min<uint64_t>(nq, threadIdx.x * per), q1 = min<uint64_t>(nq, q0 + per);
    unsigned long long tc = 0;
    uint32_t tt = 0;
    for (uint64_t q = q0; q < q1; ++q) {
        tc += a.ncand[q];
        tt += (count[q] + kM8RefineChunk - 1) / kM8RefineChunk;
    }
    sc[threadIdx.x] = tc;
    st[threadIdx.x] = tt;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const unsigned long long xc = threadIdx.x >= o ? sc[threadIdx.x - o] : 0ull;
        const uint32_t xt = threadIdx.x >= o ? st[threadIdx.x - o] : 0u;
        __syncthreads();
        sc[threadIdx.x] += xc;
        st[threadIdx.x] += xt;
        __syncthreads();
    }
    uint64_t oc = sc[threadIdx.x] - tc;
    uint32_t ot = st[threadIdx.x] - tt;
    for (uint64_t q = q0; q < q1; ++q) {
        if (with_offsets) a.coff[q] = oc;
        a.toff[q] = ot;
        oc += a.ncand[q];
        ot += (count[q] + kM8RefineChunk - 1) / kM8RefineChunk;
    }
    if (threadIdx.x == 1023) {
        if (with_offsets) a.coff[nq] = sc[1023];
        a.toff[nq] = st[1023];
        a.counters[3] = 0;
    }
}

hipError_t LaunchM8RefineTasks(hipStream_t st, const M8BatchArgs& a, const uint32_t* count, bool with_offsets) {
    hipLaunchKernelGGL(M8TaskScanKernel, dim3(1), dim3(1024), 0, st, a, count, with_offsets ? 1 : 0);
    return hipGetLastError();
}

// Set A of each query: the candidates in the lowest histogram bins (by
// coarse estimate) that hold at least set_a of them, all candidates when
// there are fewer.  tau = tau0 for A's refine.
__global__ __launch_bounds__(256) void M8SplitKernel(M8BatchArgs a) {
    const uint64_t q = blockIdx.x * 256ull + threadIdx.x;
    if (q >= a.nq) return;
    uint32_t c = 0, b = 0;
    for (; b < kM8TauBins; ++b) {
        c += a.hist[q * kM8TauBins + b];
        if (c >= a.set_a) break;
    }
    if (b == kM8TauBins) {   // fewer than set_a candidates (or none binned): all in A
        b = kM8TauBins - 1;
        c = a.ncand[q];
    }
    a.abin[q] = b;
    a.na[q] = c;
    a.tau[q] = a.first_d[q * a.k + a.k - 1];
}

hipError_t LaunchM8Split(hipStream_t st, const M8BatchArgs& a) {
    if (a.nq == 0) return hipSuccess;
    hipLaunchKernelGGL(M8SplitKernel, dim3(static_cast<uint32_t>((a.nq + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

// tau = min(tau0, the k-th exact distance after A), then each query's B
// compacted in place (one wave per query, in order, so a write never passes
// a read) to the candidates whose lower bound at tau0 is below tau: the
// bound only rises as the threshold falls, so no vector the bound keeps at
// tau is dropped.
__global__ __launch_bounds__(64) void M8CompactKernel(M8BatchArgs a) {
    const uint32_t lane = threadIdx.x;
    const bool placed = M8RecordsPlaced(a);
    for (uint64_t q = blockIdx.x; q < a.nq; q += gridDim.x) {
        const float tau0 = a.tau[q], t1 = a.first_d[q * a.k + a.k - 1];
        const float tau = t1 < tau0 ? t1 : tau0;
        const float lim = tau + 2.4414062e-4F * fabsf(tau);
        const uint64_t base = a.coff[q] + a.na[q];
        const uint32_t nb = placed ? 0u : Uniform(a.ncand[q] - a.na[q]);
        uint32_t w = placed ? a.fill_b[q] : 0u;
        for (uint32_t r0 = 0; r0 < nb; r0 += 64) {
            const uint32_t r = r0 + lane;
            uint64_t e = 0;
            bool keep = false;
            if (r < nb) {
                e = a.pool[base + r];
                keep = !(a.pdlo[base + r] >= lim);
            }
            const uint64_t m = __ballot(keep);
            if (keep) a.pool[base + w + __popcll(m & ((uint64_t{1} << lane) - 1))] = e;
            w += static_cast<uint32_t>(__popcll(m));
        }
        if (lane == 0) {
            a.tau[q] = tau;
            a.coff_b[q] = base;
            a.nfill[q] = w;
        }
    }
}

hipError_t LaunchM8Compact(hipStream_t st, const M8BatchArgs& a) {
    if (a.nq == 0) return hipSuccess;
    const uint32_t grid = static_cast<uint32_t>(a.nq < 65536 ? a.nq : 65536);
    hipLaunchKernelGGL(M8CompactKernel, dim3(grid), dim3(64), 0, st, a);
    return hipGetLastError();
}

// The candidate pool holds pool_cap entries: a query whose slots (from the
// count pass's offsets) pass its end keeps no candidates (ncand = |A| = 0,
// so nothing is placed, written or refined for it) and goes on the overflow
// list, whose queries the plain form searches at the end.  Offsets grow with
// the query index, so the overflowed queries are the last ones.
__global__ __launch_bounds__(256) void M8FitKernel(M8BatchArgs a, uint64_t pool_cap, uint32_t* __restrict__ ovf,
                                                   uint32_t* __restrict__ novf) {
    const uint64_t q = blockIdx.x * 256ull + threadIdx.x;
    if (q >= a.nq) return;
    if (a.coff[q] + a.ncand[q] > pool_cap) {
        a.ncand[q] = 0;
        a.na[q] = 0;
        ovf[atomicAdd(novf, 1u)] = static_cast<uint32_t>(q);
    }
}

hipError_t LaunchM8Fit(hipStream_t st, const M8BatchArgs& a, uint64_t pool_cap, uint32_t* ovf, uint32_t* novf) {
    if (a.nq == 0) return hipSuccess;
    hipLaunchKernelGGL(M8FitKernel, dim3(static_cast<uint32_t>((a.nq + 255) / 256)), dim3(256), 0, st, a, pool_cap,
                       ovf, novf);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Refine.  One wave takes one task (up to 64 candidates of one query), one
// lane per candidate, so the loads of 64 candidates are in flight together.
// LDS: the query's rotated vector, its u8 LUT, the level tables of the runs
// at rates <= kM8LdsLevRate.

// The exact score of one run for one vector, computed by one lane in the
// order the reference's AVX2 code uses (me8q_score_avx2.cpp: even blocks in
// one accumulator and odd blocks in another, per component, added, then the
// horizontal sum): e[ci] / o[ci] are the FMA chains of component ci over the
// even / odd blocks in block order.  Digit decode as M8HalfChainT.  Runs are
// multiples of 64 dimensions, so a run holds whole groups of eight blocks, and
// its plane bytes and fields start 8-byte aligned in their rows.
template <int T, class LevP>
__device__ __forceinline__ float M8LaneScoreT(const E8Q_GLOBAL uint8_t* __restrict__ pb,
                                              const E8Q_GLOBAL uint8_t* __restrict__ fr, LevP lev,
                                              const E8Q_GLOBAL float* __restrict__ qr, uint32_t nblk) {
    constexpr uint32_t rate = T + 1;
    constexpr uint32_t NW = T > 0 ? T : 1;
    // `nb` bits at bit `pos` of the little-endian words w (positions are
    // compile-time constants once the block loop is unrolled)
    auto bits = [](const uint64_t* w, uint32_t pos, uint32_t nb) -> uint32_t {
        const uint32_t wi = pos >> 6, off = pos & 63u;
        uint64_t x = w[wi] >> off;
        if (off + nb > 64) x |= w[wi + 1] << (64 - off);
        return static_cast<uint32_t>(x) & ((1u << nb) - 1u);
    };
    float e[8] = {0, 0, 0, 0, 0, 0, 0, 0}, o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // eight blocks at a time: their plane bytes (one 8-byte load) and their
    // fields (8 T bytes, T 8-byte loads), all loads in flight together
    for (uint32_t g8 = 0; g8 < nblk; g8 += 8) {
        const uint64_t pw = *reinterpret_cast<const E8Q_GLOBAL uint64_t*>(pb + g8);
        uint64_t fw[NW];
#pragma unroll
        for (uint32_t i = 0; i < NW; ++i)
            fw[i] = T > 0 ? *reinterpret_cast<const E8Q_GLOBAL uint64_t*>(fr + static_cast<uint64_t>(g8) * T + 8 * i)
                          : 0ull;
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b) {
            const uint32_t pbyte = static_cast<uint32_t>(pw >> (8 * b)) & 0xFFu;
            uint32_t dg[8], br;
            if constexpr (T == 0) {
#pragma unroll
                for (uint32_t ci = 0; ci < 7; ++ci) dg[ci] = (pbyte >> (7 - ci)) & 1u;
                dg[7] = __popc(pbyte >> 1) & 1u;
                br = pbyte & 1u;
            } else {
                const uint32_t base = 8 * b * T;
                uint32_t par = 0;
#pragma unroll
                for (uint32_t ci = 0; ci < 7; ++ci) {
                    const uint32_t f = bits(fw, base + T * ci, T);
                    par ^= f & 1u;
                    dg[ci] = f | (((pbyte >> (7 - ci)) & 1u) << T);
                }
                const uint32_t f7 = T > 1 ? bits(fw, base + 7 * T, T - 1) : 0u;
                dg[7] = ((f7 << 1) | par) | ((pbyte & 1u) << T);
                br = bits(fw, base + 8 * T - 1, 1);
            }
            const auto lb = lev + (br << rate);
            const E8Q_GLOBAL float* qb = qr + (g8 + b) * 8;
            if (b & 1u) {
#pragma unroll
                for (uint32_t ci = 0; ci < 8; ++ci) o[ci] = fmaf(lb[dg[ci]], qb[ci], o[ci]);
            } else {
#pragma unroll
                for (uint32_t ci = 0; ci < 8; ++ci) e[ci] = fmaf(lb[dg[ci]], qb[ci], e[ci]);
            }
        }
    }
    float v[8];
#pragma unroll
    for (uint32_t ci = 0; ci < 8; ++ci) v[ci] = e[ci] + o[ci];
    return ((v[0] + v[4]) + (v[2] + v[6])) + ((v[1] + v[5]) + (v[3] + v[7]));
}

// The score of one run by its rate, with the level tables in LDS (LevP =
// const E8Q_LDS float*) or in global memory.
template <class LevP>
__device__ float M8LaneScore(uint32_t rate, const E8Q_GLOBAL uint8_t* pb, const E8Q_GLOBAL uint8_t* fr, LevP lev,
                             const E8Q_GLOBAL float* qr, uint32_t nblk) {
    switch (rate) {
        case 1: return M8LaneScoreT<0>(pb, fr, lev, qr, nblk);
        case 2: return M8LaneScoreT<1>(pb, fr, lev, qr, nblk);
        case 3: return M8LaneScoreT<2>(pb, fr, lev, qr, nblk);
        case 4: return M8LaneScoreT<3>(pb, fr, lev, qr, nblk);
        case 5: return M8LaneScoreT<4>(pb, fr, lev, qr, nblk);
        case 6: return M8LaneScoreT<5>(pb, fr, lev, qr, nblk);
        case 7: return M8LaneScoreT<6>(pb, fr, lev, qr, nblk);
        case 8: return M8LaneScoreT<7>(pb, fr, lev, qr, nblk);
        case 9: return M8LaneScoreT<8>(pb, fr, lev, qr, nblk);
        case 10: return M8LaneScoreT<9>(pb, fr, lev, qr, nblk);
        case 11: return M8LaneScoreT<10>(pb, fr, lev, qr, nblk);
        case 12: return M8LaneScoreT<11>(pb, fr, lev, qr, nblk);
        default: return M8LaneScoreT<12>(pb, fr, lev, qr, nblk);
    }
}

// LDS of a refine workgroup: the level tables of the runs staged there
// (lev_floats floats, run r at lev_off[r]), then each wave's u8 LUT.
struct M8RefLayout {
    uint32_t lut, total;
    __host__ __device__ M8RefLayout(uint32_t PC, uint32_t LV) {
        lut = (LV * 4 + 15) & ~15u;
        total = lut + kM8RefineWaves * PC * 32;
    }
};

// Orders this wave's LDS accesses before and after it (a wave's LDS
// operations are performed in order; this keeps the compiler from moving
// them across).
__device__ __forceinline__ void WaveLdsSync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <bool kFirst>
__global__ __launch_bounds__(64 * kM8RefineWaves, 3) void M8RefineKernel(M8BatchArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const M8View& v = a.v;
    const uint32_t P = static_cast<uint32_t>(v.padded), PC = v.plane_cols, R = v.nrun, k = a.k, maxp = a.maxp;
    const M8RefLayout L(PC, a.ref_lev_floats);
    const uint32_t lane = threadIdx.x & 63, wave = Uniform(threadIdx.x >> 6);
    uint8_t* slut = smem + L.lut + wave * PC * 32;   // this wave's query LUT
    {   // the staged level tables, by the whole workgroup
        float* slev = reinterpret_cast<float*>(smem);
        for (uint32_t r = 0; r < R; ++r) {
            if (a.ref_lev_off[r] == kM8LevGlobal) continue;
            const uint32_t nq2 = 2u << v.runs[r].rate;
            for (uint32_t e = threadIdx.x; e < nq2; e += 64 * kM8RefineWaves) slev[a.ref_lev_off[r] + e] = v.runs[r].lev[e];
        }
        __syncthreads();
    }
    const float finf = __builtin_inff();
    const double dinf = __builtin_inf();
    const double dz = a.z;
    const uint32_t ntask = Uniform(a.toff[a.nq]);
    uint32_t qcur = 0xFFFFFFFFu;
    unsigned long long n_ref = 0, n_surv = 0, n_bits = 0, n_bytes = 0;
    // Tasks are taken kM8RefineGrab at a time (consecutive tasks mostly share
    // a query: one search for it, its LDS data reused).
    uint32_t task = 0, tend = 0, q = 0, qend = 0;
    for (;;) {
        if (task == tend) {
            uint32_t t0 = 0;
            if (lane == 0) t0 = atomicAdd(&a.counters[3], kM8RefineGrab);
            task = Uniform(static_cast<uint32_t>(__shfl(static_cast<int>(t0), 0)));
            if (task >= ntask) break;
            tend = min(ntask, task + kM8RefineGrab);
            uint32_t qs = 0;
            if (lane == 0) {   // the task's query: the last q with toff[q] <= task
                uint64_t lo = 0, hi = a.nq;
                while (hi - lo > 1) {
                    const uint64_t mid = (lo + hi) >> 1;
                    if (a.toff[mid] <= task) lo = mid;
                    else hi = mid;
                }
                qs = static_cast<uint32_t>(lo);
            }
            q = Uniform(static_cast<uint32_t>(__shfl(static_cast<int>(qs), 0)));
            qend = Uniform(a.toff[q + 1]);
        }
        while (task >= qend) {   // the next query with tasks (toff is non-decreasing)
            ++q;
            qend = Uniform(a.toff[q + 1]);
        }
        if (q != qcur) {
            WaveLdsSync();   // the previous query's LUT is no longer read
            for (uint32_t e = lane; e < PC * 8; e += 64)
                reinterpret_cast<uint32_t*>(slut)[e] =
                    reinterpret_cast<const uint32_t*>(a.lut + static_cast<uint64_t>(q) * PC * 32)[e];
            qcur = q;
            WaveLdsSync();
        }
        const uint32_t nc = a.nfill[q];
        const uint32_t c0 = (task - a.toff[q]) * kM8RefineChunk;
        const uint32_t c1 = min(nc, c0 + kM8RefineChunk);
        uint64_t* slots = a.pool + a.coff[q] + c0;   // this task's candidates, then its survivors
        const bool has = c0 + lane < c1;
        const uint64_t ent = has ? slots[lane] : 0ull;   // read before survivors are written over it
        const float tau = a.tau[q];
        const double taud = tau < finf ? static_cast<double>(tau) : dinf;
        const float sqk = tau < finf ? sqrtf(fmaxf(0.0F, tau)) : finf;
        const double sqdk = taud < dinf ? sqrt(fmax(taud, 0.0)) : dinf;
        const float lut_delta = a.qc[static_cast<uint64_t>(q) * 4 + 1];
        const double* rcq = a.rc + static_cast<uint64_t>(q) * R * 4;
        bool live = false;
        float dsurv = 0.0F;
        uint32_t idsurv = 0;
        if (has) {
            // the entry names the probe and the vector's batch slot: every
            // load below depends on the entry alone
            const uint32_t p = static_cast<uint32_t>(ent >> 32), slot = static_cast<uint32_t>(ent);
            const uint64_t gb = slot >> 5;
            const uint32_t l = slot & 31u;
            const uint32_t cntb = v.bcnt[gb];
            const uint32_t covered = cntb >= 8 ? (cntb & ~7u) : 0u;
            const uint8_t* prow = v.plane + gb * v.rec + static_cast<uint64_t>(l) * PC;
            // the candidate's refine record: run headers, its W / Omega bank
            // bytes, the run fields; the batch's scales per run
            const uint8_t* rr = v.rrec + static_cast<uint64_t>(slot) * v.rrec_bytes;
            const uint8_t* rbank = rr + 16 * R;
            const float4* bsc = v.bscale + gb * R;
            // algorithmic bytes: the pool entry, the plane row, per run its
            // 16-byte header, its W and Omega bank bytes and 1/32 of the
            // batch's scales
            n_bytes += 8 + PC + 19 * R;
            const double d2 = static_cast<double>(a.pdist[static_cast<uint64_t>(q) * maxp + p]);
            const double* sqd2 = a.sqd2 + (static_cast<uint64_t>(q) * maxp + p) * R;
            // The coarse terms in the reference's arithmetic, against tau, run by
            // run in run order (fp32 for the lanes the AVX2 filter covers, the
            // fp64 twin for the rest).  Every private array here is indexed by
            // compile-time constants only (guarded, unrolled run loops; Pick for
            // a run-time position) so that it stays in registers.  The terms
            // that do not involve the coarse sums come first: they fix the
            // cascade's run order, so the first run is scored before the plane
            // row is read again for the sums, while its line is still cached.
            double str[kM8BatchRuns] = {0.0, 0.0, 0.0, 0.0};
            double pool = 0.0, qadd = 0.0;
            const bool f32 = l < covered;
            if (f32) {   // the AVX2 filter's fp32 values
                float pl = 0.0F, qa = 0.0F, wpad = 0.0F;
#pragma unroll
                for (uint32_t r = 0; r < kM8BatchRuns; ++r) {
                    if (r < R) {
                        const float4 s4 = bsc[r];
                        const float wq = fmaf(static_cast<float>(rbank[2 * r]), s4.y, s4.x);
                        const float pq = fmaf(static_cast<float>(rbank[2 * r + 1]), s4.w, s4.z);
                        const float ra = static_cast<float>(sqd2[r]);
                        const float rh = sqk < ra ? sqk : ra;
                        const float t = pq * rh;
                        str[r] = t;
                        pl = fmaf(t, t, pl);
                        qa = fmaf(fabsf(wq), static_cast<float>(rcq[r * 4 + 2]), qa);
                        wpad = wpad + static_cast<float>(0.5 * static_cast<double>(s4.y) * (rcq[r * 4 + 3] + rcq[r * 4 + 2]));
                    }
                }
                pool = pl;
                qadd = qa + wpad;
            } else {   // the fp64 twin for the lanes the filter does not cover
#pragma unroll
                for (uint32_t r = 0; r < kM8BatchRuns; ++r) {
                    if (r < R) {
                        const float4 s4 = bsc[r];
                        const double wq = fma(static_cast<double>(rbank[2 * r]), static_cast<double>(s4.y), static_cast<double>(s4.x));
                        const double pq = fma(static_cast<double>(rbank[2 * r + 1]), static_cast<double>(s4.w), static_cast<double>(s4.z));
                        const double sa = sqd2[r];
                        const double rh = sqdk < sa ? sqdk : sa;
                        const double tr = pq * rh;
                        str[r] = tr;
                        pool = fma(tr, tr, pool);
                        qadd = fma(fabs(wq), rcq[r * 4 + 2], qadd);
                        qadd = fma(0.5 * static_cast<double>(s4.y), rcq[r * 4 + 3] + rcq[r * 4 + 2], qadd);
                    }
                }
            }
            // The cascade's run order for a group of one candidate: decreasing
            // t^2 / nblk, stable (GCC's insertion sort below 17 elements): a
            // sorting network on the total order (mean descending, run
            // ascending) -- the same order -- carrying each run's t; runs past R
            // sort last.
            uint32_t ord[kM8BatchRuns] = {0, 1, 2, 3};
            double sts[kM8BatchRuns] = {str[0], str[1], str[2], str[3]};   // t in cascade order
            if (R > 1) {
                double mean[kM8BatchRuns];
#pragma unroll
                for (uint32_t r = 0; r < kM8BatchRuns; ++r)
                    mean[r] = r < R ? fma(str[r], str[r], 0.0) /
                                          static_cast<double>(v.runs[r].nblk > 1 ? v.runs[r].nblk : 1)
                                    : -1.0;
                auto cswap = [&](uint32_t i, uint32_t j) {   // i < j: j first if it sorts before i
                    const bool sw = mean[j] > mean[i] || (mean[j] == mean[i] && ord[j] < ord[i]);
                    const double m0 = mean[i], t0 = sts[i];
                    const uint32_t o0 = ord[i];
                    mean[i] = sw ? mean[j] : m0;
                    sts[i] = sw ? sts[j] : t0;
                    ord[i] = sw ? ord[j] : o0;
                    mean[j] = sw ? m0 : mean[j];
                    sts[j] = sw ? t0 : sts[j];
                    ord[j] = sw ? o0 : ord[j];
                };
                cswap(0, 1);
                cswap(2, 3);
                cswap(0, 2);
                cswap(1, 3);
                cswap(1, 2);
            }
            auto pick = [](const auto (&x)[kM8BatchRuns], uint32_t i) {
                return i == 0 ? x[0] : i == 1 ? x[1] : i == 2 ? x[2] : x[3];
            };
            const uint32_t lo0 = a.ref_lev_off[0], lo1 = a.ref_lev_off[1], lo2 = a.ref_lev_off[2],
                           lo3 = a.ref_lev_off[3];
            const E8Q_GLOBAL float* qrot = G(a.rotq) + static_cast<uint64_t>(q) * P;
            // The exact score of run r (its plane bytes and fields).
            auto run_score = [&](uint32_t r) -> double {
                const M8Run& ru = v.runs[r];
                const E8Q_GLOBAL uint8_t* pb = G(prow) + ru.blk0;
                const E8Q_GLOBAL uint8_t* fr = G(rr) + ru.ref_off + v.rrec_shift;
                const uint32_t lo = r == 0 ? lo0 : r == 1 ? lo1 : r == 2 ? lo2 : lo3;
                return lo != kM8LevGlobal
                           ? M8LaneScore(ru.rate, pb, fr, (const E8Q_LDS float*)(reinterpret_cast<const float*>(smem)) + lo,
                                         qrot + ru.dim_off, ru.nblk)
                           : M8LaneScore(ru.rate, pb, fr, G(ru.lev), qrot + ru.dim_off, ru.nblk);
            };
            const double se0 = run_score(ord[0]);
            // integer coarse sums of each run (AccumulateCoarseHi over its columns)
            int32_t part[kM8BatchRuns] = {0, 0, 0, 0};
            auto word_sum = [&](uint64_t w, uint32_t col) -> int32_t {
                int32_t s = 0;
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const uint32_t byte = static_cast<uint32_t>(w >> (8 * b)) & 0xFFu;
                    const uint8_t* lc = slut + (col + b) * 32;
                    s += lc[byte & 15u] + lc[16 + (byte >> 4)];
                }
                return s;
            };
            if (PC <= 8 * kM8RefRowWords) {
                // the row's plane words in one burst of independent loads (the
                // runs cover the row in whole words, in order from column 0),
                // then each word's sum into its run
                uint64_t wv[kM8RefRowWords];
#pragma unroll
                for (uint32_t j = 0; j < kM8RefRowWords; ++j)
                    wv[j] = 8 * j < PC ? *reinterpret_cast<const E8Q_GLOBAL uint64_t*>(G(prow) + 8 * j) : 0ull;
                const uint32_t e0 = v.runs[0].blk0 + v.runs[0].nblk;
                const uint32_t e1 = R > 1 ? v.runs[1].blk0 + v.runs[1].nblk : PC;
                const uint32_t e2 = R > 2 ? v.runs[2].blk0 + v.runs[2].nblk : PC;
#pragma unroll
                for (uint32_t j = 0; j < kM8RefRowWords; ++j) {
                    if (8 * j < PC) {
                        const int32_t sw = word_sum(wv[j], 8 * j);
                        const uint32_t r = (8 * j >= e0 ? 1u : 0u) + (8 * j >= e1 ? 1u : 0u) + (8 * j >= e2 ? 1u : 0u);
                        part[0] += r == 0 ? sw : 0;
                        part[1] += r == 1 ? sw : 0;
                        part[2] += r == 2 ? sw : 0;
                        part[3] += r == 3 ? sw : 0;
                    }
                }
            } else {
#pragma unroll
                for (uint32_t r = 0; r < kM8BatchRuns; ++r) {
                    if (r < R) {
                        const uint32_t b0 = v.runs[r].blk0, b1 = b0 + v.runs[r].nblk;
                        for (uint32_t col = b0; col < b1; col += 8)
                            part[r] += word_sum(*reinterpret_cast<const uint64_t*>(prow + col), col);
                    }
                }
            }
            // each run's coarse score
            double ssc[kM8BatchRuns] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (uint32_t r = 0; r < kM8BatchRuns; ++r) {
                if (r < R)
                    ssc[r] = f32 ? static_cast<double>(fmaf(static_cast<float>(part[r]), lut_delta,
                                                            static_cast<float>(rcq[r * 4 + 1])))
                                 : fma(static_cast<double>(lut_delta), static_cast<double>(part[r]), rcq[r * 4 + 1]);
            }
            const float* f4 = reinterpret_cast<const float*>(rr);
            double dist;
            if (R == 1) {
                dist = d2 + static_cast<double>(f4[1]);
            } else {
                double dcq = d2;
#pragma unroll
                for (uint32_t r = 0; r < kM8BatchRuns; ++r) {
                    if (r < R) {
                        const double coef = f4[r * 4], h2 = f4[r * 4 + 1];
                        const double gn = f4[r * 4 + 2], vp = f4[r * 4 + 3];
                        dcq += fma(-(2.0 * coef), fma(gn, ssc[r], vp), h2);
                    }
                }
                dist = dcq;
            }
            double rem2 = pool;
            live = true;
            uint64_t bits = 0;
            const uint32_t nsteps = kFirst ? 1u : R;
            for (uint32_t oi = 0; oi < nsteps && live; ++oi) {
                const uint32_t r = pick(ord, oi);
                const M8Run& ru = v.runs[r];
                const double se = oi == 0 ? se0 : run_score(r);
                const double coef = f4[r * 4];
                if (R == 1) {
                    dist = fma(-(2.0 * coef), se, dist);
                } else {
                    const double gn = f4[r * 4 + 2], vp = f4[r * 4 + 3];
                    dist = fma(2.0 * coef, fma(gn, pick(ssc, r), vp) - se, dist);
                }
                const double t = pick(sts, oi);
                rem2 = fma(-t, t, rem2);
                if (kFirst || oi > 0) bits += static_cast<uint64_t>(ru.rate) * ru.len;
                n_bytes += static_cast<uint64_t>(ru.rate - 1) * ru.len / 8;   // the run's fields
                if (a.cascade && fma(-dz, sqrt(fmax(rem2, 0.0)), dist) - qadd >= taud) live = false;
            }
            if (kFirst) ++n_ref;
            n_bits += bits;
            if (!kFirst || R == 1) {
                dsurv = static_cast<float>(dist);
                live = live && (dsurv < tau || !(tau < finf));
                if (live) idsurv = v.pids[slot];
            }
        }
        if (kFirst && R > 1) {   // the cascade continues in the second phase
            const uint64_t m = __ballot(live);
            uint32_t base = 0;
            if (lane == 0 && m != 0) base = atomicAdd(&a.ncont[q], static_cast<uint32_t>(__popcll(m)));
            base = Uniform(static_cast<uint32_t>(__shfl(static_cast<int>(base), 0)));
            if (live) {
                a.cont[a.coff[q] + base + __popcll(m & ((uint64_t{1} << lane) - 1))] = ent;
                n_bytes += 8;
            }
            if (lane == 0) a.tsurv[task] = 0u;
            ++task;
            continue;
        }
        // survivors, compacted over the task's first slots
        const uint64_t m = __ballot(live);
        if (live) {
            slots[__popcll(m & ((uint64_t{1} << lane) - 1))] = (static_cast<uint64_t>(M8Key(dsurv)) << 32) | idsurv;
            ++n_surv;
            n_bytes += 12;   // its id read, its entry written
        }
        if (lane == 0) a.tsurv[task] = static_cast<uint32_t>(__popcll(m));
        ++task;
    }
    // per-wave totals
    for (int o = 32; o > 0; o >>= 1) {
        n_ref += static_cast<unsigned long long>(__shfl_xor(static_cast<long long>(n_ref), o));
        n_surv += static_cast<unsigned long long>(__shfl_xor(static_cast<long long>(n_surv), o));
        n_bits += static_cast<unsigned long long>(__shfl_xor(static_cast<long long>(n_bits), o));
        n_bytes += static_cast<unsigned long long>(__shfl_xor(static_cast<long long>(n_bytes), o));
    }
    if (lane == 0 && a.stats && (n_ref | n_surv | n_bytes)) {
        atomicAdd(&a.stats[2], n_ref);
        atomicAdd(&a.stats[3], n_surv);
        atomicAdd(&a.stats[4], n_bits);
        atomicAdd(&a.stats[kFirst ? 6 : 7], n_bytes);
    }
}

size_t M8RefineLds(const M8View& v, uint32_t lev_floats) { return M8RefLayout(v.plane_cols, lev_floats).total; }

hipError_t LaunchM8Refine(hipStream_t st, const M8BatchArgs& a, uint32_t grid, bool first) {
    const size_t lds = M8RefineLds(a.v, a.ref_lev_floats);
    const void* fn = first ? reinterpret_cast<const void*>(&M8RefineKernel<true>)
                           : reinterpret_cast<const void*>(&M8RefineKernel<false>);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    if (first) hipLaunchKernelGGL(M8RefineKernel<true>, dim3(grid), dim3(64 * kM8RefineWaves), lds, st, a);
    else hipLaunchKernelGGL(M8RefineKernel<false>, dim3(grid), dim3(64 * kM8RefineWaves), lds, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Top-k of each query: its threshold-pass list and its survivors (at the
// start of each of its refine tasks' pool slots), sorted by (distance, id)
// with a bitonic sort in LDS of the best k so far plus the next chunk.  One
// wave per query.

constexpr uint32_t kM8SelChunk = 1024;

__global__ __launch_bounds__(64) void M8SelectKernel(M8BatchArgs a) {
    __shared__ uint64_t buf[kM8SelChunk];
    const uint32_t lane = threadIdx.x, k = a.k;
    const float finf = __builtin_inff();
    for (uint64_t q = blockIdx.x; q < a.nq; q += gridDim.x) {
        const uint32_t t0 = Uniform(a.toff[q]), t1 = Uniform(a.toff[q + 1]);
        uint32_t ns = 0;
        for (uint32_t t = t0; t < t1; ++t) ns += a.tsurv[t];
        ns = Uniform(ns);
        const uint64_t* slots = a.pool + a.coff[q];
        const uint32_t n = k + ns;
        uint32_t ch = 64;
        while (ch < kM8SelChunk && ch < n) ch <<= 1;
        const uint32_t room = ch - k;
        for (uint32_t i = lane; i < k; i += 64) {   // the threshold-pass list
            const float d = a.first_d[q * k + i];
            buf[i] = d < finf ? (static_cast<uint64_t>(M8Key(d)) << 32) | a.first_ids[q * k + i] : ~0ull;
        }
        // survivors task by task: task t holds tsurv[t] of them at slots (t - t0) * chunk
        uint32_t t = t0, done = 0, s = 0;
        do {
            uint32_t fill = 0;
            while (fill < room && t < t1) {
                const uint32_t nt = Uniform(a.tsurv[t]);
                const uint32_t take = min(room - fill, nt - done);
                const uint64_t* src = slots + static_cast<uint64_t>(t - t0) * kM8RefineChunk + done;
                for (uint32_t i = lane; i < take; i += 64) buf[k + fill + i] = src[i];
                fill += take;
                done += take;
                if (done == nt) {
                    ++t;
                    done = 0;
                }
            }
            for (uint32_t i = fill + lane; i < room; i += 64) buf[k + i] = ~0ull;
            s += fill;
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
