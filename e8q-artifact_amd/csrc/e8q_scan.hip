// e8q_scan.hip -- the plain search kernel: one wave per query, the
// reference's per-query schedule reproduced bit for bit.
//
// Reference: E8qIndex::Search (e8q_search.cpp:215-293) with ScanClusterAvx2
// (e8q_search_avx2.cpp:407-598), CollectLanes (e8q_scan_common.hpp:87-113),
// FlushSurvivors (e8q_search.cpp:182-207), RefineCoreAvx2 (e8q_search_avx2.cpp:
// 82-403), InsertHeap/ExtractTopK (e8q_search.cpp:135-180, 295-318).
//
// Layout per query: the probed clusters are visited in probe order; each
// cluster's 32-vector batches are scored two at a time (64 lanes: coarse
// integer accumulate from the LDS query LUT + the threshold-independent part
// of the ZPR bound), then batch skip, lane mask and survivor staging against
// the live top-k threshold, refining the staged rows whenever flush_every
// rows are waiting and at the end of every cluster.  The top-k is a sorted
// (dist, id) list in LDS; its k-th entry is the reference heap's root, the
// prune threshold.
#include "e8q_device.h"
#include "e8q_kernels.h"

namespace e8qdev {

constexpr int kWave = 64;       // one wave per query
constexpr int kMaxFlush = 256;

struct StagedRow {
    uint32_t off;   // vector offset inside the current cluster
    uint32_t id;
    float g1;
    float h3;       // d2 + h3 (e8q.hpp:76-79)
};

__device__ __forceinline__ bool Better(float d1, uint32_t i1, float d2, uint32_t i2) {
    return d1 < d2 || (d1 == d2 && i1 < i2);
}

// Digit `i` (0..7) and the branch of one (lane, block) from the rowA byte and
// the (bits-1)-bit refine field (e8q_scan.cpp:212-265).
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

struct PlainShared {
    float t1[kWave];             // ZPR bound minus the slack term, per lane
    float pp[kWave];             // perp (Omega) per lane
    int32_t amax[2], amin[2];    // batch max/min of the coarse sums
    float td[kMaxK], nd[kMaxK];  // sorted top-k (dist, id), merge target
    uint32_t ti[kMaxK], ni[kMaxK];
    float cd[32];
    uint32_t ci[32];
    float limit;
};

// Refines the staged rows, up to 32 at a time, and merges them into the
// sorted top-k list.  The rows' code bytes (one rowA byte per block at
// stride 32, plus the contiguous (bits-1)-byte refine fields) are gathered
// into LDS with independent loads first; then 8 rows x 8 components per
// pass decode and accumulate in the AVX2 order (even blocks in acc, odd in
// acc2, then the 8-lane horizontal sum).
__device__ void FlushRows(PlainShared& S, const StagedRow* staged, int ns, int& cnt, int k, const IndexView& ix,
                          const float* qv, const uint8_t* arena_c, const uint8_t* ref_c, uint8_t* rowbuf,
                          const float* tab) {
    const int lane = threadIdx.x;
    const int B = ix.blocks;
    const int c = ix.bits - 1;
    const int fbytes = ix.bits >= 2 ? B * c : 0;
    const int rstride = B + fbytes;
    for (int rb = 0; rb < ns; rb += 32) {
        __syncthreads();
        const int nn = min(32, ns - rb);
        // Loads are issued four at a time per lane before their LDS stores so
        // several gathers are in flight.
        const int na = nn * B;
        for (int e0 = lane; e0 < na; e0 += 4 * kWave) {
            uint32_t v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + u * kWave;
                v[u] = 0;
                if (e < na) {
                    const int r = e / B, blk = e - r * B;
                    const uint32_t off = staged[rb + r].off;
                    v[u] = arena_c[static_cast<uint64_t>(off >> 5) * ix.batch_bytes + kHdr + blk * 32 + (off & 31)];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + u * kWave;
                if (e < na) {
                    const int r = e / B, blk = e - r * B;
                    rowbuf[r * rstride + blk] = static_cast<uint8_t>(v[u]);
                }
            }
        }
        if (fbytes > 0) {
            const int w8 = fbytes / 8;   // blocks is a multiple of 8: whole 8-byte words
            const int nw = nn * w8;
            for (int e0 = lane; e0 < nw; e0 += 4 * kWave) {
                uint2 v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int e = e0 + u * kWave;
                    v[u] = make_uint2(0, 0);
                    if (e < nw) {
                        const int r = e / w8, w = e - r * w8;
                        const uint32_t off = staged[rb + r].off;
                        const uint8_t* src = ref_c + static_cast<uint64_t>(off >> 5) * ix.ref_batch_bytes +
                                             (off & 31) * ix.refine_stride;
                        v[u] = reinterpret_cast<const uint2*>(src)[w];
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int e = e0 + u * kWave;
                    if (e < nw) {
                        const int r = e / w8, w = e - r * w8;
                        uint2* dst = reinterpret_cast<uint2*>(rowbuf + r * rstride + B + w * 8);
                        *dst = v[u];
                    }
                }
            }
        }
        __syncthreads();
        for (int sub = 0; sub < nn; sub += 8) {
            const int r = sub + (lane >> 3);
            const int i = lane & 7;
            if (r < nn) {
                const StagedRow row = staged[rb + r];
                const uint8_t* plane = rowbuf + r * rstride;
                const uint8_t* field = plane + B;
                float acc = 0.0F, acc2 = 0.0F;
#pragma unroll 4
                for (int blk = 0; blk < B; ++blk) {
                    uint32_t digit;
                    int br;
                    DecodeDigit(ix.bits, plane[blk], field + blk * c, i, &digit, &br);
                    const float lv = tab[static_cast<uint32_t>(br) * ix.tab_stride + digit];
                    if (blk & 1) acc2 = fmaf(lv, qv[blk * 8 + i], acc2);
                    else acc = fmaf(lv, qv[blk * 8 + i], acc);
                }
                const float sc = Hsum8Lanes(acc + acc2);
                if (i == 0) {
                    S.cd[r] = fmaf(-(2.0F * row.g1), sc, row.h3);
                    S.ci[r] = row.id;
                }
            }
        }
        __syncthreads();
        // Merge: each list entry moves down by the number of new rows better
        // than it; each new row lands at (# list entries better) + (# new
        // rows better).  (dist, id) is a strict total order (ids unique).
        for (int e = lane; e < cnt; e += kWave) {
            const float d = S.td[e];
            const uint32_t id = S.ti[e];
            int rank = e;
#pragma unroll 8
            for (int j = 0; j < nn; ++j) rank += Better(S.cd[j], S.ci[j], d, id) ? 1 : 0;
            if (rank < k) {
                S.nd[rank] = d;
                S.ni[rank] = id;
            }
        }
        if (lane < nn) {
            const float d = S.cd[lane];
            const uint32_t id = S.ci[lane];
            int lo = 0, hi = cnt;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (Better(S.td[mid], S.ti[mid], d, id)) lo = mid + 1;
                else hi = mid;
            }
            int rank = lo;
#pragma unroll 8
            for (int j = 0; j < nn; ++j) rank += Better(S.cd[j], S.ci[j], d, id) ? 1 : 0;
            if (rank < k) {
                S.nd[rank] = d;
                S.ni[rank] = id;
            }
        }
        __syncthreads();
        cnt = min(k, cnt + nn);
        for (int e = lane; e < cnt; e += kWave) {
            S.td[e] = S.nd[e];
            S.ti[e] = S.ni[e];
        }
        __syncthreads();
        if (lane == 0) S.limit = cnt == k ? S.td[k - 1] : __builtin_inff();
    }
    __syncthreads();
}

__device__ __forceinline__ StagedRow MakeRow(const IndexView& ix, const uint8_t* ref_c, uint64_t vb, uint32_t bj,
                                             uint32_t t, float d2) {
    const uint32_t off = bj * 32;
    const float* g1b =
        reinterpret_cast<const float*>(ref_c + static_cast<uint64_t>(bj) * ix.ref_batch_bytes + ix.refine_stride * 32);
    return StagedRow{off + t, ix.ids[vb + off + t], g1b[t], d2 + g1b[32 + t]};
}

// ---- record prefetch ------------------------------------------------------
typedef __attribute__((address_space(3))) void* LdsPtr;
typedef __attribute__((address_space(1))) void* GblPtr;

// Steps of prefetch kept in flight: ~12 KB of records per wave, 1..3 steps.
__host__ __device__ inline int PrefetchDepth(const IndexView& ix) {
    const int d = static_cast<int>(12288 / (2 * ix.batch_bytes));
    return d < 1 ? 1 : (d > 3 ? 3 : d);
}

// Copies `bytes` (a multiple of 64) global -> LDS with direct-to-LDS loads:
// one instruction per kilobyte, lane i moving bytes [16i, 16i+16) of it.
// Every instruction has at least one active lane, so each call issues exactly
// ceil(bytes / 1024) loads (the count the vmcnt waits rely on).
__device__ __forceinline__ void CopyToLdsAsync(const uint8_t* src, uint8_t* dst, uint32_t bytes, int lane) {
    for (uint32_t off = 0; off < bytes; off += 1024) {
        const uint32_t o = off + static_cast<uint32_t>(lane) * 16;
        if (o < bytes)
            __builtin_amdgcn_global_load_lds((GblPtr)(src + o), (LdsPtr)(dst + off), 16, 0, 0);
    }
}

// s_waitcnt on vmcnt only (expcnt / lgkmcnt fields at their no-wait maxima).
template <int N>
__device__ __forceinline__ void WaitVmN() {
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Waits until at most n vector-memory operations are outstanding; values
// beyond the table wait for all (always safe, only slower).
__device__ __forceinline__ void WaitVm(int n) {
    switch (n) {
        case 1: WaitVmN<1>(); break;
        case 2: WaitVmN<2>(); break;
        case 3: WaitVmN<3>(); break;
        case 4: WaitVmN<4>(); break;
        case 5: WaitVmN<5>(); break;
        case 6: WaitVmN<6>(); break;
        case 7: WaitVmN<7>(); break;
        case 8: WaitVmN<8>(); break;
        case 9: WaitVmN<9>(); break;
        case 10: WaitVmN<10>(); break;
        case 11: WaitVmN<11>(); break;
        case 12: WaitVmN<12>(); break;
        default: WaitVmN<0>(); break;
    }
}

// Producer side of the step sequence: the same walk as the consumer loop
// (probes in order, empty clusters skipped, two batches per step).
struct StepSeq {
    uint32_t p;              // next probe to open
    uint32_t b;              // next batch in the open cluster
    uint32_t nb;             // batches of the open cluster
    const uint8_t* arena_p;  // its scan records
};

__device__ __forceinline__ const uint8_t* NextStep(StepSeq& s, const IndexView& ix, const PlainArgs& a, uint64_t qi,
                                                   uint32_t np) {
    while (s.b >= s.nb) {
        if (s.p >= np) return nullptr;
        const uint32_t cid = a.probe_ids[qi * a.maxp + s.p];
        const uint64_t cnt = ix.cluster_begin[cid + 1] - ix.cluster_begin[cid];
        s.nb = static_cast<uint32_t>((cnt + 31) / 32);
        s.arena_p = ix.arena + ix.arena_begin[cid];
        s.b = 0;
        ++s.p;
    }
    const uint8_t* src = s.arena_p + static_cast<uint64_t>(s.b) * ix.batch_bytes;
    s.b += 2;
    return src;
}

// One wave per query.  Each step covers two 32-vector batches (lanes 0..31
// and 32..63): coarse integer sums from the query LUT in LDS, the
// threshold-free part of the ZPR bound, and a speculative evaluation of both
// batches against the threshold at the step's start.  When the survivors of
// the step fit the staging area without a flush, those decisions are the
// reference's decisions (the threshold cannot move before the next flush)
// and they are applied at once; otherwise the step is replayed batch by
// batch, lane by lane, exactly as CollectLanes does.
__global__ __launch_bounds__(kWave) void PlainSearchKernel(IndexView ix, PlainArgs a) {
    __shared__ PlainShared S;
    extern __shared__ float dyn[];
    float* qv = dyn;                                                    // padded floats
    uint8_t* lutc = reinterpret_cast<uint8_t*>(dyn + ix.padded);        // blocks*32 bytes
    float* ltab = reinterpret_cast<float*>(lutc + static_cast<size_t>(ix.blocks) * 32);  // level table (b <= 8)
    StagedRow* staged_rows = reinterpret_cast<StagedRow*>(ltab + (ix.bits <= 8 ? 512 : 0));
    uint8_t* rowbuf = reinterpret_cast<uint8_t*>(staged_rows + a.flush_every);
    uint8_t* ring = rowbuf + ((32 * static_cast<size_t>(ix.blocks) * ix.bits + 15) & ~size_t{15});
    const uint32_t step_bytes = 2 * ix.batch_bytes;   // one step = two scan records
    const int depth = PrefetchDepth(ix);
    const uint32_t nring = static_cast<uint32_t>(depth) + 1;
    const int ni = static_cast<int>((step_bytes + 1023) / 1024);   // load instructions per step
    const int lane = threadIdx.x;
    if (ix.bits <= 8)
        for (int i = lane; i < 512; i += kWave) ltab[i] = ix.refine_tab[i];
    const float* tab = ix.bits <= 8 ? ltab : ix.refine_tab;
    const uint32_t j = lane >> 5, t = lane & 31;
    const int k = static_cast<int>(a.k);
    const int fe = a.flush_every;
    const float inf = __builtin_inff();
    const uint32_t P = static_cast<uint32_t>(ix.padded);
    const uint32_t E = static_cast<uint32_t>(ix.blocks) * 32;
    for (uint64_t qi = blockIdx.x; qi < a.nq; qi += gridDim.x) {
        for (uint32_t i = lane; i < P; i += kWave) qv[i] = a.rotq[qi * P + i];
        for (uint32_t i = lane; i < E; i += kWave) lutc[i] = a.lut[qi * E + i];
        if (lane == 0) S.limit = inf;
        __syncthreads();
        const float* cst = a.consts + qi * 8;
        const float delta = cst[0], base = cst[1], lut_err = cst[2], score_max = cst[3];
        int cnt = 0, staged = 0;
        uint64_t n_cand = 0, n_surv = 0, n_batch = 0;
        const uint32_t np = min(static_cast<uint32_t>(ix.nlist), a.nprobe_q ? a.nprobe_q[qi] : a.nprobe_all);
        // Prefetch ring: the records of the next `depth` steps stream into LDS
        // while the current step is scored.
        StepSeq seq{0, 0, 0, nullptr};
        uint32_t issued = 0, consumed = 0;
        for (int d = 0; d < depth; ++d) {
            const uint8_t* src = NextStep(seq, ix, a, qi, np);
            if (!src) break;
            CopyToLdsAsync(src, ring + (issued % nring) * step_bytes, step_bytes, lane);
            ++issued;
        }
        for (uint32_t p = 0; p < np; ++p) {
            const uint32_t cid = a.probe_ids[qi * a.maxp + p];
            const float d2 = a.probe_d2[qi * a.maxp + p];
            const uint64_t vb = ix.cluster_begin[cid];
            const uint32_t ccount = static_cast<uint32_t>(ix.cluster_begin[cid + 1] - vb);
            if (ccount == 0) continue;
            const float sqrt_d2 = sqrtf(fmaxf(0.0F, d2));
            const uint32_t nb = (ccount + 31) / 32;
            const uint8_t* arena_c = ix.arena + ix.arena_begin[cid];
            const uint8_t* ref_c = ix.ref_arena + ix.ref_arena_begin[cid];
            for (uint32_t b0 = 0; b0 < nb; b0 += 2) {
                const uint32_t jn = min(2u, nb - b0);
                const uint32_t bj = b0 + j;
                // This step's records were requested `issued - consumed - 1`
                // steps of loads ago: wait for them, then request the step
                // `depth` ahead into the slot the previous step just freed.
                WaitVm(ni * static_cast<int>(issued - consumed - 1));
                __syncthreads();
                const uint8_t* recs = ring + (consumed % nring) * step_bytes;
                if (const uint8_t* src = NextStep(seq, ix, a, qi, np)) {
                    CopyToLdsAsync(src, ring + (issued % nring) * step_bytes, step_bytes, lane);
                    ++issued;
                }
                ++consumed;
                const float L = S.limit;
                bool pass = false;
                if (j < jn) {
                    const uint8_t* rec = recs + j * ix.batch_bytes;
                    int32_t acc = 0;
#pragma unroll 8
                    for (int blk = 0; blk < ix.blocks; ++blk) {
                        const uint32_t h = rec[kHdr + blk * 32 + t];
                        acc += lutc[blk * 32 + (h & 15)] + lutc[blk * 32 + 16 + (h >> 4)];
                    }
                    const float* sc = reinterpret_cast<const float*>(rec + 96);
                    const float pad_w_round = (0.5F * sc[3]) * (score_max + lut_err);
                    const float vbase = (sc[0] + d2) - pad_w_round;
                    const float scv = fmaf(static_cast<float>(acc), delta, base);
                    const float bt = fmaf(static_cast<float>(rec[t]), sc[1], vbase);
                    const float w = fmaf(static_cast<float>(rec[32 + t]), sc[3], sc[2]);
                    const float pp = fmaf(static_cast<float>(rec[64 + t]), sc[5], sc[4]);
                    const float dc = fmaf(-w, scv, bt);
                    const float t1 = fmaf(-fabsf(w), lut_err, dc);
                    S.t1[lane] = t1;
                    S.pp[lane] = pp;
                    int32_t mx = acc, mn = acc;
                    for (int o = 16; o > 0; o >>= 1) {
                        mx = max(mx, __shfl_xor(mx, o, 32));
                        mn = min(mn, __shfl_xor(mn, o, 32));
                    }
                    if (t == 0) {
                        S.amax[j] = mx;
                        S.amin[j] = mn;
                    }
                    const float rh = L < d2 ? sqrtf(fmaxf(0.0F, L)) : sqrt_d2;
                    bool skip = false;
                    if (!a.disable_prune && L < inf) {
                        const float pad_add = fmaf(lut_err, sc[6], pad_w_round);
                        const float perp_amax = fmaf(sc[5], 255.0F, sc[4]);
                        const float s_hi = fmaf(static_cast<float>(mx), delta, base);
                        const float s_lo = fmaf(static_cast<float>(mn), delta, base);
                        const float sa = fabsf(s_hi) < fabsf(s_lo) ? fabsf(s_lo) : fabsf(s_hi);
                        const float lb = fmaf(-perp_amax, rh * a.z, fmaf(-sc[6], sa, (sc[0] + d2) - pad_add));
                        skip = lb >= L;
                    }
                    const uint32_t n_logic = min(32u, ccount - bj * 32);
                    const float dlo = fmaf(-(a.z * rh), pp, t1);
                    pass = !skip && t < n_logic && (a.disable_prune || dlo < L);
                }
                const uint64_t bal = __ballot(pass);
                const uint32_t pm0 = static_cast<uint32_t>(bal), pm1 = static_cast<uint32_t>(bal >> 32);
                const int total = __popc(pm0) + __popc(pm1);
                if (staged + total <= fe) {
                    // No flush inside the step: apply the speculative decisions.
                    if (pass) {
                        const uint32_t m = j ? pm1 : pm0;
                        const int rank = staged + (j ? __popc(pm0) : 0) + __popc(m & ((1u << t) - 1u));
                        staged_rows[rank] = MakeRow(ix, ref_c, vb, bj, t, d2);
                    }
                    staged += total;
                    n_surv += total;
                    n_batch += jn;
                    n_cand += min(ccount - b0 * 32, jn * 32);
                } else {
                    __syncthreads();
                    // Exact replay, batch by batch (a flush lands inside this step).
                    for (uint32_t jj = 0; jj < jn; ++jj) {
                        const uint32_t bb = b0 + jj, off = bb * 32;
                        const uint32_t n_logic = min(32u, ccount - off);
                        n_cand += n_logic;
                        n_batch += 1;
                        const float limit0 = S.limit;
                        const float rh = limit0 < d2 ? sqrtf(fmaxf(0.0F, limit0)) : sqrt_d2;
                        if (!a.disable_prune && limit0 < inf) {
                            const float* sc = reinterpret_cast<const float*>(recs + jj * ix.batch_bytes + 96);
                            const float pad_w_round = (0.5F * sc[3]) * (score_max + lut_err);
                            const float pad_add = fmaf(lut_err, sc[6], pad_w_round);
                            const float perp_amax = fmaf(sc[5], 255.0F, sc[4]);
                            const float s_hi = fmaf(static_cast<float>(S.amax[jj]), delta, base);
                            const float s_lo = fmaf(static_cast<float>(S.amin[jj]), delta, base);
                            const float sa = fabsf(s_hi) < fabsf(s_lo) ? fabsf(s_lo) : fabsf(s_hi);
                            const float lb =
                                fmaf(-perp_amax, rh * a.z, fmaf(-sc[6], sa, (sc[0] + d2) - pad_add));
                            if (lb >= limit0) continue;
                        }
                        const float slack = a.z * rh;
                        const bool in32 = lane < 32;
                        const float dlo = in32 ? fmaf(-slack, S.pp[jj * 32 + lane], S.t1[jj * 32 + lane]) : inf;
                        const uint32_t lmask = n_logic >= 32 ? 0xFFFFFFFFu : ((1u << n_logic) - 1u);
                        uint32_t mask = a.disable_prune
                                            ? lmask
                                            : (static_cast<uint32_t>(__ballot(in32 && dlo < limit0)) & lmask);
                        while (mask != 0) {
                            const float Lc = S.limit;
                            const uint32_t ps =
                                a.disable_prune ? mask : (mask & static_cast<uint32_t>(__ballot(in32 && dlo < Lc)));
                            if (ps == 0) break;
                            const int room = fe - staged;
                            const int pc = __popc(ps);
                            const int take = pc <= room ? pc : room;
                            if (in32 && ((ps >> lane) & 1u)) {
                                const int rank = __popc(ps & ((1u << lane) - 1u));
                                if (rank < take) staged_rows[staged + rank] = MakeRow(ix, ref_c, vb, bb, lane, d2);
                            }
                            staged += take;
                            n_surv += take;
                            if (pc <= room) break;
                            // The next passing lane meets a full staging area: flush,
                            // then stage it without a re-test (CollectLanes' order).
                            uint32_t rest = ps;
                            for (int q = 0; q < take; ++q) rest &= rest - 1;
                            const int pl = __ffs(rest) - 1;
                            FlushRows(S, staged_rows, staged, cnt, k, ix, qv, arena_c, ref_c, rowbuf, tab);
                            if (lane == pl) staged_rows[0] = MakeRow(ix, ref_c, vb, bb, pl, d2);
                            staged = 1;
                            n_surv += 1;
                            mask &= (pl >= 31) ? 0u : ~((2u << pl) - 1u);
                        }
                    }
                }
                __syncthreads();
            }
            if (staged > 0) {
                FlushRows(S, staged_rows, staged, cnt, k, ix, qv, arena_c, ref_c, rowbuf, tab);
                staged = 0;
            }
        }
        for (int i = lane; i < k; i += kWave) {
            a.out_ids[qi * a.k + i] = i < cnt ? S.ti[i] : 0xFFFFFFFFu;
            a.out_dists[qi * a.k + i] = i < cnt ? S.td[i] : inf;
        }
        if (lane == 0 && a.stats != nullptr) {
            atomicAdd(&a.stats[0], static_cast<unsigned long long>(n_cand));
            atomicAdd(&a.stats[1], static_cast<unsigned long long>(n_surv));
            atomicAdd(&a.stats[2], static_cast<unsigned long long>(n_batch));
        }
        __syncthreads();
    }
}

size_t PlainSearchLds(const IndexView& ix, int flush_every) {
    return ix.padded * sizeof(float) + static_cast<size_t>(ix.blocks) * 32 + (ix.bits <= 8 ? 512 * sizeof(float) : 0) +
           static_cast<size_t>(flush_every) * sizeof(StagedRow) +
           ((32 * static_cast<size_t>(ix.blocks) * static_cast<size_t>(ix.bits) + 15) & ~size_t{15}) +
           (PrefetchDepth(ix) + 1) * 2 * static_cast<size_t>(ix.batch_bytes);
}

hipError_t LaunchPlainSearch(hipStream_t st, const IndexView& ix, const PlainArgs& a) {
    if (a.nq == 0) return hipSuccess;
    if (a.flush_every < 1 || a.flush_every > kMaxFlush) return hipErrorInvalidValue;
    const uint32_t grid = static_cast<uint32_t>(a.nq < 262144 ? a.nq : 262144);
    const size_t dyn = PlainSearchLds(ix, a.flush_every);
    if (dyn > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&PlainSearchKernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(dyn));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(PlainSearchKernel, dim3(grid), dim3(kWave), dyn, st, ix, a);
    return hipGetLastError();
}

// Coarse accumulate on caller-provided batches and LUTs (the integer kernel
// on its own, for the exactness tests).  lut is the full 64-byte-per-block
// reference layout.
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
