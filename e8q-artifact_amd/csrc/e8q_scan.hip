// e8q_scan.hip -- the plain search kernel: one workgroup per query, the
// reference's per-query schedule reproduced bit for bit.
//
// Reference: E8qIndex::Search (e8q_search.cpp:215-293) with ScanClusterAvx2
// (e8q_search_avx2.cpp:407-598), CollectLanes (e8q_scan_common.hpp:87-113),
// FlushSurvivors (e8q_search.cpp:182-207), RefineCoreAvx2 (e8q_search_avx2.cpp:
// 82-403), InsertHeap/ExtractTopK (e8q_search.cpp:135-180, 295-318).
//
// Layout per query: the probed clusters are visited in probe order; each
// cluster's 32-vector batches are scored 8 at a time (256 threads = 8 batches
// x 32 lanes: coarse integer accumulate + the threshold-independent part of
// the ZPR bound), then walked in order by the whole workgroup in lockstep:
// batch skip, lane mask, and survivor staging against the live top-k
// threshold, refining the staged rows (32 rows x 8 components = 256 threads)
// whenever flush_every rows are waiting and at the end of every cluster.
// The top-k is a sorted (dist, id) list in LDS; its k-th entry is the
// reference heap's root, the prune threshold.
#include "e8q_device.h"
#include "e8q_kernels.h"

namespace e8qdev {

constexpr int kPlainThreads = 256;
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
    float t1[kPlainThreads];     // ZPR bound minus the slack term, per lane
    float pp[kPlainThreads];     // perp (Omega) per lane
    int32_t amax[8], amin[8];    // batch max/min of the coarse sums
    StagedRow staged[kMaxFlush];
    float td[kMaxK], nd[kMaxK];  // sorted top-k (dist, id), merge target
    uint32_t ti[kMaxK], ni[kMaxK];
    float cd[32];
    uint32_t ci[32];
    float limit;
};

__device__ void FlushRows(PlainShared& S, int ns, int& cnt, int k, const IndexView& ix, const float* qv,
                          const uint8_t* arena_c, const uint8_t* ref_c) {
    const int tid = threadIdx.x;
    for (int rb = 0; rb < ns; rb += 32) {
        __syncthreads();
        const int nn = min(32, ns - rb);
        const int r = rb + (tid >> 3);
        const int i = tid & 7;
        if (r < ns) {
            const StagedRow row = S.staged[r];
            const uint32_t bj = row.off >> 5, lane = row.off & 31;
            const uint8_t* plane = arena_c + static_cast<uint64_t>(bj) * ix.batch_bytes + kHdr;
            const uint8_t* field = ref_c + static_cast<uint64_t>(bj) * ix.ref_batch_bytes + lane * ix.refine_stride;
            const int c = ix.bits - 1;
            float acc = 0.0F, acc2 = 0.0F;
            for (int blk = 0; blk < ix.blocks; ++blk) {
                uint32_t digit;
                int br;
                DecodeDigit(ix.bits, plane[blk * 32 + lane], field + blk * c, i, &digit, &br);
                const float lv = ix.refine_tab[static_cast<uint32_t>(br) * ix.tab_stride + digit];
                if (blk & 1) acc2 = fmaf(lv, qv[blk * 8 + i], acc2);
                else acc = fmaf(lv, qv[blk * 8 + i], acc);
            }
            const float sc = Hsum8Lanes(acc + acc2);
            if (i == 0) {
                S.cd[r - rb] = fmaf(-(2.0F * row.g1), sc, row.h3);
                S.ci[r - rb] = row.id;
            }
        }
        __syncthreads();
        // Merge the nn new candidates into the sorted top-k list.
        if (tid < cnt) {
            const float d = S.td[tid];
            const uint32_t id = S.ti[tid];
            int rank = tid;
            for (int j = 0; j < nn; ++j) rank += Better(S.cd[j], S.ci[j], d, id) ? 1 : 0;
            if (rank < k) {
                S.nd[rank] = d;
                S.ni[rank] = id;
            }
        } else if (tid >= kMaxK && tid < kMaxK + nn) {
            const int j = tid - kMaxK;
            const float d = S.cd[j];
            const uint32_t id = S.ci[j];
            int lo = 0, hi = cnt;   // # of list entries better than (d, id)
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (Better(S.td[mid], S.ti[mid], d, id)) lo = mid + 1;
                else hi = mid;
            }
            int rank = lo;
            for (int j2 = 0; j2 < nn; ++j2) rank += Better(S.cd[j2], S.ci[j2], d, id) ? 1 : 0;
            if (rank < k) {
                S.nd[rank] = d;
                S.ni[rank] = id;
            }
        }
        __syncthreads();
        cnt = min(k, cnt + nn);
        if (tid < cnt) {
            S.td[tid] = S.nd[tid];
            S.ti[tid] = S.ni[tid];
        }
        __syncthreads();
        if (tid == 0) S.limit = cnt == k ? S.td[k - 1] : __builtin_inff();
    }
    __syncthreads();
}

__global__ __launch_bounds__(kPlainThreads) void PlainSearchKernel(IndexView ix, PlainArgs a) {
    __shared__ PlainShared S;
    extern __shared__ float dyn[];
    float* qv = dyn;                                              // padded floats
    uint8_t* lutc = reinterpret_cast<uint8_t*>(dyn + ix.padded);  // blocks*32 bytes
    const int tid = threadIdx.x;
    const int lane64 = tid & 63;
    const int k = static_cast<int>(a.k);
    const float inf = __builtin_inff();
    const uint32_t P = static_cast<uint32_t>(ix.padded);
    const uint32_t E = static_cast<uint32_t>(ix.blocks) * 32;
    for (uint64_t qi = blockIdx.x; qi < a.nq; qi += gridDim.x) {
        for (uint32_t i = tid; i < P; i += blockDim.x) qv[i] = a.rotq[qi * P + i];
        for (uint32_t i = tid; i < E; i += blockDim.x) lutc[i] = a.lut[qi * E + i];
        if (tid == 0) S.limit = inf;
        __syncthreads();
        const float* cst = a.consts + qi * 8;
        const float delta = cst[0], base = cst[1], lut_err = cst[2], score_max = cst[3];
        int cnt = 0, staged = 0;
        uint64_t n_cand = 0, n_surv = 0, n_batch = 0;
        const uint32_t np = min(static_cast<uint32_t>(ix.nlist), a.nprobe_q ? a.nprobe_q[qi] : a.nprobe_all);
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
            for (uint32_t b0 = 0; b0 < nb; b0 += 8) {
                // ---- parallel: coarse sums + threshold-free bound terms
                {
                    const uint32_t j = tid >> 5, t = tid & 31, bj = b0 + j;
                    if (bj < nb) {
                        const uint8_t* rec = arena_c + static_cast<uint64_t>(bj) * ix.batch_bytes;
                        int32_t acc = 0;
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
                        S.t1[tid] = fmaf(-fabsf(w), lut_err, dc);
                        S.pp[tid] = pp;
                        int32_t mx = acc, mn = acc;
                        for (int o = 16; o > 0; o >>= 1) {
                            mx = max(mx, __shfl_xor(mx, o, 32));
                            mn = min(mn, __shfl_xor(mn, o, 32));
                        }
                        if (t == 0) {
                            S.amax[j] = mx;
                            S.amin[j] = mn;
                        }
                    }
                }
                __syncthreads();
                // ---- sequential walk over the chunk's batches (uniform control flow)
                const uint32_t jn = min(8u, nb - b0);
                for (uint32_t j = 0; j < jn; ++j) {
                    const uint32_t bj = b0 + j, off = bj * 32;
                    const uint32_t n_logic = min(32u, ccount - off);
                    n_cand += n_logic;
                    n_batch += 1;
                    const float limit0 = S.limit;
                    const float rh = limit0 < d2 ? sqrtf(fmaxf(0.0F, limit0)) : sqrt_d2;
                    if (!a.disable_prune && limit0 < inf) {
                        const float* sc =
                            reinterpret_cast<const float*>(arena_c + static_cast<uint64_t>(bj) * ix.batch_bytes + 96);
                        const float pad_w_round = (0.5F * sc[3]) * (score_max + lut_err);
                        const float pad_add = fmaf(lut_err, sc[6], pad_w_round);
                        const float perp_amax = fmaf(sc[5], 255.0F, sc[4]);
                        const float s_hi = fmaf(static_cast<float>(S.amax[j]), delta, base);
                        const float s_lo = fmaf(static_cast<float>(S.amin[j]), delta, base);
                        const float sa = fabsf(s_hi) < fabsf(s_lo) ? fabsf(s_lo) : fabsf(s_hi);
                        const float lb = fmaf(-perp_amax, rh * a.z, fmaf(-sc[6], sa, (sc[0] + d2) - pad_add));
                        if (lb >= limit0) continue;
                    }
                    const float slack = a.z * rh;
                    const bool in32 = lane64 < 32;
                    const float dlo = in32 ? fmaf(-slack, S.pp[j * 32 + lane64], S.t1[j * 32 + lane64]) : inf;
                    const uint32_t lmask = n_logic >= 32 ? 0xFFFFFFFFu : ((1u << n_logic) - 1u);
                    uint32_t mask =
                        a.disable_prune ? lmask : (static_cast<uint32_t>(__ballot(in32 && dlo < limit0)) & lmask);
                    while (mask != 0) {
                        const float L = S.limit;
                        const uint32_t pass =
                            a.disable_prune ? mask : (mask & static_cast<uint32_t>(__ballot(in32 && dlo < L)));
                        if (pass == 0) break;
                        const int room = a.flush_every - staged;
                        const int pc = __popc(pass);
                        // Lanes staged before any flush: the first min(pc, room) passing lanes.
                        const int take = pc <= room ? pc : room;
                        if (tid < 32 && ((pass >> tid) & 1u)) {
                            const int rank = __popc(pass & ((1u << tid) - 1u));
                            if (rank < take) {
                                const uint8_t* rr = ref_c + static_cast<uint64_t>(bj) * ix.ref_batch_bytes;
                                const float* g1b = reinterpret_cast<const float*>(rr + ix.refine_stride * 32);
                                const float* h3b = g1b + 32;
                                S.staged[staged + rank] = StagedRow{off + tid, ix.ids[vb + off + tid], g1b[tid],
                                                                    d2 + h3b[tid]};
                            }
                        }
                        staged += take;
                        n_surv += take;
                        if (pc <= room) {
                            mask = 0;
                            break;
                        }
                        // The next passing lane finds the staging full: flush, then
                        // stage it without a re-test (CollectLanes' order).
                        uint32_t rest = pass;
                        for (int t = 0; t < take; ++t) rest &= rest - 1;
                        const int pl = __ffs(rest) - 1;
                        FlushRows(S, staged, cnt, k, ix, qv, arena_c, ref_c);
                        staged = 0;
                        if (tid == pl) {
                            const uint8_t* rr = ref_c + static_cast<uint64_t>(bj) * ix.ref_batch_bytes;
                            const float* g1b = reinterpret_cast<const float*>(rr + ix.refine_stride * 32);
                            const float* h3b = g1b + 32;
                            S.staged[0] = StagedRow{off + pl, ix.ids[vb + off + pl], g1b[pl], d2 + h3b[pl]};
                        }
                        staged = 1;
                        n_surv += 1;
                        mask &= (pl >= 31) ? 0u : ~((2u << pl) - 1u);
                    }
                }
                __syncthreads();
            }
            if (staged > 0) {
                FlushRows(S, staged, cnt, k, ix, qv, arena_c, ref_c);
                staged = 0;
            }
        }
        for (int i = tid; i < k; i += blockDim.x) {
            a.out_ids[qi * a.k + i] = i < cnt ? S.ti[i] : 0xFFFFFFFFu;
            a.out_dists[qi * a.k + i] = i < cnt ? S.td[i] : inf;
        }
        if (tid == 0 && a.stats != nullptr) {
            atomicAdd(&a.stats[0], static_cast<unsigned long long>(n_cand));
            atomicAdd(&a.stats[1], static_cast<unsigned long long>(n_surv));
            atomicAdd(&a.stats[2], static_cast<unsigned long long>(n_batch));
        }
        __syncthreads();
    }
}

hipError_t LaunchPlainSearch(hipStream_t st, const IndexView& ix, const PlainArgs& a) {
    if (a.nq == 0) return hipSuccess;
    const uint32_t grid = static_cast<uint32_t>(a.nq < 65536 ? a.nq : 65536);
    const size_t dyn = ix.padded * sizeof(float) + static_cast<size_t>(ix.blocks) * 32;
    hipLaunchKernelGGL(PlainSearchKernel, dim3(grid), dim3(kPlainThreads), dyn, st, ix, a);
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
