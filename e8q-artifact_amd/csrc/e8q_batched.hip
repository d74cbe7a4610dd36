// e8q_batched.hip -- the batched search form: every query of a call as one
// (query, probed cluster) work list.
//
// The plain kernel reproduces the reference's per-query schedule, whose
// prune threshold is the live top-k root and so serializes each query.  This
// form fixes the threshold per query instead:
//   1. threshold pass: the plain kernel over each query's first probe only;
//      its top-k there is the query's starting list and tau_q, its k-th
//      distance, the threshold the reference holds after that cluster;
//   2. the pairs (q, p), 1 <= p < nprobe(q), are grouped by cluster and cut
//      into items (a tile of the cluster's pairs x a range of its batches);
//      a persistent kernel takes items from an atomic counter, stages the
//      item's scan records in LDS once, and tests every pair of the tile
//      against them with its fixed tau_q -- the reference's coarse sums and
//      ZPR bound, the batch skip and the lane test -- appending the passing
//      lanes to the query's candidate list;
//   3. one wave per query refines its candidates with the plain kernel's
//      refine and merges them into its starting list.
// The reference's running threshold after its first cluster is at most
// tau_q and the bound only rises as the threshold falls, so every vector the
// reference refines after its first cluster is a candidate here too: the
// result's k-th distance is never worse than the reference's.  A query whose
// candidates overflow its list is searched again by the plain kernel.
#include <algorithm>

#include "e8q_scan_dev.h"

namespace e8qdev {

__global__ __launch_bounds__(256) void Fill32Kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) p[i] = v;
}

hipError_t LaunchFill32(hipStream_t st, uint32_t* p, uint64_t n, uint32_t v) {
    if (n == 0) return hipSuccess;
    const uint64_t g = (n + 255) / 256;
    hipLaunchKernelGGL(Fill32Kernel, dim3(static_cast<uint32_t>(g < 4096 ? g : 4096)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

__device__ __forceinline__ uint32_t NprobeOf(const IndexView& ix, const BatchedArgs& a, uint64_t q) {
    return min(static_cast<uint32_t>(ix.nlist), a.nprobe_q ? a.nprobe_q[q] : a.nprobe_all);
}

// ---------------------------------------------------------------------------
// Work list

// Pairs per cluster, and the candidate / batch counts of the pairs.
__global__ __launch_bounds__(256) void PairCountKernel(const BatchedParams* __restrict__ prm) {
    const IndexView& ix = prm->ix;
    const BatchedArgs& a = prm->a;
    __shared__ unsigned long long s_cand, s_nb;
    if (threadIdx.x == 0) s_cand = s_nb = 0;
    __syncthreads();
    unsigned long long cand = 0, nbs = 0;
    const uint64_t total = a.nq * a.maxp;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < total; i += gridDim.x * 256ull) {
        const uint64_t q = i / a.maxp;
        const uint32_t p = static_cast<uint32_t>(i - q * a.maxp);
        if (p < a.p_lo || p >= NprobeOf(ix, a, q) || (a.qmask != nullptr && a.qmask[q] == 0)) continue;
        const uint32_t c = a.probe_ids[i];
        const uint32_t cnt = static_cast<uint32_t>(ix.cluster_begin[c + 1] - ix.cluster_begin[c]);
        if (cnt == 0) continue;
        atomicAdd(&a.ccount[c], 1u);
        cand += cnt;
        nbs += (cnt + 31) / 32;
    }
    atomicAdd(&s_cand, cand);
    atomicAdd(&s_nb, nbs);
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&a.stats[0], s_cand);
        atomicAdd(&a.stats[2], s_nb);
    }
}

// Exclusive scans of the pair counts and of the item counts per cluster;
// one workgroup of 1024 threads, each owning a contiguous run of clusters.
__global__ __launch_bounds__(1024) void WorkScanKernel(const BatchedParams* __restrict__ prm) {
    const IndexView& ix = prm->ix;
    const BatchedArgs& a = prm->a;
    __shared__ uint32_t sp[1024], si[1024];
    const uint32_t L = static_cast<uint32_t>(ix.nlist);
    const uint32_t per = (L + 1023) / 1024;
    const uint32_t c0 = min(L, threadIdx.x * per), c1 = min(L, c0 + per);
    auto items_of = [&](uint32_t c) -> uint32_t {
        const uint32_t n = a.ccount[c];
        if (n == 0) return 0;
        const uint32_t nb = static_cast<uint32_t>((ix.cluster_begin[c + 1] - ix.cluster_begin[c] + 31) / 32);
        return ((n + a.qtile - 1) / a.qtile) * ((nb + a.btile - 1) / a.btile);
    };
    uint32_t tp = 0, ti = 0;
    for (uint32_t c = c0; c < c1; ++c) {
        tp += a.ccount[c];
        ti += items_of(c);
    }
    sp[threadIdx.x] = tp;
    si[threadIdx.x] = ti;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {   // inclusive Hillis-Steele scan
        const uint32_t vp = threadIdx.x >= o ? sp[threadIdx.x - o] : 0;
        const uint32_t vi = threadIdx.x >= o ? si[threadIdx.x - o] : 0;
        __syncthreads();
        sp[threadIdx.x] += vp;
        si[threadIdx.x] += vi;
        __syncthreads();
    }
    uint32_t op = sp[threadIdx.x] - tp, oi = si[threadIdx.x] - ti;
    for (uint32_t c = c0; c < c1; ++c) {
        a.coff[c] = op;
        a.ioff[c] = oi;
        a.cfill[c] = 0;
        op += a.ccount[c];
        oi += items_of(c);
    }
    if (threadIdx.x == 1023) {
        a.coff[L] = sp[1023];
        a.ioff[L] = si[1023];
        a.counters[0] = si[1023];
        a.counters[1] = 0;
        a.counters[2] = 0;
    }
}

__global__ __launch_bounds__(256) void PairScatterKernel(const BatchedParams* __restrict__ prm) {
    const IndexView& ix = prm->ix;
    const BatchedArgs& a = prm->a;
    const uint64_t total = a.nq * a.maxp;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < total; i += gridDim.x * 256ull) {
        const uint64_t q = i / a.maxp;
        const uint32_t p = static_cast<uint32_t>(i - q * a.maxp);
        if (p < a.p_lo || p >= NprobeOf(ix, a, q) || (a.qmask != nullptr && a.qmask[q] == 0)) continue;
        const uint32_t c = a.probe_ids[i];
        if (ix.cluster_begin[c + 1] == ix.cluster_begin[c]) continue;
        const uint32_t pos = atomicAdd(&a.cfill[c], 1u);
        a.pairs[a.coff[c] + pos] = (q << 32) | p;
    }
}

// Items of each cluster: pair tiles x batch ranges, pair-tile major.  One
// thread per item (a head list can hold 1e5 items): its list is the last one
// whose first item precedes it.
__global__ __launch_bounds__(256) void ItemsKernel(const BatchedParams* __restrict__ prm) {
    const IndexView& ix = prm->ix;
    const BatchedArgs& a = prm->a;
    const uint64_t L = ix.nlist;
    const uint64_t total = min<uint64_t>(a.ioff[L], a.items_cap);
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < total; i += gridDim.x * 256ull) {
        uint64_t lo = 0, hi = L;   // last c in [0, L) with ioff[c] <= i
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) >> 1;
            if (a.ioff[mid] <= i) lo = mid;
            else hi = mid;
        }
        const uint32_t c = static_cast<uint32_t>(lo);
        const uint32_t n = a.ccount[c];
        const uint32_t nb = static_cast<uint32_t>((ix.cluster_begin[c + 1] - ix.cluster_begin[c] + 31) / 32);
        const uint32_t nbt = (nb + a.btile - 1) / a.btile;
        const uint32_t local = static_cast<uint32_t>(i - a.ioff[c]);
        const uint32_t qt = local / nbt, bt = local - qt * nbt;
        BatchItem it{};
        it.c = c;
        it.p0 = a.coff[c] + qt * a.qtile;
        it.np = min(a.qtile, n - qt * a.qtile);
        it.b0 = bt * a.btile;
        it.nb = min(a.btile, nb - bt * a.btile);
        a.items[i] = it;
    }
}

hipError_t LaunchBatchedWorkList(hipStream_t st, const IndexView& ix, const BatchedArgs& a, BatchedParams* dparams,
                                 bool upload) {
    hipError_t e = hipSuccess;
    if (upload) {
        const BatchedParams hp{ix, a};
        e = hipMemcpyAsync(dparams, &hp, sizeof(hp), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return e;
    }
    const uint64_t total = a.nq * a.maxp;
    const uint32_t g = static_cast<uint32_t>(std::min<uint64_t>((total + 255) / 256, 2048));
    hipLaunchKernelGGL(PairCountKernel, dim3(g), dim3(256), 0, st, dparams);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(WorkScanKernel, dim3(1), dim3(1024), 0, st, dparams);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(PairScatterKernel, dim3(g), dim3(256), 0, st, dparams);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint32_t gc = static_cast<uint32_t>(std::min<uint64_t>((a.items_cap + 255) / 256, 4096));
    hipLaunchKernelGGL(ItemsKernel, dim3(std::max<uint32_t>(gc, 1)), dim3(256), 0, st, dparams);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The persistent scan.  256 threads = four waves; each item's records are
// staged in LDS once, each wave tests one pair of the tile at a time.
constexpr int kScanThreads = 256;

size_t BatchedScanLds(const IndexView& ix, uint32_t qtile, uint32_t btile) {
    const size_t E = static_cast<size_t>(ix.blocks) * 32;
    return static_cast<size_t>(btile) * ix.batch_bytes + qtile * E + qtile * 8 * sizeof(float) +
           2 * qtile * sizeof(uint32_t);
}

template <int NQW>
__global__ __launch_bounds__(kScanThreads) void BatchedScanKernel(const BatchedParams* __restrict__ prm) {
    const IndexView& ix = prm->ix;
    const BatchedArgs& a = prm->a;
    extern __shared__ __align__(16) uint8_t dyn[];
    __shared__ uint32_t s_it;
    const uint32_t BB = Uniform(ix.batch_bytes);
    const int B = static_cast<int>(Uniform(static_cast<uint32_t>(ix.blocks)));
    const uint32_t E = static_cast<uint32_t>(B) * 32;
    const uint32_t qtile = a.qtile, btile = a.btile;
    uint8_t* recs = dyn;                                           // btile scan records
    uint8_t* luts = dyn + static_cast<size_t>(btile) * BB;         // qtile LUTs
    float* qc = reinterpret_cast<float*>(luts + static_cast<size_t>(qtile) * E);   // qtile x 8 constants
    uint32_t* pq = reinterpret_cast<uint32_t*>(qc + qtile * 8);    // query of each pair
    uint32_t* pp = pq + qtile;                                     // probe position of each pair
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t j = lane >> 5, t = lane & 31;
    const float inf = __builtin_inff();
    const float z = a.z;
    const bool prune = !a.disable_prune;
    const uint32_t k = a.k, cap = a.cap;
    const uint32_t nitems = Uniform(static_cast<uint32_t>(min<uint64_t>(*G(a.counters), a.items_cap)));
    unsigned long long staged_recs = 0;
    for (;;) {
        if (threadIdx.x == 0) s_it = atomicAdd(&a.counters[1], 1u);
        __syncthreads();
        const uint32_t it = Uniform(s_it);
        if (it >= nitems) break;
        const E8Q_GLOBAL BatchItem* item = G(a.items) + it;
        const uint32_t c = Uniform(item->c), np = Uniform(item->np), nb = Uniform(item->nb), b0 = Uniform(item->b0);
        const uint32_t p0 = Uniform(item->p0);
        const uint32_t ccnt = static_cast<uint32_t>(G(ix.cluster_begin)[c + 1] - G(ix.cluster_begin)[c]);
        const E8Q_GLOBAL uint8_t* crec = G(ix.arena) + G(ix.arena_begin)[c] + static_cast<uint64_t>(b0) * BB;
        const uint32_t rbytes = nb * BB;
        for (uint32_t o = threadIdx.x * 16; o < rbytes; o += kScanThreads * 16)
            *reinterpret_cast<f32x4*>(recs + o) = *reinterpret_cast<const E8Q_GLOBAL f32x4*>(crec + o);
        for (uint32_t i = wave; i < np; i += 4) {
            const uint64_t pr = G(a.pairs)[p0 + i];
            const uint32_t q = static_cast<uint32_t>(pr >> 32), p = static_cast<uint32_t>(pr);
            const E8Q_GLOBAL u32x2* src = reinterpret_cast<const E8Q_GLOBAL u32x2*>(G(a.lut) + static_cast<uint64_t>(q) * E);
            for (uint32_t w = lane; w < E / 8; w += 64) reinterpret_cast<u32x2*>(luts + i * E)[w] = src[w];
            if (lane == 0) {
                const E8Q_GLOBAL float* cs = G(a.consts) + static_cast<uint64_t>(q) * 8;
                const float tau = G(a.first_dists)[static_cast<uint64_t>(q) * k + k - 1];
                const float d2 = G(a.probe_d2)[static_cast<uint64_t>(q) * a.maxp + p];
                const float sd2 = sqrtf(fmaxf(0.0F, d2));
                float* o = qc + i * 8;
                o[0] = cs[0];                               // delta
                o[1] = cs[1];                               // base
                o[2] = cs[2];                               // lut_err
                o[3] = cs[3] + cs[2];                       // score_max + lut_err
                o[4] = cs[5];                               // U
                o[5] = tau;
                o[6] = z * (tau < d2 ? sqrtf(fmaxf(0.0F, tau)) : sd2);   // z * r_horizon
                o[7] = d2;
                pq[i] = q;
                pp[i] = p;
            }
        }
        __syncthreads();
        for (uint32_t i = wave; i < np; i += 4) {
            const float* o = qc + i * 8;
            const float delta = UniformF(o[0]), base = UniformF(o[1]), lut_err = UniformF(o[2]);
            const float sml = UniformF(o[3]), U = UniformF(o[4]), tau = UniformF(o[5]);
            const float zrh = UniformF(o[6]), d2 = UniformF(o[7]);
            const uint32_t q = Uniform(pq[i]), p = Uniform(pp[i]);
            const uint8_t* lut = luts + i * E;
            const bool hot = prune && tau < inf;
            for (uint32_t bb = 0; bb < nb; bb += 2) {
                const uint32_t bj = bb + j;
                const bool act = bj < nb;
                const uint8_t* rec = recs + bj * BB;
                const float* sc = reinterpret_cast<const float*>(rec + 96);
                float lbp = 0.0F, pam = 0.0F, vbase = 0.0F;
                bool uskip = !act;
                if (act) {
                    const float pad_w_round = (0.5F * sc[3]) * sml;
                    vbase = (sc[0] + d2) - pad_w_round;
                    const float pad_add = fmaf(lut_err, sc[6], pad_w_round);
                    lbp = (sc[0] + d2) - pad_add;
                    pam = fmaf(sc[5], 255.0F, sc[4]);
                    if (hot) uskip = fmaf(-pam, zrh, fmaf(-sc[6], U, lbp)) >= tau;
                }
                if (__ballot(uskip) == ~uint64_t{0}) continue;
                float t1 = inf, ppv = 0.0F;
                uint32_t sabs = 0;
                if (act) {
                    const u32x2* rw = reinterpret_cast<const u32x2*>(rec + kHdr + t * static_cast<uint32_t>(B));
                    int32_t acc = 0;
                    if constexpr (NQW > 0) {
#pragma unroll
                        for (int w = 0; w < NQW; ++w) acc += Coarse8(rw[w], lut + w * 256);
                    } else {
                        for (int w = 0; w < B / 8; ++w) acc += Coarse8(rw[w], lut + w * 256);
                    }
                    const float scv = fmaf(static_cast<float>(acc), delta, base);
                    const float bt = fmaf(static_cast<float>(rec[t]), sc[1], vbase);
                    const float w = fmaf(static_cast<float>(rec[32 + t]), sc[3], sc[2]);
                    ppv = fmaf(static_cast<float>(rec[64 + t]), sc[5], sc[4]);
                    const float dc = fmaf(-w, scv, bt);
                    t1 = fmaf(-fabsf(w), lut_err, dc);
                    sabs = __float_as_uint(fabsf(scv));
                }
                const uint32_t hm = HalfMax(sabs);
                const float sa = __uint_as_float(j ? ReadLane(hm, 63) : ReadLane(hm, 31));
                if (act) lbp = fmaf(-sc[6], sa, lbp);
                const bool skip = hot && fmaf(-pam, zrh, lbp) >= tau;
                const uint32_t n_logic = act ? min(32u, ccnt - (b0 + bj) * 32) : 0u;
                const float dlo = fmaf(-zrh, ppv, t1);
                const bool pass = act && !skip && t < n_logic && (!prune || dlo < tau);
                const uint64_t bal = __ballot(pass);
                if (bal != 0) {
                    const uint32_t n = static_cast<uint32_t>(__popcll(bal));
                    uint32_t at = 0;
                    if (lane == 0) at = atomicAdd(&a.ncand[q], n);
                    at = __shfl(at, 0);
                    if (pass) {
                        const uint32_t slot = at + static_cast<uint32_t>(__popcll(bal & ((uint64_t{1} << lane) - 1)));
                        if (slot < cap)
                            Gw(a.cand)[static_cast<uint64_t>(q) * cap + slot] =
                                (static_cast<uint64_t>(p) << 32) | ((b0 + bj) * 32 + t);
                    }
                }
            }
        }
        if (threadIdx.x == 0) staged_recs += nb;
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&a.stats[3], staged_recs);
}

template <int NQW>
static hipError_t LaunchScanT(hipStream_t st, const BatchedParams* dp, uint32_t grid, size_t dyn) {
    if (dyn > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&BatchedScanKernel<NQW>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(dyn));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(BatchedScanKernel<NQW>, dim3(grid), dim3(kScanThreads), dyn, st, dp);
    return hipGetLastError();
}

hipError_t LaunchBatchedScan(hipStream_t st, const IndexView& ix, const BatchedArgs& a, const BatchedParams* dparams,
                             uint32_t grid) {
    if (ix.blocks % 8 != 0 || a.qtile == 0 || a.btile == 0 || (a.btile & 1)) return hipErrorInvalidValue;
    const size_t dyn = BatchedScanLds(ix, a.qtile, a.btile);
    if (dyn > 160 * 1024) return hipErrorInvalidValue;
    switch (ix.blocks / 8) {
        case 2: return LaunchScanT<2>(st, dparams, grid, dyn);
        case 12: return LaunchScanT<12>(st, dparams, grid, dyn);
        default: return LaunchScanT<0>(st, dparams, grid, dyn);
    }
}

// ---------------------------------------------------------------------------
// Refine and merge, one wave per query.

// Candidate rows of one query: (probe position, vector offset) pairs.
struct CandRows {
    const uint64_t* cand;
    const uint32_t* probe_ids;   // this query's row
    const float* probe_d2;
    const IndexView* ix;
    __device__ RowLoc operator()(int i) const {
        const uint64_t v = *G(cand + i);
        const uint32_t p = static_cast<uint32_t>(v >> 32);
        const uint32_t c = *G(probe_ids + p);
        return RowLoc{ix->arena + *G(ix->arena_begin + c), ix->ref_arena + *G(ix->ref_arena_begin + c),
                      static_cast<uint32_t>(v), *G(ix->cluster_begin + c), *G(probe_d2 + p)};
    }
};

template <bool GEN>
__global__ __launch_bounds__(kWave) void BatchedRefineKernel(const BatchedParams* __restrict__ prm) {
    const IndexView& ix = prm->ix;
    const BatchedArgs& a = prm->a;
    __shared__ PlainShared S;
    extern __shared__ __align__(16) uint8_t dyn[];
    float* qv = reinterpret_cast<float*>(dyn);                                      // padded
    uint8_t* rowbuf = reinterpret_cast<uint8_t*>(qv + ix.padded);                    // kGroup rows
    const int lane = threadIdx.x;
    if (ix.bits <= 8)
        for (int i = lane; i < 512; i += kWave) S.tab[i] = G(ix.refine_tab)[i];
    const int k = static_cast<int>(a.k);
    const uint32_t P = static_cast<uint32_t>(ix.padded);
    const float inf = __builtin_inff();
    for (uint64_t q = blockIdx.x; q < a.nq; q += gridDim.x) {
        const uint32_t n = Uniform(*G(a.ncand + q));
        if (n > a.cap) {   // searched again by the plain kernel
            if (lane == 0) a.overflow[atomicAdd(&a.counters[2], 1u)] = static_cast<uint32_t>(q);
            continue;
        }
        for (uint32_t i = lane; i < P; i += kWave) qv[i] = G(a.rotq)[q * P + i];
        int cnt = 0;
        for (int i = lane; i < k; i += kWave) {
            const uint32_t id = G(a.first_ids)[q * k + i];
            S.td[i] = G(a.first_dists)[q * k + i];
            S.ti[i] = id;
        }
        for (int i = lane; i < k; i += kWave) cnt += G(a.first_ids)[q * k + i] != 0xFFFFFFFFu ? 1 : 0;
        for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
        cnt = static_cast<int>(Uniform(static_cast<uint32_t>(cnt)));
        __syncthreads();
        float limit = cnt == k ? UniformF(S.td[k - 1]) : inf;
        const CandRows rows{a.cand + q * a.cap, a.probe_ids + q * a.maxp, a.probe_d2 + q * a.maxp, &ix};
        FlushRows<GEN>(S, static_cast<int>(n), cnt, limit, &ix, k, qv, rows, rowbuf);
        for (int i = lane; i < k; i += kWave) {
            Gw(a.out_ids)[q * k + i] = i < cnt ? S.ti[i] : 0xFFFFFFFFu;
            Gw(a.out_dists)[q * k + i] = i < cnt ? S.td[i] : inf;
        }
        if (lane == 0 && a.stats != nullptr) atomicAdd(&a.stats[1], static_cast<unsigned long long>(n));
        __syncthreads();
    }
}

hipError_t LaunchBatchedRefine(hipStream_t st, const IndexView& ix, const BatchedArgs& a,
                               const BatchedParams* dparams) {
    if (a.nq == 0) return hipSuccess;
    const size_t dyn = ix.padded * sizeof(float) + kGroup * static_cast<size_t>(ix.blocks) * ix.bits + 16;
    const uint32_t grid = static_cast<uint32_t>(a.nq < 262144 ? a.nq : 262144);
    auto go = [&](auto kern) -> hipError_t {
        if (dyn > 64 * 1024) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     static_cast<int>(dyn));
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kWave), dyn, st, dparams);
        return hipGetLastError();
    };
    if (ix.bits >= 2 && ix.bits <= 5) return go(BatchedRefineKernel<false>);
    return go(BatchedRefineKernel<true>);
}

}  // namespace e8qdev
