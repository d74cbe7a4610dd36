// e8q_query.hip -- query-side kernels: FHT-Kac-4 rotation, query LUT build
// and quantization, route distances, top-nprobe selection.
//
// Parity: rotation is add/sub/mul only (exact reproduction of the
// reference's butterfly order, rotator.cpp:53-80); the LUT and route
// arithmetic reproduce the reference's fused multiply-adds one for one
// (e8q_search.cpp:20-73, kernels_avx2.cpp:8-103).
#include "e8q_device.h"
#include "e8q_kernels.h"

namespace e8qdev {

// ---------------------------------------------------------------------------
// Rotation: one workgroup per row (RotateInLds, e8q_device.h).
__global__ void RotateRowsKernel(const float* __restrict__ x, uint64_t rows, uint32_t dim, IndexView ix,
                                 float* __restrict__ out, int center) {
    extern __shared__ float buf[];
    const uint32_t P = static_cast<uint32_t>(ix.padded);
    for (uint64_t row = blockIdx.x; row < rows; row += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) buf[i] = i < dim ? x[row * dim + i] : 0.0F;
        __syncthreads();
        RotateInLds(buf, ix);
        for (uint32_t i = threadIdx.x; i < P; i += blockDim.x)
            out[row * P + i] = center ? buf[i] - ix.mean_rot[i] : buf[i];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Query LUT (PrepareQuery): per block two 16-entry tables, A over dims 4..7
// (low nibble of the rowA byte) and B over dims 0..3 (high nibble), each
// entry a fp64 FMA chain over the four dims, stored as float.  Stored
// compactly as [blk][A16 | B16]; the reference's duplicated halves hold the
// same bytes.  Then the symmetric u8 quantization with the AVX2 quantizer.
// consts[q*8 + {0..4}] = delta, base, lut_err, score_max, qn2.
__global__ void QueryPrepKernel(const float* __restrict__ rotq, uint64_t nq, IndexView ix,
                                uint8_t* __restrict__ lut, float* __restrict__ consts) {
    extern __shared__ float cf[];
    __shared__ float red[64];
    const int B = ix.blocks;
    const uint32_t P = static_cast<uint32_t>(ix.padded);
    const uint32_t E = static_cast<uint32_t>(B) * 32;
    for (uint64_t qi = blockIdx.x; qi < nq; qi += gridDim.x) {
        const float* q = rotq + qi * P;
        float hmax = 0.0F;
        for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) {
            const uint32_t blk = e >> 5, w = e & 31, v = w & 15;
            const float* qb = q + blk * 8;
            const int dim0 = w < 16 ? 4 : 0;   // A: dims 4..7, B: dims 0..3
            double acc = 0.0;
            for (int h = 0; h < 4; ++h) {
                const int bit = (v >> (3 - h)) & 1;
                acc = fma(static_cast<double>(qb[dim0 + h]), ix.gate_mid[2 * (dim0 + h) + bit], acc);
            }
            const float f = static_cast<float>(acc);
            cf[e] = f;
            hmax = fmaxf(hmax, fabsf(f));
        }
        // block max of |cf| (max is order independent)
        for (int o = 32; o > 0; o >>= 1) hmax = fmaxf(hmax, __shfl_xor(hmax, o));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = hmax;
        __syncthreads();
        float hc = 0.0F;
        for (uint32_t w = 0; w < (blockDim.x + 63) / 64; ++w) hc = fmaxf(hc, red[w]);
        const float delta = hc > 0.0F ? 2.0F * hc / 255.0F : 1.0F;
        const float qinv = 1.0F / delta;
        const float qoff = hc / delta;
        auto quant = [&](uint32_t e) {
            const float xq = fmaf(cf[e], qinv, qoff);
            const float r = rintf(xq);
            int iv = (r >= -2147483648.0F && r < 2147483648.0F) ? static_cast<int>(r) : INT32_MIN;
            return iv < 0 ? 0 : (iv > 255 ? 255 : iv);
        };
        for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) lut[qi * E + e] = static_cast<uint8_t>(quant(e));
        // U: no code's coarse score S = acc*delta + base exceeds it in
        // magnitude, since acc lies between the sums of the 2*B tables'
        // minima and maxima and S is monotone in acc (the search kernels'
        // header-only batch and cluster skips use it).
        int tmin = 0, tmax = 0;
        for (uint32_t tb = threadIdx.x; tb < 2u * B; tb += blockDim.x) {
            int mn = 255, mx = 0;
            for (uint32_t v = 0; v < 16; ++v) {
                const int iv = quant((tb >> 1) * 32 + (tb & 1) * 16 + v);
                mn = min(mn, iv);
                mx = max(mx, iv);
            }
            tmin += mn;
            tmax += mx;
        }
        for (int o = 32; o > 0; o >>= 1) {
            tmin += __shfl_xor(tmin, o);
            tmax += __shfl_xor(tmax, o);
        }
        __syncthreads();
        if ((threadIdx.x & 63) == 0) {
            red[2 * (threadIdx.x >> 6)] = __int_as_float(tmin);
            red[2 * (threadIdx.x >> 6) + 1] = __int_as_float(tmax);
        }
        __syncthreads();
        if (threadIdx.x < 8) {
            // QueryNorm2Avx2: 8 FMA chains over d = l, l+8, ..., then Hsum.
            float a = 0.0F;
            for (uint32_t d = threadIdx.x; d < P; d += 8) a = fmaf(q[d], q[d], a);
            const float qn2 = Hsum8Lanes(a);
            if (threadIdx.x == 0) {
                float* c = consts + qi * 8;
                c[0] = delta;
                c[1] = -hc * static_cast<float>(2 * B);
                c[2] = delta * (sqrtf(static_cast<float>(B)) * 2.0F);
                c[3] = static_cast<float>(B) * fmaf(delta, 255.0F, hc * 2.0F);
                c[4] = qn2;
                int smin = 0, smax = 0;
                for (uint32_t w = 0; w < (blockDim.x + 63) / 64; ++w) {
                    smin += __float_as_int(red[2 * w]);
                    smax += __float_as_int(red[2 * w + 1]);
                }
                c[5] = fmaxf(fabsf(fmaf(static_cast<float>(smin), delta, c[1])),
                             fabsf(fmaf(static_cast<float>(smax), delta, c[1])));
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Route distances d2[q][c] = (qn2 + cn2) - 2*dot, dot in the AVX2
// RouteGemv order: lane l of a 32-lane group accumulates the FMA chain over
// d = l, l+32, ...; then ((a0+a1)+(a2+a3)) per 8-lane slice and the Hsum.
// Grid: x = centroid tiles of 8 kRouteCT (eight 32-lane groups of kRouteCT
// centroids), y = query tiles of kRouteQT (staged in LDS).  A lane runs the
// chains of its kRouteCT centroids against the kRouteQT queries at once: per
// dimension step kRouteCT centroid loads and kRouteQT LDS reads feed
// kRouteCT * kRouteQT FMAs.  Each (query, centroid) pair keeps its own chain
// and reduction order.
constexpr int kRouteCT = 8;   // centroids per 32-lane group

// The 32-lane reduction of the chains, lane 0 holding the dot product.
__device__ __forceinline__ float RouteReduce32(float v) {
    v = v + __shfl_down(v, 8, 32);    // lanes 0..7: a0+a1, lanes 16..23: a2+a3
    v = v + __shfl_down(v, 16, 32);   // lanes 0..7: (a0+a1)+(a2+a3)
    v = v + __shfl_down(v, 4, 32);    // lanes 0..3: s[i]+s[i+4]
    v = v + __shfl_down(v, 2, 32);    // lanes 0..1
    v = v + __shfl_down(v, 1, 32);    // lane 0
    return v;
}
constexpr int kRoutePf = 2;   // dimension steps of centroid loads in flight

__global__ void RouteNoDimKernel(RouteArgs ra) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ra.nq * ra.nlist) return;
    const uint64_t q = i / ra.nlist, c = i % ra.nlist;
    float d2 = fmaf(-0.0F, 2.0F, ra.cn2[c] + ra.qn2[q * ra.qn2_stride]);
    if (d2 < 0.0F) d2 = 0.0F;
    ra.dist[i] = d2;
}

template <int kRouteQT>
__global__ __launch_bounds__(256) void RouteDistKernel(RouteArgs ra) {
    extern __shared__ float qs[];   // kRouteQT * D
    const uint32_t D = ra.dim;
    const uint64_t q0 = static_cast<uint64_t>(blockIdx.y) * kRouteQT;
    const int nqt = static_cast<int>(min<uint64_t>(kRouteQT, ra.nq - q0));
    for (uint32_t i = threadIdx.x; i < kRouteQT * D; i += blockDim.x) {
        const uint32_t qi = i / D;
        qs[i] = qi < static_cast<uint32_t>(nqt) ? ra.rotq[(q0 + qi) * ra.qstride + (i % D)] : 0.0F;
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 31;
    const uint64_t c0 = (static_cast<uint64_t>(blockIdx.x) * (blockDim.x / 32) + (threadIdx.x >> 5)) * kRouteCT;
    const float* cc[kRouteCT];
#pragma unroll
    for (int u = 0; u < kRouteCT; ++u) cc[u] = ra.cent + (c0 + u < ra.nlist ? c0 + u : 0) * ra.cstride;
    float acc[kRouteCT][kRouteQT];
#pragma unroll
    for (int u = 0; u < kRouteCT; ++u)
#pragma unroll
        for (int t = 0; t < kRouteQT; ++t) acc[u][t] = 0.0F;
    // The chains step d = lane, lane + 32, ...: full steps (every lane in
    // range) with the centroid values loaded kRoutePf steps ahead in a
    // register ring, then the last partial step.
    auto step = [&](uint32_t d, const float* cv) {
#pragma unroll
        for (int t = 0; t < kRouteQT; ++t) {
            const float qv = qs[t * D + d];
#pragma unroll
            for (int u = 0; u < kRouteCT; ++u) acc[u][t] = fmaf(qv, cv[u], acc[u][t]);
        }
    };
    const uint32_t nfull = D / 32;
    uint32_t m = 0;
    if (nfull >= static_cast<uint32_t>(kRoutePf)) {
        float ring[kRoutePf][kRouteCT];
#pragma unroll
        for (int j = 0; j < kRoutePf; ++j)
#pragma unroll
            for (int u = 0; u < kRouteCT; ++u) ring[j][u] = cc[u][lane + 32 * j];
        for (; m + 2 * kRoutePf <= nfull; m += kRoutePf) {   // ring: steps m .. m + kRoutePf - 1
            float next[kRoutePf][kRouteCT];   // steps m + kRoutePf .., loaded before this group's FMAs
#pragma unroll
            for (int j = 0; j < kRoutePf; ++j)
#pragma unroll
                for (int u = 0; u < kRouteCT; ++u) next[j][u] = cc[u][lane + 32 * (m + j + kRoutePf)];
#pragma unroll
            for (int j = 0; j < kRoutePf; ++j) step(lane + 32 * (m + j), ring[j]);
#pragma unroll
            for (int j = 0; j < kRoutePf; ++j)
#pragma unroll
                for (int u = 0; u < kRouteCT; ++u) ring[j][u] = next[j][u];
        }
#pragma unroll
        for (int j = 0; j < kRoutePf; ++j) step(lane + 32 * (m + j), ring[j]);
        m += kRoutePf;
    }
    for (; m < nfull; ++m) {
        const uint32_t d = lane + 32 * m;
        float cv[kRouteCT];
#pragma unroll
        for (int u = 0; u < kRouteCT; ++u) cv[u] = cc[u][d];
        step(d, cv);
    }
    if (lane + 32 * nfull < D) {
        const uint32_t d = lane + 32 * nfull;
        float cv[kRouteCT];
#pragma unroll
        for (int u = 0; u < kRouteCT; ++u) cv[u] = cc[u][d];
        step(d, cv);
    }
#pragma unroll
    for (int u = 0; u < kRouteCT; ++u) {
        const uint64_t c = c0 + u;
#pragma unroll
        for (int t = 0; t < kRouteQT; ++t) {
            const float v = RouteReduce32(acc[u][t]);
            if (lane == 0 && c < ra.nlist && t < nqt) {
                const float qn2 = ra.qn2[(q0 + t) * ra.qn2_stride];
                float d2 = fmaf(-v, 2.0F, ra.cn2[c] + qn2);
                if (d2 < 0.0F) d2 = 0.0F;
                ra.dist[(q0 + t) * ra.nlist + c] = d2;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Top-nprobe per query: the nprobe smallest keys (dist bits << 32 | id) in
// ascending key order (ivf_route.cpp:9-27's partial_sort).  Radix select
// over the 64-bit key, 8 bits per pass from the top, then a bitonic sort of
// the selected keys.  One workgroup per query; dist row staged in LDS.
constexpr int kSelThreads = 256;

__device__ __forceinline__ uint64_t KeyOf(const uint32_t* bits, uint32_t c) {
    return (static_cast<uint64_t>(bits[c]) << 32) | c;
}

__global__ void SelectProbesKernel(const float* __restrict__ dist, uint64_t nq, uint32_t nlist,
                                   const uint32_t* __restrict__ nprobe_q, uint32_t nprobe_all, uint32_t maxp,
                                   uint32_t* __restrict__ probe_ids, float* __restrict__ probe_d2) {
    extern __shared__ uint32_t sm[];
    uint32_t* bits = sm;                                   // nlist
    uint64_t* sel = reinterpret_cast<uint64_t*>(sm + ((nlist + 1) & ~1u));   // pow2 >= maxp
    __shared__ uint32_t hist[256];
    __shared__ uint32_t s_pick[2];
    __shared__ uint32_t s_cnt;
    uint32_t sortn = 1;
    while (sortn < maxp) sortn <<= 1;
    for (uint64_t qi = blockIdx.x; qi < nq; qi += gridDim.x) {
        const uint32_t m = min(nlist, nprobe_q ? nprobe_q[qi] : nprobe_all);
        for (uint32_t c = threadIdx.x; c < nlist; c += blockDim.x) bits[c] = FloatBits(dist[qi * nlist + c]);
        __syncthreads();
        // Radix select: find the m-th smallest key K (1-based).
        uint64_t prefix = 0, pmask = 0;
        uint32_t need = m;
        for (int shift = 56; shift >= 0 && m > 0; shift -= 8) {
            for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
            __syncthreads();
            for (uint32_t c = threadIdx.x; c < nlist; c += blockDim.x) {
                const uint64_t key = KeyOf(bits, c);
                if ((key & pmask) == prefix) atomicAdd(&hist[(key >> shift) & 255], 1u);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t acc = 0, b = 0;
                for (; b < 256; ++b) {
                    if (acc + hist[b] >= need) break;
                    acc += hist[b];
                }
                s_pick[0] = b;
                s_pick[1] = need - acc;
            }
            __syncthreads();
            prefix |= static_cast<uint64_t>(s_pick[0]) << shift;
            pmask |= static_cast<uint64_t>(255) << shift;
            need = s_pick[1];
            __syncthreads();
        }
        const uint64_t K = prefix;   // keys are unique: exactly m keys <= K
        if (threadIdx.x == 0) s_cnt = 0;
        for (uint32_t i = threadIdx.x; i < sortn; i += blockDim.x) sel[i] = ~0ull;
        __syncthreads();
        if (m > 0)
            for (uint32_t c = threadIdx.x; c < nlist; c += blockDim.x) {
                const uint64_t key = KeyOf(bits, c);
                if (key <= K) sel[atomicAdd(&s_cnt, 1u)] = key;
            }
        __syncthreads();
        // Bitonic sort of sortn keys (ascending).
        for (uint32_t k2 = 2; k2 <= sortn; k2 <<= 1) {
            for (uint32_t j = k2 >> 1; j > 0; j >>= 1) {
                for (uint32_t i = threadIdx.x; i < sortn; i += blockDim.x) {
                    const uint32_t l = i ^ j;
                    if (l > i) {
                        const uint64_t a = sel[i], b = sel[l];
                        const bool up = (i & k2) == 0;
                        if ((a > b) == up) {
                            sel[i] = b;
                            sel[l] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (uint32_t i = threadIdx.x; i < maxp; i += blockDim.x) {
            const bool ok = i < m;
            probe_ids[qi * maxp + i] = ok ? static_cast<uint32_t>(sel[i]) : 0xFFFFFFFFu;
            probe_d2[qi * maxp + i] = ok ? __uint_as_float(static_cast<uint32_t>(sel[i] >> 32)) : 0.0F;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Two-stage probe selection, with the same result as RouteDistKernel +
// SelectProbesKernel.  Stage 1 (RouteApproxKernel): every (query, centroid)
// distance as an fp32 MFMA contraction, a = max(0, qn2 + cn2 - 2 dot).
// Stage 2 (SelectProbesTwoStageKernel), per query: the m-th smallest key
// over a; U = the largest a + E over those m, where E = errf (qn2 + cn2)
// bounds |a - d| for d the reference-order distance (both within
// (D/2 + D/32 + 13) 2^-24 (qn2 + cn2); errf is four times that); the m
// smallest d are at most U, so every centroid with d <= U has a - E <= U:
// those candidates get d recomputed in the reference's chain order
// (RouteDistKernel's chains and reduction), the rest the largest key, and
// the exact top-m is selected from the keys as before.

typedef float f32x16r __attribute__((ext_vector_type(16)));
constexpr int kApproxPf = 4;   // float4 steps of each operand loaded ahead

// Workgroup: 32 queries x 128 centroids, one 32 x 32 tile per wave; lane l
// supplies row l % 32 and dimensions h * D/2 + s (h = l / 32) of the K = 2
// operand, so each lane streams its own row as float4 loads.
__global__ __launch_bounds__(256) void RouteApproxKernel(RouteArgs ra) {
    __shared__ float s_qn2[32];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 31, h = lane >> 5;
    const uint32_t half = ra.dim / 2, n4 = half / 4;
    const uint64_t q0 = static_cast<uint64_t>(blockIdx.y) * 32;
    if (threadIdx.x < 32) s_qn2[threadIdx.x] = ra.qn2[min<uint64_t>(q0 + threadIdx.x, ra.nq - 1) * ra.qn2_stride];
    __syncthreads();
    const uint64_t c0 = static_cast<uint64_t>(blockIdx.x) * 128 + wave * 32;
    const uint64_t qa = min<uint64_t>(q0 + r, ra.nq - 1), ca = min<uint64_t>(c0 + r, ra.nlist - 1);
    const float4* qp = reinterpret_cast<const float4*>(ra.rotq + qa * ra.qstride + h * half);
    const float4* cp = reinterpret_cast<const float4*>(ra.cent + ca * ra.cstride + h * half);
    f32x16r acc = {};
    auto step = [&](const float4& qv, const float4& cv) {
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(qv.x, cv.x, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(qv.y, cv.y, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(qv.z, cv.z, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(qv.w, cv.w, acc, 0, 0, 0);
    };
    // Two register sets of kApproxPf steps each: a set's loads are issued a
    // whole set ahead of its MFMAs (no copies between iterations).
    uint32_t s4 = 0;
    if (n4 >= 2 * kApproxPf) {
        float4 qa[kApproxPf], ca[kApproxPf], qb[kApproxPf], cb[kApproxPf];
        auto load = [&](float4* qd, float4* cd, uint32_t at) {
#pragma unroll
            for (int j = 0; j < kApproxPf; ++j) {
                qd[j] = qp[at + j];
                cd[j] = cp[at + j];
            }
        };
        auto run = [&](const float4* qd, const float4* cd) {
#pragma unroll
            for (int j = 0; j < kApproxPf; ++j) step(qd[j], cd[j]);
        };
        load(qa, ca, 0);
        for (; s4 + 3 * kApproxPf <= n4; s4 += 2 * kApproxPf) {   // sets at s4 (a) and s4 + kApproxPf (b)
            load(qb, cb, s4 + kApproxPf);
            run(qa, ca);
            load(qa, ca, s4 + 2 * kApproxPf);
            run(qb, cb);
        }
        run(qa, ca);   // the set loaded at s4
        s4 += kApproxPf;
    }
    for (; s4 < n4; ++s4) step(qp[s4], cp[s4]);
    const uint64_t c = c0 + r;   // accumulator register i: query row 8 (i / 4) + 4 h + i % 4, centroid c
    if (c >= ra.nlist) return;
    const float cn2 = ra.cn2[c];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t row = 8 * (i / 4) + 4 * h + (i % 4);
        const uint64_t q = q0 + row;
        if (q < ra.nq) ra.dist[q * ra.nlist + c] = fmaxf(0.0F, fmaf(-2.0F, acc[i], cn2 + s_qn2[row]));
    }
}

// The m-th smallest key (bits[c] << 32 | c) over nlist, m >= 1, by an 8-bit
// radix select from the top (keys are unique).
__device__ uint64_t RadixSelectKey(const uint32_t* bits, uint32_t nlist, uint32_t m, uint32_t* hist,
                                   uint32_t* s_pick) {
    uint64_t prefix = 0, pmask = 0;
    uint32_t need = m;
    for (int shift = 56; shift >= 0; shift -= 8) {
        for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
        __syncthreads();
        for (uint32_t c = threadIdx.x; c < nlist; c += blockDim.x) {
            const uint64_t key = KeyOf(bits, c);
            if ((key & pmask) == prefix) atomicAdd(&hist[(key >> shift) & 255], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t acc = 0, b = 0;
            for (; b < 256; ++b) {
                if (acc + hist[b] >= need) break;
                acc += hist[b];
            }
            s_pick[0] = b;
            s_pick[1] = need - acc;
        }
        __syncthreads();
        prefix |= static_cast<uint64_t>(s_pick[0]) << shift;
        pmask |= static_cast<uint64_t>(255) << shift;
        need = s_pick[1];
        __syncthreads();
    }
    return prefix;
}

__global__ void SelectProbesTwoStageKernel(RouteArgs ra, const uint32_t* __restrict__ nprobe_q, uint32_t nprobe_all,
                                           uint32_t maxp, uint32_t* __restrict__ probe_ids,
                                           float* __restrict__ probe_d2, float errf) {
    extern __shared__ uint32_t sm[];
    const uint32_t nlist = static_cast<uint32_t>(ra.nlist), D = ra.dim;
    uint32_t* bits = sm;                                                     // nlist
    uint64_t* sel = reinterpret_cast<uint64_t*>(sm + ((nlist + 1) & ~1u));   // pow2 >= maxp
    __shared__ uint32_t hist[256];
    __shared__ uint32_t s_pick[2];
    __shared__ uint32_t s_cnt, s_u;
    __shared__ uint32_t s_cand[kSelThreads], s_ncand;
    uint32_t sortn = 1;
    while (sortn < maxp) sortn <<= 1;
    const uint32_t lane = threadIdx.x & 31, grp = threadIdx.x >> 5, ngrp = blockDim.x >> 5;
    for (uint64_t qi = blockIdx.x; qi < ra.nq; qi += gridDim.x) {
        const uint32_t m = min(nlist, nprobe_q ? nprobe_q[qi] : nprobe_all);
        if (m > 0) {
            const float qn2 = ra.qn2[qi * ra.qn2_stride];
            for (uint32_t c = threadIdx.x; c < nlist; c += blockDim.x) bits[c] = FloatBits(ra.dist[qi * nlist + c]);
            if (threadIdx.x == 0) s_u = 0u;
            __syncthreads();
            const uint64_t K1 = RadixSelectKey(bits, nlist, m, hist, s_pick);
            // U: the largest a + E over the m smallest keys (non-negative floats
            // order as their bits)
            for (uint32_t c = threadIdx.x; c < nlist; c += blockDim.x)
                if (KeyOf(bits, c) <= K1)
                    atomicMax(&s_u, __float_as_uint(__uint_as_float(bits[c]) + errf * (qn2 + ra.cn2[c])));
            __syncthreads();
            const float U = __uint_as_float(s_u);
            const float* qrow = ra.rotq + qi * ra.qstride;
            // candidates, a chunk of kSelThreads centroids at a time: compacted,
            // then their reference-order distances, one 32-lane group each
            for (uint32_t b0 = 0; b0 < nlist; b0 += blockDim.x) {
                if (threadIdx.x == 0) s_ncand = 0;
                __syncthreads();
                const uint32_t c = b0 + threadIdx.x;
                if (c < nlist) {
                    const float a = __uint_as_float(bits[c]);
                    if (a - errf * (qn2 + ra.cn2[c]) <= U) s_cand[atomicAdd(&s_ncand, 1u)] = c;
                    else bits[c] = 0xFFFFFFFFu;
                }
                __syncthreads();
                const uint32_t nc = s_ncand;
                for (uint32_t j = grp; j < nc; j += ngrp) {
                    const uint32_t cc = s_cand[j];
                    const float* crow = ra.cent + static_cast<uint64_t>(cc) * ra.cstride;
                    float acc = 0.0F;   // the chain in order; its loads issue eight steps ahead
#pragma unroll 8
                    for (uint32_t d = lane; d < D; d += 32) acc = fmaf(qrow[d], crow[d], acc);
                    const float v = RouteReduce32(acc);
                    if (lane == 0) {
                        float d2 = fmaf(-v, 2.0F, ra.cn2[cc] + qn2);
                        if (d2 < 0.0F) d2 = 0.0F;
                        bits[cc] = FloatBits(d2);
                    }
                }
                __syncthreads();
            }
        }
        const uint64_t K = m > 0 ? RadixSelectKey(bits, nlist, m, hist, s_pick) : 0ull;
        if (threadIdx.x == 0) s_cnt = 0;
        for (uint32_t i = threadIdx.x; i < sortn; i += blockDim.x) sel[i] = ~0ull;
        __syncthreads();
        if (m > 0)
            for (uint32_t c = threadIdx.x; c < nlist; c += blockDim.x) {
                const uint64_t key = KeyOf(bits, c);
                if (key <= K) sel[atomicAdd(&s_cnt, 1u)] = key;
            }
        __syncthreads();
        for (uint32_t k2 = 2; k2 <= sortn; k2 <<= 1) {   // bitonic sort of sortn keys (ascending)
            for (uint32_t j = k2 >> 1; j > 0; j >>= 1) {
                for (uint32_t i = threadIdx.x; i < sortn; i += blockDim.x) {
                    const uint32_t l = i ^ j;
                    if (l > i) {
                        const uint64_t x = sel[i], y = sel[l];
                        const bool up = (i & k2) == 0;
                        if ((x > y) == up) {
                            sel[i] = y;
                            sel[l] = x;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (uint32_t i = threadIdx.x; i < maxp; i += blockDim.x) {
            const bool ok = i < m;
            probe_ids[qi * maxp + i] = ok ? static_cast<uint32_t>(sel[i]) : 0xFFFFFFFFu;
            probe_d2[qi * maxp + i] = ok ? __uint_as_float(static_cast<uint32_t>(sel[i] >> 32)) : 0.0F;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Launchers

hipError_t LaunchRotate(hipStream_t st, const float* x, uint64_t rows, uint32_t dim, const IndexView& ix,
                        float* out, bool center) {
    if (rows == 0) return hipSuccess;
    const uint32_t grid = static_cast<uint32_t>(rows < 65536 ? rows : 65536);
    hipLaunchKernelGGL(RotateRowsKernel, dim3(grid), dim3(256), ix.padded * sizeof(float), st, x, rows, dim, ix,
                       out, center ? 1 : 0);
    return hipGetLastError();
}

hipError_t LaunchQueryPrep(hipStream_t st, const float* rotq, uint64_t nq, const IndexView& ix, uint8_t* lut,
                           float* consts) {
    if (nq == 0) return hipSuccess;
    const uint32_t grid = static_cast<uint32_t>(nq < 65536 ? nq : 65536);
    hipLaunchKernelGGL(QueryPrepKernel, dim3(grid), dim3(256), static_cast<size_t>(ix.blocks) * 32 * sizeof(float),
                       st, rotq, nq, ix, lut, consts);
    return hipGetLastError();
}

static RouteArgs E8qRouteArgs(const float* rotq, const float* consts, uint64_t nq, const IndexView& ix, float* dist) {
    RouteArgs ra{};
    ra.rotq = rotq;
    ra.qstride = static_cast<uint32_t>(ix.padded);
    ra.qn2 = consts + 4;
    ra.qn2_stride = 8;
    ra.cent = ix.centroid_rot;
    ra.cn2 = ix.centroid_n2;
    ra.nlist = ix.nlist;
    ra.dim = static_cast<uint32_t>(ix.padded);
    ra.cstride = static_cast<uint32_t>(ix.padded);
    ra.nq = nq;
    ra.dist = dist;
    return ra;
}

hipError_t LaunchRoute(hipStream_t st, const float* rotq, const float* consts, uint64_t nq, const IndexView& ix,
                       float* dist) {
    return LaunchRouteDist(st, E8qRouteArgs(rotq, consts, nq, ix, dist));
}

hipError_t LaunchRouteProbes(hipStream_t st, const float* rotq, const float* consts, uint64_t nq, const IndexView& ix,
                             float* dist, const uint32_t* nprobe_q, uint32_t nprobe_all, uint32_t maxp,
                             uint32_t* probe_ids, float* probe_d2) {
    return LaunchRouteSelect(st, E8qRouteArgs(rotq, consts, nq, ix, dist), nprobe_q, nprobe_all, maxp, probe_ids,
                             probe_d2);
}

hipError_t LaunchRouteDist(hipStream_t st, const RouteArgs& ra) {
    if (ra.nq == 0) return hipSuccess;
    if (ra.dim == 0) {   // no coded prefix: every centroid scores its norm
        // dist = (cn2 + qn2) - 0, clamped; one thread per (query, centroid)
        const uint64_t tot = ra.nq * ra.nlist;
        hipLaunchKernelGGL(RouteNoDimKernel, dim3(static_cast<uint32_t>((tot + 255) / 256)), dim3(256), 0, st, ra);
        return hipGetLastError();
    }
    const uint64_t per_block = 8 * kRouteCT;   // eight 32-lane groups
    auto go = [&](auto kern, uint32_t qt) -> hipError_t {
        const dim3 grid(static_cast<uint32_t>((ra.nlist + per_block - 1) / per_block),
                        static_cast<uint32_t>((ra.nq + qt - 1) / qt));
        const size_t lds = qt * static_cast<size_t>(ra.dim) * sizeof(float);
        if (lds > 160 * 1024) return hipErrorInvalidValue;
        if (lds > 64 * 1024) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, ra);
        return hipGetLastError();
    };
    // 16 queries per tile; 4 above D = 2560 (the query tile lives in LDS).
    if (8ull * ra.dim * sizeof(float) <= 160 * 1024) return go(RouteDistKernel<8>, 8);
    return go(RouteDistKernel<4>, 4);
}

hipError_t LaunchRouteSelect(hipStream_t st, const RouteArgs& ra, const uint32_t* nprobe_q, uint32_t nprobe_all,
                             uint32_t maxp, uint32_t* probe_ids, float* probe_d2) {
    if (ra.nq == 0) return hipSuccess;
    auto aligned = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
    const bool two = ra.dim > 0 && ra.dim % 8 == 0 && ra.qstride % 4 == 0 && ra.cstride % 4 == 0 && aligned(ra.rotq) &&
                     aligned(ra.cent) && ra.nlist >= 256 && ra.nlist < (1ull << 31) &&
                     static_cast<uint64_t>(maxp) * 8 <= ra.nlist && ra.nq <= 65535ull * 32;
    if (!two) {
        const hipError_t e = LaunchRouteDist(st, ra);
        if (e != hipSuccess) return e;
        return LaunchSelectProbes(st, ra.dist, ra.nq, static_cast<uint32_t>(ra.nlist), nprobe_q, nprobe_all, maxp,
                                  probe_ids, probe_d2);
    }
    hipLaunchKernelGGL(RouteApproxKernel, dim3(static_cast<uint32_t>((ra.nlist + 127) / 128),
                                               static_cast<uint32_t>((ra.nq + 31) / 32)),
                       dim3(256), 0, st, ra);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // four times the bound on |a - d|: (D/2 + D/32 + 13) 2^-24 (qn2 + cn2)
    const float errf = static_cast<float>(4.0 * (ra.dim / 2.0 + ra.dim / 32.0 + 13.0) * 5.9604644775390625e-8);
    const uint32_t grid = static_cast<uint32_t>(ra.nq < 65536 ? ra.nq : 65536);
    const size_t lds = SelectProbesLds(static_cast<uint32_t>(ra.nlist), maxp);
    if (lds > 64 * 1024) {
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&SelectProbesTwoStageKernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(SelectProbesTwoStageKernel, dim3(grid), dim3(kSelThreads), lds, st, ra, nprobe_q, nprobe_all,
                       maxp, probe_ids, probe_d2, errf);
    return hipGetLastError();
}

size_t SelectProbesLds(uint32_t nlist, uint32_t maxp) {
    uint32_t sortn = 1;
    while (sortn < maxp) sortn <<= 1;
    return ((nlist + 1) & ~1u) * sizeof(uint32_t) + static_cast<size_t>(sortn) * sizeof(uint64_t);
}

hipError_t LaunchSelectProbes(hipStream_t st, const float* dist, uint64_t nq, uint32_t nlist, const uint32_t* nprobe_q,
                              uint32_t nprobe_all, uint32_t maxp, uint32_t* probe_ids, float* probe_d2) {
    if (nq == 0) return hipSuccess;
    const uint32_t grid = static_cast<uint32_t>(nq < 65536 ? nq : 65536);
    const size_t lds = SelectProbesLds(nlist, maxp);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&SelectProbesKernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(SelectProbesKernel, dim3(grid), dim3(kSelThreads), lds, st, dist, nq,
                       nlist, nprobe_q, nprobe_all, maxp, probe_ids, probe_d2);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Heaviest-first query order for the plain scans.  One wave scans one query,
// so a query with a long probed list (a Zipf head list) runs for most of the
// kernel; started last, it sets the kernel's end.  Each query's work is
// estimated by the sizes of its probed lists (begin[c + 1] - begin[c], in
// any unit), bucketed on a log scale (four buckets per octave), and the
// queries are listed by bucket, heaviest first, so the longest ones start
// with the grid.  Each query is searched alone, so the order changes no
// result.

__device__ __forceinline__ uint32_t WorkBucket(uint64_t w) {
    if (w == 0) return 0;
    const uint32_t e = 63 - static_cast<uint32_t>(__clzll(w));
    const uint32_t m = static_cast<uint32_t>(e >= 2 ? (w >> (e - 2)) & 3u : (w << (2 - e)) & 3u);
    return min(1 + 4 * e + m, kOrderBuckets - 1);
}

__global__ __launch_bounds__(256) void QueryWorkKernel(const uint64_t* __restrict__ begin, uint64_t nlist,
                                                       const uint32_t* __restrict__ probe,
                                                       const uint32_t* __restrict__ nprobe_q, uint32_t nprobe_all,
                                                       uint32_t maxp, uint64_t nq, uint32_t* __restrict__ key,
                                                       uint32_t* __restrict__ hist) {
    const uint64_t qi = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    const uint64_t np = min<uint64_t>(min<uint64_t>(nprobe_q ? nprobe_q[qi] : nprobe_all, maxp), nlist);
    uint64_t w = 0;
    for (uint64_t p = 0; p < np; ++p) {
        const uint32_t c = probe[qi * maxp + p];
        if (c < nlist) w += begin[c + 1] - begin[c];
    }
    const uint32_t b = WorkBucket(w);
    key[qi] = b;
    atomicAdd(&hist[b], 1u);
}

// hist[kOrderBuckets] the bucket sizes, cur[kOrderBuckets] zeroed.
__global__ __launch_bounds__(256) void QueryOrderKernel(uint64_t nq, const uint32_t* __restrict__ key,
                                                        const uint32_t* __restrict__ hist, uint32_t* __restrict__ cur,
                                                        uint32_t* __restrict__ order) {
    __shared__ uint32_t start[kOrderBuckets];
    if (threadIdx.x == 0) {   // bucket b starts after every heavier bucket
        uint32_t s = 0;
        for (int32_t b = kOrderBuckets - 1; b >= 0; --b) {
            start[b] = s;
            s += hist[b];
        }
    }
    __syncthreads();
    const uint64_t qi = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    const uint32_t b = key[qi];
    order[start[b] + atomicAdd(&cur[b], 1u)] = static_cast<uint32_t>(qi);
}

hipError_t LaunchQueryOrder(hipStream_t st, const uint64_t* begin, uint64_t nlist, const uint32_t* probe,
                            const uint32_t* nprobe_q, uint32_t nprobe_all, uint32_t maxp, uint64_t nq, uint32_t* key,
                            uint32_t* hist, uint32_t* order) {
    if (nq == 0) return hipSuccess;
    hipError_t e = LaunchFill32(st, hist, 2 * kOrderBuckets, 0);
    if (e != hipSuccess) return e;
    const uint32_t grid = static_cast<uint32_t>((nq + 255) / 256);
    hipLaunchKernelGGL(QueryWorkKernel, dim3(grid), dim3(256), 0, st, begin, nlist, probe, nprobe_q, nprobe_all, maxp,
                       nq, key, hist);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(QueryOrderKernel, dim3(grid), dim3(256), 0, st, nq, key, hist, hist + kOrderBuckets, order);
    return hipGetLastError();
}

}  // namespace e8qdev
