// e8q_encode.hip -- the build (encode) path on the device.
//
// Reference: E8qIndex::Build (e8q_build.cpp:87-503) and the E8 codec
// (e8q_codec.cpp:297-726).  Every decision the encoder makes is in fp64 and
// every fit is the reference's sequential fp64 FMA chain, so digits,
// branches, packed bytes, g1/h3 and the u8 banks come out bit-identical.
//
// Kernels:
//   MeanKernel        fp64 column sums in row order (one lane per dim; four
//                     waves stage rows through an LDS ring)
//   MeanFinishKernel  mean -> float -> rotated mean
//   CentroidKernel    rotate, center, fp64 ||c||^2
//   EncodeRowsKernel  one 64-lane workgroup per row: rotate, residual norm,
//                     unit, E8 nearest-lattice-point encode (one lane per
//                     8-dim block), 7 fit chains (one lane each), pack
//   BanksKernel       per 32-vector batch u8 banks (e8q_build.cpp:33-80)
#include "e8q_device.h"
#include "e8q_encoder.h"
#include "e8q_kernels.h"

namespace e8qdev {

// ---------------------------------------------------------------------------
// Mean: fp64 sum per dimension in row order (e8q_build.cpp:190-204).  A
// workgroup owns 64 consecutive dimensions.  Its four waves load 64-row
// stages (16 rows each, one coalesced 256-byte access per row) four stages
// ahead into registers and copy each stage into an LDS ring; wave 0 then runs
// the 64 dependent fp64 chains (one lane per dimension) over the stage in
// row order.
constexpr int kMeanRows = 16;   // rows per wave per stage
constexpr int kMeanStage = 4 * kMeanRows;

// Rows of stage `stage` for this wave.  Row addresses are clamped into the
// array (wave-uniform, so they stay in scalar registers) so every load is
// unconditional and nothing uses a loaded value before it is copied to LDS
// four stages later: rows past the array are never added (the last stage
// adds only its rows < n) and lanes past `dim` never store their sum.
__device__ __forceinline__ void MeanLoad(float (&g)[kMeanRows], const float* __restrict__ base, uint64_t stage,
                                         uint32_t wave, uint64_t n, uint32_t dim, uint32_t dc) {
    const uint64_t r0 = stage * kMeanStage + wave * kMeanRows;
#pragma unroll
    for (int i = 0; i < kMeanRows; ++i) {
        const uint64_t r = r0 + i;
        g[i] = base[(r < n ? r : n - 1) * dim + dc];
    }
}

__global__ __launch_bounds__(256) void MeanKernel(const float* __restrict__ base, uint64_t n, uint32_t dim,
                                                  double* __restrict__ mu) {
    __shared__ float ring[2][kMeanStage][64];
    const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t d = blockIdx.x * 64 + lane;
    const bool live = d < dim;
    const uint32_t dc = live ? d : dim - 1;
    const uint64_t nfull = n / kMeanStage, nst = (n + kMeanStage - 1) / kMeanStage;
    float g0[kMeanRows], g1[kMeanRows], g2[kMeanRows], g3[kMeanRows];
    MeanLoad(g0, base, 0, wave, n, dim, dc);
    MeanLoad(g1, base, 1, wave, n, dim, dc);
    MeanLoad(g2, base, 2, wave, n, dim, dc);
    MeanLoad(g3, base, 3, wave, n, dim, dc);
    // Every wave runs the chains (only wave 0's result is stored), so the
    // steady state has no branch.
    double acc = 0.0;
    auto publish = [&](float (&g)[kMeanRows], uint64_t s) {
        float(*slot)[64] = ring[s & 1];
#pragma unroll
        for (int i = 0; i < kMeanRows; ++i) slot[wave * kMeanRows + i][lane] = g[i];
        MeanLoad(g, base, s + 4, wave, n, dim, dc);
        __syncthreads();
        return slot;
    };
    auto add_full = [&](float (*slot)[64]) {
#pragma unroll
        for (int h = 0; h < kMeanStage; h += 16) {
            float x[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) x[i] = slot[h + i][lane];
#pragma unroll
            for (int i = 0; i < 16; ++i) acc += static_cast<double>(x[i]);
        }
    };
    auto add_rows = [&](float (*slot)[64], uint64_t s) {   // a stage that may be partial
        const int rows = static_cast<int>(min<uint64_t>(kMeanStage, n - s * kMeanStage));
        for (int i = 0; i < rows; ++i) acc += static_cast<double>(slot[i][lane]);
    };
    uint64_t s = 0;
    for (; s + 4 <= nfull; s += 4) {
        add_full(publish(g0, s));
        add_full(publish(g1, s + 1));
        add_full(publish(g2, s + 2));
        add_full(publish(g3, s + 3));
    }
    // the last (at most four) stages, the final one possibly partial
    if (s < nst) add_rows(publish(g0, s), s);
    if (s + 1 < nst) add_rows(publish(g1, s + 1), s + 1);
    if (s + 2 < nst) add_rows(publish(g2, s + 2), s + 2);
    if (s + 3 < nst) add_rows(publish(g3, s + 3), s + 3);
    if (wave == 0 && live) mu[d] = acc;
}

__global__ void MeanFinishKernel(const double* __restrict__ mu, uint64_t n, IndexView ix, float* __restrict__ mean_rot) {
    extern __shared__ float buf[];
    const uint32_t P = static_cast<uint32_t>(ix.padded);
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x)
        buf[i] = i < ix.dim ? static_cast<float>(mu[i] / static_cast<double>(n)) : 0.0F;
    __syncthreads();
    RotateInLds(buf, ix);
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) mean_rot[i] = buf[i];
}

// Rotated, centered centroids and their fp64 squared norms (e8q_build.cpp:249-272).
__global__ void CentroidKernel(const float* __restrict__ cent, IndexView ix, float* __restrict__ crot,
                               float* __restrict__ cn2) {
    extern __shared__ float buf[];
    const uint32_t P = static_cast<uint32_t>(ix.padded);
    for (uint64_t c = blockIdx.x; c < ix.nlist; c += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) buf[i] = i < ix.dim ? cent[c * ix.dim + i] : 0.0F;
        __syncthreads();
        RotateInLds(buf, ix);
        for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
            buf[i] = buf[i] - ix.mean_rot[i];
            crot[c * P + i] = buf[i];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (uint32_t i = 0; i < P; ++i) {
                const double v = static_cast<double>(buf[i]);
                s = fma(v, v, s);
            }
            cn2[c] = static_cast<float>(s);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Per-row encode.  Workgroup = one wave; rows in a grid-stride loop.
constexpr int kEncThreads = 64;
constexpr uint32_t kEncLdsLevels = 2048;   // merged tables up to this length (b <= 10) are copied to LDS

// LDS of EncodeRowsKernel: rotated row / unit, rotated centroid, digits,
// branches, then (kLdsTab) the merged level table and its digits.
__host__ __device__ inline uint32_t EncodeRowsTabOffset(uint32_t P, uint32_t blocks) {
    return (2 * P * 4 + P * 2 + blocks + 15) & ~15u;
}

template <bool kLdsTab>
__global__ __launch_bounds__(kEncThreads) void EncodeRowsKernel(IndexView ix, EncodeView ev, EncodeArgs a) {
    extern __shared__ float sm[];
    const uint32_t P = static_cast<uint32_t>(ix.padded);
    float* xb = sm;                                            // rotated row, then unit
    float* rc = sm + P;                                        // rotated centroid
    uint16_t* dg = reinterpret_cast<uint16_t*>(sm + 2 * P);    // digits
    uint8_t* brs = reinterpret_cast<uint8_t*>(dg + P);         // branches (blocks)
    // kLdsTab: the merged table (mlen levels + digits), the per-branch level
    // table (2 x 2^bits = mlen levels) and the 16 scan class means in LDS
    double* tlev = reinterpret_cast<double*>(reinterpret_cast<uint8_t*>(sm) + EncodeRowsTabOffset(P, ix.blocks));
    double* tbl = tlev + (kLdsTab ? ev.mlen : 0);
    double* tgm = tbl + (kLdsTab ? ev.mlen : 0);
    uint16_t* tdig = reinterpret_cast<uint16_t*>(tgm + (kLdsTab ? 16 : 0));
    if (kLdsTab) {
        for (uint32_t i = threadIdx.x; i < ev.mlen; i += kEncThreads) {
            tlev[i] = ev.mlev[i];
            tdig[i] = ev.mdig[i];
            tbl[i] = ev.levels[i];
        }
        if (threadIdx.x < 16) tgm[threadIdx.x] = ix.gate_mid[threadIdx.x];
    }
    const double* mlev = kLdsTab ? tlev : ev.mlev;
    const uint16_t* mdig = kLdsTab ? tdig : ev.mdig;
    const double* blev = kLdsTab ? tbl : ev.levels;
    const double* gmid = kLdsTab ? tgm : ix.gate_mid;
    __shared__ double s_fit[8];
    __shared__ double s_rn2;
    const int tid = threadIdx.x;
    const int bits = ix.bits;
    const int cbytes = bits - 1;
    const double sr = ev.scale * sqrt(static_cast<double>(P));
    const double inv_ls = 1.0 / sqrt(static_cast<double>(P));
    const uint32_t Q = 1u << bits;
    for (uint64_t r = blockIdx.x; r < a.n; r += gridDim.x) {
        const uint32_t c = a.cluster_ids[r];
        const uint32_t slot = a.slot[r];
        const uint64_t rank = slot - ix.cluster_begin[c];
        const uint32_t lane = static_cast<uint32_t>(rank & 31);
        uint8_t* rec = a.arena + ix.arena_begin[c] + (rank >> 5) * ix.batch_bytes;
        uint8_t* rrec = a.ref_arena + ix.ref_arena_begin[c] + (rank >> 5) * ix.ref_batch_bytes;
        for (uint32_t i = tid; i < P; i += kEncThreads) xb[i] = i < ix.dim ? a.base[r * ix.dim + i] : 0.0F;
        __syncthreads();
        RotateInLds(xb, ix);
        for (uint32_t i = tid; i < P; i += kEncThreads) {
            xb[i] = xb[i] - ix.mean_rot[i];
            rc[i] = ix.centroid_rot[static_cast<uint64_t>(c) * P + i];
        }
        __syncthreads();
        if (tid == 0) {   // P is a multiple of 64: eight coordinates' reads ahead of their FMAs
            double rn2 = 0.0;
            for (uint32_t d0 = 0; d0 < P; d0 += 8) {
                double rv[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) rv[j] = static_cast<double>(xb[d0 + j]) - static_cast<double>(rc[d0 + j]);
#pragma unroll
                for (int j = 0; j < 8; ++j) rn2 = fma(rv[j], rv[j], rn2);
            }
            s_rn2 = rn2;
            a.ids[slot] = static_cast<uint32_t>(r);
        }
        __syncthreads();
        const double rn2 = s_rn2;
        const double rn = sqrt(rn2);
        if (rn < 1e-20) {   // point on its centroid: everything stays zero
            __syncthreads();
            continue;
        }
        const double inv_rn = 1.0 / rn;
        for (uint32_t d = tid; d < P; d += kEncThreads)
            xb[d] = static_cast<float>((static_cast<double>(xb[d]) - static_cast<double>(rc[d])) * inv_rn);
        __syncthreads();
        for (int blk = tid; blk < ix.blocks; blk += kEncThreads) {
            uint16_t d8[8];
            int br;
            EncodeBlock(xb + blk * 8, sr, ev, mlev, mdig, d8, &br);
            for (int i = 0; i < 8; ++i) dg[blk * 8 + i] = d8[i];
            brs[blk] = static_cast<uint8_t>(br);
            // rowA byte (e8q_scan.cpp:193-205)
            uint32_t byte = 0;
            for (int i = 0; i < 8; ++i) {
                const uint32_t top = (bits == 1 && i == 7) ? static_cast<uint32_t>(br & 1)
                                                           : ((d8[i] >> (bits - 1)) & 1u);
                byte |= top << (7 - i);
            }
            rec[kHdr + lane * ix.blocks + blk] = static_cast<uint8_t>(byte);   // device layout: [lane][blk]
            if (bits >= 2) {   // refine field (e8q_codec.cpp:827-837)
                const int cb = cbytes;
                const uint64_t m = (uint64_t{1} << cb) - 1;
                uint64_t lo = 0, hi = 0;
                for (int i = 0; i < 8; ++i) {
                    const uint64_t f = i < 7 ? (d8[i] & m) : (((d8[7] & ~1u) & m) | static_cast<uint64_t>(br & 1));
                    const int pos = cb * i;
                    if (pos < 64) {
                        lo |= f << pos;
                        if (pos + cb > 64) hi |= f >> (64 - pos);
                    } else {
                        hi |= f << (pos - 64);
                    }
                }
                uint8_t* out = rrec + lane * ix.refine_stride + blk * cb;
                for (int b = 0; b < cb; ++b) out[b] = static_cast<uint8_t>(b < 8 ? (lo >> (8 * b)) : (hi >> (8 * (b - 8))));
            }
        }
        __syncthreads();
        // Seven fit chains in coordinate order (e8q_build.cpp:339-376), all on
        // lane 0 and interleaved (each chain keeps its own order; one load of
        // the operands per coordinate feeds the seven steps):
        // 0 <u,vE> 1 <c,vE> 2 ||vE||^2 3 <vE,w> 4 <u,w> 5 ||w||^2 6 <c,w>.
        if (tid == 0) {
            double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0, f4 = 0.0, f5 = 0.0, f6 = 0.0;
            for (int blk = 0; blk < ix.blocks; ++blk) {
                const uint32_t br = brs[blk];
                double vev[8], wcv[8], uv[8], ccv[8];
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) {
                    const uint32_t i = blk * 8 + kk;
                    const uint32_t digit = dg[i];
                    vev[kk] = blev[br * Q + digit] * inv_ls;
                    wcv[kk] = gmid[2 * kk + (digit >> (bits - 1))];
                    uv[kk] = static_cast<double>(xb[i]);
                    ccv[kk] = static_cast<double>(rc[i]);
                }
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) {
                    const double ve = vev[kk], wc = wcv[kk], u = uv[kk], cc = ccv[kk];
                    f0 = fma(u, ve, f0);
                    f1 = fma(ve, cc, f1);
                    f2 = fma(ve, ve, f2);
                    f3 = fma(wc, ve, f3);
                    f4 = fma(wc, u, f4);
                    // ||w||^2 at the 8th coordinate of a block: separately rounded
                    // square (the reference build hoists it out of the FMA).
                    f5 = kk == 7 ? f5 + wc * wc : fma(wc, wc, f5);
                    f6 = fma(wc, cc, f6);
                }
            }
            s_fit[0] = f0;
            s_fit[1] = f1;
            s_fit[2] = f2;
            s_fit[3] = f3;
            s_fit[4] = f4;
            s_fit[5] = f5;
            s_fit[6] = f6;
        }
        __syncthreads();
        if (tid == 0) {
            const double due = s_fit[0], dce = s_fit[1], n2e = s_fit[2], dew = s_fit[3], duw = s_fit[4];
            const double n2w = s_fit[5], dcw = s_fit[6];
            double g1 = 0.0, h3 = 0.0;
            if (n2e > 1e-30 && due > 1e-12) {
                g1 = rn / due;
                h3 = fma(2.0 * g1, dce, rn2);
            }
            float* g1b = reinterpret_cast<float*>(rrec + ix.refine_stride * 32);
            float* h3b = g1b + 32;
            g1b[lane] = static_cast<float>(g1);
            h3b[lane] = static_cast<float>(h3);
            const double g1v = static_cast<double>(static_cast<float>(g1));
            const double sqrt_dm1 = sqrt(static_cast<double>(ix.dim > 1 ? ix.dim - 1 : 1));
            const double n2w_safe = fmax(n2w, 1e-30);
            const double uw_eps = 1e-12 * sqrt(n2w_safe);
            double w_use, perp;
            if (fabs(duw) > uw_eps) {
                const double wx = 2.0 * rn / duw;
                const double g4 = g1v * 4.0;
                const double rx2 = fma(-(wx * g4), dew, fma(g1v * g4, n2e, n2w * (wx * wx)));
                w_use = wx;
                perp = sqrt(fmax(0.0, rx2)) / sqrt_dm1;
            } else {
                const double gamma = dew / n2w_safe;
                const double vp2 = fmax(fma(-n2w, gamma * gamma, n2e), 0.0);
                w_use = 2.0 * g1v * gamma;
                perp = 2.0 * g1v * sqrt(vp2) / sqrt_dm1;
            }
            a.ow[slot] = static_cast<float>(w_use);
            a.ob[slot] = static_cast<float>(fma(w_use, dcw, rn2));
            a.op[slot] = static_cast<float>(perp);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Banks: one 32-lane half-wave per batch; workgroup = 8 batches of one
// cluster.  Grid: x = cluster, loop over its batches.
__global__ __launch_bounds__(256) void BanksKernel(IndexView ix, const float* __restrict__ ob,
                                                   const float* __restrict__ ow, const float* __restrict__ op,
                                                   uint8_t* __restrict__ arena) {
    const uint64_t c = blockIdx.x;
    const uint64_t vb = ix.cluster_begin[c];
    const uint32_t cnt = static_cast<uint32_t>(ix.cluster_begin[c + 1] - vb);
    const uint32_t nb = (cnt + 31) / 32;
    const uint32_t t = threadIdx.x & 31;
    for (uint32_t b = threadIdx.x >> 5; b < nb; b += blockDim.x >> 5) {
        const uint32_t lo = b * 32;
        const uint32_t m = min(32u, cnt - lo);
        const bool live = t < m;
        const uint64_t s = vb + lo + t;
        const float vb_ = live ? ob[s] : 0.0F, vw = live ? ow[s] : 0.0F, vp = live ? op[s] : 0.0F;
        float bmin = live ? vb_ : __builtin_inff(), bmax = live ? vb_ : -__builtin_inff();
        float wmin = live ? vw : __builtin_inff(), wmax = live ? vw : -__builtin_inff();
        float pmin = live ? vp : __builtin_inff(), pmax = live ? vp : -__builtin_inff();
        for (int o = 16; o > 0; o >>= 1) {
            bmin = fminf(bmin, __shfl_xor(bmin, o, 32));
            bmax = fmaxf(bmax, __shfl_xor(bmax, o, 32));
            wmin = fminf(wmin, __shfl_xor(wmin, o, 32));
            wmax = fmaxf(wmax, __shfl_xor(wmax, o, 32));
            pmin = fminf(pmin, __shfl_xor(pmin, o, 32));
            pmax = fmaxf(pmax, __shfl_xor(pmax, o, 32));
        }
        const float wamax = fmaxf(fabsf(wmin), fabsf(wmax));
        const float bsc = bmax > bmin ? (bmax - bmin) / 255.0F : 1.0F;
        const float wsc = wmax > wmin ? (wmax - wmin) / 255.0F : 1.0F;
        const float psc = pmax > pmin ? (pmax - pmin) / 255.0F : 1.0F;
        uint8_t* bp = arena + ix.arena_begin[c] + static_cast<uint64_t>(b) * ix.batch_bytes;
        uint32_t bq = 0, wq = 0, pq = 0;
        if (live) {
            int v = static_cast<int>((vb_ - bmin) / bsc);
            bq = static_cast<uint32_t>(v < 0 ? 0 : (v > 255 ? 255 : v));
            v = static_cast<int>(roundf((vw - wmin) / wsc));
            wq = static_cast<uint32_t>(v < 0 ? 0 : (v > 255 ? 255 : v));
            v = static_cast<int>(ceilf((vp - pmin) / psc));
            pq = static_cast<uint32_t>(v < 0 ? 0 : (v > 255 ? 255 : v));
        }
        bp[t] = static_cast<uint8_t>(bq);
        bp[32 + t] = static_cast<uint8_t>(wq);
        bp[64 + t] = static_cast<uint8_t>(pq);
        float* sc = reinterpret_cast<float*>(bp + 96);
        if (t < 8) {
            const float vals[8] = {bmin, bsc, wmin, wsc, pmin, psc, wamax, 0.0F};
            sc[t] = vals[t];
        }
    }
}

// Debug: encode caller-provided unit vectors (no fits, no packing).
__global__ void EncodeUnitsKernel(EncodeView ev, const float* __restrict__ unit, uint64_t rows, uint32_t P, double sr,
                                  uint16_t* __restrict__ digits, uint8_t* __restrict__ branch) {
    const uint32_t blocks = P / 8;
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= rows * blocks) return;
    const uint64_t r = g / blocks;
    const uint32_t blk = static_cast<uint32_t>(g % blocks);
    uint16_t d8[8];
    int br;
    EncodeBlock(unit + r * P + blk * 8, sr, ev, d8, &br);
    for (int i = 0; i < 8; ++i) digits[r * P + blk * 8 + i] = d8[i];
    branch[r * blocks + blk] = static_cast<uint8_t>(br);
}

// ---------------------------------------------------------------------------
// Scan-record rowA layout: the E8AX file stores it [blk][lane] (32 bytes per
// block); the device stores it [lane][blk] so that one lane's code bytes are
// one contiguous load.  Headers are not touched.  One workgroup per record,
// which reads its whole rowA into LDS before writing, so src == dst is fine.
__global__ __launch_bounds__(256) void RecordLayoutKernel(const uint8_t* src, uint8_t* dst, uint64_t nrec,
                                                          uint32_t batch_bytes, uint32_t blocks, int to_device) {
    extern __shared__ uint8_t rbuf[];
    const uint32_t R = 32u * blocks;
    for (uint64_t r = blockIdx.x; r < nrec; r += gridDim.x) {
        const uint8_t* s = src + r * batch_bytes + kHdr;
        uint8_t* d = dst + r * batch_bytes + kHdr;
        for (uint32_t i = threadIdx.x; i < R; i += blockDim.x) rbuf[i] = s[i];
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < R; i += blockDim.x) {
            if (to_device) {
                const uint32_t lane = i / blocks, blk = i - lane * blocks;
                d[i] = rbuf[blk * 32 + lane];
            } else {
                const uint32_t blk = i >> 5, lane = i & 31;
                d[i] = rbuf[lane * blocks + blk];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Cluster summaries of the scan-record headers; one wave per cluster.
__global__ __launch_bounds__(64) void ClusterSummaryKernel(IndexView ix, float4* __restrict__ csum) {
    const uint64_t c = blockIdx.x;
    const uint32_t cnt = static_cast<uint32_t>(ix.cluster_begin[c + 1] - ix.cluster_begin[c]);
    const uint32_t nb = (cnt + 31) / 32;
    float s0 = __builtin_inff(), s3 = -__builtin_inff(), s6 = -__builtin_inff(), pm = -__builtin_inff();
    for (uint32_t b = threadIdx.x; b < nb; b += 64) {
        const float* sc = reinterpret_cast<const float*>(ix.arena + ix.arena_begin[c] +
                                                         static_cast<uint64_t>(b) * ix.batch_bytes + 96);
        s0 = fminf(s0, sc[0]);
        s3 = fmaxf(s3, sc[3]);
        s6 = fmaxf(s6, sc[6]);
        pm = fmaxf(pm, fmaf(sc[5], 255.0F, sc[4]));
    }
    for (int o = 32; o > 0; o >>= 1) {
        s0 = fminf(s0, __shfl_xor(s0, o));
        s3 = fmaxf(s3, __shfl_xor(s3, o));
        s6 = fmaxf(s6, __shfl_xor(s6, o));
        pm = fmaxf(pm, __shfl_xor(pm, o));
    }
    if (threadIdx.x == 0) csum[c] = make_float4(s0, s3, s6, pm);
}

// ---------------------------------------------------------------------------
// Launchers

hipError_t LaunchClusterSummary(hipStream_t st, const IndexView& ix, float4* csum) {
    if (ix.nlist == 0) return hipSuccess;
    hipLaunchKernelGGL(ClusterSummaryKernel, dim3(static_cast<uint32_t>(ix.nlist)), dim3(64), 0, st, ix, csum);
    return hipGetLastError();
}

hipError_t LaunchRecordLayout(hipStream_t st, const IndexView& ix, const uint8_t* src, uint8_t* dst, uint64_t nrec,
                              bool to_device) {
    if (nrec == 0) return hipSuccess;
    const uint32_t grid = static_cast<uint32_t>(nrec < 65536 ? nrec : 65536);
    hipLaunchKernelGGL(RecordLayoutKernel, dim3(grid), dim3(256), 32 * static_cast<size_t>(ix.blocks), st, src, dst,
                       nrec, ix.batch_bytes, static_cast<uint32_t>(ix.blocks), to_device ? 1 : 0);
    return hipGetLastError();
}

hipError_t LaunchMean(hipStream_t st, const float* base, uint64_t n, uint32_t dim, double* mu) {
    hipLaunchKernelGGL(MeanKernel, dim3((dim + 63) / 64), dim3(256), 0, st, base, n, dim, mu);
    return hipGetLastError();
}

hipError_t LaunchMeanFinish(hipStream_t st, const double* mu, uint64_t n, const IndexView& ix, float* mean_rot) {
    hipLaunchKernelGGL(MeanFinishKernel, dim3(1), dim3(256), ix.padded * sizeof(float), st, mu, n, ix, mean_rot);
    return hipGetLastError();
}

hipError_t LaunchCentroids(hipStream_t st, const float* centroids, const IndexView& ix, float* centroid_rot,
                           float* centroid_n2) {
    const uint32_t grid = static_cast<uint32_t>(ix.nlist < 65536 ? ix.nlist : 65536);
    hipLaunchKernelGGL(CentroidKernel, dim3(grid), dim3(256), ix.padded * sizeof(float), st, centroids, ix,
                       centroid_rot, centroid_n2);
    return hipGetLastError();
}

hipError_t LaunchEncodeRows(hipStream_t st, const IndexView& ix, const EncodeView& ev, const EncodeArgs& a) {
    if (a.n == 0) return hipSuccess;
    const uint32_t P = static_cast<uint32_t>(ix.padded);
    const bool tab = ev.merged && ev.mlen <= kEncLdsLevels;
    const size_t lds =
        EncodeRowsTabOffset(P, ix.blocks) + (tab ? ev.mlen * (2 * sizeof(double) + sizeof(uint16_t)) + 16 * sizeof(double) : 0) + 16;
    const uint32_t grid = static_cast<uint32_t>(a.n < 131072 ? a.n : 131072);
    if (tab) hipLaunchKernelGGL(EncodeRowsKernel<true>, dim3(grid), dim3(kEncThreads), lds, st, ix, ev, a);
    else hipLaunchKernelGGL(EncodeRowsKernel<false>, dim3(grid), dim3(kEncThreads), lds, st, ix, ev, a);
    return hipGetLastError();
}

hipError_t LaunchBanks(hipStream_t st, const IndexView& ix, const float* ob, const float* ow, const float* op,
                       uint8_t* arena) {
    if (ix.nlist == 0) return hipSuccess;
    hipLaunchKernelGGL(BanksKernel, dim3(static_cast<uint32_t>(ix.nlist)), dim3(256), 0, st, ix, ob, ow, op, arena);
    return hipGetLastError();
}

hipError_t LaunchEncodeUnits(hipStream_t st, const IndexView& ix, const EncodeView& ev, const float* unit,
                             uint64_t rows, double scale, uint16_t* digits, uint8_t* branch) {
    const uint32_t P = static_cast<uint32_t>(ix.padded);
    const uint64_t total = rows * (P / 8);
    if (total == 0) return hipSuccess;
    const double sr = scale * sqrt(static_cast<double>(P));
    hipLaunchKernelGGL(EncodeUnitsKernel, dim3(static_cast<uint32_t>((total + 63) / 64)), dim3(64), 0, st, ev, unit,
                       rows, P, sr, digits, branch);
    return hipGetLastError();
}

}  // namespace e8qdev
