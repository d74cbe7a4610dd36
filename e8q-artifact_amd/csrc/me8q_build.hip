// me8q_build.hip -- ME8Q build on the device: rotated centroids, the
// per-row mixed-rate encode (a26) and the FastScan layout (a27).
//
//   M8CentroidKernel  per list: each coded run's slice rotated by its own
//                     FHT-Kac-4 (me8q_search.cpp:45-72), |c'|^2 over the
//                     padded row (ivf_common.hpp:43-54) and per-run |c'_r|^2
//                     (me8q_fast.cpp:76-86).
//   M8EncodeKernel    per row: residual, dropped-tail energy, per-run
//                     rotation, E8 encode of the run's unit residual at
//                     FixedScale(rate), bit-tight pack, fp64 fits coef / h2
//                     (me8q_build.cpp:203-309).
//   M8FastKernel      per (vector, run): top-bit plane bytes, (rate-1)-bit
//                     refine fields, ZPR fit gain / vpc / perp from the
//                     stored code (me8q_fast.cpp:88-221).
//   M8BanksKernel     per (batch, run): u8 banks + scales (me8q_fast.cpp:
//                     230-301).
// Every fp64 chain runs on one thread in the reference's coordinate order;
// the multiply-adds the reference build contracts are written as fma.
#include <hip/hip_runtime.h>

#include "e8q_encoder.h"
#include "e8q_kernels.h"
#include "me8q_device.h"
#include "me8q_kernels.h"

namespace e8qdev {

__global__ __launch_bounds__(256) void M8CentroidKernel(M8View v) {
    extern __shared__ float cb[];   // padded
    const uint32_t P = static_cast<uint32_t>(v.padded);
    for (uint64_t c = blockIdx.x; c < v.nlist; c += gridDim.x) {
        for (uint32_t d = threadIdx.x; d < P; d += blockDim.x) cb[d] = v.centroids[c * P + d];
        __syncthreads();
        for (uint32_t r = 0; r < v.nrun; ++r) {
            const M8Run& ru = v.runs[r];
            RotateInLds(cb + ru.dim_off, ru.len, ru.trunc, ru.fac, ru.signs);
        }
        // dropped segments keep zeros in the rotated row
        for (uint32_t s = 0; s < v.nseg; ++s) {
            const M8Seg sg = v.segs[s];
            if (sg.bits != 0) continue;
            for (uint32_t d = threadIdx.x; d < sg.len; d += blockDim.x) cb[sg.off + d] = 0.0F;
        }
        __syncthreads();
        for (uint32_t d = threadIdx.x; d < P; d += blockDim.x) v.cent_rot[c * P + d] = cb[d];
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (uint32_t d = 0; d < P; ++d) s = fma(static_cast<double>(cb[d]), static_cast<double>(cb[d]), s);
            v.cn2[c] = static_cast<float>(s);
        }
        if (threadIdx.x >= 32 && threadIdx.x - 32 < v.nrun) {
            const M8Run& ru = v.runs[threadIdx.x - 32];
            double a = 0.0;
            for (uint32_t t = 0; t < ru.len; ++t) {
                const double x = cb[ru.dim_off + t];
                a = fma(x, x, a);
            }
            v.rc_n2[c * v.nrun + (threadIdx.x - 32)] = a;
        }
        __syncthreads();
    }
}

// Encode: one 64-thread workgroup per row (grid-stride).  LDS: residual row
// (padded floats), the current run's unit residual, its digits and branches.
constexpr int kM8EncThreads = 64;

__global__ __launch_bounds__(kM8EncThreads) void M8EncodeKernel(M8View v, const float* __restrict__ base,
                                                                const uint32_t* __restrict__ cid,
                                                                const uint32_t* __restrict__ rank, uint64_t n) {
    extern __shared__ float sm[];
    const uint32_t P = static_cast<uint32_t>(v.padded);
    float* res = sm;                                            // padded
    float* unit = sm + P;                                       // max run length (<= padded)
    uint16_t* dg = reinterpret_cast<uint16_t*>(sm + 2 * P);     // padded
    uint8_t* brs = reinterpret_cast<uint8_t*>(dg + P);          // padded / 8
    __shared__ double s_drop, s_n2, s_fit[2];
    const int tid = threadIdx.x;
    for (uint64_t row = blockIdx.x; row < n; row += gridDim.x) {
        const uint32_t c = cid[row];
        const uint64_t rk = rank[row];
        const uint64_t gi = v.list_begin[c] + rk;
        uint8_t* code = v.codes + gi * v.stride;
        for (uint32_t d = tid; d < P; d += kM8EncThreads)
            res[d] = d < v.dim ? base[row * v.dim + d] - v.centroids[static_cast<uint64_t>(c) * P + d] : 0.0F;
        __syncthreads();
        if (tid == 0) {
            // energy of the dropped dims, from the unrotated residual
            double drop = 0.0;
            if (v.drop_energy)
                for (uint32_t s = 0; s < v.nseg; ++s) {
                    const M8Seg sg = v.segs[s];
                    if (sg.bits != 0) continue;
                    for (uint32_t j = 0; j < sg.len; ++j) {
                        const double x = res[sg.off + j];
                        drop = fma(x, x, drop);
                    }
                }
            s_drop = drop;
            v.ids[gi] = static_cast<uint32_t>(row);
        }
        __syncthreads();
        for (uint32_t r = 0; r < v.nrun; ++r) {
            const M8Run& ru = v.runs[r];
            RotateInLds(res + ru.dim_off, ru.len, ru.trunc, ru.fac, ru.signs);
        }
        for (uint32_t r = 0; r < v.nrun; ++r) {
            const M8Run& ru = v.runs[r];
            const float* x = res + ru.dim_off;
            if (tid == 0) {   // run lengths are multiples of 64: eight reads ahead of their FMAs
                double n2 = 0.0;
                for (uint32_t j0 = 0; j0 < ru.len; j0 += 8) {
                    double xv[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) xv[i] = static_cast<double>(x[j0 + i]);
#pragma unroll
                    for (int i = 0; i < 8; ++i) n2 = fma(xv[i], xv[i], n2);
                }
                s_n2 = n2;
            }
            __syncthreads();
            const double n2 = s_n2;
            const double nrm = sqrt(n2);
            const bool zero = !(nrm > 0.0);
            if (!zero) {
                const double inv = 1.0 / nrm;
                for (uint32_t j = tid; j < ru.len; j += kM8EncThreads)
                    unit[j] = static_cast<float>(static_cast<double>(x[j]) * inv);
            }
            __syncthreads();
            const double sr = ru.ev.scale * sqrt(static_cast<double>(ru.len));
            for (uint32_t b = tid; b < ru.nblk; b += kM8EncThreads) {
                uint16_t d8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                int br = 0;
                if (!zero) EncodeBlock(unit + b * 8, sr, ru.ev, d8, &br);
                uint32_t d32[8];
                for (int i = 0; i < 8; ++i) {
                    dg[b * 8 + i] = d8[i];
                    d32[i] = d8[i];
                }
                brs[b] = static_cast<uint8_t>(br);
                M8Pack(d32, static_cast<uint32_t>(br), ru.rate, code + ru.code_off + b * ru.rate);
            }
            __syncthreads();
            if (tid == 0) {   // the two fits, each one chain in coordinate order, interleaved on one lane
                const float* cr = v.cent_rot + static_cast<uint64_t>(c) * P + ru.dim_off;
                const uint32_t q = 1u << ru.rate;
                double a0 = 0.0, a1 = 0.0;
                for (uint32_t b = 0; b < ru.nblk; ++b) {
                    const float* lev = ru.lev + brs[b] * q;
                    double lv[8], xv[8], cv[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        lv[i] = static_cast<double>(lev[dg[b * 8 + i]]);
                        xv[i] = static_cast<double>(x[b * 8 + i]);
                        cv[i] = static_cast<double>(cr[b * 8 + i]);
                    }
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        a0 = fma(xv[i], lv[i], a0);
                        a1 = fma(cv[i], lv[i], a1);
                    }
                }
                s_fit[0] = a0;
                s_fit[1] = a1;
            }
            __syncthreads();
            if (tid == 0) {
                const double dr = s_fit[0], dc = s_fit[1];
                double coef = dr > 1e-12 ? n2 / dr : 0.0;
                if (!isfinite(coef)) coef = 0.0;
                double h2 = fma(2.0 * coef, dc, n2);
                if (r + 1 == v.nrun) h2 += s_drop;
                float* fac = reinterpret_cast<float*>(code + r * 8);
                fac[0] = static_cast<float>(coef);
                fac[1] = static_cast<float>(isfinite(h2) ? h2 : 0.0);
            }
            __syncthreads();
        }
    }
}

// Fast layout: one thread per (vector, run).
__global__ __launch_bounds__(256) void M8FastKernel(M8View v, float* __restrict__ perp) {
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= v.n * v.nrun) return;
    const uint64_t gi = t / v.nrun;
    const uint32_t r = static_cast<uint32_t>(t % v.nrun);
    const uint32_t c = M8ListOf(v.list_begin, v.nlist, gi);
    const uint64_t i = gi - v.list_begin[c];
    const M8Run& ru = v.runs[r];
    const uint8_t* code = v.codes + gi * v.stride;
    uint8_t* rrow = v.refine + gi * v.rstride;
    uint8_t* pb = v.plane + (v.batch_begin[c] + (i >> 5)) * v.rec + (i & 31) * v.plane_cols + ru.blk0;
    const float* rc = v.cent_rot + static_cast<uint64_t>(c) * v.padded + ru.dim_off;
    const uint32_t bits = ru.rate, q = 1u << bits;
    double dew = 0.0, n2w = 0.0, n2e = 0.0, dce = 0.0, dcw = 0.0;
    for (uint32_t b = 0; b < ru.nblk; ++b) {
        uint32_t d[8], br;
        M8Unpack(code + ru.code_off + b * bits, bits, d, &br);
        if (bits >= 2) M8Pack(d, br, bits - 1, rrow + ru.ref_off + b * (bits - 1));
        uint32_t byte = 0;
        for (int k = 0; k < 8; ++k) {
            const uint32_t top = (bits == 1 && k == 7) ? (br & 1u) : ((d[k] >> (bits - 1)) & 1u);
            byte |= top << (7 - k);
        }
        pb[b] = static_cast<uint8_t>(byte);
        const float* lev = ru.lev + br * q;
        for (int k = 0; k < 8; ++k) {
            const double ve = lev[d[k]];
            const double wc = (bits == 1 && k == 7) ? 0.0 : static_cast<double>((d[k] >> (bits - 1)) ? ru.pbar1 : ru.pbar0);
            const double cc = rc[b * 8 + k];
            dew = fma(ve, wc, dew);
            n2w = fma(wc, wc, n2w);
            n2e = fma(ve, ve, n2e);
            dce = fma(cc, ve, dce);
            dcw = fma(cc, wc, dcw);
        }
    }
    const float* fac = reinterpret_cast<const float*>(code + r * 8);
    const double coef = fac[0], h2 = fac[1];
    const double rn2 = fmax(fma(-(2.0 * coef), dce, h2), 0.0);
    const double rn = sqrt(rn2);
    const double slm1 = sqrt(static_cast<double>(ru.len > 1 ? ru.len - 1 : 1));
    const double ac = fabs(coef);
    float gain = 0.0F, vpc = 0.0F, pp = 0.0F;
    if (n2e > 1e-30 && ac > 1e-30) {
        const double ven = sqrt(n2e);
        const double n2ws = fmax(n2w, 1e-30);
        const double duw = dew / ven;
        const double eps = 1e-12 * sqrt(n2ws);
        double wx, nrx;
        if (fabs(duw) > eps) {
            wx = 2.0 * rn / duw;
            const double c4 = 4.0 * coef;
            const double rx2 = fma(-(c4 * wx), dew, fma(c4 * coef, n2e, (wx * wx) * n2w));
            nrx = sqrt(fmax(0.0, rx2));
        } else {
            const double g = dew / n2ws;
            wx = 2.0 * coef * g;
            const double vp2 = fmax(fma(-(g * g), n2w, n2e), 0.0);
            nrx = 2.0 * ac * sqrt(vp2);
        }
        gain = static_cast<float>(wx / (2.0 * ac));
        vpc = static_cast<float>(fma(-wx, dcw, h2 - rn2) / (2.0 * ac));
        pp = static_cast<float>((nrx / slm1) / (2.0 * ac));
    } else if (n2w > 1e-30) {
        const double g = dew / n2w;
        const double vp2 = fmax(fma(-(g * g), n2w, n2e), 0.0);
        gain = static_cast<float>(g);
        vpc = static_cast<float>(fma(-g, dcw, dce));
        pp = static_cast<float>(sqrt(vp2) / slm1);
    }
    float* f4 = reinterpret_cast<float*>(rrow + r * 16);
    f4[0] = fac[0];
    f4[1] = fac[1];
    f4[2] = gain;
    f4[3] = vpc;
    perp[gi * v.nrun + r] = pp;
}

// Banks: one thread per (batch, run).
__global__ __launch_bounds__(256) void M8BanksKernel(M8View v, const float* __restrict__ perp, uint64_t nbatch) {
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nbatch * v.nrun) return;
    const uint64_t gb = t / v.nrun;
    const uint32_t r = static_cast<uint32_t>(t % v.nrun);
    const uint32_t c = M8ListOf(v.batch_begin, v.nlist, gb);
    const uint64_t b = gb - v.batch_begin[c];
    const uint64_t cnt = v.list_begin[c + 1] - v.list_begin[c];
    const uint64_t row0 = b * 32;
    const uint32_t nb = static_cast<uint32_t>(min<uint64_t>(32, cnt - row0));
    uint8_t* bp = v.plane + gb * v.rec + 32u * v.plane_cols + r * kM8BankBytes;
    float vmin[3] = {0.0F, 0.0F, 0.0F}, vmax[3] = {0.0F, 0.0F, 0.0F};
    auto vals = [&](uint32_t l, float out[3]) {
        const uint64_t gi = v.list_begin[c] + row0 + l;
        const float* f4 = reinterpret_cast<const float*>(v.refine + gi * v.rstride + r * 16);
        const float cf = f4[0], h2 = f4[1], gain = f4[2], vpc = f4[3];
        out[0] = fmaf(-(2.0F * cf), vpc, h2);
        out[1] = 2.0F * cf * gain;
        out[2] = perp[gi * v.nrun + r] * 2.0F * fabsf(cf);
    };
    for (uint32_t l = 0; l < nb; ++l) {
        float x[3];
        vals(l, x);
        for (int k = 0; k < 3; ++k) {
            if (l == 0) {
                vmin[k] = vmax[k] = x[k];
            } else {
                vmin[k] = x[k] < vmin[k] ? x[k] : vmin[k];
                vmax[k] = vmax[k] < x[k] ? x[k] : vmax[k];
            }
        }
    }
    float sc[3];
    for (int k = 0; k < 3; ++k) sc[k] = vmax[k] > vmin[k] ? (vmax[k] - vmin[k]) / 255.0F : 1.0F;
    for (uint32_t l = 0; l < nb; ++l) {
        float x[3];
        vals(l, x);
        const int a0 = static_cast<int>((x[0] - vmin[0]) / sc[0]);
        const int a1 = static_cast<int>(lroundf((x[1] - vmin[1]) / sc[1]));
        const int a2 = static_cast<int>(ceilf((x[2] - vmin[2]) / sc[2]));
        bp[l] = static_cast<uint8_t>(min(max(a0, 0), 255));
        bp[32 + l] = static_cast<uint8_t>(min(max(a1, 0), 255));
        bp[64 + l] = static_cast<uint8_t>(min(max(a2, 0), 255));
    }
    float* scl = reinterpret_cast<float*>(bp + 128);
    scl[0] = vmin[0];
    scl[1] = sc[0];
    scl[2] = vmin[1];
    scl[3] = sc[1];
    scl[4] = vmin[2];
    scl[5] = sc[2];
    scl[6] = 0.0F;
    scl[7] = 0.0F;
}

// The batched form's refine records (rrec) and batch scales (bscale) from the
// refine rows and the plane records' banks: one thread per 8-byte word of a
// record.  A record is the refine row's float4 header, the vector's W and
// Omega bank bytes per run (zero padded to rrec_shift bytes), the refine
// row's fields, zeros to the end of its last 128-byte line.
// One 8-byte word of the refine record of batch slot s = gb * 32 + l (zero
// past the list's vectors).
__global__ __launch_bounds__(256) void M8RefineRecKernel(M8View v, uint64_t nbatch) {
    const uint64_t words = v.rrec_bytes / 8;
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nbatch * 32 * words) return;
    const uint64_t s = t / words, gb = s >> 5;
    const uint32_t l = static_cast<uint32_t>(s & 31);
    const uint32_t o = static_cast<uint32_t>(t % words) * 8;
    const uint32_t R = v.nrun, hdr = 16 * R;
    const uint32_t fields = v.rstride_fields;   // bytes of the fields after the header
    const uint32_t c = M8ListOf(v.batch_begin, v.nlist, gb);
    const uint64_t i = (gb - v.batch_begin[c]) * 32 + l;   // position in list c
    uint64_t w = 0;
    if (i < v.list_begin[c + 1] - v.list_begin[c]) {
        const uint64_t gi = v.list_begin[c] + i;
        const uint8_t* row = v.refine + gi * v.rstride;
        if (o < hdr) {
            w = *reinterpret_cast<const uint64_t*>(row + o);
        } else if (o < hdr + v.rrec_shift) {
            const uint8_t* banks = v.plane + gb * v.rec + 32u * v.plane_cols;
            for (uint32_t k = 0; k < 8; ++k) {
                const uint32_t j = o - hdr + k;   // byte j of the bank area: run j / 2, W (even) or Omega (odd)
                if (j < 2 * R)
                    w |= static_cast<uint64_t>(banks[(j / 2) * kM8BankBytes + 32 * (1 + (j & 1)) + l]) << (8 * k);
            }
        } else if (o < hdr + v.rrec_shift + fields) {
            w = *reinterpret_cast<const uint64_t*>(row + o - v.rrec_shift);
        }
    }
    *reinterpret_cast<uint64_t*>(v.rrec + s * v.rrec_bytes + o) = w;
}

// Each batch's vector count and each batch slot's id.
__global__ __launch_bounds__(256) void M8SlotKernel(M8View v, uint64_t nbatch) {
    const uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= nbatch * 32) return;
    const uint64_t gb = s >> 5;
    const uint32_t c = M8ListOf(v.batch_begin, v.nlist, gb);
    const uint64_t cnt = v.list_begin[c + 1] - v.list_begin[c], i = (gb - v.batch_begin[c]) * 32 + (s & 31);
    v.pids[s] = i < cnt ? v.ids[v.list_begin[c] + i] : 0xFFFFFFFFu;
    if ((s & 31) == 0) {
        const uint64_t left = cnt - (i < cnt ? i : cnt);
        v.bcnt[gb] = static_cast<uint8_t>(left < 32 ? left : 32);
    }
}

__global__ __launch_bounds__(256) void M8BatchScaleKernel(M8View v, uint64_t nbatch) {
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nbatch * v.nrun) return;
    const uint64_t gb = t / v.nrun;
    const uint32_t r = static_cast<uint32_t>(t % v.nrun);
    const float* scl = reinterpret_cast<const float*>(v.plane + gb * v.rec + 32u * v.plane_cols + r * kM8BankBytes + 128);
    v.bscale[t] = make_float4(scl[2], scl[3], scl[4], scl[5]);
}

// ---------------------------------------------------------------------------
// Launchers

hipError_t LaunchM8RefineRecords(hipStream_t st, const M8View& v, uint64_t nbatch) {
    const uint64_t t1 = nbatch * 32 * (v.rrec_bytes / 8);
    if (t1 > 0) {
        hipLaunchKernelGGL(M8RefineRecKernel, dim3(static_cast<uint32_t>((t1 + 255) / 256)), dim3(256), 0, st, v,
                           nbatch);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(M8SlotKernel, dim3(static_cast<uint32_t>((nbatch * 32 + 255) / 256)), dim3(256), 0, st, v,
                           nbatch);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const uint64_t t2 = nbatch * v.nrun;
    if (t2 > 0)
        hipLaunchKernelGGL(M8BatchScaleKernel, dim3(static_cast<uint32_t>((t2 + 255) / 256)), dim3(256), 0, st, v,
                           nbatch);
    return hipGetLastError();
}

hipError_t LaunchM8Centroids(hipStream_t st, const M8View& v) {
    if (v.nlist == 0) return hipSuccess;
    const uint32_t grid = static_cast<uint32_t>(v.nlist < 65536 ? v.nlist : 65536);
    hipLaunchKernelGGL(M8CentroidKernel, dim3(grid), dim3(256), v.padded * sizeof(float), st, v);
    return hipGetLastError();
}

size_t M8EncodeLds(const M8View& v) {
    return 2 * v.padded * sizeof(float) + v.padded * sizeof(uint16_t) + v.padded / 8 + 16;
}

hipError_t LaunchM8Encode(hipStream_t st, const M8View& v, const float* base, const uint32_t* cid,
                          const uint32_t* rank, uint64_t n) {
    if (n == 0) return hipSuccess;
    const size_t lds = M8EncodeLds(v);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&M8EncodeKernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    const uint32_t grid = static_cast<uint32_t>(n < 262144 ? n : 262144);
    hipLaunchKernelGGL(M8EncodeKernel, dim3(grid), dim3(kM8EncThreads), lds, st, v, base, cid, rank, n);
    return hipGetLastError();
}

hipError_t LaunchM8Fast(hipStream_t st, const M8View& v, float* perp, uint64_t nbatch) {
    const uint64_t t1 = v.n * v.nrun;
    if (t1 > 0) {
        hipLaunchKernelGGL(M8FastKernel, dim3(static_cast<uint32_t>((t1 + 255) / 256)), dim3(256), 0, st, v, perp);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const uint64_t t2 = nbatch * v.nrun;
    if (t2 > 0)
        hipLaunchKernelGGL(M8BanksKernel, dim3(static_cast<uint32_t>((t2 + 255) / 256)), dim3(256), 0, st, v, perp,
                           nbatch);
    return hipGetLastError();
}

}  // namespace e8qdev
