// e8q_device.h -- device-side views and helpers shared by the E8Q kernels.
//
// Floating point: every kernel TU is compiled with -ffp-contract=off, so a
// multiply-add is fused only where it is written as fmaf()/fma(); those are
// exactly the places the reference fuses (explicit _mm256_fmadd/_fnmadd in
// its AVX2 kernels, and GCC's contraction of its generic C++ at x86-64-v3).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace e8qdev {

constexpr int kHdr = 128;         // scan-record header bytes (e8q.hpp:193-197)
constexpr int kMaxK = 128;        // e8q.hpp:44
constexpr int kMaxBits = 13;

// Read-only device view of one index; passed by value to every kernel.
struct IndexView {
    uint64_t n, dim, padded, nlist;
    int bits, blocks;
    uint32_t batch_bytes;      // scan record, roundup64(128 + 32*blocks)
    uint32_t ref_batch_bytes;  // refine record, roundup64(32*stride + 256)
    uint32_t refine_stride;    // blocks*(bits-1) bytes per lane (0 at b == 1)
    uint32_t tab_stride;       // max(256, 2^bits)
    uint32_t trunc;            // rotator FHT window
    float fac;                 // 1/sqrt(trunc)
    const uint8_t* signs;
    const float* mean_rot;
    const float* centroid_rot;
    const float* centroid_n2;
    const uint64_t* cluster_begin;
    const uint64_t* arena_begin;
    const uint64_t* ref_arena_begin;
    const uint8_t* arena;
    const uint8_t* ref_arena;
    const uint32_t* ids;
    const float* refine_tab;
    const double* gate_mid;    // [16]
    // Per cluster, over its scan records: {min base_min (sc0), max W_step
    // (sc3), max |W|max (sc6), max Omega_max = Omega_min + 255 Omega_step}.
    // The batch-skip bound is monotone in each, so these give a lower bound
    // for every batch of the cluster at once.
    const float4* csum;
};

// Encoder tables (device): merged sorted level table for b >= 2, the four
// single-class tables for b == 1, and the per-branch level tables for fits.
struct EncodeView {
    int bits;
    int merged;                // 1: period-4 interleave verified
    uint32_t mlen;             // 2^(bits+1)
    double scale;              // s(b)
    const double* mlev;        // merged sorted levels
    const uint16_t* mdig;      // their digits
    const double* levels;      // [2][2^bits]
    const double* cls_lev;     // b == 1: [branch][parity] single level
    const uint16_t* cls_dig;
};

__device__ __forceinline__ uint32_t FloatBits(float f) { return __float_as_uint(f); }

// AVX horizontal-sum order over 8 values held by 8 consecutive lanes
// (lane base+i holds v[i]): ((v0+v4)+(v2+v6)) + ((v1+v5)+(v3+v7)).  Every
// lane of the 8-group returns the sum.  simd_helpers.hpp:12-18.
__device__ __forceinline__ float Hsum8Lanes(float v) {
    const float s4 = v + __shfl_xor(v, 4, 8);        // lane i<4: v[i]+v[i+4]
    const float s2 = s4 + __shfl_xor(s4, 2, 8);      // lane 0: s4[0]+s4[2], lane 1: s4[1]+s4[3]
    const float s1 = s2 + __shfl_xor(s2, 1, 8);      // (s4[0]+s4[2]) + (s4[1]+s4[3])
    return __shfl(s1, 0, 8);
}

// ---------------------------------------------------------------------------
// Rotation: one workgroup per row; the row lives in LDS for the 4 rounds of
// {sign flip, FHT over the power-of-two window, 1/sqrt(window)} with a Kac
// walk between rounds when the padded length exceeds the window.
// P = rotated length (a multiple of 64), T = its power-of-two window,
// fac = 1/sqrt(T), signs = 4*P/8 bytes.
__device__ inline void RotateInLds(float* buf, uint32_t P, uint32_t T, float fac, const uint8_t* signs) {
    const bool kac = T != P;
    const uint32_t per_round = P / 8;
    for (int r = 0; r < 4; ++r) {
        const uint8_t* sg = signs + r * per_round;
        for (uint32_t i = threadIdx.x; i < P; i += blockDim.x)
            if ((sg[i >> 3] >> (i & 7)) & 1) buf[i] = -buf[i];
        __syncthreads();
        float* win = buf + ((!kac || (r & 1) == 0) ? 0 : P - T);
        for (uint32_t h = 1, lg = 0; h < T; h <<= 1, ++lg) {   // h = 2^lg: pair j, j + h of butterfly t
            for (uint32_t t = threadIdx.x; t < T / 2; t += blockDim.x) {
                const uint32_t j = ((t >> lg) << (lg + 1)) | (t & (h - 1));
                const float a = win[j], b = win[j + h];
                win[j] = a + b;
                win[j + h] = a - b;
            }
            __syncthreads();
        }
        for (uint32_t i = threadIdx.x; i < T; i += blockDim.x) win[i] = win[i] * fac;
        __syncthreads();
        if (kac) {
            const uint32_t half = P / 2;
            for (uint32_t i = threadIdx.x; i < half; i += blockDim.x) {
                const float a = buf[i], b = buf[i + half];
                buf[i] = a + b;
                buf[i + half] = a - b;
            }
            __syncthreads();
        }
    }
    if (kac) {
        for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) buf[i] = buf[i] * 0.25F;
        __syncthreads();
    }
}

__device__ inline void RotateInLds(float* buf, const IndexView& ix) {
    RotateInLds(buf, static_cast<uint32_t>(ix.padded), ix.trunc, ix.fac, ix.signs);
}

}  // namespace e8qdev
