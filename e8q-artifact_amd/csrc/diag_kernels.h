// diag_kernels.h -- launchers of the diagnostics kernels (diag.hip): the
// reference's Estimate / ProbeBoundW / ExhaustiveTopK (E8Q) and Estimate /
// ExhaustiveTopK / SearchIVF (ME8Q) on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "e8q_device.h"
#include "me8q_device.h"

namespace e8qdev {

constexpr uint32_t kDiagMaxK = 1024;   // top-k held by one workgroup

// inv[id] = 1 + the last position holding id (0: none); inv cleared first.
hipError_t LaunchDiagInverseIds(hipStream_t st, const uint32_t* ids, uint64_t n, uint32_t* inv);

// E8Q.  rotq: centered rotations of the queries (LaunchRotate).  `bad` is set
// for a pair whose id no stored vector carries.
hipError_t LaunchE8Estimate(hipStream_t st, const IndexView& ix, const float* rotq, const uint32_t* qidx,
                            const uint32_t* ids, uint64_t npair, const uint32_t* inv, float* out, uint32_t* bad);
// rotx: centered rotations of the pairs' base rows; levels: [2][2^bits] fp64.
// NaN for a pair with qidx >= nq or id >= n (as the binding's batch form).
hipError_t LaunchE8ProbeBoundW(hipStream_t st, const IndexView& ix, const double* levels, const float* rotq,
                               const float* rotx, const uint32_t* qidx, const uint32_t* ids, uint64_t nq,
                               uint64_t npair, const uint32_t* inv, double* out, uint32_t* bad);
// d2c[q * nlist + c]: the list distances ExhaustiveTopK uses.
hipError_t LaunchE8ListD2(hipStream_t st, const IndexView& ix, const float* rotq, uint64_t nq, float* d2c);
// dist[(q - q0) * n + i]: every stored vector's distance, queries q0..q0+nqc.
hipError_t LaunchE8Exhaustive(hipStream_t st, const IndexView& ix, const float* rotq, const float* d2c, uint64_t q0,
                              uint64_t nqc, float* dist);

// ME8Q.  queries: raw (nq x dim); rotq: per-run rotations (LaunchM8QueryPrep).
hipError_t LaunchM8Estimate(hipStream_t st, const M8View& v, const float* queries, const float* rotq,
                            const uint32_t* qidx, const uint32_t* ids, uint64_t npair, const uint32_t* inv, float* out,
                            uint32_t* bad);
hipError_t LaunchM8ListD2(hipStream_t st, const M8View& v, const float* queries, uint64_t nq, double* d2c);
hipError_t LaunchM8Exhaustive(hipStream_t st, const M8View& v, const float* rotq, const double* d2c, uint64_t q0,
                              uint64_t nqc, float* dist);
// SearchIVF's candidates: off [nq][take+1] probe offsets, count [nq]; then
// per query (q - q0) at dist/cid + (q - q0) * cap in probe order.
hipError_t LaunchM8IvfOffsets(hipStream_t st, const M8View& v, const uint32_t* probe, uint32_t maxp, uint32_t take,
                              uint64_t nq, uint64_t* off, uint64_t* count);
hipError_t LaunchM8Ivf(hipStream_t st, const M8View& v, const float* rotq, const uint32_t* probe, const float* pd,
                       uint32_t maxp, uint32_t take, const uint64_t* off, uint64_t q0, uint64_t nqc, uint64_t cap,
                       float* dist, uint32_t* cid);

// Top-k of queries q0..q0+nqc over their candidates (dist + (q - q0) *
// stride, ids idv + (q - q0) * id_stride, count[q] of them or m_fixed):
// push = false by (distance, id), push = true in PushCandidate's order.
// Outputs at [q * k], UINT32_MAX / +inf padded.  k <= kDiagMaxK.
hipError_t LaunchDiagSelect(hipStream_t st, bool push, const float* dist, const uint32_t* idv, uint64_t id_stride,
                            const uint64_t* count, uint64_t m_fixed, uint64_t stride, uint64_t q0, uint64_t nqc,
                            uint32_t k, uint32_t* out_ids, float* out_d);

}  // namespace e8qdev
