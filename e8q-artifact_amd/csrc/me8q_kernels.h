// me8q_kernels.h -- host-callable launchers of the ME8Q kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "me8q_device.h"

namespace e8qdev {

// ---- build (me8q_build.hip)
hipError_t LaunchM8Centroids(hipStream_t st, const M8View& v);
size_t M8EncodeLds(const M8View& v);
// rank[row] = position of `row` inside its list (the reference's list_rank,
// me8q_build.cpp:154-163).
hipError_t LaunchM8Encode(hipStream_t st, const M8View& v, const float* base, const uint32_t* cid,
                          const uint32_t* rank, uint64_t n);
// perp: n*nrun float scratch; the plane buffer must be zero beforehand.
hipError_t LaunchM8Fast(hipStream_t st, const M8View& v, float* perp, uint64_t nbatch);

// ---- search (me8q_search.hip)
// Per-query prep: padded, per-run rotated query (rotq [nq][padded]), the
// quantized class-mean LUT (lut [nq][plane_cols*32]: per plane column 16
// low-nibble then 16 high-nibble bytes), qc [nq][4] = {lut_min, lut_delta,
// |q|^2 over the route prefix, bits per row} and rc [nq][nrun][4] doubles =
// {|q_r|^2, lut_min_x, qerr, score_max}.
hipError_t LaunchM8QueryPrep(hipStream_t st, const M8View& v, const float* queries, uint64_t nq, float* rotq,
                             uint8_t* lut, float* qc, double* rc);
// sqd2 [nq][maxp][nrun] = sqrt(max(|q_r - c_r|^2, 0)) per probed list.
hipError_t LaunchM8RunDist(hipStream_t st, const M8View& v, const float* rotq, const double* rc,
                           const uint32_t* probe, const uint32_t* nprobe_q, uint32_t nprobe_all, uint32_t maxp,
                           uint64_t nq, double* sqd2);

struct M8ScanArgs {
    M8View v;
    const float* rotq;
    const uint8_t* lut;
    const float* qc;
    const double* rc;
    const uint32_t* probe;
    const float* pdist;
    const double* sqd2;
    const uint32_t* nprobe_q;
    uint32_t nprobe_all, maxp;
    uint64_t nq;
    uint32_t k;
    float z;
    uint32_t* out_ids;
    float* out_d;
    uint64_t* qstats;   // [nq][8] candidates, survivors, bits_scored, bits_total, pruned, batches read,
                        //         candidates refined, refine groups
    uint32_t lev_floats;   // total size of the runs' level tables
};
size_t M8ScanLds(const M8View& v, uint32_t k, uint32_t lev_floats);
hipError_t LaunchM8Scan(hipStream_t st, const M8ScanArgs& a);

}  // namespace e8qdev
