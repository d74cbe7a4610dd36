// e8q_kernels.h -- kernel launchers (host-callable) of the E8Q device library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "e8q_device.h"

namespace e8qdev {

struct PlainArgs {
    uint64_t nq;
    const float* rotq;        // [nq][padded] rotated, centered queries
    const uint8_t* lut;       // [nq][blocks*32] compact u8 LUT
    const float* consts;      // [nq][8]: delta, base, lut_err, score_max, qn2
    const uint32_t* probe_ids;  // [nq][maxp]
    const float* probe_d2;      // [nq][maxp]
    uint32_t maxp;
    const uint32_t* nprobe_q;   // optional [nq]
    uint32_t nprobe_all;
    uint64_t k;
    float z;
    int disable_prune;
    int flush_every;
    uint32_t* out_ids;          // [nq][k]
    float* out_dists;
    unsigned long long* stats;  // [4] candidates, survivors, batches, records read
};

hipError_t LaunchRotate(hipStream_t st, const float* x, uint64_t rows, uint32_t dim, const IndexView& ix, float* out,
                        bool center);
hipError_t LaunchQueryPrep(hipStream_t st, const float* rotq, uint64_t nq, const IndexView& ix, uint8_t* lut,
                           float* consts);
hipError_t LaunchRoute(hipStream_t st, const float* rotq, const float* consts, uint64_t nq, const IndexView& ix,
                       float* dist);
size_t SelectProbesLds(uint32_t nlist, uint32_t maxp);
hipError_t LaunchSelectProbes(hipStream_t st, const float* dist, uint64_t nq, uint32_t nlist,
                              const uint32_t* nprobe_q, uint32_t nprobe_all, uint32_t maxp, uint32_t* probe_ids,
                              float* probe_d2);
// The plain scan reads its parameters from device memory (`dparams`, filled
// by the launcher on the stream) so that rarely used fields are loaded where
// they are used instead of occupying scalar registers for the whole kernel.
struct PlainParams {
    IndexView ix;
    PlainArgs a;
};
hipError_t LaunchPlainSearch(hipStream_t st, const IndexView& ix, const PlainArgs& a, PlainParams* dparams);
hipError_t LaunchCoarseAccumulate(hipStream_t st, const uint8_t* rowa, uint64_t nbatch, const uint8_t* lut, int blocks,
                                  int32_t* out);

// Build (encode) path.
struct EncodeArgs {
    const float* base;          // [n][dim] device
    const uint32_t* cluster_ids;  // [n]
    const uint32_t* slot;       // [n] cluster-major slot of each row
    uint64_t n;
    uint8_t* arena;
    uint8_t* ref_arena;
    uint32_t* ids;
    float* ob;                  // [n] by slot: bank inputs base / W / Omega
    float* ow;
    float* op;
};

hipError_t LaunchMean(hipStream_t st, const float* base, uint64_t n, uint32_t dim, double* mu);
hipError_t LaunchMeanFinish(hipStream_t st, const double* mu, uint64_t n, const IndexView& ix, float* mean_rot);
hipError_t LaunchCentroids(hipStream_t st, const float* centroids, const IndexView& ix, float* centroid_rot,
                           float* centroid_n2);
hipError_t LaunchEncodeRows(hipStream_t st, const IndexView& ix, const EncodeView& ev, const EncodeArgs& a);
hipError_t LaunchBanks(hipStream_t st, const IndexView& ix, const float* ob, const float* ow, const float* op,
                       uint8_t* arena);
// Scan records [nrec] between the E8AX rowA layout ([blk][lane]) and the
// device layout ([lane][blk]); src may equal dst.
hipError_t LaunchRecordLayout(hipStream_t st, const IndexView& ix, const uint8_t* src, uint8_t* dst, uint64_t nrec,
                              bool to_device);
// Fills ix.csum from the scan-record headers (after build or upload).
hipError_t LaunchClusterSummary(hipStream_t st, const IndexView& ix, float4* csum);
hipError_t LaunchEncodeUnits(hipStream_t st, const IndexView& ix, const EncodeView& ev, const float* unit,
                             uint64_t rows, double scale, uint16_t* digits, uint8_t* branch);

}  // namespace e8qdev
