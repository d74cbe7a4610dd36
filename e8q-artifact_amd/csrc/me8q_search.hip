// me8q_search.hip -- ME8Q FastScan search on the device (a13-a20):
//
//   M8QueryPrepKernel  per query: zero-pad, rotate each coded run with its
//                      own FHT-Kac-4 (me8q_search.cpp:26-43), the class-mean
//                      LUT per plane column and its u8 quantization
//                      (me8q_fast.cpp:627-727, fastscan_avx2.cpp:8-51), the
//                      route prefix norm and the per-run constants.
//   M8RunDistKernel    per (query, probed list, run): d2_r = |q_r|^2 +
//                      |c_r|^2 - 2 q_r.c_r in fp64 (me8q_fast.cpp:756-770).
//   M8ScanKernel       one wave per query, lists in probe order: per run the
//                      integer coarse sums of two batches at once (a17), the
//                      fp32 coarse prune of whole groups of 8 lanes (a18,
//                      me8q_prune_avx2.cpp:9-98), the fp64 lane test, the
//                      four-candidate exact cascade (a19, me8q_fast.cpp:
//                      856-979) and the sorted top-k (a20, me8q_search.cpp:
//                      121-134).
//
// The reference's schedule is sequential per query: the k-th distance only
// changes when four staged candidates are refined, so the kernel tests all
// remaining kept lanes of a batch at once against the current threshold,
// stages the first four that pass (lane order), refines them, and tests
// the lanes after the last staged one again with the new threshold.  That
// is the reference's lane-by-lane loop, result for result.
#include <hip/hip_runtime.h>

#include "e8q_kernels.h"
#include "me8q_device.h"
#include "me8q_kernels.h"

namespace e8qdev {

constexpr int kM8PrepThreads = 256;
constexpr uint32_t kM8PassWords = 4;   // 8-byte words per lane of a scoring pass held in registers

__global__ __launch_bounds__(kM8PrepThreads) void M8QueryPrepKernel(M8View v, const float* __restrict__ queries,
                                                                    uint64_t nq, float* __restrict__ rotq,
                                                                    uint8_t* __restrict__ lut, float* __restrict__ qc,
                                                                    double* __restrict__ rc) {
    extern __shared__ float sm[];
    const uint32_t P = static_cast<uint32_t>(v.padded);
    const uint32_t PC = v.plane_cols;
    float* buf = sm;            // padded
    float* lf = sm + P;         // PC * 32 float LUT (16 low-nibble, 16 high-nibble per column)
    __shared__ float s_lo[kM8PrepThreads], s_hi[kM8PrepThreads];
    const int tid = threadIdx.x;
    for (uint64_t qi = blockIdx.x; qi < nq; qi += gridDim.x) {
        for (uint32_t d = tid; d < P; d += kM8PrepThreads) buf[d] = d < v.dim ? queries[qi * v.dim + d] : 0.0F;
        __syncthreads();
        for (uint32_t r = 0; r < v.nrun; ++r) {
            const M8Run& ru = v.runs[r];
            RotateInLds(buf + ru.dim_off, ru.len, ru.trunc, ru.fac, ru.signs);
        }
        for (uint32_t s = 0; s < v.nseg; ++s) {
            const M8Seg sg = v.segs[s];
            if (sg.bits != 0) continue;
            for (uint32_t d = tid; d < sg.len; d += kM8PrepThreads) buf[sg.off + d] = 0.0F;
        }
        __syncthreads();
        for (uint32_t d = tid; d < P; d += kM8PrepThreads) rotq[qi * P + d] = buf[d];
        // Class-mean LUT.  `qb4*a4 + qb5*a5 + qb6*a6 + qb7*a7` as the
        // reference build evaluates it: the first product fused into the
        // first add, then one fused multiply-add per further term.
        for (uint32_t e = tid; e < PC * 16; e += kM8PrepThreads) {
            const uint32_t col = e >> 4, val = e & 15;
            uint32_t r = 0;
            while (r + 1 < v.nrun && v.runs[r + 1].blk0 <= col) ++r;
            const M8Run& ru = v.runs[r];
            const float p0 = ru.pbar0, p1 = ru.pbar1;
            const float p70 = ru.rate == 1 ? 0.0F : p0, p71 = ru.rate == 1 ? 0.0F : p1;
            const float* qb = buf + ru.dim_off + (col - ru.blk0) * 8;
            const float a4 = (val >> 3) & 1 ? p1 : p0, a5 = (val >> 2) & 1 ? p1 : p0, a6 = (val >> 1) & 1 ? p1 : p0;
            const float a7 = val & 1 ? p71 : p70;
            const float a3 = val & 1 ? p1 : p0;
            lf[col * 32 + val] = fmaf(qb[7], a7, fmaf(qb[6], a6, fmaf(qb[4], a4, qb[5] * a5)));
            lf[col * 32 + 16 + val] = fmaf(qb[3], a3, fmaf(qb[2], a6, fmaf(qb[0], a4, qb[1] * a5)));
        }
        __syncthreads();
        float lo = __builtin_inff(), hi = -__builtin_inff();
        for (uint32_t e = tid; e < PC * 32; e += kM8PrepThreads) {
            lo = fminf(lo, lf[e]);
            hi = fmaxf(hi, lf[e]);
        }
        s_lo[tid] = lo;
        s_hi[tid] = hi;
        __syncthreads();
        for (int w = kM8PrepThreads / 2; w > 0; w >>= 1) {
            if (tid < w) {
                s_lo[tid] = fminf(s_lo[tid], s_lo[tid + w]);
                s_hi[tid] = fmaxf(s_hi[tid], s_hi[tid + w]);
            }
            __syncthreads();
        }
        lo = PC ? s_lo[0] : 0.0F;
        hi = PC ? s_hi[0] : 0.0F;
        const float delta = hi > lo ? (hi - lo) / 255.0F : 1.0F;
        for (uint32_t e = tid; e < PC * 32; e += kM8PrepThreads) {
            const float t = (lf[e] - lo) / delta + 0.5F;
            const int iv = static_cast<int>(t);
            lut[qi * PC * 32 + e] = static_cast<uint8_t>(min(max(iv, 0), 255));
        }
        if (tid == 0) {
            // |q|^2 over the route prefix in the AVX2 order (8 FMA chains).
            float acc8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (uint32_t d = 0; d < v.route_dim; d += 8)
                for (int l = 0; l < 8; ++l) acc8[l] = fmaf(buf[d + l], buf[d + l], acc8[l]);
            const float qn2 = ((acc8[0] + acc8[4]) + (acc8[2] + acc8[6])) + ((acc8[1] + acc8[5]) + (acc8[3] + acc8[7]));
            uint32_t bits_row = 0;
            for (uint32_t r = 0; r < v.nrun; ++r) bits_row += v.runs[r].rate * v.runs[r].len;
            qc[qi * 4 + 0] = lo;
            qc[qi * 4 + 1] = delta;
            qc[qi * 4 + 2] = qn2;
            qc[qi * 4 + 3] = static_cast<float>(bits_row);
        }
        if (tid >= 64 && static_cast<uint32_t>(tid - 64) < v.nrun) {
            const uint32_t r = tid - 64;
            const M8Run& ru = v.runs[r];
            double a = 0.0;
            for (uint32_t t = 0; t < ru.len; ++t) {
                const double x = buf[ru.dim_off + t];
                a = fma(x, x, a);
            }
            const double lmx = static_cast<double>(lo) * (2.0 * static_cast<double>(ru.nblk));
            const float sq = static_cast<float>(sqrt(static_cast<double>(ru.nblk > 1 ? ru.nblk : 1)));
            const double qerr = static_cast<double>(2.0F * sq * delta);
            const double hv = fma(2.0 * static_cast<double>(ru.nblk) * 255.0, static_cast<double>(delta), lmx);
            double* o = rc + (qi * v.nrun + r) * 4;
            o[0] = a;
            o[1] = lmx;
            o[2] = qerr;
            o[3] = fmax(fabs(lmx), fabs(hv));
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void M8RunDistKernel(M8View v, const float* __restrict__ rotq,
                                                       const double* __restrict__ rc,
                                                       const uint32_t* __restrict__ probe,
                                                       const uint32_t* __restrict__ nprobe_q, uint32_t nprobe_all,
                                                       uint32_t maxp, uint64_t nq, double* __restrict__ sqd2) {
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t per_q = static_cast<uint64_t>(maxp) * v.nrun;
    if (t >= nq * per_q) return;
    const uint64_t qi = t / per_q;
    const uint32_t p = static_cast<uint32_t>((t % per_q) / v.nrun);
    const uint32_t r = static_cast<uint32_t>(t % v.nrun);
    const uint64_t np = min<uint64_t>(nprobe_q ? nprobe_q[qi] : nprobe_all, v.nlist);
    if (p >= np) return;
    const uint32_t c = probe[qi * maxp + p];
    const M8Run& ru = v.runs[r];
    const float* qr = rotq + qi * v.padded + ru.dim_off;
    const float* cr = v.cent_rot + static_cast<uint64_t>(c) * v.padded + ru.dim_off;
    double dot = 0.0;
    for (uint32_t j = 0; j < ru.len; ++j) dot = fma(static_cast<double>(qr[j]), static_cast<double>(cr[j]), dot);
    const double d2r = fma(-2.0, dot, rc[(qi * v.nrun + r) * 4] + v.rc_n2[static_cast<uint64_t>(c) * v.nrun + r]);
    sqd2[t] = sqrt(fmax(d2r, 0.0));
}

// ---------------------------------------------------------------------------
// A run's scoring constants, copied to LDS once per workgroup: the scoring
// passes read them per lane, for a different run in each 16-lane group.
struct M8RunTab {
    uint32_t rate, nblk, blk0, ref_off;
    uint32_t dim_off, lo_off;   // first dim; its level table's offset in the LDS tables
    uint32_t nwp, nw;           // 8-byte words of its plane bytes, and of plane bytes + fields
    const float* lev;           // its level table in global memory
    uint64_t bits;              // rate x len: the bits a cascade step scores
};
static_assert(sizeof(M8RunTab) == 48, "LDS run table entry");
typedef const __attribute__((address_space(3))) M8RunTab* LdsRun;

// Scan: LDS layout of one wave (one query).
struct M8ScanLayout {
    uint32_t q, lut, td, ti, acc, fdc, fpool, fqa, fsc, ftr, tsc, ttr, ssc, str, sord, smean, rq, rqf, ssq,
        stage, srend, slev, sS32, plist, runbuf, hdr, idl, rtab, total;
    __host__ __device__ M8ScanLayout(uint32_t P, uint32_t PC, uint32_t R, uint32_t K, uint32_t RS, uint32_t LV,
                                     uint32_t SB) {
        uint32_t o = 0;
        auto take = [&](uint32_t bytes, uint32_t align) {
            o = (o + align - 1) & ~(align - 1);
            const uint32_t at = o;
            o += bytes;
            return at;
        };
        const uint32_t R1 = R > 0 ? R : 1;
        q = take(P * 4, 16);
        lut = take(PC * 32 + 16, 16);
        td = take(K * 4, 16);
        ti = take(K * 4, 16);
        acc = take(R1 * 64 * 4, 16);
        fdc = take(64 * 4, 16);
        fpool = take(64 * 4, 16);
        fqa = take(64 * 4, 16);
        fsc = take(R1 * 64 * 4, 16);
        ftr = take(R1 * 64 * 4, 16);
        tsc = take(R1 * 8 * 8, 16);
        ttr = take(R1 * 8 * 8, 16);
        ssc = take(4 * R1 * 8, 16);
        str = take(4 * R1 * 8, 16);
        sord = take(R1 * 4, 16);
        smean = take(R1 * 8, 16);
        rq = take(R1 * 3 * 8, 16);      // lut_min_x, qerr, score_max
        rqf = take(R1 * 2 * 4, 16);     // float lut_min_x, float qerr
        ssq = take(R1 * 8, 16);         // sqrt(d2_r) of the current list
        srend = take((R1 + 1) * 4, 16); // last plane column + 1 of each run
        slev = take(LV * 4, 16);        // the runs' level tables, [2][2^rate] each
        sS32 = take(32 * R1 * 4, 16);   // exact run scores of a batch's passing lanes
        plist = take(32, 16);           // those lanes, in order
        stage = take(4 * 64, 16);       // staged candidates: lane,row,pool,qadd,dist,died,bits
        runbuf = take(4 * SB, 16);      // per scored (lane, run) pair of a pass: its plane bytes, then its fields
        hdr = take(32 * R1 * 16, 16);   // a batch's passing lanes' refine row headers (R float4)
        idl = take(32 * 4, 16);         // ... and their ids
        rtab = take(R1 * sizeof(M8RunTab), 16);   // the runs' scoring constants
        total = (o + 15) & ~15u;
    }
};

struct M8Stage {         // one staged candidate
    uint32_t lane, row, died, pad;
    double pool, qadd, dist;
    uint64_t bits;
    uint64_t pad2[2];
};

// Inserts (d, id) into the sorted list (td, ti) of `cnt` entries (<= K):
// a full list rejects d unless it is below the worst; a new item goes before
// equal ones (lower_bound by distance).  Called by the whole wave with
// uniform arguments.
__device__ void M8Push(float* td, uint32_t* ti, uint32_t& cnt, uint32_t K, float d, uint32_t id) {
    const int lane = threadIdx.x;
    if (cnt >= K && !(d < td[K - 1])) return;
    if (cnt >= K) cnt = K - 1;
    uint32_t pos = 0;
    for (uint32_t e0 = 0; e0 < cnt; e0 += 64) {
        const uint32_t e = e0 + lane;
        pos += __popcll(__ballot(e < cnt && td[e] < d));
    }
    if (cnt > pos) {
        for (int32_t s = static_cast<int32_t>((cnt - 1) & ~63u); s >= static_cast<int32_t>(pos & ~63u); s -= 64) {
            const uint32_t e = static_cast<uint32_t>(s) + lane;
            const bool mv = e >= pos && e < cnt;
            float x = 0.0F;
            uint32_t y = 0;
            if (mv) {
                x = td[e];
                y = ti[e];
            }
            __syncthreads();
            if (mv) {
                td[e + 1] = x;
                ti[e + 1] = y;
            }
            __syncthreads();
        }
    }
    if (lane == 0) {
        td[pos] = d;
        ti[pos] = id;
    }
    __syncthreads();
    ++cnt;
}

__global__ __launch_bounds__(64) void M8ScanKernel(M8ScanArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const M8View& v = a.v;
    const uint32_t P = static_cast<uint32_t>(v.padded), PC = v.plane_cols, R = v.nrun, K = a.k;
    const M8ScanLayout L(P, PC, R, K, v.rstride, a.lev_floats, a.run_slot);
    float* q = reinterpret_cast<float*>(smem + L.q);
    uint8_t* lut = smem + L.lut;
    float* td = reinterpret_cast<float*>(smem + L.td);
    uint32_t* ti = reinterpret_cast<uint32_t*>(smem + L.ti);
    int32_t* sacc = reinterpret_cast<int32_t*>(smem + L.acc);
    float* fdc = reinterpret_cast<float*>(smem + L.fdc);
    float* fpool = reinterpret_cast<float*>(smem + L.fpool);
    float* fqa = reinterpret_cast<float*>(smem + L.fqa);
    float* fsc = reinterpret_cast<float*>(smem + L.fsc);
    float* ftr = reinterpret_cast<float*>(smem + L.ftr);
    double* tsc = reinterpret_cast<double*>(smem + L.tsc);
    double* ttr = reinterpret_cast<double*>(smem + L.ttr);
    double* ssc = reinterpret_cast<double*>(smem + L.ssc);
    double* str = reinterpret_cast<double*>(smem + L.str);
    uint32_t* sord = reinterpret_cast<uint32_t*>(smem + L.sord);
    double* smean = reinterpret_cast<double*>(smem + L.smean);
    double* rq = reinterpret_cast<double*>(smem + L.rq);
    float* rqf = reinterpret_cast<float*>(smem + L.rqf);
    double* ssq = reinterpret_cast<double*>(smem + L.ssq);
    M8Stage* stg = reinterpret_cast<M8Stage*>(smem + L.stage);
    uint32_t* srend = reinterpret_cast<uint32_t*>(smem + L.srend);
    float* slev = reinterpret_cast<float*>(smem + L.slev);
    float* sS32 = reinterpret_cast<float*>(smem + L.sS32);
    uint8_t* plist = smem + L.plist;
    uint8_t* runbuf = smem + L.runbuf;
    float* hdrl = reinterpret_cast<float*>(smem + L.hdr);
    uint32_t* idl = reinterpret_cast<uint32_t*>(smem + L.idl);
    // level tables of every run, copied once per workgroup; run r's table
    // starts at slev + lev_off(r)
    // Level tables of the runs at rates <= kM8LdsLevRate (at most 256
    // floats each) live in LDS, run after run; wider tables are read from
    // global memory.
    {
        uint32_t o = 0;
        for (uint32_t r = 0; r < R; ++r) {
            if (v.runs[r].rate > kM8LdsLevRate) continue;
            const uint32_t nq2 = 2u << v.runs[r].rate;
            for (uint32_t e = threadIdx.x; e < nq2; e += 64) slev[o + e] = v.runs[r].lev[e];
            o += nq2;
        }
    }
    M8RunTab* rtab = reinterpret_cast<M8RunTab*>(smem + L.rtab);
    for (uint32_t r = threadIdx.x; r < R; r += 64) {
        const M8Run& ru = v.runs[r];
        uint32_t lo = 0;
        for (uint32_t r2 = 0; r2 < r; ++r2)
            if (v.runs[r2].rate <= kM8LdsLevRate) lo += 2u << v.runs[r2].rate;
        M8RunTab t;
        t.rate = ru.rate;
        t.nblk = ru.nblk;
        t.blk0 = ru.blk0;
        t.ref_off = ru.ref_off;
        t.dim_off = ru.dim_off;
        t.lo_off = lo;
        t.nwp = ru.nblk / 8;
        t.nw = t.nwp + ru.nblk * (ru.rate - 1) / 8;
        t.lev = ru.lev;
        t.bits = static_cast<uint64_t>(ru.rate) * ru.len;
        rtab[r] = t;
    }
    const LdsRun rtl = (LdsRun)(rtab);
    const int lane = threadIdx.x;
    const float finf = __builtin_inff();
    const double dinf = __builtin_inf();
    const float z = a.z;
    const double dz = z;

    const uint64_t nrun_q = a.qsel ? (a.nsel_dev ? min<uint64_t>(a.nsel, *a.nsel_dev) : a.nsel) : a.nq;
    for (uint64_t qj = blockIdx.x; qj < nrun_q; qj += gridDim.x) {
        const uint64_t qi = a.qsel ? a.qsel[qj] : qj;
        for (uint32_t d = lane; d < P; d += 64) q[d] = a.rotq[qi * P + d];
        for (uint32_t e = lane; e < PC * 8; e += 64)
            reinterpret_cast<uint32_t*>(lut)[e] = reinterpret_cast<const uint32_t*>(a.lut + qi * PC * 32)[e];
        for (uint32_t r = lane; r <= R; r += 64) srend[r] = r < R ? v.runs[r].blk0 + v.runs[r].nblk : 0xFFFFFFFFu;
        for (uint32_t r = lane; r < R; r += 64) {
            const double* o = a.rc + (qi * R + r) * 4;
            rq[r * 3 + 0] = o[1];
            rq[r * 3 + 1] = o[2];
            rq[r * 3 + 2] = o[3];
            rqf[r * 2 + 0] = static_cast<float>(o[1]);
            rqf[r * 2 + 1] = static_cast<float>(o[2]);
        }
        const float lut_delta = a.qc[qi * 4 + 1];
        const uint64_t bits_row = static_cast<uint64_t>(a.qc[qi * 4 + 3]);
        uint32_t cnt = 0;
        uint64_t n_cand = 0, n_surv = 0, n_bits = 0, n_btot = 0, n_batch = 0, n_staged = 0, n_flush = 0;
        __syncthreads();
        const uint64_t np = min<uint64_t>(a.nprobe_q ? a.nprobe_q[qi] : a.nprobe_all, v.nlist);
        uint64_t used = 0, p_stop = np, b_stop = 0;   // the threshold pass's budget and where it stopped
        for (uint64_t p = 0; p < np; ++p) {
            const uint32_t c = a.probe[qi * a.maxp + p];
            const float d2f = a.pdist[qi * a.maxp + p];
            const double d2 = d2f;
            const uint64_t lb = v.list_begin[c];
            const uint64_t cntc = v.list_begin[c + 1] - lb;
            if (R == 0) {   // no coded run: every row estimates to the list distance
                for (uint64_t i = 0; i < cntc; ++i) M8Push(td, ti, cnt, K, d2f, v.ids[lb + i]);
                n_cand += cntc;
                n_surv += cntc;
                continue;
            }
            if (cntc == 0) continue;
            for (uint32_t r = lane; r < R; r += 64) ssq[r] = a.sqd2[(qi * a.maxp + p) * R + r];
            const uint64_t nbl = (cntc + 31) >> 5;
            const uint64_t nb = a.budget ? min<uint64_t>(nbl, a.budget - used) : nbl;
            const uint64_t bat0 = v.batch_begin[c];
            n_batch += nb;
            for (uint64_t b0 = 0; b0 < nb; b0 += 2) {
                {   // integer coarse sums of batches b0 (lanes 0..31) and b0+1 (32..63)
                    const uint64_t b = b0 + (lane >> 5);
                    const uint32_t l = lane & 31;
                    if (b < nb) {
                        const uint8_t* rowp = v.plane + (bat0 + b) * v.rec + l * PC;
                        uint32_t r = 0, rend = srend[0];
                        int32_t s = 0;
                        auto sum8 = [&](uint2 w, uint32_t col) {   // eight columns' lookups into the current run
                            if (col == rend) {
                                sacc[r * 64 + lane] = s;
                                s = 0;
                                ++r;
                                rend = srend[r];
                            }
                            const uint8_t* lc = lut + col * 32;
#pragma unroll
                            for (int k2 = 0; k2 < 4; ++k2) {
                                const uint32_t x0 = (w.x >> (8 * k2)) & 0xFFu, x1 = (w.y >> (8 * k2)) & 0xFFu;
                                s += lc[k2 * 32 + (x0 & 15)] + lc[k2 * 32 + 16 + (x0 >> 4)];
                                s += lc[(4 + k2) * 32 + (x1 & 15)] + lc[(4 + k2) * 32 + 16 + (x1 >> 4)];
                            }
                        };
                        if (PC <= 128) {
                            // the whole row in one burst of independent loads
                            uint2 w[16];
#pragma unroll
                            for (uint32_t u = 0; u < 16; ++u)
                                w[u] = 8 * u < PC ? *reinterpret_cast<const uint2*>(rowp + 8 * u) : uint2{0u, 0u};
#pragma unroll
                            for (uint32_t u = 0; u < 16; ++u)
                                if (8 * u < PC) sum8(w[u], 8 * u);
                        } else {
                            // runs occupy consecutive plane columns, 8 per word pair: four
                            // word pairs are loaded before any is summed
                            for (uint32_t j0 = 0; j0 < PC; j0 += 32) {
                                uint2 w[4];
#pragma unroll
                                for (int u = 0; u < 4; ++u) {
                                    const uint32_t col = j0 + 8 * u;
                                    w[u] = col < PC ? *reinterpret_cast<const uint2*>(rowp + col) : uint2{0u, 0u};
                                }
#pragma unroll
                                for (int u = 0; u < 4; ++u)
                                    if (j0 + 8 * u < PC) sum8(w[u], j0 + 8 * u);
                            }
                        }
                        sacc[r * 64 + lane] = s;
                    }
                }
                __syncthreads();
                for (uint32_t hh = 0; hh < 2 && b0 + hh < nb; ++hh) {
                    const uint64_t bb = b0 + hh;
                    const uint32_t cntb = static_cast<uint32_t>(min<uint64_t>(32, cntc - bb * 32));
                    const uint32_t covered = cntb >= 8 ? (cntb & ~7u) : 0u;
                    const uint8_t* rec = v.plane + (bat0 + bb) * v.rec;
                    const uint8_t* banks = rec + 32u * PC;
                    const bool mine = static_cast<uint32_t>(lane >> 5) == hh;
                    const uint32_t l = lane & 31;
                    const float distk = cnt < K ? finf : td[K - 1];
                    const float sqk = distk < finf ? sqrtf(fmaxf(0.0F, distk)) : finf;
                    bool kept = false;
                    if (mine && l < covered) {   // fp32 coarse prune (AVX2 filter)
                        float dc = d2f, pool = 0.0F, qa = 0.0F, wpad = 0.0F;
                        for (uint32_t r = 0; r < R; ++r) {
                            const uint8_t* bp = banks + r * kM8BankBytes;
                            const float* sc = reinterpret_cast<const float*>(bp + 128);
                            const float s_c = fmaf(static_cast<float>(sacc[r * 64 + lane]), lut_delta, rqf[r * 2]);
                            const float bq = fmaf(static_cast<float>(bp[l]), sc[1], sc[0]);
                            const float wq = fmaf(static_cast<float>(bp[32 + l]), sc[3], sc[2]);
                            const float pq = fmaf(static_cast<float>(bp[64 + l]), sc[5], sc[4]);
                            dc = fmaf(-wq, s_c, dc);
                            dc = dc + bq;
                            const float ra = static_cast<float>(ssq[r]);
                            const float rh = sqk < ra ? sqk : ra;
                            const float t = pq * rh;
                            fsc[r * 64 + lane] = s_c;
                            ftr[r * 64 + lane] = t;
                            pool = fmaf(t, t, pool);
                            qa = fmaf(fabsf(wq), rqf[r * 2 + 1], qa);
                            const float wp = static_cast<float>(0.5 * static_cast<double>(sc[3]) *
                                                                (rq[r * 3 + 2] + rq[r * 3 + 1]));
                            wpad = wpad + wp;
                        }
                        const float qw = qa + wpad;
                        const float slack = fmaf(z, sqrtf(pool), qw);
                        fdc[lane] = dc;
                        fpool[lane] = pool;
                        fqa[lane] = qw;
                        kept = !(dc - slack >= distk);
                    } else if (mine && l < cntb) {
                        kept = true;
                    }
                    n_cand += cntb;
                    n_btot += static_cast<uint64_t>(cntb) * bits_row;
                    const uint32_t keep = static_cast<uint32_t>(__ballot(kept) >> (32 * hh));
                    __syncthreads();
                    // The lane test of a kept lane against the current k-th distance:
                    // fp32 filter values for the lanes the filter covered, the fp64
                    // twin for the rest (their slack depends on the threshold).
                    auto lane_test = [&](double distk_d, double& pool2, double& qadd) -> bool {
                        double dcq;
                        if (l < covered) {
                            dcq = fdc[lane];
                            pool2 = fpool[lane];
                            qadd = fqa[lane];
                        } else {   // fp64 twin of the filter for the lanes it does not cover
                            dcq = d2;
                            pool2 = 0.0;
                            qadd = 0.0;
                            const double sqdk = distk_d < dinf ? sqrt(fmax(distk_d, 0.0)) : dinf;
                            const uint32_t tl = l - covered;
                            for (uint32_t r = 0; r < R; ++r) {
                                const uint8_t* bp = banks + r * kM8BankBytes;
                                const float* scl = reinterpret_cast<const float*>(bp + 128);
                                const double scr = fma(static_cast<double>(lut_delta),
                                                       static_cast<double>(sacc[r * 64 + lane]), rq[r * 3]);
                                const double bq = fma(static_cast<double>(bp[l]), static_cast<double>(scl[1]),
                                                      static_cast<double>(scl[0]));
                                const double wq = fma(static_cast<double>(bp[32 + l]), static_cast<double>(scl[3]),
                                                      static_cast<double>(scl[2]));
                                const double pq = fma(static_cast<double>(bp[64 + l]), static_cast<double>(scl[5]),
                                                      static_cast<double>(scl[4]));
                                dcq += fma(-wq, scr, bq);
                                const double sa = ssq[r];
                                const double rh = sqdk < sa ? sqdk : sa;
                                const double tr = pq * rh;
                                tsc[r * 8 + tl] = scr;
                                ttr[r * 8 + tl] = tr;
                                pool2 = fma(tr, tr, pool2);
                                qadd = fma(fabs(wq), rq[r * 3 + 1], qadd);
                                qadd = fma(0.5 * static_cast<double>(scl[3]), rq[r * 3 + 2] + rq[r * 3 + 1], qadd);
                            }
                        }
                        return !(dcq - fma(dz, sqrt(pool2), qadd) >= distk_d);
                    };
                    // Lanes passing now: the threshold only falls while this batch is
                    // refined and the test only gets stricter as it falls, so every
                    // lane staged below is among them.  Their exact run scores are
                    // computed here at once (they do not depend on the threshold):
                    // per (lane, run) pair 16 lanes, the even-block and the odd-block
                    // FMA chains on two 8-lane groups, added and summed in the AVX order.
                    uint32_t p0;
                    {
                        bool pass0 = false;
                        if (mine && ((keep >> l) & 1u)) {
                            double pl, qd;
                            pass0 = lane_test(cnt < K ? dinf : static_cast<double>(td[K - 1]), pl, qd);
                        }
                        p0 = static_cast<uint32_t>(__ballot(pass0) >> (32 * hh));
                        if (pass0) {
                            plist[__popc(p0 & ((1u << l) - 1u))] = static_cast<uint8_t>(l);
                            // the lane's refine row header and id, read by its cascade and
                            // push if it is staged (every staged lane passes now)
                            const float4* h =
                                reinterpret_cast<const float4*>(v.refine + (lb + bb * 32 + l) * v.rstride);
                            for (uint32_t r = 0; r < R; ++r) reinterpret_cast<float4*>(hdrl + l * R * 4)[r] = h[r];
                            idl[l] = v.ids[lb + bb * 32 + l];
                        }
                        __syncthreads();
                        const uint32_t np0 = __popc(p0);
                        const uint32_t npairs = R * np0;
                        const uint32_t half = (lane >> 3) & 1u, ci = lane & 7u;
                        const uint32_t g = static_cast<uint32_t>(lane) >> 4, wl = lane & 15u;
                        typedef const __attribute__((address_space(3))) uint8_t* LdsBytes;
                        typedef const __attribute__((address_space(3))) float* LdsFloats;
                        const LdsRun rt = rtl;
                        uint8_t* slot = runbuf + g * a.run_slot;
                        // Per pass four (lane, run) pairs, 16 lanes each: the pair's plane
                        // bytes and fields go to its LDS slot (8-byte words; runs start
                        // 8-byte aligned in both rows), then its two FMA chains read them
                        // there.  Slots of at most 16 x kM8PassWords words: each lane holds
                        // its words of the next pass in registers, loaded before this
                        // pass's chains run, so their latency hides behind them.
                        const bool pf = a.run_slot <= 16 * 8 * kM8PassWords;
                        auto pair_of = [&](uint32_t t0, uint32_t& r, uint32_t& lj) -> bool {
                            const uint32_t pair = t0 + g;
                            if (pair >= npairs) return false;
                            r = pair / np0;
                            lj = plist[pair - r * np0];
                            return true;
                        };
                        auto words_of = [&](uint32_t r, uint32_t lj, const uint64_t*& sp, const uint64_t*& sf) {
                            sp = reinterpret_cast<const uint64_t*>(rec + lj * PC + rt[r].blk0);
                            sf = reinterpret_cast<const uint64_t*>(v.refine + (lb + bb * 32 + lj) * v.rstride +
                                                                   rt[r].ref_off);
                        };
                        uint64_t pw[kM8PassWords];
                        auto fetch = [&](uint32_t t0) {
                            uint32_t r = 0, lj = 0, nwp = 0, nw = 0;
                            const uint64_t *sp = nullptr, *sf = nullptr;
                            if (pair_of(t0, r, lj)) {
                                nwp = rt[r].nwp;
                                nw = rt[r].nw;
                                words_of(r, lj, sp, sf);
                            }
#pragma unroll
                            for (uint32_t i2 = 0; i2 < kM8PassWords; ++i2) {
                                const uint32_t w = wl + 16 * i2;
                                pw[i2] = w < nw ? (w < nwp ? sp[w] : sf[w - nwp]) : 0ull;
                            }
                        };
                        if (pf) fetch(0);
                        for (uint32_t t0 = 0; t0 < npairs; t0 += 4) {
                            float acc = 0.0F;
                            uint32_t lj = 0, r = 0;
                            const bool ok = pair_of(t0, r, lj);
                            if (ok) {
                                uint64_t* sw = reinterpret_cast<uint64_t*>(slot);
                                const uint32_t nw = rt[r].nw;
                                if (pf) {
#pragma unroll
                                    for (uint32_t i2 = 0; i2 < kM8PassWords; ++i2)
                                        if (wl + 16 * i2 < nw) sw[wl + 16 * i2] = pw[i2];
                                } else {
                                    const uint32_t nwp = rt[r].nwp;
                                    const uint64_t *sp, *sf;
                                    words_of(r, lj, sp, sf);
                                    for (uint32_t w = wl; w < nw; w += 16) sw[w] = w < nwp ? sp[w] : sf[w - nwp];
                                }
                            }
                            __syncthreads();
                            if (pf && t0 + 4 < npairs) fetch(t0 + 4);
                            if (ok) {
                                const LdsBytes pb = (LdsBytes)(slot);
                                const uint32_t nblk = rt[r].nblk;
                                acc = M8HalfChain(rt[r].rate, pb, pb + nblk, (LdsFloats)(slev) + rt[r].lo_off, rt[r].lev,
                                                  (LdsFloats)(q) + rt[r].dim_off + ci, nblk, half, ci);
                            }
                            const float other = __shfl_xor(acc, 8);
                            const float sc8 = Hsum8Lanes(acc + other);
                            if (ci == 0 && half == 0 && ok) sS32[lj * R + r] = sc8;
                            __syncthreads();   // the slots are refilled by the next pass
                        }
                    }
                    uint32_t nst = 0, pos = 0;
                    for (;;) {
                        const uint32_t rem = pos >= 32 ? 0u : keep & p0 & ~((1u << pos) - 1u);
                        if (rem == 0) break;
                        const double distk_d = cnt < K ? dinf : static_cast<double>(td[K - 1]);
                        bool pass = false;
                        double pool2 = 0.0, qadd = 0.0;
                        if (mine && ((rem >> l) & 1u)) pass = lane_test(distk_d, pool2, qadd);
                        const uint32_t pm = static_cast<uint32_t>(__ballot(pass) >> (32 * hh));
                        const uint32_t need = 4 - nst;
                        uint32_t sel = 0, m = pm;
                        for (uint32_t t = 0; t < need && m; ++t) {
                            const uint32_t low = m & (0u - m);
                            sel |= low;
                            m ^= low;
                        }
                        if (mine && ((sel >> l) & 1u)) {
                            const uint32_t slot = nst + __popc(sel & ((1u << l) - 1u));
                            M8Stage& s = stg[slot];
                            s.lane = l;
                            s.row = static_cast<uint32_t>(bb * 32 + l);
                            s.pool = pool2;
                            s.qadd = qadd;
                            for (uint32_t r = 0; r < R; ++r) {
                                ssc[slot * R + r] = l < covered ? static_cast<double>(fsc[r * 64 + lane]) : tsc[r * 8 + (l - covered)];
                                str[slot * R + r] = l < covered ? static_cast<double>(ftr[r * 64 + lane]) : ttr[r * 8 + (l - covered)];
                            }
                        }
                        nst += __popc(sel);
                        __syncthreads();
                        const bool full = nst == 4;
                        if (full || (nst > 0 && pm == sel)) {
                            // refine the staged candidates (flush4)
                            const uint32_t n = nst;
                            n_staged += n;
                            ++n_flush;
                            // The cascade's run order: mean t^2 / nblk of each run over the
                            // group (lane r sums run r's in candidate order), then an
                            // insertion sort, descending, stable (GCC's std::sort below 17
                            // elements).  With no NaN mean that order is each run's rank:
                            // the runs with a larger mean, or an equal one and a lower index.
                            if (R > 1) {
                                double m = 0.0;
                                const bool mr = static_cast<uint32_t>(lane) < R;
                                if (mr) {
                                    for (uint32_t j = 0; j < n; ++j) m = fma(str[j * R + lane], str[j * R + lane], m);
                                    const uint32_t nbk = rtl[lane].nblk;
                                    m /= static_cast<double>(nbk > 1 ? nbk : 1);
                                    smean[lane] = m;
                                }
                                const bool nan_any = __ballot(mr && m != m) != 0;
                                __syncthreads();
                                if (!nan_any) {
                                    if (mr) {
                                        uint32_t rank = 0;
                                        for (uint32_t r2 = 0; r2 < R; ++r2) {
                                            const double m2 = smean[r2];
                                            rank += (m2 > m || (m2 == m && r2 < static_cast<uint32_t>(lane))) ? 1u : 0u;
                                        }
                                        sord[rank] = lane;
                                    }
                                } else if (lane == 0) {
                                    for (uint32_t r = 0; r < R; ++r) sord[r] = r;
                                    for (uint32_t i = 1; i < R; ++i) {
                                        const uint32_t x = sord[i];
                                        uint32_t j = i;
                                        while (j > 0 && smean[x] > smean[sord[j - 1]]) {
                                            sord[j] = sord[j - 1];
                                            --j;
                                        }
                                        sord[j] = x;
                                    }
                                }
                            } else if (lane == 0) {
                                sord[0] = 0;
                            }
                            __syncthreads();
                            const double distk_c = cnt < K ? dinf : static_cast<double>(td[K - 1]);
                            if (static_cast<uint32_t>(lane) < n) {   // the cascade, one lane per candidate
                                const uint32_t j = lane;
                                M8Stage& s = stg[j];
                                const float* f4 = hdrl + s.lane * R * 4;
                                double dist;
                                if (R == 1) {
                                    dist = d2 + static_cast<double>(f4[1]);
                                } else {
                                    double dc = d2;
                                    for (uint32_t r = 0; r < R; ++r) {
                                        const double coef = f4[r * 4], h2 = f4[r * 4 + 1];
                                        const double gn = f4[r * 4 + 2], vp = f4[r * 4 + 3];
                                        dc += fma(-(2.0 * coef), fma(gn, ssc[j * R + r], vp), h2);
                                    }
                                    dist = dc;
                                }
                                double rem2 = s.pool;
                                bool live = true;
                                uint64_t bs = 0;
                                for (uint32_t oi = 0; oi < R && live; ++oi) {
                                    const uint32_t r = sord[oi];
                                    const double se = sS32[s.lane * R + r];
                                    const double coef = f4[r * 4];
                                    if (R == 1) {
                                        dist = fma(-(2.0 * coef), se, dist);
                                    } else {
                                        const double gn = f4[r * 4 + 2], vp = f4[r * 4 + 3];
                                        dist = fma(2.0 * coef, fma(gn, ssc[j * R + r], vp) - se, dist);
                                    }
                                    const double t = str[j * R + r];
                                    rem2 = fma(-t, t, rem2);
                                    bs += rtl[r].bits;
                                    if (fma(-dz, sqrt(fmax(rem2, 0.0)), dist) - s.qadd >= distk_c) live = false;
                                }
                                s.dist = dist;
                                s.died = live ? 0u : 1u;
                                s.bits = bs;
                            }
                            __syncthreads();
                            for (uint32_t j = 0; j < n; ++j) {
                                const M8Stage s = stg[j];
                                n_bits += s.bits;
                                if (s.died) continue;
                                ++n_surv;
                                M8Push(td, ti, cnt, K, static_cast<float>(s.dist), idl[s.lane]);
                            }
                            __syncthreads();
                            pos = full ? 32 - __clz(sel) : 32;
                            nst = 0;
                        } else {
                            break;
                        }
                    }
                    __syncthreads();
                }
            }
            if (a.budget) {
                used += nb;
                if (used >= a.budget) {
                    p_stop = nb == nbl ? p + 1 : p;
                    b_stop = nb == nbl ? 0 : nb;
                    break;
                }
            }
        }
        if (lane == 0 && a.cover) a.cover[qi] = (p_stop << 32) | b_stop;
        for (uint32_t i = lane; i < K; i += 64) {
            a.out_ids[qi * K + i] = i < cnt ? ti[i] : 0xFFFFFFFFu;
            a.out_d[qi * K + i] = i < cnt ? td[i] : finf;
        }
        if (lane == 0 && a.qstats) {
            uint64_t* o = a.qstats + qi * 8;
            o[0] = n_cand;
            o[1] = n_surv;
            o[2] = n_bits;
            o[3] = n_btot;
            o[4] = n_cand - n_surv;
            o[5] = n_batch;
            o[6] = n_staged;
            o[7] = n_flush;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Launchers


hipError_t LaunchM8QueryPrep(hipStream_t st, const M8View& v, const float* queries, uint64_t nq, float* rotq,
                             uint8_t* lut, float* qc, double* rc) {
    if (nq == 0) return hipSuccess;
    const size_t lds = (v.padded + static_cast<size_t>(v.plane_cols) * 32) * sizeof(float);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&M8QueryPrepKernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    const uint32_t grid = static_cast<uint32_t>(nq < 65536 ? nq : 65536);
    hipLaunchKernelGGL(M8QueryPrepKernel, dim3(grid), dim3(kM8PrepThreads), lds, st, v, queries, nq, rotq, lut, qc,
                       rc);
    return hipGetLastError();
}

hipError_t LaunchM8RunDist(hipStream_t st, const M8View& v, const float* rotq, const double* rc,
                           const uint32_t* probe, const uint32_t* nprobe_q, uint32_t nprobe_all, uint32_t maxp,
                           uint64_t nq, double* sqd2) {
    const uint64_t tot = nq * maxp * v.nrun;
    if (tot == 0) return hipSuccess;
    hipLaunchKernelGGL(M8RunDistKernel, dim3(static_cast<uint32_t>((tot + 255) / 256)), dim3(256), 0, st, v, rotq, rc,
                       probe, nprobe_q, nprobe_all, maxp, nq, sqd2);
    return hipGetLastError();
}

size_t M8ScanLds(const M8View& v, uint32_t k, uint32_t lev_floats, uint32_t run_slot) {
    return M8ScanLayout(static_cast<uint32_t>(v.padded), v.plane_cols, v.nrun, k, v.rstride, lev_floats, run_slot)
        .total;
}

hipError_t LaunchM8Scan(hipStream_t st, const M8ScanArgs& a) {
    const uint64_t nrun = a.qsel ? a.nsel : a.nq;
    if (nrun == 0) return hipSuccess;
    const size_t lds = M8ScanLds(a.v, a.k, a.lev_floats, a.run_slot);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&M8ScanKernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    // with the count on the device, most of nsel are usually not searched:
    // a smaller grid strides over them
    const uint64_t gmax = a.nsel_dev ? 4096 : 1048576;
    const uint32_t grid = static_cast<uint32_t>(nrun < gmax ? nrun : gmax);
    hipLaunchKernelGGL(M8ScanKernel, dim3(grid), dim3(64), lds, st, a);
    return hipGetLastError();
}

}  // namespace e8qdev
