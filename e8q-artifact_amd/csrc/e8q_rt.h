// e8q_rt.h -- host-side plumbing shared by the C ABI translation units of
// libe8q_gpu.so (E8Q: e8q_runtime.hip, ME8Q: me8q_runtime.hip): the
// thread-local error channel, owned device buffers, the device context and
// the encoder-table preparation.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/e8q_gpu.h"

namespace e8qrt {

inline thread_local std::string g_err;

struct Fail {
    e8q_status st;
    std::string msg;
};

[[noreturn]] inline void Throw(e8q_status st, const std::string& m) { throw Fail{st, m}; }

inline void Check(hipError_t e, const char* what) {
    if (e != hipSuccess) Throw(e == hipErrorOutOfMemory ? E8Q_ENOMEM : E8Q_ERUNTIME,
                               std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
e8q_status Guard(F&& f) {
    // The launchers report launch failures through hipGetLastError, which
    // also returns the last failure of any earlier HIP call on this thread,
    // including calls made by other libraries; drop that before the call.
    (void)hipGetLastError();
    try {
        f();
        return E8Q_OK;
    } catch (const Fail& x) {
        g_err = x.msg;
        return x.st;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return E8Q_ENOMEM;
    } catch (const std::exception& x) {
        g_err = x.what();
        return E8Q_ERUNTIME;
    }
}

// Owned device allocation.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            reset();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b) {
        reset();
        if (b == 0) b = 1;
        Check(hipMalloc(&p, b), "hipMalloc");
        bytes = b;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

}  // namespace e8qrt

struct e8q_ctx {
    using DBuf = e8qrt::DBuf;
    int device = 0;
    int cus = 0;      // compute units (the persistent batched scan's grid)
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};   // search timing
    static constexpr int kKernelEvents = 16;
    hipEvent_t kev[kKernelEvents] = {};   // per-launch timing of one kernel's launches in a search
    DBuf ws;   // per-call workspace, grown on demand (never by a call that fits it)
    uint64_t reserved = 0;       // bytes of the largest e8q_ctx_reserve (0: none)
    uint64_t m8_pool_per_q = 0;  // batched ME8Q pool entries per query a past call needed (unreserved contexts)

    // Replay of a search's device sequence as a HIP graph, per call shape.
    // A shape's first call runs eagerly and registers the shape; its second
    // call captures the sequence, later ones replay it.
    struct GraphEntry {
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
        uint64_t last_use = 0;
    };
    static constexpr size_t kMaxGraphs = 16;
    bool graphs = false;
    std::vector<GraphEntry> gcache;
    uint64_t gclock = 0, captures = 0, replays = 0;
    GraphEntry* FindGraph(const std::vector<uint64_t>& key) {
        for (auto& g : gcache)
            if (g.key == key) {
                g.last_use = ++gclock;
                return &g;
            }
        return nullptr;
    }
    GraphEntry* AddGraph(const std::vector<uint64_t>& key) {
        if (gcache.size() >= kMaxGraphs) {   // evict the least recently used
            auto lru = std::min_element(gcache.begin(), gcache.end(),
                                        [](const GraphEntry& a, const GraphEntry& b) { return a.last_use < b.last_use; });
            if (lru->exec) (void)hipGraphExecDestroy(lru->exec);
            gcache.erase(lru);
        }
        gcache.push_back(GraphEntry{key, nullptr, ++gclock});
        return &gcache.back();
    }
    void ClearGraphs() {
        for (auto& g : gcache)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        gcache.clear();
    }

    // Shape of the batched scan: fixed by the caller (nonzero fix_*), else
    // chosen by timing candidates on the first batched call per (index
    // shape, power-of-two query-count bucket).
    uint32_t fix_qtile = 0, fix_btile = 0, fix_grid = 0;
    struct Tuned {
        uint64_t shape[4];   // batch bytes, blocks, nlist, query bucket
        uint32_t qtile, btile, grid;
    };
    std::vector<Tuned> tuned;
    uint32_t last_qtile = 0, last_btile = 0, last_grid = 0;
    uint64_t tunings = 0;

    // Grids of the batched ME8Q form's persistent kernels (the MFMA
    // contraction, the exact refine): fixed by the caller (nonzero), else
    // chosen by timing candidates on the first batched search per (index
    // shape, query bucket, k and nprobe buckets).  The table of tuned grids
    // is the context's own unless it shares another's (e8q_ctx_share_tuning:
    // a server's workers tune each bucket once between them).
    uint32_t m8_fix_mfma = 0, m8_fix_refine = 0;
    struct M8Tuned {
        uint64_t shape[6];
        uint32_t mfma_grid, refine_grid;
    };
    struct M8TuneTable {
        std::mutex mu;
        std::vector<M8Tuned> v;
    };
    std::shared_ptr<M8TuneTable> m8_table = std::make_shared<M8TuneTable>();
    uint32_t m8_last_mfma = 0, m8_last_refine = 0, m8_last_budget = 0, m8_last_set_a = 0;
    uint64_t m8_tunings = 0;

    // Bump allocator over ws for one call.
    struct Carve {
        std::vector<std::pair<void**, size_t>> want;
        template <class T>
        void add(T** p, size_t count) { want.push_back({reinterpret_cast<void**>(p), count * sizeof(T)}); }
        size_t bytes() const {
            size_t total = 0;
            for (auto& w : want) total += (w.second + 255) & ~size_t{255};
            return total;
        }
    };
    void Bind() { e8qrt::Check(hipSetDevice(device), "hipSetDevice"); }
    void Take(Carve& c) {
        const size_t total = c.bytes();
        if (ws.bytes < total) ws.alloc(total + total / 4);
        char* base = ws.as<char>();
        size_t off = 0;
        for (auto& w : c.want) {
            *w.first = base + off;
            off += (w.second + 255) & ~size_t{255};
        }
    }
};

namespace e8qrt {

template <class T>
void Upload(DBuf& b, const T* src, size_t count, hipStream_t st) {
    b.alloc(count * sizeof(T));
    if (count) Check(hipMemcpyAsync(b.p, src, count * sizeof(T), hipMemcpyHostToDevice, st), "upload");
}

// Encoder tables from the per-branch level tables (e8q_codec.cpp:125-352):
// the merged sorted table with its period-4 class interleave verified, and
// the four (branch, parity) single-class tables sorted by (level, digit).
struct EncTabHost {
    std::vector<double> mlev, cl;
    std::vector<uint16_t> mdig, cd;
    bool merged = false;
};

inline EncTabHost MakeEncoderTables(const double* levels, int bits) {
    EncTabHost t;

    const uint32_t q = 1u << bits;
    struct E {
        double v;
        uint16_t d;
        int br;
    };
    std::vector<E> m;
    m.reserve(2 * q);
    for (uint32_t d = 0; d < q; ++d) {
        m.push_back({levels[d], static_cast<uint16_t>(d), 0});
        m.push_back({levels[q + d], static_cast<uint16_t>(d), 1});
    }
    std::stable_sort(m.begin(), m.end(), [](const E& a, const E& b) { return a.v < b.v; });
    bool alt = true;
    for (size_t i = 0; i < m.size(); ++i) {
        if (i > 0 && !(m[i].v > m[i - 1].v)) alt = false;
        if (m[i].br != static_cast<int>(i & 1) || (m[i].d & 1u) != ((i >> 1) & 1u)) alt = false;
    }
    std::vector<double>& mlev = t.mlev;
    std::vector<uint16_t>& mdig = t.mdig;
    mlev.resize(m.size());
    mdig.resize(m.size());
    for (size_t i = 0; i < m.size(); ++i) {
        mlev[i] = m[i].v;
        mdig[i] = m[i].d;
    }
    const uint32_t half = q / 2;
    std::vector<double>& cl = t.cl;
    std::vector<uint16_t>& cd = t.cd;
    cl.resize(4 * half);
    cd.resize(4 * half);
    for (int br = 0; br < 2; ++br)
        for (int par = 0; par < 2; ++par) {
            std::vector<std::pair<double, uint16_t>> e;
            for (uint32_t d = par; d < q; d += 2) e.push_back({levels[br * q + d], static_cast<uint16_t>(d)});
            std::sort(e.begin(), e.end());
            for (uint32_t i = 0; i < half; ++i) {
                cl[(br * 2 + par) * half + i] = e[i].first;
                cd[(br * 2 + par) * half + i] = e[i].second;
            }
        }
    t.merged = alt;
    return t;
}

}  // namespace e8qrt
