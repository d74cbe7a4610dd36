// serve.h -- a continuous-batching search server over a device-resident index.
//
// The reference has no serving layer (its binding runs one blocking search
// call per Python request).  Here requests of different sizes, k and nprobe
// arrive at any time from any thread; W worker threads, each with its own
// device context (stream + scratch workspace reserved once), take the oldest
// waiting request and every later waiting request with the same k, up to the
// batch limit and up to what the worker's workspace holds (admission by
// memory), and run them as one device search with per-query nprobe.  While one
// worker's batch runs, the others' batches copy and compute on their own
// streams.  Each query is searched by its own wavefront, so a query's result
// does not depend on the other queries of its batch: a request gets the same
// ids and distances alone and in any mix.
#pragma once

#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <exception>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace e8qlib::serve {

// One device search over a batch: nq queries (host, dim floats each), their
// nprobe values, outputs nq*k.
struct BatchView {
    const float* queries = nullptr;
    size_t nq = 0, k = 0;
    const uint32_t* nprobe = nullptr;
    uint32_t max_nprobe = 0;
    uint32_t* ids = nullptr;
    float* dists = nullptr;
};

// What one device search reports back to the server.
struct BatchReport {
    uint64_t overflowed = 0;   // queries the batched ME8Q form handed to the plain form
    double device_ms = 0.0;    // device time of the whole search (HIP events)
};

// What a worker runs on: its own device context over a shared index.
class Backend {
   public:
    virtual ~Backend() = default;
    virtual size_t dim() const = 0;
    // Device scratch of one search of this shape.
    virtual uint64_t WorkspaceBytes(size_t nq, size_t k, size_t max_nprobe) const = 0;
    // Grows the worker's workspace to `bytes` once, before serving.
    virtual void Reserve(uint64_t bytes) = 0;
    // From now on tuned launch shapes are shared with `first` (a worker of
    // the same server).
    virtual void ShareTuning(Backend& first) { (void)first; }
    virtual BatchReport Search(const BatchView& b) = 0;   // blocking
};

struct Options {
    int workers = 2;
    size_t max_batch_queries = 4096;   // queries per device search
    size_t max_k = 128;
    size_t max_nprobe = 256;
    // Device workspace per worker; 0 = what a full batch at max_k and
    // max_nprobe needs.  A request that cannot fit it alone is rejected.
    uint64_t workspace_bytes = 0;
    // Queries waiting to be admitted; Submit blocks while the queue is full.
    size_t max_queued_queries = size_t{1} << 20;
    // Device form of an ME8Q index's batches: E8Q_MODE_PLAIN (0) or
    // E8Q_MODE_BATCHED (1).  An E8Q index always runs the plain form.
    int mode = 0;
};

struct Completed {
    std::vector<uint32_t> ids;
    std::vector<float> dists;
    size_t nq = 0, k = 0;
    double latency_us = 0.0;   // submit to results
    double queue_us = 0.0;     // submit to batch start
    size_t batch_queries = 0;  // size of the batch it ran in
    size_t batch_requests = 0;
};

struct Stats {
    uint64_t requests = 0, queries = 0, batches = 0, rejected = 0;
    uint64_t max_batch_queries_seen = 0;
    uint64_t max_workspace_used = 0;   // largest WorkspaceBytes of a formed batch
    uint64_t workspace_bytes = 0;      // per worker
    uint64_t overflowed = 0;           // queries answered by the plain form after a pool overflow
    double device_ms = 0.0;            // device time of every batch, summed
    double batch_ms = 0.0;             // wall time of every batch's search call, summed
};

class Server {
   public:
    Server(std::vector<std::unique_ptr<Backend>> workers, const Options& opt);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    // Queues a request (the queries are copied).  Throws std::invalid_argument
    // for a request that can never be admitted; blocks while the queue is full.
    uint64_t Submit(const float* queries, size_t nq, size_t k, size_t nprobe);
    // Blocks until the request is done; rethrows a device failure.
    Completed Wait(uint64_t ticket);
    bool Done(uint64_t ticket);
    Stats stats();
    // Stops accepting work, finishes what is queued, joins the workers.
    void Close();
    uint64_t workspace_bytes() const { return ws_bytes_; }

   private:
    struct Request {
        uint64_t ticket = 0;
        std::vector<float> q;
        size_t nq = 0, k = 0, nprobe = 0;
        double t_submit = 0.0;
    };
    struct Slot {
        bool done = false;
        Completed out;
        std::exception_ptr err;
    };
    void WorkerLoop(int w);
    bool Fits(size_t nq, size_t k, size_t maxp) const;

    std::vector<std::unique_ptr<Backend>> workers_;
    Options opt_;
    uint64_t ws_bytes_ = 0;
    size_t dim_ = 0;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_, cv_space_;
    std::deque<Request> queue_;
    size_t queued_queries_ = 0;
    std::unordered_map<uint64_t, Slot> slots_;
    uint64_t next_ticket_ = 1;
    bool closing_ = false;
    Stats stats_;
    std::vector<std::thread> threads_;
};

double NowUs();

}  // namespace e8qlib::serve

namespace e8qlib {
class E8qIndex;
namespace me8q {
class Me8qIndex;
}
namespace serve {
// Servers whose workers each own a device context over the given index (the
// index must not be rebuilt or reloaded while the server runs).
std::unique_ptr<Server> MakeE8qServer(const E8qIndex& ix, const Options& opt);
std::unique_ptr<Server> MakeMe8qServer(const me8q::Me8qIndex& ix, const Options& opt);
// Host-only workers for testing the scheduling and the admission rules.
std::unique_ptr<Server> MakeFakeServer(size_t dim, uint64_t bytes_per_query, double delay_us, const Options& opt);
}  // namespace serve
}  // namespace e8qlib
