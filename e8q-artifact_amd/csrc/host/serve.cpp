// serve.cpp -- continuous-batching server (serve.h) and its device backends.
#include "serve.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <thread>
#include <type_traits>

#include "e8q_gpu.h"
#include "e8q_index.h"
#include "me8q_index.h"

namespace e8qlib::serve {

double NowUs() {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

Server::Server(std::vector<std::unique_ptr<Backend>> workers, const Options& opt)
    : workers_(std::move(workers)), opt_(opt) {
    if (workers_.empty()) throw std::invalid_argument("Server: at least one worker");
    if (opt_.max_batch_queries == 0) throw std::invalid_argument("Server: max_batch_queries must be positive");
    dim_ = workers_[0]->dim();
    ws_bytes_ = opt_.workspace_bytes
                    ? opt_.workspace_bytes
                    : workers_[0]->WorkspaceBytes(opt_.max_batch_queries, opt_.max_k, opt_.max_nprobe);
    for (auto& w : workers_) w->Reserve(ws_bytes_);
    for (size_t i = 1; i < workers_.size(); ++i) workers_[i]->ShareTuning(*workers_[0]);
    stats_.workspace_bytes = ws_bytes_;
    for (size_t i = 0; i < workers_.size(); ++i) threads_.emplace_back([this, i] { WorkerLoop(static_cast<int>(i)); });
}

Server::~Server() { Close(); }

bool Server::Fits(size_t nq, size_t k, size_t maxp) const {
    return nq <= opt_.max_batch_queries && workers_[0]->WorkspaceBytes(nq, k, maxp) <= ws_bytes_;
}

uint64_t Server::Submit(const float* queries, size_t nq, size_t k, size_t nprobe) {
    if (nq == 0 || k == 0 || nprobe == 0) throw std::invalid_argument("Server::Submit: nq, k and nprobe must be positive");
    if (!queries) throw std::invalid_argument("Server::Submit: null queries");
    if (!Fits(nq, k, nprobe)) {
        std::lock_guard<std::mutex> lk(mu_);
        ++stats_.rejected;
        throw std::invalid_argument("Server::Submit: request exceeds the serving memory budget (" +
                                    std::to_string(workers_[0]->WorkspaceBytes(nq, k, nprobe)) + " > " +
                                    std::to_string(ws_bytes_) + " bytes) or the batch limit");
    }
    Request r;
    r.q.assign(queries, queries + nq * dim_);
    r.nq = nq;
    r.k = k;
    r.nprobe = nprobe;
    std::unique_lock<std::mutex> lk(mu_);
    cv_space_.wait(lk, [&] { return closing_ || queued_queries_ + nq <= opt_.max_queued_queries || queue_.empty(); });
    if (closing_) throw std::runtime_error("Server::Submit: server closed");
    r.ticket = next_ticket_++;
    r.t_submit = NowUs();
    const uint64_t t = r.ticket;
    slots_[t];
    queued_queries_ += nq;
    ++stats_.requests;
    stats_.queries += nq;
    queue_.push_back(std::move(r));
    lk.unlock();
    cv_work_.notify_one();
    return t;
}

void Server::WorkerLoop(int w) {
    Backend& be = *workers_[static_cast<size_t>(w)];
    std::vector<float> q;
    std::vector<uint32_t> npq, ids;
    std::vector<float> dists;
    for (;;) {
        std::vector<Request> batch;
        size_t nq = 0, k = 0, maxp = 0;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_work_.wait(lk, [&] { return closing_ || !queue_.empty(); });
            if (queue_.empty()) return;   // closing and drained
            // The oldest request heads the batch; later ones with its k join in
            // arrival order while the batch still fits the limit and the
            // workspace.  Requests with another k wait for the next batch (the
            // oldest waiting request always heads the next one).
            k = queue_.front().k;
            for (auto it = queue_.begin(); it != queue_.end();) {
                if (it->k != k) {
                    ++it;
                    continue;
                }
                const size_t mp = std::max(maxp, it->nprobe);
                if (!batch.empty() && !Fits(nq + it->nq, k, mp)) break;
                nq += it->nq;
                maxp = mp;
                queued_queries_ -= it->nq;
                batch.push_back(std::move(*it));
                it = queue_.erase(it);
            }
            ++stats_.batches;
            stats_.max_batch_queries_seen = std::max<uint64_t>(stats_.max_batch_queries_seen, nq);
            stats_.max_workspace_used = std::max<uint64_t>(stats_.max_workspace_used, be.WorkspaceBytes(nq, k, maxp));
        }
        cv_space_.notify_all();
        const double t_start = NowUs();
        q.resize(nq * dim_);
        npq.resize(nq);
        ids.resize(nq * k);
        dists.resize(nq * k);
        size_t off = 0;
        for (const Request& r : batch) {
            std::memcpy(q.data() + off * dim_, r.q.data(), r.nq * dim_ * sizeof(float));
            std::fill(npq.begin() + static_cast<long>(off), npq.begin() + static_cast<long>(off + r.nq),
                      static_cast<uint32_t>(r.nprobe));
            off += r.nq;
        }
        std::exception_ptr err;
        BatchReport rep;
        try {
            BatchView bv;
            bv.queries = q.data();
            bv.nq = nq;
            bv.k = k;
            bv.nprobe = npq.data();
            bv.max_nprobe = static_cast<uint32_t>(maxp);
            bv.ids = ids.data();
            bv.dists = dists.data();
            rep = be.Search(bv);
        } catch (...) {
            err = std::current_exception();
        }
        const double t_end = NowUs();
        {
            std::lock_guard<std::mutex> lk(mu_);
            stats_.overflowed += rep.overflowed;
            stats_.device_ms += rep.device_ms;
            stats_.batch_ms += (t_end - t_start) / 1e3;
            off = 0;
            for (const Request& r : batch) {
                Slot& s = slots_[r.ticket];
                s.done = true;
                s.err = err;
                if (!err) {
                    s.out.ids.assign(ids.begin() + static_cast<long>(off * k),
                                     ids.begin() + static_cast<long>((off + r.nq) * k));
                    s.out.dists.assign(dists.begin() + static_cast<long>(off * k),
                                       dists.begin() + static_cast<long>((off + r.nq) * k));
                }
                s.out.nq = r.nq;
                s.out.k = k;
                s.out.latency_us = t_end - r.t_submit;
                s.out.queue_us = t_start - r.t_submit;
                s.out.batch_queries = nq;
                s.out.batch_requests = batch.size();
                off += r.nq;
            }
        }
        cv_done_.notify_all();
    }
}

Completed Server::Wait(uint64_t ticket) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = slots_.find(ticket);
    if (it == slots_.end()) throw std::invalid_argument("Server::Wait: unknown ticket");
    cv_done_.wait(lk, [&] { return slots_[ticket].done; });
    Slot s = std::move(slots_[ticket]);
    slots_.erase(ticket);
    if (s.err) std::rethrow_exception(s.err);
    return std::move(s.out);
}

bool Server::Done(uint64_t ticket) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = slots_.find(ticket);
    return it != slots_.end() && it->second.done;
}

Stats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    return stats_;
}

void Server::Close() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (closing_ && threads_.empty()) return;
        closing_ = true;
    }
    cv_work_.notify_all();
    cv_space_.notify_all();
    for (auto& t : threads_)
        if (t.joinable()) t.join();
    threads_.clear();
}

}  // namespace e8qlib::serve

// ---------------------------------------------------------------------------
// Backends

namespace e8qlib::serve {
namespace {

int DeviceOrdinal() {
    const char* e = std::getenv("E8Q_DEVICE");
    return e ? std::atoi(e) : 0;
}

void CheckSt(e8q_status st) {
    if (st == E8Q_OK) return;
    const std::string msg = e8q_last_error();
    if (st == E8Q_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error("e8q device: " + msg);
}

// A worker's own device context over a shared, read-only index.
template <class Index>
class DeviceBackend final : public Backend {
   public:
    DeviceBackend(const Index& ix, int mode) : ix_(ix), mode_(mode) { CheckSt(e8q_ctx_create(DeviceOrdinal(), &ctx_)); }
    ~DeviceBackend() override {
        if (ctx_) (void)e8q_ctx_destroy(ctx_);
    }
    size_t dim() const override { return ix_.dim(); }
    uint64_t WorkspaceBytes(size_t nq, size_t k, size_t max_nprobe) const override {
        if constexpr (std::is_same_v<Index, me8q::Me8qIndex>) return ix_.SearchWorkspaceBytes(nq, k, max_nprobe, mode_);
        else return ix_.SearchWorkspaceBytes(nq, k, max_nprobe);
    }
    void Reserve(uint64_t bytes) override { CheckSt(e8q_ctx_reserve(ctx_, bytes)); }
    void ShareTuning(Backend& first) override {
        CheckSt(e8q_ctx_share_tuning(ctx_, static_cast<DeviceBackend&>(first).ctx_));
    }
    BatchReport Search(const BatchView& b) override {
        BatchReport r;
        if constexpr (std::is_same_v<Index, me8q::Me8qIndex>) {
            e8q_m8_stats st{};
            ix_.SearchOnContext(ctx_, b.queries, b.nq, b.k, b.nprobe, b.ids, b.dists, mode_, &st);
            r.overflowed = st.overflowed;
            r.device_ms = st.total_ms;
        } else {
            e8q_search_stats st{};
            ix_.SearchOnContext(ctx_, b.queries, b.nq, b.k, b.nprobe, b.ids, b.dists, &st);
            r.device_ms = st.total_ms;
        }
        return r;
    }

   private:
    const Index& ix_;
    int mode_ = 0;
    e8q_ctx* ctx_ = nullptr;
};

}  // namespace

std::unique_ptr<Server> MakeE8qServer(const E8qIndex& ix, const Options& opt) {
    std::vector<std::unique_ptr<Backend>> w;
    for (int i = 0; i < opt.workers; ++i) w.push_back(std::make_unique<DeviceBackend<E8qIndex>>(ix, 0));
    return std::make_unique<Server>(std::move(w), opt);
}

std::unique_ptr<Server> MakeMe8qServer(const me8q::Me8qIndex& ix, const Options& opt) {
    std::vector<std::unique_ptr<Backend>> w;
    for (int i = 0; i < opt.workers; ++i) w.push_back(std::make_unique<DeviceBackend<me8q::Me8qIndex>>(ix, opt.mode));
    return std::make_unique<Server>(std::move(w), opt);
}

// Host-only backend for testing the scheduler: a search of q fills ids with
// bits of its first coordinate + j and dists with q[0] + j after sleeping
// `delay_us`; its workspace is bytes_per_query per query.
class FakeBackend final : public Backend {
   public:
    FakeBackend(size_t dim, uint64_t bytes_per_query, double delay_us)
        : dim_(dim), bpq_(bytes_per_query), delay_us_(delay_us) {}
    size_t dim() const override { return dim_; }
    uint64_t WorkspaceBytes(size_t nq, size_t k, size_t max_nprobe) const override {
        return nq * (bpq_ + 8 * k + 4 * max_nprobe);
    }
    void Reserve(uint64_t) override {}
    BatchReport Search(const BatchView& b) override {
        if (delay_us_ > 0) std::this_thread::sleep_for(std::chrono::microseconds(static_cast<int64_t>(delay_us_)));
        for (size_t i = 0; i < b.nq; ++i) {
            const float x = b.queries[i * dim_];
            uint32_t bits;
            std::memcpy(&bits, &x, 4);
            for (size_t j = 0; j < b.k; ++j) {
                b.ids[i * b.k + j] = bits + static_cast<uint32_t>(j) + b.nprobe[i];
                b.dists[i * b.k + j] = x + static_cast<float>(j);
            }
        }
        return BatchReport{};
    }

   private:
    size_t dim_;
    uint64_t bpq_;
    double delay_us_;
};

std::unique_ptr<Server> MakeFakeServer(size_t dim, uint64_t bytes_per_query, double delay_us, const Options& opt) {
    std::vector<std::unique_ptr<Backend>> w;
    for (int i = 0; i < opt.workers; ++i) w.push_back(std::make_unique<FakeBackend>(dim, bytes_per_query, delay_us));
    return std::make_unique<Server>(std::move(w), opt);
}

}  // namespace e8qlib::serve
