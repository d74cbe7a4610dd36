// serve_bindings.cpp -- `e8qlib.Server`: the continuous-batching serving loop
// (serve.h) over an E8qIndex or Me8qIndex of this module.
//
//   srv = e8qlib.Server(index, workers=2, max_batch=4096, max_k=128,
//                       max_nprobe=256, workspace_bytes=0, max_queued=1<<20)
//   t = srv.submit(queries, k, nprobe)        # returns at once (blocks only on a full queue)
//   ids, dists, info = srv.wait(t)            # info: latency_us, queue_us, batch_queries, ...
//   srv.stats(); srv.close()
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <memory>
#include <stdexcept>
#include <string>

#include "e8q_gpu.h"
#include "e8q_index.h"
#include "me8q_index.h"
#include "serve.h"

namespace py = pybind11;

namespace {

using e8qlib::serve::Options;
using e8qlib::serve::Server;

class PyServer {
   public:
    PyServer(py::object index, int workers, size_t max_batch, size_t max_k, size_t max_nprobe, uint64_t workspace_bytes,
             size_t max_queued, const std::string& mode)
        : index_(std::move(index)) {
        const std::string kind = py::str(index_.attr("__class__").attr("__name__"));
        const uintptr_t ptr = index_.attr("_native")().cast<uintptr_t>();
        if (ptr == 0) throw std::runtime_error("Server: the index must be built or loaded");
        Options opt = MakeOptions(workers, max_batch, max_k, max_nprobe, workspace_bytes, max_queued);
        if (mode == "batched") opt.mode = E8Q_MODE_BATCHED;
        else if (mode != "plain") throw std::invalid_argument("mode must be 'plain' or 'batched'");
        if (kind == "E8qIndex") {
            if (opt.mode != E8Q_MODE_PLAIN) throw std::invalid_argument("Server: an E8qIndex is served by the plain form");
            if (max_k > e8qlib::kMaxK) throw std::invalid_argument("max_k must be <= 128 for E8qIndex");
            py::gil_scoped_release nogil;
            srv_ = e8qlib::serve::MakeE8qServer(*reinterpret_cast<const e8qlib::E8qIndex*>(ptr), opt);
        } else if (kind == "Me8qIndex") {
            py::gil_scoped_release nogil;
            srv_ = e8qlib::serve::MakeMe8qServer(*reinterpret_cast<const e8qlib::me8q::Me8qIndex*>(ptr), opt);
        } else {
            throw std::invalid_argument("Server: index must be an e8qlib.E8qIndex or e8qlib.Me8qIndex");
        }
    }

    // Host-only workers (for tests of the scheduler and the admission rules).
    static PyServer Sim(size_t dim, uint64_t bytes_per_query, double delay_us, int workers, size_t max_batch,
                        size_t max_k, size_t max_nprobe, uint64_t workspace_bytes, size_t max_queued) {
        PyServer s;
        s.srv_ = e8qlib::serve::MakeFakeServer(
            dim, bytes_per_query, delay_us, MakeOptions(workers, max_batch, max_k, max_nprobe, workspace_bytes, max_queued));
        return s;
    }

    uint64_t submit(py::handle queries, size_t k, size_t nprobe) {
        auto q = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(queries);
        if (!q || q.ndim() != 2) throw std::invalid_argument("queries must be a 2D NumPy array");
        Server& s = live();
        const float* p = q.data();
        const size_t nq = static_cast<size_t>(q.shape(0));
        py::gil_scoped_release nogil;
        return s.Submit(p, nq, k, nprobe);
    }

    py::tuple wait(uint64_t ticket) {
        Server& s = live();
        e8qlib::serve::Completed c;
        {
            py::gil_scoped_release nogil;
            c = s.Wait(ticket);
        }
        const std::vector<ssize_t> shape{static_cast<ssize_t>(c.nq), static_cast<ssize_t>(c.k)};
        py::array_t<uint32_t> ids(shape);
        py::array_t<float> dists(shape);
        std::copy(c.ids.begin(), c.ids.end(), ids.mutable_data());
        std::copy(c.dists.begin(), c.dists.end(), dists.mutable_data());
        py::dict info;
        info["latency_us"] = c.latency_us;
        info["queue_us"] = c.queue_us;
        info["batch_queries"] = c.batch_queries;
        info["batch_requests"] = c.batch_requests;
        return py::make_tuple(ids, dists, info);
    }

    bool done(uint64_t ticket) { return live().Done(ticket); }

    py::dict stats() {
        const e8qlib::serve::Stats st = live().stats();
        py::dict d;
        d["requests"] = st.requests;
        d["queries"] = st.queries;
        d["batches"] = st.batches;
        d["rejected"] = st.rejected;
        d["max_batch_queries"] = st.max_batch_queries_seen;
        d["max_workspace_used"] = st.max_workspace_used;
        d["workspace_bytes"] = st.workspace_bytes;
        d["overflowed"] = st.overflowed;
        d["device_ms"] = st.device_ms;
        d["batch_ms"] = st.batch_ms;
        return d;
    }

    void close() {
        if (srv_) {
            py::gil_scoped_release nogil;
            srv_->Close();
        }
    }

   private:
    PyServer() = default;
    static Options MakeOptions(int workers, size_t max_batch, size_t max_k, size_t max_nprobe, uint64_t ws,
                               size_t max_queued) {
        if (workers < 1) throw std::invalid_argument("workers must be positive");
        Options o;
        o.workers = workers;
        o.max_batch_queries = max_batch;
        o.max_k = max_k;
        o.max_nprobe = max_nprobe;
        o.workspace_bytes = ws;
        o.max_queued_queries = max_queued;
        return o;
    }
    Server& live() {
        if (!srv_) throw std::runtime_error("Server is closed");
        return *srv_;
    }
    py::object index_;   // keeps the index alive while the server runs
    std::unique_ptr<Server> srv_;
};

}  // namespace

void RegisterServe(py::module_& m) {
    py::class_<PyServer>(m, "Server")
        .def(py::init<py::object, int, size_t, size_t, size_t, uint64_t, size_t, const std::string&>(),
             py::arg("index"), py::arg("workers") = 2, py::arg("max_batch") = 4096, py::arg("max_k") = 128,
             py::arg("max_nprobe") = 256, py::arg("workspace_bytes") = 0, py::arg("max_queued") = size_t{1} << 20,
             py::arg("mode") = "plain")
        .def_static("_sim", &PyServer::Sim, py::arg("dim"), py::arg("bytes_per_query"), py::arg("delay_us"),
                    py::arg("workers") = 2, py::arg("max_batch") = 4096, py::arg("max_k") = 128,
                    py::arg("max_nprobe") = 256, py::arg("workspace_bytes") = 0,
                    py::arg("max_queued") = size_t{1} << 20)
        .def("submit", &PyServer::submit, py::arg("queries"), py::arg("k"), py::arg("nprobe"))
        .def("wait", &PyServer::wait, py::arg("ticket"))
        .def("done", &PyServer::done, py::arg("ticket"))
        .def("stats", &PyServer::stats)
        .def("close", &PyServer::close);
}
