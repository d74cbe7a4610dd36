// e8q_bindings.cpp -- pybind11 module `e8qlib._e8qlib`: the reference's
// Python surface (python_bindings/e8q_bindings.cpp:22-344) over the device
// E8qIndex.  Same constructor, method and argument names, same validation
// order and messages, same sentinel-prefilled (ids uint32, dists float32)
// outputs and stats dict.  Extra keyword arguments: `mode` ("plain" |
// "batched" | "mfma") and `nprobe_per_query`.  `num_threads` is accepted and unused:
// the device sizes its own parallelism.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "e8q_index.h"

namespace py = pybind11;

namespace {

template <typename T>
py::array_t<T, py::array::c_style | py::array::forcecast> Ensure(py::handle v, const char* name, int ndim) {
    auto a = py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(v);
    if (!a) throw std::invalid_argument(std::string(name) + " must be a NumPy array");
    if (a.ndim() != ndim)
        throw std::invalid_argument(std::string(name) + (ndim == 2 ? " must be a 2D NumPy array" : " must be a 1D NumPy array"));
    return a;
}

e8qlib::SearchMode ParseMode(const std::string& m) {
    if (m == "plain") return e8qlib::SearchMode::kPlain;
    if (m == "batched") return e8qlib::SearchMode::kBatched;
    if (m == "mfma") return e8qlib::SearchMode::kMfma;
    throw std::invalid_argument("mode must be 'plain', 'batched' or 'mfma'");
}

class PyE8qIndex {
   public:
    PyE8qIndex(size_t dim, size_t nbits, uint64_t seed) : dim_(dim), nbits_(nbits), seed_(seed) {
        if (dim == 0) throw std::invalid_argument("dim must be positive");
        if (nbits < 1 || nbits > 13) throw std::invalid_argument("nbits must be in [1, 13]");
    }

    void build(py::handle data, py::handle centroids, py::handle cluster_ids, size_t num_threads,
               float prune_alpha) {
        auto d = Ensure<float>(data, "data", 2);
        auto c = Ensure<float>(centroids, "centroids", 2);
        auto ids = Ensure<uint32_t>(cluster_ids, "cluster_ids", 1);
        if (static_cast<size_t>(d.shape(1)) != dim_) throw std::invalid_argument("data dimension does not match index dim");
        if (static_cast<size_t>(c.shape(1)) != dim_)
            throw std::invalid_argument("centroid dimension does not match index dim");
        if (ids.shape(0) != d.shape(0))
            throw std::invalid_argument("cluster_ids length must match number of rows in data");
        auto idx = std::make_unique<e8qlib::E8qIndex>();
        e8qlib::E8qBuildOptions opt;
        opt.prune_alpha = prune_alpha;
        {
            py::gil_scoped_release nogil;
            idx->Build(d.data(), c.data(), ids.data(), static_cast<size_t>(d.shape(0)), dim_,
                       static_cast<size_t>(c.shape(0)), static_cast<int>(nbits_), seed_, num_threads, opt);
        }
        index_ = std::move(idx);
    }

    // Build from rows already in device memory (base [n][dim] and centroids
    // [num_clusters][dim] float32 device pointers, cluster ids both as a host
    // array and as a device pointer): no host copy of the base.
    void build_device(uintptr_t base_ptr, uintptr_t centroids_ptr, size_t num_clusters, py::handle cluster_ids,
                      uintptr_t cluster_ids_ptr, float prune_alpha) {
        auto ids = Ensure<uint32_t>(cluster_ids, "cluster_ids", 1);
        if (!base_ptr || !centroids_ptr || !cluster_ids_ptr) throw std::invalid_argument("null device pointer");
        auto idx = std::make_unique<e8qlib::E8qIndex>();
        e8qlib::E8qBuildOptions opt;
        opt.prune_alpha = prune_alpha;
        {
            py::gil_scoped_release nogil;
            idx->BuildFromDevice(reinterpret_cast<const float*>(base_ptr), reinterpret_cast<const float*>(centroids_ptr),
                                 ids.data(), reinterpret_cast<const uint32_t*>(cluster_ids_ptr),
                                 static_cast<size_t>(ids.shape(0)), dim_, num_clusters, static_cast<int>(nbits_), seed_,
                                 opt);
        }
        index_ = std::move(idx);
    }

    py::tuple search(py::handle queries, size_t k, size_t nprobe, size_t num_threads, bool return_stats,
                     const std::string& mode, py::object nprobe_per_query) {
        (void)num_threads;
        auto q = Ensure<float>(queries, "queries", 2);
        if (!index_) throw std::runtime_error("E8qIndex must be built or loaded before search");
        if (static_cast<size_t>(q.shape(1)) != dim_) throw std::invalid_argument("query dimension does not match index dim");
        if (k == 0) throw std::invalid_argument("k must be positive");
        if (k > e8qlib::kMaxK) throw std::invalid_argument("k must be <= " + std::to_string(e8qlib::kMaxK));
        if (nprobe == 0) throw std::invalid_argument("nprobe must be positive");
        const e8qlib::SearchMode m = ParseMode(mode);
        const size_t nq = static_cast<size_t>(q.shape(0));
        py::array_t<uint32_t, py::array::c_style | py::array::forcecast> npq;
        const uint32_t* npq_ptr = nullptr;
        if (!nprobe_per_query.is_none()) {
            npq = Ensure<uint32_t>(nprobe_per_query, "nprobe_per_query", 1);
            if (static_cast<size_t>(npq.shape(0)) != nq)
                throw std::invalid_argument("nprobe_per_query length must match number of queries");
            for (size_t i = 0; i < nq; ++i)
                if (npq.data()[i] == 0) throw std::invalid_argument("nprobe must be positive");
            npq_ptr = npq.data();
        }
        const std::vector<ssize_t> shape{static_cast<ssize_t>(nq), static_cast<ssize_t>(k)};
        py::array_t<uint32_t> ids(shape);
        py::array_t<float> dists(shape);
        std::fill(ids.mutable_data(), ids.mutable_data() + ids.size(), std::numeric_limits<uint32_t>::max());
        std::fill(dists.mutable_data(), dists.mutable_data() + dists.size(), std::numeric_limits<float>::infinity());
        e8qlib::SearchStats st;
        if (nq > 0) {
            uint32_t* ip = ids.mutable_data();
            float* dp = dists.mutable_data();
            py::gil_scoped_release nogil;
            index_->SearchBatch(q.data(), nq, k, nprobe, npq_ptr, ip, dp, &st, m);
        }
        if (!return_stats) return py::make_tuple(ids, dists);
        py::dict s;
        s["candidates"] = st.candidates;
        s["survivors"] = st.survivors;
        s["rotate_us"] = 0.0;
        s["route_us"] = 0.0;
        s["scan_us"] = 0.0;
        s["refine_us"] = 0.0;
        s["survivor_rate"] =
            st.candidates > 0 ? static_cast<double>(st.survivors) / static_cast<double>(st.candidates) : 0.0;
        s["batches"] = st.batches;
        s["scanned"] = st.scanned;
        s["device_scan_ms"] = st.device_scan_ms;
        s["device_total_ms"] = st.device_total_ms;
        return py::make_tuple(ids, dists, s);
    }

    // Search over device-resident buffers (e.g. torch tensors' data_ptr()):
    // queries [nq, dim] float32, out_ids [nq, k] uint32, out_dists [nq, k]
    // float32, all on the index's device and ready (caller synchronized).
    py::dict search_device(uintptr_t queries, size_t nq, size_t k, size_t nprobe, uintptr_t out_ids,
                           uintptr_t out_dists, const std::string& mode, py::object nprobe_per_query) {
        if (!index_) throw std::runtime_error("E8qIndex must be built or loaded before search");
        if (k == 0) throw std::invalid_argument("k must be positive");
        if (k > e8qlib::kMaxK) throw std::invalid_argument("k must be <= " + std::to_string(e8qlib::kMaxK));
        if (nprobe == 0) throw std::invalid_argument("nprobe must be positive");
        const e8qlib::SearchMode m = ParseMode(mode);
        py::array_t<uint32_t, py::array::c_style | py::array::forcecast> npq;
        const uint32_t* npq_ptr = nullptr;
        if (!nprobe_per_query.is_none()) {
            npq = Ensure<uint32_t>(nprobe_per_query, "nprobe_per_query", 1);
            if (static_cast<size_t>(npq.shape(0)) != nq)
                throw std::invalid_argument("nprobe_per_query length must match number of queries");
            npq_ptr = npq.data();
        }
        e8qlib::SearchStats st;
        {
            py::gil_scoped_release nogil;
            index_->SearchBatch(reinterpret_cast<const float*>(queries), nq, k, nprobe, npq_ptr,
                                reinterpret_cast<uint32_t*>(out_ids), reinterpret_cast<float*>(out_dists), &st, m,
                                true, true);
        }
        py::dict s;
        s["candidates"] = st.candidates;
        s["survivors"] = st.survivors;
        s["batches"] = st.batches;
        s["scanned"] = st.scanned;
        s["device_scan_ms"] = st.device_scan_ms;
        s["device_total_ms"] = st.device_total_ms;
        return s;
    }

    void set_prune_alpha(float alpha) {
        if (!index_) throw std::runtime_error("E8qIndex must be built or loaded before set_prune_alpha");
        if (!(alpha > 0.0F)) throw std::invalid_argument("prune_alpha must be positive");
        index_->SetPruneAlpha(alpha);
    }

    void set_disable_prune(bool on) {
        if (!index_) throw std::runtime_error("E8qIndex must be built or loaded before set_disable_prune");
        index_->SetDisablePrune(on);
    }

    // The reference binding's estimate (e8q_bindings.cpp:171-180).
    float estimate(py::handle query, uint32_t id) const {
        if (!index_) throw std::runtime_error("E8qIndex must be built or loaded before estimate");
        auto q = Ensure<float>(query, "query", 1);
        if (static_cast<size_t>(q.shape(0)) != dim_) throw std::invalid_argument("query dimension does not match index dim");
        py::gil_scoped_release nogil;
        return index_->Estimate(q.data(), id);
    }

    // The reference binding's batched ZPR bound probe (e8q_bindings.cpp:
    // 181-234): pair i probes base_rows[i] (the build row of ids[i]) with
    // queries[qidx[i]]; NaN for a pair out of range.  One device call for all
    // pairs (num_threads is accepted for the reference's signature).
    py::array_t<double> probe_bound_w(py::handle queries, py::handle base_rows, py::handle qidx, py::handle ids,
                                      size_t num_threads) const {
        (void)num_threads;
        if (!index_) throw std::runtime_error("E8qIndex must be built or loaded before probe_bound_w");
        auto q = Ensure<float>(queries, "queries", 2);
        auto b = Ensure<float>(base_rows, "base_rows", 2);
        auto qi = Ensure<uint32_t>(qidx, "qidx", 1);
        auto ii = Ensure<uint32_t>(ids, "ids", 1);
        if (static_cast<size_t>(q.shape(1)) != dim_) throw std::invalid_argument("queries dimension does not match index dim");
        if (static_cast<size_t>(b.shape(1)) != dim_)
            throw std::invalid_argument("base_rows dimension does not match index dim");
        const size_t m = static_cast<size_t>(ii.shape(0));
        if (static_cast<size_t>(qi.shape(0)) != m || static_cast<size_t>(b.shape(0)) != m)
            throw std::invalid_argument("qidx, ids and base_rows must have the same length");
        py::array_t<double> out(static_cast<py::ssize_t>(m));
        if (m == 0) return out;
        py::gil_scoped_release nogil;
        index_->ProbeBoundWBatch(q.data(), static_cast<size_t>(q.shape(0)), b.data(), qi.data(), ii.data(), m,
                                 out.mutable_data());
        return out;
    }

    py::list exhaustive_topk(py::handle query, size_t k) const {
        if (!index_) throw std::runtime_error("E8qIndex must be built or loaded before search");
        auto q = Ensure<float>(query, "query", 1);
        if (static_cast<size_t>(q.shape(0)) != dim_) throw std::invalid_argument("query dimension does not match index dim");
        std::vector<std::pair<float, uint32_t>> out;
        index_->ExhaustiveTopK(q.data(), k, out);
        py::list l;
        for (auto& p : out) l.append(py::make_tuple(p.first, p.second));
        return l;
    }

    void save(const std::string& path) const {
        if (!index_) throw std::runtime_error("E8qIndex must be built or loaded before save");
        index_->Save(path.c_str());
    }

    static PyE8qIndex load(const std::string& path) {
        PyE8qIndex w(1, 1, 0);
        auto idx = std::make_unique<e8qlib::E8qIndex>();
        idx->Load(path.c_str());
        w.dim_ = idx->dim();
        w.nbits_ = static_cast<size_t>(idx->bits());
        w.index_ = std::move(idx);
        return w;
    }

    size_t dim() const { return dim_; }
    size_t nbits() const { return nbits_; }
    bool is_built() const { return index_ != nullptr; }
    size_t n() const { return index_ ? index_->n() : 0; }
    size_t num_clusters() const { return index_ ? index_->num_clusters() : 0; }
    size_t memory_bytes() const { return index_ ? index_->memory_bytes() : 0; }
    // The C++ index behind this object (for the serving loop).
    uintptr_t native() const { return reinterpret_cast<uintptr_t>(index_.get()); }
    size_t device_bytes() const { return index_ ? index_->device_bytes() : 0; }
    bool disable_prune() const { return index_ ? index_->disable_prune() : false; }
    float prune_alpha() const { return index_ ? index_->prune_alpha() : 0.0F; }

   private:
    size_t dim_, nbits_;
    uint64_t seed_;
    std::shared_ptr<e8qlib::E8qIndex> index_;
};

}  // namespace

void RegisterMe8q(py::module_& m);   // me8q_bindings.cpp
void RegisterServe(py::module_& m);  // serve_bindings.cpp

PYBIND11_MODULE(_e8qlib, m) {
    m.doc() = "E8Q / ME8Q IVF indexes on MI355X (gfx950)";
    RegisterMe8q(m);
    py::class_<PyE8qIndex>(m, "E8qIndex")
        .def(py::init<size_t, size_t, uint64_t>(), py::arg("dim"), py::arg("nbits"), py::arg("seed") = 0)
        .def("build", &PyE8qIndex::build, py::arg("data"), py::arg("centroids"), py::arg("cluster_ids"),
             py::arg("num_threads") = 1, py::arg("prune_alpha") = 0.0F)
        .def("build_device", &PyE8qIndex::build_device, py::arg("base_ptr"), py::arg("centroids_ptr"),
             py::arg("num_clusters"), py::arg("cluster_ids"), py::arg("cluster_ids_ptr"), py::arg("prune_alpha") = 0.0F)
        .def("search", &PyE8qIndex::search, py::arg("queries"), py::arg("k"), py::arg("nprobe"),
             py::arg("num_threads") = 1, py::arg("return_stats") = false, py::arg("mode") = "plain",
             py::arg("nprobe_per_query") = py::none())
        .def("search_device", &PyE8qIndex::search_device, py::arg("queries_ptr"), py::arg("nq"), py::arg("k"),
             py::arg("nprobe"), py::arg("out_ids_ptr"), py::arg("out_dists_ptr"), py::arg("mode") = "plain",
             py::arg("nprobe_per_query") = py::none())
        .def("exhaustive_topk", &PyE8qIndex::exhaustive_topk, py::arg("query"), py::arg("k"))
        .def("estimate", &PyE8qIndex::estimate, py::arg("query"), py::arg("id"))
        .def("probe_bound_w", &PyE8qIndex::probe_bound_w, py::arg("queries"), py::arg("base"), py::arg("qidx"),
             py::arg("ids"), py::arg("num_threads") = 1)
        .def("set_disable_prune", &PyE8qIndex::set_disable_prune, py::arg("on"))
        .def("set_prune_alpha", &PyE8qIndex::set_prune_alpha, py::arg("alpha"))
        .def("save", &PyE8qIndex::save, py::arg("path"))
        .def_static("load", &PyE8qIndex::load, py::arg("path"))
        .def_property_readonly("dim", &PyE8qIndex::dim)
        .def_property_readonly("nbits", &PyE8qIndex::nbits)
        .def_property_readonly("is_built", &PyE8qIndex::is_built)
        .def_property_readonly("n", &PyE8qIndex::n)
        .def_property_readonly("num_clusters", &PyE8qIndex::num_clusters)
        .def_property_readonly("memory_bytes", &PyE8qIndex::memory_bytes)
        .def_property_readonly("device_bytes", &PyE8qIndex::device_bytes)
        .def_property_readonly("disable_prune", &PyE8qIndex::disable_prune)
        .def_property_readonly("prune_alpha", &PyE8qIndex::prune_alpha)
        .def("_native", &PyE8qIndex::native);
    // Launch options of the process-wide device context.
    m.def("set_search_graphs", &e8qlib::SetSearchGraphs, py::arg("on"),
          "Replay repeated search shapes as HIP graphs (first call eager, second captures).");
    m.def("set_batched_shape", &e8qlib::SetBatchedShape, py::arg("qtile") = 0, py::arg("btile") = 0,
          py::arg("grid") = 0, "Fix the batched scan's (pairs, batches) per item and grid; zeros: autotuned.");
    m.def("search_context_info", [] {
        const e8qlib::ContextInfo i = e8qlib::GetContextInfo();
        py::dict d;
        d["qtile"] = i.qtile;
        d["btile"] = i.btile;
        d["grid"] = i.grid;
        d["graphs"] = i.graphs;
        d["tunings"] = i.tunings;
        d["captures"] = i.captures;
        d["replays"] = i.replays;
        d["me8q_mfma_grid"] = i.m8_mfma_grid;
        d["me8q_refine_grid"] = i.m8_refine_grid;
        d["me8q_tunings"] = i.m8_tunings;
        return d;
    });
    m.def("set_me8q_grids", &e8qlib::SetMe8qGrids, py::arg("mfma") = 0, py::arg("refine") = 0,
          "Fix the batched ME8Q form's MFMA-contraction and refine grids; zeros: autotuned.");
    RegisterServe(m);
}
