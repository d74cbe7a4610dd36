// me8q_bindings.cpp -- the `Me8qIndex` class of the pybind11 module
// `e8qlib._e8qlib`: the reference's Python surface (python_bindings/
// me8q_bindings.cpp:22-315) over the device Me8qIndex.  Same constructor,
// method and argument names, validation order and messages, the same
// sentinel-prefilled outputs and stats dict.  Extra: `search(...,
// nprobe_per_query=array)`, `search_device` for device-resident queries and
// outputs, `device_bytes`.  `num_threads` is accepted and unused.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "me8q_index.h"

namespace py = pybind11;

namespace {

template <typename T>
py::array_t<T, py::array::c_style | py::array::forcecast> EnsureArr(py::handle v, const char* name, int ndim) {
    auto a = py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(v);
    if (!a) throw std::invalid_argument(std::string(name) + " must be a NumPy array");
    if (a.ndim() != ndim)
        throw std::invalid_argument(std::string(name) + (ndim == 2 ? " must be a 2D NumPy array" : " must be a 1D NumPy array"));
    return a;
}

using e8qlib::me8q::Me8qIndex;

int ParseMode(const std::string& mode) {
    if (mode == "plain") return 0;
    if (mode == "batched") return 1;
    throw std::invalid_argument("mode must be 'plain' or 'batched'");
}

class PyMe8qIndex {
   public:
    PyMe8qIndex(size_t dim, float avg_bits, uint64_t seed) : dim_(dim), avg_bits_(avg_bits), seed_(seed) {
        if (dim == 0) throw std::invalid_argument("dim must be positive");
        if (avg_bits <= 0.0F) throw std::invalid_argument("avg_bits must be positive");
    }

    void build(py::handle data, py::handle centroids, py::handle cluster_ids, size_t num_threads, py::handle variance,
               bool drop_energy, float prune_alpha, const std::string& plan_spec, double extra_budget_bits) {
        auto d = EnsureArr<float>(data, "data", 2);
        auto c = EnsureArr<float>(centroids, "centroids", 2);
        auto ids = EnsureArr<uint32_t>(cluster_ids, "cluster_ids", 1);
        if (static_cast<size_t>(d.shape(1)) != dim_) throw std::invalid_argument("data dimension does not match index dim");
        if (static_cast<size_t>(c.shape(1)) != dim_)
            throw std::invalid_argument("centroid dimension does not match index dim");
        if (ids.shape(0) != d.shape(0))
            throw std::invalid_argument("cluster_ids length must match number of rows in data");
        py::array_t<float, py::array::c_style | py::array::forcecast> var;
        const float* varp = nullptr;
        if (!variance.is_none()) {
            var = EnsureArr<float>(variance, "variance", 1);
            if (static_cast<size_t>(var.shape(0)) != dim_) throw std::invalid_argument("variance length must match dim");
            varp = var.data();
        }
        e8qlib::me8q::BuildOptions opt;
        opt.avg_bits = avg_bits_;
        opt.threads = num_threads;
        opt.drop_energy = drop_energy;
        opt.prune_alpha = prune_alpha;
        opt.plan_spec = plan_spec;
        opt.extra_budget_bits = extra_budget_bits;
        if (seed_ != 0) opt.rot_seed = seed_;
        auto idx = std::make_unique<Me8qIndex>();
        {
            py::gil_scoped_release nogil;
            idx->Build(d.data(), dim_, c.data(), dim_, ids.data(), static_cast<size_t>(d.shape(0)), dim_,
                       static_cast<size_t>(c.shape(0)), opt, varp);
        }
        index_ = std::move(idx);
        built_ = true;
    }

    // build() with the rows already in device memory (base_ptr: n x dim
    // float32, cluster_ids_ptr: n uint32); centroids, cluster_ids and the
    // variance (or a plan_spec) as host arrays.
    void build_device(uintptr_t base_ptr, py::handle centroids, py::handle cluster_ids, uintptr_t cluster_ids_ptr,
                      py::handle variance, bool drop_energy, float prune_alpha, const std::string& plan_spec,
                      double extra_budget_bits) {
        auto c = EnsureArr<float>(centroids, "centroids", 2);
        auto ids = EnsureArr<uint32_t>(cluster_ids, "cluster_ids", 1);
        if (!base_ptr || !cluster_ids_ptr) throw std::invalid_argument("null device pointer");
        if (static_cast<size_t>(c.shape(1)) != dim_)
            throw std::invalid_argument("centroid dimension does not match index dim");
        py::array_t<float, py::array::c_style | py::array::forcecast> var;
        const float* varp = nullptr;
        if (!variance.is_none()) {
            var = EnsureArr<float>(variance, "variance", 1);
            if (static_cast<size_t>(var.shape(0)) != dim_) throw std::invalid_argument("variance length must match dim");
            varp = var.data();
        }
        e8qlib::me8q::BuildOptions opt;
        opt.avg_bits = avg_bits_;
        opt.drop_energy = drop_energy;
        opt.prune_alpha = prune_alpha;
        opt.plan_spec = plan_spec;
        opt.extra_budget_bits = extra_budget_bits;
        if (seed_ != 0) opt.rot_seed = seed_;
        auto idx = std::make_unique<Me8qIndex>();
        {
            py::gil_scoped_release nogil;
            idx->BuildFromDevice(reinterpret_cast<const float*>(base_ptr), c.data(), ids.data(),
                                 reinterpret_cast<const uint32_t*>(cluster_ids_ptr), static_cast<size_t>(ids.shape(0)),
                                 dim_, static_cast<size_t>(c.shape(0)), opt, varp);
        }
        index_ = std::move(idx);
        built_ = true;
    }

    py::tuple search(py::handle queries, size_t k, size_t nprobe, size_t num_threads, bool return_stats,
                     py::object nprobe_per_query, const std::string& mode) {
        (void)num_threads;
        const int m = ParseMode(mode);
        auto q = EnsureArr<float>(queries, "queries", 2);
        if (!built_ || !index_) throw std::runtime_error("Me8qIndex must be built or loaded before search");
        if (static_cast<size_t>(q.shape(1)) != dim_) throw std::invalid_argument("query dimension does not match index dim");
        if (k == 0) throw std::invalid_argument("k must be positive");
        if (nprobe == 0) throw std::invalid_argument("nprobe must be positive");
        if (k > e8qlib::me8q::kMaxDeviceK) throw std::invalid_argument("k must be <= 1024 on the device");
        const size_t nq = static_cast<size_t>(q.shape(0));
        py::array_t<uint32_t, py::array::c_style | py::array::forcecast> npq;
        const uint32_t* npq_ptr = nullptr;
        if (!nprobe_per_query.is_none()) {
            npq = EnsureArr<uint32_t>(nprobe_per_query, "nprobe_per_query", 1);
            if (static_cast<size_t>(npq.shape(0)) != nq)
                throw std::invalid_argument("nprobe_per_query length must match number of queries");
            npq_ptr = npq.data();
        }
        const std::vector<ssize_t> shape{static_cast<ssize_t>(nq), static_cast<ssize_t>(k)};
        py::array_t<uint32_t> ids(shape);
        py::array_t<float> dists(shape);
        std::fill(ids.mutable_data(), ids.mutable_data() + ids.size(), std::numeric_limits<uint32_t>::max());
        std::fill(dists.mutable_data(), dists.mutable_data() + dists.size(), std::numeric_limits<float>::infinity());
        e8qlib::me8q::SearchStats st;
        if (nq > 0) {
            py::gil_scoped_release nogil;
            index_->SearchDevice(q.data(), nq, k, nprobe, npq_ptr, ids.mutable_data(), dists.mutable_data(), &st,
                                 false, false, nullptr, m);
        }
        if (!return_stats) return py::make_tuple(ids, dists);
        const double dn = nq > 0 ? static_cast<double>(nq) : 1.0;
        py::dict s;
        s["candidates"] = st.candidates;
        s["survivors"] = st.survivors;
        s["bits_scored"] = st.bits_scored;
        s["bits_total"] = st.bits_total;
        s["pruned"] = st.pruned;
        s["rotate_us"] = 0.0;
        s["route_us"] = 0.0;
        s["scan_us"] = 1000.0 * st.device_scan_ms / dn;
        s["refine_us"] = 0.0;
        s["lut_us"] = 0.0;
        s["filter_us"] = 0.0;
        s["lane_us"] = 0.0;
        s["rest_us"] = 0.0;
        s["survivor_rate"] =
            st.candidates > 0 ? static_cast<double>(st.survivors) / static_cast<double>(st.candidates) : 0.0;
        s["device_scan_ms"] = st.device_scan_ms;
        s["device_total_ms"] = st.device_total_ms;
        s["batches"] = st.batches;
        s["refined"] = st.refined;
        s["form"] = st.form == 1 ? "batched" : "plain";
        s["overflowed"] = st.overflowed;
        s["pairs_scored"] = st.pairs_scored;
        s["items"] = st.items;
        s["refine_first_ms"] = st.refine_ms[0];
        s["refine_cascade_ms"] = st.refine_ms[1];
        s["refine_first_launches"] = st.refine_launches[0];
        s["refine_cascade_launches"] = st.refine_launches[1];
        s["refine_first_bytes"] = st.refine_bytes[0];
        s["refine_cascade_bytes"] = st.refine_bytes[1];
        return py::make_tuple(ids, dists, s);
    }

    py::dict search_device(uintptr_t q_ptr, size_t nq, size_t k, size_t nprobe, uintptr_t ids_ptr,
                           uintptr_t dists_ptr, py::object nprobe_per_query, const std::string& mode) {
        const int m = ParseMode(mode);
        if (!built_ || !index_) throw std::runtime_error("Me8qIndex must be built or loaded before search");
        if (k == 0) throw std::invalid_argument("k must be positive");
        if (nprobe == 0) throw std::invalid_argument("nprobe must be positive");
        py::array_t<uint32_t, py::array::c_style | py::array::forcecast> npq;
        const uint32_t* npq_ptr = nullptr;
        if (!nprobe_per_query.is_none()) {
            npq = EnsureArr<uint32_t>(nprobe_per_query, "nprobe_per_query", 1);
            if (static_cast<size_t>(npq.shape(0)) != nq)
                throw std::invalid_argument("nprobe_per_query length must match number of queries");
            npq_ptr = npq.data();
        }
        e8qlib::me8q::SearchStats st;
        {
            py::gil_scoped_release nogil;
            index_->SearchDevice(reinterpret_cast<const float*>(q_ptr), nq, k, nprobe, npq_ptr,
                                 reinterpret_cast<uint32_t*>(ids_ptr), reinterpret_cast<float*>(dists_ptr), &st, true,
                                 true, nullptr, m);
        }
        py::dict s;
        s["candidates"] = st.candidates;
        s["survivors"] = st.survivors;
        s["bits_scored"] = st.bits_scored;
        s["bits_total"] = st.bits_total;
        s["pruned"] = st.pruned;
        s["batches"] = st.batches;
        s["refined"] = st.refined;
        s["groups"] = st.groups;
        s["device_scan_ms"] = st.device_scan_ms;
        s["device_total_ms"] = st.device_total_ms;
        s["form"] = st.form == 1 ? "batched" : "plain";
        s["overflowed"] = st.overflowed;
        s["pairs_scored"] = st.pairs_scored;
        s["items"] = st.items;
        s["refine_first_ms"] = st.refine_ms[0];
        s["refine_cascade_ms"] = st.refine_ms[1];
        s["refine_first_launches"] = st.refine_launches[0];
        s["refine_cascade_launches"] = st.refine_launches[1];
        s["refine_first_bytes"] = st.refine_bytes[0];
        s["refine_cascade_bytes"] = st.refine_bytes[1];
        return s;
    }

    // The reference binding's estimate (me8q_bindings.cpp:209-218).
    float estimate(py::handle query, uint32_t id) const {
        if (!built_ || !index_) throw std::runtime_error("Me8qIndex must be built or loaded before estimate");
        auto q = EnsureArr<float>(query, "query", 1);
        if (static_cast<size_t>(q.shape(0)) != dim_)
            throw std::invalid_argument("query dimension does not match index dim");
        py::gil_scoped_release nogil;
        return index_->Estimate(q.data(), id);
    }

    // Me8qIndex::ExhaustiveTopK (me8q_search.cpp:238-266) of each query row,
    // k wide, UINT32_MAX / +inf padded.
    py::tuple exhaustive_topk(py::handle queries, size_t k) const {
        if (!built_ || !index_) throw std::runtime_error("Me8qIndex must be built or loaded before search");
        auto q = EnsureArr<float>(queries, "queries", 2);
        if (static_cast<size_t>(q.shape(1)) != dim_)
            throw std::invalid_argument("queries dimension does not match index dim");
        const size_t nq = static_cast<size_t>(q.shape(0));
        py::array_t<uint32_t> ids({nq, k});
        py::array_t<float> dists({nq, k});
        if (nq > 0 && k > 0) {
            py::gil_scoped_release nogil;
            index_->ExhaustiveTopKBatch(q.data(), nq, k, ids.mutable_data(), dists.mutable_data());
        }
        return py::make_tuple(ids, dists);
    }

    // Me8qIndex::SearchIVF (me8q_search.cpp:136-236) of each query row.
    py::tuple search_ivf(py::handle queries, size_t k, size_t nprobe, bool return_stats) const {
        if (!built_ || !index_) throw std::runtime_error("Me8qIndex must be built or loaded before search");
        auto q = EnsureArr<float>(queries, "queries", 2);
        if (static_cast<size_t>(q.shape(1)) != dim_)
            throw std::invalid_argument("queries dimension does not match index dim");
        const size_t nq = static_cast<size_t>(q.shape(0));
        py::array_t<uint32_t> ids({nq, k});
        py::array_t<float> dists({nq, k});
        std::fill(ids.mutable_data(), ids.mutable_data() + nq * k, std::numeric_limits<uint32_t>::max());
        std::fill(dists.mutable_data(), dists.mutable_data() + nq * k, std::numeric_limits<float>::infinity());
        std::vector<uint64_t> st(nq * 6, 0);
        {
            py::gil_scoped_release nogil;
            index_->SearchIVFBatch(q.data(), nq, k, nprobe, ids.mutable_data(), dists.mutable_data(), st.data());
        }
        if (!return_stats) return py::make_tuple(ids, dists);
        py::dict s;
        uint64_t tot[6] = {0, 0, 0, 0, 0, 0};
        for (size_t i = 0; i < nq; ++i)
            for (int f = 0; f < 6; ++f) tot[f] += st[i * 6 + f];
        s["candidates"] = tot[0];
        s["clusters"] = tot[1];
        s["bits_scored"] = tot[2];
        s["bits_total"] = tot[3];
        s["survivors"] = tot[4];
        s["pruned"] = tot[5];
        return py::make_tuple(ids, dists, s);
    }

    void save(const std::string& path) const {
        if (!built_ || !index_) throw std::runtime_error("Me8qIndex must be built or loaded before save");
        index_->Save(path.c_str());
    }

    static PyMe8qIndex load(const std::string& path) {
        PyMe8qIndex w(1, 4.0F, 0);
        auto idx = std::make_unique<Me8qIndex>();
        idx->Load(path.c_str());
        w.dim_ = idx->dim();
        w.avg_bits_ = static_cast<float>(idx->bits_per_dim());
        w.index_ = std::move(idx);
        w.built_ = true;
        return w;
    }

    size_t dim() const { return dim_; }
    float avg_bits() const { return avg_bits_; }
    bool is_built() const { return built_; }
    size_t n() const { return index_ ? index_->n() : 0; }
    size_t num_clusters() const { return index_ ? index_->num_clusters() : 0; }
    double bits_per_dim() const { return index_ ? index_->bits_per_dim() : 0.0; }
    size_t memory_bytes() const { return index_ ? index_->memory_bytes() : 0; }
    size_t device_bytes() const { return index_ ? index_->device_bytes() : 0; }
    // The C++ index behind this object (for the serving loop).
    uintptr_t native() const { return reinterpret_cast<uintptr_t>(index_.get()); }
    std::vector<std::pair<size_t, int>> plan() const {
        std::vector<std::pair<size_t, int>> out;
        if (!index_) return out;
        for (const auto& s : index_->shared_plan()) out.emplace_back(s.len, s.bits);
        return out;
    }

   private:
    size_t dim_ = 0;
    float avg_bits_ = 4.0F;
    uint64_t seed_ = 0;
    bool built_ = false;
    std::unique_ptr<Me8qIndex> index_;
};

}  // namespace

void RegisterMe8q(py::module_& m) {
    py::class_<PyMe8qIndex>(m, "Me8qIndex")
        .def(py::init<size_t, float, uint64_t>(), py::arg("dim"), py::arg("avg_bits") = 4.0F, py::arg("seed") = 0)
        .def("build", &PyMe8qIndex::build, py::arg("data"), py::arg("centroids"), py::arg("cluster_ids"),
             py::arg("num_threads") = 1, py::arg("variance") = py::none(), py::arg("drop_energy") = true,
             py::arg("prune_alpha") = 0.0F, py::arg("plan_spec") = "", py::arg("extra_budget_bits") = 0.0)
        .def("build_device", &PyMe8qIndex::build_device, py::arg("base_ptr"), py::arg("centroids"),
             py::arg("cluster_ids"), py::arg("cluster_ids_ptr"), py::arg("variance") = py::none(),
             py::arg("drop_energy") = true, py::arg("prune_alpha") = 0.0F, py::arg("plan_spec") = "",
             py::arg("extra_budget_bits") = 0.0)
        .def("search", &PyMe8qIndex::search, py::arg("queries"), py::arg("k"), py::arg("nprobe"),
             py::arg("num_threads") = 1, py::arg("return_stats") = false, py::arg("nprobe_per_query") = py::none(),
             py::arg("mode") = "plain")
        .def("search_device", &PyMe8qIndex::search_device, py::arg("queries_ptr"), py::arg("nq"), py::arg("k"),
             py::arg("nprobe"), py::arg("out_ids_ptr"), py::arg("out_dists_ptr"),
             py::arg("nprobe_per_query") = py::none(), py::arg("mode") = "plain")
        .def("estimate", &PyMe8qIndex::estimate, py::arg("query"), py::arg("id"))
        .def("exhaustive_topk", &PyMe8qIndex::exhaustive_topk, py::arg("queries"), py::arg("k"))
        .def("search_ivf", &PyMe8qIndex::search_ivf, py::arg("queries"), py::arg("k"), py::arg("nprobe"),
             py::arg("return_stats") = false)
        .def("save", &PyMe8qIndex::save, py::arg("path"))
        .def_static("load", &PyMe8qIndex::load, py::arg("path"))
        .def_property_readonly("dim", &PyMe8qIndex::dim)
        .def_property_readonly("avg_bits", &PyMe8qIndex::avg_bits)
        .def_property_readonly("is_built", &PyMe8qIndex::is_built)
        .def_property_readonly("n", &PyMe8qIndex::n)
        .def_property_readonly("num_clusters", &PyMe8qIndex::num_clusters)
        .def_property_readonly("bits_per_dim", &PyMe8qIndex::bits_per_dim)
        .def_property_readonly("memory_bytes", &PyMe8qIndex::memory_bytes)
        .def_property_readonly("device_bytes", &PyMe8qIndex::device_bytes)
        .def_property_readonly("plan", &PyMe8qIndex::plan)
        .def("_native", &PyMe8qIndex::native);
}
