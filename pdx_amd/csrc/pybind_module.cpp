// pybind11 module `pdx_amd.compiled`: the Python face of the host index API, with the same class,
// constructor arguments, methods and errors as the reference's `pdxearch.compiled`
// (python/lib.cpp:6-60, include/pdx/lib/lib.hpp:16-175), plus batched GPU search.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "host/pdx_index.hpp"

namespace py = pybind11;

namespace {

PDX::DistanceMetric to_metric(uint8_t m) {  // lib.hpp:16-27
    switch (m) {
        case 0: return PDX::DistanceMetric::L2SQ;
        case 1: return PDX::DistanceMetric::COSINE;
        case 2: return PDX::DistanceMetric::IP;
        default: throw std::runtime_error("Unknown distance metric: " + std::to_string(m));
    }
}

using Result = std::pair<py::array_t<uint32_t>, py::array_t<float>>;

Result to_numpy(const std::vector<PDX::KNNCandidate>& r) {
    py::array_t<uint32_t> ids(r.size());
    py::array_t<float> dists(r.size());
    auto pi = ids.mutable_unchecked<1>();
    auto pd = dists.mutable_unchecked<1>();
    for (size_t i = 0; i < r.size(); ++i) {
        pi(i) = r[i].index;
        pd(i) = r[i].distance;
    }
    return {ids, dists};
}

class PyPDXIndex {
  public:
    std::unique_ptr<PDX::IPDXIndex> index;
    int servers = 0;  // running servers hold the device copy; changes to the index wait for them to close

    void require_no_server(const char* what) const {
        if (servers)
            throw std::runtime_error(std::string(what) + ": close the running server(s) of this index first");
    }

    PyPDXIndex() = default;
    PyPDXIndex(const std::string& type, uint32_t d, uint8_t metric, uint32_t seed, uint32_t num_clusters,
               uint32_t num_meso_clusters, bool normalize, float sampling_fraction, uint32_t kmeans_iters,
               bool hierarchical_indexing, uint32_t n_threads) {
        PDX::PDXIndexConfig c{};
        c.num_dimensions = d;
        c.distance_metric = to_metric(metric);
        c.seed = seed;
        c.num_clusters = num_clusters;
        c.num_meso_clusters = num_meso_clusters;
        c.normalize = normalize;
        c.sampling_fraction = sampling_fraction;
        c.kmeans_iters = kmeans_iters;
        c.hierarchical_indexing = hierarchical_indexing;
        c.n_threads = n_threads;
        if (type == "pdx_f32") index = std::make_unique<PDX::PDXIndexF32>(c);
        else if (type == "pdx_u8") index = std::make_unique<PDX::PDXIndexU8>(c);
        else if (type == "pdx_tree_f32") index = std::make_unique<PDX::PDXTreeIndexF32>(c);
        else if (type == "pdx_tree_u8") index = std::make_unique<PDX::PDXTreeIndexU8>(c);
        else
            throw std::runtime_error("Unknown index type: " + type +
                                     ". Valid types: pdx_f32, pdx_u8, pdx_tree_f32, pdx_tree_u8");
    }

    static PyPDXIndex Load(const std::string& path) {
        PyPDXIndex p;
        p.index = PDX::LoadPDXIndex(path);
        return p;
    }

    void BuildIndex(const py::array_t<float, py::array::c_style | py::array::forcecast>& data) {
        require_no_server("build_index");
        auto b = data.request();
        if (b.ndim != 2) throw std::runtime_error("data must be a 2D numpy array (n_embeddings x dimensions)");
        if ((uint32_t)b.shape[1] != index->GetNumDimensions() && index->GetNumDimensions() != 0)
            throw std::invalid_argument("data has " + std::to_string(b.shape[1]) + " columns, index expects " +
                                        std::to_string(index->GetNumDimensions()));
        py::gil_scoped_release nogil;
        index->BuildIndex(static_cast<const float*>(b.ptr), (size_t)b.shape[0]);
    }

    void check_query(const py::buffer_info& b) const {
        if (b.ndim != 1) throw std::runtime_error("query must be a 1D numpy array");
        if ((uint32_t)b.shape[0] != index->GetNumDimensions())
            throw std::invalid_argument("query has " + std::to_string(b.shape[0]) + " dimensions, index expects " +
                                        std::to_string(index->GetNumDimensions()));
    }

    Result Search(const py::array_t<float, py::array::c_style | py::array::forcecast>& q, uint32_t knn) const {
        auto b = q.request();
        check_query(b);
        std::vector<PDX::KNNCandidate> r;
        {
            py::gil_scoped_release nogil;
            r = index->Search(static_cast<const float*>(b.ptr), knn);
        }
        return to_numpy(r);
    }

    Result FilteredSearch(const py::array_t<float, py::array::c_style | py::array::forcecast>& q, uint32_t knn,
                          const py::array_t<size_t, py::array::c_style | py::array::forcecast>& rows) const {
        auto b = q.request();
        check_query(b);
        auto rb = rows.request();
        if (rb.ndim != 1) throw std::runtime_error("row_ids must be a 1D numpy array");
        const size_t* rp = static_cast<const size_t*>(rb.ptr);
        std::vector<size_t> passing(rp, rp + rb.shape[0]);
        std::vector<PDX::KNNCandidate> r;
        {
            py::gil_scoped_release nogil;
            r = index->FilteredSearch(static_cast<const float*>(b.ptr), knn, passing);
        }
        return to_numpy(r);
    }

    // Batched: queries (nq x d), knn scalar or per-query array, nprobe scalar (None = current) or
    // per-query array. Returns (ids [nq, kmax], dists [nq, kmax], counts [nq]).
    py::tuple SearchBatch(const py::array_t<float, py::array::c_style | py::array::forcecast>& q, py::object knn,
                          py::object nprobe) const {
        auto b = q.request();
        if (b.ndim != 2) throw std::runtime_error("queries must be a 2D numpy array");
        const size_t nq = (size_t)b.shape[0];
        std::vector<uint32_t> kv, npv;
        uint32_t kscalar = 0;
        if (py::isinstance<py::int_>(knn)) {
            kscalar = knn.cast<uint32_t>();
        } else {
            kv = knn.cast<std::vector<uint32_t>>();
            if (kv.size() != nq) throw std::invalid_argument("knn array must have one entry per query");
            for (uint32_t k : kv) kscalar = std::max(kscalar, k);
        }
        if (!nprobe.is_none()) {
            if (py::isinstance<py::int_>(nprobe)) index->SetNProbe(nprobe.cast<uint32_t>());
            else {
                npv = nprobe.cast<std::vector<uint32_t>>();
                if (npv.size() != nq) throw std::invalid_argument("nprobe array must have one entry per query");
            }
        }
        if (kscalar == 0) throw std::invalid_argument("knn must be >= 1");
        py::array_t<uint32_t> ids({nq, (size_t)kscalar});
        py::array_t<float> dists({nq, (size_t)kscalar});
        py::array_t<uint32_t> counts(nq);
        {
            auto pi = ids.mutable_data();
            auto pd = dists.mutable_data();
            auto pc = counts.mutable_data();
            const float* qp = static_cast<const float*>(b.ptr);
            py::gil_scoped_release nogil;
            index->SearchBatch(qp, nq, kscalar, kv.empty() ? nullptr : kv.data(), npv.empty() ? nullptr : npv.data(),
                               kscalar, pi, pd, pc);
        }
        return py::make_tuple(ids, dists, counts);
    }

    void Append(size_t row_id, const py::array_t<float, py::array::c_style | py::array::forcecast>& e) {
        require_no_server("append");
        auto b = e.request();
        if (b.ndim != 1) throw std::runtime_error("embedding must be a 1D numpy array");
        if ((uint32_t)b.shape[0] != index->GetNumDimensions())
            throw std::invalid_argument("embedding has the wrong number of dimensions");
        index->Append(row_id, static_cast<const float*>(b.ptr));
    }

    py::array_t<uint32_t> GetClusterRowIds(uint32_t c) const {
        auto v = index->GetClusterRowIds(c);
        py::array_t<uint32_t> r(v.size());
        std::memcpy(r.mutable_data(), v.data(), v.size() * 4);
        return r;
    }

    size_t DeviceBytes() const { return pdxg_index_device_bytes(index->DeviceHandle()); }
};

// Continuous batching over an index's device copy (pdxg_server_*, include/pdx_amd.h): requests of one
// query each, with their own k and nprobe, are batched as slots free up. Not in the reference, whose
// Python API serves one query per call.
class PyServer {
  public:
    PyServer(PyPDXIndex& ix, uint32_t max_batch, uint32_t max_k, uint32_t max_nprobe, uint32_t max_slots,
             size_t scratch_budget_bytes, uint32_t pipeline_min)
        : owner_(&ix), d_(ix.index->GetNumDimensions()), max_k_(max_k) {
        pdxg_server_config c{};
        c.max_batch = max_batch;
        c.max_k = max_k;
        c.max_nprobe = max_nprobe;
        c.max_slots = max_slots;
        c.scratch_budget_bytes = scratch_budget_bytes;
        c.pipeline_min = pipeline_min;
        PDX::detail::check(pdxg_server_create(ix.index->DeviceHandle(), &c, &sv_));
        owner_->servers++;
    }
    ~PyServer() { close(); }
    void close() {
        if (!sv_) return;
        {
            py::gil_scoped_release nogil;
            pdxg_server_destroy(sv_);
        }
        sv_ = nullptr;
        owner_->servers--;
    }
    void live() const {
        if (!sv_) throw std::runtime_error("server is closed");
    }
    void Submit(const py::array_t<float, py::array::c_style | py::array::forcecast>& q, uint32_t knn, uint32_t nprobe,
                uint64_t tag) {
        live();
        auto b = q.request();
        if (b.ndim != 1 || (uint32_t)b.shape[0] != d_)
            throw std::invalid_argument("query must be a 1D numpy array of " + std::to_string(d_) + " floats");
        PDX::detail::check(pdxg_server_submit(sv_, static_cast<const float*>(b.ptr), knn, nprobe, tag));
    }
    // [(tag, ids, distances, latency_ms), ...] of up to max_results completed requests, waiting up to
    // timeout_ms for the first one.
    py::list Poll(uint32_t max_results, int timeout_ms) {
        live();
        std::vector<uint64_t> tags(max_results);
        std::vector<uint32_t> ids((size_t)max_results * max_k_), cnt(max_results);
        std::vector<float> dists((size_t)max_results * max_k_), lat(max_results);
        uint32_t got = 0;
        {
            py::gil_scoped_release nogil;
            PDX::detail::check(pdxg_server_poll(sv_, max_results, timeout_ms, tags.data(), max_k_, ids.data(),
                                                dists.data(), cnt.data(), lat.data(), &got));
        }
        py::list out;
        for (uint32_t r = 0; r < got; ++r) {
            std::vector<PDX::KNNCandidate> res(cnt[r]);
            for (uint32_t i = 0; i < cnt[r]; ++i) res[i] = {ids[(size_t)r * max_k_ + i], dists[(size_t)r * max_k_ + i]};
            auto [pi, pd] = to_numpy(res);
            out.append(py::make_tuple(tags[r], pi, pd, lat[r]));
        }
        return out;
    }
    py::dict Stats() {
        live();
        pdxg_server_stats st{};
        PDX::detail::check(pdxg_server_get_stats(sv_, &st));
        py::dict d;
        d["submitted"] = st.submitted;
        d["completed"] = st.completed;
        d["batches"] = st.batches;
        d["slots"] = st.slots;
        d["scratch_bytes"] = st.scratch_bytes;
        d["gpu_ms"] = st.gpu_ms;
        d["host_ms"] = st.host_ms;
        d["turnaround_ms"] = st.turnaround_ms;
        d["summary_capacity_bytes"] = st.summary_capacity_bytes;
        d["admitted_summary_bytes"] = st.admitted_summary_bytes;
        d["footprint_limited_batches"] = st.footprint_limited_batches;
        return d;
    }
    // Scan plans per batch size from a sample of queries with their k and nprobe (the server must be
    // idle). Returns [[largest batch, meso block, data block, two-phase meso 0/1, scan order], ...].
    py::list Autotune(const py::array_t<float, py::array::c_style | py::array::forcecast>& q,
                      const std::vector<uint32_t>& k, const std::vector<uint32_t>& nprobe) {
        live();
        auto b = q.request();
        if (b.ndim != 2 || (uint32_t)b.shape[1] != d_) throw std::invalid_argument("queries must be n x d");
        const uint32_t n = (uint32_t)b.shape[0];
        if (k.size() != n || nprobe.size() != n) throw std::invalid_argument("k and nprobe need one entry per query");
        {
            py::gil_scoped_release nogil;
            PDX::detail::check(pdxg_server_autotune(sv_, static_cast<const float*>(b.ptr), k.data(), nprobe.data(), n));
        }
        std::vector<uint32_t> rows(5 * 16);
        uint32_t nr = 0;
        PDX::detail::check(pdxg_server_get_plans(sv_, rows.data(), 16, &nr));
        py::list out;
        for (uint32_t r = 0; r < nr; ++r)
            out.append(std::vector<uint32_t>(rows.begin() + 5 * r, rows.begin() + 5 * r + 5));
        return out;
    }

  private:
    PyPDXIndex* owner_;
    uint32_t d_, max_k_;
    pdxg_server* sv_ = nullptr;
};

}  // namespace

PYBIND11_MODULE(compiled, m) {
    m.doc() = "PDX search on MI355X: the pdxearch index API over hand-written HIP kernels";
    py::class_<PyPDXIndex>(m, "PDXIndex")
        .def(py::init<const std::string&, uint32_t, uint8_t, uint32_t, uint32_t, uint32_t, bool, float, uint32_t, bool,
                      uint32_t>(),
             py::arg("index_type"), py::arg("num_dimensions"), py::arg("distance_metric") = 0, py::arg("seed") = 42,
             py::arg("num_clusters") = 0, py::arg("num_meso_clusters") = 0, py::arg("normalize") = false,
             py::arg("sampling_fraction") = 0.0f, py::arg("kmeans_iters") = 10, py::arg("hierarchical_indexing") = true,
             py::arg("n_threads") = 0)
        .def("build_index", &PyPDXIndex::BuildIndex, py::arg("data"))
        .def("search", &PyPDXIndex::Search, py::arg("query"), py::arg("knn"))
        .def("filtered_search", &PyPDXIndex::FilteredSearch, py::arg("query"), py::arg("knn"), py::arg("row_ids"))
        .def("search_batch", &PyPDXIndex::SearchBatch, py::arg("queries"), py::arg("knn"),
             py::arg("nprobe") = py::none())
        .def("set_nprobe", [](PyPDXIndex& s, uint32_t n) { s.index->SetNProbe(n); }, py::arg("nprobe"))
        .def("save", [](PyPDXIndex& s, const std::string& p) {
                s.require_no_server("save");
                s.index->Save(p);
            }, py::arg("path"))
        .def("get_num_dimensions", [](const PyPDXIndex& s) { return s.index->GetNumDimensions(); })
        .def("get_num_clusters", [](const PyPDXIndex& s) { return s.index->GetNumClusters(); })
        .def("get_cluster_size", [](const PyPDXIndex& s, uint32_t c) { return s.index->GetClusterSize(c); },
             py::arg("cluster_id"))
        .def("get_cluster_row_ids", &PyPDXIndex::GetClusterRowIds, py::arg("cluster_id"))
        .def("get_in_memory_size_in_bytes", [](const PyPDXIndex& s) { return s.index->GetInMemorySizeInBytes(); })
        .def("get_device_size_in_bytes", &PyPDXIndex::DeviceBytes)
        .def("append", &PyPDXIndex::Append, py::arg("row_id"), py::arg("embedding"))
        .def("delete", [](PyPDXIndex& s, size_t r) {
                s.require_no_server("delete");
                s.index->Delete(r);
            }, py::arg("row_id"));
    py::class_<PyServer>(m, "Server")
        .def(py::init<PyPDXIndex&, uint32_t, uint32_t, uint32_t, uint32_t, size_t, uint32_t>(), py::arg("index"),
             py::arg("max_batch") = 1024, py::arg("max_k") = 100, py::arg("max_nprobe") = 0, py::arg("max_slots") = 2,
             py::arg("scratch_budget_bytes") = 0, py::arg("pipeline_min") = 0, py::keep_alive<1, 2>())
        .def("submit", &PyServer::Submit, py::arg("query"), py::arg("knn"), py::arg("nprobe"), py::arg("tag"))
        .def("poll", &PyServer::Poll, py::arg("max_results") = 1024, py::arg("timeout_ms") = 100)
        .def("stats", &PyServer::Stats)
        .def("autotune", &PyServer::Autotune, py::arg("queries"), py::arg("knn"), py::arg("nprobe"))
        .def("close", &PyServer::close);
    m.def("load_index", &PyPDXIndex::Load, py::arg("path"), "Load a PDX index from a single file (auto-detects type).");
}
