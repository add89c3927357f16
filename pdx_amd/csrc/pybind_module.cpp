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
        .def("save", [](PyPDXIndex& s, const std::string& p) { s.index->Save(p); }, py::arg("path"))
        .def("get_num_dimensions", [](const PyPDXIndex& s) { return s.index->GetNumDimensions(); })
        .def("get_num_clusters", [](const PyPDXIndex& s) { return s.index->GetNumClusters(); })
        .def("get_cluster_size", [](const PyPDXIndex& s, uint32_t c) { return s.index->GetClusterSize(c); },
             py::arg("cluster_id"))
        .def("get_cluster_row_ids", &PyPDXIndex::GetClusterRowIds, py::arg("cluster_id"))
        .def("get_in_memory_size_in_bytes", [](const PyPDXIndex& s) { return s.index->GetInMemorySizeInBytes(); })
        .def("get_device_size_in_bytes", &PyPDXIndex::DeviceBytes)
        .def("append", &PyPDXIndex::Append, py::arg("row_id"), py::arg("embedding"))
        .def("delete", [](PyPDXIndex& s, size_t r) { s.index->Delete(r); }, py::arg("row_id"));
    m.def("load_index", &PyPDXIndex::Load, py::arg("path"), "Load a PDX index from a single file (auto-detects type).");
}
