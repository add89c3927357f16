// Host-side index API of the MI355X PDX path: the reference's C++ interface (include/pdx/index.hpp:
// PDXIndexConfig, IPDXIndex, PDXIndex<Q>, PDXTreeIndex<Q>, LoadPDXIndex) with the same names,
// argument meaning and exceptions, keeping clusters in host memory in the reference layout and a
// device-resident copy behind the C-ABI of include/pdx_amd.h. Every search runs on the GPU; if the
// device library or the GPU is unavailable, Search throws.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "pdx_amd.h"
#include "pdx_host_util.hpp"

namespace PDX {

static constexpr size_t PDX_MAX_DIMS = 16384;     // common.hpp:44
static constexpr uint32_t H_DIM_SIZE = 64;         // common.hpp:45
static constexpr uint32_t U8_INTERLEAVE_SIZE = 4;  // common.hpp:46

enum class DistanceMetric { L2SQ, COSINE, IP };  // common.hpp:74
enum Quantization { F32, U8 };                   // common.hpp:76 (F16/BF are unused there)
enum class PDXIndexType : uint8_t { PDX_F32 = 0, PDX_U8 = 1, PDX_TREE_F32 = 2, PDX_TREE_U8 = 3 };

struct KNNCandidate {  // common.hpp:115-118
    uint32_t index;
    float distance;
};

struct PDXDimensionSplit {
    uint32_t horizontal_dimensions;
    uint32_t vertical_dimensions;
};

// common.hpp:133-159
inline PDXDimensionSplit GetPDXDimensionSplit(uint32_t d) {
    float prop = d <= 128 ? 0.25f : 0.75f;
    uint32_t h = static_cast<uint32_t>(static_cast<float>(d) * prop);
    uint32_t v = d - h;
    if (h % H_DIM_SIZE > 0) {
        h = ((h + H_DIM_SIZE / 2) / H_DIM_SIZE) * H_DIM_SIZE;
        v = d - h;
    }
    if (!v) {
        h = H_DIM_SIZE;
        v = d - h;
    }
    if (d <= H_DIM_SIZE) {
        h = 0;
        v = d;
    }
    return {h, v};
}

// common.hpp:191-199
inline uint32_t ComputeNumberOfClusters(uint32_t n) {
    if (n < 500000) return static_cast<uint32_t>(std::ceil(2 * std::sqrt((double)n)));
    if (n < 2500000) return static_cast<uint32_t>(std::ceil(4 * std::sqrt((double)n)));
    return static_cast<uint32_t>(std::ceil(8 * std::sqrt((double)n)));
}

inline bool DistanceMetricRequiresNormalization(DistanceMetric m) {
    return m == DistanceMetric::COSINE || m == DistanceMetric::IP;
}

// index.hpp:27-73 (same fields, defaults and messages)
struct PDXIndexConfig {
    uint32_t num_dimensions;
    DistanceMetric distance_metric = DistanceMetric::L2SQ;
    uint32_t seed = 42;
    uint32_t num_clusters = 0;
    uint32_t num_meso_clusters = 0;
    bool normalize = false;
    float sampling_fraction = 0.0f;
    uint32_t kmeans_iters = 10;
    bool hierarchical_indexing = true;
    uint32_t n_threads = 0;

    void Validate() const {
        if (num_dimensions == 0 || num_dimensions > PDX_MAX_DIMS)
            throw std::invalid_argument("num_dimensions must be between 1 and " + std::to_string(PDX_MAX_DIMS) +
                                        ", got " + std::to_string(num_dimensions));
        if (sampling_fraction < 0.0f || sampling_fraction > 1.0f)
            throw std::invalid_argument("sampling_fraction must be between 0.0 and 1.0, got " +
                                        std::to_string(sampling_fraction));
        if (num_meso_clusters > 0 && num_clusters > 0 && num_meso_clusters >= num_clusters)
            throw std::invalid_argument("num_meso_clusters (" + std::to_string(num_meso_clusters) +
                                        ") must be smaller than num_clusters (" + std::to_string(num_clusters) + ")");
        if (kmeans_iters == 0 || kmeans_iters >= 100)
            throw std::invalid_argument("kmeans_iters must be between 1 and 99, got " + std::to_string(kmeans_iters));
    }
    void ValidateNumEmbeddings(size_t n) const {
        if (num_clusters > 0 && num_clusters > n)
            throw std::invalid_argument("num_clusters (" + std::to_string(num_clusters) + ") exceeds num_embeddings (" +
                                        std::to_string(n) + ")");
    }
};

// ---------------------------------------------------------------------------------------------
// Host storage in the reference layout
// ---------------------------------------------------------------------------------------------
template <Quantization Q>
struct Cluster {  // cluster.hpp:17-427
    using data_t = typename std::conditional<Q == U8, uint8_t, float>::type;
    static constexpr float CAPACITY_THRESHOLD = 1.3f;
    static constexpr float MIN_CAPACITY_THRESHOLD = 0.5f;
    static constexpr uint32_t MIN_MAX_CAPACITY = 256;

    uint32_t num_embeddings = 0;
    uint32_t used_capacity = 0;
    uint32_t max_capacity = 0;
    uint32_t min_capacity = 0;
    uint32_t num_dimensions = 0;
    uint32_t id = 0;
    uint32_t mesocluster_id = 0;
    std::vector<uint32_t> indices;  // row id per slot
    std::vector<data_t> data;       // max_capacity * d, PDX layout
    std::set<uint32_t> tombstones;  // deleted slots (ordered: the lowest slot is reused first)

    Cluster() = default;
    Cluster(uint32_t n, uint32_t d)
        : Cluster(n, std::max(static_cast<uint32_t>(n * CAPACITY_THRESHOLD), MIN_MAX_CAPACITY), d) {}
    Cluster(uint32_t n, uint32_t max_cap, uint32_t d)
        : num_embeddings(n), used_capacity(n), max_capacity(max_cap),
          min_capacity(static_cast<uint32_t>(n * MIN_CAPACITY_THRESHOLD)), num_dimensions(d), indices(max_cap, 0),
          data((size_t)max_cap * d, data_t(0)) {}

    // Element ranges of the PDX layout: (first dim, width) in storage order (layout.hpp:21-87).
    std::vector<std::pair<uint32_t, uint32_t>> Regions() const {
        auto s = GetPDXDimensionSplit(num_dimensions);
        std::vector<std::pair<uint32_t, uint32_t>> r;
        if (Q == F32) {
            for (uint32_t j = 0; j < s.vertical_dimensions; ++j) r.push_back({j, 1});
        } else {
            uint32_t j = 0;
            for (; j + U8_INTERLEAVE_SIZE <= s.vertical_dimensions; j += U8_INTERLEAVE_SIZE) r.push_back({j, U8_INTERLEAVE_SIZE});
            if (j < s.vertical_dimensions) r.push_back({j, s.vertical_dimensions - j});
        }
        for (uint32_t h = 0; h < s.horizontal_dimensions; h += H_DIM_SIZE) r.push_back({s.vertical_dimensions + h, H_DIM_SIZE});
        return r;
    }

    void InsertEmbedding(uint32_t slot, uint32_t row_id, const data_t* emb) {  // cluster.hpp:341-384
        for (auto [start, width] : Regions())
            std::memcpy(&data[(size_t)start * max_capacity + (size_t)slot * width], emb + start, width * sizeof(data_t));
        indices[slot] = row_id;
    }
    void ReadEmbedding(uint32_t slot, data_t* out) const {  // cluster.hpp:299-337
        for (auto [start, width] : Regions())
            std::memcpy(out + start, &data[(size_t)start * max_capacity + (size_t)slot * width], width * sizeof(data_t));
    }
    void CopySlot(uint32_t src, uint32_t dst) {
        std::vector<data_t> tmp(num_dimensions);
        ReadEmbedding(src, tmp.data());
        InsertEmbedding(dst, indices[src], tmp.data());
    }
    uint32_t AppendEmbedding(uint32_t row_id, const data_t* emb) {  // cluster.hpp:121-146
        uint32_t slot = used_capacity;
        bool reused = false;
        if (!tombstones.empty()) {
            slot = *tombstones.begin();
            tombstones.erase(tombstones.begin());
            reused = true;
        }
        if (slot >= max_capacity)
            throw std::runtime_error("AppendEmbedding: cluster buffer overflow (used_capacity=" +
                                     std::to_string(used_capacity) + ", max_capacity=" + std::to_string(max_capacity) + ")");
        InsertEmbedding(slot, row_id, emb);
        num_embeddings++;
        if (!reused) used_capacity++;
        return slot;
    }
    void DeleteEmbedding(uint32_t slot) {  // cluster.hpp:148-154
        tombstones.insert(slot);
        num_embeddings--;
    }
    // cluster.hpp:260-294: fill holes from the tail; returns (row_id, new slot) of moved rows.
    std::vector<std::pair<uint32_t, uint32_t>> CompactCluster() {
        std::vector<std::pair<uint32_t, uint32_t>> moves;
        auto drain_tail = [&] {
            while (used_capacity > 0 && tombstones.count(used_capacity - 1)) {
                tombstones.erase(used_capacity - 1);
                indices[used_capacity - 1] = 0;
                used_capacity--;
            }
        };
        drain_tail();
        while (!tombstones.empty()) {
            uint32_t hole = *tombstones.begin();
            tombstones.erase(tombstones.begin());
            uint32_t last = used_capacity - 1;
            CopySlot(last, hole);
            moves.emplace_back(indices[hole], hole);
            indices[last] = 0;
            used_capacity--;
            drain_tail();
        }
        return moves;
    }
    size_t GetInMemorySizeInBytes() const {
        return sizeof(*this) + num_embeddings * sizeof(uint32_t) + (size_t)num_embeddings * num_dimensions * sizeof(data_t);
    }
    void SaveCompact(std::ostream& out) const {  // cluster.hpp:185-221
        for (auto [start, width] : Regions())
            out.write(reinterpret_cast<const char*>(&data[(size_t)start * max_capacity]),
                      (std::streamsize)((size_t)width * num_embeddings * sizeof(data_t)));
    }
    void LoadCompact(const char*& p, const char* end) {  // cluster.hpp:225-255
        for (auto [start, width] : Regions()) {
            size_t bytes = (size_t)width * num_embeddings * sizeof(data_t);
            if (p + bytes > end) throw std::runtime_error("index file truncated");
            std::memcpy(&data[(size_t)start * max_capacity], p, bytes);
            p += bytes;
        }
    }
};

template <Quantization Q>
struct IVF {  // ivf_wrapper.hpp:15-179
    uint32_t num_dimensions = 0;
    uint64_t total_num_embeddings = 0;
    uint32_t num_clusters = 0;
    uint32_t num_vertical_dimensions = 0;
    uint32_t num_horizontal_dimensions = 0;
    std::vector<Cluster<Q>> clusters;
    std::vector<size_t> cluster_offsets;
    size_t total_capacity = 0;
    size_t max_cluster_capacity = 0;
    bool is_normalized = false;
    std::vector<float> centroids;
    float quantization_scale = 1.0f, quantization_scale_squared = 1.0f, inverse_quantization_scale_squared = 1.0f,
          quantization_base = 0.0f;

    void SetDims(uint32_t d) {
        num_dimensions = d;
        auto s = GetPDXDimensionSplit(d);
        num_vertical_dimensions = s.vertical_dimensions;
        num_horizontal_dimensions = s.horizontal_dimensions;
    }
    void SetQuantization(float base, float scale) {
        quantization_base = base;
        quantization_scale = scale;
        quantization_scale_squared = scale * scale;
        inverse_quantization_scale_squared = 1.0f / (scale * scale);
    }
    void ComputeClusterOffsets() {
        cluster_offsets.assign(num_clusters, 0);
        total_capacity = max_cluster_capacity = 0;
        for (uint32_t i = 0; i < num_clusters; ++i) {
            cluster_offsets[i] = total_capacity;
            total_capacity += clusters[i].max_capacity;
            max_cluster_capacity = std::max<size_t>(max_cluster_capacity, clusters[i].max_capacity);
        }
    }
    void SaveClusters(std::ostream& out) const {
        for (auto& c : clusters) {
            out.write(reinterpret_cast<const char*>(&c.num_embeddings), 4);
            out.write(reinterpret_cast<const char*>(&c.max_capacity), 4);
        }
        for (auto& c : clusters) c.SaveCompact(out);
        for (auto& c : clusters) out.write(reinterpret_cast<const char*>(c.indices.data()), 4 * (std::streamsize)c.num_embeddings);
    }
    void LoadClusters(const char*& p, const char* end, uint32_t nc) {
        if (p + 8ull * nc > end) throw std::runtime_error("index file truncated");
        std::vector<uint32_t> hdr(2 * (size_t)nc);
        std::memcpy(hdr.data(), p, 8ull * nc);
        p += 8ull * nc;
        num_clusters = nc;
        clusters.clear();
        clusters.reserve(nc);
        for (uint32_t i = 0; i < nc; ++i) {
            if (hdr[2 * i] > hdr[2 * i + 1]) throw std::runtime_error("cluster size exceeds capacity");
            clusters.emplace_back(hdr[2 * i], hdr[2 * i + 1], num_dimensions);
            clusters[i].id = i;
            clusters[i].LoadCompact(p, end);
        }
        for (auto& c : clusters) {
            size_t b = 4ull * c.num_embeddings;
            if (p + b > end) throw std::runtime_error("index file truncated");
            std::memcpy(c.indices.data(), p, b);
            p += b;
        }
    }
};

template <Quantization Q>
struct IVFTree : IVF<Q> {  // ivf_wrapper.hpp:181-401
    IVF<F32> l0;
};

// ---------------------------------------------------------------------------------------------
// Interface (index.hpp:165-189) plus batched GPU entry points
// ---------------------------------------------------------------------------------------------
class IPDXIndex {
  public:
    virtual ~IPDXIndex() = default;
    virtual std::vector<KNNCandidate> Search(const float* query_embedding, size_t knn) const = 0;
    virtual std::vector<KNNCandidate> FilteredSearch(const float* query_embedding, size_t knn,
                                                     const std::vector<size_t>& passing_row_ids) const = 0;
    virtual void BuildIndex(const float* embeddings, size_t num_embeddings) = 0;
    virtual void SetNProbe(uint32_t n_probe) const = 0;
    virtual void Save(const std::string& path) = 0;
    virtual void Restore(const std::string& path) = 0;
    virtual uint32_t GetNumDimensions() const = 0;
    virtual uint32_t GetNumClusters() const = 0;
    virtual uint32_t GetClusterSize(uint32_t cluster_id) const = 0;
    virtual std::vector<uint32_t> GetClusterRowIds(uint32_t cluster_id) const = 0;
    virtual size_t GetInMemorySizeInBytes() const = 0;
    virtual void Append(size_t, const float*) {
        throw std::runtime_error("Append is not supported by this index type. Use PDXTreeIndex.");
    }
    virtual void Delete(size_t) {
        throw std::runtime_error("Delete is not supported by this index type. Use PDXTreeIndex.");
    }
    // Batched search of nq queries (row-major nq x d). k / nprobe: per-query arrays (or nullptr:
    // knn / the current nprobe for every query). Results: out_ids / out_dists rows of `stride`,
    // out_counts per query. Not in the reference: the GPU path's batch entry point.
    virtual void SearchBatch(const float* queries, size_t nq, size_t knn, const uint32_t* k, const uint32_t* nprobe,
                             uint32_t stride, uint32_t* out_ids, float* out_dists, uint32_t* out_counts,
                             uint64_t* out_bytes = nullptr) const = 0;
    virtual pdxg_index* DeviceHandle() const = 0;
};

namespace detail {

inline void check(pdxg_status s) {
    if (s != PDXG_OK) {
        std::string msg = pdxg_last_error();
        if (s == PDXG_ERR_INVALID) throw std::invalid_argument(msg);
        throw std::runtime_error("GPU path: " + msg);
    }
}

inline int device_ordinal() {
    const char* e = std::getenv("PDX_AMD_DEVICE");
    return e ? std::atoi(e) : 0;
}

template <Quantization Q>
struct ClusterDescs {
    std::vector<pdxg_cluster_desc> descs;
    std::vector<std::vector<uint32_t>> tombs;
    void add(const Cluster<Q>& c) {
        tombs.emplace_back(c.tombstones.begin(), c.tombstones.end());
        pdxg_cluster_desc d{};
        d.num_embeddings = c.num_embeddings;
        d.used_capacity = c.used_capacity;
        d.max_capacity = c.max_capacity;
        d.data = c.data.data();
        d.ids = c.indices.data();
        d.tombstones = tombs.back().data();
        d.num_tombstones = (uint32_t)tombs.back().size();
        descs.push_back(d);
    }
};

// Device copy of one index; recreated after structural changes, patched per cluster otherwise.
class DeviceIndex {
  public:
    ~DeviceIndex() { reset(); }
    void reset() {
        if (h_) pdxg_index_destroy(h_);
        h_ = nullptr;
    }
    pdxg_index* get() const { return h_; }

    template <Quantization Q>
    void create(PDXIndexType type, const IVF<Q>& L, const IVF<F32>* l0, const std::vector<float>& rotation) {
        reset();
        ClusterDescs<Q> c1;
        for (auto& c : L.clusters) {
            c1.tombs.reserve(L.clusters.size());
            c1.add(c);
        }
        ClusterDescs<F32> c0;
        if (l0) {
            c0.tombs.reserve(l0->clusters.size());
            for (auto& c : l0->clusters) c0.add(c);
        }
        pdxg_index_desc d{};
        d.index_type = static_cast<uint32_t>(type);
        d.num_dimensions = L.num_dimensions;
        d.num_vertical_dimensions = L.num_vertical_dimensions;
        d.num_horizontal_dimensions = L.num_horizontal_dimensions;
        d.is_normalized = L.is_normalized;
        d.quantization_base = L.quantization_base;
        d.quantization_scale = L.quantization_scale;
        d.rotation = rotation.data();
        d.l1.num_clusters = L.num_clusters;
        d.l1.clusters = c1.descs.data();
        d.l1.centroids = L.centroids.size() == (size_t)L.num_clusters * L.num_dimensions ? L.centroids.data() : nullptr;
        if (l0) {
            d.l0.num_clusters = l0->num_clusters;
            d.l0.clusters = c0.descs.data();
            d.l0.centroids = l0->centroids.data();
        }
        check(pdxg_index_create(device_ordinal(), &d, &h_));
    }

    template <Quantization Q>
    void update(uint32_t level, uint32_t cluster_id, const Cluster<Q>& c) {
        if (!h_) return;
        ClusterDescs<Q> cd;
        cd.tombs.reserve(1);
        cd.add(c);
        check(pdxg_index_update_cluster(h_, level, cluster_id, &cd.descs[0]));
    }

  private:
    pdxg_index* h_ = nullptr;
};

inline std::vector<char> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

template <class T>
T take(const char*& p, const char* end) {
    if (p + sizeof(T) > end) throw std::runtime_error("index file truncated");
    T v;
    std::memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    return v;
}

// ScalarQuantizer::QuantizeEmbedding (quantizers/scalar.hpp:76-94)
inline void quantize(const float* x, float base, float scale, uint8_t* out, uint32_t d) {
    for (uint32_t i = 0; i < d; ++i) {
        int r = static_cast<int>(std::round((x[i] - base) * scale));
        out[i] = r > 255 ? 255 : (r < 0 ? 0 : static_cast<uint8_t>(r));
    }
}

}  // namespace detail

// Shared machinery of the flat and two-level indexes.
template <Quantization Q, bool TREE>
class PDXIndexBase : public IPDXIndex {
  public:
    using data_t = typename Cluster<Q>::data_t;
    using level_t = typename std::conditional<TREE, IVFTree<Q>, IVF<Q>>::type;
    static constexpr uint32_t DELETED_MARKER = std::numeric_limits<uint32_t>::max();

    PDXIndexBase() = default;
    explicit PDXIndexBase(PDXIndexConfig cfg) : config(cfg) {
        config.Validate();
        rotation = host::random_rotation(config.num_dimensions, config.seed);
    }

    static constexpr PDXIndexType IndexType() {
        if (TREE) return Q == F32 ? PDXIndexType::PDX_TREE_F32 : PDXIndexType::PDX_TREE_U8;
        return Q == F32 ? PDXIndexType::PDX_F32 : PDXIndexType::PDX_U8;
    }

    void SetNProbe(uint32_t n) const override { nprobe = n; }
    uint32_t GetNProbe() const { return nprobe; }
    uint32_t GetNumDimensions() const override { return index.num_dimensions; }
    uint32_t GetNumClusters() const override { return index.num_clusters; }
    uint32_t GetClusterSize(uint32_t c) const override { return index.clusters.at(c).num_embeddings; }
    std::vector<uint32_t> GetClusterRowIds(uint32_t c) const override {
        const auto& cl = index.clusters.at(c);
        std::vector<uint32_t> out;
        out.reserve(cl.num_embeddings);
        for (uint32_t i = 0; i < cl.used_capacity; ++i)
            if (!cl.tombstones.count(i)) out.push_back(cl.indices[i]);
        return out;
    }
    size_t GetInMemorySizeInBytes() const override {
        size_t s = sizeof(*this) + rotation.size() * sizeof(float) + index.num_dimensions * sizeof(float);
        for (auto& c : index.clusters) s += c.GetInMemorySizeInBytes();
        s += index.centroids.capacity() * sizeof(float);
        if constexpr (TREE) {
            for (auto& c : index.l0.clusters) s += c.GetInMemorySizeInBytes();
            s += index.l0.centroids.capacity() * sizeof(float);
        }
        s += row_id_cluster_mapping.capacity() * sizeof(std::pair<uint32_t, uint32_t>);
        return s;
    }
    const level_t& GetIndex() const { return index; }
    const std::vector<float>& GetRotation() const { return rotation; }

    pdxg_index* DeviceHandle() const override {
        EnsureDevice();
        return device.get();
    }

    std::vector<KNNCandidate> Search(const float* q, size_t knn) const override {
        if (knn == 0) throw std::invalid_argument("knn must be >= 1");
        uint32_t k = static_cast<uint32_t>(knn);
        std::vector<uint32_t> ids(k);
        std::vector<float> dists(k);
        uint32_t cnt = 0;
        SearchBatch(q, 1, knn, nullptr, nullptr, k, ids.data(), dists.data(), &cnt);
        std::vector<KNNCandidate> res(cnt);
        for (uint32_t i = 0; i < cnt; ++i) res[i] = {ids[i], dists[i]};
        return res;
    }

    void SearchBatch(const float* queries, size_t nq, size_t knn, const uint32_t* k, const uint32_t* np,
                     uint32_t stride, uint32_t* out_ids, float* out_dists, uint32_t* out_counts,
                     uint64_t* out_bytes = nullptr) const override {
        RequireBuilt();
        EnsureDevice();
        uint32_t kmax = static_cast<uint32_t>(knn);
        if (k)
            for (size_t i = 0; i < nq; ++i) kmax = std::max(kmax, k[i]);
        if (kmax == 0) throw std::invalid_argument("knn must be >= 1");
        if (stride < kmax) throw std::invalid_argument("stride must be >= the largest k");
        uint32_t npmax = nprobe;
        if (np) {
            npmax = 1;
            for (size_t i = 0; i < nq; ++i) npmax = np[i] == 0 ? 0 : (npmax == 0 ? 0 : std::max(npmax, np[i]));
        }
        pdxg_search_args a{};
        a.num_queries = static_cast<uint32_t>(nq);
        a.queries = queries;
        a.k = k;
        a.nprobe = np;
        a.k_all = static_cast<uint32_t>(knn);
        a.nprobe_all = nprobe;
        a.k_max = kmax;
        a.nprobe_max = npmax;
        a.out_stride = stride;
        a.out_ids = out_ids;
        a.out_distances = out_dists;
        a.out_counts = out_counts;
        a.out_bytes = out_bytes;
        detail::check(pdxg_search(device.get(), &a));
    }

    std::vector<KNNCandidate> FilteredSearch(const float* q, size_t knn,
                                             const std::vector<size_t>& passing_row_ids) const override {
        RequireBuilt();
        EnsureDevice();
        if (knn == 0) throw std::invalid_argument("knn must be >= 1");
        // PredicateEvaluator (index.hpp:436-446, 833-845)
        std::vector<uint8_t> sel(index.total_capacity, 0);
        for (size_t row : passing_row_ids) {
            if (row >= row_id_cluster_mapping.size()) continue;
            auto [c, slot] = row_id_cluster_mapping[row];
            if (c == DELETED_MARKER) continue;
            sel[index.cluster_offsets[c] + slot] = 1;
        }
        std::vector<uint32_t> ids(knn), idist(knn);
        std::vector<float> dists(knn);
        uint32_t cnt = 0;
        detail::check(pdxg_filtered_search(device.get(), q, 0, (uint32_t)knn, nprobe, sel.data(), ids.data(),
                                           dists.data(), idist.data(), &cnt));
        std::vector<KNNCandidate> res(cnt);
        for (uint32_t i = 0; i < cnt; ++i) res[i] = {ids[i], dists[i]};
        return res;
    }

    void BuildIndex(const float* embeddings, size_t n) override {
        std::vector<size_t> rows(n);
        std::iota(rows.begin(), rows.end(), size_t{0});
        BuildIndex(rows.data(), embeddings, n);
    }

    // index.hpp:329-394 / 646-757
    void BuildIndex(const size_t* row_ids, const float* embeddings, size_t n) {
        if (n == 0) throw std::invalid_argument("num_embeddings must be positive");
        config.ValidateNumEmbeddings(n);
        const uint32_t d = config.num_dimensions;
        uint32_t nlist = config.num_clusters ? config.num_clusters : ComputeNumberOfClusters((uint32_t)n);
        nlist = std::min<uint32_t>(nlist, (uint32_t)n);
        const bool norm = config.normalize || DistanceMetricRequiresNormalization(config.distance_metric);
        const unsigned th = host::num_threads(config.n_threads);
        // NormalizeAndRotate (index.hpp:75-99)
        std::vector<float> src;
        const float* rin = embeddings;
        if (norm) {
            src.resize(n * d);
            host::parallel_for(n, th, [&](size_t b, size_t e) {
                for (size_t i = b; i < e; ++i) host::normalize_row(embeddings + i * d, &src[i * d], d);
            });
            rin = src.data();
        }
        std::vector<float> pre(n * (size_t)d);
        host::matmul(rin, n, d, rotation.data(), d, pre.data(), th);
        src.clear();
        src.shrink_to_fit();
        index = level_t{};
        index.SetDims(d);
        index.total_num_embeddings = n;
        index.is_normalized = norm;
        if constexpr (Q == U8) {  // ComputeQuantizationParams (scalar.hpp:60-74)
            float lo = std::numeric_limits<float>::max(), hi = std::numeric_limits<float>::lowest();
            for (float v : pre) {
                lo = std::min(lo, v);
                hi = std::max(hi, v);
            }
            float range = hi - lo;
            index.SetQuantization(lo, range > 0 ? 255.0f / range : 1.0f);
        }
        auto km = host::kmeans(pre.data(), n, d, nlist, config.kmeans_iters, config.seed, config.sampling_fraction,
                               norm, th);
        index.num_clusters = nlist;
        index.centroids = std::move(km.centroids);
        Populate(index, km, pre.data(), row_ids);
        if constexpr (TREE) {
            uint32_t n0 = config.num_meso_clusters ? config.num_meso_clusters
                                                   : static_cast<uint32_t>(std::sqrt((double)nlist));
            n0 = std::max(1u, std::min(n0, nlist));
            auto k0 = host::kmeans(index.centroids.data(), nlist, d, n0, 10, config.seed, 1.0f, norm, th);
            index.l0 = IVF<F32>{};
            index.l0.SetDims(d);
            index.l0.total_num_embeddings = nlist;
            index.l0.is_normalized = norm;
            index.l0.num_clusters = n0;
            index.l0.centroids = std::move(k0.centroids);
            std::vector<size_t> l0_rows(nlist);
            std::iota(l0_rows.begin(), l0_rows.end(), size_t{0});
            Populate(index.l0, k0, index.centroids.data(), l0_rows.data());
            for (uint32_t mc = 0; mc < n0; ++mc)
                for (uint64_t l1 : k0.assignments[mc]) index.clusters[l1].mesocluster_id = mc;
        }
        BuildRowIdClusterMapping();
        device.reset();
        built = true;
    }

    // index.hpp:213-239 / 483-514
    void Save(const std::string& path) override {
        RequireBuilt();
        bool moved = false;
        for (uint32_t c = 0; c < index.num_clusters; ++c) {
            for (auto [row, slot] : index.clusters[c].CompactCluster()) {
                row_id_cluster_mapping[row] = {c, slot};
                moved = true;
            }
        }
        if constexpr (TREE)
            for (auto& c : index.l0.clusters) moved |= !c.CompactCluster().empty();
        if (moved) device.reset();
        std::ofstream out(path, std::ios::binary);
        if (!out) throw std::runtime_error("cannot write " + path);
        uint8_t t = static_cast<uint8_t>(IndexType());
        out.write(reinterpret_cast<const char*>(&t), 1);
        uint32_t d = index.num_dimensions;
        out.write(reinterpret_cast<const char*>(&d), 4);
        out.write(reinterpret_cast<const char*>(&d), 4);
        out.write(reinterpret_cast<const char*>(rotation.data()), (std::streamsize)(rotation.size() * 4));
        out.write(reinterpret_cast<const char*>(&index.num_dimensions), 4);
        out.write(reinterpret_cast<const char*>(&index.num_vertical_dimensions), 4);
        out.write(reinterpret_cast<const char*>(&index.num_horizontal_dimensions), 4);
        out.write(reinterpret_cast<const char*>(&index.num_clusters), 4);
        if constexpr (TREE) {
            out.write(reinterpret_cast<const char*>(&index.l0.num_clusters), 4);
            index.l0.SaveClusters(out);
        }
        index.SaveClusters(out);
        char norm = index.is_normalized;
        out.write(&norm, 1);
        if constexpr (TREE)
            out.write(reinterpret_cast<const char*>(index.l0.centroids.data()), (std::streamsize)(index.l0.centroids.size() * 4));
        else
            out.write(reinterpret_cast<const char*>(index.centroids.data()), (std::streamsize)(index.centroids.size() * 4));
        if constexpr (Q == U8) {
            out.write(reinterpret_cast<const char*>(&index.quantization_base), 4);
            out.write(reinterpret_cast<const char*>(&index.quantization_scale), 4);
        }
        if (!out) throw std::runtime_error("write failed: " + path);
    }

    // index.hpp:241-263 / 516-543
    void Restore(const std::string& path) override {
        auto buf = detail::read_file(path);
        const char* p = buf.data();
        const char* end = p + buf.size();
        uint8_t t = detail::take<uint8_t>(p, end);
        if (t != static_cast<uint8_t>(IndexType())) throw std::runtime_error("index type mismatch in " + path);
        uint32_t rows = detail::take<uint32_t>(p, end), cols = detail::take<uint32_t>(p, end);
        if ((size_t)rows * cols * 4 > (size_t)(end - p)) throw std::runtime_error("index file truncated");
        std::vector<float> m((size_t)rows * cols);
        std::memcpy(m.data(), p, m.size() * 4);
        p += m.size() * 4;
        index = level_t{};
        uint32_t d = detail::take<uint32_t>(p, end);
        index.SetDims(d);
        uint32_t v = detail::take<uint32_t>(p, end), h = detail::take<uint32_t>(p, end);
        if (v != index.num_vertical_dimensions || h != index.num_horizontal_dimensions)
            throw std::runtime_error("dimension split in file disagrees with GetPDXDimensionSplit");
        if (rows != d || cols != d)
            throw std::runtime_error("only a dense d x d rotation is supported (the DCT variant needs FFTW)");
        uint32_t n1 = detail::take<uint32_t>(p, end);
        if constexpr (TREE) {
            uint32_t n0 = detail::take<uint32_t>(p, end);
            index.l0.SetDims(d);
            index.l0.LoadClusters(p, end, n0);
        }
        index.LoadClusters(p, end, n1);
        bool norm = detail::take<uint8_t>(p, end) != 0;
        index.is_normalized = norm;
        if constexpr (TREE) {
            index.l0.is_normalized = norm;
            size_t cb = (size_t)index.l0.num_clusters * d * 4;
            if (p + cb > end) throw std::runtime_error("index file truncated");
            index.l0.centroids.resize((size_t)index.l0.num_clusters * d);
            std::memcpy(index.l0.centroids.data(), p, cb);
            p += cb;
            for (uint32_t mc = 0; mc < index.l0.num_clusters; ++mc) {
                auto& c = index.l0.clusters[mc];
                for (uint32_t s = 0; s < c.num_embeddings; ++s) {
                    if (c.indices[s] >= n1) throw std::runtime_error("meso-cluster member is not an L1 cluster");
                    index.clusters[c.indices[s]].mesocluster_id = mc;
                }
            }
            index.l0.ComputeClusterOffsets();
        } else {
            size_t cb = (size_t)n1 * d * 4;
            if (p + cb > end) throw std::runtime_error("index file truncated");
            index.centroids.resize((size_t)n1 * d);
            std::memcpy(index.centroids.data(), p, cb);
            p += cb;
        }
        if constexpr (Q == U8) {
            float base = detail::take<float>(p, end), scale = detail::take<float>(p, end);
            index.SetQuantization(base, scale);
        }
        index.ComputeClusterOffsets();
        uint64_t tot = 0;
        for (auto& c : index.clusters) tot += c.num_embeddings;
        index.total_num_embeddings = tot;
        rotation = std::move(m);
        config.num_dimensions = d;
        BuildRowIdClusterMapping();
        device.reset();
        built = true;
    }

    // Row location (cluster, slot) of a row id; (DELETED_MARKER, DELETED_MARKER) once deleted.
    std::pair<uint32_t, uint32_t> GetRowLocation(size_t row_id) const { return row_id_cluster_mapping.at(row_id); }

  protected:
    PDXIndexConfig config{};
    level_t index;
    std::vector<float> rotation;
    mutable uint32_t nprobe = 0;
    mutable detail::DeviceIndex device;
    std::vector<std::pair<uint32_t, uint32_t>> row_id_cluster_mapping;
    bool built = false;

    void RequireBuilt() const {
        if (!built) throw std::runtime_error("index is empty: call BuildIndex or Restore first");
    }

    void EnsureDevice() const {
        if (device.get()) return;
        if constexpr (TREE) device.create<Q>(IndexType(), index, &index.l0, rotation);
        else device.create<Q>(IndexType(), index, nullptr, rotation);
    }

    // PopulateIVFClusters (index.hpp:101-163): members in assignment order, quantised for U8.
    template <Quantization QL>
    void Populate(IVF<QL>& L, const host::KMeansResult& km, const float* src, const size_t* row_ids) {
        const uint32_t d = L.num_dimensions;
        L.clusters.clear();
        L.clusters.reserve(L.num_clusters);
        for (uint32_t c = 0; c < L.num_clusters; ++c) {
            L.clusters.emplace_back((uint32_t)km.assignments[c].size(), d);
            L.clusters[c].id = c;
        }
        host::parallel_for(L.num_clusters, host::num_threads(config.n_threads), [&](size_t b, size_t e) {
            std::vector<typename Cluster<QL>::data_t> tmp(d);
            for (size_t c = b; c < e; ++c) {
                auto& cl = L.clusters[c];
                uint32_t pos = 0;
                for (uint64_t r : km.assignments[c]) {
                    if constexpr (QL == U8) detail::quantize(src + r * d, L.quantization_base, L.quantization_scale, tmp.data(), d);
                    else std::memcpy(tmp.data(), src + r * d, d * 4);
                    cl.InsertEmbedding(pos++, static_cast<uint32_t>(row_ids[r]), tmp.data());
                }
            }
        });
        L.ComputeClusterOffsets();
    }

    void BuildRowIdClusterMapping() {  // index.hpp:416-430
        size_t total = 0;
        for (auto& c : index.clusters) total += c.num_embeddings;
        row_id_cluster_mapping.assign(total, {DELETED_MARKER, DELETED_MARKER});
        for (uint32_t c = 0; c < index.num_clusters; ++c) {
            const auto& cl = index.clusters[c];
            for (uint32_t s = 0; s < cl.used_capacity; ++s) {
                if (cl.tombstones.count(s)) continue;
                uint32_t row = cl.indices[s];
                if (row >= row_id_cluster_mapping.size()) row_id_cluster_mapping.resize(row + 1, {DELETED_MARKER, DELETED_MARKER});
                row_id_cluster_mapping[row] = {c, s};
            }
        }
    }
};

template <Quantization Q>
class PDXIndex : public PDXIndexBase<Q, false> {  // index.hpp:191-447
  public:
    using PDXIndexBase<Q, false>::PDXIndexBase;
    void Append(size_t, const float*) override {
        throw std::runtime_error("Append is not implemented in PDXIndex. Use PDXTreeIndex instead.");
    }
    void Delete(size_t) override {
        throw std::runtime_error("Delete is not implemented in PDXIndex. Use PDXTreeIndex instead.");
    }
};

template <Quantization Q>
class PDXTreeIndex : public PDXIndexBase<Q, true> {  // index.hpp:449-1729
    using Base = PDXIndexBase<Q, true>;
    using Base::config;
    using Base::device;
    using Base::index;
    using Base::row_id_cluster_mapping;
    using Base::DELETED_MARKER;

  public:
    using Base::Base;

    uint32_t GetTopLevelNumClusters() const { return index.l0.num_clusters; }

    // index.hpp:581-616. The nearest L1 cluster comes from the meso-level search with
    // max(1, n_meso / 8) probes and k = 1 on the transformed vector.
    void Append(size_t row_id, const float* embedding) override {
        this->RequireBuilt();
        if (row_id != row_id_cluster_mapping.size())
            throw std::invalid_argument("Append: row_id " + std::to_string(row_id) + " is not sequential (expected " +
                                        std::to_string(row_id_cluster_mapping.size()) + ")");
        this->EnsureDevice();
        const uint32_t d = index.num_dimensions;
        std::vector<float> rot(d);
        std::vector<uint8_t> qu(d);
        detail::check(pdxg_preprocess(device.get(), embedding, 1, 0, rot.data(), Q == U8 ? qu.data() : nullptr));
        uint32_t nprobe0 = std::max(1u, GetTopLevelNumClusters() / 8);
        uint32_t cid = 0, cnt = 0;
        float dist = 0;
        detail::check(pdxg_search_meso(device.get(), rot.data(), 1, 1, nprobe0, &cid, &dist, &cnt));
        if (cnt == 0) throw std::runtime_error("Append: no candidate cluster found");
        auto& cl = index.clusters[cid];
        uint32_t slot;
        if constexpr (Q == U8) slot = cl.AppendEmbedding((uint32_t)row_id, qu.data());
        else slot = cl.AppendEmbedding((uint32_t)row_id, rot.data());
        row_id_cluster_mapping.emplace_back(cid, slot);
        index.total_num_embeddings++;
        CheckClusterHealth(cid);
    }

    // index.hpp:619-638
    void Delete(size_t row_id) override {
        this->RequireBuilt();
        if (row_id >= row_id_cluster_mapping.size())
            throw std::invalid_argument("Delete: row_id " + std::to_string(row_id) + " is not in the index");
        auto [cid, slot] = row_id_cluster_mapping[row_id];
        if (cid == DELETED_MARKER)
            throw std::invalid_argument("Delete: row_id " + std::to_string(row_id) + " was already deleted");
        index.clusters[cid].DeleteEmbedding(slot);
        row_id_cluster_mapping[row_id] = {DELETED_MARKER, DELETED_MARKER};
        index.total_num_embeddings--;
        CheckClusterHealth(cid);
    }

  private:
    // index.hpp:1241-1255. A full cluster is compacted when it holds tombstones; a full cluster
    // without tombstones gets a larger buffer (the reference splits it with k-means, which this
    // build does not do). Under-full clusters are left in place (the reference merges them).
    void CheckClusterHealth(uint32_t cid) {
        auto& cl = index.clusters[cid];
        if (cl.used_capacity == cl.max_capacity) {
            if (cl.num_embeddings < cl.used_capacity) {
                for (auto [row, s] : cl.CompactCluster()) row_id_cluster_mapping[row] = {cid, s};
            } else {
                Grow(cl);
                index.ComputeClusterOffsets();
                device.reset();  // capacities changed: re-upload on next use
                return;
            }
        }
        device.template update<Q>(1, cid, cl);
    }

    void Grow(Cluster<Q>& cl) {
        uint32_t ncap = std::max(static_cast<uint32_t>(cl.max_capacity * Cluster<Q>::CAPACITY_THRESHOLD), cl.max_capacity + 1);
        Cluster<Q> g(cl.num_embeddings, ncap, cl.num_dimensions);
        g.used_capacity = cl.used_capacity;
        g.min_capacity = cl.min_capacity;
        g.id = cl.id;
        g.mesocluster_id = cl.mesocluster_id;
        std::vector<typename Cluster<Q>::data_t> tmp(cl.num_dimensions);
        for (uint32_t s = 0; s < cl.used_capacity; ++s) {
            cl.ReadEmbedding(s, tmp.data());
            g.InsertEmbedding(s, cl.indices[s], tmp.data());
        }
        g.tombstones = cl.tombstones;
        cl = std::move(g);
    }
};

using PDXIndexF32 = PDXIndex<F32>;
using PDXIndexU8 = PDXIndex<U8>;
using PDXTreeIndexF32 = PDXTreeIndex<F32>;
using PDXTreeIndexU8 = PDXTreeIndex<U8>;

// index.hpp:1736-1760
inline std::unique_ptr<IPDXIndex> LoadPDXIndex(const std::string& path) {
    uint8_t t;
    {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw std::runtime_error("cannot open " + path);
        f.read(reinterpret_cast<char*>(&t), 1);
        if (!f) throw std::runtime_error("index file truncated");
    }
    std::unique_ptr<IPDXIndex> idx;
    switch (static_cast<PDXIndexType>(t)) {
        case PDXIndexType::PDX_F32: idx = std::make_unique<PDXIndexF32>(); break;
        case PDXIndexType::PDX_U8: idx = std::make_unique<PDXIndexU8>(); break;
        case PDXIndexType::PDX_TREE_F32: idx = std::make_unique<PDXTreeIndexF32>(); break;
        case PDXIndexType::PDX_TREE_U8: idx = std::make_unique<PDXTreeIndexU8>(); break;
        default: throw std::runtime_error("Unknown PDX index type: " + std::to_string(static_cast<int>(t)));
    }
    idx->Restore(path);
    return idx;
}

}  // namespace PDX
