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

    // Calls f(first dim, width) for each element range of the PDX layout in storage order
    // (layout.hpp:21-87): F32 vertical dims one by one, U8 vertical dims in groups of 4 (and a
    // tail of v % 4), then the 64-dim horizontal blocks.
    template <class F>
    void ForEachRegion(F&& f) const {
        const auto s = GetPDXDimensionSplit(num_dimensions);
        const uint32_t v = s.vertical_dimensions;
        if (Q == F32) {
            for (uint32_t j = 0; j < v; ++j) f(j, 1u);
        } else {
            uint32_t j = 0;
            for (; j + U8_INTERLEAVE_SIZE <= v; j += U8_INTERLEAVE_SIZE) f(j, U8_INTERLEAVE_SIZE);
            if (j < v) f(j, v - j);
        }
        for (uint32_t h = 0; h < s.horizontal_dimensions; h += H_DIM_SIZE) f(v + h, H_DIM_SIZE);
    }

    void InsertEmbedding(uint32_t slot, uint32_t row_id, const data_t* emb) {  // cluster.hpp:341-384
        ForEachRegion([&](uint32_t start, uint32_t width) {
            std::memcpy(&data[(size_t)start * max_capacity + (size_t)slot * width], emb + start, width * sizeof(data_t));
        });
        indices[slot] = row_id;
    }
    void ReadEmbedding(uint32_t slot, data_t* out) const {  // cluster.hpp:299-337
        ForEachRegion([&](uint32_t start, uint32_t width) {
            std::memcpy(out + start, &data[(size_t)start * max_capacity + (size_t)slot * width], width * sizeof(data_t));
        });
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
        ForEachRegion([&](uint32_t start, uint32_t width) {
            out.write(reinterpret_cast<const char*>(&data[(size_t)start * max_capacity]),
                      (std::streamsize)((size_t)width * num_embeddings * sizeof(data_t)));
        });
    }
    void LoadCompact(const char*& p, const char* end) {  // cluster.hpp:225-255
        ForEachRegion([&](uint32_t start, uint32_t width) {
            size_t bytes = (size_t)width * num_embeddings * sizeof(data_t);
            if (p + bytes > end) throw std::runtime_error("index file truncated");
            std::memcpy(&data[(size_t)start * max_capacity], p, bytes);
            p += bytes;
        });
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

// ScalarQuantizer::DequantizeEmbedding (quantizers/scalar.hpp:96-106)
inline void dequantize(const uint8_t* q, float base, float scale, float* out, uint32_t d) {
    for (uint32_t i = 0; i < d; ++i) out[i] = static_cast<float>(q[i]) / scale + base;
}

// F32 L2 over d dims in the 16-lane order of the reference's Horizontal kernel
// (distance_computers/avx512_computers.hpp:47-79): lane l accumulates dims l, l+16, ... with fma,
// lanes then combine (i, i+8), (i, i+4), (i, i+1), (i, i+2).
inline float l2sq16(const float* a, const float* b, uint32_t d) {
    float acc[16] = {};
    for (uint32_t base = 0; base < d; base += 16) {
        const uint32_t lim = std::min<uint32_t>(16, d - base);
        for (uint32_t l = 0; l < lim; ++l) {
            const float df = a[base + l] - b[base + l];
            acc[l] = std::fmaf(df, df, acc[l]);
        }
    }
    float x[8], r[4];
    for (int i = 0; i < 8; ++i) x[i] = acc[i] + acc[i + 8];
    for (int i = 0; i < 4; ++i) r[i] = x[i] + x[i + 4];
    return (r[0] + r[1]) + (r[2] + r[3]);
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
        // Compaction changes a cluster whenever it holds tombstones, also when they all sit at its
        // tail and no row moves (used_capacity shrinks): the device copy is then re-created.
        bool changed = false;
        for (uint32_t c = 0; c < index.num_clusters; ++c) {
            changed |= !index.clusters[c].tombstones.empty();
            for (auto [row, slot] : index.clusters[c].CompactCluster()) row_id_cluster_mapping[row] = {c, slot};
        }
        if constexpr (TREE)
            for (auto& c : index.l0.clusters) {
                changed |= !c.tombstones.empty();
                c.CompactCluster();
            }
        if (changed) device.reset();
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
            // The file holds no L1 centroids (ivf_wrapper.hpp:314-380); each one is the vector its
            // meso cluster stores for that L1 cluster, so they are read back from there (the
            // maintenance's splits and merges and the data level's own probe need them).
            index.centroids.assign((size_t)n1 * d, 0.0f);
            for (uint32_t mc = 0; mc < index.l0.num_clusters; ++mc) {
                auto& c = index.l0.clusters[mc];
                for (uint32_t s = 0; s < c.num_embeddings; ++s) {
                    if (c.indices[s] >= n1) throw std::runtime_error("meso-cluster member is not an L1 cluster");
                    index.clusters[c.indices[s]].mesocluster_id = mc;
                    c.ReadEmbedding(s, &index.centroids[(size_t)c.indices[s] * d]);
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
    using data_t = typename Base::data_t;

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
        restructured_ = false;
        CheckClusterHealth(cid);
        SyncDevice(cid);
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
        restructured_ = false;
        CheckClusterHealth(cid);
        SyncDevice(cid);
    }

  private:
    using dist_t = typename std::conditional<Q == U8, uint32_t, float>::type;
    static constexpr size_t kMaxSplitNeighbors = 32;  // index.hpp:948 (max_neighbors)
    static constexpr uint32_t kSplitIters = 4;        // index.hpp:1036-1046, 1330-1342

    bool restructured_ = false;  // a split or merge renumbered clusters or changed capacities
    uint64_t l1_merges_ = 0;     // L1 merges so far (each renumbers the last cluster)

    // A cluster changed in place is patched on the device; after a split or merge the device copy
    // is re-created at the next search (cluster count, capacities and meso membership changed).
    void SyncDevice(uint32_t cid) {
        if (restructured_) {
            device.reset();
            restructured_ = false;
        } else {
            device.template update<Q>(1, cid, index.clusters[cid]);
        }
    }

    bool NormalizedCentroids() const {
        return config.normalize || DistanceMetricRequiresNormalization(config.distance_metric);
    }
    const float* Centroid(uint32_t cid) const { return &index.centroids[(size_t)cid * index.num_dimensions]; }

    // Row-major copies of the first n slots (the cluster has no tombstones below n).
    std::vector<data_t> RowMajor(const Cluster<Q>& c, uint32_t n) const {
        std::vector<data_t> out((size_t)n * c.num_dimensions);
        for (uint32_t s = 0; s < n; ++s) c.ReadEmbedding(s, &out[(size_t)s * c.num_dimensions]);
        return out;
    }
    // index.hpp:861-881
    std::vector<float> Dequantized(const data_t* raw, size_t n) const {
        const uint32_t d = index.num_dimensions;
        std::vector<float> out(n * d);
        if constexpr (Q == U8) {
            for (size_t i = 0; i < n; ++i)
                detail::dequantize(raw + i * d, index.quantization_base, index.quantization_scale, &out[i * d], d);
        } else if (n) {
            std::memcpy(out.data(), raw, n * d * sizeof(float));
        }
        return out;
    }
    // index.hpp:884-895
    uint32_t QuantizeAndAppend(uint32_t cid, uint32_t row_id, const float* emb) {
        auto& c = index.clusters[cid];
        if constexpr (Q == U8) {
            std::vector<uint8_t> qv(index.num_dimensions);
            detail::quantize(emb, index.quantization_base, index.quantization_scale, qv.data(), index.num_dimensions);
            return c.AppendEmbedding(row_id, qv.data());
        } else {
            return c.AppendEmbedding(row_id, emb);
        }
    }
    // index.hpp:920-942: mean of the accumulated sum (the k-means centroid when the group is
    // empty), normalised when the index is
    std::vector<float> CentroidMean(const std::vector<float>& sum, size_t count, const float* fallback) const {
        const uint32_t d = index.num_dimensions;
        std::vector<float> out(d);
        if (count == 0) {
            std::memcpy(out.data(), fallback, d * sizeof(float));
        } else {
            const float inv = 1.0f / static_cast<float>(count);
            for (uint32_t j = 0; j < d; ++j) out[j] = sum[j] * inv;
        }
        if (NormalizedCentroids()) host::normalize_row(out.data(), out.data(), d);
        return out;
    }
    // index.hpp:945-981: the other L1 clusters of the meso cluster, at most the 32 nearest
    std::vector<uint32_t> NeighborClusterIds(uint32_t cid, uint32_t meso, const float* centroid) const {
        std::vector<uint32_t> ids;
        const auto& m = index.l0.clusters[meso];
        for (uint32_t p = 0; p < m.used_capacity; ++p)
            if (!m.tombstones.count(p) && m.indices[p] != cid) ids.push_back(m.indices[p]);
        if (ids.size() > kMaxSplitNeighbors) {
            std::vector<std::pair<float, uint32_t>> nd;
            for (uint32_t n : ids) nd.push_back({detail::l2sq16(centroid, Centroid(n), index.num_dimensions), n});
            std::nth_element(nd.begin(), nd.begin() + kMaxSplitNeighbors, nd.end(),
                             [](const auto& a, const auto& b) { return a.first < b.first; });
            ids.clear();
            for (size_t i = 0; i < kMaxSplitNeighbors; ++i) ids.push_back(nd[i].second);
        }
        return ids;
    }
    // index.hpp:983-996
    uint32_t PositionInMeso(uint32_t cid, uint32_t meso) const {
        const auto& m = index.l0.clusters[meso];
        for (uint32_t p = 0; p < m.used_capacity; ++p)
            if (!m.tombstones.count(p) && m.indices[p] == cid) return p;
        throw std::runtime_error("FindPositionInMesoCluster: L1 cluster " + std::to_string(cid) +
                                 " not found in L0 meso-cluster " + std::to_string(meso));
    }
    uint32_t LiveMesoMembers(uint32_t meso) const { return index.l0.clusters[meso].num_embeddings; }
    // Index into `cand` of the nearest candidate (L2, ties to the first) for each of n rows
    // (the reference's batch_computer::FindNearestNeighbor).
    std::vector<uint32_t> Nearest(const float* embs, size_t n, const std::vector<const float*>& cand) const {
        const uint32_t d = index.num_dimensions;
        std::vector<uint32_t> out(n, 0);
        for (size_t i = 0; i < n; ++i) {
            float best = std::numeric_limits<float>::infinity();
            for (uint32_t c = 0; c < cand.size(); ++c) {
                const float v = detail::l2sq16(embs + i * d, cand[c], d);
                if (v < best) {
                    best = v;
                    out[i] = c;
                }
            }
        }
        return out;
    }
    // index.hpp:1691-1727: the distance of every used slot of a cluster to one vector, vertical
    // dims then the 64-dim horizontal blocks (exact integers for U8).
    std::vector<dist_t> DistancesToCluster(const data_t* query, const Cluster<Q>& c) const {
        const uint32_t d = index.num_dimensions, v = index.num_vertical_dimensions;
        std::vector<dist_t> out(c.used_capacity);
        std::vector<data_t> row(d);
        for (uint32_t s = 0; s < c.used_capacity; ++s) {
            c.ReadEmbedding(s, row.data());
            if constexpr (Q == U8) {
                uint32_t acc = 0;
                for (uint32_t j = 0; j < d; ++j) {
                    const int df = (int)query[j] - (int)row[j];
                    acc += (uint32_t)(df * df);
                }
                out[s] = acc;
            } else {
                float acc = 0.0f;
                for (uint32_t j = 0; j < v; ++j) {
                    const float df = query[j] - row[j];
                    acc = std::fmaf(df, df, acc);
                }
                for (uint32_t hb = v; hb < d; hb += H_DIM_SIZE) acc += detail::l2sq16(query + hb, row.data() + hb, H_DIM_SIZE);
                out[s] = acc;
            }
        }
        return out;
    }
    // A new L1 cluster holding `count` row-major embeddings (StoreClusterEmbeddings).
    Cluster<Q> NewCluster(uint32_t id, uint32_t meso, const std::vector<data_t>& embs, const std::vector<uint32_t>& ids) const {
        const uint32_t d = index.num_dimensions;
        Cluster<Q> c(static_cast<uint32_t>(ids.size()), d);
        c.id = id;
        c.mesocluster_id = meso;
        for (size_t i = 0; i < ids.size(); ++i) c.InsertEmbedding((uint32_t)i, ids[i], &embs[i * d]);
        return c;
    }

    // index.hpp:1241-1255
    void CheckClusterHealth(uint32_t cid, bool allow_merges = true) {
        auto& cl = index.clusters[cid];
        if (cl.used_capacity == cl.max_capacity) {
            if (cl.num_embeddings < cl.used_capacity) {  // compacting is cheaper than splitting
                for (auto [row, s] : cl.CompactCluster()) row_id_cluster_mapping[row] = {cid, s};
            } else {
                SplitCluster(cid);
            }
        } else if (allow_merges && cl.num_embeddings <= cl.min_capacity) {
            DestroyAndMergeCluster(cid);
        }
    }

    // index.hpp:1257-1311. A cluster that is the only live member of its meso cluster, or the only
    // cluster, is kept: its rows would have nowhere to go (the reference has no guard there).
    void DestroyAndMergeCluster(uint32_t cid) {
        const uint32_t d = index.num_dimensions;
        const uint32_t meso = index.clusters[cid].mesocluster_id;
        if (index.num_clusters <= 1 || LiveMesoMembers(meso) <= 1) return;
        restructured_ = true;
        l1_merges_++;
        auto& cl = index.clusters[cid];
        cl.CompactCluster();
        const uint32_t n = cl.num_embeddings;
        std::vector<uint32_t> rows(cl.indices.begin(), cl.indices.begin() + n);
        std::vector<data_t> raw = RowMajor(cl, n);
        std::vector<float> embs = Dequantized(raw.data(), n);
        index.l0.clusters[meso].DeleteEmbedding(PositionInMeso(cid, meso));
        index.l0.total_num_embeddings--;
        const uint32_t last = index.num_clusters - 1;
        if (cid != last) {  // swap-and-pop
            index.clusters[cid] = std::move(index.clusters[last]);
            auto& mv = index.clusters[cid];
            mv.id = cid;
            std::memcpy(&index.centroids[(size_t)cid * d], Centroid(last), d * sizeof(float));
            for (uint32_t s = 0; s < mv.used_capacity; ++s)
                if (!mv.tombstones.count(s)) row_id_cluster_mapping[mv.indices[s]] = {cid, s};
            auto& mm = index.l0.clusters[mv.mesocluster_id];
            mm.indices[PositionInMeso(last, mv.mesocluster_id)] = cid;
        }
        index.clusters.pop_back();
        index.centroids.resize((size_t)last * d);
        index.num_clusters--;
        index.ComputeClusterOffsets();
        index.l0.ComputeClusterOffsets();
        ReassignEmbeddings(rows.data(), embs.data(), n, meso, false);
        CheckL0ClusterHealth(meso);
    }

    // index.hpp:1314-1611 (the cluster is full and has no tombstones). A 2-means split of its
    // members; a member closer to another centroid of the meso cluster than to both halves is
    // reassigned; members of the (up to 32) neighbouring clusters that are not closer to their own
    // centroid than to both halves move to the nearer half.
    void SplitCluster(uint32_t cid) {
        restructured_ = true;
        const uint32_t d = index.num_dimensions;
        const uint32_t meso = index.clusters[cid].mesocluster_id;
        const uint32_t n = index.clusters[cid].num_embeddings;
        std::vector<data_t> raw = RowMajor(index.clusters[cid], n);
        std::vector<float> embs = Dequantized(raw.data(), n);
        std::vector<float> old_c(Centroid(cid), Centroid(cid) + d);
        std::vector<uint32_t> neighbors = NeighborClusterIds(cid, meso, old_c.data());
        // ComputeKMeans(..., 2 clusters, sampling 1.0, 4 iterations, 1 thread). The reference asks
        // for angular (spherical) k-means at every split; on vectors that are not normalised its
        // unit-norm centroids would lie far from the data and leave both halves empty, so angular
        // k-means is used when the index's vectors are normalised and plain Lloyd otherwise.
        auto km = host::kmeans(embs.data(), n, d, 2, kSplitIters, config.seed, 1.0f, NormalizedCentroids(), 1);
        const float* ca = km.centroids.data();
        const float* cb = ca + d;
        std::vector<uint32_t> ga, gb, rest;
        for (uint32_t i = 0; i < n; ++i) {
            const float* e = &embs[(size_t)i * d];
            const float dold = detail::l2sq16(e, old_c.data(), d);
            const float da = detail::l2sq16(e, ca, d), db = detail::l2sq16(e, cb, d);
            const float mn = std::min(da, db);
            bool elsewhere = false;
            if (mn > dold)
                for (uint32_t c : neighbors)
                    if (detail::l2sq16(e, Centroid(c), d) < mn) {
                        elsewhere = true;
                        break;
                    }
            if (elsewhere) rest.push_back(i);
            else (da <= db ? ga : gb).push_back(i);
        }
        std::vector<data_t> ea, eb;
        std::vector<uint32_t> ia, ib;
        std::vector<float> sa(d, 0.0f), sb(d, 0.0f);
        auto gather = [&](const std::vector<uint32_t>& g, std::vector<data_t>& eo, std::vector<uint32_t>& io,
                          std::vector<float>& sum) {
            for (uint32_t i : g) {
                eo.insert(eo.end(), &raw[(size_t)i * d], &raw[(size_t)(i + 1) * d]);
                io.push_back(index.clusters[cid].indices[i]);
                for (uint32_t j = 0; j < d; ++j) sum[j] += embs[(size_t)i * d + j];
            }
        };
        gather(ga, ea, ia, sa);
        gather(gb, eb, ib, sb);
        std::vector<float> rest_embs;
        std::vector<uint32_t> rest_ids;
        for (uint32_t i : rest) {
            rest_embs.insert(rest_embs.end(), &embs[(size_t)i * d], &embs[(size_t)(i + 1) * d]);
            rest_ids.push_back(index.clusters[cid].indices[i]);
        }
        // members of the neighbours that are nearer to a half than to their own centroid move
        std::vector<data_t> qa(d), qb(d), qo(d);
        if constexpr (Q == U8) {
            detail::quantize(ca, index.quantization_base, index.quantization_scale, qa.data(), d);
            detail::quantize(cb, index.quantization_base, index.quantization_scale, qb.data(), d);
        } else {
            std::memcpy(qa.data(), ca, d * sizeof(float));
            std::memcpy(qb.data(), cb, d * sizeof(float));
        }
        std::vector<data_t> row(d);
        std::vector<float> rowf(d);
        for (uint32_t nb : neighbors) {
            auto& nc = index.clusters[nb];
            if constexpr (Q == U8)
                detail::quantize(Centroid(nb), index.quantization_base, index.quantization_scale, qo.data(), d);
            else
                std::memcpy(qo.data(), Centroid(nb), d * sizeof(float));
            auto down = DistancesToCluster(qo.data(), nc);
            auto dta = DistancesToCluster(qa.data(), nc);
            auto dtb = DistancesToCluster(qb.data(), nc);
            for (uint32_t p = 0; p < nc.used_capacity; ++p) {
                if (nc.tombstones.count(p)) continue;
                if (down[p] < dta[p] && down[p] < dtb[p]) continue;
                nc.ReadEmbedding(p, row.data());
                if constexpr (Q == U8) detail::dequantize(row.data(), index.quantization_base, index.quantization_scale, rowf.data(), d);
                else std::memcpy(rowf.data(), row.data(), d * sizeof(float));
                const bool to_a = dta[p] <= dtb[p];
                const uint32_t rid = nc.indices[p];
                nc.DeleteEmbedding(p);
                (to_a ? ea : eb).insert((to_a ? ea : eb).end(), row.begin(), row.end());
                (to_a ? ia : ib).push_back(rid);
                auto& sum = to_a ? sa : sb;
                for (uint32_t j = 0; j < d; ++j) sum[j] += rowf[j];
            }
        }
        std::vector<float> ta = CentroidMean(sa, ia.size(), ca), tb = CentroidMean(sb, ib.size(), cb);
        const uint32_t bid = index.num_clusters;
        index.clusters[cid] = NewCluster(cid, meso, ea, ia);
        index.clusters.push_back(NewCluster(bid, meso, eb, ib));
        index.num_clusters++;
        std::memcpy(&index.centroids[(size_t)cid * d], ta.data(), d * sizeof(float));
        index.centroids.insert(index.centroids.end(), tb.begin(), tb.end());
        for (size_t i = 0; i < ia.size(); ++i) row_id_cluster_mapping[ia[i]] = {cid, (uint32_t)i};
        for (size_t i = 0; i < ib.size(); ++i) row_id_cluster_mapping[ib[i]] = {bid, (uint32_t)i};
        // the meso cluster: the old centroid out, both halves in. It may split in between; it is not
        // merged there (a merge would renumber the meso cluster the second half goes to), only by
        // the check at the end.
        {
            auto& m = index.l0.clusters[meso];
            m.DeleteEmbedding(PositionInMeso(cid, meso));
            m.CompactCluster();
            m.AppendEmbedding(cid, ta.data());
        }
        CheckL0ClusterHealth(meso, false);
        index.l0.clusters[meso].AppendEmbedding(bid, tb.data());
        index.l0.total_num_embeddings++;
        if (!rest_ids.empty()) ReassignEmbeddings(rest_ids.data(), rest_embs.data(), (uint32_t)rest_ids.size(), meso);
        index.ComputeClusterOffsets();
        index.l0.ComputeClusterOffsets();
        CheckL0ClusterHealth(meso);
    }

    // index.hpp:1616-1685: each row to the nearest live L1 cluster of the meso cluster.
    void ReassignEmbeddings(const uint32_t* rows, const float* embs, uint32_t n, uint32_t meso, bool allow_merges = true) {
        const uint32_t d = index.num_dimensions;
        auto candidates = [&] {
            std::vector<uint32_t> cand;
            const auto& m = index.l0.clusters[meso];
            for (uint32_t p = 0; p < m.used_capacity; ++p)
                if (!m.tombstones.count(p)) cand.push_back(m.indices[p]);
            if (cand.empty()) throw std::runtime_error("ReassignEmbeddings: meso-cluster has no live L1 cluster");
            return cand;
        };
        std::vector<uint32_t> cand;
        std::vector<uint32_t> a;
        uint64_t merges_seen = ~0ull;
        for (uint32_t i = 0; i < n; ++i) {
            // nearest candidates for the remaining rows; recomputed only when a merge triggered by
            // an earlier row renumbered clusters (splits keep every existing cluster id)
            if (merges_seen != l1_merges_) {
                cand = candidates();
                std::vector<const float*> cp;
                for (uint32_t c : cand) cp.push_back(Centroid(c));
                a = Nearest(embs + (size_t)i * d, n - i, cp);
                a.insert(a.begin(), i, 0u);
                merges_seen = l1_merges_;
            }
            const uint32_t best = cand[a[i]];
            const uint32_t pos = QuantizeAndAppend(best, rows[i], embs + (size_t)i * d);
            row_id_cluster_mapping[rows[i]] = {best, pos};
            CheckClusterHealth(best, allow_merges);
        }
    }

    // index.hpp:1004-1015
    void CheckL0ClusterHealth(uint32_t mid, bool allow_merges = true) {
        auto& m = index.l0.clusters[mid];
        if (m.used_capacity == m.max_capacity) {
            if (m.num_embeddings < m.used_capacity) m.CompactCluster();
            else SplitL0Cluster(mid);
        } else if (allow_merges && m.num_embeddings <= m.min_capacity) {
            DestroyAndMergeL0Cluster(mid);
        }
    }

    // index.hpp:1017-1137: 2-means of the member L1 centroids.
    void SplitL0Cluster(uint32_t mid) {
        restructured_ = true;
        const uint32_t d = index.num_dimensions;
        const uint32_t n = index.l0.clusters[mid].num_embeddings;
        std::vector<uint32_t> l1(index.l0.clusters[mid].indices.begin(), index.l0.clusters[mid].indices.begin() + n);
        std::vector<float> cents((size_t)n * d);
        for (uint32_t i = 0; i < n; ++i) std::memcpy(&cents[(size_t)i * d], Centroid(l1[i]), d * sizeof(float));
        auto km = host::kmeans(cents.data(), n, d, 2, kSplitIters, config.seed, 1.0f, NormalizedCentroids(), 1);
        std::vector<uint32_t> ids[2];
        std::vector<float> embs[2], sums[2] = {std::vector<float>(d, 0.0f), std::vector<float>(d, 0.0f)};
        for (int g = 0; g < 2; ++g)
            for (uint64_t i : km.assignments[g]) {
                ids[g].push_back(l1[i]);
                embs[g].insert(embs[g].end(), &cents[i * d], &cents[(i + 1) * d]);
                for (uint32_t j = 0; j < d; ++j) sums[g][j] += cents[i * d + j];
            }
        std::vector<float> ta = CentroidMean(sums[0], ids[0].size(), km.centroids.data());
        std::vector<float> tb = CentroidMean(sums[1], ids[1].size(), km.centroids.data() + d);
        const uint32_t bid = index.l0.num_clusters;
        auto make = [&](uint32_t id, int g) {
            Cluster<F32> c(static_cast<uint32_t>(ids[g].size()), d);
            c.id = id;
            for (size_t i = 0; i < ids[g].size(); ++i) c.InsertEmbedding((uint32_t)i, ids[g][i], &embs[g][i * d]);
            return c;
        };
        index.l0.clusters[mid] = make(mid, 0);
        index.l0.clusters.push_back(make(bid, 1));
        index.l0.num_clusters++;
        std::memcpy(&index.l0.centroids[(size_t)mid * d], ta.data(), d * sizeof(float));
        index.l0.centroids.insert(index.l0.centroids.end(), tb.begin(), tb.end());
        for (uint32_t c : ids[0]) index.clusters[c].mesocluster_id = mid;
        for (uint32_t c : ids[1]) index.clusters[c].mesocluster_id = bid;
        index.l0.ComputeClusterOffsets();
    }

    // index.hpp:1139-1190. The last meso cluster is kept (its members would have nowhere to go).
    void DestroyAndMergeL0Cluster(uint32_t mid) {
        if (index.l0.num_clusters <= 1) return;
        restructured_ = true;
        const uint32_t d = index.num_dimensions;
        auto& m = index.l0.clusters[mid];
        m.CompactCluster();
        const uint32_t n = m.num_embeddings;
        std::vector<uint32_t> l1(m.indices.begin(), m.indices.begin() + n);
        std::vector<float> cents((size_t)n * d);
        for (uint32_t i = 0; i < n; ++i) std::memcpy(&cents[(size_t)i * d], Centroid(l1[i]), d * sizeof(float));
        const uint32_t last = index.l0.num_clusters - 1;
        if (mid != last) {  // swap-and-pop
            index.l0.clusters[mid] = std::move(index.l0.clusters[last]);
            index.l0.clusters[mid].id = mid;
            std::memcpy(&index.l0.centroids[(size_t)mid * d], &index.l0.centroids[(size_t)last * d], d * sizeof(float));
            auto& mv = index.l0.clusters[mid];
            for (uint32_t p = 0; p < mv.used_capacity; ++p)
                if (!mv.tombstones.count(p)) index.clusters[mv.indices[p]].mesocluster_id = mid;
        }
        index.l0.clusters.pop_back();
        index.l0.centroids.resize((size_t)last * d);
        index.l0.num_clusters--;
        index.l0.total_num_embeddings -= n;
        ReassignEmbeddingsL0(l1.data(), cents.data(), n);
        index.l0.ComputeClusterOffsets();
    }

    // index.hpp:1192-1239: each L1 centroid to the nearest meso centroid.
    void ReassignEmbeddingsL0(const uint32_t* l1, const float* cents, uint32_t n) {
        const uint32_t d = index.num_dimensions;
        std::vector<const float*> cp;
        for (uint32_t c = 0; c < index.l0.num_clusters; ++c) cp.push_back(&index.l0.centroids[(size_t)c * d]);
        std::vector<uint32_t> a = Nearest(cents, n, cp);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t t = a[i];
            index.l0.clusters[t].AppendEmbedding(l1[i], cents + (size_t)i * d);
            index.l0.total_num_embeddings++;
            index.clusters[l1[i]].mesocluster_id = t;
            CheckL0ClusterHealth(t, false);
        }
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
