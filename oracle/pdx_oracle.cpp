// Oracle: a CPU restatement of the PDX search path, used ONLY as test infrastructure
// (parity checker for the HIP path and the timed CPU baseline in bench.py). Nothing in
// the product (pdx_amd/) links, imports or executes this file.
//
// It restates, function by function, the reference search semantics:
//   * constants and the dimension split ........ include/pdx/common.hpp:45-49,133-159,191-199
//   * ADSampling ratio table / thresholds ....... include/pdx/pruners/adsampling.hpp:91-98,147-157
//   * U8 threshold conversion ................... include/pdx/searcher.hpp:75-91
//   * query normalisation / quantisation ........ include/pdx/quantizers/scalar.hpp:23-40,76-94
//   * query rotation y = x.M .................... include/pdx/pruners/adsampling.hpp:225-245
//   * distance kernels .......................... include/pdx/distance_computers/scalar_computers.hpp:19-134
//                                                 (F32 Horizontal in the AVX-512 reduction order,
//                                                  avx512_computers.hpp:47-79)
//   * IVF probe ................................. include/pdx/searcher.hpp:181-215
//   * Start / Warmup / Prune / Merge / Result ... include/pdx/searcher.hpp:218-585
//   * Search / FilteredSearch drivers ........... include/pdx/searcher.hpp:596-856
//   * two-level probe ........................... include/pdx/index.hpp:545-566
//   * on-disk format ............................ include/pdx/ivf_wrapper.hpp:89-131,218-312;
//                                                 include/pdx/cluster.hpp:225-255; index.hpp:241-263
//
// Deliberate, documented choices where the reference leaves behaviour implementation-defined:
//   * ties: candidates are ordered by (float distance, row id); the reference's
//     std::priority_queue / partial_sort order equal distances arbitrarily (common.hpp:120-126).
//   * F32 Horizontal (64-dim blocks and full-d centroid distances) uses the AVX-512 order:
//     16 fma lanes (lane l sums dims l, l+16, ...), then lanes (i,i+8), (i,i+4), (i,i+1), (i,i+2).
//   * F32 Vertical is one fma chain per vector in dimension order (scalar_computers.hpp:39-41
//     with contraction), exactly as every reference ISA does it.
//   * query normalisation sums squares in the same 16-lane order; rotation is one fma chain per
//     output in input-dimension order (the reference's BLAS order is library-specific).
//   * SQ8 distances are exact integers (scalar_computers.hpp:77-134 semantics).
// Build with -ffp-contract=off: every fused multiply-add here is an explicit std::fmaf.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <numeric>
#include <queue>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

namespace orc {

constexpr uint32_t H_DIM = 64;          // common.hpp:45
constexpr uint32_t U8_GROUP = 4;        // common.hpp:46
constexpr uint32_t FETCH[20] = {16,  16,  32,  32,   32,   32,   64,   64,   64,    64,
                                128, 128, 128, 128,  256,  256,  512,  1024, 2048, 16384};  // common.hpp:47-49
constexpr float AGGRESSIVENESS = 1.5f;  // common.hpp:67
constexpr float SELECTIVITY = 0.80f;    // searcher.hpp:57

struct Split {
    uint32_t h, v;
};

// common.hpp:133-159
Split dimension_split(uint32_t d) {
    float prop = d <= 128 ? 0.25f : 0.75f;
    uint32_t h = static_cast<uint32_t>(static_cast<float>(d) * prop);
    uint32_t v = d - h;
    if (h % H_DIM > 0) {
        h = ((h + H_DIM / 2) / H_DIM) * H_DIM;
        v = d - h;
    }
    if (!v) {
        h = H_DIM;
        v = d - h;
    }
    if (d <= H_DIM) {
        h = 0;
        v = d;
    }
    return {h, v};
}

// adsampling.hpp:147-157 (float / uint32 division, then two double factors, cast to float)
float ratio_at(size_t i, uint32_t D) {
    if (i == 0) return 1.0f;
    if (i == D) return 1.0f;
    double f = static_cast<double>(static_cast<float>(i) / static_cast<float>(D));
    double a = 1.0 + static_cast<double>(AGGRESSIVENESS) / std::sqrt(static_cast<double>(i));
    return static_cast<float>(f * a * a);
}

// 16-lane fma accumulation + AVX-512 reduction tree (avx512_computers.hpp:47-79).
float l2sq16(const float* a, const float* b, size_t d) {
    float acc[16];
    for (int l = 0; l < 16; ++l) acc[l] = 0.0f;
    for (size_t base = 0; base < d; base += 16) {
        size_t lim = std::min<size_t>(16, d - base);
        for (size_t l = 0; l < lim; ++l) {
            float df = a[base + l] - b[base + l];
            acc[l] = std::fmaf(df, df, acc[l]);
        }
    }
    float x[8], r[4];
    for (int i = 0; i < 8; ++i) x[i] = acc[i] + acc[i + 8];
    for (int i = 0; i < 4; ++i) r[i] = x[i] + x[i + 4];
    return (r[0] + r[1]) + (r[2] + r[3]);
}

// Sum of squares in the same order (used by query normalisation).
float sumsq16(const float* a, size_t d) {
    float acc[16];
    for (int l = 0; l < 16; ++l) acc[l] = 0.0f;
    for (size_t base = 0; base < d; base += 16) {
        size_t lim = std::min<size_t>(16, d - base);
        for (size_t l = 0; l < lim; ++l) acc[l] = std::fmaf(a[base + l], a[base + l], acc[l]);
    }
    float x[8], r[4];
    for (int i = 0; i < 8; ++i) x[i] = acc[i] + acc[i + 8];
    for (int i = 0; i < 4; ++i) r[i] = x[i] + x[i + 4];
    return (r[0] + r[1]) + (r[2] + r[3]);
}

uint32_t l2sq_u8(const uint8_t* a, const uint8_t* b, size_t d) {  // scalar_computers.hpp:123-134
    uint32_t s = 0;
    for (size_t i = 0; i < d; ++i) {
        int df = static_cast<int>(a[i]) - static_cast<int>(b[i]);
        s += static_cast<uint32_t>(df * df);
    }
    return s;
}

struct Cluster {
    uint32_t n_live = 0;   // num_embeddings (cluster.hpp:89)
    uint32_t n_used = 0;   // used_capacity (cluster.hpp:91)
    uint32_t cap = 0;      // max_capacity = stride (cluster.hpp:93)
    std::vector<uint8_t> data;  // strided PDX buffer, cap * d elements
    std::vector<uint32_t> ids;  // row ids per slot
    std::unordered_set<uint32_t> tombstones;  // slot indices
};

struct Level {
    uint32_t d = 0, v = 0, h = 0, nc = 0;
    bool u8 = false;
    bool normalized = false;
    std::vector<Cluster> clusters;
    std::vector<float> centroids;  // nc * d (may be empty: tree L1 after load)
    std::vector<size_t> cluster_offsets;  // ivf_wrapper.hpp:76-87
    size_t total_capacity = 0;
    size_t max_cap = 0;
    float base = 0.0f, scale = 1.0f, scale_sq = 1.0f, inv_scale_sq = 1.0f;

    size_t elem() const { return u8 ? 1 : 4; }
    void compute_offsets() {
        cluster_offsets.assign(nc, 0);
        total_capacity = 0;
        max_cap = 0;
        for (uint32_t c = 0; c < nc; ++c) {
            cluster_offsets[c] = total_capacity;
            total_capacity += clusters[c].cap;
            max_cap = std::max<size_t>(max_cap, clusters[c].cap);
        }
    }
};

struct Index {
    uint8_t type = 0;  // common.hpp:78: 0 F32, 1 U8, 2 TREE_F32, 3 TREE_U8
    uint32_t mrows = 0, mcols = 0;
    std::vector<float> matrix;  // row-major mrows x mcols
    Level l1;                   // data clusters
    Level l0;                   // meso clusters (tree only, always F32)
    std::vector<float> ratios;  // [d]
    bool is_tree() const { return type >= 2; }
};

struct Reader {
    const char* p;
    const char* end;
    template <class T>
    T get() {
        if (p + sizeof(T) > end) throw std::runtime_error("index file truncated");
        T v;
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        return v;
    }
    void bytes(void* dst, size_t n) {
        if (p + n > end) throw std::runtime_error("index file truncated");
        std::memcpy(dst, p, n);
        p += n;
    }
};

// cluster.hpp:225-255: compact file data into the max_capacity-strided buffer.
void load_cluster_data(Reader& r, Cluster& c, const Level& L) {
    const size_t es = L.elem();
    const size_t S = c.cap, n = c.n_live;
    c.data.assign(S * L.d * es, 0);
    uint8_t* data = c.data.data();
    if (!L.u8) {
        for (uint32_t j = 0; j < L.v; ++j) r.bytes(data + (size_t)j * S * es, n * es);
    } else {
        uint32_t j = 0;
        for (; j + U8_GROUP <= L.v; j += U8_GROUP) r.bytes(data + (size_t)j * S, n * U8_GROUP);
        if (j < L.v) r.bytes(data + (size_t)j * S, n * (L.v - j));
    }
    uint8_t* hb = data + (size_t)S * L.v * es;
    for (uint32_t j = 0; j < L.h; j += H_DIM) {
        r.bytes(hb, n * H_DIM * es);
        hb += (size_t)S * H_DIM * es;
    }
}

void load_level_clusters(Reader& r, Level& L, const uint32_t* headers) {
    L.clusters.resize(L.nc);
    for (uint32_t c = 0; c < L.nc; ++c) {
        Cluster& cl = L.clusters[c];
        cl.n_live = cl.n_used = headers[2 * c];
        cl.cap = headers[2 * c + 1];
        if (cl.n_live > cl.cap) throw std::runtime_error("cluster size exceeds capacity");
        load_cluster_data(r, cl, L);
    }
    for (uint32_t c = 0; c < L.nc; ++c) {
        Cluster& cl = L.clusters[c];
        cl.ids.assign(cl.cap, 0);
        r.bytes(cl.ids.data(), sizeof(uint32_t) * cl.n_live);
    }
}

std::unique_ptr<Index> load_index(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    Reader r{buf.data(), buf.data() + buf.size()};
    auto idx = std::make_unique<Index>();
    idx->type = r.get<uint8_t>();
    if (idx->type > 3) throw std::runtime_error("Unknown PDX index type");
    idx->mrows = r.get<uint32_t>();
    idx->mcols = r.get<uint32_t>();
    idx->matrix.resize((size_t)idx->mrows * idx->mcols);
    r.bytes(idx->matrix.data(), idx->matrix.size() * 4);
    const bool u8 = (idx->type == 1 || idx->type == 3);
    Level& L = idx->l1;
    L.u8 = u8;
    if (!idx->is_tree()) {  // ivf_wrapper.hpp:89-131
        L.d = r.get<uint32_t>();
        L.v = r.get<uint32_t>();
        L.h = r.get<uint32_t>();
        L.nc = r.get<uint32_t>();
        std::vector<uint32_t> hdr(2 * (size_t)L.nc);
        r.bytes(hdr.data(), hdr.size() * 4);
        load_level_clusters(r, L, hdr.data());
        L.normalized = r.get<uint8_t>() != 0;
        L.centroids.resize((size_t)L.nc * L.d);
        r.bytes(L.centroids.data(), L.centroids.size() * 4);
    } else {  // ivf_wrapper.hpp:218-312
        uint32_t d = r.get<uint32_t>(), v = r.get<uint32_t>(), h = r.get<uint32_t>();
        uint32_t n1 = r.get<uint32_t>(), n0 = r.get<uint32_t>();
        Level& M = idx->l0;
        M.d = d; M.v = v; M.h = h; M.nc = n0; M.u8 = false;
        std::vector<uint32_t> hdr0(2 * (size_t)n0);
        r.bytes(hdr0.data(), hdr0.size() * 4);
        load_level_clusters(r, M, hdr0.data());
        L.d = d; L.v = v; L.h = h; L.nc = n1;
        std::vector<uint32_t> hdr1(2 * (size_t)n1);
        r.bytes(hdr1.data(), hdr1.size() * 4);
        load_level_clusters(r, L, hdr1.data());
        bool norm = r.get<uint8_t>() != 0;
        L.normalized = M.normalized = norm;
        M.centroids.resize((size_t)n0 * d);
        r.bytes(M.centroids.data(), M.centroids.size() * 4);
        M.compute_offsets();
    }
    if (u8) {
        L.base = r.get<float>();
        L.scale = r.get<float>();
        L.scale_sq = L.scale * L.scale;             // ivf_wrapper.hpp:127
        L.inv_scale_sq = 1.0f / L.scale_sq;         // ivf_wrapper.hpp:128
    }
    L.compute_offsets();
    Split s = dimension_split(L.d);
    if (s.h != L.h || s.v != L.v) throw std::runtime_error("header split disagrees with rule");
    idx->ratios.resize(L.d);
    for (uint32_t i = 0; i < L.d; ++i) idx->ratios[i] = ratio_at(i, L.d);
    return idx;
}

// ---------------------------------------------------------------------------------------------
// Query preprocessing
// ---------------------------------------------------------------------------------------------

// scalar.hpp:23-40. A zero vector is left unchanged (the reference leaves the output unwritten).
void normalize(const float* src, float* out, uint32_t d) {
    float sum = sumsq16(src, d);
    if (sum == 0.0f) {
        std::memcpy(out, src, sizeof(float) * d);
        return;
    }
    float norm = std::sqrt(sum);
    for (uint32_t i = 0; i < d; ++i) out[i] = src[i] / norm;
}

// adsampling.hpp:225-245: y = x . M with M row-major (y_i = sum_j x_j M[j][i]).
void rotate(const Index& idx, const float* x, float* y) {
    const uint32_t d = idx.l1.d;
    if (idx.mrows != d || idx.mcols != d)
        throw std::runtime_error("only dense d x d rotations are supported (no DCT)");
    const float* M = idx.matrix.data();
    for (uint32_t i = 0; i < d; ++i) {
        float acc = 0.0f;
        for (uint32_t j = 0; j < d; ++j) acc = std::fmaf(x[j], M[(size_t)j * d + i], acc);
        y[i] = acc;
    }
}

void preprocess(const Index& idx, const float* raw, float* out) {  // searcher.hpp:601-613
    const uint32_t d = idx.l1.d;
    if (idx.l1.normalized) {
        std::vector<float> tmp(d);
        normalize(raw, tmp.data(), d);
        rotate(idx, tmp.data(), out);
    } else {
        rotate(idx, raw, out);
    }
}

// scalar.hpp:76-94
void quantize(const float* x, float base, float scale, uint8_t* out, uint32_t d) {
    for (uint32_t i = 0; i < d; ++i) {
        int r = static_cast<int>(std::round((x[i] - base) * scale));
        out[i] = r > 255 ? 255 : (r < 0 ? 0 : static_cast<uint8_t>(r));
    }
}

// ---------------------------------------------------------------------------------------------
// Search state
// ---------------------------------------------------------------------------------------------

struct Cand {
    float dist;
    uint32_t id;
    uint32_t idist;  // exact integer distance for SQ8 levels (0 for F32)
};
inline bool key_less(const Cand& a, const Cand& b) {
    return a.dist < b.dist || (a.dist == b.dist && a.id < b.id);
}
struct HeapLess {
    bool operator()(const Cand& a, const Cand& b) const { return key_less(a, b); }
};
using Heap = std::priority_queue<Cand, std::vector<Cand>, HeapLess>;

struct Stats {
    uint64_t bytes = 0;  // algorithmic bytes read (cluster data + ids + centroids)
};

template <bool U8>
struct Scan {
    using dist_t = typename std::conditional<U8, uint32_t, float>::type;
    const Level& L;
    const std::vector<float>& ratios;
    const float* qf = nullptr;    // F32 query
    const uint8_t* qu = nullptr;  // U8 query
    Stats* st;
    std::vector<dist_t> dist;
    std::vector<uint32_t> pos;
    uint32_t k;

    Scan(const Level& L, const std::vector<float>& ratios, Stats* st, uint32_t k)
        : L(L), ratios(ratios), st(st), dist(L.max_cap), pos(L.max_cap), k(k) {}

    static dist_t mask_value() { return std::numeric_limits<dist_t>::max() / 2; }

    float to_float(dist_t v) const {  // searcher.hpp:275-278,552-555
        if constexpr (U8) return static_cast<float>(v) * L.inv_scale_sq;
        else return v;
    }

    // adsampling.hpp:91-98 + searcher.hpp:75-91
    dist_t threshold(const Heap& heap, uint32_t cur) const {
        float ratio = cur == L.d ? 1.0f : ratios[cur];
        float tf = heap.top().dist * ratio;
        if constexpr (U8) {
            float scaled = tf * L.scale_sq;
            return scaled >= static_cast<float>(std::numeric_limits<uint32_t>::max())
                       ? std::numeric_limits<uint32_t>::max()
                       : static_cast<uint32_t>(scaled);
        } else {
            return tf;
        }
    }

    // Vertical (scalar_computers.hpp:19-44 / 77-121) over dims [a,b) for slots (all or gathered).
    void vertical(const Cluster& c, uint32_t a, uint32_t b, size_t nvec, const uint32_t* gather) {
        const size_t S = c.cap;
        if constexpr (!U8) {
            const float* data = reinterpret_cast<const float*>(c.data.data());
            for (uint32_t j = a; j < b; ++j) {
                const float* row = data + (size_t)j * S;
                const float qj = qf[j];
                for (size_t t = 0; t < nvec; ++t) {
                    size_t i = gather ? gather[t] : t;
                    float df = qj - row[i];
                    dist[i] = std::fmaf(df, df, dist[i]);
                }
            }
        } else {
            const uint8_t* data = c.data.data();
            uint32_t j = a;
            for (; j + 4 <= b; j += 4) {
                const uint8_t* grp = data + (size_t)j * S;
                for (size_t t = 0; t < nvec; ++t) {
                    size_t i = gather ? gather[t] : t;
                    const uint8_t* x = grp + 4 * i;
                    int d0 = qu[j] - x[0], d1 = qu[j + 1] - x[1], d2 = qu[j + 2] - x[2],
                        d3 = qu[j + 3] - x[3];
                    dist[i] += static_cast<uint32_t>(d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3);
                }
            }
            if (j < b) {
                const uint32_t rem = b - j;
                const uint8_t* grp = data + (size_t)j * S;
                for (size_t t = 0; t < nvec; ++t) {
                    size_t i = gather ? gather[t] : t;
                    for (uint32_t u = 0; u < rem; ++u) {
                        int df = qu[j + u] - grp[i * rem + u];
                        dist[i] += static_cast<uint32_t>(df * df);
                    }
                }
            }
        }
    }

    // Horizontal 64-dim block hb (element offset into the horizontal region) for slot i.
    dist_t horizontal(const Cluster& c, uint32_t blk, size_t i) const {
        const size_t S = c.cap;
        const size_t off = (size_t)L.v * S + (size_t)blk * S + i * H_DIM;
        if constexpr (!U8) {
            const float* data = reinterpret_cast<const float*>(c.data.data());
            return l2sq16(qf + L.v + blk, data + off, H_DIM);
        } else {
            return l2sq_u8(qu + L.v + blk, c.data.data() + off, H_DIM);
        }
    }

    void mask_tombstones(const Cluster& c) {  // searcher.hpp:169-179
        for (uint32_t t : c.tombstones) dist[t] = mask_value();
    }

    void push_m_smallest(const Cluster& c, size_t n, size_t m, Heap& heap) {
        std::vector<uint32_t> order(n);
        std::iota(order.begin(), order.end(), 0u);
        auto cand = [&](uint32_t i) { return Cand{to_float(dist[i]), c.ids[i], U8 ? (uint32_t)dist[i] : 0u}; };
        std::partial_sort(order.begin(), order.begin() + m, order.end(),
                          [&](uint32_t a, uint32_t b) { return key_less(cand(a), cand(b)); });
        for (size_t t = 0; t < m; ++t) heap.push(cand(order[t]));
        st->bytes += 4 * m;
    }

    // searcher.hpp:218-281
    void start(const Cluster& c, Heap& heap) {
        const size_t n = c.n_used;
        std::fill(dist.begin(), dist.begin() + n, dist_t{0});
        vertical(c, 0, L.v, n, nullptr);
        for (uint32_t hb = 0; hb < L.h; hb += H_DIM)
            for (size_t i = 0; i < n; ++i) dist[i] += horizontal(c, hb, i);
        st->bytes += (uint64_t)n * L.d * L.elem();
        mask_tombstones(c);
        size_t m = std::min<size_t>(k - heap.size(), n);
        push_m_smallest(c, n, m, heap);
    }

    // searcher.hpp:284-372
    void filtered_start(const Cluster& c, Heap& heap, const uint8_t* sel, uint32_t passing) {
        const size_t n = c.n_used;
        std::fill(dist.begin(), dist.begin() + n, dist_t{0});
        size_t np = 0;
        for (size_t i = 0; i < n; ++i) {
            pos[np] = static_cast<uint32_t>(i);
            np += sel[i] == 1;
        }
        float selection = static_cast<float>(passing) / static_cast<float>(n);
        for (uint32_t hb = 0; hb < L.h; hb += H_DIM)
            for (size_t t = 0; t < np; ++t) dist[pos[t]] += horizontal(c, hb, pos[t]);
        st->bytes += (uint64_t)np * L.h * L.elem();
        if (selection > (1 - SELECTIVITY)) {
            vertical(c, 0, L.v, n, nullptr);
            st->bytes += (uint64_t)n * L.v * L.elem();
        } else {
            vertical(c, 0, L.v, np, pos.data());
            st->bytes += (uint64_t)np * L.v * L.elem();
        }
        size_t m = std::min<size_t>(k - heap.size(), passing);
        for (size_t i = 0; i < n; ++i)
            if (sel[i] == 0) dist[i] = mask_value();
        mask_tombstones(c);
        push_m_smallest(c, n, m, heap);
    }

    // searcher.hpp:376-435. Returns false when the filtered variant skips straight to Prune.
    void warmup(const Cluster& c, const Heap& heap, uint32_t& cur, const uint8_t* sel, uint32_t passing) {
        const size_t n = c.n_used;
        cur = 0;
        size_t t = 0;
        size_t need = static_cast<size_t>(std::ceil(SELECTIVITY * static_cast<float>(n)));
        std::fill(dist.begin(), dist.begin() + n, dist_t{0});
        mask_tombstones(c);
        size_t n_pruned = 0;
        if (sel) {
            float selection = static_cast<float>(passing) / static_cast<float>(n);
            for (size_t i = 0; i < n; ++i)
                if (sel[i] == 0) dist[i] = mask_value();
            if (selection < (1 - SELECTIVITY)) return;
        }
        dist_t tau = threshold(heap, cur);
        while (n_pruned < need && cur < L.v) {
            uint32_t end = std::min<uint32_t>(cur + FETCH[t], L.v);
            vertical(c, cur, end, n, nullptr);
            st->bytes += (uint64_t)n * (end - cur) * L.elem();
            cur = end;
            ++t;
            tau = threshold(heap, cur);
            n_pruned = 0;
            for (size_t i = 0; i < n; ++i) n_pruned += dist[i] >= tau;
        }
    }

    // searcher.hpp:438-540
    size_t prune(const Cluster& c, const Heap& heap, uint32_t cur, const uint8_t* sel) {
        const size_t n = c.n_used;
        dist_t tau = threshold(heap, cur);
        mask_tombstones(c);
        size_t np = 0;
        for (size_t i = 0; i < n; ++i) {
            pos[np] = static_cast<uint32_t>(i);
            np += (dist[i] < tau) && (!sel || sel[i] == 1);
        }
        uint32_t cv = cur, hd = 0;
        auto filter = [&]() {
            size_t cnt = np;
            np = 0;
            for (size_t t = 0; t < cnt; ++t) {
                pos[np] = pos[t];
                np += dist[pos[t]] < tau;
            }
        };
        while (L.h && np && hd < L.h) {
            for (size_t t = 0; t < np; ++t) dist[pos[t]] += horizontal(c, hd, pos[t]);
            st->bytes += (uint64_t)np * H_DIM * L.elem();
            hd += H_DIM;
            cur += H_DIM;
            tau = threshold(heap, cur);
            filter();
        }
        while (np && cv < L.v) {
            uint32_t end = std::min<uint32_t>(cv + H_DIM, L.v);
            vertical(c, cv, end, np, pos.data());
            st->bytes += (uint64_t)np * (end - cv) * L.elem();
            cur = std::min<uint32_t>(cur + H_DIM, L.d);
            cv = end;
            tau = threshold(heap, cur);
            filter();
            if (cur == L.d) break;
        }
        return np;
    }

    // searcher.hpp:542-566
    void merge(const Cluster& c, size_t np, Heap& heap) {
        for (size_t t = 0; t < np; ++t) {
            uint32_t i = pos[t];
            Cand cd{to_float(dist[i]), c.ids[i], U8 ? (uint32_t)dist[i] : 0u};
            if (heap.size() < k || key_less(cd, heap.top())) {
                if (heap.size() >= k) heap.pop();
                heap.push(cd);
            }
        }
        st->bytes += 4 * np;
    }
};

// searcher.hpp:181-215 (ties by cluster index)
std::vector<uint32_t> probe_order(const Level& L, const float* q, size_t nprobe, Stats* st) {
    std::vector<float> dc(L.nc);
    for (uint32_t c = 0; c < L.nc; ++c) dc[c] = l2sq16(q, L.centroids.data() + (size_t)c * L.d, L.d);
    st->bytes += (uint64_t)L.nc * L.d * 4;
    std::vector<uint32_t> order(L.nc);
    std::iota(order.begin(), order.end(), 0u);
    auto lt = [&](uint32_t a, uint32_t b) { return dc[a] < dc[b] || (dc[a] == dc[b] && a < b); };
    size_t m = std::min<size_t>(nprobe, L.nc);
    std::partial_sort(order.begin(), order.begin() + m, order.end(), lt);
    order.resize(m);
    return order;
}

// searcher.hpp:568-585
std::vector<Cand> result_from_heap(Heap& heap, uint32_t k) {
    if (!heap.empty() && heap.top().dist == std::numeric_limits<float>::max()) heap.pop();
    size_t n = std::min<size_t>(heap.size(), k);
    std::vector<Cand> out(n);
    for (size_t i = n; i > 0; --i) {
        out[i - 1] = heap.top();
        heap.pop();
    }
    return out;
}

struct Selection {  // db_mock/predicate_evaluator.hpp:9-31, built as in index.hpp:436-446
    std::vector<uint8_t> sel;      // total_capacity
    std::vector<uint32_t> passing; // per cluster
};

// searcher.hpp:596-724 (order == nullptr: IVF probe) and 726-856 (selection != nullptr).
// `q` is the already preprocessed (normalised + rotated) query.
template <bool U8>
std::vector<Cand> search_level(const Level& L, const std::vector<float>& ratios, const float* q,
                               uint32_t k, size_t nprobe, const std::vector<uint32_t>* order,
                               const Selection* filt, Stats* st) {
    Heap heap;
    size_t nvisit = (nprobe == 0 || nprobe > L.nc) ? L.nc : nprobe;
    std::vector<uint32_t> ord;
    if (order && !filt) {
        nvisit = std::min(nvisit, order->size());
        ord.assign(order->begin(), order->begin() + nvisit);
    } else {
        if (L.centroids.size() != (size_t)L.nc * L.d)
            throw std::runtime_error("level has no centroids (two-level index restored from file)");
        ord = probe_order(L, q, nvisit, st);
    }
    Scan<U8> sc(L, ratios, st, k);
    std::vector<uint8_t> qq;
    if constexpr (U8) {
        qq.resize(L.d);
        quantize(q, L.base, L.scale, qq.data(), L.d);
        sc.qu = qq.data();
    } else {
        sc.qf = q;
    }
    for (size_t ci = 0; ci < ord.size(); ++ci) {
        const uint32_t cid = ord[ci];
        const Cluster& c = L.clusters[cid];
        const uint8_t* sel = nullptr;
        uint32_t passing = 0;
        if (filt) {
            sel = filt->sel.data() + L.cluster_offsets[cid];
            passing = filt->passing[cid];
            if (passing == 0) continue;
        }
        if (c.n_live == 0) continue;
        uint32_t cur = 0;
        if (heap.size() < k) {
            if (filt) sc.filtered_start(c, heap, sel, passing);
            else sc.start(c, heap);
            continue;
        }
        sc.warmup(c, heap, cur, sel, passing);
        size_t np = sc.prune(c, heap, cur, sel);
        if (np) sc.merge(c, np, heap);
    }
    return result_from_heap(heap, k);
}

std::vector<Cand> search_level_any(const Level& L, const std::vector<float>& ratios, const float* q,
                                   uint32_t k, size_t nprobe, const std::vector<uint32_t>* order,
                                   const Selection* filt, Stats* st) {
    return L.u8 ? search_level<true>(L, ratios, q, k, nprobe, order, filt, st)
                : search_level<false>(L, ratios, q, k, nprobe, order, filt, st);
}

// PDXIndex::Search (index.hpp:269-271) and PDXTreeIndex::Search (index.hpp:545-566).
std::vector<Cand> search(const Index& idx, const float* q_pre, uint32_t k, uint32_t nprobe, Stats* st) {
    if (!idx.is_tree()) return search_level_any(idx.l1, idx.ratios, q_pre, k, nprobe, nullptr, nullptr, st);
    uint32_t np = nprobe == 0 ? idx.l1.nc : nprobe;
    uint32_t np0 = idx.l0.nc;
    if (np < idx.l1.nc / 2) np0 /= 2;
    auto top = search_level<false>(idx.l0, idx.ratios, q_pre, np, np0, nullptr, nullptr, st);
    std::vector<uint32_t> order(top.size());
    for (size_t i = 0; i < top.size(); ++i) order[i] = top[i].id;
    return search_level_any(idx.l1, idx.ratios, q_pre, k, np, &order, nullptr, st);
}

}  // namespace orc

// =============================================================================================
// C API for ctypes
// =============================================================================================

static thread_local std::string g_err;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_load(const char* path) {
    try {
        return orc::load_index(path).release();
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_free(void* h) { delete static_cast<orc::Index*>(h); }

// info: [type, d, v, h, n_l1, n_l0, normalized, max_cap_l1, max_cap_l0]
void orc_info(void* h, uint32_t* out) {
    auto* idx = static_cast<orc::Index*>(h);
    out[0] = idx->type;
    out[1] = idx->l1.d;
    out[2] = idx->l1.v;
    out[3] = idx->l1.h;
    out[4] = idx->l1.nc;
    out[5] = idx->l0.nc;
    out[6] = idx->l1.normalized;
    out[7] = (uint32_t)idx->l1.max_cap;
    out[8] = (uint32_t)idx->l0.max_cap;
}

void orc_quant_params(void* h, float* out) {
    auto* idx = static_cast<orc::Index*>(h);
    out[0] = idx->l1.base;
    out[1] = idx->l1.scale;
    out[2] = idx->l1.scale_sq;
    out[3] = idx->l1.inv_scale_sq;
}

void orc_split(uint32_t d, uint32_t* h, uint32_t* v) {
    auto s = orc::dimension_split(d);
    *h = s.h;
    *v = s.v;
}

void orc_ratios(uint32_t d, float* out) {
    for (uint32_t i = 0; i < d; ++i) out[i] = orc::ratio_at(i, d);
}

float orc_l2sq16(const float* a, const float* b, uint32_t d) { return orc::l2sq16(a, b, d); }
uint32_t orc_l2sq_u8(const uint8_t* a, const uint8_t* b, uint32_t d) { return orc::l2sq_u8(a, b, d); }

// F32 distance of n row-major vectors in the order Start gives a vector it scans whole
// (searcher.hpp:218-281): the vertical fma chain over dims [0, v), then each 64-dim horizontal
// block added in turn (distance_computers: vertical, then horizontal in the AVX-512 order).
void orc_start_distances(const float* q, const float* x, uint32_t n, uint32_t d, uint32_t v, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* row = x + (size_t)i * d;
        float acc = 0.0f;
        for (uint32_t j = 0; j < v; ++j) {
            const float df = q[j] - row[j];
            acc = std::fmaf(df, df, acc);
        }
        for (uint32_t hb = v; hb < d; hb += orc::H_DIM) acc += orc::l2sq16(q + hb, row + hb, orc::H_DIM);
        out[i] = acc;
    }
}

void orc_normalize(const float* x, float* out, uint32_t d) { orc::normalize(x, out, d); }

void orc_quantize(const float* x, float base, float scale, uint8_t* out, uint32_t d) {
    orc::quantize(x, base, scale, out, d);
}

// Normalise (if the index is normalised) and rotate nq raw queries.
int orc_preprocess(void* h, const float* raw, uint32_t nq, float* out) {
    try {
        auto* idx = static_cast<orc::Index*>(h);
        const uint32_t d = idx->l1.d;
        for (uint32_t q = 0; q < nq; ++q) orc::preprocess(*idx, raw + (size_t)q * d, out + (size_t)q * d);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Batched search. queries: nq x d (raw unless `transformed`); k/nprobe per query; outputs with
// row stride `kstride`. bytes (optional): algorithmic bytes per query. n_threads: worker threads.
int orc_search(void* h, const float* queries, uint32_t nq, int transformed, const uint32_t* k,
               const uint32_t* nprobe, uint32_t kstride, uint32_t* out_ids, float* out_dists,
               uint32_t* out_idists, uint32_t* out_counts, uint64_t* out_bytes, int n_threads) {
    auto* idx = static_cast<orc::Index*>(h);
    const uint32_t d = idx->l1.d;
    std::atomic<uint32_t> next{0};
    std::atomic<int> failed{0};
    std::string err;
    auto worker = [&]() {
        std::vector<float> qp(d);
        for (;;) {
            uint32_t q = next.fetch_add(1);
            if (q >= nq || failed.load()) break;
            try {
                const float* qr = queries + (size_t)q * d;
                if (transformed) std::memcpy(qp.data(), qr, sizeof(float) * d);
                else orc::preprocess(*idx, qr, qp.data());
                orc::Stats st;
                if (k[q] > kstride) throw std::runtime_error("k exceeds kstride");
                auto res = orc::search(*idx, qp.data(), k[q], nprobe[q], &st);
                out_counts[q] = (uint32_t)res.size();
                for (size_t i = 0; i < res.size(); ++i) {
                    out_ids[(size_t)q * kstride + i] = res[i].id;
                    out_dists[(size_t)q * kstride + i] = res[i].dist;
                    if (out_idists) out_idists[(size_t)q * kstride + i] = res[i].idist;
                }
                if (out_bytes) out_bytes[q] = st.bytes;
            } catch (const std::exception& e) {
                if (!failed.exchange(1)) err = e.what();
            }
        }
    };
    int nt = std::max(1, n_threads);
    if (nt == 1) {
        worker();
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) th.emplace_back(worker);
        for (auto& t : th) t.join();
    }
    if (failed.load()) {
        g_err = err;
        return -1;
    }
    return 0;
}

// Meso-level search alone (two-level indexes): the top-level PDXearch<F32> over the L1 centroids,
// as PDXTreeIndex::Append runs it to pick a cluster (index.hpp:596-607: nprobe over the meso
// clusters, k results, already transformed vector). Row stride of the outputs is k.
int orc_search_meso(void* h, const float* queries_transformed, uint32_t nq, uint32_t k, uint32_t nprobe,
                    uint32_t* out_ids, float* out_dists, uint32_t* out_counts) {
    try {
        auto* idx = static_cast<orc::Index*>(h);
        if (!idx->is_tree()) throw std::runtime_error("not a two-level index");
        const uint32_t d = idx->l1.d;
        for (uint32_t q = 0; q < nq; ++q) {
            orc::Stats st;
            auto res = orc::search_level<false>(idx->l0, idx->ratios, queries_transformed + (size_t)q * d, k, nprobe,
                                                nullptr, nullptr, &st);
            out_counts[q] = (uint32_t)res.size();
            for (size_t i = 0; i < res.size(); ++i) {
                out_ids[(size_t)q * k + i] = res[i].id;
                out_dists[(size_t)q * k + i] = res[i].dist;
            }
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Filtered search of one level-1 probe (PDXIndex::FilteredSearch, index.hpp:273-280).
// `slot_sel`: selection flag per (cluster offset + slot), total_capacity entries.
int orc_filtered_search(void* h, const float* query, int transformed, uint32_t k, uint32_t nprobe,
                        const uint8_t* slot_sel, uint32_t* out_ids, float* out_dists,
                        uint32_t* out_idists, uint32_t* out_count) {
    try {
        auto* idx = static_cast<orc::Index*>(h);
        const orc::Level& L = idx->l1;
        std::vector<float> qp(L.d);
        if (transformed) std::memcpy(qp.data(), query, sizeof(float) * L.d);
        else orc::preprocess(*idx, query, qp.data());
        orc::Selection s;
        s.sel.assign(slot_sel, slot_sel + L.total_capacity);
        s.passing.assign(L.nc, 0);
        for (uint32_t c = 0; c < L.nc; ++c)
            for (uint32_t i = 0; i < L.clusters[c].cap; ++i) s.passing[c] += s.sel[L.cluster_offsets[c] + i] == 1;
        orc::Stats st;
        auto res = orc::search_level_any(L, idx->ratios, qp.data(), k, nprobe, nullptr, &s, &st);
        *out_count = (uint32_t)res.size();
        for (size_t i = 0; i < res.size(); ++i) {
            out_ids[i] = res[i].id;
            out_dists[i] = res[i].dist;
            if (out_idists) out_idists[i] = res[i].idist;
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// In-memory construction (same content as a file, without the file): one cluster description per
// cluster, strided PDX buffers with stride = max_capacity (cluster.hpp:17-105).
struct orc_cluster {
    uint32_t n_live, n_used, cap;
    const void* data;
    const uint32_t* ids;
    const uint32_t* tombstones;
    uint32_t n_tombstones;
};

static void fill_level(orc::Level& L, uint32_t nc, const orc_cluster* cs, const float* cent) {
    L.nc = nc;
    L.clusters.resize(nc);
    for (uint32_t c = 0; c < nc; ++c) {
        const orc_cluster& x = cs[c];
        orc::Cluster& cl = L.clusters[c];
        cl.n_live = x.n_live;
        cl.n_used = x.n_used;
        cl.cap = x.cap;
        const size_t bytes = (size_t)x.cap * L.d * L.elem();
        cl.data.assign(static_cast<const uint8_t*>(x.data), static_cast<const uint8_t*>(x.data) + bytes);
        cl.ids.assign(x.cap, 0);
        std::memcpy(cl.ids.data(), x.ids, sizeof(uint32_t) * x.n_used);
        for (uint32_t t = 0; t < x.n_tombstones; ++t) cl.tombstones.insert(x.tombstones[t]);
    }
    if (cent) L.centroids.assign(cent, cent + (size_t)nc * L.d);
    L.compute_offsets();
}

void* orc_create(uint32_t type, uint32_t d, uint32_t normalized, float base, float scale, const float* rotation,
                 uint32_t nc1, const orc_cluster* c1, const float* cent1, uint32_t nc0, const orc_cluster* c0,
                 const float* cent0) {
    try {
        auto idx = std::make_unique<orc::Index>();
        idx->type = (uint8_t)type;
        idx->mrows = idx->mcols = d;
        idx->matrix.assign(rotation, rotation + (size_t)d * d);
        orc::Split s = orc::dimension_split(d);
        const bool u8 = type == 1 || type == 3;
        orc::Level& L = idx->l1;
        L.d = d; L.v = s.v; L.h = s.h; L.u8 = u8; L.normalized = normalized != 0;
        fill_level(L, nc1, c1, cent1);
        if (u8) {
            L.base = base;
            L.scale = scale;
            L.scale_sq = scale * scale;
            L.inv_scale_sq = 1.0f / L.scale_sq;
        }
        if (type >= 2) {
            orc::Level& M = idx->l0;
            M.d = d; M.v = s.v; M.h = s.h; M.u8 = false; M.normalized = L.normalized;
            fill_level(M, nc0, c0, cent0);
        }
        idx->ratios.resize(d);
        for (uint32_t i = 0; i < d; ++i) idx->ratios[i] = orc::ratio_at(i, d);
        return idx.release();
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// Level-1 cluster offsets (cluster_offsets of ivf_wrapper.hpp:76-87), nc entries + total.
void orc_cluster_offsets(void* h, uint64_t* out) {
    auto* idx = static_cast<orc::Index*>(h);
    for (uint32_t c = 0; c < idx->l1.nc; ++c) out[c] = idx->l1.cluster_offsets[c];
    out[idx->l1.nc] = idx->l1.total_capacity;
}

// Row ids of every slot of level-1 cluster c (n_used entries) and its sizes.
uint32_t orc_cluster_ids(void* h, uint32_t c, uint32_t* out) {
    auto* idx = static_cast<orc::Index*>(h);
    const auto& cl = idx->l1.clusters[c];
    std::memcpy(out, cl.ids.data(), sizeof(uint32_t) * cl.n_used);
    return cl.n_used;
}

// Tombstone slot `slot` of level-1 cluster c (Cluster::DeleteEmbedding, cluster.hpp:148-154).
int orc_delete_slot(void* h, uint32_t c, uint32_t slot) {
    auto* idx = static_cast<orc::Index*>(h);
    auto& cl = idx->l1.clusters[c];
    if (slot >= cl.n_used || cl.tombstones.count(slot)) return -1;
    cl.tombstones.insert(slot);
    cl.n_live--;
    return 0;
}

}  // extern "C"
