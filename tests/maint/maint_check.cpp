// Test infrastructure: drives the host index's maintenance (PDXTreeIndex Append / Delete with
// cluster splits and merges at both levels, index.hpp:1004-1685) on the CPU and checks the index's
// structural invariants after every operation. The C-ABI calls the host index makes are replaced
// here by CPU stand-ins: preprocessing (normalise + rotate) and the meso-level cluster choice of an
// append (the nearest stored L1 centroid); the device copy itself is not modelled. Built with
// AddressSanitizer by tests/test_maintenance_host.py.
#include <cstdio>
#include <map>
#include <random>

#include "pdx_index.hpp"

namespace {
struct StubIndex {
    uint32_t d = 0;
    bool normalized = false;
    float base = 0.0f, scale = 1.0f;
    std::vector<float> rotation;
    std::vector<float> l1_centroids;  // from the meso level's stored vectors
    std::vector<uint32_t> l1_ids;
};
thread_local std::string g_err;
}  // namespace

extern "C" {
const char* pdxg_last_error(void) { return g_err.c_str(); }
pdxg_status pdxg_index_create(int, const pdxg_index_desc* desc, pdxg_index** out) {
    auto* s = new StubIndex;
    s->d = desc->num_dimensions;
    s->normalized = desc->is_normalized != 0;
    s->base = desc->quantization_base;
    s->scale = desc->quantization_scale;
    s->rotation.assign(desc->rotation, desc->rotation + (size_t)s->d * s->d);
    const auto& L0 = desc->l0;
    const uint32_t v = desc->num_vertical_dimensions;
    for (uint32_t c = 0; c < L0.num_clusters; ++c) {
        const auto& cd = L0.clusters[c];
        std::set<uint32_t> dead(cd.tombstones, cd.tombstones + cd.num_tombstones);
        const float* data = static_cast<const float*>(cd.data);
        for (uint32_t sl = 0; sl < cd.used_capacity; ++sl) {
            if (dead.count(sl)) continue;
            for (uint32_t j = 0; j < s->d; ++j) {
                size_t off = j < v ? (size_t)j * cd.max_capacity + sl
                                   : (size_t)v * cd.max_capacity + (size_t)((j - v) / 64) * 64 * cd.max_capacity +
                                         (size_t)sl * 64 + (j - v) % 64;
                s->l1_centroids.push_back(data[off]);
            }
            s->l1_ids.push_back(cd.ids[sl]);
        }
    }
    *out = reinterpret_cast<pdxg_index*>(s);
    return PDXG_OK;
}
pdxg_status pdxg_index_destroy(pdxg_index* ix) {
    delete reinterpret_cast<StubIndex*>(ix);
    return PDXG_OK;
}
pdxg_status pdxg_index_update_cluster(pdxg_index*, uint32_t, uint32_t, const pdxg_cluster_desc* cd) {
    if (cd->num_embeddings > cd->used_capacity || cd->used_capacity > cd->max_capacity) {
        g_err = "inconsistent cluster";
        return PDXG_ERR_INVALID;
    }
    return PDXG_OK;
}
pdxg_status pdxg_preprocess(pdxg_index* ix, const float* q, uint32_t nq, uint32_t transformed, float* out_rot,
                            uint8_t* out_q) {
    auto* s = reinterpret_cast<StubIndex*>(ix);
    for (uint32_t i = 0; i < nq; ++i) {
        std::vector<float> x(q + (size_t)i * s->d, q + (size_t)(i + 1) * s->d);
        if (!transformed) {
            if (s->normalized) PDX::host::normalize_row(x.data(), x.data(), s->d);
            for (uint32_t o = 0; o < s->d; ++o) {
                float acc = 0.0f;
                for (uint32_t j = 0; j < s->d; ++j) acc += x[j] * s->rotation[(size_t)j * s->d + o];
                out_rot[(size_t)i * s->d + o] = acc;
            }
        } else {
            std::memcpy(out_rot + (size_t)i * s->d, x.data(), s->d * 4);
        }
        if (out_q) PDX::detail::quantize(out_rot + (size_t)i * s->d, s->base, s->scale, out_q + (size_t)i * s->d, s->d);
    }
    return PDXG_OK;
}
pdxg_status pdxg_search_meso(pdxg_index* ix, const float* q, uint32_t nq, uint32_t k, uint32_t, uint32_t* out_ids,
                             float* out_d, uint32_t* out_counts) {
    auto* s = reinterpret_cast<StubIndex*>(ix);
    for (uint32_t i = 0; i < nq; ++i) {
        float best = std::numeric_limits<float>::infinity();
        uint32_t bi = 0;
        for (size_t c = 0; c < s->l1_ids.size(); ++c) {
            float v = PDX::detail::l2sq16(q + (size_t)i * s->d, &s->l1_centroids[c * s->d], s->d);
            if (v < best) {
                best = v;
                bi = s->l1_ids[c];
            }
        }
        out_ids[(size_t)i * k] = bi;
        out_d[(size_t)i * k] = best;
        out_counts[i] = s->l1_ids.empty() ? 0 : 1;
    }
    return PDXG_OK;
}
pdxg_status pdxg_search(pdxg_index*, const pdxg_search_args*) {
    g_err = "search is not modelled";
    return PDXG_ERR_UNSUPPORTED;
}
pdxg_status pdxg_filtered_search(pdxg_index*, const float*, uint32_t, uint32_t, uint32_t, const uint8_t*, uint32_t*,
                                 float*, uint32_t*, uint32_t*) {
    g_err = "search is not modelled";
    return PDXG_ERR_UNSUPPORTED;
}
}  // extern "C"

static int g_failures = 0;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            std::fprintf(stderr, "FAIL %s: ", #cond);      \
            std::fprintf(stderr, __VA_ARGS__);             \
            std::fprintf(stderr, "\n");                    \
            if (++g_failures > 20) std::exit(1);           \
        }                                                  \
    } while (0)

// Every invariant the index structure must keep across maintenance.
template <class IDX>
void check_invariants(const IDX& ix, const std::vector<char>& live, const char* when) {
    const auto& L = ix.GetIndex();
    const uint32_t d = L.num_dimensions, nc = L.num_clusters, n0 = L.l0.num_clusters;
    CHECK(L.clusters.size() == nc && L.centroids.size() == (size_t)nc * d, "%s: L1 sizes", when);
    CHECK(L.l0.clusters.size() == n0 && L.l0.centroids.size() == (size_t)n0 * d, "%s: L0 sizes", when);
    size_t total_live = 0;
    for (uint32_t c = 0; c < nc; ++c) {
        const auto& cl = L.clusters[c];
        CHECK(cl.id == c, "%s: cluster %u id %u", when, c, cl.id);
        CHECK(cl.used_capacity <= cl.max_capacity, "%s: cluster %u over capacity", when, c);
        CHECK(cl.num_embeddings + cl.tombstones.size() == cl.used_capacity, "%s: cluster %u counts", when, c);
        CHECK(cl.mesocluster_id < n0, "%s: cluster %u meso %u", when, c, cl.mesocluster_id);
        for (uint32_t s = 0; s < cl.used_capacity; ++s) {
            if (cl.tombstones.count(s)) continue;
            uint32_t row = cl.indices[s];
            CHECK(row < live.size() && live[row], "%s: dead row %u in cluster %u", when, row, c);
            auto loc = ix.GetRowLocation(row);
            CHECK(loc.first == c && loc.second == s, "%s: row %u at (%u,%u), mapping (%u,%u)", when, row, c, s,
                  loc.first, loc.second);
            total_live++;
        }
    }
    size_t want = 0;
    for (size_t r = 0; r < live.size(); ++r) {
        if (!live[r]) {
            CHECK(ix.GetRowLocation(r).first == std::numeric_limits<uint32_t>::max(), "%s: deleted row %zu mapped",
                  when, r);
            continue;
        }
        want++;
        auto loc = ix.GetRowLocation(r);
        CHECK(loc.first < nc, "%s: live row %zu unmapped", when, r);
    }
    CHECK(total_live == want, "%s: %zu live slots for %zu live rows", when, total_live, want);
    CHECK(L.total_num_embeddings == want, "%s: total_num_embeddings %lu != %zu", when,
          (unsigned long)L.total_num_embeddings, want);
    // the meso level holds each L1 cluster exactly once, in its mesocluster_id, with its centroid
    std::vector<int> seen(nc, 0);
    for (uint32_t m = 0; m < n0; ++m) {
        const auto& mc = L.l0.clusters[m];
        CHECK(mc.id == m, "%s: meso %u id", when, m);
        CHECK(mc.num_embeddings + mc.tombstones.size() == mc.used_capacity, "%s: meso %u counts", when, m);
        std::vector<float> v(d);
        for (uint32_t s = 0; s < mc.used_capacity; ++s) {
            if (mc.tombstones.count(s)) continue;
            uint32_t c = mc.indices[s];
            CHECK(c < nc, "%s: meso %u holds cluster %u of %u", when, m, c, nc);
            if (c >= nc) continue;
            seen[c]++;
            CHECK(L.clusters[c].mesocluster_id == m, "%s: cluster %u in meso %u, mesocluster_id %u", when, c, m,
                  L.clusters[c].mesocluster_id);
            mc.ReadEmbedding(s, v.data());
            CHECK(std::memcmp(v.data(), &L.centroids[(size_t)c * d], d * 4) == 0, "%s: centroid of %u", when, c);
        }
    }
    for (uint32_t c = 0; c < nc; ++c) CHECK(seen[c] == 1, "%s: cluster %u in %d meso clusters", when, c, seen[c]);
}

template <class IDX>
int run(PDX::DistanceMetric metric, bool normalize, uint32_t seed, uint32_t nlist, uint32_t nmeso, int burst,
        bool want_meso_split) {
    const uint32_t d = 32, n = 3000;
    std::mt19937 rng(seed);
    std::normal_distribution<float> nd;
    std::vector<float> cent(8 * d), x((size_t)n * d);
    for (auto& v : cent) v = nd(rng);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < d; ++j) x[(size_t)i * d + j] = cent[(i % 8) * d + j] + 0.3f * nd(rng);
    PDX::PDXIndexConfig cfg{d};
    cfg.distance_metric = metric;
    cfg.normalize = normalize;
    cfg.num_clusters = nlist;
    cfg.num_meso_clusters = nmeso;
    cfg.seed = seed;
    cfg.kmeans_iters = 4;
    IDX ix(cfg);
    ix.BuildIndex(x.data(), n);
    std::vector<char> live(n, 1);
    check_invariants(ix, live, "build");
    const uint32_t nc0 = ix.GetNumClusters(), n00 = ix.GetIndex().l0.num_clusters;
    uint32_t max_nc = nc0, max_n0 = n00;
    // a burst of appends around one point: its cluster fills and splits, repeatedly
    std::vector<float> e(d);
    size_t row = n;
    for (int i = 0; i < burst; ++i, ++row) {
        for (uint32_t j = 0; j < d; ++j) e[j] = cent[j] + 0.05f * nd(rng) + (i % 7) * 0.02f;
        ix.Append(row, e.data());
        live.push_back(1);
        if (i % 397 == 0) check_invariants(ix, live, "append burst");
        max_nc = std::max(max_nc, ix.GetNumClusters());
        max_n0 = std::max(max_n0, ix.GetIndex().l0.num_clusters);
    }
    check_invariants(ix, live, "after appends");
    const uint32_t nc1 = ix.GetNumClusters();
    // delete most rows: clusters fall to half their size and merge
    std::vector<size_t> order(live.size());
    std::iota(order.begin(), order.end(), 0);
    std::shuffle(order.begin(), order.end(), rng);
    uint32_t min_nc = nc1;
    for (size_t t = 0; t < order.size() * 85 / 100; ++t) {
        ix.Delete(order[t]);
        live[order[t]] = 0;
        if (t % 131 == 0) check_invariants(ix, live, "deletes");
        min_nc = std::min(min_nc, ix.GetNumClusters());
    }
    check_invariants(ix, live, "after deletes");
    // and a mixed tail
    for (int i = 0; i < 1500; ++i) {
        if (i % 3) {
            for (uint32_t j = 0; j < d; ++j) e[j] = cent[((i / 3) % 8) * d + j] + 0.3f * nd(rng);
            ix.Append(row++, e.data());
            live.push_back(1);
        } else {
            size_t r = rng() % live.size();
            if (live[r]) {
                ix.Delete(r);
                live[r] = 0;
            }
        }
    }
    check_invariants(ix, live, "mixed");
    std::printf("clusters %u -> max %u -> %u after appends -> min %u after deletes -> %u; meso %u -> max %u -> %u\n",
                nc0, max_nc, nc1, min_nc, ix.GetNumClusters(), n00, max_n0, ix.GetIndex().l0.num_clusters);
    CHECK(max_nc > nc0, "no L1 split happened");
    CHECK(min_nc < nc1, "no L1 merge happened");
    if (want_meso_split) CHECK(max_n0 > n00, "no meso split happened");
    return 0;
}

int main() {
    // L1 splits and merges, meso merges
    run<PDX::PDXTreeIndexF32>(PDX::DistanceMetric::L2SQ, false, 11, 40, 6, 2500, false);
    run<PDX::PDXTreeIndexF32>(PDX::DistanceMetric::COSINE, true, 12, 40, 6, 2500, false);
    run<PDX::PDXTreeIndexU8>(PDX::DistanceMetric::L2SQ, false, 14, 40, 6, 2500, false);
    run<PDX::PDXTreeIndexU8>(PDX::DistanceMetric::IP, true, 15, 60, 3, 2500, false);
    // one meso cluster of 197 L1 clusters (capacity 256): enough L1 splits split it
    run<PDX::PDXTreeIndexF32>(PDX::DistanceMetric::L2SQ, false, 13, 197, 1, 14000, true);
    if (g_failures) {
        std::fprintf(stderr, "%d failures\n", g_failures);
        return 1;
    }
    std::printf("maintenance invariants hold\n");
    return 0;
}
