// C-ABI of the MI355X PDX search path (include/pdx_amd.h): device residency of an index,
// per-call workspaces and the launch sequence of the search kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pdx_amd.h"
#include "pdx_kernels.hpp"
#include "pdx_meso.hpp"

namespace {

using namespace pdxg;

thread_local std::string g_last_error;

struct Error : std::runtime_error {
    pdxg_status code;
    Error(pdxg_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            throw Error(e_ == hipErrorOutOfMemory ? PDXG_ERR_NOMEM : PDXG_ERR_HIP,               \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                      \
    } while (0)

void require(bool ok, const std::string& msg, pdxg_status code = PDXG_ERR_INVALID) {
    if (!ok) throw Error(code, msg);
}

template <class F>
pdxg_status guarded(F&& f) {
    try {
        f();
        return PDXG_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return PDXG_ERR_INVALID;
    }
}

// On-chip top-k buffers are sized by the largest k of a call; beyond this the GPU path reports
// PDXG_ERR_UNSUPPORTED rather than silently degrading.
constexpr uint32_t kMaxK = 2048;
constexpr uint32_t kMaxProbeSort = 8192;  // k_probe_select sorts next_pow2(nc) keys in LDS

// adsampling.hpp:147-157
float ratio_at(size_t i, uint32_t D) {
    if (i == 0 || i == D) return 1.0f;
    double f = static_cast<double>(static_cast<float>(i) / static_cast<float>(D));
    double a = 1.0 + 1.5 / std::sqrt(static_cast<double>(i));
    return static_cast<float>(f * a * a);
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool owned = true;  // false: a view into memory the buffer does not own (never freed)
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
        owned = true;
    }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        n = count;
    }
    void view(T* ptr, size_t count) {
        release();
        p = ptr;
        n = count;
        owned = false;
    }
    void ensure(size_t count) {
        if (count > n || !p) alloc(std::max(count, n * 5 / 4));
    }
    void upload(const T* h, size_t count, size_t off = 0) {
        if (count) HIP_CHECK(hipMemcpy(p + off, h, count * sizeof(T), hipMemcpyHostToDevice));
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Device stride alignment of a cluster's PDX buffer, in vectors: 32 makes every dim-major row
// (4 bytes per vector: an F32 dim or a U8 group of 4 dims) start on a 128-byte line.
constexpr uint32_t kStrideAlign = 32;

struct Level {
    uint32_t nc = 0, d = 0, v = 0, h = 0;
    bool u8 = false;
    bool has_centroids = false;
    std::vector<uint32_t> n_used, n_live, cap, n_tomb;
    std::vector<uint32_t> stride;  // device stride of each cluster: cap rounded up to kStrideAlign
    std::vector<uint64_t> data_off, slot_off, tomb_off;
    uint64_t data_bytes = 0, total_slots = 0, tomb_words = 0;
    uint64_t tomb_total = 0;
    // [m] = sum of the m largest n_used, each rounded up to a multiple of 4: the summarised
    // vectors a query visiting m clusters may need in the two-phase meso scan (16-byte aligned
    // component rows)
    std::vector<uint64_t> top_prefix;
    DevBuf<uint8_t> data;
    DevBuf<uint64_t> d_data_off, d_slot_off, d_tomb_off;
    DevBuf<uint32_t> ids, d_n_used, d_n_live, d_stride, tomb, d_n_tomb;
    DevBuf<float> centroids;

    size_t elem() const { return u8 ? 1 : 4; }
    uint32_t max_used() const {
        uint32_t m = 1;
        for (uint32_t x : n_used) m = std::max(m, x);
        return m;
    }
    size_t device_bytes() const {
        return data.bytes() + d_data_off.bytes() + d_slot_off.bytes() + d_tomb_off.bytes() + ids.bytes() +
               d_n_used.bytes() + d_n_live.bytes() + d_stride.bytes() + tomb.bytes() + d_n_tomb.bytes() +
               centroids.bytes();
    }

    void check_cluster(const pdxg_cluster_desc& c) const {
        require(c.max_capacity > 0, "cluster max_capacity must be positive");
        require(c.used_capacity <= c.max_capacity, "used_capacity exceeds max_capacity");
        require(c.num_embeddings <= c.used_capacity, "num_embeddings exceeds used_capacity");
        require(c.num_embeddings + c.num_tombstones == c.used_capacity,
                "num_embeddings + tombstones must equal used_capacity");
        require(c.used_capacity == 0 || (c.data && c.ids), "cluster data/ids missing");
        require((uint64_t)(c.max_capacity + kStrideAlign) * d < (1ull << 32), "cluster buffer too large");
        for (uint32_t t = 0; t < c.num_tombstones; ++t)
            require(c.tombstones[t] < c.used_capacity, "tombstone slot out of range");
    }

    // The cluster's PDX buffer (stride max_capacity) re-laid with stride[c]: each row of the layout
    // (an F32 vertical dim, a U8 group of 4 vertical dims or the U8 tail group of v % 4 dims, a
    // 64-dim horizontal block) keeps its per-vector layout and starts at its dim position times
    // the new stride; the padding is zero.
    void upload_cluster(uint32_t c, const pdxg_cluster_desc& cd) {
        if (cd.used_capacity) {
            const size_t e = elem(), S = cd.max_capacity, T = stride[c];
            std::vector<uint8_t> buf(T * d * e, 0);
            const uint8_t* src = static_cast<const uint8_t*>(cd.data);
            auto row = [&](size_t p, size_t w) { std::memcpy(buf.data() + p * T * e, src + p * S * e, S * w * e); };
            if (u8) {
                const uint32_t rem = v % 4;
                for (uint32_t p = 0; p + 4 <= v; p += 4) row(p, 4);
                if (rem) row(v - rem, rem);
            } else {
                for (uint32_t p = 0; p < v; ++p) row(p, 1);
            }
            for (uint32_t p = v; p < d; p += kHDim) row(p, kHDim);
            HIP_CHECK(hipMemcpy(data.p + data_off[c], buf.data(), buf.size(), hipMemcpyHostToDevice));
            ids.upload(cd.ids, cd.used_capacity, slot_off[c]);
        }
        const uint32_t words = (cd.max_capacity + 31) / 32;
        std::vector<uint32_t> bm(words, 0u);
        for (uint32_t t = 0; t < cd.num_tombstones; ++t) bm[cd.tombstones[t] >> 5] |= 1u << (cd.tombstones[t] & 31);
        tomb.upload(bm.data(), words, tomb_off[c]);
        n_used[c] = cd.used_capacity;
        n_live[c] = cd.num_embeddings;
        n_tomb[c] = cd.num_tombstones;
    }

    void create(const pdxg_level_desc& L, uint32_t dd, uint32_t vv, uint32_t hh, bool is_u8) {
        nc = L.num_clusters;
        d = dd;
        v = vv;
        h = hh;
        u8 = is_u8;
        require(nc > 0, "level has no clusters");
        require(L.clusters != nullptr, "level clusters missing");
        n_used.resize(nc);
        n_live.resize(nc);
        cap.resize(nc);
        stride.resize(nc);
        n_tomb.resize(nc);
        data_off.resize(nc);
        slot_off.resize(nc);
        tomb_off.resize(nc);
        data_bytes = total_slots = tomb_words = 0;
        for (uint32_t c = 0; c < nc; ++c) {
            const auto& cd = L.clusters[c];
            check_cluster(cd);
            cap[c] = cd.max_capacity;
            stride[c] = (cd.max_capacity + kStrideAlign - 1) / kStrideAlign * kStrideAlign;
            data_off[c] = data_bytes;
            data_bytes += ((size_t)stride[c] * d * elem() + 255) / 256 * 256;
            slot_off[c] = total_slots;
            total_slots += cd.max_capacity;
            tomb_off[c] = tomb_words;
            tomb_words += (cd.max_capacity + 31) / 32;
        }
        data.alloc(data_bytes);
        ids.alloc(total_slots);
        tomb.alloc(tomb_words);
        for (uint32_t c = 0; c < nc; ++c) upload_cluster(c, L.clusters[c]);
        d_data_off.alloc(nc);
        d_data_off.upload(data_off.data(), nc);
        d_slot_off.alloc(nc);
        d_slot_off.upload(slot_off.data(), nc);
        d_tomb_off.alloc(nc);
        d_tomb_off.upload(tomb_off.data(), nc);
        d_stride.alloc(nc);
        d_stride.upload(stride.data(), nc);
        d_n_used.alloc(nc);
        d_n_live.alloc(nc);
        d_n_tomb.alloc(nc);
        sync_counts();
        has_centroids = L.centroids != nullptr;
        if (has_centroids) {
            centroids.alloc((size_t)nc * d);
            centroids.upload(L.centroids, (size_t)nc * d);
        }
    }

    void sync_counts() {
        d_n_used.upload(n_used.data(), nc);
        d_n_live.upload(n_live.data(), nc);
        d_n_tomb.upload(n_tomb.data(), nc);
        tomb_total = 0;
        for (uint32_t x : n_tomb) tomb_total += x;
        std::vector<uint32_t> sz(n_used);
        for (uint32_t& x : sz) x = (x + 3u) & ~3u;
        std::sort(sz.begin(), sz.end(), std::greater<uint32_t>());
        top_prefix.assign(nc + 1, 0);
        for (uint32_t c = 0; c < nc; ++c) top_prefix[c + 1] = top_prefix[c] + sz[c];
    }

    LevelDev view(float scale_sq, float inv_scale_sq) const {
        LevelDev x;
        x.data = data.p;
        x.data_off = d_data_off.p;
        x.ids = ids.p;
        x.slot_off = d_slot_off.p;
        x.n_used = d_n_used.p;
        x.n_live = d_n_live.p;
        x.stride = d_stride.p;
        x.tomb = tomb.p;
        x.tomb_off = d_tomb_off.p;
        x.n_tomb = d_n_tomb.p;
        x.centroids = centroids.p;
        x.nc = nc;
        x.d = d;
        x.v = v;
        x.h = h;
        x.scale_sq = scale_sq;
        x.inv_scale_sq = inv_scale_sq;
        return x;
    }
};

struct LdsPlan {
    uint32_t bytes = 0;
    uint32_t o_q, o_dist, o_pos, o_h, o_hp, o_t, o_tp, o_o, o_op, o_ck, o_cp, o_red, o_tau, o_ring, o_r16;
};

LdsPlan plan_scan_lds(const Level& L, uint32_t kmax, uint32_t block, bool summ) {
    LdsPlan p{};
    uint32_t off = 0;
    auto take = [&](uint32_t bytes) {
        off = (off + 15) / 16 * 16;
        uint32_t o = off;
        off += bytes;
        return o;
    };
    const uint32_t nw = (L.v + 3) / 4 + L.h / 4;
    p.o_q = take(summ ? 16 : (L.u8 ? nw * 4 : L.d * 4));  // the replay reads the query from global memory
    const uint32_t capn = L.max_used();
    p.o_dist = take(capn * 4);
    p.o_pos = take(capn * 2);  // 16-bit slots
    p.o_h = take(kmax * 8);
    p.o_t = take(kmax * 8);
    p.o_o = take(kmax * 8);
    p.o_ck = take(block * 8);
    p.o_hp = take(kmax * 4);
    p.o_tp = take(kmax * 4);
    p.o_op = take(kmax * 4);
    p.o_cp = take(block * 4);
    p.o_red = take(64);
    // block thresholds; the replay adds its checkpoint counts and ends
    p.o_tau = take((L.h / kHDim + (L.v + kHDim - 1) / kHDim + 1 + (summ ? 2 * kMaxCheckpoints + 1 : 0)) * 4);
    p.o_ring = take(kRing * sizeof(ClusterMeta));
    p.o_r16 = take((L.d / 16 + 1) * 4);  // ratios at multiples of 16 dims
    p.bytes = (off + 15) / 16 * 16;
    return p;
}

}  // namespace

struct pdxg_workspace {
    int device = 0;
    const pdxg_index* index = nullptr;  // the index the workspace was planned for
    bool timing = false;
    DevBuf<unsigned long long> stage_bytes;  // [meso scan, data scan] of the last search when timing
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ev_summ = nullptr;  // end of the two-phase meso scan's first phase
    bool last_summ = false;        // the last timed search ran the two-phase meso scan
    ~pdxg_workspace() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (ev_summ) (void)hipEventDestroy(ev_summ);
    }
    void mark(int i, hipStream_t st) {
        if (timing) (void)hipEventRecord(ev[i], st);
    }
    uint32_t max_queries = 0, max_k = 0, max_nprobe = 0;
    uint32_t k0_max = 0, order_stride = 0;
    uint32_t block[2] = {0, 0};  // scan workgroup size per level (meso, data); 0 = heuristic
    // order of the per-query scans: 0 query order, 1 longest walk first (k_lpt_order), 2 longest
    // first through a persistent grid's work queue
    uint32_t order_mode = 1;
    DevBuf<uint32_t> perm, queue;
    DevBuf<float> qnorm, qrot, probe;
    DevBuf<uint8_t> qu8;
    DevBuf<uint32_t> k1, nv1, k0, nv0, order, ids0, cnt0;
    // two-phase meso scan (two-level indexes): partial-distance summaries of the visited meso
    // clusters. A query's region holds the padded sizes of the meso clusters it visits; `region`
    // bounds that for any query (the nv0_max largest clusters) and the buffer holds summ_cap
    // summarised vectors (max_queries x region unless a server gave the slot less).
    bool summ_ok = false;   // buffers allocated
    bool summ_use = true;   // plan: use them when the index allows
    uint32_t nv0_max = 0, nw = 0, ns = 0;
    unsigned long long region = 0, summ_cap = 0;
    // optional: the batch's per-query nprobe on the host (a server's staging), so that the batch's
    // summaries are checked against summ_cap by each query's own bound instead of `region`
    const uint32_t* host_nprobe = nullptr;
    DevBuf<uint32_t> m_cnt, m_cur, m_lstart, m_cstart, m_tstart, m_list, m_soff;
    DevBuf<unsigned long long> m_qbase;
    DevBuf<float> summ;
    size_t bytes() const {
        return perm.bytes() + queue.bytes() + qnorm.bytes() + qrot.bytes() + probe.bytes() + qu8.bytes() + k1.bytes() + nv1.bytes() + k0.bytes() +
               nv0.bytes() + order.bytes() + ids0.bytes() + cnt0.bytes() + m_cnt.bytes() + m_cur.bytes() +
               m_lstart.bytes() + m_cstart.bytes() + m_tstart.bytes() + m_list.bytes() + m_soff.bytes() +
               m_qbase.bytes() + summ.bytes();
    }
};

struct pdxg_index {
    int device = 0;
    uint32_t type = 0, d = 0, v = 0, h = 0;
    bool normalized = false;
    float qbase = 0.0f, qscale = 1.0f, qscale_sq = 1.0f, inv_qscale_sq = 1.0f;
    Level l1, l0;
    DevBuf<float> rotation, ratios;
    int max_lds = 65536;
    int num_cus = 1;
    std::mutex mu;  // serialises the host-buffer API (shared workspace)
    // searches plan and launch under a shared lock; pdxg_index_update_cluster takes it exclusively
    // and waits for the device before it rewrites cluster data, counts or tombstones
    std::shared_mutex rw;
    std::unique_ptr<pdxg_workspace> ws;
    DevBuf<float> h_queries, h_dists;
    DevBuf<uint32_t> h_k, h_np, h_ids, h_idists, h_counts;
    DevBuf<unsigned long long> h_bytes;
    DevBuf<uint8_t> sel;
    DevBuf<uint32_t> npass, f_order;
    DevBuf<float> f_probe;

    bool tree() const { return type == PDXG_TREE_F32 || type == PDXG_TREE_U8; }
    bool u8() const { return type == PDXG_U8 || type == PDXG_TREE_U8; }
};

namespace {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        HIP_CHECK(hipGetDevice(&prev));
        if (prev != dev) HIP_CHECK(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

uint32_t k0_capacity(const pdxg_index* ix, uint32_t max_nprobe) {
    if (!ix->tree()) return 0;
    return (max_nprobe == 0 || max_nprobe > ix->l1.nc) ? ix->l1.nc : max_nprobe;
}

// Summaries larger than this per workspace fall back to the pruned meso scan.
constexpr size_t kMaxSummaryBytes = 24ull << 30;

// Warmup checkpoints of a scan over v vertical dims (the cumulative fetch sizes clipped to v).
uint32_t warmup_checkpoints(uint32_t v) {
    static const uint32_t fetch[20] = {16, 16, 32, 32, 32, 32, 64, 64, 64, 64,
                                       128, 128, 128, 128, 256, 256, 512, 1024, 2048, 16384};
    uint32_t n = 0, end = 0;
    while (end < v) {
        require(n < 20, "vertical dimensions beyond the warmup schedule");
        end = std::min(end + fetch[n], v);
        ++n;
    }
    return n;
}

void plan_workspace(const pdxg_index* ix, uint32_t nq, uint32_t max_k, uint32_t max_nprobe, pdxg_workspace& ws) {
    ws.max_queries = nq;
    ws.max_k = max_k;
    ws.max_nprobe = max_nprobe;
    ws.k0_max = k0_capacity(ix, max_nprobe);
    ws.order_stride = ix->tree() ? ix->l0.nc : ((max_nprobe == 0 || max_nprobe > ix->l1.nc) ? ix->l1.nc : max_nprobe);
    ws.summ_ok = false;
    if (ix->tree()) {
        // meso clusters visited per query (k_setup): all of them, or half when nprobe < nlist / 2
        const uint32_t n0 = ix->l0.nc, n1 = ix->l1.nc;
        const bool halve = max_nprobe != 0 && max_nprobe < n1 / 2 && n0 / 2 != 0;
        ws.nv0_max = halve ? n0 / 2 : n0;
        ws.region = ix->l0.top_prefix[ws.nv0_max];
        ws.summ_cap = (unsigned long long)nq * ws.region;
        ws.nw = warmup_checkpoints(ix->l0.v);
        ws.ns = ws.nw + ix->l0.h / kHDim;
        ws.summ_ok = (size_t)ws.summ_cap * ws.ns * 4 <= kMaxSummaryBytes && ws.summ_cap < (1ull << 32) &&
                     (uint64_t)nq * ix->d < (1ull << 32);
    }
}

size_t workspace_bytes_for(const pdxg_index* ix, uint32_t nq, uint32_t max_k, uint32_t max_nprobe) {
    pdxg_workspace ws;
    plan_workspace(ix, nq, max_k, max_nprobe, ws);
    const size_t d = ix->d;
    const size_t nc_probe = ix->tree() ? ix->l0.nc : ix->l1.nc;
    size_t b = 0;
    b += 2 * (size_t)nq * d * 4;          // qnorm, qrot
    b += (size_t)nq * nc_probe * 4;        // probe
    b += ix->u8() ? (size_t)nq * d : 0;    // qu8
    b += 5 * (size_t)nq * 4;               // k1 nv1 k0 nv0 perm
    b += (size_t)nq * ws.order_stride * 4; // order
    b += (size_t)nq * ws.k0_max * 4 + (size_t)nq * 4;
    if (ws.summ_ok) {
        b += 5 * ((size_t)ix->l0.nc + 1) * 4;        // counts, cursors, entry / item / tail item offsets
        b += (size_t)nq * ws.nv0_max * 4;            // entries
        b += (size_t)nq * ws.order_stride * 4;       // summary offsets
        b += (size_t)(nq + 1) * 8;                   // region starts
        b += (size_t)ws.summ_cap * ws.ns * 4;        // summaries
    }
    return b;
}

void alloc_workspace(const pdxg_index* ix, pdxg_workspace& ws) {
    const size_t nq = ws.max_queries, d = ix->d;
    const size_t nc_probe = ix->tree() ? ix->l0.nc : ix->l1.nc;
    ws.qnorm.alloc(nq * d);
    ws.qrot.alloc(nq * d);
    ws.probe.alloc(nq * nc_probe);
    if (ix->u8()) ws.qu8.alloc(nq * d);
    ws.k1.alloc(nq);
    ws.nv1.alloc(nq);
    ws.k0.alloc(nq);
    ws.nv0.alloc(nq);
    ws.perm.alloc(nq);
    ws.queue.alloc(2);
    ws.order.alloc(nq * ws.order_stride);
    if (ix->tree()) {
        ws.ids0.alloc(nq * ws.k0_max);
        ws.cnt0.alloc(nq);
    }
    if (ws.summ_ok) {
        const size_t n0 = ix->l0.nc;
        ws.m_cnt.alloc(n0 + 1);
        ws.m_cur.alloc(n0 + 1);
        ws.m_lstart.alloc(n0 + 1);
        ws.m_cstart.alloc(n0 + 1);
        ws.m_tstart.alloc(n0 + 1);
        ws.m_list.alloc(nq * ws.nv0_max);
        ws.m_soff.alloc(nq * ws.order_stride);
        ws.m_qbase.alloc(nq + 1);
        ws.summ.alloc((size_t)ws.summ_cap * ws.ns);
    }
}

template <bool U8, bool FILT, uint32_t BLOCK, bool SUMM = false>
void launch_scan_block(const pdxg_index* ix, const Level& L, ScanArgs a, uint32_t kmax, hipStream_t st) {
    require(L.max_used() < kPosNone, "cluster too large for the scan's 16-bit slot lists", PDXG_ERR_UNSUPPORTED);
    LdsPlan p = plan_scan_lds(L, kmax, BLOCK, SUMM);
    require(p.bytes <= (uint32_t)ix->max_lds,
            "cluster capacity / k too large for the on-chip scan buffers (" + std::to_string(p.bytes) + " B LDS)",
            PDXG_ERR_UNSUPPORTED);
    a.o_q = p.o_q;
    a.o_dist = p.o_dist;
    a.o_pos = p.o_pos;
    a.o_h = p.o_h;
    a.o_hp = p.o_hp;
    a.o_t = p.o_t;
    a.o_tp = p.o_tp;
    a.o_o = p.o_o;
    a.o_op = p.o_op;
    a.o_ck = p.o_ck;
    a.o_cp = p.o_cp;
    a.o_red = p.o_red;
    a.o_tau = p.o_tau;
    a.o_ring = p.o_ring;
    a.o_r16 = p.o_r16;
    if constexpr (!FILT) {
        if (a.queue) {  // persistent: as many workgroups as are resident at once
            require(p.bytes + 64 <= (uint32_t)ix->max_lds, "no LDS left for the work-queue slot", PDXG_ERR_UNSUPPORTED);
            auto kern = k_scan<U8, FILT, BLOCK, SUMM, true>;
            if (p.bytes > 65536)
                HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.bytes));
            if (a.nq == 0) return;
            int per_cu = 0;
            HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), BLOCK,
                                                                   p.bytes));
            require(per_cu >= 1, "scan workgroup does not fit a CU", PDXG_ERR_UNSUPPORTED);
            const uint32_t grid = std::min<uint32_t>(a.nq, (uint32_t)per_cu * (uint32_t)ix->num_cus);
            hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), p.bytes, st, a);
            HIP_CHECK(hipGetLastError());
            return;
        }
    }
    require(!a.queue, "no persistent filtered scan", PDXG_ERR_UNSUPPORTED);
    auto kern = k_scan<U8, FILT, BLOCK, SUMM, false>;
    if (p.bytes > 65536)
        HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)p.bytes));
    if (a.nq == 0) return;
    hipLaunchKernelGGL(kern, dim3(a.nq), dim3(BLOCK), p.bytes, st, a);
    HIP_CHECK(hipGetLastError());
}

constexpr uint32_t kBlocks[5] = {64, 128, 256, 512, 1024};

// Workgroup size of a scan launch when none was tuned: about 4 slots per thread for batches that
// fill the GPU (at most 256 threads, so several queries share a CU), about 2 slots per thread
// for small batches (more loads in flight per query).
uint32_t heuristic_block(const Level& L, uint32_t nq) {
    uint64_t tot = 0;
    for (uint32_t x : L.n_used) tot += x;
    const uint32_t mean = (uint32_t)std::max<uint64_t>(1, tot / std::max(1u, L.nc));
    const bool small = nq < 64;
    uint32_t want = 64;
    while (want < (small ? mean / 2 : mean / 4) && want < (small ? 1024u : 256u)) want <<= 1;
    return want;
}

template <bool U8, bool FILT, bool SUMM = false>
void launch_scan(const pdxg_index* ix, const Level& L, ScanArgs a, uint32_t kmax, hipStream_t st, uint32_t block = 0) {
    if (block == 0) block = heuristic_block(L, a.nq);
    if constexpr (FILT) {  // filtered search: single queries, two sizes
        if (block >= 512) launch_scan_block<U8, FILT, 1024>(ix, L, a, kmax, st);
        else launch_scan_block<U8, FILT, 256>(ix, L, a, kmax, st);
    } else {
        switch (block) {
            case 64: launch_scan_block<U8, FILT, 64, SUMM>(ix, L, a, kmax, st); break;
            case 128: launch_scan_block<U8, FILT, 128, SUMM>(ix, L, a, kmax, st); break;
            case 256: launch_scan_block<U8, FILT, 256, SUMM>(ix, L, a, kmax, st); break;
            case 512: launch_scan_block<U8, FILT, 512, SUMM>(ix, L, a, kmax, st); break;
            default: launch_scan_block<U8, FILT, 1024, SUMM>(ix, L, a, kmax, st); break;
        }
    }
}

// Phase 1 of the two-phase meso scan: visit lists inverted per meso cluster, then the
// partial-distance summaries of every visited (query, meso member) pair.
// The two-phase meso scan runs when the workspace holds its buffers and the plan asks for it, the
// meso level has no tombstones (the replay does not mask) and the meso clusters a query may visit
// still fit the per-query summary region planned at workspace creation.
// Meso clusters a query visits (k_setup's rule, index.hpp:545-566), nprobe as the call clamps it.
uint32_t meso_visits(const pdxg_index* ix, uint32_t nprobe, uint32_t np_cap) {
    const uint32_t n0 = ix->l0.nc, n1 = ix->l1.nc;
    uint32_t np = nprobe;
    if (np_cap != 0 && (np == 0 || np > np_cap)) np = np_cap;
    const uint32_t np1 = np == 0 ? n1 : np;
    uint32_t np0 = n0;
    if (np1 < n1 / 2) np0 /= 2;
    return (np0 == 0 || np0 > n0) ? n0 : np0;
}

// Summarised vectors a query may need: the padded sizes of the largest meso clusters it can visit.
unsigned long long meso_summary_bound(const pdxg_index* ix, uint32_t nprobe, uint32_t np_cap) {
    return ix->l0.top_prefix[meso_visits(ix, nprobe, np_cap)];
}

bool summary_applies(const pdxg_index* ix, const pdxg_workspace& ws, uint32_t nq, uint32_t np_cap) {
    if (!(ws.summ_ok && ws.summ_use && ix->l0.tomb_total == 0)) return false;
    if (ws.host_nprobe) {  // the batch's own bound (a server admitted it by that bound)
        unsigned long long need = 0;
        for (uint32_t q = 0; q < nq; ++q) need += meso_summary_bound(ix, ws.host_nprobe[q], np_cap);
        return need <= ws.summ_cap;
    }
    return ix->l0.top_prefix[ws.nv0_max] <= ws.region && (unsigned long long)nq * ws.region <= ws.summ_cap;
}

void launch_meso_summary(const pdxg_index* ix, pdxg_workspace& ws, uint32_t nq, hipStream_t st) {
    const Level& L = ix->l0;
    MesoArgs m{};
    m.L = L.view(1.0f, 1.0f);
    m.qf = ws.qrot.p;
    m.order = ws.order.p;
    m.nvisit = ws.nv0.p;
    m.stride = ws.order_stride;
    m.nq = nq;
    m.nv_max = ws.nv0_max;
    m.cnt = ws.m_cnt.p;
    m.cur = ws.m_cur.p;
    m.lstart = ws.m_lstart.p;
    m.cstart = ws.m_cstart.p;
    m.tstart = ws.m_tstart.p;
    m.list = ws.m_list.p;
    m.soff = ws.m_soff.p;
    m.summ = ws.summ.p;
    m.summ_cap = ws.summ_cap;
    m.qbase = ws.m_qbase.p;
    m.nw = ws.nw;
    m.ns = ws.ns;
    m.pairs = ws.timing && ws.stage_bytes.p ? ws.stage_bytes.p + 2 : nullptr;
    require((uint64_t)nq * ws.order_stride < (1ull << 32), "too many queries x meso clusters for one search");
    hipLaunchKernelGGL(k_fill_u32, dim3((L.nc + 255) / 256), dim3(256), 0, st, ws.m_cnt.p, L.nc, 0u);
    const size_t pairs = (size_t)nq * ws.nv0_max;
    const dim3 gp((unsigned)((pairs + kListPer - 1) / kListPer));
    require(L.nc <= kListHistMax, "too many meso clusters for the visit-list histogram", PDXG_ERR_UNSUPPORTED);
    hipLaunchKernelGGL(k_meso_count, gp, dim3(kListThreads), L.nc * 4, st, m);
    hipLaunchKernelGGL(k_meso_offsets, dim3(1), dim3(kScanThreads), 0, st, m);
    hipLaunchKernelGGL(k_meso_scatter, gp, dim3(kListThreads), L.nc * 4, st, m);
    hipLaunchKernelGGL(k_meso_soff, dim3(nq), dim3(kWave), 0, st, m);
    hipLaunchKernelGGL(k_meso_qbase, dim3(1), dim3(kScanThreads), 0, st, m);
    // work items (counted on the device): sum over clusters of ceil(entries / kMesoQ) x chunks,
    // where a cluster's entries are at most nq and all entries at most `pairs`; the grid covers
    // this bound, one item per wave
    size_t chunks = 0, ch_max = 1;
    for (uint32_t c = 0; c < L.nc; ++c) {
        const size_t ch = std::max<size_t>(1, (L.n_used[c] + kWave - 1) / kWave);
        chunks += ch;
        ch_max = std::max(ch_max, ch);
    }
    const size_t items = std::min(((pairs + kMesoQ - 1) / kMesoQ + L.nc) * ch_max,
                                  ((size_t)nq + kMesoQ - 1) / kMesoQ * chunks);
    const size_t wpx = ((items + 7) / 8 + kMesoWaves - 1) / kMesoWaves;  // workgroups per XCD
    require(8 * wpx * kMesoWaves * kWave < (1ull << 32), "too many meso work items for one dispatch",
            PDXG_ERR_UNSUPPORTED);
    hipLaunchKernelGGL(k_meso_summary, dim3((uint32_t)(8 * wpx)), dim3(kMesoWaves * kWave), 0, st, m);
    // packed tail items (counted on the device): at most one per 8-entry group of the clusters
    // whose last chunk holds at most 32 vectors, each run as kTailParts waves
    size_t packed = 0;
    for (uint32_t c = 0; c < L.nc; ++c) packed += meso_tail_groups(L.n_used[c]) > 1 ? 1 : 0;
    if (packed) {
        const size_t titems = kTailParts * std::min((pairs + kMesoQ - 1) / kMesoQ + packed,
                                                    ((size_t)nq + kMesoQ - 1) / kMesoQ * packed);
        const size_t tpx = ((titems + 7) / 8 + kMesoWaves - 1) / kMesoWaves;
        require(8 * tpx * kMesoWaves * kWave < (1ull << 32), "too many meso tail items for one dispatch",
                PDXG_ERR_UNSUPPORTED);
        hipLaunchKernelGGL(k_meso_tails, dim3((uint32_t)(8 * tpx)), dim3(kMesoWaves * kWave), 0, st, m);
    }
    HIP_CHECK(hipGetLastError());
}

void launch_probe(const pdxg_index* ix, const Level& L, const float* qrot, uint32_t nq, const uint32_t* nvisit,
                  uint32_t* order, uint32_t order_stride, float* probe, hipStream_t st) {
    require(L.has_centroids, "level has no centroids on device", PDXG_ERR_UNSUPPORTED);
    require(L.nc <= kMaxProbeSort, "too many centroids for the on-chip probe sort", PDXG_ERR_UNSUPPORTED);
    if (nq == 0) return;
    // queries per workgroup: 8 while their LDS copy fits in 64 KB, fewer for small batches / large d
    const uint32_t ngy = (L.nc + kProbeCPB - 1) / kProbeCPB;
    if (nq >= 64 && 8u * ix->d * 4 <= 65536) {
        dim3 g1((nq + 7) / 8, ngy);
        hipLaunchKernelGGL(k_probe_dist<8>, g1, dim3(256), 8 * ix->d * 4, st, qrot, L.centroids.p, probe, ix->d, L.nc, nq);
    } else if (nq >= 16 && 4u * ix->d * 4 <= 65536) {
        dim3 g1((nq + 3) / 4, ngy);
        hipLaunchKernelGGL(k_probe_dist<4>, g1, dim3(256), 4 * ix->d * 4, st, qrot, L.centroids.p, probe, ix->d, L.nc, nq);
    } else {
        dim3 g1(nq, ngy);
        hipLaunchKernelGGL(k_probe_dist<1>, g1, dim3(256), ix->d * 4, st, qrot, L.centroids.p, probe, ix->d, L.nc, nq);
    }
    HIP_CHECK(hipGetLastError());
    uint32_t n2 = 1;
    while (n2 < L.nc) n2 <<= 1;
    if (n2 * 8 > 65536)
        HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_probe_select),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)(n2 * 8)));
    hipLaunchKernelGGL(k_probe_select, dim3(nq), dim3(256), n2 * 8, st, probe, L.nc, n2, nvisit, order, order_stride,
                       nq);
    HIP_CHECK(hipGetLastError());
}

void launch_preprocess(const pdxg_index* ix, pdxg_workspace& ws, const float* queries, uint32_t nq,
                       bool transformed, hipStream_t st) {
    if (nq == 0) return;
    const float* src = queries;
    if (!transformed && ix->normalized) {
        hipLaunchKernelGGL(k_normalize, dim3(nq), dim3(256), 0, st, queries, ws.qnorm.p, ix->d, nq);
        HIP_CHECK(hipGetLastError());
        src = ws.qnorm.p;
    }
    uint8_t* yq = ix->u8() ? ws.qu8.p : static_cast<uint8_t*>(nullptr);
    // large batches: register-blocked tiles, when they still give every CU two workgroups (at
    // d = 768 a batch of 1024 makes only 192 tiles, and the 16-query kernel's 768 run faster)
    const size_t tiles = (size_t)((ix->d + kRotBI - 1) / kRotBI) * ((nq + kRotBQ - 1) / kRotBQ);
    if (!transformed && nq >= kRotTiledMin && tiles >= 2 * (size_t)ix->num_cus) {
        dim3 g((ix->d + kRotBI - 1) / kRotBI, (nq + kRotBQ - 1) / kRotBQ);
        hipLaunchKernelGGL(k_rotate_tiled, g, dim3(256), 0, st, src, ix->rotation.p, ws.qrot.p, yq, ix->d, nq,
                           ix->qbase, ix->qscale);
    } else if (!transformed && nq <= kRotSmallQ && (size_t)ix->d * 4 <= 65536) {  // a few queries: wide
        dim3 g((ix->d + 63) / 64, nq);
        hipLaunchKernelGGL(k_rotate_small, g, dim3(64), ix->d * 4, st, src, ix->rotation.p, ws.qrot.p, yq, ix->d,
                           ix->qbase, ix->qscale);
    } else {
        dim3 g((ix->d + kRotTI - 1) / kRotTI, (nq + kRotTQ - 1) / kRotTQ);
        hipLaunchKernelGGL(k_rotate, g, dim3(256), 0, st, src, ix->rotation.p, ws.qrot.p, yq, ix->d, nq, ix->qbase,
                           ix->qscale, transformed ? 1 : 0);
    }
    HIP_CHECK(hipGetLastError());
}

// The full search sequence on device buffers.
void search_device(pdxg_index* ix, pdxg_workspace& ws, const pdxg_search_args& a, hipStream_t st) {
    std::shared_lock<std::shared_mutex> reading(ix->rw);
    const uint32_t nq = a.num_queries;
    require(nq <= ws.max_queries, "more queries than the workspace holds");
    require(a.k_max >= 1 && a.k_max <= ws.max_k, "k_max must be in [1, workspace max_k]");
    require(a.k_max <= kMaxK, "k above " + std::to_string(kMaxK) + " is not supported on the GPU path",
            PDXG_ERR_UNSUPPORTED);
    require(a.out_stride >= a.k_max, "out_stride must be >= k_max");
    require(a.out_ids && a.out_counts, "out_ids / out_counts required");
    const uint32_t np_cap = (a.nprobe_max == 0) ? 0 : a.nprobe_max;
    require(ws.max_nprobe == 0 || (np_cap != 0 && np_cap <= ws.max_nprobe), "nprobe_max exceeds the workspace");
    if (!a.k) require(a.k_all >= 1, "k must be >= 1");
    if (!a.nprobe) require(a.nprobe_all == np_cap || np_cap == 0 || a.nprobe_all <= np_cap, "nprobe_all > nprobe_max");
    if (nq == 0) return;

    ws.mark(0, st);
    unsigned long long* sb = nullptr;
    if (ws.timing) {
        if (!ws.stage_bytes.p) ws.stage_bytes.alloc(3);
        sb = ws.stage_bytes.p;
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, sb, 3u, 0ull);
        HIP_CHECK(hipGetLastError());
    }
    launch_preprocess(ix, ws, a.queries, nq, a.queries_transformed != 0, st);

    SetupArgs sa{};
    sa.k_in = a.k;
    sa.nprobe_in = a.nprobe;
    sa.k_all = a.k_all;
    sa.nprobe_all = a.nprobe_all;
    sa.nq = nq;
    sa.tree = ix->tree();
    sa.nc1 = ix->l1.nc;
    sa.nc0 = ix->tree() ? ix->l0.nc : 0;
    sa.k_cap = a.k_max;
    sa.np_cap = np_cap;
    sa.order_cap = ws.order_stride;
    sa.k0_cap = ws.k0_max;
    sa.k1 = ws.k1.p;
    sa.nvisit1 = ws.nv1.p;
    sa.k0 = ws.k0.p;
    sa.nvisit0 = ws.nv0.p;
    sa.bytes = reinterpret_cast<unsigned long long*>(a.out_bytes);
    sa.probe_bytes = (uint64_t)(ix->tree() ? ix->l0.nc : ix->l1.nc) * ix->d * 4;  // centroids streamed per query
    hipLaunchKernelGGL(k_setup, dim3((nq + 255) / 256), dim3(256), 0, st, sa);
    HIP_CHECK(hipGetLastError());

    const float* qrot = ws.qrot.p;

    ScanArgs base{};
    if (ws.order_mode >= 1 && nq > 1) {  // the data-level visit count orders both levels' scans
        hipLaunchKernelGGL(k_lpt_order, dim3(1), dim3(kLptThreads), 0, st, ws.nv1.p, nq, ws.perm.p);
        HIP_CHECK(hipGetLastError());
        base.perm = ws.perm.p;
    }
    uint32_t* queues[2] = {nullptr, nullptr};  // work-queue counters of the meso and data scans
    if (ws.order_mode == 2) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, ws.queue.p, 2u, 0u);
        HIP_CHECK(hipGetLastError());
        queues[0] = ws.queue.p;
        queues[1] = ws.queue.p + 1;
    }
    base.ratios = ix->ratios.p;
    base.out_bytes = reinterpret_cast<unsigned long long*>(a.out_bytes);
    base.nq = nq;

    ws.mark(1, st);
    if (!ix->tree()) {
        launch_probe(ix, ix->l1, qrot, nq, ws.nv1.p, ws.order.p, ws.order_stride, ws.probe.p, st);
        ws.mark(2, st);
        ws.mark(3, st);
        ScanArgs s = base;
        s.L = ix->l1.view(ix->qscale_sq, ix->inv_qscale_sq);
        s.qf = qrot;
        s.qu = ws.qu8.p;
        s.order = ws.order.p;
        s.order_stride = ws.order_stride;
        s.nvisit = ws.nv1.p;
        s.order_count = nullptr;
        s.kq = ws.k1.p;
        s.out_ids = a.out_ids;
        s.out_dists = a.out_distances;
        s.out_idists = a.out_int_distances;
        s.out_counts = a.out_counts;
        s.out_stride = a.out_stride;
        s.stage_bytes = sb ? sb + 1 : nullptr;
        s.queue = queues[1];
        if (ix->u8()) launch_scan<true, false>(ix, ix->l1, s, a.k_max, st, ws.block[1]);
        else launch_scan<false, false>(ix, ix->l1, s, a.k_max, st, ws.block[1]);
        ws.mark(4, st);
    } else {
        require(ws.k0_max <= kMaxK,
                "nprobe above " + std::to_string(kMaxK) + " (or 0 with that many clusters) is not supported on the "
                "GPU path", PDXG_ERR_UNSUPPORTED);
        launch_probe(ix, ix->l0, qrot, nq, ws.nv0.p, ws.order.p, ws.order_stride, ws.probe.p, st);
        ws.mark(2, st);
        ScanArgs s0 = base;
        s0.L = ix->l0.view(1.0f, 1.0f);
        s0.qf = qrot;
        s0.order = ws.order.p;
        s0.order_stride = ws.order_stride;
        s0.nvisit = ws.nv0.p;
        s0.kq = ws.k0.p;
        s0.out_ids = ws.ids0.p;
        s0.out_dists = nullptr;
        s0.out_idists = nullptr;
        s0.out_counts = ws.cnt0.p;
        s0.out_stride = ws.k0_max;
        s0.stage_bytes = sb;
        s0.queue = queues[0];
        const bool summ = summary_applies(ix, ws, nq, np_cap);
        ws.last_summ = summ;
        if (summ) {
            launch_meso_summary(ix, ws, nq, st);
            if (ws.timing) (void)hipEventRecord(ws.ev_summ, st);
            s0.summ = ws.summ.p;
            s0.soff = ws.m_soff.p;
            s0.qbase = ws.m_qbase.p;
            s0.nw = ws.nw;
            s0.ns = ws.ns;
            launch_scan<false, false, true>(ix, ix->l0, s0, ws.k0_max, st, ws.block[0]);
        } else {
            launch_scan<false, false>(ix, ix->l0, s0, ws.k0_max, st, ws.block[0]);
        }
        ws.mark(3, st);
        ScanArgs s1 = base;
        s1.L = ix->l1.view(ix->qscale_sq, ix->inv_qscale_sq);
        s1.qf = qrot;
        s1.qu = ws.qu8.p;
        s1.order = ws.ids0.p;
        s1.order_stride = ws.k0_max;
        s1.nvisit = ws.nv1.p;
        s1.order_count = ws.cnt0.p;
        s1.kq = ws.k1.p;
        s1.out_ids = a.out_ids;
        s1.out_dists = a.out_distances;
        s1.out_idists = a.out_int_distances;
        s1.out_counts = a.out_counts;
        s1.out_stride = a.out_stride;
        s1.stage_bytes = sb ? sb + 1 : nullptr;
        s1.queue = queues[1];
        if (ix->u8()) launch_scan<true, false>(ix, ix->l1, s1, a.k_max, st, ws.block[1]);
        else launch_scan<false, false>(ix, ix->l1, s1, a.k_max, st, ws.block[1]);
        ws.mark(4, st);
    }
}

pdxg_workspace* host_workspace(pdxg_index* ix, uint32_t nq, uint32_t max_k, uint32_t max_nprobe) {
    auto& ws = ix->ws;
    pdxg_workspace want;
    plan_workspace(ix, nq, max_k, max_nprobe, want);
    bool ok = ws && ws->max_queries >= nq && ws->max_k >= max_k && ws->k0_max >= want.k0_max &&
              ws->order_stride >= want.order_stride &&
              (ws->max_nprobe == 0 || (max_nprobe != 0 && max_nprobe <= ws->max_nprobe));
    if (!ok) {
        uint32_t mq = std::max(nq, ws ? ws->max_queries : 0u);
        uint32_t mk = std::max(max_k, ws ? ws->max_k : 0u);
        uint32_t mp = (max_nprobe == 0 || (ws && ws->max_nprobe == 0)) ? 0u
                                                                       : std::max(max_nprobe, ws ? ws->max_nprobe : 0u);
        ws.reset(new pdxg_workspace);
        ws->device = ix->device;
        ws->index = ix;
        plan_workspace(ix, mq, mk, mp, *ws);
        alloc_workspace(ix, *ws);
    }
    return ws.get();
}

// Runs the given search with every scan workgroup size, level by level, and keeps the fastest
// (for the meso level also the two-phase scan against the pruned one).
void autotune_workspace(pdxg_index* ix, pdxg_workspace* ws, const pdxg_search_args* args, hipStream_t st) {
    {
        const bool was_timing = ws->timing;
        if (!was_timing) {
            for (auto& e : ws->ev)
                if (!e) HIP_CHECK(hipEventCreate(&e));
            if (!ws->ev_summ) HIP_CHECK(hipEventCreate(&ws->ev_summ));
            ws->timing = true;
        }
        // tune the meso level first (two-level indexes), then the data level; each candidate runs
        // twice and the second run's stage time counts (results do not depend on the block size)
        // (the meso level also tries the two-phase scan against the pruned one)
        for (int lvl = ix->tree() ? 0 : 1; lvl <= 1; ++lvl) {
            float best = 1e30f;
            uint32_t best_b = 0;
            bool best_summ = false;
            const bool summ_before = ws->summ_use;
            for (int summ = (lvl == 0 && ws->summ_ok) ? 1 : 0; summ >= 0; --summ) {
                if (lvl == 0) ws->summ_use = summ != 0;
                for (uint32_t b : kBlocks) {
                    ws->block[lvl] = b;
                    float ms[4] = {0, 0, 0, 0};
                    bool ok = true;
                    for (int rep = 0; rep < 2 && ok; ++rep) {
                        try {
                            search_device(ix, *ws, *args, st);
                            HIP_CHECK(hipEventSynchronize(ws->ev[4]));
                            for (int i = 0; i < 4; ++i)
                                HIP_CHECK(hipEventElapsedTime(&ms[i], ws->ev[i], ws->ev[i + 1]));
                        } catch (const Error& e) {
                            if (e.code != PDXG_ERR_UNSUPPORTED) throw;
                            ok = false;  // e.g. the LDS plan of this block size does not fit
                        }
                    }
                    const float t = lvl == 0 ? ms[2] : ms[3];
                    if (ok && t < best) {
                        best = t;
                        best_b = b;
                        best_summ = summ != 0;
                    }
                }
            }
            require(best_b != 0, "no scan block size fits this index", PDXG_ERR_UNSUPPORTED);
            ws->block[lvl] = best_b;
            ws->summ_use = lvl == 0 ? best_summ : summ_before;
        }
        // order of the per-query scans (both levels): query order, longest walk first, longest
        // first through a persistent work queue
        // (the modes differ by a few percent at most: each keeps the best of three runs after a
        // warm-up run)
        float best_t = 1e30f;
        uint32_t best_mode = 1;
        for (uint32_t mode = 0; mode < 3; ++mode) {
            ws->order_mode = mode;
            float ms = 1e30f;
            bool ok = true;
            for (int rep = 0; rep < 4 && ok; ++rep) {
                try {
                    search_device(ix, *ws, *args, st);
                    HIP_CHECK(hipEventSynchronize(ws->ev[4]));
                    float t = 0.0f;
                    HIP_CHECK(hipEventElapsedTime(&t, ws->ev[2], ws->ev[4]));
                    if (rep > 0) ms = std::min(ms, t);
                } catch (const Error& e) {
                    if (e.code != PDXG_ERR_UNSUPPORTED) throw;
                    ok = false;
                }
            }
            if (ok && ms < best_t) {
                best_t = ms;
                best_mode = mode;
            }
        }
        ws->order_mode = best_mode;
        ws->timing = was_timing;
    }
}

}  // namespace

extern "C" {

const char* pdxg_last_error(void) { return g_last_error.c_str(); }

pdxg_status pdxg_device_info(int device, char* buf, size_t len) {
    return guarded([&] {
        hipDeviceProp_t p;
        HIP_CHECK(hipGetDeviceProperties(&p, device));
        int lds = 0;
        HIP_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
        if (buf && len) snprintf(buf, len, "%s %s %d CUs %d B LDS", p.gcnArchName, p.name, p.multiProcessorCount, lds);
    });
}

pdxg_status pdxg_index_create(int device, const pdxg_index_desc* desc, pdxg_index** out) {
    return guarded([&] {
        require(desc && out, "null argument");
        *out = nullptr;
        require(desc->index_type <= PDXG_TREE_U8, "unknown index type");
        require(desc->num_dimensions > 0, "num_dimensions must be positive");
        require(desc->num_vertical_dimensions + desc->num_horizontal_dimensions == desc->num_dimensions,
                "vertical + horizontal dimensions must equal num_dimensions");
        require(desc->num_horizontal_dimensions % kHDim == 0, "horizontal dimensions must be a multiple of 64");
        require(desc->rotation != nullptr, "rotation matrix required (dense d x d)");
        DeviceGuard g(device);
        std::unique_ptr<pdxg_index> ix(new pdxg_index);
        ix->device = device;
        ix->type = desc->index_type;
        ix->d = desc->num_dimensions;
        ix->v = desc->num_vertical_dimensions;
        ix->h = desc->num_horizontal_dimensions;
        ix->normalized = desc->is_normalized != 0;
        if (ix->u8()) {
            ix->qbase = desc->quantization_base;
            ix->qscale = desc->quantization_scale;
            ix->qscale_sq = ix->qscale * ix->qscale;
            ix->inv_qscale_sq = 1.0f / ix->qscale_sq;
        }
        HIP_CHECK(hipDeviceGetAttribute(&ix->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
        HIP_CHECK(hipDeviceGetAttribute(&ix->num_cus, hipDeviceAttributeMultiprocessorCount, device));
        ix->l1.create(desc->l1, ix->d, ix->v, ix->h, ix->u8());
        if (ix->tree()) {
            // the L0 search returns these ids as the L1 visit order: they must name L1 clusters
            for (uint32_t c = 0; c < desc->l0.num_clusters; ++c) {
                const auto& cd = desc->l0.clusters[c];
                for (uint32_t s = 0; s < cd.used_capacity; ++s)
                    require(cd.ids[s] < desc->l1.num_clusters, "meso-cluster member id is not an L1 cluster");
            }
            ix->l0.create(desc->l0, ix->d, ix->v, ix->h, false);
            require(ix->l0.has_centroids, "two-level index requires the meso-cluster centroids");
        } else {
            require(ix->l1.has_centroids, "IVF index requires centroids");
        }
        ix->rotation.alloc((size_t)ix->d * ix->d);
        ix->rotation.upload(desc->rotation, (size_t)ix->d * ix->d);
        std::vector<float> r(ix->d);
        for (uint32_t i = 0; i < ix->d; ++i) r[i] = ratio_at(i, ix->d);
        ix->ratios.alloc(ix->d);
        ix->ratios.upload(r.data(), ix->d);
        HIP_CHECK(hipDeviceSynchronize());
        *out = ix.release();
    });
}

pdxg_status pdxg_index_destroy(pdxg_index* index) {
    return guarded([&] {
        if (!index) return;
        DeviceGuard g(index->device);
        delete index;
    });
}

size_t pdxg_index_device_bytes(const pdxg_index* ix) {
    if (!ix) return 0;
    return ix->l1.device_bytes() + ix->l0.device_bytes() + ix->rotation.bytes() + ix->ratios.bytes();
}

pdxg_status pdxg_index_update_cluster(pdxg_index* ix, uint32_t level, uint32_t cluster_id,
                                      const pdxg_cluster_desc* cd) {
    return guarded([&] {
        require(ix && cd, "null argument");
        require(level <= 1, "level must be 0 or 1");
        require(level == 1 || ix->tree(), "flat indexes have no level 0");
        DeviceGuard g(ix->device);
        std::lock_guard<std::mutex> lk(ix->mu);
        std::unique_lock<std::shared_mutex> writing(ix->rw);
        Level& L = level == 1 ? ix->l1 : ix->l0;
        require(cluster_id < L.nc, "cluster id out of range");
        require(cd->max_capacity == L.cap[cluster_id], "max_capacity change requires re-creating the index");
        L.check_cluster(*cd);
        // searches already launched (on any stream, e.g. a server's slots) read this cluster
        HIP_CHECK(hipDeviceSynchronize());
        L.upload_cluster(cluster_id, *cd);
        L.sync_counts();
        HIP_CHECK(hipDeviceSynchronize());
    });
}

size_t pdxg_workspace_bytes(const pdxg_index* index, uint32_t max_queries, uint32_t max_k, uint32_t max_nprobe) {
    if (!index) return 0;
    return workspace_bytes_for(index, max_queries, max_k, max_nprobe);
}

pdxg_status pdxg_workspace_create(pdxg_index* ix, uint32_t max_queries, uint32_t max_k, uint32_t max_nprobe,
                                  pdxg_workspace** out) {
    return guarded([&] {
        require(ix && out, "null argument");
        require(max_queries > 0 && max_k > 0, "max_queries and max_k must be positive");
        *out = nullptr;
        DeviceGuard g(ix->device);
        std::unique_ptr<pdxg_workspace> ws(new pdxg_workspace);
        ws->device = ix->device;
        ws->index = ix;
        plan_workspace(ix, max_queries, max_k, max_nprobe, *ws);
        alloc_workspace(ix, *ws);
        *out = ws.release();
    });
}

pdxg_status pdxg_workspace_destroy(pdxg_workspace* ws) {
    return guarded([&] {
        if (!ws) return;
        DeviceGuard g(ws->device);
        delete ws;
    });
}

pdxg_status pdxg_workspace_enable_timing(pdxg_workspace* ws, int enable) {
    return guarded([&] {
        require(ws != nullptr, "null argument");
        DeviceGuard g(ws->device);
        if (enable) {
            for (auto& e : ws->ev)
                if (!e) HIP_CHECK(hipEventCreate(&e));
            if (!ws->ev_summ) HIP_CHECK(hipEventCreate(&ws->ev_summ));
        }
        ws->timing = enable != 0;
    });
}

pdxg_status pdxg_workspace_stage_times(pdxg_workspace* ws, float* out) {
    return guarded([&] {
        require(ws && out, "null argument");
        require(ws->timing, "timing is not enabled on this workspace");
        DeviceGuard g(ws->device);
        HIP_CHECK(hipEventSynchronize(ws->ev[4]));
        for (int i = 0; i < 4; ++i) HIP_CHECK(hipEventElapsedTime(&out[i], ws->ev[i], ws->ev[i + 1]));
    });
}

pdxg_status pdxg_workspace_meso_times(pdxg_workspace* ws, float* out) {
    return guarded([&] {
        require(ws && out, "null argument");
        require(ws->timing, "timing is not enabled on this workspace");
        DeviceGuard g(ws->device);
        HIP_CHECK(hipEventSynchronize(ws->ev[4]));
        out[0] = out[1] = 0.0f;
        if (ws->last_summ) {
            HIP_CHECK(hipEventElapsedTime(&out[0], ws->ev[2], ws->ev_summ));
            HIP_CHECK(hipEventElapsedTime(&out[1], ws->ev_summ, ws->ev[3]));
        } else {
            HIP_CHECK(hipEventElapsedTime(&out[1], ws->ev[2], ws->ev[3]));
        }
    });
}

pdxg_status pdxg_workspace_stage_bytes(pdxg_workspace* ws, uint64_t* out) {
    return guarded([&] {
        require(ws && out, "null argument");
        require(ws->timing && ws->stage_bytes.p, "timing is not enabled on this workspace");
        DeviceGuard g(ws->device);
        HIP_CHECK(hipEventSynchronize(ws->ev[4]));
        HIP_CHECK(hipMemcpy(out, ws->stage_bytes.p, 16, hipMemcpyDeviceToHost));
    });
}

pdxg_status pdxg_workspace_meso_pairs(pdxg_workspace* ws, uint64_t* out) {
    return guarded([&] {
        require(ws && out, "null argument");
        require(ws->timing && ws->stage_bytes.p, "timing is not enabled on this workspace");
        DeviceGuard g(ws->device);
        HIP_CHECK(hipEventSynchronize(ws->ev[4]));
        *out = 0;
        if (ws->last_summ) HIP_CHECK(hipMemcpy(out, ws->stage_bytes.p + 2, 8, hipMemcpyDeviceToHost));
    });
}

pdxg_status pdxg_workspace_set_plan(pdxg_workspace* ws, uint32_t block_meso, uint32_t block_data) {
    return guarded([&] {
        require(ws != nullptr, "null argument");
        auto ok = [](uint32_t b) { return b == 0 || b == 64 || b == 128 || b == 256 || b == 512 || b == 1024; };
        require(ok(block_meso) && ok(block_data), "scan block sizes must be 0, 64, 128, 256, 512 or 1024");
        ws->block[0] = block_meso;
        ws->block[1] = block_data;
    });
}

pdxg_status pdxg_workspace_set_meso_summary(pdxg_workspace* ws, int enable) {
    return guarded([&] {
        require(ws != nullptr, "null argument");
        ws->summ_use = enable != 0;
    });
}

pdxg_status pdxg_workspace_set_order(pdxg_workspace* ws, int mode) {
    return guarded([&] {
        require(ws, "null argument");
        require(mode >= 0 && mode <= 2, "order mode must be 0, 1 or 2");
        ws->order_mode = static_cast<uint32_t>(mode);
    });
}

int pdxg_workspace_order(const pdxg_workspace* ws) { return ws ? static_cast<int>(ws->order_mode) : -1; }

int pdxg_workspace_meso_summary(const pdxg_workspace* ws) {
    return ws && ws->index && summary_applies(ws->index, *ws, ws->max_queries, ws->max_nprobe) ? 1 : 0;
}

pdxg_status pdxg_workspace_get_plan(pdxg_workspace* ws, uint32_t* out2) {
    return guarded([&] {
        require(ws && out2, "null argument");
        out2[0] = ws->block[0];
        out2[1] = ws->block[1];
    });
}

pdxg_status pdxg_workspace_autotune(pdxg_index* ix, pdxg_workspace* ws, const pdxg_search_args* args, void* stream) {
    return guarded([&] {
        require(ix && ws && args, "null argument");
        DeviceGuard g(ix->device);
        autotune_workspace(ix, ws, args, static_cast<hipStream_t>(stream));
    });
}

pdxg_status pdxg_search_device(pdxg_index* ix, pdxg_workspace* ws, const pdxg_search_args* args, void* stream) {
    return guarded([&] {
        require(ix && ws && args, "null argument");
        DeviceGuard g(ix->device);
        search_device(ix, *ws, *args, static_cast<hipStream_t>(stream));
    });
}

// ---------------------------------------------------------------------------------------------
// Search without allocation: the per-call workspace of pdxg_search_device carved out of one
// fixed device array. Each device holds its own copy of `scratch` (loaded with the code object)
// until the process ends; host code reaches it only through the address hipGetSymbolAddress
// returns on that device.
// ---------------------------------------------------------------------------------------------
namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on `device`, looked up once per device; the device must be current.
unsigned char* scratch_on(int device) {
    require(device >= 0 && device < kMaxScratchDevices, "device ordinal out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr)
            throw Error(PDXG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    return static_cast<unsigned char*>(g_scratch_ptr[device]);
}

// Points the workspace's buffers (as alloc_workspace sizes them, without the two-phase meso
// summaries) into [base, base + cap); returns the bytes the layout needs (views past `cap` are
// never used: the caller rejects the call first).
size_t carve_workspace(const pdxg_index* ix, pdxg_workspace& ws, bool need_qnorm, unsigned char* base) {
    size_t off = 0;
    auto take = [&](auto& buf, size_t count) {
        using T = typename std::remove_reference<decltype(*buf.p)>::type;
        off = (off + 255) & ~size_t(255);
        count = std::max<size_t>(count, 1);
        buf.view(reinterpret_cast<T*>(base + off), count);
        off += count * sizeof(T);
    };
    const size_t nq = ws.max_queries, d = ix->d;
    const size_t nc_probe = ix->tree() ? ix->l0.nc : ix->l1.nc;
    if (need_qnorm) take(ws.qnorm, nq * d);
    take(ws.qrot, nq * d);
    take(ws.probe, nq * nc_probe);
    if (ix->u8()) take(ws.qu8, nq * d);
    take(ws.k1, nq);
    take(ws.nv1, nq);
    take(ws.k0, nq);
    take(ws.nv0, nq);
    take(ws.perm, nq);
    take(ws.queue, 2);
    take(ws.order, nq * ws.order_stride);
    if (ix->tree()) {
        take(ws.ids0, nq * ws.k0_max);
        take(ws.cnt0, nq);
    }
    return off;
}

}  // namespace

pdxg_status pdxg_search_device_noalloc(pdxg_index* ix, const pdxg_search_args* args, void* stream) {
    return guarded([&] {
        require(ix && args, "null argument");
        DeviceGuard g(ix->device);
        unsigned char* base = scratch_on(ix->device);
        const pdxg_search_args& a = *args;
        require(a.k_max >= 1, "k_max must be >= 1");
        pdxg_workspace ws;  // host-side description only: every buffer is a view into `scratch`
        ws.device = ix->device;
        ws.index = ix;
        plan_workspace(ix, std::max(a.num_queries, 1u), a.k_max, a.nprobe_max, ws);
        ws.summ_ok = false;  // the meso level runs the pruned scan (its summaries need far more room)
        const bool need_qnorm = ix->normalized && !a.queries_transformed;
        const size_t need = carve_workspace(ix, ws, need_qnorm, base);
        require(need <= kScratchBytes,
                "this search needs " + std::to_string(need) + " bytes of scratch; pdxg_search_device_noalloc holds " +
                    std::to_string(kScratchBytes) + " (use fewer queries per call, or pdxg_search_device)",
                PDXG_ERR_UNSUPPORTED);
        // every buffer is written before it is read within the call (the work-queue counters by a
        // fill kernel), so the previous call's contents never leak into this one
        search_device(ix, ws, a, static_cast<hipStream_t>(stream));
    });
}

pdxg_status pdxg_search(pdxg_index* ix, const pdxg_search_args* args) {
    return guarded([&] {
        require(ix && args, "null argument");
        DeviceGuard g(ix->device);
        std::lock_guard<std::mutex> lk(ix->mu);
        const pdxg_search_args& a = *args;
        const uint32_t nq = a.num_queries;
        require(a.queries || nq == 0, "queries required");
        uint32_t kmax = a.k_max;
        if (a.k) {
            for (uint32_t q = 0; q < nq; ++q) require(a.k[q] >= 1, "k must be >= 1");
        }
        require(kmax >= 1, "k_max must be >= 1");
        if (nq == 0) return;
        pdxg_workspace* ws = host_workspace(ix, nq, kmax, a.nprobe_max);
        const size_t d = ix->d;
        ix->h_queries.ensure((size_t)nq * d);
        ix->h_queries.upload(a.queries, (size_t)nq * d);
        pdxg_search_args da = a;
        da.queries = ix->h_queries.p;
        if (a.k) {
            ix->h_k.ensure(nq);
            ix->h_k.upload(a.k, nq);
            da.k = ix->h_k.p;
        }
        if (a.nprobe) {
            ix->h_np.ensure(nq);
            ix->h_np.upload(a.nprobe, nq);
            da.nprobe = ix->h_np.p;
        }
        const size_t no = (size_t)nq * a.out_stride;
        ix->h_ids.ensure(no);
        ix->h_dists.ensure(no);
        ix->h_counts.ensure(nq);
        da.out_ids = ix->h_ids.p;
        da.out_distances = ix->h_dists.p;
        da.out_counts = ix->h_counts.p;
        da.out_int_distances = nullptr;
        if (a.out_int_distances) {
            ix->h_idists.ensure(no);
            da.out_int_distances = ix->h_idists.p;
        }
        da.out_bytes = nullptr;
        if (a.out_bytes) {
            ix->h_bytes.ensure(nq);
            da.out_bytes = reinterpret_cast<uint64_t*>(ix->h_bytes.p);
        }
        search_device(ix, *ws, da, nullptr);
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(a.out_counts, ix->h_counts.p, nq * 4, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(a.out_ids, ix->h_ids.p, no * 4, hipMemcpyDeviceToHost));
        if (a.out_distances) HIP_CHECK(hipMemcpy(a.out_distances, ix->h_dists.p, no * 4, hipMemcpyDeviceToHost));
        if (a.out_int_distances)
            HIP_CHECK(hipMemcpy(a.out_int_distances, ix->h_idists.p, no * 4, hipMemcpyDeviceToHost));
        if (a.out_bytes) HIP_CHECK(hipMemcpy(a.out_bytes, ix->h_bytes.p, nq * 8, hipMemcpyDeviceToHost));
    });
}

pdxg_status pdxg_preprocess(pdxg_index* ix, const float* queries, uint32_t nq, uint32_t transformed,
                            float* out_rot, uint8_t* out_q) {
    return guarded([&] {
        require(ix && queries && out_rot, "null argument");
        DeviceGuard g(ix->device);
        std::lock_guard<std::mutex> lk(ix->mu);
        if (nq == 0) return;
        pdxg_workspace* ws = host_workspace(ix, nq, ix->ws ? ix->ws->max_k : 1u, ix->ws ? ix->ws->max_nprobe : 1u);
        ix->h_queries.ensure((size_t)nq * ix->d);
        ix->h_queries.upload(queries, (size_t)nq * ix->d);
        launch_preprocess(ix, *ws, ix->h_queries.p, nq, transformed != 0, nullptr);
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(out_rot, ws->qrot.p, (size_t)nq * ix->d * 4, hipMemcpyDeviceToHost));
        if (out_q && ix->u8()) HIP_CHECK(hipMemcpy(out_q, ws->qu8.p, (size_t)nq * ix->d, hipMemcpyDeviceToHost));
    });
}

pdxg_status pdxg_search_meso(pdxg_index* ix, const float* queries, uint32_t nq, uint32_t k, uint32_t nprobe,
                             uint32_t* out_ids, float* out_dists, uint32_t* out_counts) {
    return guarded([&] {
        require(ix && queries && out_ids && out_counts, "null argument");
        require(ix->tree(), "meso search needs a two-level index");
        require(k >= 1 && k <= kMaxK, "k must be in [1, 2048]", PDXG_ERR_UNSUPPORTED);
        DeviceGuard g(ix->device);
        std::lock_guard<std::mutex> lk(ix->mu);
        if (nq == 0) return;
        pdxg_workspace* ws = host_workspace(ix, nq, ix->ws ? ix->ws->max_k : 1u, ix->ws ? ix->ws->max_nprobe : 1u);
        ix->h_queries.ensure((size_t)nq * ix->d);
        ix->h_queries.upload(queries, (size_t)nq * ix->d);
        launch_preprocess(ix, *ws, ix->h_queries.p, nq, true, nullptr);
        const uint32_t n0 = ix->l0.nc;
        const uint32_t nv = (nprobe == 0 || nprobe > n0) ? n0 : nprobe;  // searcher.hpp:614-616
        std::vector<uint32_t> hk(nq, k), hnv(nq, nv);
        ws->k0.upload(hk.data(), nq);
        ws->nv0.upload(hnv.data(), nq);
        launch_probe(ix, ix->l0, ws->qrot.p, nq, ws->nv0.p, ws->order.p, ws->order_stride, ws->probe.p, nullptr);
        const size_t no = (size_t)nq * k;
        ix->h_ids.ensure(no);
        ix->h_dists.ensure(no);
        ix->h_counts.ensure(nq);
        ScanArgs s{};
        s.L = ix->l0.view(1.0f, 1.0f);
        s.ratios = ix->ratios.p;
        s.qf = ws->qrot.p;
        s.order = ws->order.p;
        s.order_stride = ws->order_stride;
        s.nvisit = ws->nv0.p;
        s.kq = ws->k0.p;
        s.out_ids = ix->h_ids.p;
        s.out_dists = ix->h_dists.p;
        s.out_counts = ix->h_counts.p;
        s.out_stride = k;
        s.nq = nq;
        launch_scan<false, false>(ix, ix->l0, s, k, nullptr);
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(out_counts, ix->h_counts.p, (size_t)nq * 4, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(out_ids, ix->h_ids.p, no * 4, hipMemcpyDeviceToHost));
        if (out_dists) HIP_CHECK(hipMemcpy(out_dists, ix->h_dists.p, no * 4, hipMemcpyDeviceToHost));
    });
}

pdxg_status pdxg_filtered_search(pdxg_index* ix, const float* query, uint32_t transformed, uint32_t k,
                                 uint32_t nprobe, const uint8_t* selection, uint32_t* out_ids, float* out_dists,
                                 uint32_t* out_idists, uint32_t* out_count) {
    return guarded([&] {
        require(ix && query && selection && out_ids && out_count, "null argument");
        require(k >= 1, "k must be >= 1");
        require(k <= kMaxK, "k too large for the GPU path", PDXG_ERR_UNSUPPORTED);
        DeviceGuard g(ix->device);
        std::lock_guard<std::mutex> lk(ix->mu);
        Level& L = ix->l1;
        require(L.has_centroids, "filtered search needs the data-level centroids", PDXG_ERR_UNSUPPORTED);
        pdxg_workspace* ws = host_workspace(ix, 1, std::max(k, ix->ws ? ix->ws->max_k : 1u), 0);
        ix->h_queries.ensure(ix->d);
        ix->h_queries.upload(query, ix->d);
        launch_preprocess(ix, *ws, ix->h_queries.p, 1, transformed != 0, nullptr);
        // selection vector and per-cluster passing counts (PredicateEvaluator)
        std::vector<uint32_t> npass(L.nc, 0);
        for (uint32_t c = 0; c < L.nc; ++c)
            for (uint32_t i = 0; i < L.cap[c]; ++i) npass[c] += selection[L.slot_off[c] + i] == 1;
        ix->sel.ensure(L.total_slots);
        ix->sel.upload(selection, L.total_slots);
        ix->npass.ensure(L.nc);
        ix->npass.upload(npass.data(), L.nc);
        const uint32_t nv = (nprobe == 0 || nprobe > L.nc) ? L.nc : nprobe;
        std::vector<uint32_t> hk(1, k), hnv(1, nv);
        ws->k1.upload(hk.data(), 1);
        ws->nv1.upload(hnv.data(), 1);
        ix->f_order.ensure(L.nc);
        ix->f_probe.ensure(L.nc);
        launch_probe(ix, L, ws->qrot.p, 1, ws->nv1.p, ix->f_order.p, L.nc, ix->f_probe.p, nullptr);
        ix->h_ids.ensure(k);
        ix->h_dists.ensure(k);
        ix->h_idists.ensure(k);
        ix->h_counts.ensure(1);
        ScanArgs s{};
        s.L = L.view(ix->qscale_sq, ix->inv_qscale_sq);
        s.ratios = ix->ratios.p;
        s.qf = ws->qrot.p;
        s.qu = ws->qu8.p;
        s.order = ix->f_order.p;
        s.order_stride = L.nc;
        s.nvisit = ws->nv1.p;
        s.kq = ws->k1.p;
        s.out_ids = ix->h_ids.p;
        s.out_dists = ix->h_dists.p;
        s.out_idists = ix->h_idists.p;
        s.out_counts = ix->h_counts.p;
        s.out_stride = k;
        s.sel = ix->sel.p;
        s.n_pass = ix->npass.p;
        s.nq = 1;
        if (ix->u8()) launch_scan<true, true>(ix, L, s, k, nullptr);
        else launch_scan<false, true>(ix, L, s, k, nullptr);
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(out_count, ix->h_counts.p, 4, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(out_ids, ix->h_ids.p, (size_t)*out_count * 4, hipMemcpyDeviceToHost));
        if (out_dists) HIP_CHECK(hipMemcpy(out_dists, ix->h_dists.p, (size_t)*out_count * 4, hipMemcpyDeviceToHost));
        if (out_idists) HIP_CHECK(hipMemcpy(out_idists, ix->h_idists.p, (size_t)*out_count * 4, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"

#include "pdx_server.inc"
#include "pdx_dense.inc"
#include "pdx_graph.inc"
#include "pdx_build.inc"
