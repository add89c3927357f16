// Two-phase meso-level scan of two-level indexes.
//
// The meso level is searched with k = nprobe over about half of the meso clusters per query
// (index.hpp:545-566); with k that large ADSampling prunes little, so the pruned scan streams
// most of every visited meso cluster once per query. The two phases below compute the same
// results with the reads shared across queries:
//
//   phase 1  invert the per-query visit lists into per-cluster lists of (query, rank) entries and
//            compute, for every visited (query, vector) pair, every partial distance the scan can
//            compare: the vertical fma chain at each warmup checkpoint (the cumulative
//            DIMENSIONS_FETCHING_SIZES clipped to v, common.hpp:47-49) and each 64-dim horizontal
//            block distance in the 16-lane order. A work item is (cluster, 8 entries, 64-vector
//            chunk) on one wave; the 8 queries are wave-uniform and read through the scalar cache,
//            the chunk's vectors map one per lane (coalesced vertical reads).
//   phase 2  k_scan<..., SUMM = true>: the unchanged Start / Warmup / Prune / Merge control flow
//            per query, reading those values instead of the vectors.
//
// Every value the replay compares is produced by the same float operations in the same order as
// in the pruned scan, so thresholds, survivors, heap contents and byte counts are identical.
#pragma once
#include "pdx_kernels.hpp"

namespace pdxg {

constexpr uint32_t kMesoQ = 8;      // queries per work item (one wave)
constexpr uint32_t kMesoWaves = 4;  // work items per workgroup

// Groups of kMesoQ entries that share one wave on the last 64-vector chunk of a cluster of n
// vectors: 4 when that chunk holds at most 16 vectors, 2 when at most 32 (a packed tail item, lanes
// split between the groups), else 1 (the chunk is an ordinary work item).
__host__ __device__ constexpr uint32_t meso_tail_groups(uint32_t n) {
    return n == 0 ? 1u : (n - (n - 1) / 64 * 64 <= 16 ? 4u : (n - (n - 1) / 64 * 64 <= 32 ? 2u : 1u));
}
static_assert(kWave == 64, "meso_tail_groups assumes 64-lane waves");
static_assert(meso_tail_groups(64) == 1 && meso_tail_groups(65) == 4 && meso_tail_groups(96) == 2 &&
                  meso_tail_groups(97) == 1 && meso_tail_groups(16) == 4 && meso_tail_groups(17) == 2,
              "tail sizes");

struct MesoArgs {
    LevelDev L;                 // the meso level (F32)
    const float* qf;            // rotated queries [nq x d]
    const uint32_t* order;      // meso visit order [nq x stride]
    const uint32_t* nvisit;     // meso clusters visited per query
    uint32_t stride, nq, nv_max;
    uint32_t* cnt;              // [nc] entries per cluster
    uint32_t* cur;              // [nc] scatter cursors
    uint32_t* lstart;           // [nc + 1] first entry of each cluster
    uint32_t* cstart;           // [nc + 1] first work item of each cluster; a cluster's items are
                                // (8-entry group, 64-vector chunk) pairs, chunk-minor; a last
                                // chunk of at most 32 vectors is left to the packed tail items
    uint32_t* tstart;           // [nc + 1] first packed tail item of each cluster (meso_tail_groups)
    uint32_t* list;             // entries q * stride + rank, grouped by cluster
    uint32_t* soff;             // [nq x stride] first summarised vector of (query, rank) in its region
    unsigned long long* qbase;  // [nq + 1] first summarised vector of each query's region (the
                                // running sum of the queries' padded visited-cluster sizes)
    float* summ;                // summ_cap vectors x ns floats
    unsigned long long summ_cap;
    uint32_t nw, ns;            // warmup checkpoints; floats per summarised vector
    unsigned long long* pairs;  // optional: summarised (query, vector) pairs of the launch
};

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Visit-list inversion. A workgroup takes kListPer consecutive (query, rank) entries; the
// per-cluster counts go to an LDS histogram ([nc] words of dynamic LDS) and reach the global
// counters with one atomic per (workgroup, cluster), instead of one per entry on a few hundred
// contended addresses.
constexpr uint32_t kListThreads = 256;
constexpr uint32_t kListPerThread = 8;
constexpr uint32_t kListPer = kListThreads * kListPerThread;
constexpr uint32_t kListHistMax = 16384;  // largest nc for the LDS histogram (64 KB); the meso probe allows 8192

__global__ __launch_bounds__(kListThreads) void k_meso_count(MesoArgs a) {
    extern __shared__ uint32_t hist[];
    const uint32_t nc = a.L.nc;
    const size_t total = (size_t)a.nq * a.nv_max;
    const size_t e0 = (size_t)blockIdx.x * kListPer;
    for (uint32_t i = threadIdx.x; i < nc; i += kListThreads) hist[i] = 0;
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kListPerThread; ++k) {
        const size_t t = e0 + k * kListThreads + threadIdx.x;
        if (t >= total) break;
        const uint32_t q = static_cast<uint32_t>(t / a.nv_max), r = static_cast<uint32_t>(t % a.nv_max);
        if (r < a.nvisit[q]) atomicAdd(&hist[a.order[(size_t)q * a.stride + r]], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nc; i += kListThreads)
        if (hist[i]) atomicAdd(&a.cnt[i], hist[i]);
}

// Single-workgroup kernels of the meso phase run with kScanThreads threads: a small workgroup finds
// room on a CU at once while kernels of other batches in flight hold the rest (a 1024-thread one
// waited for a whole CU's worth of free wave slots).
constexpr uint32_t kScanThreads = 256;

// Entry and work-item offsets per cluster (one workgroup).
__global__ __launch_bounds__(kScanThreads) void k_meso_offsets(MesoArgs a) {
    __shared__ uint32_t red[kScanThreads / kWave + 1];
    const uint32_t nc = a.L.nc;
    uint32_t lbase = 0, cbase = 0, tbase = 0;
    for (uint32_t m0 = 0; m0 < nc; m0 += kScanThreads) {
        const uint32_t m = m0 + threadIdx.x;
        const uint32_t c = m < nc ? a.cnt[m] : 0u;
        uint32_t lt, ct, tt;
        const uint32_t lx = block_excl_sum<kScanThreads>(c, red, &lt);
        const uint32_t n = m < nc ? a.L.n_used[m] : 0u;
        const uint32_t nch = m < nc ? max(1u, (n + kWave - 1) / kWave) : 0u;
        const uint32_t tg = meso_tail_groups(n), groups = (c + kMesoQ - 1) / kMesoQ;
        const uint32_t cx = block_excl_sum<kScanThreads>(groups * (tg > 1 ? nch - 1 : nch), red, &ct);
        const uint32_t tx = block_excl_sum<kScanThreads>(tg > 1 ? (groups + tg - 1) / tg : 0u, red, &tt);
        if (m < nc) {
            a.lstart[m] = lbase + lx;
            a.cstart[m] = cbase + cx;
            a.tstart[m] = tbase + tx;
            a.cur[m] = 0;
        }
        lbase += lt;
        cbase += ct;
        tbase += tt;
    }
    if (threadIdx.x == 0) {
        a.lstart[nc] = lbase;
        a.cstart[nc] = cbase;
        a.tstart[nc] = tbase;
    }
}

// Entries land in their cluster's list in no particular order (each entry names its query and
// rank, and phase 1 writes each value to that query's own summary region). A workgroup ranks its
// entries in LDS and reserves one range per cluster with a single global atomic.
__global__ __launch_bounds__(kListThreads) void k_meso_scatter(MesoArgs a) {
    extern __shared__ uint32_t hist[];
    const uint32_t nc = a.L.nc;
    const size_t total = (size_t)a.nq * a.nv_max;
    const size_t e0 = (size_t)blockIdx.x * kListPer;
    for (uint32_t i = threadIdx.x; i < nc; i += kListThreads) hist[i] = 0;
    __syncthreads();
    uint32_t mm[kListPerThread], val[kListPerThread], rank[kListPerThread];
#pragma unroll
    for (uint32_t k = 0; k < kListPerThread; ++k) {
        mm[k] = 0xFFFFFFFFu;
        const size_t t = e0 + k * kListThreads + threadIdx.x;
        if (t >= total) continue;
        const uint32_t q = static_cast<uint32_t>(t / a.nv_max), r = static_cast<uint32_t>(t % a.nv_max);
        if (r >= a.nvisit[q]) continue;
        mm[k] = a.order[(size_t)q * a.stride + r];
        val[k] = q * a.stride + r;
        rank[k] = atomicAdd(&hist[mm[k]], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nc; i += kListThreads)
        if (hist[i]) hist[i] = atomicAdd(&a.cur[i], hist[i]);
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kListPerThread; ++k)
        if (mm[k] != 0xFFFFFFFFu) a.list[a.lstart[mm[k]] + hist[mm[k]] + rank[k]] = val[k];
}

// Summary offsets of a query's visits: running sum of the visited clusters' sizes, each rounded
// up to a multiple of 4 so that every component row of a summary block starts 16-byte aligned
// (one wave per query; the loads of up to 4 x 64 visits are issued before their in-order scan).
// The query's total goes to qbase[q]; k_meso_qbase turns those into region starts.
__global__ __launch_bounds__(kWave) void k_meso_soff(MesoArgs a) {
    __shared__ uint32_t red[2];
    const uint32_t q = blockIdx.x;
    const uint32_t nv = a.nvisit[q];
    const uint32_t* ord = a.order + (size_t)q * a.stride;
    uint32_t base = 0, pairs = 0;
    for (uint32_t r0 = 0; r0 < nv; r0 += 4 * kWave) {
        uint32_t c[4], x[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t r = r0 + u * kWave + threadIdx.x;
            c[u] = r < nv ? ord[r] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) x[u] = r0 + u * kWave + threadIdx.x < nv ? a.L.n_used[c[u]] : 0u;
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t r = r0 + u * kWave + threadIdx.x;
            uint32_t tot;
            const uint32_t ex = block_excl_sum<kWave>((x[u] + 3u) & ~3u, red, &tot);
            if (r < nv) a.soff[(size_t)q * a.stride + r] = base + ex;
            base += tot;
            pairs += x[u];
        }
    }
    if (threadIdx.x == 0) a.qbase[q] = base;
    if (a.pairs) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) pairs += __shfl_xor(pairs, off);
        if (threadIdx.x == 0) atomicAdd(a.pairs, (unsigned long long)pairs);
    }
}

// Region starts: exclusive running sum of the per-query totals in qbase[0 .. nq), in place, and
// the grand total in qbase[nq] (one workgroup; each thread takes 4 consecutive queries of a chunk).
__global__ __launch_bounds__(kScanThreads) void k_meso_qbase(MesoArgs a) {
    __shared__ uint32_t red[kScanThreads / kWave + 1];
    constexpr uint32_t kPer = 4;
    unsigned long long base = 0;
    for (uint32_t q0 = 0; q0 < a.nq; q0 += kScanThreads * kPer) {
        const uint32_t qt = q0 + threadIdx.x * kPer;
        uint32_t x[kPer], sum = 0;
#pragma unroll
        for (uint32_t u = 0; u < kPer; ++u) {
            x[u] = qt + u < a.nq ? static_cast<uint32_t>(a.qbase[qt + u]) : 0u;  // one query's total < 2^32
            sum += x[u];
        }
        uint32_t tot;
        unsigned long long ex = base + block_excl_sum<kScanThreads>(sum, red, &tot);
#pragma unroll
        for (uint32_t u = 0; u < kPer; ++u) {
            if (qt + u < a.nq) a.qbase[qt + u] = ex;
            ex += x[u];
        }
        base += tot;
    }
    if (threadIdx.x == 0) a.qbase[a.nq] = base;
}

// Query rows are read with wave-uniform addresses through the scalar cache.
typedef const __attribute__((address_space(4))) float const_float;
__device__ __forceinline__ const_float* to_const(const float* p) {
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wold-style-cast"
    return (const_float*)p;
#pragma clang diagnostic pop
}

// Phase 1 body for NQ (compile-time) queries of one work item on one wave: the cluster's vectors
// one per lane in chunks of 64, every query's chain per lane.
template <uint32_t NQ>
__device__ __forceinline__ void meso_summarize(const MesoArgs& a, const float* data, uint32_t S, uint32_t n,
                                               const_float* qbase, const uint32_t (&qo)[kMesoQ],
                                               const uint32_t (&ob)[kMesoQ], uint32_t i_begin) {
    const uint32_t lane = threadIdx.x % kWave;
    const uint32_t sn = (n + 3u) & ~3u;  // component stride of the summary block
    const uint32_t v = a.L.v, nh = a.L.h / kHDim;
    const bool aligned = ((size_t)v * S) % 4 == 0;  // 16-byte aligned horizontal blocks
    // one 64-bit base per query and constant offsets within a step, so the query values come in
    // as multi-dword scalar loads with no per-value address arithmetic
    const_float* qp[NQ];
#pragma unroll
    for (uint32_t u = 0; u < NQ; ++u) qp[u] = qbase + qo[u];
    {  // one 64-vector chunk per work item
        const uint32_t i0 = i_begin;
        const uint32_t i = i0 + lane;
        const bool valid = i < n;
        const uint32_t ii = valid ? i : n - 1;
        const float* col = data + ii;
        float acc[NQ];
#pragma unroll
        for (uint32_t u = 0; u < NQ; ++u) acc[u] = 0.0f;
        // vertical loads in flight per step: 16 for items of one or two queries (a single-query
        // search), within the registers the 8-query item needs; 8 otherwise
        constexpr uint32_t VS = NQ <= 2 ? 16 : 8;
        uint32_t j = 0, end = 0;
        for (uint32_t t = 0; t < a.nw; ++t) {
            end = min(end + c_fetch[t], v);
            if constexpr (VS > 8) {
                for (; j + VS <= end; j += VS) {
                    float x[VS];
#pragma unroll
                    for (uint32_t s = 0; s < VS; ++s) x[s] = col[(size_t)(j + s) * S];
#pragma unroll
                    for (uint32_t u = 0; u < NQ; ++u) {
                        const_float* pj = qp[u] + j;
#pragma unroll
                        for (uint32_t s = 0; s < VS; ++s) {
                            const float df = pj[s] - x[s];
                            acc[u] = __builtin_fmaf(df, df, acc[u]);
                        }
                    }
                }
            }
            for (; j + 8 <= end; j += 8) {
                float x[8];
#pragma unroll
                for (uint32_t s = 0; s < 8; ++s) x[s] = col[(size_t)(j + s) * S];
#pragma unroll
                for (uint32_t u = 0; u < NQ; ++u) {
                    const_float* pj = qp[u] + j;
#pragma unroll
                    for (uint32_t s = 0; s < 8; ++s) {
                        const float df = pj[s] - x[s];
                        acc[u] = __builtin_fmaf(df, df, acc[u]);
                    }
                }
            }
            for (; j + 4 <= end; j += 4) {
                float x[4];
#pragma unroll
                for (uint32_t s = 0; s < 4; ++s) x[s] = col[(size_t)(j + s) * S];
#pragma unroll
                for (uint32_t u = 0; u < NQ; ++u) {
                    const_float* pj = qp[u] + j;
#pragma unroll
                    for (uint32_t s = 0; s < 4; ++s) {
                        const float df = pj[s] - x[s];
                        acc[u] = __builtin_fmaf(df, df, acc[u]);
                    }
                }
            }
            for (; j < end; ++j) {
                const float x = col[(size_t)j * S];
#pragma unroll
                for (uint32_t u = 0; u < NQ; ++u) {
                    const float df = qp[u][j] - x;
                    acc[u] = __builtin_fmaf(df, df, acc[u]);
                }
            }
            if (valid) {
#pragma unroll
                for (uint32_t u = 0; u < NQ; ++u)
                    if (ob[u] != ~0u) a.summ[(size_t)ob[u] * a.ns + (size_t)t * sn + i] = acc[u];
            }
        }
        for (uint32_t b = 0; b < nh; ++b) {
            const float* blk = data + (size_t)(v + b * kHDim) * S + (size_t)ii * kHDim;
            float xb[kHDim];
            if (aligned) {
#pragma unroll
                for (uint32_t r = 0; r < kHDim / 4; ++r) {
                    const float4 t4 = reinterpret_cast<const float4*>(blk)[r];
                    xb[4 * r] = t4.x;
                    xb[4 * r + 1] = t4.y;
                    xb[4 * r + 2] = t4.z;
                    xb[4 * r + 3] = t4.w;
                }
            } else {
#pragma unroll
                for (uint32_t l = 0; l < kHDim; ++l) xb[l] = blk[l];
            }
            for (uint32_t u = 0; u < NQ; ++u) {
                float p[16];
#pragma unroll
                for (uint32_t hf = 0; hf < 2; ++hf) {  // 32 query values (scalar registers) at a time
                    const_float* qh = qp[u] + v + b * kHDim + 32 * hf;
#pragma unroll
                    for (uint32_t r = 0; r < 2; ++r)
#pragma unroll
                        for (uint32_t l = 0; l < 16; ++l) {
                            const float df = qh[16 * r + l] - xb[32 * hf + 16 * r + l];
                            p[l] = __builtin_fmaf(df, df, (hf == 0 && r == 0) ? 0.0f : p[l]);
                        }
                }
                const float hv = tree16_regs(p);
                if (valid && ob[u] != ~0u) a.summ[(size_t)ob[u] * a.ns + (size_t)(a.nw + b) * sn + i] = hv;
            }
        }
    }
}

// Phase 1 work item. Summary block of (query, rank), component-major: value t of vector
// i at [t * sn + i], sn = n rounded up to a multiple of 4; t < nw: vertical chain over dims
// [0, end_t); t = nw + b: horizontal block b.
__device__ __forceinline__ void meso_item(const MesoArgs& a, uint32_t item) {
    const uint32_t nc = a.L.nc;
    uint32_t lo = 0, hi = nc;          // cstart[lo] <= item < cstart[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.cstart[mid] <= item) lo = mid;
        else hi = mid;
    }
    const uint32_t m = lo;
    const uint32_t n = a.L.n_used[m];
    if (n == 0) return;
    // chunks of this cluster in ordinary items (its last one is a packed tail item's when small;
    // the cluster has an item here, so at least one chunk remains)
    const uint32_t nch = (n + kWave - 1) / kWave - (meso_tail_groups(n) > 1 ? 1u : 0u);
    const uint32_t local = item - a.cstart[m];
    const uint32_t e0 = a.lstart[m] + (local / nch) * kMesoQ;
    const uint32_t i_begin = (local % nch) * kWave;
    const uint32_t ne = min(kMesoQ, a.lstart[m + 1] - e0);
    const uint32_t d = a.L.d, S = a.L.stride[m];
    const float* data = reinterpret_cast<const float*>(a.L.data + a.L.data_off[m]);

    // per query: row offset of the rotated query and first summarised vector of its block
    // (region index q * region + soff < 2^32, checked by the host); ~0u = not stored
    const_float* qbase = to_const(a.qf);
    uint32_t qo[kMesoQ], ob[kMesoQ];
#pragma unroll
    for (uint32_t u = 0; u < kMesoQ; ++u) {
        const bool ok = u < ne;
        const uint32_t e = __builtin_amdgcn_readfirstlane(a.list[e0 + (ok ? u : 0u)]);
        const uint32_t q = e / a.stride;
        const uint32_t so = __builtin_amdgcn_readfirstlane(a.soff[e]);
        qo[u] = q * d;
        // the query's region is [qbase[q], qbase[q + 1]) and lies within the buffer (the host
        // admitted the batch by a bound on it); checked again so no store can leave either
        const unsigned long long r0 = a.qbase[q], r1 = a.qbase[q + 1];
        ob[u] = (ok && r0 + so + ((n + 3u) & ~3u) <= r1 && r1 <= a.summ_cap) ? static_cast<uint32_t>(r0 + so) : ~0u;
    }
    switch (ne) {  // the last item of a cluster may hold fewer queries
        case 8: meso_summarize<8>(a, data, S, n, qbase, qo, ob, i_begin); break;
        case 7: meso_summarize<7>(a, data, S, n, qbase, qo, ob, i_begin); break;
        case 6: meso_summarize<6>(a, data, S, n, qbase, qo, ob, i_begin); break;
        case 5: meso_summarize<5>(a, data, S, n, qbase, qo, ob, i_begin); break;
        case 4: meso_summarize<4>(a, data, S, n, qbase, qo, ob, i_begin); break;
        case 3: meso_summarize<3>(a, data, S, n, qbase, qo, ob, i_begin); break;
        case 2: meso_summarize<2>(a, data, S, n, qbase, qo, ob, i_begin); break;
        default: meso_summarize<1>(a, data, S, n, qbase, qo, ob, i_begin); break;
    }
}

// One item per wave. The dispatcher deals workgroups to the 8 XCDs in turn (workgroup b runs on
// XCD b % 8), so XCD x takes the contiguous items [x * per, (x + 1) * per) of the count made on
// the device: a cluster's items share one XCD's L2 and every XCD gets the same number of items
// however loose the host's bound (the grid) was. Waves past the count leave at once. (A
// grid-stride loop would need fewer waves but costs the body a register: past 80 a wave per SIMD
// is lost, measured clearly slower.)
__global__ __launch_bounds__(256) void k_meso_summary(MesoArgs a) {
    const uint32_t total = a.cstart[a.L.nc];
    const uint32_t per = (total + 7) / 8;
    const uint32_t local = blockIdx.x / 8 * kMesoWaves + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint32_t item = blockIdx.x % 8 * per + local;
    if (local < per && item < total) meso_item(a, item);
}

// Packed tail items. The last chunk of a cluster holding at most 32 (16) vectors is computed for 2
// (4) groups of kMesoQ entries on one wave: lane l takes vector l mod W of the chunk (W = 32 or 16)
// for the entries of group l / W, so no lane idles where an ordinary item would leave 32-63 lanes
// of the chunk empty. The groups' query values differ between the lanes, so each lane loads a
// slab of 16 dims of its group's query (lane l holds dim l mod 16 of the slab; W is a multiple of
// 16, so a row of 16 lanes belongs to one group) and the subtraction takes dim t of the slab from
// lane t of its row by a DPP row broadcast, folded into the subtraction (v_sub_f32_dpp
// row_newbcast). Every lane runs every instruction (the DPP reads need the whole row); only the
// stores are predicated. Each value comes from the same operations in the same order as in
// meso_summarize.
template <uint32_t T>
__device__ __forceinline__ float row_bcast(float v) {
    // row_newbcast always reads a lane of the row, so no lane takes the "old" value: none is given
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x150 + T, 0xF, 0xF, true));
}

// Vertical dims [lo, hi) of a 16-dim slab (wave-uniform bounds), in dim order.
template <uint32_t T>
__device__ __forceinline__ void tail_vslab(float (&acc)[kMesoQ], const float (&vq)[kMesoQ], const float (&x)[16],
                                           uint32_t lo, uint32_t hi) {
    if (T >= lo && T < hi) {
#pragma unroll
        for (uint32_t u = 0; u < kMesoQ; ++u) {
            const float df = row_bcast<T>(vq[u]) - x[T];
            acc[u] = __builtin_fmaf(df, df, acc[u]);
        }
    }
    if constexpr (T < 15) tail_vslab<T + 1>(acc, vq, x, lo, hi);
}

// One 64-dim horizontal block in the 16-lane order: p[l] over dims l, l + 16, l + 32, l + 48;
// vq[r] holds dims [16 r, 16 r + 16) of the block for this lane's row.
template <uint32_t L>
__device__ __forceinline__ void tail_hblock(float (&p)[16], const float (&vq)[4], const float (&xb)[kHDim]) {
    float df = row_bcast<L>(vq[0]) - xb[L];
    p[L] = __builtin_fmaf(df, df, 0.0f);
    df = row_bcast<L>(vq[1]) - xb[16 + L];
    p[L] = __builtin_fmaf(df, df, p[L]);
    df = row_bcast<L>(vq[2]) - xb[32 + L];
    p[L] = __builtin_fmaf(df, df, p[L]);
    df = row_bcast<L>(vq[3]) - xb[48 + L];
    p[L] = __builtin_fmaf(df, df, p[L]);
    if constexpr (L < 15) tail_hblock<L + 1>(p, vq, xb);
}

// A packed tail item runs as kTailParts waves: part 0 the vertical chain and its checkpoints,
// parts 1 .. kTailParts - 1 a third each of the horizontal blocks (independent of each other), so
// the few packed items still fill the GPU.
constexpr uint32_t kTailParts = 4;

__device__ __forceinline__ void meso_tail_item(const MesoArgs& a, uint32_t item, uint32_t part) {
    const uint32_t nc = a.L.nc;
    uint32_t lo = 0, hi = nc;  // tstart[lo] <= item < tstart[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.tstart[mid] <= item) lo = mid;
        else hi = mid;
    }
    const uint32_t m = lo;
    const uint32_t n = a.L.n_used[m];  // > 0: the cluster has a packed tail item
    const uint32_t tg = meso_tail_groups(n), W = kWave / tg;
    const uint32_t i0 = (n - 1) / kWave * kWave, mt = n - i0;  // the last chunk: vectors [i0, n)
    const uint32_t lane = threadIdx.x % kWave, col = lane % W;
    const uint32_t l0 = a.lstart[m], l1 = a.lstart[m + 1];
    const uint32_t e0 = l0 + ((item - a.tstart[m]) * tg + lane / W) * kMesoQ;  // this lane's group
    const uint32_t ne = e0 < l1 ? min(kMesoQ, l1 - e0) : 0u;
    const uint32_t i = i0 + min(col, mt - 1);
    const bool valid = col < mt;
    const uint32_t d = a.L.d, S = a.L.stride[m], v = a.L.v, nh = a.L.h / kHDim;
    const uint32_t sn = (n + 3u) & ~3u;
    const float* data = reinterpret_cast<const float*>(a.L.data + a.L.data_off[m]);

    // per lane: row offset of each query of its group and its summary block (~0u = not stored)
    uint32_t qo[kMesoQ], ob[kMesoQ];
#pragma unroll
    for (uint32_t u = 0; u < kMesoQ; ++u) {
        const bool ok = u < ne;
        const uint32_t e = a.list[ok ? e0 + u : l0];
        const uint32_t q = e / a.stride;
        qo[u] = q * d;
        const unsigned long long r0 = a.qbase[q], r1 = a.qbase[q + 1], so = a.soff[e];
        ob[u] = (ok && r0 + so + sn <= r1 && r1 <= a.summ_cap) ? static_cast<uint32_t>(r0 + so) : ~0u;
    }
    const uint32_t sl = lane & 15;  // this lane's dim within a slab

    if (part == 0) {
        // vertical: the fma chain over dims [0, end_t), stored at every checkpoint
        const float* colp = data + i;
        float acc[kMesoQ];
#pragma unroll
        for (uint32_t u = 0; u < kMesoQ; ++u) acc[u] = 0.0f;
        uint32_t j = 0, end = 0;
        for (uint32_t t = 0; t < a.nw; ++t) {
            end = min(end + c_fetch[t], v);
            while (j < end) {
                const uint32_t jb = j & ~15u, top = min(end - jb, 16u);
                float vq[kMesoQ];
#pragma unroll
                for (uint32_t u = 0; u < kMesoQ; ++u) vq[u] = a.qf[qo[u] + min(jb + sl, d - 1)];
                float x[16];
#pragma unroll
                for (uint32_t s = 0; s < 16; ++s) x[s] = s < top ? colp[(size_t)(jb + s) * S] : 0.0f;
                tail_vslab<0>(acc, vq, x, j - jb, top);
                j = jb + top;
            }
            if (valid) {
#pragma unroll
                for (uint32_t u = 0; u < kMesoQ; ++u)
                    if (ob[u] != ~0u) a.summ[(size_t)ob[u] * a.ns + (size_t)t * sn + i] = acc[u];
            }
        }
        return;
    }
    // horizontal blocks of this part
    const bool aligned = ((size_t)v * S) % 4 == 0;
    const uint32_t b0 = (part - 1) * nh / (kTailParts - 1), b1 = part * nh / (kTailParts - 1);
    for (uint32_t b = b0; b < b1; ++b) {
        const float* blk = data + (size_t)(v + b * kHDim) * S + (size_t)i * kHDim;
        float xb[kHDim];
        if (aligned) {
#pragma unroll
            for (uint32_t r = 0; r < kHDim / 4; ++r) {
                const float4 t4 = reinterpret_cast<const float4*>(blk)[r];
                xb[4 * r] = t4.x;
                xb[4 * r + 1] = t4.y;
                xb[4 * r + 2] = t4.z;
                xb[4 * r + 3] = t4.w;
            }
        } else {
#pragma unroll
            for (uint32_t l = 0; l < kHDim; ++l) xb[l] = blk[l];
        }
        const uint32_t hb = v + b * kHDim + sl;
#pragma unroll
        for (uint32_t u = 0; u < kMesoQ; ++u) {
            float vq[4];
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r) vq[r] = a.qf[qo[u] + hb + 16 * r];
            float p[16];
            tail_hblock<0>(p, vq, xb);
            float hv = tree16_regs(p);
            // keeps the block's arithmetic ahead of the predicated store (sunk into it, the
            // subtractions could no longer take the row broadcasts, which must run on every lane)
            asm volatile("" : "+v"(hv));
            if (valid && ob[u] != ~0u) a.summ[(size_t)ob[u] * a.ns + (size_t)(a.nw + b) * sn + i] = hv;
        }
    }
}

// One part of a packed tail item per wave (the parts of an item adjacent), dealt to the XCDs as in
// k_meso_summary.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_meso_tails(MesoArgs a) {
    const uint32_t total = a.tstart[a.L.nc] * kTailParts;
    const uint32_t per = (total + 7) / 8;
    const uint32_t local = blockIdx.x / 8 * kMesoWaves + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint32_t item = blockIdx.x % 8 * per + local;
    if (local < per && item < total) meso_tail_item(a, item / kTailParts, item % kTailParts);
}

}  // namespace pdxg
