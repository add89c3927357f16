// HIP kernels of the PDX search path for gfx950.
//
//   k_setup          per-query k / probe counts          (searcher.hpp:614-627, index.hpp:545-566)
//   k_normalize      query L2 normalisation              (quantizers/scalar.hpp:23-40)
//   k_rotate         y = x . M, fused SQ8 quantisation   (pruners/adsampling.hpp:225-245, scalar.hpp:76-94)
//   k_probe_dist     query x centroid L2 distances       (searcher.hpp:181-195)
//   k_probe_select   ascending top-nprobe per query      (searcher.hpp:196-214)
//   k_scan           Start / Warmup / Prune / Merge over a query's ordered clusters
//                    (searcher.hpp:218-566, 596-856), one workgroup per query
#pragma once
#include <float.h>

#include "pdx_device.hpp"

namespace pdxg {

// DIMENSIONS_FETCHING_SIZES, common.hpp:47-49
__constant__ uint32_t c_fetch[20] = {16,  16,  32,  32,  32,  32,  64,   64,   64,   64,
                                     128, 128, 128, 128, 256, 256, 512, 1024, 2048, 16384};

// Device view of one IVF level (IVF<Q>, ivf_wrapper.hpp:15-179). All clusters live in one pool;
// cluster c's PDX buffer starts at data + data_off[c] (bytes) with stride stride[c] vectors: its
// max_capacity rounded up to a multiple of 32, so that every dim-major row starts on a 128-byte
// line (the file's layout, with the cluster's max_capacity as the stride, otherwise unchanged).
struct LevelDev {
    const uint8_t* data;
    const uint64_t* data_off;
    const uint32_t* ids;       // slot row ids, cluster c at slot_off[c]
    const uint64_t* slot_off;  // running sum of max_capacity (cluster_offsets)
    const uint32_t* n_used;
    const uint32_t* n_live;
    const uint32_t* stride;
    const uint32_t* tomb;      // tombstone bitmap, cluster c at tomb_off[c] (words)
    const uint64_t* tomb_off;
    const uint32_t* n_tomb;
    const float* centroids;    // nc x d
    uint32_t nc, d, v, h;
    float scale_sq, inv_scale_sq;
};

struct ScanArgs {
    LevelDev L;
    const float* ratios;        // ADSampling ratio table [d]
    const float* qf;            // F32 queries [nq x d]
    const uint8_t* qu;          // U8 queries [nq x d]
    const uint32_t* order;      // cluster visit order [nq x order_stride]
    uint32_t order_stride;
    const uint32_t* nvisit;     // clusters to visit per query
    const uint32_t* order_count;  // optional per-query cap on nvisit (length of `order` rows)
    const uint32_t* kq;         // k per query
    uint32_t* out_ids;
    float* out_dists;
    uint32_t* out_idists;
    uint32_t* out_counts;
    uint32_t out_stride;
    unsigned long long* out_bytes;
    unsigned long long* stage_bytes;  // optional: algorithmic bytes of this launch, summed over queries
    const uint8_t* sel;         // filtered search: selection byte per slot (cluster at slot_off)
    const uint32_t* n_pass;     // filtered search: passing slots per cluster
    // two-phase meso scan: per (query, visited cluster) partial-distance summaries
    const float* summ;          // per query a region of its visited clusters' padded sizes x ns floats
    const uint32_t* soff;       // [nq x order_stride] first summarised vector of (query, rank)
    const unsigned long long* qbase;  // [nq + 1] first summarised vector of each query's region
    uint32_t nw, ns;            // warmup checkpoints; floats per summarised vector (nw + h/64)
    uint32_t nq;
    const uint32_t* perm;       // optional launch order: workgroup b runs query perm[b]
    uint32_t* queue;            // optional work-queue counter (persistent launch): queries are
                                // taken in perm order by atomically incrementing it
    // dynamic LDS layout (byte offsets)
    uint32_t o_q, o_dist, o_pos, o_h, o_hp, o_t, o_tp, o_o, o_op, o_ck, o_cp, o_red, o_tau, o_ring, o_r16;
};

// Per-cluster parameters of the next kRing visited clusters, staged in LDS in one batch so that
// a cluster's scan starts without a dependent global round trip.
struct ClusterMeta {
    unsigned long long data_off, slot_off, tomb_off;
    uint32_t c, n_live, n_used, stride, n_tomb, aux;  // aux: passing slots (filtered) / summary offset
};
constexpr uint32_t kRing = 16;
constexpr uint32_t kPosNone = 0xFFFFu;  // a survivor slot that was pruned

// Largest number of warmup checkpoints (entries of DIMENSIONS_FETCHING_SIZES used).
constexpr uint32_t kMaxCheckpoints = 20;

template <uint32_t BLOCK>
__device__ __forceinline__ uint32_t block_sum(uint32_t x, uint32_t* s_red) {
    constexpr uint32_t NW = BLOCK / kWave;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) s_red[w] = x;
    __syncthreads();
    uint32_t tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < NW; ++i) tot += s_red[i];
    __syncthreads();
    return tot;
}

// ------------------------------------------------------------------------------------------
// Per-query parameters
// ------------------------------------------------------------------------------------------
struct SetupArgs {
    const uint32_t* k_in;
    const uint32_t* nprobe_in;
    uint32_t k_all, nprobe_all, nq;
    uint32_t tree;
    uint32_t nc1, nc0;
    uint32_t k_cap;      // largest k the scan buffers hold
    uint32_t np_cap;     // declared largest nprobe (0: unbounded)
    uint32_t order_cap;  // flat: row length of the probe order buffer
    uint32_t k0_cap;     // tree: row length of the L0 result buffer
    uint32_t* k1;        // final k
    uint32_t* nvisit1;   // data-level clusters to visit
    uint32_t* k0;        // tree: L0 search k (= nprobe)
    uint32_t* nvisit0;   // tree: meso clusters to visit
    unsigned long long* bytes;      // optional per-query byte counters, initialised here
    unsigned long long probe_bytes; // centroid bytes streamed by the first probe
};

__global__ void k_setup(SetupArgs a) {
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.nq) return;
    // Values outside the declared maxima are clamped so that no buffer can be overrun.
    uint32_t k = a.k_in ? a.k_in[q] : a.k_all;
    k = k < 1 ? 1 : (k > a.k_cap ? a.k_cap : k);
    uint32_t np = a.nprobe_in ? a.nprobe_in[q] : a.nprobe_all;
    if (a.np_cap != 0 && (np == 0 || np > a.np_cap)) np = a.np_cap;
    a.k1[q] = k;
    if (a.bytes) a.bytes[q] = a.probe_bytes;
    if (!a.tree) {  // searcher.hpp:614-616
        uint32_t nv = (np == 0 || np > a.nc1) ? a.nc1 : np;
        a.nvisit1[q] = nv > a.order_cap ? a.order_cap : nv;
        return;
    }
    // index.hpp:545-566: nprobe 0 means all L1 clusters; halve the meso probes when visiting
    // fewer than half of the L1 clusters; the L0 search returns the top-nprobe L1 ids.
    uint32_t np1 = np == 0 ? a.nc1 : np;
    uint32_t np0 = a.nc0;
    if (np1 < a.nc1 / 2) np0 /= 2;
    a.k0[q] = np1 > a.k0_cap ? a.k0_cap : np1;
    a.nvisit0[q] = (np0 == 0 || np0 > a.nc0) ? a.nc0 : np0;
    a.nvisit1[q] = np1 > a.nc1 ? a.nc1 : np1;
}

// Longest-first launch order of the per-query scans: queries sorted by descending visit count
// (counting sort, counts above kLptBuckets - 1 share the top bucket; order within a bucket is
// unspecified). Workgroups are dispatched in launch order, so the longest walks start first and
// the short ones fill in behind them. One workgroup of kLptThreads threads: small, so that it finds
// room on a CU at once while kernels of other batches in flight hold the rest.
constexpr uint32_t kLptBuckets = 4096;
constexpr uint32_t kLptThreads = 256;
__global__ __launch_bounds__(kLptThreads) void k_lpt_order(const uint32_t* __restrict__ key, uint32_t nq,
                                                           uint32_t* __restrict__ perm) {
    __shared__ uint32_t cur[kLptBuckets];
    __shared__ uint32_t red[kLptThreads / kWave];
    constexpr uint32_t per = kLptBuckets / kLptThreads;
    for (uint32_t b = threadIdx.x; b < kLptBuckets; b += kLptThreads) cur[b] = 0;
    __syncthreads();
    // bucket kLptBuckets - 1 - min(key, kLptBuckets - 1): ascending bucket = descending key
    for (uint32_t q = threadIdx.x; q < nq; q += kLptThreads)
        atomicAdd(&cur[kLptBuckets - 1 - min(key[q], kLptBuckets - 1)], 1u);
    __syncthreads();
    uint32_t c[per], sum = 0;
#pragma unroll
    for (uint32_t u = 0; u < per; ++u) {
        c[u] = cur[threadIdx.x * per + u];
        sum += c[u];
    }
    uint32_t total;
    uint32_t off = block_excl_sum<kLptThreads>(sum, red, &total);
#pragma unroll
    for (uint32_t u = 0; u < per; ++u) {
        cur[threadIdx.x * per + u] = off;
        off += c[u];
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < nq; q += kLptThreads)
        perm[atomicAdd(&cur[kLptBuckets - 1 - min(key[q], kLptBuckets - 1)], 1u)] = q;
}

__global__ void k_fill_u64(unsigned long long* p, uint32_t n, unsigned long long v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------------------------
// Query preprocessing
// ------------------------------------------------------------------------------------------
// One workgroup per query; the 16 lanes of row 0 sum squares in the 16-lane order.
__global__ void k_normalize(const float* __restrict__ x, float* __restrict__ out, uint32_t d, uint32_t nq) {
    __shared__ float s_norm;
    __shared__ int s_zero;
    const uint32_t q = blockIdx.x;
    if (q >= nq) return;
    const float* xq = x + (size_t)q * d;
    if (threadIdx.x < 16) {
        // lane l sums dims l, l + 16, ... in that order; 8 loads are in flight ahead of their fmas
        float acc = 0.0f;
        uint32_t j = threadIdx.x;
        for (; j + 7 * 16 < d; j += 8 * 16) {
            float v[8];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) v[u] = xq[j + 16 * u];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) acc = __builtin_fmaf(v[u], v[u], acc);
        }
        for (; j < d; j += 16) acc = __builtin_fmaf(xq[j], xq[j], acc);
        acc = tree16(acc);
        if (threadIdx.x == 0) {
            s_zero = acc == 0.0f;
            s_norm = __builtin_sqrtf(acc);
        }
    }
    __syncthreads();
    const float nrm = s_norm;
    const bool zero = s_zero;
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) out[(size_t)q * d + j] = zero ? xq[j] : xq[j] / nrm;
}

// y[q][i] = sum_j x[q][j] * M[j][i] as one fma chain in j order per output; optional SQ8
// quantisation of y: clamp(round((y - base) * scale), 0, 255) with round-half-away-from-zero.
constexpr uint32_t kRotTQ = 16;   // queries per workgroup
constexpr uint32_t kRotTI = 64;   // outputs per workgroup
constexpr uint32_t kRotTJ = 64;   // staged input dims
__global__ __launch_bounds__(256) void k_rotate(const float* __restrict__ x, const float* __restrict__ M,
                                                float* __restrict__ y, uint8_t* __restrict__ yq, uint32_t d,
                                                uint32_t nq, float base, float scale, int identity) {
    __shared__ float xs[kRotTQ][kRotTJ + 1];
    const uint32_t tx = threadIdx.x % kRotTI, ty = threadIdx.x / kRotTI;  // ty in 0..3
    const uint32_t i = blockIdx.x * kRotTI + tx;
    const uint32_t q0 = blockIdx.y * kRotTQ;
    float acc[kRotTQ / 4];
#pragma unroll
    for (uint32_t r = 0; r < kRotTQ / 4; ++r) acc[r] = 0.0f;
    if (!identity) {
        for (uint32_t j0 = 0; j0 < d; j0 += kRotTJ) {
            for (uint32_t t = threadIdx.x; t < kRotTQ * kRotTJ; t += 256) {
                uint32_t qq = t / kRotTJ, jj = t % kRotTJ;
                uint32_t q = q0 + qq, j = j0 + jj;
                xs[qq][jj] = (q < nq && j < d) ? x[(size_t)q * d + j] : 0.0f;
            }
            __syncthreads();
            const uint32_t jl = min(kRotTJ, d - j0);
            if (i < d) {
                for (uint32_t jj = 0; jj < jl; ++jj) {
                    float m = M[(size_t)(j0 + jj) * d + i];
#pragma unroll
                    for (uint32_t r = 0; r < kRotTQ / 4; ++r) acc[r] = __builtin_fmaf(xs[ty + 4 * r][jj], m, acc[r]);
                }
            }
            __syncthreads();
        }
    }
    if (i >= d) return;
#pragma unroll
    for (uint32_t r = 0; r < kRotTQ / 4; ++r) {
        uint32_t q = q0 + ty + 4 * r;
        if (q >= nq) continue;
        float v = identity ? x[(size_t)q * d + i] : acc[r];
        if (y) y[(size_t)q * d + i] = v;
        if (yq) {
            float t = (v - base) * scale;
            int rr = static_cast<int>(roundf(t));
            yq[(size_t)q * d + i] = static_cast<uint8_t>(rr > 255 ? 255 : (rr < 0 ? 0 : rr));
        }
    }
}

// ------------------------------------------------------------------------------------------
// IVF probe
// ------------------------------------------------------------------------------------------
// dist[q][c] = l2sq16(q, centroid c). One workgroup per (QT queries, 64 centroids): the QT queries
// sit in LDS and each 16-lane row streams one centroid at a time, reusing every centroid load for
// all QT queries (QT accumulators per lane, each in the 16-lane order of l2sq16).
// Rotation of a few queries (nq <= kRotSmallQ): one workgroup of 64 threads per (64 outputs, query),
// the query staged in LDS (d floats, dynamic), each thread one output's fma chain over the input
// dims in order with the column loads of M issued 64 ahead. Bit-identical to k_rotate; it spreads a
// single query over d / 64 workgroups instead of leaving each chain behind a dependent load.
constexpr uint32_t kRotSmallQ = 8;
__global__ __launch_bounds__(64) void k_rotate_small(const float* __restrict__ x, const float* __restrict__ M,
                                                     float* __restrict__ y, uint8_t* __restrict__ yq, uint32_t d,
                                                     float base, float scale) {
    extern __shared__ float xq[];  // [d]
    const uint32_t q = blockIdx.y, i = blockIdx.x * 64 + threadIdx.x;
    for (uint32_t j = threadIdx.x; j < d; j += 64) xq[j] = x[(size_t)q * d + j];
    __syncthreads();
    if (i >= d) return;
    const float* col = M + i;
    float acc = 0.0f;
    uint32_t j = 0;
    for (; j + 64 <= d; j += 64) {
        float m[64];
#pragma unroll
        for (uint32_t u = 0; u < 64; ++u) m[u] = col[(size_t)(j + u) * d];
#pragma unroll
        for (uint32_t u = 0; u < 64; ++u) acc = __builtin_fmaf(xq[j + u], m[u], acc);
    }
    for (; j + 8 <= d; j += 8) {
        float m[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) m[u] = col[(size_t)(j + u) * d];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) acc = __builtin_fmaf(xq[j + u], m[u], acc);
    }
    for (; j < d; ++j) acc = __builtin_fmaf(xq[j], col[(size_t)j * d], acc);
    if (y) y[(size_t)q * d + i] = acc;
    if (yq) {
        const float t = (acc - base) * scale;
        const int rr = static_cast<int>(roundf(t));
        yq[(size_t)q * d + i] = static_cast<uint8_t>(rr > 255 ? 255 : (rr < 0 ? 0 : rr));
    }
}

// Batched rotation (nq >= kRotTiledMin and at least two tiles per CU): 64 queries x 64 outputs
// per workgroup, a 4 x 4 tile per thread (queries ty + 16r, outputs 4tx + c), with a 32-dim chunk
// of x and of M staged in LDS, so each M
// element is read from memory once per 64 queries and each pair of LDS reads feeds 16 fmas. Every
// output is still one fma chain over the input dims in order (bit-identical to k_rotate).
constexpr uint32_t kRotBQ = 64, kRotBI = 64, kRotBJ = 32;
// below this many queries the 16-query kernel spreads better (measured at d = 1024: 64 queries
// 124 us vs 227 us tiled; 4096 queries 316 us vs 273 us tiled)
constexpr uint32_t kRotTiledMin = 1024;
__global__ __launch_bounds__(256) void k_rotate_tiled(const float* __restrict__ x, const float* __restrict__ M,
                                                      float* __restrict__ y, uint8_t* __restrict__ yq, uint32_t d,
                                                      uint32_t nq, float base, float scale) {
    __shared__ float xs[kRotBQ][kRotBJ + 1];
    __shared__ __align__(16) float ms[kRotBJ][kRotBI];
    const uint32_t tx = threadIdx.x % 16, ty = threadIdx.x / 16;  // outputs 4tx..4tx+3, queries ty + 16r
    const uint32_t i0 = blockIdx.x * kRotBI, q0 = blockIdx.y * kRotBQ;
    float acc[4][4];
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r)
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) acc[r][c] = 0.0f;
    // the next chunk's x and M values are loaded into registers while the current one is used
    constexpr uint32_t kPX = kRotBQ * kRotBJ / 256, kPM = kRotBJ * kRotBI / 256;
    float px[kPX], pm[kPM];
    auto fetch = [&](uint32_t j0) {
#pragma unroll
        for (uint32_t u = 0; u < kPX; ++u) {
            const uint32_t t = threadIdx.x + 256 * u;
            const uint32_t qq = t / kRotBJ, jj = t % kRotBJ, q = q0 + qq, j = j0 + jj;
            px[u] = (q < nq && j < d) ? x[(size_t)q * d + j] : 0.0f;
        }
#pragma unroll
        for (uint32_t u = 0; u < kPM; ++u) {
            const uint32_t t = threadIdx.x + 256 * u;
            const uint32_t jj = t / kRotBI, ii = t % kRotBI, j = j0 + jj, i = i0 + ii;
            pm[u] = (j < d && i < d) ? M[(size_t)j * d + i] : 0.0f;
        }
    };
    fetch(0);
    for (uint32_t j0 = 0; j0 < d; j0 += kRotBJ) {
#pragma unroll
        for (uint32_t u = 0; u < kPX; ++u) {
            const uint32_t t = threadIdx.x + 256 * u;
            xs[t / kRotBJ][t % kRotBJ] = px[u];
        }
#pragma unroll
        for (uint32_t u = 0; u < kPM; ++u) {
            const uint32_t t = threadIdx.x + 256 * u;
            ms[t / kRotBI][t % kRotBI] = pm[u];
        }
        __syncthreads();
        if (j0 + kRotBJ < d) fetch(j0 + kRotBJ);
        const uint32_t jl = min(kRotBJ, d - j0);  // only the real dims enter the chains
#pragma unroll 8
        for (uint32_t jj = 0; jj < jl; ++jj) {
            const float4 m = *reinterpret_cast<const float4*>(&ms[jj][4 * tx]);
            float xv[4];
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r) xv[r] = xs[ty + 16 * r][jj];
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r) {
                acc[r][0] = __builtin_fmaf(xv[r], m.x, acc[r][0]);
                acc[r][1] = __builtin_fmaf(xv[r], m.y, acc[r][1]);
                acc[r][2] = __builtin_fmaf(xv[r], m.z, acc[r][2]);
                acc[r][3] = __builtin_fmaf(xv[r], m.w, acc[r][3]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r) {
        const uint32_t q = q0 + ty + 16 * r;
        if (q >= nq) continue;
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
            const uint32_t i = i0 + 4 * tx + c;
            if (i >= d) continue;
            const float v = acc[r][c];
            if (y) y[(size_t)q * d + i] = v;
            if (yq) {
                const float t = (v - base) * scale;
                const int rr = static_cast<int>(roundf(t));
                yq[(size_t)q * d + i] = static_cast<uint8_t>(rr > 255 ? 255 : (rr < 0 ? 0 : rr));
            }
        }
    }
}

constexpr uint32_t kProbeCPB = 64;
template <uint32_t QT>
__global__ __launch_bounds__(256) void k_probe_dist(const float* __restrict__ qf, const float* __restrict__ cent,
                                                    float* __restrict__ out, uint32_t d, uint32_t nc, uint32_t nq) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    float* qs = reinterpret_cast<float*>(lds_raw);  // [QT][d]
    const uint32_t q0 = blockIdx.x * QT;
    const uint32_t nqt = min(QT, nq - q0);
    for (uint32_t t = threadIdx.x; t < QT * d; t += blockDim.x) {
        uint32_t qq = t / d, j = t % d;
        qs[t] = qq < nqt ? qf[(size_t)(q0 + qq) * d + j] : 0.0f;
    }
    __syncthreads();
    const uint32_t g = threadIdx.x / 16, l = threadIdx.x % 16, G = blockDim.x / 16;
    const uint32_t c0 = blockIdx.y * kProbeCPB;
    // a row takes centroids cc and cc + G together, so each query value read from LDS feeds both
    for (uint32_t cc = g; cc < kProbeCPB; cc += 2 * G) {
        const uint32_t c = c0 + cc, c2 = c + G;
        if (c >= nc) break;  // uniform across the 16-lane row
        const bool two = cc + G < kProbeCPB && c2 < nc;  // uniform across the row
        const float* cr = cent + (size_t)c * d;
        const float* cr2 = cent + (size_t)(two ? c2 : c) * d;
        float acc[QT], acc2[QT];
#pragma unroll
        for (uint32_t r = 0; r < QT; ++r) acc[r] = acc2[r] = 0.0f;
        // lane l takes dims l, l + 16, ... in that order; 8 loads per centroid are in flight at once
        uint32_t j = l;
        for (; j + 7 * 16 < d; j += 8 * 16) {
            float xv[8], xw[8];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) {
                xv[u] = cr[j + 16 * u];
                xw[u] = cr2[j + 16 * u];
            }
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u)
#pragma unroll
                for (uint32_t r = 0; r < QT; ++r) {
                    const float qv = qs[r * d + j + 16 * u];
                    const float df = qv - xv[u], dg = qv - xw[u];
                    acc[r] = __builtin_fmaf(df, df, acc[r]);
                    acc2[r] = __builtin_fmaf(dg, dg, acc2[r]);
                }
        }
        for (; j < d; j += 16) {
            const float x = cr[j], y = cr2[j];
#pragma unroll
            for (uint32_t r = 0; r < QT; ++r) {
                const float qv = qs[r * d + j];
                const float df = qv - x, dg = qv - y;
                acc[r] = __builtin_fmaf(df, df, acc[r]);
                acc2[r] = __builtin_fmaf(dg, dg, acc2[r]);
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < QT; ++r) {
            float s = tree16(acc[r]);
            if (l == 0 && r < nqt) out[(size_t)(q0 + r) * nc + c] = s;
        }
        if (two) {
#pragma unroll
            for (uint32_t r = 0; r < QT; ++r) {
                float s = tree16(acc2[r]);
                if (l == 0 && r < nqt) out[(size_t)(q0 + r) * nc + c2] = s;
            }
        }
    }
}

// First nvisit[q] cluster ids by ascending (distance, id). Bitonic sort of n2 >= nc keys in LDS.
__global__ __launch_bounds__(256) void k_probe_select(const float* __restrict__ dist, uint32_t nc, uint32_t n2,
                                                      const uint32_t* __restrict__ nvisit, uint32_t* __restrict__ order,
                                                      uint32_t order_stride, uint32_t nq) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    uint64_t* keys = reinterpret_cast<uint64_t*>(lds_raw);
    const uint32_t q = blockIdx.x;
    if (q >= nq) return;
    for (uint32_t c = threadIdx.x; c < n2; c += 256)
        keys[c] = c < nc ? make_key(dist[(size_t)q * nc + c], c) : kKeyInf;
    __syncthreads();
    bitonic_sort<256>(keys, nullptr, n2);
    const uint32_t nv = nvisit[q];
    for (uint32_t t = threadIdx.x; t < nv; t += 256) order[(size_t)q * order_stride + t] = key_id(keys[t]);
}

// ------------------------------------------------------------------------------------------
// The pruned PDX scan
// ------------------------------------------------------------------------------------------
template <bool U8>
struct DistT {
    using type = float;
};
template <>
struct DistT<true> {
    using type = uint32_t;
};


// SUMM: the meso-level replay. Every partial distance the scan compares (warmup checkpoints
// and 64-dim horizontal blocks) was computed beforehand by k_meso_summary in the same operation
// order, so the replay reads those values instead of the vectors; the vertical remainder after
// an early warmup exit still reads the vectors.
template <bool U8, bool FILT, uint32_t BLOCK, bool SUMM = false>
struct Scanner {
    static_assert(!SUMM || (!U8 && !FILT), "the summary replay is F32 and unfiltered");
    using dist_t = typename DistT<U8>::type;
    static constexpr uint32_t G = BLOCK / 16;  // 16-lane rows per workgroup
    static constexpr uint32_t kVT = 4;  // vectors per thread per vertical tile

    const ScanArgs& a;
    uint8_t* lds;
    const float* qf;  // F32 query [d]: in LDS, or (replay, which reads it only for a vertical
                      // remainder after an early warmup exit) its row in global memory
    uint32_t* qw;     // U8 query words: [nvw] vertical then [h/4] horizontal
    dist_t* dist;
    uint16_t* pos;    // survivor slots (a cluster's LDS distance array keeps it far below 2^16 slots)
    uint64_t *H, *T, *O, *CK;
    uint32_t *HP, *TP, *OP, *CP, *red;
    dist_t* taus;     // thresholds of the current prune steps
    const float* rat;  // ADSampling ratio table [d] (global)
    float* r16;        // LDS copy of the ratios at multiples of 16 dims, where every checkpoint and,
                       // when v is a multiple of 16, every prune step ends: [j] = rat[16 j]
    ClusterMeta* ring;
    uint32_t* wcnt;   // SUMM: pruned counts per warmup checkpoint
    uint32_t* cpe;    // warmup checkpoint ends (cumulative fetch sizes clipped to v)
    uint32_t d, v, h, nvw;
    uint32_t hl = 0;
    unsigned long long bytes = 0;        // uniform: identical in every thread
    unsigned long long bytes_local = 0;  // this thread's share (summed over the block at the end)

    // per-cluster state
    const uint8_t* base;
    const uint32_t* cid;
    const uint32_t* tbm;
    const uint8_t* csel;
    const float* sb;  // SUMM: summary block of the current (query, cluster), component-major
    uint32_t sn;      // SUMM: its component stride, n rounded up to a multiple of 4
    uint32_t n, S, ntomb;
    bool hvec;        // horizontal blocks of this cluster are 16-byte aligned

    __device__ Scanner(const ScanArgs& args, uint8_t* l) : a(args), lds(l) {
        qf = reinterpret_cast<const float*>(lds + a.o_q);
        qw = reinterpret_cast<uint32_t*>(lds + a.o_q);
        dist = reinterpret_cast<dist_t*>(lds + a.o_dist);
        pos = reinterpret_cast<uint16_t*>(lds + a.o_pos);
        H = reinterpret_cast<uint64_t*>(lds + a.o_h);
        HP = reinterpret_cast<uint32_t*>(lds + a.o_hp);
        T = reinterpret_cast<uint64_t*>(lds + a.o_t);
        TP = reinterpret_cast<uint32_t*>(lds + a.o_tp);
        O = reinterpret_cast<uint64_t*>(lds + a.o_o);
        OP = reinterpret_cast<uint32_t*>(lds + a.o_op);
        CK = reinterpret_cast<uint64_t*>(lds + a.o_ck);
        CP = reinterpret_cast<uint32_t*>(lds + a.o_cp);
        red = reinterpret_cast<uint32_t*>(lds + a.o_red);
        taus = reinterpret_cast<dist_t*>(lds + a.o_tau);
        rat = a.ratios;
        ring = reinterpret_cast<ClusterMeta*>(lds + a.o_ring);
        r16 = reinterpret_cast<float*>(lds + a.o_r16);
        d = a.L.d;
        v = a.L.v;
        h = a.L.h;
        nvw = (v + 3) / 4;
        wcnt = reinterpret_cast<uint32_t*>(taus) + (h / kHDim + (v + kHDim - 1) / kHDim + 1);
        cpe = wcnt + kMaxCheckpoints + 1;
    }

    static __device__ __forceinline__ dist_t mask_value() {
        if constexpr (U8) return 0xFFFFFFFFu / 2u;
        else return FLT_MAX / 2.0f;
    }

    __device__ __forceinline__ float to_float(dist_t x) const {
        if constexpr (U8) return static_cast<float>(x) * a.L.inv_scale_sq;
        else return x;
    }

    __device__ __forceinline__ bool tombstoned(uint32_t i) const {
        return ntomb != 0 && ((tbm[i >> 5] >> (i & 31)) & 1u);
    }

    // GetPruningThreshold: adsampling.hpp:91-98 with the U8 conversion of searcher.hpp:75-91.
    __device__ __forceinline__ dist_t threshold(uint32_t k, uint32_t cur) const {
        float top = key_dist(H[k - 1]);
        float ratio = cur == d ? 1.0f : (cur % 16 == 0 ? r16[cur / 16] : rat[cur]);
        float tf = top * ratio;
        if constexpr (U8) {
            float sc = tf * a.L.scale_sq;
            return sc >= 4294967296.0f ? 0xFFFFFFFFu : static_cast<uint32_t>(sc);
        } else {
            return tf;
        }
    }

    __device__ __forceinline__ uint32_t qbyte(uint32_t j) const { return (qw[j >> 2] >> (8 * (j & 3))) & 0xFFu; }

    __device__ void load_query(uint32_t q) {
        for (uint32_t j = threadIdx.x; j < (d + 15) / 16; j += BLOCK) r16[j] = rat[16 * j];
        if (SUMM && threadIdx.x == 0) {
            uint32_t e = 0;
            for (uint32_t t = 0; t < kMaxCheckpoints; ++t) {
                e = min(e + c_fetch[t], v);
                cpe[t] = e;
            }
        }
        if constexpr (!U8) {
            if constexpr (SUMM) {
                qf = a.qf + (size_t)q * d;
            } else {
                float* ql = reinterpret_cast<float*>(lds + a.o_q);
                for (uint32_t j = threadIdx.x; j < d; j += BLOCK) ql[j] = a.qf[(size_t)q * d + j];
            }
        } else {
            const uint8_t* src = a.qu + (size_t)q * d;
            for (uint32_t w = threadIdx.x; w < nvw; w += BLOCK) {
                uint32_t word = 0;
                for (uint32_t t = 0; t < 4; ++t) {
                    uint32_t j = 4 * w + t;
                    word |= (j < v ? static_cast<uint32_t>(src[j]) : 0u) << (8 * t);
                }
                qw[w] = word;
            }
            for (uint32_t w = threadIdx.x; w < h / 4; w += BLOCK) {
                uint32_t word = 0;
                for (uint32_t t = 0; t < 4; ++t) word |= static_cast<uint32_t>(src[v + 4 * w + t]) << (8 * t);
                qw[nvw + w] = word;
            }
        }
        __syncthreads();
    }

    // Slot of element t of this thread's vertical tile at i0: a tile covers VT * BLOCK slots
    // (VT a multiple of 4), four consecutive slots per thread in each 4*BLOCK-slot part, so one
    // 16-byte load reads their values (F32) or words (U8) of a dim (group).
    template <uint32_t VT>
    static __device__ __forceinline__ uint32_t tile_slot(uint32_t i0, uint32_t t) {
        static_assert(VT % 4 == 0, "vertical tiles hold whole quads of slots");
        return i0 + (t >> 2) * 4 * BLOCK + 4 * threadIdx.x + (t & 3);
    }

    // A vertical tile's distances from / into dist[] (slots >= n are clamped on the way in and
    // dropped on the way out): quads inside the cluster move as one 16-byte LDS access (o_dist is
    // 16-byte aligned).
    template <uint32_t VT>
    __device__ __forceinline__ void load_tile(dist_t (&acc)[VT], uint32_t i0) const {
#pragma unroll
        for (uint32_t r = 0; r < VT / 4; ++r) {
            const uint32_t s = tile_slot<VT>(i0, 4 * r);
            if (s + 3 < n) {
                const uint4 x = *reinterpret_cast<const uint4*>(dist + s);
                acc[4 * r] = __builtin_bit_cast(dist_t, x.x);
                acc[4 * r + 1] = __builtin_bit_cast(dist_t, x.y);
                acc[4 * r + 2] = __builtin_bit_cast(dist_t, x.z);
                acc[4 * r + 3] = __builtin_bit_cast(dist_t, x.w);
            } else {
#pragma unroll
                for (uint32_t e = 0; e < 4; ++e) acc[4 * r + e] = dist[min(s + e, n - 1)];
            }
        }
    }
    template <uint32_t VT>
    __device__ __forceinline__ void store_tile(const dist_t (&acc)[VT], uint32_t i0) const {
#pragma unroll
        for (uint32_t r = 0; r < VT / 4; ++r) {
            const uint32_t s = tile_slot<VT>(i0, 4 * r);
            if (s + 3 < n) {
                *reinterpret_cast<uint4*>(dist + s) = make_uint4(
                    __builtin_bit_cast(uint32_t, acc[4 * r]), __builtin_bit_cast(uint32_t, acc[4 * r + 1]),
                    __builtin_bit_cast(uint32_t, acc[4 * r + 2]), __builtin_bit_cast(uint32_t, acc[4 * r + 3]));
            } else {
#pragma unroll
                for (uint32_t e = 0; e < 4; ++e)
                    if (s + e < n) dist[s + e] = acc[4 * r + e];
            }
        }
    }

    // Vertical (scalar_computers.hpp:19-44 / 77-121) over dims [da, db) for the VT slots
    // tile_slot(i0, t) of this thread; slots >= n are clamped to a valid slot and ignored.
    template <uint32_t VT = kVT>
    __device__ __forceinline__ void vertical_tile(dist_t (&acc)[VT], uint32_t i0, uint32_t da, uint32_t db) const {
        constexpr uint32_t VQ = VT / 4;
        // first slot of each of this thread's quads; a quad holding a valid slot lies inside its
        // dim's (group's) row (device strides are multiples of 32 slots, rows 256-byte aligned),
        // the others are clamped to the last quad and ignored
        uint32_t q4[VQ];
#pragma unroll
        for (uint32_t r = 0; r < VQ; ++r) q4[r] = min(tile_slot<VT>(i0, 4 * r), (n - 1) & ~3u);
        if constexpr (!U8) {
            const float* data = reinterpret_cast<const float*>(base);
            auto val4 = [&](uint32_t j, uint32_t r) {
                return *reinterpret_cast<const float4*>(data + (size_t)j * S + q4[r]);
            };
            // each slot's chain runs over the dims in order
            auto fma4 = [&](uint32_t r, const float4& x, float qj) {
                float df = qj - x.x;
                acc[4 * r] = __builtin_fmaf(df, df, acc[4 * r]);
                df = qj - x.y;
                acc[4 * r + 1] = __builtin_fmaf(df, df, acc[4 * r + 1]);
                df = qj - x.z;
                acc[4 * r + 2] = __builtin_fmaf(df, df, acc[4 * r + 2]);
                df = qj - x.w;
                acc[4 * r + 3] = __builtin_fmaf(df, df, acc[4 * r + 3]);
            };
            uint32_t j = da;
            // 8 dims x VT vectors of loads in flight per round trip
            for (; j + 8 <= db; j += 8) {
                float4 x[8][VQ];
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u)
#pragma unroll
                    for (uint32_t r = 0; r < VQ; ++r) x[u][r] = val4(j + u, r);
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) {
                    const float qj = qf[j + u];
#pragma unroll
                    for (uint32_t r = 0; r < VQ; ++r) fma4(r, x[u][r], qj);
                }
            }
            for (; j + 4 <= db; j += 4) {
                float4 x[4][VQ];
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u)
#pragma unroll
                    for (uint32_t r = 0; r < VQ; ++r) x[u][r] = val4(j + u, r);
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) {
                    const float qj = qf[j + u];
#pragma unroll
                    for (uint32_t r = 0; r < VQ; ++r) fma4(r, x[u][r], qj);
                }
            }
            for (; j < db; ++j) {
                const float qj = qf[j];
#pragma unroll
                for (uint32_t r = 0; r < VQ; ++r) fma4(r, val4(j, r), qj);
            }
        } else {
            uint32_t idx[VT];
#pragma unroll
            for (uint32_t t = 0; t < VT; ++t) idx[t] = min(tile_slot<VT>(i0, t), n - 1);
            auto word4 = [&](uint32_t j, uint32_t r) {
                return *reinterpret_cast<const uint4*>(base + (size_t)j * S + 4 * q4[r]);
            };
            auto add4 = [&](uint32_t r, const uint4& x, uint32_t qv, uint32_t qs) {
                acc[4 * r] += sqdist4(qv, x.x, qs);
                acc[4 * r + 1] += sqdist4(qv, x.y, qs);
                acc[4 * r + 2] += sqdist4(qv, x.z, qs);
                acc[4 * r + 3] += sqdist4(qv, x.w, qs);
            };
            uint32_t j = da;
            // 32 dims (8 words) x VT vectors of loads in flight per round trip (VT <= 4: wider
            // tiles take 16-dim rounds)
            for (; VT <= 4 && j + 32 <= db; j += 32) {
                uint4 x[8][VQ];
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u)
#pragma unroll
                    for (uint32_t r = 0; r < VQ; ++r) x[u][r] = word4(j + 4 * u, r);
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) {
                    const uint32_t w = (j >> 2) + u;
                    const uint32_t qv = qw[w], qs = sumsq4(qv);
#pragma unroll
                    for (uint32_t r = 0; r < VQ; ++r) add4(r, x[u][r], qv, qs);
                }
            }
            for (; j + 16 <= db; j += 16) {
                uint4 x[4][VQ];
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u)
#pragma unroll
                    for (uint32_t r = 0; r < VQ; ++r) x[u][r] = word4(j + 4 * u, r);
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) {
                    const uint32_t w = (j >> 2) + u;
                    const uint32_t qv = qw[w], qs = sumsq4(qv);
#pragma unroll
                    for (uint32_t r = 0; r < VQ; ++r) add4(r, x[u][r], qv, qs);
                }
            }
            for (; j + 4 <= db; j += 4) {
                const uint32_t w = j >> 2;
                const uint32_t qv = qw[w], qs = sumsq4(qv);
#pragma unroll
                for (uint32_t r = 0; r < VQ; ++r) add4(r, word4(j, r), qv, qs);
            }
            if (j < db) {  // tail group of rem < 4 dims, stored as rem bytes per vector
                const uint32_t rem = db - j;
#pragma unroll
                for (uint32_t t = 0; t < VT; ++t)
                    for (uint32_t u = 0; u < rem; ++u) {
                        int df = static_cast<int>(qbyte(j + u)) -
                                 static_cast<int>(base[(size_t)j * S + (size_t)idx[t] * rem + u]);
                        acc[t] += static_cast<uint32_t>(df * df);
                    }
            }
        }
    }

    // Vertical over dims [da, db) for one gathered slot i (VerticalPruning). Every dim is a
    // separate (strided) load, so 8 F32 dims or 8 U8 words are loaded before the chain uses them.
    __device__ __forceinline__ dist_t vertical_one(dist_t acc, uint32_t i, uint32_t da, uint32_t db) const {
        if constexpr (!U8) {
            const float* data = reinterpret_cast<const float*>(base);
            uint32_t j = da;
            for (; j + 8 <= db; j += 8) {
                float x[8];
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) x[u] = data[(size_t)(j + u) * S + i];
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) {
                    const float df = qf[j + u] - x[u];
                    acc = __builtin_fmaf(df, df, acc);
                }
            }
            for (; j < db; ++j) {
                float df = qf[j] - data[(size_t)j * S + i];
                acc = __builtin_fmaf(df, df, acc);
            }
        } else {
            uint32_t j = da;
            for (; j + 32 <= db; j += 32) {
                uint32_t x[8];
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) x[u] = *reinterpret_cast<const uint32_t*>(base + (size_t)(j + 4 * u) * S + 4 * i);
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) {
                    const uint32_t w = (j >> 2) + u;
                    acc += sqdist4(qw[w], x[u], sumsq4(qw[w]));
                }
            }
            for (; j + 4 <= db; j += 4) {
                const uint32_t w = j >> 2;
                acc += sqdist4(qw[w], *reinterpret_cast<const uint32_t*>(base + (size_t)j * S + 4 * i), sumsq4(qw[w]));
            }
            if (j < db) {
                const uint32_t rem = db - j;
                for (uint32_t u = 0; u < rem; ++u) {
                    int df = static_cast<int>(qbyte(j + u)) - static_cast<int>(base[(size_t)j * S + (size_t)i * rem + u]);
                    acc += static_cast<uint32_t>(df * df);
                }
            }
        }
        return acc;
    }

    // Horizontal 64-dim block `hd` of slot i, computed by the 16 lanes of this thread's row
    // (valid in lane 0 of the row).
    __device__ __forceinline__ dist_t horizontal16(uint32_t i, uint32_t hd, uint32_t l) const {
        if constexpr (!U8) {
            const float* blk = reinterpret_cast<const float*>(base) + (size_t)(v + hd) * S + (size_t)i * kHDim;
            const float* qh = qf + v + hd;
            float x0 = blk[l], x1 = blk[l + 16], x2 = blk[l + 32], x3 = blk[l + 48];
            float acc = 0.0f, df;
            df = qh[l] - x0;
            acc = __builtin_fmaf(df, df, acc);
            df = qh[l + 16] - x1;
            acc = __builtin_fmaf(df, df, acc);
            df = qh[l + 32] - x2;
            acc = __builtin_fmaf(df, df, acc);
            df = qh[l + 48] - x3;
            acc = __builtin_fmaf(df, df, acc);
            return tree16(acc);
        } else {
            const uint8_t* blk = base + (size_t)(v + hd) * S + (size_t)i * kHDim;
            const uint32_t w = nvw + hd / 4 + l;
            uint32_t x = *reinterpret_cast<const uint32_t*>(blk + 4 * l);
            return sum16_u(sqdist4(qw[w], x, sumsq4(qw[w])));
        }
    }

    // The same block distance computed by one thread: the 16 lane partial sums in registers and
    // the same reduction tree (F32), or the exact integer sum (U8).
    __device__ __forceinline__ dist_t horizontal_thread(uint32_t i, uint32_t hd) const {
        if constexpr (SUMM) {
            return sb[(size_t)(a.nw + hd / kHDim) * sn + i];
        } else if constexpr (!U8) {
            const float* blk = reinterpret_cast<const float*>(base) + (size_t)(v + hd) * S + (size_t)i * kHDim;
            const float* qh = qf + v + hd;
            float p[16];
            if constexpr (BLOCK < 1024) {  // half a block in flight per round trip (32 registers)
#pragma unroll 1
                for (uint32_t hb = 0; hb < 2; ++hb) {
                    float x[32];
                    if (hvec) {
#pragma unroll
                        for (uint32_t c = 0; c < 8; ++c) {
                            const float4 t4 = reinterpret_cast<const float4*>(blk + 32 * hb)[c];
                            x[4 * c] = t4.x;
                            x[4 * c + 1] = t4.y;
                            x[4 * c + 2] = t4.z;
                            x[4 * c + 3] = t4.w;
                        }
                    } else {
#pragma unroll
                        for (uint32_t l = 0; l < 32; ++l) x[l] = blk[32 * hb + l];
                    }
#pragma unroll
                    for (uint32_t r = 0; r < 2; ++r)
#pragma unroll
                        for (uint32_t l = 0; l < 16; ++l) {
                            const float df = qh[32 * hb + 16 * r + l] - x[16 * r + l];
                            p[l] = __builtin_fmaf(df, df, (hb == 0 && r == 0) ? 0.0f : p[l]);
                        }
                }
                return tree16_regs(p);
            }
#pragma unroll 1
            for (uint32_t r = 0; r < 4; ++r) {  // dims 16r .. 16r+15: one step of each lane's chain
                float x[16];
                if (hvec) {
#pragma unroll
                    for (uint32_t c = 0; c < 4; ++c) {
                        const float4 t4 = reinterpret_cast<const float4*>(blk + 16 * r)[c];
                        x[4 * c] = t4.x;
                        x[4 * c + 1] = t4.y;
                        x[4 * c + 2] = t4.z;
                        x[4 * c + 3] = t4.w;
                    }
                } else {
#pragma unroll
                    for (uint32_t l = 0; l < 16; ++l) x[l] = blk[16 * r + l];
                }
#pragma unroll
                for (uint32_t l = 0; l < 16; ++l) {
                    const float df = qh[16 * r + l] - x[l];
                    p[l] = __builtin_fmaf(df, df, r == 0 ? 0.0f : p[l]);
                }
            }
            return tree16_regs(p);
        } else {
            uint32_t x[16];
            fetch_block_u8(i, hd, x);
            return block_dist_u8(x, hd);
        }
    }

    // U8 horizontal block of slot i as 16 packed words, and its exact distance to the query.
    __device__ __forceinline__ void fetch_block_u8(uint32_t i, uint32_t hd, uint32_t (&x)[16]) const {
        const uint8_t* blk = base + (size_t)(v + hd) * S + (size_t)i * kHDim;
        if (hvec) {
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r) {
                const uint4 t4 = reinterpret_cast<const uint4*>(blk)[r];
                x[4 * r] = t4.x;
                x[4 * r + 1] = t4.y;
                x[4 * r + 2] = t4.z;
                x[4 * r + 3] = t4.w;
            }
        } else {
#pragma unroll
            for (uint32_t w = 0; w < 16; ++w)
                x[w] = static_cast<uint32_t>(blk[4 * w]) | (static_cast<uint32_t>(blk[4 * w + 1]) << 8) |
                       (static_cast<uint32_t>(blk[4 * w + 2]) << 16) | (static_cast<uint32_t>(blk[4 * w + 3]) << 24);
        }
    }
    __device__ __forceinline__ uint32_t block_dist_u8(const uint32_t (&x)[16], uint32_t hd) const {
        const uint32_t w0 = nvw + hd / 4;
        uint32_t acc = 0;
#pragma unroll
        for (uint32_t w = 0; w < 16; ++w) acc += sqdist4(qw[w0 + w], x[w], sumsq4(qw[w0 + w]));
        return acc;
    }

    __device__ __forceinline__ uint64_t key_of(uint32_t i) const {
        return make_key(to_float(dist[i]), cid[i]);
    }
    __device__ __forceinline__ uint32_t pay_of(uint32_t i) const {
        if constexpr (U8) return dist[i];
        else return 0u;
    }
    __device__ __forceinline__ bool selected(uint32_t i) const {
        if constexpr (FILT) return csel[i] == 1;
        else return true;
    }

    // Push the m smallest slots (all n, keys (distance, id)) onto the heap: the partial_sort
    // + push of Start / FilteredStart (searcher.hpp:256-280, 347-371).
    __device__ void push_m_smallest(uint32_t m) {
        uint32_t tl = 0;
        for (uint32_t i0 = 0; i0 < n; i0 += BLOCK) {
            uint32_t i = i0 + threadIdx.x;
            bool valid = i < n;
            uint64_t key = valid ? key_of(i) : kKeyInf;
            uint32_t pay = valid ? pay_of(i) : 0u;
            tl = list_insert_chunk<BLOCK>(valid, key, pay, T, TP, tl, m, CK, CP, O, OP, red);
        }
        uint32_t nl = merge_sorted<BLOCK>(H, HP, hl, T, TP, tl, O, OP, hl + tl);
        for (uint32_t i = threadIdx.x; i < nl; i += BLOCK) {
            H[i] = O[i];
            HP[i] = OP[i];
        }
        __syncthreads();
        hl = nl;
        bytes += 4ull * m;
    }

    __device__ void mask_tombstones_all() {
        if (ntomb == 0) return;
        for (uint32_t i = threadIdx.x; i < n; i += BLOCK)
            if (tombstoned(i)) dist[i] = mask_value();
        __syncthreads();
    }

    // searcher.hpp:218-281
    __device__ void start(uint32_t k) {
        if constexpr (SUMM) {
            const uint32_t nw = a.nw, nh = h / kHDim;
            for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
                float dd = nw ? sb[(size_t)(nw - 1) * sn + i] : 0.0f;
                for (uint32_t b = 0; b < nh; ++b) dd += sb[(size_t)(nw + b) * sn + i];
                dist[i] = dd;
            }
            __syncthreads();
            bytes += (unsigned long long)n * d * 4;
            mask_tombstones_all();
            push_m_smallest(min(k - hl, n));
            return;
        }
        for (uint32_t i0 = 0; i0 < n; i0 += BLOCK * kVT) {
            dist_t acc[kVT];
#pragma unroll
            for (uint32_t t = 0; t < kVT; ++t) acc[t] = dist_t(0);
            vertical_tile(acc, i0, 0, v);
            store_tile<kVT>(acc, i0);
        }
        __syncthreads();
        if (h) {
            // one thread per vector; the block distances of a vector are independent, so their
            // loads are issued together (U8: two or four blocks per round trip) and added in
            // block order
            const uint32_t nhs = h / kHDim;
            for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
                dist_t dd = dist[i];
                if constexpr (U8) {
                    constexpr uint32_t NB = BLOCK == 64 ? 4 : 2;  // more vectors per thread with one wave
                    for (uint32_t b = 0; b < nhs; b += NB) {
                        uint32_t x[NB][16];
#pragma unroll
                        for (uint32_t u = 0; u < NB; ++u)
                            if (b + u < nhs) fetch_block_u8(i, (b + u) * kHDim, x[u]);
#pragma unroll
                        for (uint32_t u = 0; u < NB; ++u)
                            if (b + u < nhs) dd += block_dist_u8(x[u], (b + u) * kHDim);
                    }
                } else {
                    for (uint32_t hd = 0; hd < h; hd += kHDim) dd += horizontal_thread(i, hd);
                }
                dist[i] = dd;
            }
            __syncthreads();
        }
        bytes += (unsigned long long)n * d * (U8 ? 1 : 4);
        mask_tombstones_all();
        uint32_t m = min(k - hl, n);
        push_m_smallest(m);
    }

    // searcher.hpp:284-372
    __device__ void filtered_start(uint32_t k, uint32_t passing) {
        for (uint32_t i = threadIdx.x; i < n; i += BLOCK) dist[i] = dist_t(0);
        uint32_t np = 0;
        for (uint32_t i0 = 0; i0 < n; i0 += BLOCK) {
            uint32_t i = i0 + threadIdx.x;
            bool p = i < n && selected(i);
            uint32_t s;
            uint32_t r = block_excl_scan<BLOCK>(p, red, &s);
            if (p) pos[np + r] = static_cast<uint16_t>(i);
            np += s;
        }
        __syncthreads();
        const float selection = static_cast<float>(passing) / static_cast<float>(n);
        if (h) {
            const uint32_t g = threadIdx.x / 16, l = threadIdx.x % 16;
            for (uint32_t j = g; j < np; j += G) {
                uint32_t i = pos[j];
                dist_t dd = dist[i];
                for (uint32_t hd = 0; hd < h; hd += kHDim) dd += horizontal16(i, hd, l);
                if (l == 0) dist[i] = dd;
            }
            __syncthreads();
            bytes += (unsigned long long)np * h * (U8 ? 1 : 4);
        }
        if (selection > (1.0f - 0.8f)) {
            for (uint32_t i0 = 0; i0 < n; i0 += BLOCK * kVT) {
                dist_t acc[kVT];
                load_tile<kVT>(acc, i0);
                vertical_tile(acc, i0, 0, v);
                store_tile<kVT>(acc, i0);
            }
            bytes += (unsigned long long)n * v * (U8 ? 1 : 4);
        } else {
            for (uint32_t j = threadIdx.x; j < np; j += BLOCK) {
                uint32_t i = pos[j];
                dist[i] = vertical_one(dist[i], i, 0, v);
            }
            bytes += (unsigned long long)np * v * (U8 ? 1 : 4);
        }
        __syncthreads();
        uint32_t m = min(k - hl, passing);
        for (uint32_t i = threadIdx.x; i < n; i += BLOCK)
            if (!selected(i) || tombstoned(i)) dist[i] = mask_value();
        __syncthreads();
        push_m_smallest(m);
    }

    // searcher.hpp:376-435. Returns the dimension index reached.
    // One tile of a warmup step: the vertical dims [cur, end) of slots tile_slot(i0, u) (u < VT);
    // returns the wave's count of dist >= tau over the tile.
    template <uint32_t VT>
    __device__ __forceinline__ uint32_t warmup_tile(uint32_t i0, uint32_t cur, uint32_t end, dist_t tau) {
        dist_t acc[VT];
        load_tile<VT>(acc, i0);
        vertical_tile<VT>(acc, i0, cur, end);
        store_tile<VT>(acc, i0);
        uint32_t cnt = 0;
#pragma unroll
        for (uint32_t u = 0; u < VT; ++u) {
            const uint32_t i = tile_slot<VT>(i0, u);
            cnt += static_cast<uint32_t>(__popcll(__ballot(i < n && acc[u] >= tau)));
        }
        return cnt;
    }

    __device__ uint32_t warmup(uint32_t k, uint32_t passing) {
        for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
            dist_t x = dist_t(0);
            if (tombstoned(i)) x = mask_value();
            if constexpr (FILT) {
                if (!selected(i)) x = mask_value();
            }
            dist[i] = x;
        }
        __syncthreads();
        if constexpr (FILT) {
            const float selection = static_cast<float>(passing) / static_cast<float>(n);
            if (selection < (1.0f - 0.8f)) return 0;
        }
        const uint32_t need = static_cast<uint32_t>(ceilf(0.8f * static_cast<float>(n)));
        if constexpr (SUMM) {
            // every checkpoint value is at hand: count all of them (ballots, four checkpoints
            // in flight), then take the first checkpoint where the loop below would stop
            const uint32_t nw = a.nw;
            if (nw == 0) return 0;
            for (uint32_t t = threadIdx.x; t < nw; t += BLOCK) wcnt[t] = 0;
            __syncthreads();
            const uint32_t lane = threadIdx.x & (kWave - 1);
            constexpr uint32_t RI = 3, RW = 8;  // slots per thread and checkpoints held in registers
            if (BLOCK == 64 && n <= RI * BLOCK && nw <= RW) {  // one-wave replays (large batches)
                // every value of the cluster in flight at once and kept: the chosen checkpoint's
                // distances come from registers
                float x[RI][RW];
#pragma unroll
                for (uint32_t r = 0; r < RI; ++r) {
                    const uint32_t i = r * BLOCK + threadIdx.x;
#pragma unroll
                    for (uint32_t u = 0; u < RW; ++u) x[r][u] = (u < nw && i < n) ? sb[(size_t)u * sn + i] : 0.0f;
                }
                uint32_t cnt[RW];
#pragma unroll
                for (uint32_t u = 0; u < RW; ++u) {
                    cnt[u] = 0;
                    if (u < nw) {
                        const dist_t tau = threshold(k, cpe[u]);
#pragma unroll
                        for (uint32_t r = 0; r < RI; ++r)
                            cnt[u] += static_cast<uint32_t>(
                                __popcll(__ballot(r * BLOCK + threadIdx.x < n && x[r][u] >= tau)));
                    }
                }
                if (lane == 0) {
#pragma unroll
                    for (uint32_t u = 0; u < RW; ++u)
                        if (u < nw && cnt[u]) atomicAdd(&wcnt[u], cnt[u]);
                }
                __syncthreads();
                const uint64_t hit = __ballot(lane < nw && wcnt[lane] >= need);
                const uint32_t tw = hit ? static_cast<uint32_t>(__ffsll(static_cast<long long>(hit)) - 1) : nw - 1;
                const uint32_t prev = cpe[tw];
                bytes += (unsigned long long)n * prev * 4;
#pragma unroll
                for (uint32_t r = 0; r < RI; ++r) {
                    float sel = 0.0f;
#pragma unroll
                    for (uint32_t u = 0; u < RW; ++u) sel = u == tw ? x[r][u] : sel;
                    const uint32_t i = r * BLOCK + threadIdx.x;
                    if (i < n) dist[i] = sel;
                }
                __syncthreads();
                return prev;
            }
            for (uint32_t t0 = 0; t0 < nw; t0 += 8) {
                dist_t tau[8];
                uint32_t cnt[8];
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) {
                    cnt[u] = 0;
                    if (t0 + u < nw) tau[u] = threshold(k, cpe[t0 + u]);
                }
                for (uint32_t i0 = 0; i0 < n; i0 += BLOCK) {
                    const uint32_t i = i0 + threadIdx.x;
                    float x[8];
#pragma unroll
                    for (uint32_t u = 0; u < 8; ++u) x[u] = (t0 + u < nw && i < n) ? sb[(size_t)(t0 + u) * sn + i] : 0.0f;
#pragma unroll
                    for (uint32_t u = 0; u < 8; ++u)
                        cnt[u] += static_cast<uint32_t>(__popcll(__ballot(t0 + u < nw && i < n && x[u] >= tau[u])));
                }
                if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
                    for (uint32_t u = 0; u < 8; ++u)
                        if (t0 + u < nw && cnt[u]) atomicAdd(&wcnt[t0 + u], cnt[u]);
                }
            }
            __syncthreads();
            // first checkpoint whose count reaches `need` (the loop below stops there), else the last
            const uint64_t hit = __ballot(lane < nw && wcnt[lane] >= need);
            const uint32_t tw = hit ? static_cast<uint32_t>(__ffsll(static_cast<long long>(hit)) - 1) : nw - 1;
            const uint32_t prev = cpe[tw];
            bytes += (unsigned long long)n * prev * 4;
            const float* vt = sb + (size_t)tw * sn;  // just read: served by the caches
            for (uint32_t i = threadIdx.x; i < n; i += BLOCK) dist[i] = vt[i];
            __syncthreads();
            return prev;
        }
        uint32_t cur = 0, t = 0, npruned = 0;
        while (npruned < need && cur < v) {
            const uint32_t end = min(cur + c_fetch[t], v);
            const dist_t tau = threshold(k, end);
            uint32_t cnt = 0;  // wave-uniform
            bool one_round = false;
            if constexpr (U8 && BLOCK == 64) {
                // a 16-dim U8 step of a one-wave scan over a cluster up to twice the tile: one round
                // of 32 loads instead of two tiles one after the other
                if (n > BLOCK * kVT && n <= BLOCK * 2 * kVT && end - cur <= 16) {
                    cnt = warmup_tile<2 * kVT>(0, cur, end, tau);
                    one_round = true;
                }
            }
            if (!one_round)
                for (uint32_t i0 = 0; i0 < n; i0 += BLOCK * kVT) cnt += warmup_tile<kVT>(i0, cur, end, tau);
            npruned = block_sum_uniform<BLOCK>(cnt, red);
            bytes += (unsigned long long)n * (end - cur) * (U8 ? 1 : 4);
            cur = end;
            ++t;
        }
        return cur;
    }

    // Remaining steps of one survivor i (partial distance x over the warmup dims): horizontal
    // blocks, then the vertical remainder in 64-dim chunks, each followed by its threshold test
    // (taus[]); stops at the first failed test. Returns whether i survives; x is its distance.
    __device__ __forceinline__ bool walk(uint32_t i, dist_t& x, uint32_t cur0) {
        constexpr unsigned long long elem = U8 ? 1 : 4;
        const uint32_t nhs = h / kHDim;
        bool alive = true;
        uint32_t t = 0;
        if constexpr (SUMM) {
            // stored block distances: 16 loads in flight ahead of the tests
            for (uint32_t b0 = 0; b0 < nhs && alive; b0 += 16) {
                float hv[16];
#pragma unroll
                for (uint32_t u = 0; u < 16; ++u) hv[u] = b0 + u < nhs ? sb[(size_t)(a.nw + b0 + u) * sn + i] : 0.0f;
#pragma unroll
                for (uint32_t u = 0; u < 16; ++u) {
                    if (b0 + u < nhs && alive) {
                        x += hv[u];
                        bytes_local += kHDim * elem;
                        alive = x < taus[t];
                        ++t;
                    }
                }
            }
        } else if constexpr (U8) {
            // the next block is in flight while the current one is tested
            uint32_t nx[16];
            if (nhs) fetch_block_u8(i, 0, nx);
#pragma unroll 1
            for (uint32_t b = 0; b < nhs; ++b) {
                uint32_t cx[16];
#pragma unroll
                for (uint32_t w = 0; w < 16; ++w) cx[w] = nx[w];
                if (b + 1 < nhs) fetch_block_u8(i, (b + 1) * kHDim, nx);
                x += block_dist_u8(cx, b * kHDim);
                bytes_local += kHDim * elem;
                if (!(x < taus[t])) return false;
                ++t;
            }
        } else {
            for (uint32_t hd = 0; hd < h; hd += kHDim, ++t) {
                x += horizontal_thread(i, hd);
                bytes_local += kHDim * elem;
                if (!(x < taus[t])) return false;
            }
        }
        if (!alive) return false;
        for (uint32_t cv = cur0; cv < v; cv += kHDim, ++t) {
            const uint32_t end = min(cv + kHDim, v);
            x = vertical_one(x, i, cv, end);
            bytes_local += (end - cv) * elem;
            if (!(x < taus[t])) return false;
        }
        return true;
    }

    // Prune and Merge (searcher.hpp:438-566). The survivors of the test at `cur` are compacted
    // once; each survivor's remaining steps depend only on its own partial distance and the step
    // thresholds (fixed for the cluster), so every thread walks its survivors to the end with no
    // barrier between steps. The survivors' keys stay in registers (kR per thread) for the merge.
    static constexpr uint32_t kR = 4;
    __device__ void prune_merge(uint32_t k, uint32_t cur) {
        dist_t tau = threshold(k, cur);
        uint32_t np = 0;
        for (uint32_t i0 = 0; i0 < n; i0 += BLOCK) {
            uint32_t i = i0 + threadIdx.x;
            bool valid = i < n;
            dist_t x = dist_t(0);
            if (valid) {
                x = dist[i];
                if (tombstoned(i)) {
                    x = mask_value();
                    dist[i] = x;
                }
            }
            bool p = valid && x < tau && selected(i);
            uint32_t s;
            uint32_t r = block_excl_scan<BLOCK>(p, red, &s);
            if (p) pos[np + r] = static_cast<uint16_t>(i);
            np += s;
        }
        if (np == 0) return;
        const uint32_t nhs = h / kHDim, nvs = (v - cur + kHDim - 1) / kHDim;
        for (uint32_t t = threadIdx.x; t < nhs + nvs; t += BLOCK)
            taus[t] = threshold(k, t < nhs ? cur + (t + 1) * kHDim : min(cur + h + (t - nhs + 1) * kHDim, d));
        __syncthreads();
        uint64_t key[kR];
        uint32_t pay[kR];
        uint32_t alive_cnt = 0;
        uint32_t r0 = 0;  // register slots already walked
        if constexpr (SUMM) {
            // no vertical remainder (the warmup ran to the last checkpoint): the stored block
            // distances of the thread's first WS survivors are loaded together, then tested in
            // order exactly as walk() does
            constexpr uint32_t WS = 2, WB = 12;
            if (BLOCK == 64 && cur == v && nhs <= WB) {  // one-wave replays (large batches)
                float hv[WS][WB];
                uint32_t si[WS];
#pragma unroll
                for (uint32_t r = 0; r < WS; ++r) {
                    const uint32_t j = threadIdx.x + r * BLOCK;
                    si[r] = j < np ? pos[j] : 0u;
#pragma unroll
                    for (uint32_t u = 0; u < WB; ++u)
                        hv[r][u] = (j < np && u < nhs) ? sb[(size_t)(a.nw + u) * sn + si[r]] : 0.0f;
                }
#pragma unroll
                for (uint32_t r = 0; r < WS; ++r) {
                    key[r] = kKeyInf;
                    pay[r] = 0;
                    if (threadIdx.x + r * BLOCK < np) {
                        dist_t x = dist[si[r]];
                        bool alive = true;
#pragma unroll
                        for (uint32_t u = 0; u < WB; ++u) {
                            if (u < nhs && alive) {
                                x += hv[r][u];
                                bytes_local += kHDim * 4;
                                alive = x < taus[u];
                            }
                        }
                        if (alive) {
                            key[r] = make_key(to_float(x), cid[si[r]]);
                            ++alive_cnt;
                        }
                    }
                }
                r0 = WS;
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < kR; ++r) {
            if (r < r0) continue;
            key[r] = kKeyInf;
            pay[r] = 0;
            const uint32_t j = threadIdx.x + r * BLOCK;
            if (j < np) {
                const uint32_t i = pos[j];
                const uint32_t id = cid[i];
                dist_t x = dist[i];
                if (walk(i, x, cur)) {
                    key[r] = make_key(to_float(x), id);
                    if constexpr (U8) pay[r] = x;
                    ++alive_cnt;
                }
            }
        }
        for (uint32_t j = threadIdx.x + kR * BLOCK; j < np; j += BLOCK) {  // beyond the register slots
            const uint32_t i = pos[j];
            dist_t x = dist[i];
            if (walk(i, x, cur)) {
                dist[i] = x;
                ++alive_cnt;
            } else {
                pos[j] = kPosNone;
            }
        }
        bytes_local += 4ull * alive_cnt;  // Merge reads one distance per survivor
        __syncthreads();
#pragma unroll
        for (uint32_t r = 0; r < kR; ++r)
            if (r * BLOCK < np)
                hl = list_insert_chunk<BLOCK>(key[r] != kKeyInf, key[r], pay[r], H, HP, hl, k, CK, CP, O, OP, red);
        for (uint32_t j0 = kR * BLOCK; j0 < np; j0 += BLOCK) {
            const uint32_t j = j0 + threadIdx.x;
            const uint32_t i = j < np ? pos[j] : kPosNone;
            const bool valid = i != kPosNone;
            hl = list_insert_chunk<BLOCK>(valid, valid ? key_of(i) : kKeyInf, valid ? pay_of(i) : 0u, H, HP, hl, k, CK,
                                          CP, O, OP, red);
        }
    }

    __device__ void run(uint32_t q) {
        load_query(q);
        const uint32_t k = a.kq[q];
        uint32_t nvis = a.nvisit[q];
        if (a.order_count) nvis = min(nvis, a.order_count[q]);
        const uint32_t* ord = a.order + (size_t)q * a.order_stride;
        for (uint32_t ci = 0; ci < nvis; ++ci) {
            if (ci % kRing == 0) {  // stage the parameters of the next kRing clusters
                __syncthreads();
                const uint32_t cnt = min(kRing, nvis - ci);
                for (uint32_t t = threadIdx.x; t < cnt; t += BLOCK) {
                    ClusterMeta mm;
                    const uint32_t c = ord[ci + t];
                    mm.c = c;
                    mm.n_live = a.L.n_live[c];
                    mm.n_used = a.L.n_used[c];
                    mm.stride = a.L.stride[c];
                    mm.n_tomb = a.L.n_tomb[c];
                    mm.data_off = a.L.data_off[c];
                    mm.slot_off = a.L.slot_off[c];
                    mm.tomb_off = a.L.tomb_off[c];
                    mm.aux = 0;
                    if constexpr (FILT) mm.aux = a.n_pass[c];
                    if constexpr (SUMM) mm.aux = a.soff[(size_t)q * a.order_stride + ci + t];
                    ring[t] = mm;
                }
                __syncthreads();
            }
            const ClusterMeta& mm = ring[ci % kRing];
            uint32_t passing = 0;
            if constexpr (FILT) {
                passing = mm.aux;
                if (passing == 0) continue;
                csel = a.sel + mm.slot_off;
            }
            if (mm.n_live == 0) continue;
            n = mm.n_used;
            S = mm.stride;
            base = a.L.data + mm.data_off;
            cid = a.L.ids + mm.slot_off;
            ntomb = mm.n_tomb;
            tbm = a.L.tomb + mm.tomb_off;
            hvec = ((size_t)v * S * (U8 ? 1 : 4)) % 16 == 0;
            if constexpr (SUMM) {
                sb = a.summ + ((size_t)a.qbase[q] + mm.aux) * a.ns;
                sn = (n + 3u) & ~3u;
            }
            if (hl < k) {
                if constexpr (FILT) filtered_start(k, passing);
                else start(k);
                continue;
            }
            prune_merge(k, warmup(k, passing));
        }
        // this thread's share of the algorithmic bytes, summed over the workgroup
        {
            unsigned long long x = bytes_local;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
            __syncthreads();
            if ((threadIdx.x & (kWave - 1)) == 0) CK[threadIdx.x / kWave] = x;
            __syncthreads();
            for (uint32_t w = 0; w < BLOCK / kWave; ++w) bytes += CK[w];
        }
        // searcher.hpp:568-585
        uint32_t cnt = hl;
        if (cnt > 0 && key_dist(H[cnt - 1]) == FLT_MAX) --cnt;
        cnt = min(cnt, k);
        for (uint32_t i = threadIdx.x; i < cnt; i += BLOCK) {
            a.out_ids[(size_t)q * a.out_stride + i] = key_id(H[i]);
            if (a.out_dists) a.out_dists[(size_t)q * a.out_stride + i] = key_dist(H[i]);
            if (a.out_idists) a.out_idists[(size_t)q * a.out_stride + i] = HP[i];
        }
        if (threadIdx.x == 0) {
            a.out_counts[q] = cnt;
            if (a.out_bytes) a.out_bytes[q] += bytes;
            if (a.stage_bytes) atomicAdd(a.stage_bytes, bytes);
        }
    }
};

// One workgroup per query (PERSIST = false); or a persistent grid whose workgroups take queries
// from the work queue a.queue (in perm order) until it is empty: every workgroup reaches the exit
// once the counter passes nq. A fresh Scanner per query keeps no state across the loop (the
// persistent instantiation is separate so the one-query kernel's code is unchanged by it).
template <bool U8, bool FILT, uint32_t BLOCK, bool SUMM = false, bool PERSIST = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(BLOCK == 64 ? 4 : 1))) void k_scan(ScanArgs a) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    if constexpr (PERSIST) {
        __shared__ uint32_t next;
        for (;;) {
            __syncthreads();  // the previous query's results are written
            if (threadIdx.x == 0) next = atomicAdd(a.queue, 1u);
            __syncthreads();
            const uint32_t i = next;
            if (i >= a.nq) return;
            Scanner<U8, FILT, BLOCK, SUMM> s(a, lds_raw);
            s.run(a.perm ? a.perm[i] : i);
        }
    } else {
        if (blockIdx.x >= a.nq) return;
        Scanner<U8, FILT, BLOCK, SUMM> s(a, lds_raw);
        s.run(a.perm ? a.perm[blockIdx.x] : blockIdx.x);
    }
}

}  // namespace pdxg
