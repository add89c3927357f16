// Dense low-precision first pass with an exact FP32 rerank.
//
// A dense copy of an F32 index holds every used slot as a row-major row in the rotated space:
// FP32 (for the rerank), BF16 or FP8 e4m3 (for the MFMA GEMM) and the squared norm accumulated in
// FP64. A batched search computes S = -2 q.x on the matrix cores (pdx_mma.hpp, FP32 accumulation),
// forms ||q||^2 + ||x||^2 + S in the GEMM's epilogue, keeps the R = k * oversample smallest per
// query, and reranks the R candidates with the exact FP32 PDX distance: the vertical fma chain over the
// first v dims, then the 64-dim horizontal blocks in the 16-lane order -- the value the PDX scan
// computes for a vector it does not prune (searcher.hpp:596-724 via distance_computers).
//
// Residual coding (when the level has centroids): the low-precision rows hold r = x - c, c the
// centroid of the row's cluster (an appended row takes its nearest centroid), and
// q.x = q.c + q.r with q.c from a BF16 MFMA GEMM over the centroids. r is about as small as the
// cluster's spread, so the low-precision error of q.r is that much smaller than the error of a
// low-precision q.x (this is what lets FP8 e4m3 rank neighbours whose distances differ by less
// than its rounding of whole rows).
//
// A row filter (the bitmap form of FilteredSearch's passing row ids, index.hpp:833-845) and
// deleted rows are applied by compacting the passing rows (in row order) into a packed
// low-precision matrix before the GEMM, so the GEMM's work scales with the passing rows.
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_fp8.h>

#include "pdx_kernels.hpp"

namespace pdxg {

constexpr uint32_t kDenseBF16 = 0, kDenseFP8 = 1;
constexpr float kFp8Range = 240.0f;  // scaled magnitudes stay below e4m3's largest finite 448
constexpr uint32_t kSelBuf = 256;    // per-wave gather buffer of the top-R select
constexpr uint32_t kSelMax = 256;    // largest R
constexpr uint32_t kCompactRows = 1024;  // rows per workgroup of the compaction kernels

__device__ __forceinline__ void atomic_max_pos(float* p, float v) {  // v >= 0: bit order = value order
    atomicMax(reinterpret_cast<unsigned int*>(p), __float_as_uint(v));
}

__device__ __forceinline__ double block_sum_f64_256(double x, double* s) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x / 64] = x;
    __syncthreads();
    double t = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) t += s[w];
    __syncthreads();
    return t;
}

__device__ __forceinline__ float wave_max(float m) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    return m;
}

// One workgroup (256) per used slot (blockIdx.x = slot, blockIdx.y = cluster): the FP32 row from
// the PDX layout (layout.hpp:21-87), its FP64 squared norm, id, liveness, cluster and the largest
// magnitude of what the low-precision copy will hold (x, or x - c with centroids `cent`).
__global__ __launch_bounds__(256) void k_dense_rows(LevelDev L, const unsigned long long* row_off, const float* cent,
                                                    float* x32, float* xn, uint32_t* ids, uint8_t* live,
                                                    uint32_t* rcl, float* maxabs) {
    __shared__ double red[4];
    const uint32_t c = blockIdx.y, i = blockIdx.x;
    const uint32_t n = L.n_used[c];
    if (i >= n) return;
    const uint32_t S = L.stride[c], d = L.d, v = L.v;
    const float* data = reinterpret_cast<const float*>(L.data + L.data_off[c]);
    const unsigned long long r = row_off[c] + i;
    double acc = 0.0;
    float mx = 0.0f;
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) {
        const uint32_t hj = j - v;
        const float x = j < v ? data[(size_t)j * S + i]
                              : data[(size_t)(v + (hj & ~63u)) * S + (size_t)i * kHDim + (hj & 63u)];
        x32[r * d + j] = x;
        acc += (double)x * (double)x;
        mx = fmaxf(mx, fabsf(cent ? x - cent[(size_t)c * d + j] : x));
    }
    const double t = block_sum_f64_256(acc, red);
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) atomic_max_pos(maxabs, mx);
    if (threadIdx.x == 0) {
        xn[r] = static_cast<float>(t);
        rcl[r] = c;
        const uint32_t id = L.ids[L.slot_off[c] + i];
        ids[r] = id;
        const bool tomb = L.n_tomb[c] != 0 && ((L.tomb[L.tomb_off[c] + (i >> 5)] >> (i & 31)) & 1u);
        live[r] = tomb ? 0 : 1;
    }
}

// id -> row map of the live rows.
__global__ void k_dense_map(const uint32_t* ids, const uint8_t* live, unsigned long long nr, uint32_t* row_of_id) {
    const unsigned long long r = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nr && live[r]) row_of_id[ids[r]] = static_cast<uint32_t>(r);
}

// Largest id among rows [0, nr).
__global__ void k_dense_max_id(const uint32_t* ids, unsigned long long nr, uint32_t* out) {
    const unsigned long long r = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = r < nr ? ids[r] : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), off)));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// Low-precision copy of rows [r0, r0 + nr): x (or x - c of the row's cluster when cent is given)
// as BF16 (unscaled) or FP8 e4m3 scaled by *sx, in rows of ld >= d elements whose tail is zero.
// One workgroup per row, grid-stride over rows (a dispatch holds at most 2^32 work-items per
// dimension).
__global__ __launch_bounds__(256) void k_dense_convert(const float* __restrict__ x32, void* xlp, unsigned long long r0,
                                                       unsigned long long nr, uint32_t d, uint32_t ld, uint32_t prec,
                                                       const float* sx, const float* __restrict__ cent,
                                                       const uint32_t* __restrict__ rcl) {
    const float scale = prec == kDenseBF16 ? 1.0f : *sx;
    for (unsigned long long rr = blockIdx.x; rr < nr; rr += gridDim.x) {
        const unsigned long long r = r0 + rr;
        const float* x = x32 + r * d;
        const float* c = cent ? cent + (size_t)rcl[r] * d : nullptr;
        for (uint32_t j = threadIdx.x; j < ld; j += blockDim.x) {
            const float y = j < d ? (c ? x[j] - c[j] : x[j]) : 0.0f;
            if (prec == kDenseBF16) reinterpret_cast<__hip_bfloat16*>(xlp)[r * ld + j] = __float2bfloat16(y);
            else reinterpret_cast<__hip_fp8_storage_t*>(xlp)[r * ld + j] =
                __hip_cvt_float_to_fp8(y * scale, __HIP_SATFINITE, __HIP_E4M3);
        }
    }
}

// BF16 copy of nr rows of d floats into rows of ld >= d elements whose tail is zero (grid-stride).
__global__ void k_to_bf16_rows(const float* __restrict__ src, unsigned long long nr, uint32_t d, uint32_t ld,
                               __hip_bfloat16* __restrict__ dst) {
    const unsigned long long n = nr * ld, step = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += step) {
        const unsigned long long r = e / ld;
        const uint32_t j = static_cast<uint32_t>(e % ld);
        dst[e] = __float2bfloat16(j < d ? src[r * d + j] : 0.0f);
    }
}

// FP32 squared norm of each centroid (FP64 accumulation), one workgroup per centroid.
__global__ __launch_bounds__(256) void k_dense_cnorm(const float* cent, uint32_t d, float* cn) {
    __shared__ double red[4];
    double acc = 0.0;
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) {
        const double x = cent[(size_t)blockIdx.x * d + j];
        acc += x * x;
    }
    const double t = block_sum_f64_256(acc, red);
    if (threadIdx.x == 0) cn[blockIdx.x] = static_cast<float>(t);
}

// Cluster of each appended row: argmin_c ||c||^2 - 2 x.c (G holds -2 x.c from a BF16 GEMM,
// column-major: column t = row t over the nc centroids); ties go to the smaller cluster. One wave
// per row. The anchor only sets which residual is coded: any centroid gives the same exact
// results, a near one a small residual.
__global__ __launch_bounds__(64) void k_dense_anchor(const float* __restrict__ G, const float* __restrict__ cn,
                                                     uint32_t nc, unsigned long long r0, uint32_t* rcl) {
    const uint32_t t = blockIdx.x, lane = threadIdx.x;
    uint64_t best = kKeyInf;
    for (uint32_t c = lane; c < nc; c += 64) {
        const float v = cn[c] + G[(size_t)t * nc + c];
        // order-preserving map of a float to an unsigned key (negative values included)
        const uint32_t b = __float_as_uint(v);
        const uint32_t o = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
        const uint64_t key = (static_cast<uint64_t>(o) << 32) | c;
        best = key < best ? key : best;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(best, off);
        best = o < best ? o : best;
    }
    if (lane == 0) rcl[r0 + t] = static_cast<uint32_t>(best & 0xFFFFFFFFu);
}

// FP8 scales from a largest magnitude: s = kFp8Range / m (values are multiplied by s before the
// conversion) and 1 / s (the GEMM's operand scale).
__global__ void k_dense_scales(const float* maxabs, float* s, float* inv_s) {
    const float m = *maxabs;
    const float sc = m > 0.0f ? kFp8Range / m : 1.0f;
    *s = sc;
    *inv_s = 1.0f / sc;
}

// Query statistics, one workgroup per query: ||q||^2 (FP64 accumulation) and, for FP8, the query's
// own scale s_q = kFp8Range / max |q_i| (so a query's FP8 form does not depend on its batch) and
// the factor (1 / s_q) * (1 / sx) that turns its GEMM column back into q.x units.
__global__ __launch_bounds__(256) void k_dense_qstats(const float* q, uint32_t d, uint32_t prec, const float* inv_sx,
                                                      float* qn, float* qsc, float* qfac) {
    __shared__ double red[4];
    __shared__ float mred[4];
    const float* row = q + (size_t)blockIdx.x * d;
    double acc = 0.0;
    float mx = 0.0f;
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) {
        const float x = row[j];
        acc += (double)x * (double)x;
        mx = fmaxf(mx, fabsf(x));
    }
    const double t = block_sum_f64_256(acc, red);
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) mred[threadIdx.x / 64] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < blockDim.x / 64; ++w) mx = fmaxf(mx, mred[w]);
        qn[blockIdx.x] = static_cast<float>(t);
        if (prec == kDenseBF16) {
            qsc[blockIdx.x] = 1.0f;
            qfac[blockIdx.x] = 1.0f;
        } else {
            const float sc = mx > 0.0f ? kFp8Range / mx : 1.0f;
            qsc[blockIdx.x] = sc;
            qfac[blockIdx.x] = (1.0f / sc) * *inv_sx;
        }
    }
}

// Row filter: pass = live && (no selection || selection[id] == 1). Per workgroup of kCompactRows
// rows: the passing count (k_dense_count) and then, after the offsets, the passing rows in row
// order with their squared norms (k_dense_compact).
__device__ __forceinline__ bool dense_pass(const uint8_t* live, const uint32_t* ids, const uint8_t* sel,
                                           unsigned long long r, unsigned long long nr) {
    return r < nr && live[r] && (!sel || sel[ids[r]] == 1);
}

__global__ __launch_bounds__(256) void k_dense_count(const uint8_t* live, const uint32_t* ids, const uint8_t* sel,
                                                     unsigned long long nr, uint32_t* counts) {
    __shared__ uint32_t red[8];
    const unsigned long long base = (unsigned long long)blockIdx.x * kCompactRows;
    uint32_t c = 0;
#pragma unroll
    for (uint32_t t = 0; t < kCompactRows / 256; ++t) c += dense_pass(live, ids, sel, base + t * 256 + threadIdx.x, nr);
    uint32_t tot;
    (void)block_excl_sum<256>(c, red, &tot);
    if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

// Exclusive offsets of the per-workgroup counts (one workgroup of 1024); offs[nb] = total.
__global__ __launch_bounds__(1024) void k_dense_offsets(const uint32_t* counts, uint32_t nb, uint32_t* offs) {
    __shared__ uint32_t red[17];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
        const uint32_t b = b0 + threadIdx.x;
        const uint32_t x = b < nb ? counts[b] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_sum<1024>(x, red, &tot);
        if (b < nb) offs[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) offs[nb] = carry;
}

__global__ __launch_bounds__(256) void k_dense_compact(const uint8_t* live, const uint32_t* ids, const uint8_t* sel,
                                                       const float* xn, const uint32_t* rcl, unsigned long long nr,
                                                       const uint32_t* offs, uint32_t* rowlist, float* cbias,
                                                       uint32_t* ccl) {
    __shared__ uint32_t red[8];
    const unsigned long long base = (unsigned long long)blockIdx.x * kCompactRows;
    uint32_t pos = offs[blockIdx.x];
#pragma unroll 1
    for (uint32_t t = 0; t < kCompactRows / 256; ++t) {
        const unsigned long long r = base + t * 256 + threadIdx.x;
        const bool p = dense_pass(live, ids, sel, r, nr);
        uint32_t tot;
        const uint32_t ex = block_excl_scan<256>(p, red, &tot);
        if (p) {
            rowlist[pos + ex] = static_cast<uint32_t>(r);
            cbias[pos + ex] = xn[r];
            ccl[pos + ex] = rcl[r];
        }
        pos += tot;
    }
}

// Packed low-precision rows: dst[p] = src[rowlist[p]] for p < n (row_bytes a multiple of 16).
// Grid-stride over the 16-byte words.
__global__ __launch_bounds__(256) void k_dense_gather(const uint8_t* __restrict__ src, const uint32_t* rowlist,
                                                      unsigned long long n, uint32_t row_bytes,
                                                      uint8_t* __restrict__ dst) {
    const uint32_t per_row = row_bytes / 16;
    const unsigned long long total = n * per_row, step = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += step) {
        const unsigned long long p = e / per_row;
        const uint32_t w = static_cast<uint32_t>(e % per_row);
        const uint4* s = reinterpret_cast<const uint4*>(src + (size_t)rowlist[p] * row_bytes);
        reinterpret_cast<uint4*>(dst + p * row_bytes)[w] = s[w];
    }
}

// Unfiltered bias: ||x||^2 for live rows, +inf for deleted ones.
__global__ void k_dense_bias(const float* xn, const uint8_t* live, float* bias, unsigned long long nr) {
    const unsigned long long r = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nr) bias[r] = live[r] ? xn[r] : __builtin_inff();
}

// Queries to BF16, or to FP8 scaled by their own s_q, in rows of ld >= d elements whose tail is
// zero (total = queries x ld).
__global__ void k_dense_qconvert(const float* q, unsigned long long total, uint32_t d, uint32_t ld, uint32_t prec,
                                 const float* qsc, void* qlp) {
    const unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const unsigned long long t = e / ld;
    const uint32_t j = static_cast<uint32_t>(e % ld);
    const float y = j < d ? q[t * d + j] : 0.0f;
    if (prec == kDenseBF16) reinterpret_cast<__hip_bfloat16*>(qlp)[e] = __float2bfloat16(y);
    else reinterpret_cast<__hip_fp8_storage_t*>(qlp)[e] = __hip_cvt_float_to_fp8(y * qsc[t], __HIP_SATFINITE, __HIP_E4M3);
}

// Running top-R of a stream of keys (distance bits << 32 | row) held by one wave. Keys below the
// current R-th key are gathered in LDS and merged (bitonic sort of gathered + list) when the
// gather buffer fills. The result is the exact R smallest keys of the stream, independent of the
// order keys are offered in.
struct WaveTopR {
    uint64_t* buf;  // LDS: [kSelBuf + kSelMax] gather area, then the sorted list
    uint32_t R, len, nb;
    uint64_t kth;  // current R-th key (all ones while fewer than R)

    __device__ void init(uint64_t* lds, uint32_t r) {
        buf = lds;
        R = r;
        len = 0;
        nb = 0;
        kth = kKeyInf;
    }
    __device__ uint64_t* list() const { return buf + kSelBuf; }
    __device__ void merge() {
        const uint32_t lane = threadIdx.x & 63;
        const uint32_t tot = nb + len;
        uint32_t n2 = 64;
        while (n2 < tot) n2 <<= 1;
        for (uint32_t t = lane; t < len; t += 64) buf[nb + t] = buf[kSelBuf + t];
        __syncthreads();
        for (uint32_t t = tot + lane; t < n2; t += 64) buf[t] = kKeyInf;
        __syncthreads();
        for (uint32_t kk = 2; kk <= n2; kk <<= 1)
            for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
                for (uint32_t t = lane; t < (n2 >> 1); t += 64) {
                    const uint32_t a = (t / jj) * 2 * jj + (t % jj), b = a + jj;
                    const bool asc = (a & kk) == 0;
                    const uint64_t x = buf[a], y = buf[b];
                    if ((x > y) == asc) {
                        buf[a] = y;
                        buf[b] = x;
                    }
                }
                __syncthreads();
            }
        len = min(tot, R);
        for (uint32_t t = lane; t < len; t += 64) buf[kSelBuf + t] = buf[t];
        __syncthreads();
        nb = 0;
        if (len == R) kth = buf[kSelBuf + R - 1];
    }
    // All 64 lanes call with their own key (valid = false for none).
    __device__ __forceinline__ void offer(uint64_t key, bool valid) {
        const uint32_t lane = threadIdx.x & 63;
        const bool take = valid && key < kth;
        const uint64_t bal = __ballot(take);
        if (take) buf[nb + static_cast<uint32_t>(__popcll(bal & ((1ull << lane) - 1)))] = key;
        nb += static_cast<uint32_t>(__popcll(bal));
        if (nb > kSelBuf - 64) {
            __syncthreads();
            merge();
        }
    }
    __device__ void finish() {
        if (nb) {
            __syncthreads();
            merge();
        }
    }
};

// One wave per query: the running list merged with the nseg partial lists of the first pass.
__global__ __launch_bounds__(64) void k_dense_merge(uint64_t* running, uint32_t* run_len, const uint64_t* part,
                                                    uint32_t nseg, uint32_t R) {
    __shared__ uint64_t lds[kSelBuf + kSelMax];
    const uint32_t q = blockIdx.x, lane = threadIdx.x;
    WaveTopR top;
    top.init(lds, R);
    uint64_t* L = running + (size_t)q * R;
    const uint32_t n0 = run_len[q];
    for (uint32_t t = 0; t < n0; t += 64) top.offer(t + lane < n0 ? L[t + lane] : kKeyInf, t + lane < n0);
    const uint64_t* P = part + (size_t)q * nseg * R;
    const uint32_t tot = nseg * R;
    for (uint32_t t = 0; t < tot; t += 64) {
        const uint64_t key = t + lane < tot ? P[t + lane] : kKeyInf;
        top.offer(key, key != kKeyInf);
    }
    top.finish();
    for (uint32_t t = lane; t < top.len; t += 64) L[t] = top.list()[t];
    if (lane == 0) run_len[q] = top.len;
}

// Exact FP32 PDX distance of each candidate (one lane per candidate), then the k smallest
// (distance, id) pairs. One wave per query. Candidate rows are row-major FP32.
__global__ __launch_bounds__(64) void k_dense_rerank(const float* __restrict__ qrot, const float* __restrict__ x32,
                                                     const uint32_t* __restrict__ ids, const uint64_t* running,
                                                     const uint32_t* run_len, uint32_t R, uint32_t d, uint32_t v,
                                                     const uint32_t* kq, uint32_t k_all, uint32_t out_stride,
                                                     uint32_t* out_ids, float* out_dists, uint32_t* out_counts) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    float* qs = reinterpret_cast<float*>(lds_raw);  // [d]
    uint64_t* keys = reinterpret_cast<uint64_t*>(lds_raw + ((d * 4 + 15) / 16) * 16);  // [kSelMax]
    const uint32_t q = blockIdx.x, lane = threadIdx.x;
    for (uint32_t j = lane; j < d; j += 64) qs[j] = qrot[(size_t)q * d + j];
    const uint32_t nc = run_len[q];
    __syncthreads();
    uint32_t n2 = 64;
    while (n2 < nc) n2 <<= 1;
    for (uint32_t t = lane; t < n2; t += 64) {
        uint64_t key = kKeyInf;
        if (t < nc) {
            const uint32_t row = static_cast<uint32_t>(running[(size_t)q * R + t]);
            const float* x = x32 + (size_t)row * d;
            float acc = 0.0f;
            for (uint32_t j = 0; j < v; ++j) {
                const float df = qs[j] - x[j];
                acc = __builtin_fmaf(df, df, acc);
            }
            for (uint32_t hb = v; hb < d; hb += kHDim) {
                float p[16];
#pragma unroll
                for (uint32_t l = 0; l < 16; ++l) {
                    const float df = qs[hb + l] - x[hb + l];
                    p[l] = df * df;
                }
#pragma unroll
                for (uint32_t rr = 1; rr < 4; ++rr)
#pragma unroll
                    for (uint32_t l = 0; l < 16; ++l) {
                        const float df = qs[hb + 16 * rr + l] - x[hb + 16 * rr + l];
                        p[l] = __builtin_fmaf(df, df, p[l]);
                    }
                acc += tree16_regs(p);
            }
            key = make_key(acc, ids[row]);
        }
        keys[t] = key;
    }
    __syncthreads();
    for (uint32_t kk = 2; kk <= n2; kk <<= 1)
        for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
            for (uint32_t t = lane; t < n2 / 2; t += 64) {
                const uint32_t a = (t / jj) * 2 * jj + (t % jj), b = a + jj;
                const bool asc = (a & kk) == 0;
                const uint64_t x = keys[a], y = keys[b];
                if ((x > y) == asc) {
                    keys[a] = y;
                    keys[b] = x;
                }
            }
            __syncthreads();
        }
    const uint32_t k = kq ? kq[q] : k_all;
    const uint32_t cnt = min(k, nc);
    for (uint32_t t = lane; t < cnt; t += 64) {
        out_ids[(size_t)q * out_stride + t] = key_id(keys[t]);
        if (out_dists) out_dists[(size_t)q * out_stride + t] = key_dist(keys[t]);
    }
    if (lane == 0) out_counts[q] = cnt;
}

// Appended rows (already normalised and rotated): FP32 copy, norm, id, liveness, id -> row.
__global__ __launch_bounds__(256) void k_dense_append(const float* rows, const uint32_t* new_ids, uint32_t d,
                                                      unsigned long long r0, float* x32, float* xn, uint32_t* ids,
                                                      uint8_t* live, uint32_t* row_of_id) {
    __shared__ double red[4];
    const unsigned long long r = r0 + blockIdx.x;
    double acc = 0.0;
    for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) {
        const float x = rows[(size_t)blockIdx.x * d + j];
        x32[r * d + j] = x;
        acc += (double)x * (double)x;
    }
    const double t = block_sum_f64_256(acc, red);
    if (threadIdx.x == 0) {
        xn[r] = static_cast<float>(t);
        const uint32_t id = new_ids[blockIdx.x];
        ids[r] = id;
        live[r] = 1;
        row_of_id[id] = static_cast<uint32_t>(r);
    }
}

// Deletion by id: the row goes dead and its id unmapped (ids without a live row are ignored).
__global__ void k_dense_delete(const uint32_t* del_ids, uint32_t n, uint32_t id_cap, uint32_t* row_of_id,
                               uint8_t* live) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t id = del_ids[t];
    if (id >= id_cap) return;
    const uint32_t r = row_of_id[id];
    if (r != ~0u) {
        live[r] = 0;
        row_of_id[id] = ~0u;
    }
}

}  // namespace pdxg
