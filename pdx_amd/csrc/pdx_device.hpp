// Device-side building blocks of the PDX search kernels (gfx950, wave64).
//
// Arithmetic contract shared with the CPU oracle:
//   * F32 "horizontal" distances (64-dim PDX blocks, full-d centroid distances, query norms) use
//     16 fma lanes (lane l owns dims l, l+16, ...) reduced as (i,i+8), (i,i+4), (i,i+1), (i,i+2) --
//     the AVX-512 order of include/pdx/distance_computers/avx512_computers.hpp:47-79.
//   * F32 "vertical" distances are one fma chain per vector in dimension order
//     (scalar_computers.hpp:19-44).
//   * U8 distances are exact unsigned integers (scalar_computers.hpp:77-134), computed with
//     v_dot4_u32_u8 through  sum (q-x)^2 = sum q^2 + sum x^2 - 2 sum q x  in wrapping u32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pdxg {

constexpr uint32_t kWave = 64;
constexpr uint32_t kHDim = 64;  // H_DIM_SIZE, common.hpp:45
constexpr uint64_t kKeyInf = ~0ull;

__device__ __forceinline__ uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// Candidate key: non-negative float distance bits (monotone as u32) then row id.
__device__ __forceinline__ uint64_t make_key(float dist, uint32_t id) {
    return (static_cast<uint64_t>(f2u(dist)) << 32) | id;
}
__device__ __forceinline__ float key_dist(uint64_t k) { return u2f(static_cast<uint32_t>(k >> 32)); }
__device__ __forceinline__ uint32_t key_id(uint64_t k) { return static_cast<uint32_t>(k); }

// DPP row rotate-right by N within each 16-lane row: lane l receives lane (l - N) mod 16, so
// rotating by 16 - M reads lane l + M.
template <int N>
__device__ __forceinline__ float row_ror(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x120 + N, 0xf, 0xf, false));
}
template <int N>
__device__ __forceinline__ uint32_t row_ror_u(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x120 + N, 0xf, 0xf, false));
}

// AVX-512 reduction tree over the 16 lanes of a DPP row; lane 0 of the row holds the result:
// x_i = s_i + s_{i+8}; r_i = x_i + x_{i+4}; (r_0 + r_1) + (r_2 + r_3).
__device__ __forceinline__ float tree16(float acc) {
    acc = acc + row_ror<8>(acc);   // + lane l+8
    acc = acc + row_ror<12>(acc);  // + lane l+4
    acc = acc + row_ror<15>(acc);  // + lane l+1
    acc = acc + row_ror<14>(acc);  // + lane l+2
    return acc;
}
// The same tree for 16 per-lane partial sums held by one thread (lane l's sum in p[l]); equal to
// lane 0 of tree16 bit for bit.
__device__ __forceinline__ float tree16_regs(const float (&p)[16]) {
    float x[8], r[4];
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = p[i] + p[i + 8];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = x[i] + x[i + 4];
    return (r[0] + r[1]) + (r[2] + r[3]);
}
__device__ __forceinline__ uint32_t sum16_u(uint32_t acc) {
    acc += row_ror_u<8>(acc);
    acc += row_ror_u<4>(acc);
    acc += row_ror_u<2>(acc);
    acc += row_ror_u<1>(acc);
    return acc;
}

// sum over dims of (q-x)^2 for 4 packed u8 pairs, exact (wrapping u32 identity).
__device__ __forceinline__ uint32_t sqdist4(uint32_t q, uint32_t x, uint32_t qq) {
    uint32_t xx = __builtin_amdgcn_udot4(x, x, qq, false);
    uint32_t qx = __builtin_amdgcn_udot4(q, x, 0u, false);
    return xx - 2u * qx;
}
__device__ __forceinline__ uint32_t sumsq4(uint32_t q) { return __builtin_amdgcn_udot4(q, q, 0u, false); }

// ------------------------------------------------------------------------------------------
// Block-wide helpers. All threads of the block must call them (uniform control flow).
// s_red: LDS scratch of at least BLOCK/64 + 1 words.
// ------------------------------------------------------------------------------------------
template <uint32_t BLOCK>
__device__ __forceinline__ uint32_t block_count(bool p, uint32_t* s_red) {
    constexpr uint32_t NW = BLOCK / kWave;
    const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    uint64_t b = __ballot(p);
    if (lane == 0) s_red[w] = static_cast<uint32_t>(__popcll(b));
    __syncthreads();
    uint32_t tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < NW; ++i) tot += s_red[i];
    __syncthreads();
    return tot;
}

// Block-wide sum of a wave-uniform value (each wave contributes its value once).
template <uint32_t BLOCK>
__device__ __forceinline__ uint32_t block_sum_uniform(uint32_t wave_value, uint32_t* s_red) {
    constexpr uint32_t NW = BLOCK / kWave;
    if constexpr (NW == 1) return wave_value;
    const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) s_red[w] = wave_value;
    __syncthreads();
    uint32_t tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < NW; ++i) tot += s_red[i];
    __syncthreads();
    return tot;
}

// Exclusive prefix count of p in thread order; *total receives the block-wide count.
template <uint32_t BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan(bool p, uint32_t* s_red, uint32_t* total) {
    constexpr uint32_t NW = BLOCK / kWave;
    const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    uint64_t b = __ballot(p);
    uint64_t lt = lane == 0 ? 0ull : (b & ((~0ull) >> (64 - lane)));
    uint32_t r = static_cast<uint32_t>(__popcll(lt));
    if (lane == 0) s_red[w] = static_cast<uint32_t>(__popcll(b));
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < NW; ++i) {
        uint32_t c = s_red[i];
        off += (i < w) ? c : 0u;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return off + r;
}

// Exclusive prefix sum of x in thread order; *total receives the block-wide sum.
template <uint32_t BLOCK>
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t x, uint32_t* s_red, uint32_t* total) {
    constexpr uint32_t NW = BLOCK / kWave;
    const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    uint32_t inc = x;
#pragma unroll
    for (uint32_t off = 1; off < kWave; off <<= 1) {
        uint32_t y = __shfl_up(inc, off);
        if (lane >= off) inc += y;
    }
    if (lane == kWave - 1) s_red[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < NW; ++i) {
        uint32_t c = s_red[i];
        off += (i < w) ? c : 0u;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return off + inc - x;
}

// In-LDS bitonic sort (ascending) of n2 (power of two) keys with payloads.
template <uint32_t BLOCK>
__device__ __forceinline__ void bitonic_sort(uint64_t* key, uint32_t* pay, uint32_t n2) {
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < (n2 >> 1); t += BLOCK) {
                uint32_t i = (t / j) * 2 * j + (t % j);
                uint32_t p = i + j;
                bool asc = (i & k) == 0;
                uint64_t a = key[i], b = key[p];
                if ((a > b) == asc) {
                    key[i] = b;
                    key[p] = a;
                    if (pay) {
                        uint32_t x = pay[i];
                        pay[i] = pay[p];
                        pay[p] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Number of entries of the ascending array a[0..n) that are < key.
__device__ __forceinline__ uint32_t lower_bound(const uint64_t* a, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Merge two ascending lists (unique keys) into out, keeping the first `cap` entries.
// Returns min(cap, na + nb). Ends with a barrier; out must not alias a or b.
template <uint32_t BLOCK>
__device__ __forceinline__ uint32_t merge_sorted(const uint64_t* a, const uint32_t* ap, uint32_t na,
                                                 const uint64_t* b, const uint32_t* bp, uint32_t nb,
                                                 uint64_t* out, uint32_t* outp, uint32_t cap) {
    for (uint32_t i = threadIdx.x; i < na; i += BLOCK) {
        uint32_t r = i + lower_bound(b, nb, a[i]);
        if (r < cap) {
            out[r] = a[i];
            outp[r] = ap[i];
        }
    }
    for (uint32_t i = threadIdx.x; i < nb; i += BLOCK) {
        uint32_t r = i + lower_bound(a, na, b[i]);
        if (r < cap) {
            out[r] = b[i];
            outp[r] = bp[i];
        }
    }
    __syncthreads();
    uint32_t n = na + nb;
    return n < cap ? n : cap;
}

constexpr uint32_t kRankSortMax = 256;

// Bounded sorted list in LDS (keys L / payload LP, length len, capacity cap) receiving one
// optional candidate per thread. Keeps the `cap` smallest keys seen -- identical to pushing the
// candidates one by one into a max-heap of size cap with a strict-less replacement test
// (searcher.hpp:542-566), because keys are unique. Scratch: ck/cp (BLOCK entries), ok/op (cap).
template <uint32_t BLOCK>
__device__ __forceinline__ uint32_t list_insert_chunk(bool valid, uint64_t key, uint32_t pay, uint64_t* L,
                                                      uint32_t* LP, uint32_t len, uint32_t cap, uint64_t* ck,
                                                      uint32_t* cp, uint64_t* ok, uint32_t* op,
                                                      uint32_t* s_red) {
    if (len == cap && cap > 0) valid = valid && key < L[cap - 1];
    if (cap == 0) valid = false;
    uint32_t s = 0;
    uint32_t r = block_excl_scan<BLOCK>(valid, s_red, &s);
    if (s == 0) return len;
    if (s <= kRankSortMax) {
        // few candidates: each one's rank is the number of smaller keys (keys are unique), then
        // an in-place scatter; 3 barriers instead of the sort network's log^2 stages
        if (valid) {
            ck[r] = key;
            cp[r] = pay;
        }
        __syncthreads();
        uint32_t rank = 0;
        if (valid)
            for (uint32_t t = 0; t < s; ++t) rank += ck[t] < key;
        __syncthreads();
        if (valid) {
            ck[rank] = key;
            cp[rank] = pay;
        }
        __syncthreads();
    } else {
        uint32_t n2 = 1;
        while (n2 < s) n2 <<= 1;
        if (valid) {
            ck[r] = key;
            cp[r] = pay;
        }
        for (uint32_t i = s + threadIdx.x; i < n2; i += BLOCK) {
            ck[i] = kKeyInf;
            cp[i] = 0;
        }
        __syncthreads();
        bitonic_sort<BLOCK>(ck, cp, n2);
    }
    uint32_t nl = merge_sorted<BLOCK>(L, LP, len, ck, cp, s, ok, op, cap);
    for (uint32_t i = threadIdx.x; i < nl; i += BLOCK) {
        L[i] = ok[i];
        LP[i] = op[i];
    }
    __syncthreads();
    return nl;
}

}  // namespace pdxg
