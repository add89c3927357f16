// Hand-written gfx950 MFMA kernels of the dense first pass: the low-precision GEMM of rows x
// queries (BF16 v_mfma_f32_32x32x16_bf16, or FP8 e4m3 v_mfma_scale_f32_32x32x64_f8f6f4 with unit
// scales; FP32 accumulation), with the top-R selection fused into its epilogue so the score matrix
// never reaches HBM, and the same GEMM writing -2 a.b (the -2 q.c centroid scores).
//
// Operand layout (measured on the MI355X with an in-bounds probe of both builtins): lane l holds
// row (l mod 32) of A and column (l mod 32) of B; element j of A in lane half g = l / 32 is paired
// with element j of B in the same half, so any K order applied to both operands gives the same
// product. Here lane half g takes bytes [32 g, 32 g + 32) of every 64-byte K step of its row (FP8:
// the 32 K values of one MFMA; BF16: 16 K values, 8 per MFMA of the step's two). The 32 x 32 FP32
// result: lane l, element i = row 8 (i / 4) + 4 g + (i mod 4), column l mod 32.
//
// Workgroup: 512 threads (8 waves, one workgroup per CU) own a tile of 256 queries and walk
// 256-row tiles of a row segment; wave w computes rows 64 (w mod 4) .. +64 x queries
// 128 (w / 4) .. +128 as 2 x 4 MFMA blocks. K streams through a ring of four 32 KiB LDS stages
// (64 bytes of K for the 256 rows and the 256 queries) filled by direct global -> LDS loads
// (global_load_lds_dwordx4), three stages in flight while one is multiplied, one barrier per stage.
// Within a stage row the four 16-byte pieces are XOR-swizzled by the row number so the fragment
// reads (32 rows x 16 bytes per half-wave) hit distinct LDS banks.
//
// Row strides (row_bytes) are multiples of 64 with zero padding past the data: the padded K adds
// 0 x 0 products, so the result is that of the unpadded rows.
#pragma once
#include <hip/hip_bf16.h>

#include "pdx_kernels.hpp"

namespace pdxg {

typedef float mma_f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 mma_bf16x8 __attribute__((ext_vector_type(8)));
typedef int mma_i32x8 __attribute__((ext_vector_type(8)));
typedef int mma_i32x4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kMmaThreads = 512;
constexpr uint32_t kMmaRows = 256;                 // rows per tile
constexpr uint32_t kMmaQ = 256;                    // queries per workgroup
constexpr uint32_t kMmaKB = 64;                    // K bytes per stage (and the row-stride multiple)
constexpr uint32_t kMmaStages = 4;                 // LDS ring depth
constexpr uint32_t kMmaBOff = kMmaRows * kMmaKB;   // B's part of a stage
constexpr uint32_t kMmaStageBytes = (kMmaRows + kMmaQ) * kMmaKB;   // 32 KiB
constexpr uint32_t kMmaLds = kMmaStages * kMmaStageBytes;          // 128 KiB
constexpr uint32_t kMmaLoads = 4;                  // direct loads per thread per stage (2 A, 2 B)
constexpr uint32_t kMmaCap = 512;                  // candidate list per (query, segment)
constexpr uint32_t kMmaKPL = kMmaCap / 64;         // keys per lane of the wave sort

static_assert(kMmaStageBytes == kMmaThreads * 16 * kMmaLoads, "one 16-byte piece per thread per load");

// FP8: one scaled MFMA per 64-byte step (format 0 = e4m3 for both operands, scale exponents 127 =
// 1.0). BF16: two MFMAs, the first on bytes 0..15 of the lane's 32, the second on bytes 16..31.
template <uint32_t PREC>
__device__ __forceinline__ mma_f32x16 mma_step(const mma_i32x8& a, const mma_i32x8& b, mma_f32x16 c) {
    if constexpr (PREC == 1) {
        return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, 127, 0, 127);
    } else {
        const mma_i32x4 a0 = {a[0], a[1], a[2], a[3]}, a1 = {a[4], a[5], a[6], a[7]};
        const mma_i32x4 b0 = {b[0], b[1], b[2], b[3]}, b1 = {b[4], b[5], b[6], b[7]};
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mma_bf16x8, a0), __builtin_bit_cast(mma_bf16x8, b0),
                                                    c, 0, 0, 0);
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mma_bf16x8, a1),
                                                       __builtin_bit_cast(mma_bf16x8, b1), c, 0, 0, 0);
    }
}

struct MmaArgs {
    const uint8_t* A;          // rows [m x row_bytes], low precision
    const uint8_t* B;          // queries / second operand [n x row_bytes]
    uint32_t m, n, row_bytes;  // row_bytes: a multiple of kMmaKB, zero past the data
    // write mode: out[row * out_rs + col * out_cs] = -2 a.b  (col = B row)
    float* out;
    uint32_t out_rs, out_cs;
    // top-R mode
    const float* bias;         // per row: ||x||^2 (+inf: dead)
    const uint32_t* rowlist;   // packed rows -> row number (or null: row0 + r)
    unsigned long long row0;
    const float* qn;           // ||q||^2
    const float* qfac;         // GEMM units -> q.x units
    const float* qc;           // -2 q.c per (centroid, query): qc[c * n + q], or null
    const uint32_t* cl;        // cluster per row (with qc)
    uint32_t seg_rows, nseg, R;
    uint64_t* lists;           // [n][nseg][kMmaCap] candidate keys
    uint64_t* part;            // [n][nseg][R] top-R per (query, segment), padded with all ones
    uint64_t* bound;           // [n] per query, shared by its segments (all ones at launch), or null
};

// Swizzle of the 16-byte pieces of a stage row: piece p of row r sits in slot p ^ ((r / 4) mod 4).
// A fragment read (16-byte pieces of 32 consecutive rows, one piece index per half-wave) then has
// no LDS bank conflict: measured on the MI355X with the SQ_LDS_BANK_CONFLICT counter, which reads
// zero for this pattern and half the active cycles for ((r / 2) ^ (r / 8)) mod 4.
__device__ __forceinline__ uint32_t mma_swz(uint32_t r) { return (r >> 2) & 3u; }

// s_waitcnt with vmcnt = n (n < 16), expcnt and lgkmcnt left open.
template <uint32_t N>
__device__ __forceinline__ void mma_wait_vm() {
    static_assert(N < 16, "vmcnt immediate");
    __builtin_amdgcn_s_waitcnt(0xF70 | N);
}

// Direct loads of stage k0 (64 bytes of K of the tile's 256 rows and 256 queries) into ring slot
// `buf`. Each load instruction fills 1 KiB of LDS (16 rows x 4 pieces, lane i -> byte 16 i), so a
// lane loads the global piece that the swizzle puts at its LDS position. Rows past the matrix read
// its last row (their products are never used).
__device__ __forceinline__ void mma_issue(const MmaArgs& a, uint8_t* ring, uint32_t buf, uint32_t r0, uint32_t q0,
                                          uint32_t k0) {
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63, rin = lane >> 2, slot = lane & 3;
#pragma unroll
    for (uint32_t j = 0; j < kMmaLoads; ++j) {
        const bool isb = j >= 2;
        const uint32_t grp = w * 2 + (j & 1);
        const uint32_t r = grp * 16 + rin;
        const uint32_t p = slot ^ mma_swz(r);
        const uint32_t gr = isb ? min(q0 + r, a.n - 1) : min(r0 + r, a.m - 1);
        const uint8_t* src = (isb ? a.B : a.A) + (size_t)gr * a.row_bytes + k0 + p * 16;
        uint8_t* dst = ring + buf * kMmaStageBytes + (isb ? kMmaBOff : 0u) + grp * 1024;
        __builtin_amdgcn_global_load_lds(
            (__attribute__((address_space(1))) void*)(const_cast<uint8_t*>(src)),
            (__attribute__((address_space(3))) void*)(dst), 16, 0, 0);
    }
}

// Fragments of one stage from LDS. The swizzle depends on the row only through its value mod 32,
// so every fragment of a lane has the same piece slots: lane half g finds piece 2g at x and piece
// 2g + 1 at y (x, y: that lane's addresses in the first fragment), and fragment f sits 2 KiB
// (32 rows) after fragment 0. Each block issues its reads and waits for them: its outputs are
// early-clobber (no output may share a register with an address that a later read of the block
// still needs) and no result is in flight when it ends. The compiler does not see these LDS
// reads, which keeps it from draining the direct loads of later stages before them (the stage's
// own loads are waited for explicitly).
//
// mma_frags4: fragments 0 and 1 of A (ax, ay) and of B (bx, by); mma_frags2: fragments 2 and 3
// of B.
__device__ __forceinline__ void mma_frags4(uint32_t ax, uint32_t ay, uint32_t bx, uint32_t by, mma_i32x8 (&af)[2],
                                           mma_i32x8& b0, mma_i32x8& b1) {
    mma_i32x4 x[8];
    asm volatile(
        "ds_read_b128 %0, %8\n\tds_read_b128 %1, %9\n\t"
        "ds_read_b128 %2, %8 offset:2048\n\tds_read_b128 %3, %9 offset:2048\n\t"
        "ds_read_b128 %4, %10\n\tds_read_b128 %5, %11\n\t"
        "ds_read_b128 %6, %10 offset:2048\n\tds_read_b128 %7, %11 offset:2048\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(x[0]), "=&v"(x[1]), "=&v"(x[2]), "=&v"(x[3]), "=&v"(x[4]), "=&v"(x[5]), "=&v"(x[6]), "=&v"(x[7])
        : "v"(ax), "v"(ay), "v"(bx), "v"(by)
        : "memory");
    af[0] = __builtin_shufflevector(x[0], x[1], 0, 1, 2, 3, 4, 5, 6, 7);
    af[1] = __builtin_shufflevector(x[2], x[3], 0, 1, 2, 3, 4, 5, 6, 7);
    b0 = __builtin_shufflevector(x[4], x[5], 0, 1, 2, 3, 4, 5, 6, 7);
    b1 = __builtin_shufflevector(x[6], x[7], 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ __forceinline__ void mma_frags2(uint32_t bx, uint32_t by, mma_i32x8& b2, mma_i32x8& b3) {
    mma_i32x4 x[4];
    asm volatile(
        "ds_read_b128 %0, %4 offset:4096\n\tds_read_b128 %1, %5 offset:4096\n\t"
        "ds_read_b128 %2, %4 offset:6144\n\tds_read_b128 %3, %5 offset:6144\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(x[0]), "=&v"(x[1]), "=&v"(x[2]), "=&v"(x[3])
        : "v"(bx), "v"(by)
        : "memory");
    b2 = __builtin_shufflevector(x[0], x[1], 0, 1, 2, 3, 4, 5, 6, 7);
    b3 = __builtin_shufflevector(x[2], x[3], 0, 1, 2, 3, 4, 5, 6, 7);
}

// First stages of a row tile into the ring (after retiring every vector memory operation the
// caller left: the main loop's counts assume the ring's loads are the only ones in flight, and
// stores count too on this architecture).
__device__ __forceinline__ void mma_prologue(const MmaArgs& a, uint8_t* ring, uint32_t r0, uint32_t q0) {
    mma_wait_vm<0>();
    const uint32_t nst = a.row_bytes / kMmaKB;
    for (uint32_t s = 0; s < 3 && s < nst; ++s) mma_issue(a, ring, s, r0, q0, s * kMmaKB);
}

// acc[rb][qb] += A(rows of this wave) . B(queries of this wave) over one row tile, all of K, after
// mma_prologue of the same tile (vector memory operations issued in between must have retired
// before the first stage's count is waited for; the compiler's own waits for their results do
// that). Leaves no direct load in flight; the caller's next barrier frees the ring.
template <uint32_t PREC>
__device__ __forceinline__ void mma_tile(const MmaArgs& a, uint8_t* ring, uint32_t r0, uint32_t q0,
                                         mma_f32x16 (&acc)[2][4]) {
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 5, l32 = lane & 31;
    const uint32_t wr = (w & 3) * 64, wq = (w >> 2) * 128;
    const uint32_t nst = a.row_bytes / kMmaKB;
    // per-lane LDS offsets (within a stage) of the two pieces of the wave's first A and B fragment
    static_assert(kMmaRows % 32 == 0 && kMmaBOff % 2048 == 0, "fragment rows keep their swizzle");
    const uint32_t s0 = (2 * g) ^ mma_swz(l32);
    const uint32_t oax = (wr + l32) * kMmaKB + s0 * 16, oay = (wr + l32) * kMmaKB + (s0 ^ 1) * 16;
    const uint32_t obx = kMmaBOff + (wq + l32) * kMmaKB + s0 * 16, oby = kMmaBOff + (wq + l32) * kMmaKB + (s0 ^ 1) * 16;
    const uint32_t ring_addr = static_cast<uint32_t>(
        reinterpret_cast<uintptr_t>((__attribute__((address_space(3))) uint8_t*)(ring)));
    for (uint32_t c = 0; c < nst; ++c) {
        // stage c has landed once only the loads of the (up to two) stages after it are pending
        const uint32_t ahead = min(2u, nst - 1 - c);
        if (ahead == 2) mma_wait_vm<2 * kMmaLoads>();
        else if (ahead == 1) mma_wait_vm<kMmaLoads>();
        else mma_wait_vm<0>();
        __builtin_amdgcn_s_barrier();  // every wave's part of stage c is in; stage c - 1 is read
        if (c + 3 < nst) mma_issue(a, ring, (c + 3) & 3, r0, q0, (c + 3) * kMmaKB);
        const uint32_t base = ring_addr + (c & 3) * kMmaStageBytes;
        mma_i32x8 af[2], b0, b1;
        mma_frags4(base + oax, base + oay, base + obx, base + oby, af, b0, b1);
#pragma unroll
        for (uint32_t rb = 0; rb < 2; ++rb) {
            acc[rb][0] = mma_step<PREC>(af[rb], b0, acc[rb][0]);
            acc[rb][1] = mma_step<PREC>(af[rb], b1, acc[rb][1]);
        }
        mma_i32x8 b2, b3;
        mma_frags2(base + obx, base + oby, b2, b3);
#pragma unroll
        for (uint32_t rb = 0; rb < 2; ++rb) {
            acc[rb][2] = mma_step<PREC>(af[rb], b2, acc[rb][2]);
            acc[rb][3] = mma_step<PREC>(af[rb], b3, acc[rb][3]);
        }
    }
}

// out[row * out_rs + col * out_cs] = -2 A(row) . B(col) for every row < m and column < n. Grid:
// (row tiles of 256, column tiles of 256).
template <uint32_t PREC>
__global__ __launch_bounds__(kMmaThreads) void k_mma_dot(MmaArgs a) {
    extern __shared__ __align__(16) uint8_t mma_lds[];
    const uint32_t r0 = blockIdx.x * kMmaRows, q0 = blockIdx.y * kMmaQ;
    mma_f32x16 acc[2][4];
#pragma unroll
    for (uint32_t rb = 0; rb < 2; ++rb)
#pragma unroll
        for (uint32_t qb = 0; qb < 4; ++qb)
#pragma unroll
            for (uint32_t i = 0; i < 16; ++i) acc[rb][qb][i] = 0.0f;
    mma_prologue(a, mma_lds, r0, q0);
    mma_tile<PREC>(a, mma_lds, r0, q0, acc);
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 5, l32 = lane & 31;
    const uint32_t wr = (w & 3) * 64, wq = (w >> 2) * 128;
#pragma unroll
    for (uint32_t rb = 0; rb < 2; ++rb)
#pragma unroll
        for (uint32_t qb = 0; qb < 4; ++qb) {
            const uint32_t col = q0 + wq + qb * 32 + l32;
#pragma unroll
            for (uint32_t i = 0; i < 16; ++i) {
                const uint32_t row = r0 + wr + rb * 32 + 8 * (i / 4) + 4 * g + (i % 4);
                if (row < a.m && col < a.n)
                    a.out[(size_t)row * a.out_rs + (size_t)col * a.out_cs] = -2.0f * acc[rb][qb][i];
            }
        }
}

// Ascending bitonic sort of the kMmaKPL x 64 keys held by a wave (key index = lane * kMmaKPL + j),
// unrolled at compile time: every stage's partner is a fixed register (stride < kMmaKPL) or a
// fixed lane xor (stride >= kMmaKPL), so no key array is indexed at run time.
template <uint32_t SIZE, uint32_t STRIDE>
__device__ __forceinline__ void wave_sort_step(uint64_t (&k)[kMmaKPL], uint32_t lane) {
    if constexpr (STRIDE >= kMmaKPL) {  // partner in lane ^ (STRIDE / kMmaKPL), same j
        constexpr uint32_t lx = STRIDE / kMmaKPL;
        const bool lower = ((lane * kMmaKPL) & STRIDE) == 0;
#pragma unroll
        for (uint32_t j = 0; j < kMmaKPL; ++j) {
            const bool asc = ((lane * kMmaKPL + j) & SIZE) == 0;
            const uint64_t o = ((uint64_t)(uint32_t)__shfl_xor((int)(k[j] >> 32), lx) << 32) |
                               (uint32_t)__shfl_xor((int)(uint32_t)k[j], lx);
            const uint64_t mn = k[j] < o ? k[j] : o, mx = k[j] < o ? o : k[j];
            k[j] = (lower == asc) ? mn : mx;
        }
    } else {  // partner j ^ STRIDE in this lane
#pragma unroll
        for (uint32_t j = 0; j < kMmaKPL; ++j) {
            constexpr uint32_t st = STRIDE;
            if ((j & st) == 0) {
                const uint32_t p = j | st;
                const bool asc = ((lane * kMmaKPL + j) & SIZE) == 0;
                const uint64_t x = k[j], y = k[p];
                const bool sw = asc ? (x > y) : (x < y);
                k[j] = sw ? y : x;
                k[p] = sw ? x : y;
            }
        }
    }
    if constexpr (STRIDE > 1) wave_sort_step<SIZE, STRIDE / 2>(k, lane);
}

template <uint32_t SIZE>
__device__ __forceinline__ void wave_sort_stage(uint64_t (&k)[kMmaKPL], uint32_t lane) {
    wave_sort_step<SIZE, SIZE / 2>(k, lane);
    if constexpr (SIZE < kMmaCap) wave_sort_stage<SIZE * 2>(k, lane);
}

__device__ __forceinline__ void wave_sort(uint64_t (&k)[kMmaKPL]) { wave_sort_stage<2>(k, threadIdx.x & 63); }

// Keeps the keys of one candidate list (one wave; `cnt` <= kMmaCap keys at src) that are <= bound,
// and of those the R smallest: written to dst[0..) -- in no particular order when at most R are
// kept, else the R smallest in ascending order -- and, with `pad`, all ones up to dst[R). src
// and dst may be the same list: every key is in registers before the first store. Returns the
// kept count (<= R); *kth = the largest kept key when R are kept, else the bound (wave-uniform).
__device__ __forceinline__ uint32_t mma_prune(const uint64_t* src, uint32_t cnt, uint32_t R, uint64_t bound,
                                              uint64_t* dst, bool pad, uint64_t* kth) {
    const uint32_t lane = threadIdx.x & 63;
    uint64_t k[kMmaKPL];
    bool keep[kMmaKPL];
    uint32_t pos[kMmaKPL], total = 0;
#pragma unroll
    for (uint32_t j = 0; j < kMmaKPL; ++j) {
        const uint32_t e = lane * kMmaKPL + j;
        k[j] = e < cnt ? src[e] : kKeyInf;
        keep[j] = e < cnt && k[j] <= bound;
    }
#pragma unroll
    for (uint32_t j = 0; j < kMmaKPL; ++j) {  // kept keys numbered key slot by key slot
        const uint64_t b = __ballot(keep[j]);
        pos[j] = total + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(b >> 32),
                                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), 0u));
        total += static_cast<uint32_t>(__popcll(b));
    }
    if (total <= R) {
        uint64_t mx = 0;
#pragma unroll
        for (uint32_t j = 0; j < kMmaKPL; ++j)
            if (keep[j]) {
                dst[pos[j]] = k[j];
                mx = k[j] > mx ? k[j] : mx;
            }
        if (pad)
            for (uint32_t t = total + lane; t < R; t += 64) dst[t] = kKeyInf;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t o = ((uint64_t)(uint32_t)__shfl_xor((int)(mx >> 32), off) << 32) |
                               (uint32_t)__shfl_xor((int)(uint32_t)mx, off);
            mx = o > mx ? o : mx;
        }
        *kth = total == R ? mx : bound;
        return total;
    }
#pragma unroll
    for (uint32_t j = 0; j < kMmaKPL; ++j) k[j] = keep[j] ? k[j] : kKeyInf;
    wave_sort(k);
    uint64_t kr = kKeyInf;
#pragma unroll
    for (uint32_t j = 0; j < kMmaKPL; ++j) {
        const uint32_t e = lane * kMmaKPL + j;
        if (e < R) dst[e] = k[j];
        if (e == R - 1) kr = k[j];
    }
    const uint32_t sl = (R - 1) / kMmaKPL;
    *kth = ((uint64_t)(uint32_t)__shfl((int)(kr >> 32), sl) << 32) | (uint32_t)__shfl((int)(uint32_t)kr, sl);
    return R;
}

// The query's shared bound (all ones when there is none): an atomic load, so a bound another
// workgroup is lowering is never read torn.
__device__ __forceinline__ uint64_t mma_bound(const MmaArgs& a, uint32_t q) {
    return a.bound ? __hip_atomic_load(a.bound + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kKeyInf;
}

// Starting accumulator of a first-pass element: its constant over the GEMM scale (-inf for a
// dead row; a finite constant stays finite -- a huge one only needs to stay huge).
__device__ __forceinline__ float mma_init(float base, float inv) {
    return base < __builtin_inff() ? fminf(fmaxf(base * inv, -3.0e38f), 3.0e38f) : -__builtin_inff();
}

// Fused first pass: for each query of the workgroup's 256 and each row of its segment, the key of
// dist = max(0, ||q||^2 + bias[r] (+ qc[cl[r], q]) + qfac[q] * (-2 q.x)) enters the query's
// candidate list when it is below the query's current R-th key; a list that could not take
// another tile is pruned. At the end each (query, segment) list's keys go to part (padded with all
// ones). Pruning uses the smaller of the list's own R-th key and the query's shared bound, the
// smallest R-th key any segment of the query has published: a segment holding R keys at or below
// a key shows that key is not beaten by anything above it, so every key of the query's overall R
// smallest stays at or below the bound, and the segments' lists together hold exactly those R
// (under (distance, row) order) plus possibly others, which the merge discards. Grid: (segments, query tiles of 256); with the
// segment count a multiple of 8 every query tile of a segment runs on the same XCD (workgroups
// are dealt to the 8 XCDs round-robin), so its rows come from HBM once into that XCD's L2.
//
// The per-element constant goes into the accumulator before the K loop: with s = -2 qfac[q],
// acc starts at (||q||^2 + bias[r] + qc[cl[r], q]) / s, so dist = max(0, s * acc) at the end (a
// dead or out-of-segment row starts at -inf and ends at +inf). Its loads issue together while the
// tile's first stages load, so the epilogue is a multiply and a compare per element. s is never
// zero: qfac = 0 only when every low-precision row is zero, and then every product is zero and
// s = -1 gives the same distances.
template <uint32_t PREC>
__global__ __launch_bounds__(kMmaThreads) void k_mma_topr(MmaArgs a) {
    extern __shared__ __align__(16) uint8_t mma_lds[];
    __shared__ uint64_t kth[kMmaQ];
    __shared__ uint32_t cnt[kMmaQ];
    __shared__ float qconst[3][kMmaQ];  // per query: ||q||^2, 1 / s, s (kept in LDS, not registers,
                                        // so the K loop has its registers to itself)
    const uint32_t seg = blockIdx.x, q0 = blockIdx.y * kMmaQ;
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 5, l32 = lane & 31;
    const uint32_t wr = (w & 3) * 64, wq = (w >> 2) * 128;
    for (uint32_t t = threadIdx.x; t < kMmaQ; t += kMmaThreads) {
        kth[t] = kKeyInf;
        cnt[t] = 0;
        const uint32_t qs = min(q0 + t, a.n - 1);  // queries past the batch compute on the last one
        const float f = -2.0f * a.qfac[qs];
        const float sc = f < 0.0f ? f : -1.0f;
        qconst[0][t] = a.qn[qs];
        qconst[1][t] = 1.0f / sc;
        qconst[2][t] = sc;
    }
    __syncthreads();
    const uint32_t r_begin = seg * a.seg_rows, r_end = min(a.m, r_begin + a.seg_rows);
    for (uint32_t r0 = r_begin; r0 < r_end; r0 += kMmaRows) {
        mma_f32x16 acc[2][4];
        {
            float qq[4], inv[4];
            uint32_t qs[4];
#pragma unroll
            for (uint32_t qb = 0; qb < 4; ++qb) {
                const uint32_t ql = wq + qb * 32 + l32;
                qq[qb] = qconst[0][ql];
                inv[qb] = qconst[1][ql];
                qs[qb] = min(q0 + ql, a.n - 1);
            }
            // branch-free (indices clamped into range, out-of-segment rows masked after their
            // load), in chunks of 4 rows x 4 queries so the loads in flight stay few
            if (a.qc) {
#pragma unroll
                for (uint32_t rb = 0; rb < 2; ++rb)
#pragma unroll
                    for (uint32_t i4 = 0; i4 < 16; i4 += 4) {
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (uint32_t i = i4; i < i4 + 4; ++i) {
                            const uint32_t r = r0 + wr + rb * 32 + 8 * (i / 4) + 4 * g + (i % 4);
                            const uint32_t rc = min(r, r_end - 1);
                            const float bv = a.bias[rc];
                            const float* qcr = a.qc + (size_t)a.cl[rc] * a.n;
                            const float b = r < r_end ? bv : __builtin_inff();
#pragma unroll
                            for (uint32_t qb = 0; qb < 4; ++qb)
                                acc[rb][qb][i] = mma_init(qq[qb] + (b + qcr[qs[qb]]), inv[qb]);
                        }
                    }
            } else {
#pragma unroll
                for (uint32_t rb = 0; rb < 2; ++rb)
#pragma unroll
                    for (uint32_t i4 = 0; i4 < 16; i4 += 4) {
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (uint32_t i = i4; i < i4 + 4; ++i) {
                            const uint32_t r = r0 + wr + rb * 32 + 8 * (i / 4) + 4 * g + (i % 4);
                            const float bv = a.bias[min(r, r_end - 1)];
                            const float b = r < r_end ? bv : __builtin_inff();
#pragma unroll
                            for (uint32_t qb = 0; qb < 4; ++qb) acc[rb][qb][i] = mma_init(qq[qb] + b, inv[qb]);
                        }
                    }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        mma_prologue(a, mma_lds, r0, q0);
        mma_tile<PREC>(a, mma_lds, r0, q0, acc);
        // the epilogue's per-lane values come from a lane index the compiler cannot hoist above
        // the K loop, so none of its addresses occupies a register across that loop
        uint32_t le = threadIdx.x & 63;
        asm volatile("" : "+v"(le));
        const uint32_t ge = le >> 5, l32e = le & 31;
        uint64_t kq[4];
        float sc[4];
        uint32_t qid[4];
#pragma unroll
        for (uint32_t qb = 0; qb < 4; ++qb) {
            qid[qb] = q0 + wq + qb * 32 + l32e;
            const uint64_t kl = kth[wq + qb * 32 + l32e], gb = mma_bound(a, min(qid[qb], a.n - 1));
            kq[qb] = gb < kl ? gb : kl;
            sc[qb] = qconst[2][wq + qb * 32 + l32e];
        }
#pragma unroll
        for (uint32_t rb = 0; rb < 2; ++rb) {
#pragma unroll
            for (uint32_t i = 0; i < 16; ++i) {
#pragma unroll
                for (uint32_t qb = 0; qb < 4; ++qb) {
                    const float dist = fmaxf(sc[qb] * acc[rb][qb][i], 0.0f);
                    // keys order by distance first: a key below kq needs dist <= kq's distance
                    if (dist < __builtin_inff() && __float_as_uint(dist) <= static_cast<uint32_t>(kq[qb] >> 32) &&
                        qid[qb] < a.n) {
                        const uint32_t r = r0 + wr + rb * 32 + 8 * (i / 4) + 4 * ge + (i % 4);
                        const uint32_t row = a.rowlist ? a.rowlist[r] : static_cast<uint32_t>(a.row0 + r);
                        const uint64_t key = make_key(dist, row);
                        if (key < kq[qb]) {
                            const uint32_t slot = atomicAdd(&cnt[wq + qb * 32 + l32e], 1u);
                            a.lists[((size_t)qid[qb] * a.nseg + seg) * kMmaCap + slot] = key;
                        }
                    }
                }
            }
        }
        __threadfence_block();
        __syncthreads();
        // lists that could not take another tile keep their keys within the bound, at most the R
        // smallest (one wave per list); a list holding R publishes its R-th key as the query's bound
        for (uint32_t ql = w; ql < kMmaQ; ql += kMmaThreads / 64) {
            if (cnt[ql] + kMmaRows > kMmaCap && q0 + ql < a.n) {
                const uint32_t q = q0 + ql;
                uint64_t* L = a.lists + ((size_t)q * a.nseg + seg) * kMmaCap;
                const uint64_t kl = kth[ql], gb = mma_bound(a, q);
                uint64_t kr;
                const uint32_t kept = mma_prune(L, cnt[ql], a.R, gb < kl ? gb : kl, L, false, &kr);
                if (lane == 0) {
                    cnt[ql] = kept;
                    kth[ql] = kr;
                    if (a.bound && kept == a.R) atomicMin(reinterpret_cast<unsigned long long*>(a.bound + q), kr);
                }
            }
        }
        __threadfence_block();
        __syncthreads();
    }
    // every list: its keys within the bound, at most the R smallest, into part (padded)
    for (uint32_t ql = w; ql < kMmaQ; ql += kMmaThreads / 64) {
        const uint32_t q = q0 + ql;
        if (q >= a.n) continue;
        const uint64_t kl = kth[ql], gb = mma_bound(a, q);
        uint64_t kr;
        const uint32_t kept = mma_prune(a.lists + ((size_t)q * a.nseg + seg) * kMmaCap, cnt[ql], a.R, gb < kl ? gb : kl,
                                        a.part + ((size_t)q * a.nseg + seg) * a.R, true, &kr);
        if (a.bound && kept == a.R && lane == 0) atomicMin(reinterpret_cast<unsigned long long*>(a.bound + q), kr);
    }
}

}  // namespace pdxg
