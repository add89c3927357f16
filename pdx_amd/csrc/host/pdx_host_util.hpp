// Host-side numeric helpers for index construction: seeded rotation (FP64 Householder QR),
// multithreaded row-block products and a seeded Lloyd k-means.
//
// Construction stays on the host, as in the reference (index.hpp:329-394, 640-757); the search
// path never calls into this file.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

namespace PDX {
namespace host {

inline unsigned num_threads(unsigned requested = 0) {
    if (requested) return requested;
    unsigned n = std::thread::hardware_concurrency();
    return n ? std::min(n, 32u) : 4u;
}

// Run f(begin, end) over [0, n) split into contiguous chunks on worker threads.
inline void parallel_for(size_t n, unsigned threads, const std::function<void(size_t, size_t)>& f) {
    threads = std::max(1u, std::min<unsigned>(threads, (unsigned)((n + 63) / 64)));
    if (threads <= 1 || n < 128) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    size_t chunk = (n + threads - 1) / threads;
    for (unsigned t = 0; t < threads; ++t) {
        size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back(f, b, e);
    }
    for (auto& t : th) t.join();
}

// Random orthogonal d x d matrix: Gaussian entries from std::mt19937(seed) with
// std::normal_distribution<float> (as pruners/adsampling.hpp:44-54 draws them), orthogonalised by
// a Householder QR in double precision; returns Q row-major as float.
inline std::vector<float> random_rotation(uint32_t d, uint32_t seed) {
    std::mt19937 gen(seed);
    std::normal_distribution<float> nd;
    std::vector<double> A((size_t)d * d);  // column-major A(i,j) = A[j*d + i], filled row by row
    for (uint32_t i = 0; i < d; ++i)
        for (uint32_t j = 0; j < d; ++j) A[(size_t)j * d + i] = nd(gen);
    std::vector<double> vs((size_t)d * d, 0.0);  // Householder vectors, column k = v_k
    std::vector<double> tau(d, 0.0);
    for (uint32_t k = 0; k < d; ++k) {
        double* col = &A[(size_t)k * d];
        double norm = 0.0;
        for (uint32_t i = k; i < d; ++i) norm += col[i] * col[i];
        norm = std::sqrt(norm);
        double* v = &vs[(size_t)k * d];
        if (norm == 0.0) continue;
        double alpha = col[k] > 0 ? -norm : norm;
        for (uint32_t i = k; i < d; ++i) v[i] = col[i];
        v[k] -= alpha;
        double vn = 0.0;
        for (uint32_t i = k; i < d; ++i) vn += v[i] * v[i];
        if (vn == 0.0) continue;
        tau[k] = 2.0 / vn;
        // apply H_k = I - tau v v^T to the trailing columns
        for (uint32_t j = k; j < d; ++j) {
            double* cj = &A[(size_t)j * d];
            double s = 0.0;
            for (uint32_t i = k; i < d; ++i) s += v[i] * cj[i];
            s *= tau[k];
            for (uint32_t i = k; i < d; ++i) cj[i] -= s * v[i];
        }
    }
    // Q = H_0 H_1 ... H_{d-1} applied to I, built backwards.
    std::vector<double> Q((size_t)d * d, 0.0);  // column-major
    for (uint32_t i = 0; i < d; ++i) Q[(size_t)i * d + i] = 1.0;
    for (int k = (int)d - 1; k >= 0; --k) {
        const double* v = &vs[(size_t)k * d];
        if (tau[k] == 0.0) continue;
        for (uint32_t j = (uint32_t)k; j < d; ++j) {
            double* qj = &Q[(size_t)j * d];
            double s = 0.0;
            for (uint32_t i = k; i < d; ++i) s += v[i] * qj[i];
            s *= tau[k];
            for (uint32_t i = k; i < d; ++i) qj[i] -= s * v[i];
        }
    }
    std::vector<float> out((size_t)d * d);
    for (uint32_t i = 0; i < d; ++i)
        for (uint32_t j = 0; j < d; ++j) out[(size_t)i * d + j] = static_cast<float>(Q[(size_t)j * d + i]);
    return out;
}

// out (n x m) = X (n x d) . B (d x m), all row-major float; 4 x 16 register tiles.
inline void matmul(const float* X, size_t n, uint32_t d, const float* B, uint32_t m, float* out, unsigned threads) {
    const uint32_t mp = (m + 15) / 16 * 16;
    std::vector<float> Bp((size_t)d * mp, 0.0f);
    for (uint32_t j = 0; j < d; ++j) std::memcpy(&Bp[(size_t)j * mp], B + (size_t)j * m, m * sizeof(float));
    parallel_for((n + 3) / 4, threads, [&](size_t b, size_t e) {
        for (size_t rb = b; rb < e; ++rb) {
            const size_t r0 = rb * 4;
            const uint32_t nr = (uint32_t)std::min<size_t>(4, n - r0);
            for (uint32_t c0 = 0; c0 < mp; c0 += 16) {
                float acc[4][16] = {};
                for (uint32_t j = 0; j < d; ++j) {
                    const float* bp = &Bp[(size_t)j * mp + c0];
                    for (uint32_t r = 0; r < 4; ++r) {
                        const float xv = r < nr ? X[(r0 + r) * d + j] : 0.0f;
#pragma GCC unroll 16
                        for (uint32_t c = 0; c < 16; ++c) acc[r][c] += xv * bp[c];
                    }
                }
                for (uint32_t r = 0; r < nr; ++r)
                    for (uint32_t c = 0; c < 16 && c0 + c < m; ++c) out[(r0 + r) * m + c0 + c] = acc[r][c];
            }
        }
    });
}

// Nearest centroid (L2, ties -> lowest index) of every row of X.
inline void assign_nearest(const float* X, size_t n, uint32_t d, const float* C, uint32_t k,
                           std::vector<uint32_t>& out, unsigned threads) {
    out.assign(n, 0);
    std::vector<float> Ct((size_t)d * k);
    std::vector<float> cn(k, 0.0f);
    for (uint32_t c = 0; c < k; ++c)
        for (uint32_t j = 0; j < d; ++j) {
            float x = C[(size_t)c * d + j];
            Ct[(size_t)j * k + c] = x;
            cn[c] += x * x;
        }
    const size_t block = 256;
    std::vector<float> dots;
    parallel_for((n + block - 1) / block, threads, [&](size_t b, size_t e) {
        std::vector<float> dd(block * (size_t)k);
        for (size_t bb = b; bb < e; ++bb) {
            size_t r0 = bb * block, nr = std::min(block, n - r0);
            matmul(X + r0 * d, nr, d, Ct.data(), k, dd.data(), 1);
            for (size_t r = 0; r < nr; ++r) {
                const float* row = &dd[r * k];
                float best = cn[0] - 2.0f * row[0];
                uint32_t bi = 0;
                for (uint32_t c = 1; c < k; ++c) {
                    float v = cn[c] - 2.0f * row[c];
                    if (v < best) {
                        best = v;
                        bi = c;
                    }
                }
                out[r0 + r] = bi;
            }
        }
    });
}

inline void normalize_row(const float* x, float* out, uint32_t d) {
    double s = 0.0;
    for (uint32_t j = 0; j < d; ++j) s += (double)x[j] * x[j];
    if (s == 0.0) {
        std::memmove(out, x, d * sizeof(float));
        return;
    }
    float nrm = static_cast<float>(std::sqrt(s));
    for (uint32_t j = 0; j < d; ++j) out[j] = x[j] / nrm;
}

struct KMeansResult {
    std::vector<float> centroids;              // k x d
    std::vector<std::vector<uint64_t>> assignments;  // centroid -> member rows (ascending)
};

constexpr size_t MIN_EMBEDDINGS_TO_SAMPLE = 30720;  // clustering.hpp:22

// Seeded Lloyd k-means over a sampled training set (spherical when `angular`), then a full
// assignment. sampling_fraction 0 means 1.0 below MIN_EMBEDDINGS_TO_SAMPLE rows, else 0.3
// (clustering.hpp:58-63).
inline KMeansResult kmeans(const float* X, size_t n, uint32_t d, uint32_t k, uint32_t iters, uint32_t seed,
                           float sampling_fraction, bool angular, unsigned threads) {
    KMeansResult res;
    res.assignments.resize(k);
    if (k == 1) {  // clustering.hpp:47-53
        res.centroids.assign(X, X + d);
        for (uint64_t i = 0; i < n; ++i) res.assignments[0].push_back(i);
        return res;
    }
    std::mt19937_64 rng(seed);
    float frac = sampling_fraction > 0.0f ? sampling_fraction : (n < MIN_EMBEDDINGS_TO_SAMPLE ? 1.0f : 0.3f);
    size_t ns = std::max<size_t>(k, std::min<size_t>(n, (size_t)(n * (double)frac)));
    std::vector<uint64_t> perm(n);
    for (size_t i = 0; i < n; ++i) perm[i] = i;
    std::shuffle(perm.begin(), perm.end(), rng);
    std::vector<float> train(ns * d);
    for (size_t i = 0; i < ns; ++i) std::memcpy(&train[i * d], X + perm[i] * d, d * sizeof(float));
    std::vector<float> C((size_t)k * d);
    for (uint32_t c = 0; c < k; ++c) std::memcpy(&C[(size_t)c * d], &train[(size_t)c * d], d * sizeof(float));
    std::vector<uint32_t> a;
    for (uint32_t it = 0; it < iters; ++it) {
        assign_nearest(train.data(), ns, d, C.data(), k, a, threads);
        std::vector<double> sums((size_t)k * d, 0.0);
        std::vector<uint64_t> cnt(k, 0);
        for (size_t i = 0; i < ns; ++i) {
            cnt[a[i]]++;
            double* s = &sums[(size_t)a[i] * d];
            const float* x = &train[i * d];
            for (uint32_t j = 0; j < d; ++j) s[j] += x[j];
        }
        std::uniform_int_distribution<size_t> pick(0, ns - 1);
        for (uint32_t c = 0; c < k; ++c) {
            float* cc = &C[(size_t)c * d];
            if (cnt[c] == 0) {
                std::memcpy(cc, &train[pick(rng) * d], d * sizeof(float));
            } else {
                for (uint32_t j = 0; j < d; ++j) cc[j] = static_cast<float>(sums[(size_t)c * d + j] / (double)cnt[c]);
            }
            if (angular) normalize_row(cc, cc, d);
        }
    }
    assign_nearest(X, n, d, C.data(), k, a, threads);
    for (size_t i = 0; i < n; ++i) res.assignments[a[i]].push_back(i);
    res.centroids = std::move(C);
    return res;
}

}  // namespace host
}  // namespace PDX
