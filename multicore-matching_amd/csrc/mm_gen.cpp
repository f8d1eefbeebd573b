// Seeded synthetic CSR generators: random geometric graphs, R-MAT and 2-D grids.
//
// All randomness comes from a counter-based mixer (SplitMix64's finaliser,
// S. Vigna), so every value depends only on (seed, index) and the output is
// identical for any thread count.  Rows come out strictly ascending by
// neighbour id and without self-loops; edge weights are mirrored exactly:
// w(u,v) = w(v,u) = (k + 1) / 2^24 with k = 24 bits of a hash of
// (seed, min(u,v), max(u,v)), i.e. uniform in (0, 1] and exact in float32.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <new>
#include <thread>
#include <vector>

#include "../../include/mm_match.h"
#include "mm_internal.h"

namespace {

inline uint64_t mix64(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

inline uint64_t hash2(uint64_t seed, uint64_t counter) { return mix64(mix64(seed) ^ counter); }

inline double unit_double(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

inline float edge_weight(uint64_t wseed, int32_t u, int32_t v)
{
    const uint32_t lo = (uint32_t)(u < v ? u : v), hi = (uint32_t)(u < v ? v : u);
    const uint64_t h = hash2(wseed, ((uint64_t)lo << 32) | hi);
    const uint32_t k = (uint32_t)(h >> 40);  // 24 bits
    return (float)(k + 1u) * (1.0f / 16777216.0f);
}

// Allocates the output arrays of a CSR with the given row lengths.
int32_t alloc_csr(int32_t n, const std::vector<int64_t> &offs, int32_t weighted, mm_csr *out)
{
    const int64_t nnz = offs[(size_t)n];
    out->n = n;
    out->nnz = nnz;
    out->ranges = (int32_t *)std::malloc(sizeof(int32_t) * 2 * (size_t)(n > 0 ? n : 1));
    out->neighbours = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    out->weights = weighted ? (float *)std::malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1)) : nullptr;
    if (!out->ranges || !out->neighbours || (weighted && !out->weights)) {
        mm_csr_free(out);
        return mm::set_error(MM_ERR_NOMEM, "generator: host allocation failed");
    }
    for (int32_t i = 0; i < n; ++i) {
        out->ranges[2 * i] = (int32_t)offs[(size_t)i];
        out->ranges[2 * i + 1] = (int32_t)offs[(size_t)i + 1];
    }
    return MM_OK;
}

void fill_weights(mm_csr *out, uint64_t seed)
{
    if (!out->weights) return;
    const uint64_t wseed = mix64(seed ^ 0x5745494748545321ull);
    mm::parallel_for(out->n, [&](int64_t b, int64_t e, unsigned) {
        for (int64_t i = b; i < e; ++i)
            for (int32_t j = out->ranges[2 * i]; j < out->ranges[2 * i + 1]; ++j)
                out->weights[j] = edge_weight(wseed, (int32_t)i, out->neighbours[j]);
    });
}

}  // namespace

extern "C" void mm_csr_free(mm_csr *csr)
{
    if (!csr) return;
    std::free(csr->ranges);
    std::free(csr->neighbours);
    std::free(csr->weights);
    csr->ranges = nullptr;
    csr->neighbours = nullptr;
    csr->weights = nullptr;
    csr->n = 0;
    csr->nnz = 0;
}

extern "C" int32_t mm_gen_rgg(int32_t n, double radius_scale, uint64_t seed, int32_t weighted,
                              mm_csr *out)
{
    if (!out || n < 0 || !(radius_scale > 0.0))
        return mm::set_error(MM_ERR_INVALID, "mm_gen_rgg: bad arguments");
    std::memset(out, 0, sizeof(*out));
    const double r = n > 1 ? radius_scale * std::sqrt(std::log((double)n) / (double)n) : 1.0;
    const double r2 = r * r;
    std::vector<double> px((size_t)n), py((size_t)n);
    const uint64_t pseed = mix64(seed ^ 0x504f494e54530000ull);
    mm::parallel_for(n, [&](int64_t b, int64_t e, unsigned) {
        for (int64_t i = b; i < e; ++i) {
            px[(size_t)i] = unit_double(hash2(pseed, 2 * (uint64_t)i));
            py[(size_t)i] = unit_double(hash2(pseed, 2 * (uint64_t)i + 1));
        }
    });
    // Square cells of side >= r: neighbours lie in the 3x3 block of cells.
    int64_t nc = r < 1.0 ? (int64_t)std::floor(1.0 / r) : 1;
    if (nc < 1) nc = 1;
    if (nc > 65536) nc = 65536;
    const int64_t ncells = nc * nc;
    auto cell_of = [&](int64_t i) {
        int64_t cx = (int64_t)(px[(size_t)i] * (double)nc), cy = (int64_t)(py[(size_t)i] * (double)nc);
        cx = std::min(cx, nc - 1);
        cy = std::min(cy, nc - 1);
        return cy * nc + cx;
    };
    std::vector<int64_t> cstart((size_t)ncells + 1, 0);
    std::vector<int32_t> cell((size_t)n), order((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        cell[(size_t)i] = (int32_t)cell_of(i);
        ++cstart[(size_t)cell[(size_t)i] + 1];
    }
    for (int64_t c = 0; c < ncells; ++c) cstart[(size_t)c + 1] += cstart[(size_t)c];
    {
        std::vector<int64_t> cur(cstart.begin(), cstart.end() - 1);
        for (int64_t i = 0; i < n; ++i) order[(size_t)cur[(size_t)cell[(size_t)i]]++] = (int32_t)i;
    }
    // Visits the neighbours of i (unsorted); returns their count.
    auto scan = [&](int64_t i, int32_t *dst) {
        const int64_t c = cell[(size_t)i], cx = c % nc, cy = c / nc;
        int64_t cnt = 0;
        for (int64_t dy = -1; dy <= 1; ++dy) {
            const int64_t yy = cy + dy;
            if (yy < 0 || yy >= nc) continue;
            for (int64_t dx = -1; dx <= 1; ++dx) {
                const int64_t xx = cx + dx;
                if (xx < 0 || xx >= nc) continue;
                const int64_t cc = yy * nc + xx;
                for (int64_t k = cstart[(size_t)cc]; k < cstart[(size_t)cc + 1]; ++k) {
                    const int32_t j = order[(size_t)k];
                    if (j == i) continue;
                    const double ddx = px[(size_t)i] - px[(size_t)j], ddy = py[(size_t)i] - py[(size_t)j];
                    if (ddx * ddx + ddy * ddy < r2) {
                        if (dst) dst[cnt] = j;
                        ++cnt;
                    }
                }
            }
        }
        return cnt;
    };
    std::vector<int64_t> offs((size_t)n + 1, 0);
    mm::parallel_for(n, [&](int64_t b, int64_t e, unsigned) {
        for (int64_t i = b; i < e; ++i) offs[(size_t)i + 1] = scan(i, nullptr);
    });
    for (int64_t i = 0; i < n; ++i) offs[(size_t)i + 1] += offs[(size_t)i];
    if (offs[(size_t)n] > INT32_MAX) return mm::set_error(MM_ERR_INVALID, "mm_gen_rgg: nnz exceeds int32");
    int32_t rc = alloc_csr(n, offs, weighted, out);
    if (rc != MM_OK) return rc;
    mm::parallel_for(n, [&](int64_t b, int64_t e, unsigned) {
        for (int64_t i = b; i < e; ++i) {
            int32_t *row = out->neighbours + offs[(size_t)i];
            const int64_t cnt = scan(i, row);
            std::sort(row, row + cnt);
        }
    });
    fill_weights(out, seed);
    return MM_OK;
}

extern "C" int32_t mm_gen_grid(int32_t rows, int32_t cols, uint64_t seed, int32_t weighted, mm_csr *out)
{
    if (!out || rows < 0 || cols < 0 || (int64_t)rows * cols > INT32_MAX)
        return mm::set_error(MM_ERR_INVALID, "mm_gen_grid: bad arguments");
    std::memset(out, 0, sizeof(*out));
    const int32_t n = rows * cols;
    std::vector<int64_t> offs((size_t)n + 1, 0);
    for (int32_t v = 0; v < n; ++v) {
        const int32_t r = v / cols, c = v % cols;
        offs[(size_t)v + 1] = offs[(size_t)v] + (r > 0) + (c > 0) + (c + 1 < cols) + (r + 1 < rows);
    }
    int32_t rc = alloc_csr(n, offs, weighted, out);
    if (rc != MM_OK) return rc;
    for (int32_t v = 0; v < n; ++v) {
        const int32_t r = v / cols, c = v % cols;
        int32_t *p = out->neighbours + offs[(size_t)v];
        if (r > 0) *p++ = v - cols;
        if (c > 0) *p++ = v - 1;
        if (c + 1 < cols) *p++ = v + 1;
        if (r + 1 < rows) *p++ = v + cols;
    }
    fill_weights(out, seed);
    return MM_OK;
}

extern "C" int32_t mm_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                               uint64_t seed, int32_t weighted, mm_csr *out)
{
    if (!out || scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return mm::set_error(MM_ERR_INVALID, "mm_gen_rmat: bad arguments");
    std::memset(out, 0, sizeof(*out));
    const int32_t n = (int32_t)(1u << scale);
    const int64_t samples = (int64_t)edge_factor << scale;
    const double ab = a + b, abc = a + b + c;
    const uint64_t eseed = mix64(seed ^ 0x524d415445444745ull);
    std::vector<int32_t> src((size_t)samples), dst((size_t)samples);
    mm::parallel_for(samples, [&](int64_t lo, int64_t hi, unsigned) {
        for (int64_t e = lo; e < hi; ++e) {
            uint32_t u = 0, v = 0;
            for (int32_t l = 0; l < scale; ++l) {
                const double p = unit_double(hash2(eseed, ((uint64_t)e << 5) | (uint64_t)l));
                const uint32_t bu = p >= ab, bv = (p >= a && p < ab) || p >= abc;
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
            src[(size_t)e] = (int32_t)u;
            dst[(size_t)e] = (int32_t)v;
        }
    });
    // Count both directions (self-loops dropped), scatter, then sort and
    // deduplicate every row; the result does not depend on scatter order.
    std::vector<std::atomic<int64_t>> deg((size_t)n);
    for (auto &d : deg) d.store(0, std::memory_order_relaxed);
    mm::parallel_for(samples, [&](int64_t lo, int64_t hi, unsigned) {
        for (int64_t e = lo; e < hi; ++e) {
            const int32_t u = src[(size_t)e], v = dst[(size_t)e];
            if (u == v) continue;
            deg[(size_t)u].fetch_add(1, std::memory_order_relaxed);
            deg[(size_t)v].fetch_add(1, std::memory_order_relaxed);
        }
    });
    std::vector<int64_t> raw((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; ++i) raw[(size_t)i + 1] = raw[(size_t)i] + deg[(size_t)i].load();
    std::vector<int32_t> adj((size_t)raw[(size_t)n]);
    std::vector<std::atomic<int64_t>> cur((size_t)n);
    for (int32_t i = 0; i < n; ++i) cur[(size_t)i].store(raw[(size_t)i], std::memory_order_relaxed);
    mm::parallel_for(samples, [&](int64_t lo, int64_t hi, unsigned) {
        for (int64_t e = lo; e < hi; ++e) {
            const int32_t u = src[(size_t)e], v = dst[(size_t)e];
            if (u == v) continue;
            adj[(size_t)cur[(size_t)u].fetch_add(1, std::memory_order_relaxed)] = v;
            adj[(size_t)cur[(size_t)v].fetch_add(1, std::memory_order_relaxed)] = u;
        }
    });
    src.clear(); src.shrink_to_fit();
    dst.clear(); dst.shrink_to_fit();
    std::vector<int64_t> offs((size_t)n + 1, 0);
    mm::parallel_for(n, [&](int64_t lo, int64_t hi, unsigned) {
        for (int64_t i = lo; i < hi; ++i) {
            int32_t *row = adj.data() + raw[(size_t)i];
            const int64_t len = raw[(size_t)i + 1] - raw[(size_t)i];
            std::sort(row, row + len);
            offs[(size_t)i + 1] = std::unique(row, row + len) - row;
        }
    });
    for (int32_t i = 0; i < n; ++i) offs[(size_t)i + 1] += offs[(size_t)i];
    if (offs[(size_t)n] > INT32_MAX) return mm::set_error(MM_ERR_INVALID, "mm_gen_rmat: nnz exceeds int32");
    int32_t rc = alloc_csr(n, offs, weighted, out);
    if (rc != MM_OK) return rc;
    mm::parallel_for(n, [&](int64_t lo, int64_t hi, unsigned) {
        for (int64_t i = lo; i < hi; ++i)
            std::memcpy(out->neighbours + offs[(size_t)i], adj.data() + raw[(size_t)i],
                        sizeof(int32_t) * (size_t)(offs[(size_t)i + 1] - offs[(size_t)i]));
    });
    fill_weights(out, seed);
    return MM_OK;
}
