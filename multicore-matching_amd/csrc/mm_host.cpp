// Host-side support of the matcher library: thread-local error state and a
// small static-partition thread pool used by the generators and CSR checks.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <atomic>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mm_match.h"
#include "mm_internal.h"

namespace {
thread_local std::string g_last_error;

std::atomic<unsigned> g_host_threads{0};

unsigned n_workers()
{
    const unsigned forced = g_host_threads.load();
    if (forced) return forced > 64 ? 64 : forced;
    const char *env = std::getenv("MM_GEN_THREADS");
    unsigned t = env ? (unsigned)std::atoi(env) : std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    return t > 64 ? 64 : t;
}
}  // namespace

namespace mm {

int32_t set_error(int32_t code, const char *msg)
{
    g_last_error = msg ? msg : "";
    return code;
}

int32_t set_errorf(int32_t code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

void set_host_threads(unsigned n) { g_host_threads.store(n); }

void parallel_for_dynamic(int64_t total, int64_t grain,
                          const std::function<void(int64_t, int64_t, unsigned)> &body)
{
    const unsigned nt = n_workers();
    if (total <= 0) return;
    if (grain < 1) grain = 1;
    if (nt == 1 || total <= grain) {
        body(0, total, 0);
        return;
    }
    std::atomic<int64_t> next(0);
    auto worker = [&](unsigned t) {
        for (;;) {
            const int64_t b = next.fetch_add(grain);
            if (b >= total) break;
            body(b, std::min(total, b + grain), t);
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t) pool.emplace_back(worker, t);
    for (auto &th : pool) th.join();
}

void parallel_for(int64_t total, const std::function<void(int64_t, int64_t, unsigned)> &body)
{
    const unsigned nt = n_workers();
    if (total <= 0) return;
    if (nt == 1 || total < 4096) {
        body(0, total, 0);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t chunk = (total + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const int64_t b = (int64_t)t * chunk, e = std::min(total, b + chunk);
        if (b >= e) break;
        pool.emplace_back(body, b, e, t);
    }
    for (auto &th : pool) th.join();
}

}  // namespace mm

extern "C" const char *mm_last_error(void) { return g_last_error.c_str(); }

// Result metrics of the reference's GraphMatching (src/matchcpu.cu:64-97,
// 127-246), restated for the drop-in boundary: weight and size of a match
// array, and its validity/maximality counts.
extern "C" int32_t mm_get_weight(int32_t n, const int32_t *ranges, const int32_t *neighbours, const float *weights,
                                 const int32_t *match, double *weight_out, int64_t *size_out)
{
    if (n < 0 || (n > 0 && (!ranges || !neighbours || !match)) || !weight_out || !size_out)
        return mm::set_error(MM_ERR_INVALID, "mm_get_weight: bad arguments");
    std::vector<double> wsum(64, 0.0);
    std::vector<int64_t> ssum(64, 0);
    mm::parallel_for(n, [&](int64_t b, int64_t e, unsigned t) {
        double w = 0.0;
        int64_t s = 0;
        for (int64_t i = b; i < e; ++i) {
            const int32_t m = match[i];
            if (m < 4) continue;
            for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1]; ++j) {
                if (match[neighbours[j]] == m) {
                    w += weights ? (double)weights[j] : 1.0;
                    s += 1;
                    break;
                }
            }
        }
        wsum[t] = w;
        ssum[t] = s;
    });
    double w = 0.0;
    int64_t s = 0;
    for (int t = 0; t < 64; ++t) {
        w += wsum[t];
        s += ssum[t];
    }
    *weight_out = w / 2.0;
    *size_out = s;
    return MM_OK;
}

extern "C" int32_t mm_test_matching(int32_t n, const int32_t *ranges, const int32_t *neighbours,
                                    const int32_t *match, int64_t *counts)
{
    if (n < 0 || (n > 0 && (!ranges || !neighbours || !match)) || !counts)
        return mm::set_error(MM_ERR_INVALID, "mm_test_matching: bad arguments");
    // testMatching sorts the vertices by value and walks the matched ones
    // from i1 = the next unvisited vertex, i2 = the one after it: a different
    // value in i2 makes i1 "matched to itself" and the walk resumes after i2;
    // otherwise (i1, i2) is checked for adjacency and the rest of the group
    // is "matched to more than one vertex".  Groups of equal values are thus
    // entered at their first or (after a single-member step) second member:
    // entry e' = [size == 1 + e] for the next group.  Groups come from slots
    // (value 4 + k, k < n, has slot k, so slot order is value order) and, for
    // larger values, which only invalid arrays hold, from sorted (value,
    // vertex) pairs; within a group members are in ascending vertex order.
    std::vector<int32_t> cnt((size_t)n, 0), mem((size_t)n * 3, -1);
    std::vector<std::pair<int32_t, int32_t>> outside;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t m = match[i];
        if (m < 4) continue;
        const int64_t k = (int64_t)m - 4;
        if (k >= n) {
            outside.emplace_back(m, i);
            continue;
        }
        const int32_t c = cnt[(size_t)k]++;
        if (c < 3) mem[(size_t)k * 3 + c] = i;
    }
    int64_t under = 0, multi = 0, invalid = 0, matchable = 0;
    auto adjacent = [&](int32_t a, int32_t b) {
        for (int32_t j = ranges[2 * a]; j < ranges[2 * a + 1]; ++j)
            if (neighbours[j] == b) return true;
        return false;
    };
    int entry = 0;
    // A group of c members whose first three (ascending) are at[0..2].
    auto walk = [&](int64_t c, const int32_t *at) {
        if (entry < c) {
            if (entry == c - 1) {
                ++under;
            } else {
                const int32_t a = at[entry], b = at[entry + 1];
                if (!adjacent(a, b) || !adjacent(b, a)) ++invalid;
                multi += c - entry - 2;
            }
        }
        entry = c == 1 + entry ? 1 : 0;
    };
    for (int32_t k = 0; k < n; ++k)
        if (cnt[(size_t)k]) walk(cnt[(size_t)k], mem.data() + (size_t)k * 3);
    std::sort(outside.begin(), outside.end());
    for (size_t a = 0; a < outside.size();) {
        size_t b = a;
        while (b < outside.size() && outside[b].first == outside[a].first) ++b;
        int32_t at[3] = {-1, -1, -1};
        for (size_t q = a; q < b && q < a + 3; ++q) at[q - a] = outside[q].second;
        walk((int64_t)(b - a), at);
        a = b;
    }
    for (int32_t i = 0; i < n; ++i) {
        if (match[i] >= 4) continue;
        for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1]; ++j) {
            const int32_t u = neighbours[j];
            if (u != i && match[u] < 4) { ++matchable; break; }
        }
    }
    counts[0] = under;
    counts[1] = multi;
    counts[2] = invalid;
    counts[3] = matchable;
    return (under == 0 && multi == 0 && invalid == 0) ? MM_OK : MM_ERR_INVALID;
}
