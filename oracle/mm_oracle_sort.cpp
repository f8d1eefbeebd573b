/*
 * CPU oracle, part 2 (TEST INFRASTRUCTURE ONLY): the matchers whose results
 * depend on libstdc++'s std::sort.  Type 5, the edge-sorted greedy matcher,
 * src/matchcpu.cu:747-810, and type 1, static minimum degree, 428-516 (pinned
 * by closed forms and validity/maximality in tests/test_host_matchers.py).
 * Type 5:  It sorts the undirected edge list
 * by weight, descending, with the same libstdc++ std::sort the reference
 * calls (the sort is not stable, so tie order is the library's), then matches
 * greedily.  Edges are (x, y, w) triples; the caller builds them in the order
 * the reference's Graph::edges would hold.
 */
#include <algorithm>
#include <cstdint>
#include <vector>

namespace {
struct OrcEdge {
    int32_t x, y;
    float w;
};
struct ByWeightDesc {
    bool operator()(const OrcEdge &a, const OrcEdge &b) const { return a.w > b.w; }
};
}  // namespace

extern "C" void orc_cpu_weighted_edge(int32_t n, int64_t n_edges, const int32_t *ex,
                                      const int32_t *ey, const float *ew, int32_t *match)
{
    std::vector<OrcEdge> order((size_t)n_edges);
    for (int64_t k = 0; k < n_edges; ++k) order[(size_t)k] = OrcEdge{ex[k], ey[k], ew[k]};
    std::sort(order.begin(), order.end(), ByWeightDesc());
    for (int32_t i = 0; i < n; ++i) match[i] = 0;
    for (const OrcEdge &e : order) {
        if (match[e.x] < 4 && match[e.y] < 4) {
            const int32_t m = 4 + (e.x < e.y ? e.x : e.y);
            match[e.x] = m;
            match[e.y] = m;
        }
    }
}

/* Type 1, one-sided static minimum degree (src/matchcpu.cu:428-516): the
 * identity order sorted by row length with std::sort (not stable: ties fall
 * where libstdc++ leaves them, as in the reference), then each unmatched
 * vertex in that order takes its first unmatched neighbour. */
extern "C" void orc_cpu_stat_min_deg(int32_t n, const int32_t *ranges, const int32_t *nbrs, int32_t *match)
{
    std::vector<int64_t> len((size_t)n);
    std::vector<int32_t> order((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        len[(size_t)i] = (int64_t)ranges[2 * i + 1] - ranges[2 * i];
        order[(size_t)i] = i;
        match[i] = 0;
    }
    std::sort(order.begin(), order.end(), [&len](int32_t a, int32_t b) { return len[(size_t)a] < len[(size_t)b]; });
    for (const int32_t i : order) {
        if (match[i] >= 4) continue;
        for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1]; ++j) {
            const int32_t u = nbrs[j];
            if (u != i && match[u] < 4) {
                match[i] = match[u] = 4 + (i < u ? i : u);
                break;
            }
        }
    }
}

