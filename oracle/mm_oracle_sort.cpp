/*
 * CPU oracle, part 2 (TEST INFRASTRUCTURE ONLY): the type-5 edge-sorted
 * greedy matcher, src/matchcpu.cu:747-810.  It sorts the undirected edge list
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
