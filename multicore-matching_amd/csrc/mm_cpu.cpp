// Host-side matchers of the drop-in interface, for drivers that compare the
// GPU matchers with the reference's CPU ones (bin/match.cpp:136-158):
//   type 0  GraphMatchingCPURandom        src/matchcpu.cu:259-303
//   type 4  GraphMatchingCPUWeighted      src/matchcpu.cu:695-745
//   type 5  GraphMatchingCPUWeightedEdge  src/matchcpu.cu:747-810
//   10, 11  GraphMatchingTBB{Random,Weighted}, src/matchtbb.cu:498-623: the
//           hash-seeded rounds on host threads, maximal, one rand() per round
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cassert>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "graph.h"
#include "matchcpu.h"
#include "matchgpu.h"
#include "matchtbb.h"
#include "mm_hash.h"
#include "mm_internal.h"

using namespace std;

namespace mtc
{

void recordHostEvent(hipEvent_t &ev)
{
    RandStreamGuard guard;  // the HIP runtime must not move the caller's rand() stream
    int n = 0;
    if (ev && hipGetDeviceCount(&n) == hipSuccess && n > 0) {
        (void)hipEventRecord(ev, 0);
        (void)hipEventSynchronize(ev);
    }
}

namespace
{

float weightAt(const Graph &g, int j)
{
    return g.neighbourWeights.size() == g.neighbours.size() ? g.neighbourWeights[(size_t)j] : 1.0f;
}

}  // namespace

GraphMatchingCPURandom::GraphMatchingCPURandom(const Graph &g) : GraphMatching(g) {}
GraphMatchingCPURandom::~GraphMatchingCPURandom() {}

// Vertex order; each unmatched vertex takes its first unmatched neighbour.
void GraphMatchingCPURandom::performMatching(vector<int> &match, hipEvent_t &t1, hipEvent_t &t2) const
{
    assert((int)match.size() == graph.nrVertices);
    recordHostEvent(t1);
    for (int i = 0; i < graph.nrVertices; ++i) {
        if (isMatched(match[(size_t)i])) continue;
        const int2 r = graph.neighbourRanges[(size_t)i];
        for (int j = r.x; j < r.y; ++j) {
            const int u = graph.neighbours[(size_t)j];
            if (u != i && !isMatched(match[(size_t)u])) {
                match[(size_t)i] = match[(size_t)u] = matchVal(i, u);
                break;
            }
        }
    }
    recordHostEvent(t2);
}

GraphMatchingCPUWeighted::GraphMatchingCPUWeighted(const Graph &g) : GraphMatching(g) {}
GraphMatchingCPUWeighted::~GraphMatchingCPUWeighted() {}

// Vertex order; each unmatched vertex takes its heaviest unmatched neighbour
// (strictly heavier than -1 and than the best so far: the first on ties).
void GraphMatchingCPUWeighted::performMatching(vector<int> &match, hipEvent_t &t1, hipEvent_t &t2) const
{
    assert((int)match.size() == graph.nrVertices);
    recordHostEvent(t1);
    for (int i = 0; i < graph.nrVertices; ++i) {
        if (isMatched(match[(size_t)i])) continue;
        const int2 r = graph.neighbourRanges[(size_t)i];
        float best = -1.0f;
        int pick = -1;
        for (int j = r.x; j < r.y; ++j) {
            const int u = graph.neighbours[(size_t)j];
            if (u == i || isMatched(match[(size_t)u])) continue;
            const float w = weightAt(graph, j);
            if (w > best) {
                best = w;
                pick = u;
            }
        }
        if (pick >= 0) match[(size_t)i] = match[(size_t)pick] = matchVal(i, pick);
    }
    recordHostEvent(t2);
}

GraphMatchingCPUWeightedEdge::GraphMatchingCPUWeightedEdge(const Graph &g) : GraphMatching(g) {}
GraphMatchingCPUWeightedEdge::~GraphMatchingCPUWeightedEdge() {}

// Edges by weight, heaviest first (std::sort on a copy of Graph::edges, so
// ties fall as the library's unstable sort leaves them), matched greedily.
void GraphMatchingCPUWeightedEdge::performMatching(vector<int> &match, hipEvent_t &t1, hipEvent_t &t2) const
{
    assert((int)match.size() == graph.nrVertices);
    vector<Edge> order(graph.edges);
    if (order.empty() && !graph.neighbours.empty()) {
        // A CSR built without its edge list: each undirected edge once, rows in order.
        for (int i = 0; i < graph.nrVertices; ++i) {
            const int2 r = graph.neighbourRanges[(size_t)i];
            for (int j = r.x; j < r.y; ++j)
                if (i < graph.neighbours[(size_t)j]) order.push_back(Edge(i, graph.neighbours[(size_t)j], weightAt(graph, j)));
        }
    }
    recordHostEvent(t1);
    std::sort(order.begin(), order.end(), [](const Edge &a, const Edge &b) { return a.w > b.w; });
    for (const Edge &e : order) {
        if (isMatched(match[(size_t)e.x]) || isMatched(match[(size_t)e.y])) continue;
        match[(size_t)e.x] = match[(size_t)e.y] = matchVal(e.x, e.y);
    }
    recordHostEvent(t2);
}

void setHostThreads(int n) { mm::set_host_threads(n > 0 ? (unsigned)n : 0u); }

GraphMatchingTBB::GraphMatchingTBB(const Graph &g, const unsigned int &barrier) : GraphMatching(g), selectBarrier(barrier)
{
}
GraphMatchingTBB::~GraphMatchingTBB() {}

// The rounds of the reference's maximal matchers: select (hash colour), request
// (first red neighbour, or the heaviest with the first on ties), respond (the
// first / heaviest blue proposer) and match, each a parallel loop over the
// vertices that writes only its own vertex's slots; repeated while a select
// saw an active vertex, at most NR_MAX_MATCH_ROUNDS - 1 rounds.
void GraphMatchingTBB::rounds(vector<int> &match, bool weighted, hipEvent_t &t1, hipEvent_t &t2) const
{
    assert((int)match.size() == graph.nrVertices);
    const int n = graph.nrVertices;
    match.assign((size_t)n, 0);
    vector<int> req((size_t)n, 0);
    int *m = match.data(), *q = req.data();
    const int2 *rg = graph.neighbourRanges.data();
    const int *nb = graph.neighbours.data();
    const bool has_w = graph.neighbourWeights.size() == graph.neighbours.size();
    const float *w = has_w ? graph.neighbourWeights.data() : nullptr;
    const uint32_t barrier = selectBarrier;
    recordHostEvent(t1);
    int count = 0;
    bool keep = true;
    while (keep && ++count < NR_MAX_MATCH_ROUNDS) {
        const uint32_t seed = (uint32_t)rand();
        std::atomic<int> active(0);
        mm::parallel_for(n, [&](int64_t b, int64_t e, unsigned) {
            int any = 0;
            for (int64_t i = b; i < e; ++i)
                if (m[i] < 2) {
                    any = 1;
                    m[i] = mm_select_hash((uint32_t)i, seed) < barrier ? 0 : 1;
                }
            if (any) active.store(1, std::memory_order_relaxed);
        });
        keep = active.load() != 0;
        mm::parallel_for(n, [&](int64_t b, int64_t e, unsigned) {
            for (int64_t i = b; i < e; ++i) {
                if (m[i] != 0) {
                    q[i] = n;
                    continue;
                }
                bool dead = true;
                int pick = n;
                float best = -1.0f;
                for (int j = rg[i].x; j < rg[i].y; ++j) {
                    const int u = nb[j];
                    const int mu = m[u];
                    if (mu >= 4) continue;
                    dead = false;
                    if (mu != 1) continue;
                    if (!weighted) {
                        pick = u;
                        break;
                    }
                    const float wj = w ? w[j] : 1.0f;
                    if (wj > best) {
                        best = wj;
                        pick = u;
                    }
                }
                q[i] = pick < n ? pick : n + (dead ? 1 : 0);
            }
        });
        mm::parallel_for(n, [&](int64_t b, int64_t e, unsigned) {
            for (int64_t i = b; i < e; ++i) {
                if (m[i] != 1) continue;
                int pick = n;
                float best = -1.0f;
                for (int j = rg[i].x; j < rg[i].y; ++j) {
                    const int u = nb[j];
                    if (m[u] != 0 || q[u] != (int)i) continue;
                    if (!weighted) {
                        pick = u;
                        break;
                    }
                    const float wj = w ? w[j] : 1.0f;
                    if (wj > best) {
                        best = wj;
                        pick = u;
                    }
                }
                if (pick < n) q[i] = pick;
            }
        });
        mm::parallel_for(n, [&](int64_t b, int64_t e, unsigned) {
            for (int64_t i = b; i < e; ++i) {
                const int r = q[i];
                if (r == n + 1) m[i] = 2;
                else if (r < n && q[r] == (int)i) m[i] = matchVal((int)i, r);
            }
        });
    }
    recordHostEvent(t2);
}

GraphMatchingTBBRandom::GraphMatchingTBBRandom(const Graph &g, const unsigned int &b) : GraphMatchingTBB(g, b) {}
GraphMatchingTBBRandom::~GraphMatchingTBBRandom() {}
void GraphMatchingTBBRandom::performMatching(vector<int> &match, hipEvent_t &t1, hipEvent_t &t2) const
{
    rounds(match, false, t1, t2);
}

GraphMatchingTBBWeighted::GraphMatchingTBBWeighted(const Graph &g, const unsigned int &b) : GraphMatchingTBB(g, b) {}
GraphMatchingTBBWeighted::~GraphMatchingTBBWeighted() {}
void GraphMatchingTBBWeighted::performMatching(vector<int> &match, hipEvent_t &t1, hipEvent_t &t2) const
{
    rounds(match, true, t1, t2);
}

}  // namespace mtc
