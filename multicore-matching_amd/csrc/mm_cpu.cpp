// Host-side matchers of the drop-in interface, for drivers that compare the
// GPU matchers with the reference's CPU ones (bin/match.cpp:136-158):
//   type 0  GraphMatchingCPURandom        src/matchcpu.cu:259-303
//   type 1  GraphMatchingCPUStatMinDeg    src/matchcpu.cu:428-516
//   type 2  GraphMatchingCPUKarpSipser    src/matchcpu.cu:518-682
//   type 3  GraphMatchingCPUMinDeg        src/matchcpu.cu:304-426
//   type 4  GraphMatchingCPUWeighted      src/matchcpu.cu:695-745
//   type 5  GraphMatchingCPUWeightedEdge  src/matchcpu.cu:747-810
//   10, 11  GraphMatchingTBB{Random,Weighted}, src/matchtbb.cu:498-623: the
//           hash-seeded rounds on host threads, maximal, one rand() per round
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <numeric>
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
#include "../../include/mm_match.h"

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

namespace
{

// First unmatched neighbour of i other than i itself, in row order; -1 if none.
int firstFree(const Graph &g, const vector<int> &match, int i)
{
    const int2 r = g.neighbourRanges[(size_t)i];
    for (int j = r.x; j < r.y; ++j) {
        const int u = g.neighbours[(size_t)j];
        if (u != i && !GraphMatching::isMatched(match[(size_t)u])) return u;
    }
    return -1;
}

int rowLength(const Graph &g, int i)
{
    const int2 r = g.neighbourRanges[(size_t)i];
    return r.y - r.x;
}

// Vertices keyed by a small integer, smallest key first and, among equal
// keys, in order of (re)insertion: the order a std::multimap<int, int> gives
// (equal keys are inserted at the end of their range).  One FIFO list per
// key, linked through per-vertex arrays, so insert and erase are O(1).
class KeyQueue
{
    public:
        KeyQueue(int n, int maxKey) : head(maxKey + 1, -1), tail(maxKey + 1, -1), prev(n, -1), next(n, -1), key(n, -1) {}

        bool empty() const { return count == 0; }
        int keyOf(int v) const { return key[(size_t)v]; }
        bool contains(int v) const { return key[(size_t)v] >= 0; }

        void insert(int v, int k)
        {
            key[(size_t)v] = k;
            prev[(size_t)v] = tail[(size_t)k];
            next[(size_t)v] = -1;
            if (tail[(size_t)k] >= 0) next[(size_t)tail[(size_t)k]] = v;
            else head[(size_t)k] = v;
            tail[(size_t)k] = v;
            if (k < lo) lo = k;
            ++count;
        }

        void erase(int v)
        {
            const int k = key[(size_t)v];
            if (k < 0) return;
            if (prev[(size_t)v] >= 0) next[(size_t)prev[(size_t)v]] = next[(size_t)v];
            else head[(size_t)k] = next[(size_t)v];
            if (next[(size_t)v] >= 0) prev[(size_t)next[(size_t)v]] = prev[(size_t)v];
            else tail[(size_t)k] = prev[(size_t)v];
            key[(size_t)v] = -1;
            --count;
        }

        // The first vertex of the smallest key (the queue must not be empty).
        int front()
        {
            while (head[(size_t)lo] < 0) ++lo;
            return head[(size_t)lo];
        }

    private:
        vector<int> head, tail, prev, next, key;
        int lo = 0;
        long count = 0;
};

}  // namespace

GraphMatchingCPUMinDeg::GraphMatchingCPUMinDeg(const Graph &g) : GraphMatching(g) {}
GraphMatchingCPUMinDeg::~GraphMatchingCPUMinDeg() {}

// Two-sided dynamic minimum degree: repeatedly take the vertex of least
// current degree (ties: the one whose degree changed least recently, then
// vertex order), match it to its unmatched neighbour of least current degree
// (first in row order on ties), and lower the degrees of both endpoints'
// unmatched neighbours, once per row entry; a vertex whose degree reaches 0
// leaves the queue unmatched.  Degrees start at the row lengths.
void GraphMatchingCPUMinDeg::performMatching(vector<int> &match, hipEvent_t &t1, hipEvent_t &t2) const
{
    assert((int)match.size() == graph.nrVertices);
    const int n = graph.nrVertices;
    int maxDeg = 0;
    for (int i = 0; i < n; ++i) maxDeg = std::max(maxDeg, rowLength(graph, i));
    recordHostEvent(t1);
    KeyQueue q(n, maxDeg);
    for (int i = 0; i < n; ++i)
        if (rowLength(graph, i) > 0) q.insert(i, rowLength(graph, i));
    auto lower = [&](int centre, int skip) {
        const int2 r = graph.neighbourRanges[(size_t)centre];
        for (int j = r.x; j < r.y; ++j) {
            const int u = graph.neighbours[(size_t)j];
            if (u == skip || isMatched(match[(size_t)u]) || !q.contains(u)) continue;
            const int d = q.keyOf(u);
            q.erase(u);
            if (d > 1) q.insert(u, d - 1);
        }
    };
    while (!q.empty()) {
        const int i = q.front();
        const int2 r = graph.neighbourRanges[(size_t)i];
        int best = n + 1, pick = -1;
        for (int j = r.x; j < r.y; ++j) {
            const int u = graph.neighbours[(size_t)j];
            if (u == i || isMatched(match[(size_t)u]) || !q.contains(u)) continue;
            if (q.keyOf(u) < best) {
                best = q.keyOf(u);
                pick = u;
            }
        }
        q.erase(i);
        if (pick < 0) continue;  // only self-loops or stale entries left: i stays unmatched
        match[(size_t)i] = match[(size_t)pick] = matchVal(i, pick);
        q.erase(pick);
        lower(i, i);
        lower(pick, i);
    }
    recordHostEvent(t2);
}

GraphMatchingCPUStatMinDeg::GraphMatchingCPUStatMinDeg(const Graph &g) : GraphMatching(g) {}
GraphMatchingCPUStatMinDeg::~GraphMatchingCPUStatMinDeg() {}

// One-sided static minimum degree: vertices in order of row length (std::sort
// of the identity order, so ties fall where libstdc++'s unstable sort leaves
// them, as in the reference), each unmatched one taking its first unmatched
// neighbour.
void GraphMatchingCPUStatMinDeg::performMatching(vector<int> &match, hipEvent_t &t1, hipEvent_t &t2) const
{
    assert((int)match.size() == graph.nrVertices);
    const int n = graph.nrVertices;
    vector<int> order((size_t)n);
    recordHostEvent(t1);
    std::iota(order.begin(), order.end(), 0);
    const int2 *rg = graph.neighbourRanges.data();
    std::sort(order.begin(), order.end(), [rg](int a, int b) {
        return (int64_t)rg[a].y - rg[a].x < (int64_t)rg[b].y - rg[b].x;
    });
    for (const int i : order) {
        if (isMatched(match[(size_t)i])) continue;
        const int u = firstFree(graph, match, i);
        if (u >= 0) match[(size_t)i] = match[(size_t)u] = matchVal(i, u);
    }
    recordHostEvent(t2);
}

GraphMatchingCPUKarpSipser::GraphMatchingCPUKarpSipser(const Graph &g) : GraphMatching(g) {}
GraphMatchingCPUKarpSipser::~GraphMatchingCPUKarpSipser() {}

// One-sided Karp--Sipser.  deg starts at the row lengths.  Vertices of degree
// 1 wait in a FIFO of singletons, the others (degree > 1) in a list in vertex
// order.  Singletons are matched to their first unmatched neighbour, after
// which every row entry of that partner lowers its target's degree.  When no
// singleton is left, ONE pass over the list matches each unmatched vertex of
// nonzero degree to its first unmatched neighbour and lowers the degrees of
// both endpoints' row targets; every vertex the pass reaches leaves the list.
// In both phases an unmatched vertex whose degree drops to exactly 1 moves
// from the list to the singletons (served after the pass).
void GraphMatchingCPUKarpSipser::performMatching(vector<int> &match, hipEvent_t &t1, hipEvent_t &t2) const
{
    assert((int)match.size() == graph.nrVertices);
    const int n = graph.nrVertices;
    vector<int> deg((size_t)n), nxt((size_t)n, -1), prv((size_t)n, -1), fifo;
    vector<char> listed((size_t)n, 0);
    recordHostEvent(t1);
    fifo.reserve((size_t)n);
    size_t fifoHead = 0;
    int first = -1, last = -1;
    for (int i = 0; i < n; ++i) {
        deg[(size_t)i] = rowLength(graph, i);
        if (deg[(size_t)i] == 1) fifo.push_back(i);
        else if (deg[(size_t)i] > 1) {
            listed[(size_t)i] = 1;
            prv[(size_t)i] = last;
            if (last >= 0) nxt[(size_t)last] = i;
            else first = i;
            last = i;
        }
    }
    auto unlist = [&](int v) {
        if (!listed[(size_t)v]) return;
        listed[(size_t)v] = 0;
        if (prv[(size_t)v] >= 0) nxt[(size_t)prv[(size_t)v]] = nxt[(size_t)v];
        else first = nxt[(size_t)v];
        if (nxt[(size_t)v] >= 0) prv[(size_t)nxt[(size_t)v]] = prv[(size_t)v];
        else last = prv[(size_t)v];
    };
    auto lower = [&](int centre) {
        const int2 r = graph.neighbourRanges[(size_t)centre];
        for (int j = r.x; j < r.y; ++j) {
            const int u = graph.neighbours[(size_t)j];
            if (--deg[(size_t)u] == 1 && !isMatched(match[(size_t)u])) {
                unlist(u);
                fifo.push_back(u);
            }
        }
    };
    while (fifoHead < fifo.size() || first >= 0) {
        while (fifoHead < fifo.size()) {
            const int i = fifo[fifoHead++];
            if (isMatched(match[(size_t)i]) || deg[(size_t)i] < 1) continue;
            const int u = firstFree(graph, match, i);
            if (u < 0) continue;  // only self-loops left: i stays unmatched
            match[(size_t)i] = match[(size_t)u] = matchVal(i, u);
            lower(u);
        }
        for (int h = first; h >= 0;) {
            const int i = h;
            if (!isMatched(match[(size_t)i]) && deg[(size_t)i] != 0) {
                const int u = firstFree(graph, match, i);
                if (u >= 0) {
                    match[(size_t)i] = match[(size_t)u] = matchVal(i, u);
                    lower(i);
                    lower(u);
                }
            }
            h = nxt[(size_t)i];  // read after the lowering above may have unlisted later vertices
            unlist(i);
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

// Sequential host matchers behind the C ABI (types 0-5 of the reference's
// getMatcher, bin/match.cpp:136-143): the CSR is copied into a Graph and the
// class of that type runs on it.  Type 5 takes its edges from the rows (i < j,
// row order), as the drop-in does for a Graph without an edge list.
extern "C" int32_t mm_host_match(int32_t type, int32_t n, int64_t nnz, const int32_t *ranges, const int32_t *neighbours,
                                 const float *weights, int32_t *match_out)
{
    using namespace mtc;
    if (n < 0 || nnz < 0 || (n > 0 && (!ranges || !match_out)) || (nnz > 0 && !neighbours))
        return mm::set_error(MM_ERR_INVALID, "mm_host_match: bad arguments");
    if (type < 0 || type > 5) return mm::set_error(MM_ERR_INVALID, "mm_host_match: type must be 0..5");
    Graph g;
    g.nrVertices = n;
    g.nrEdges = (int)(nnz / 2);
    g.neighbourRanges.resize((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        const int32_t b = ranges[2 * (size_t)i], e = ranges[2 * (size_t)i + 1];
        if (b < 0 || e < b || e > nnz) return mm::set_error(MM_ERR_INVALID, "mm_host_match: bad row range");
        g.neighbourRanges[(size_t)i] = make_int2(b, e);
    }
    g.neighbours.assign(neighbours, neighbours + nnz);
    for (const int u : g.neighbours)
        if (u < 0 || u >= n) return mm::set_error(MM_ERR_INVALID, "mm_host_match: neighbour id out of range");
    if (weights) g.neighbourWeights.assign(weights, weights + nnz);
    GraphMatching *m = nullptr;
    switch (type) {
    case 0: m = new GraphMatchingCPURandom(g); break;
    case 1: m = new GraphMatchingCPUStatMinDeg(g); break;
    case 2: m = new GraphMatchingCPUKarpSipser(g); break;
    case 3: m = new GraphMatchingCPUMinDeg(g); break;
    case 4: m = new GraphMatchingCPUWeighted(g); break;
    default: m = new GraphMatchingCPUWeightedEdge(g); break;
    }
    vector<int> match = m->initialMatching();
    hipEvent_t t1 = nullptr, t2 = nullptr;
    m->performMatching(match, t1, t2);
    delete m;
    std::copy(match.begin(), match.end(), match_out);
    return MM_OK;
}

