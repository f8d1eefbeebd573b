// C++ drop-in classes over the C ABI: mtc::Graph helpers, the GraphMatching
// base (src/matchcpu.cu:32-246) and the GraphMatchingGPU family
// (src/matchgpu.cu:39-851), so a match.cpp-style driver keeps its calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cassert>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <iostream>
#include <vector>

#include "../../include/mm_match.h"
#include "graph.h"
#include "matchcpu.h"
#include "matchgpu.h"

using namespace std;

namespace mtc
{

Edge::Edge() : x(0), y(0), w(0.0f) {}
Edge::Edge(const int &_x, const int &_y, const float &_w) : x(_x), y(_y), w(_w) {}
Edge::~Edge() {}

Graph::Graph() : nrVertices(0), nrVertexWeights(0), nrEdges(0) {}
Graph::~Graph() {}

bool Graph::empty() const
{
    return nrVertices == 0 || nrEdges == 0 || neighbours.empty() || neighbourRanges.empty() || edges.empty();
}

void Graph::clear()
{
    nrVertices = 0;
    nrVertexWeights = 0;
    nrEdges = 0;
    neighbourRanges.clear();
    vertexWeights.clear();
    neighbours.clear();
    neighbourWeights.clear();
    edges.clear();
}

// Relabels the vertices with a rand()-driven permutation, as
// src/graph.cu:513-548: rows move with their vertex, ids inside rows are
// renamed in place (so rows keep their order and are no longer sorted).
vector<int> Graph::random_shuffle()
{
    vector<int> permutation(nrVertices);
    for (int i = 0; i < nrVertices; ++i) permutation[i] = i;
    std::random_shuffle(permutation.begin(), permutation.end());
    vector<int2> ranges(nrVertices);
    for (int i = 0; i < nrVertices; ++i) ranges[i] = neighbourRanges[permutation[i]];
    neighbourRanges.swap(ranges);
    vector<int> inverse(nrVertices);
    for (int i = 0; i < nrVertices; ++i) inverse[permutation[i]] = i;
    for (int &v : neighbours) v = inverse[v];
    for (Edge &e : edges) {
        e.x = inverse[e.x];
        e.y = inverse[e.y];
    }
    return permutation;
}

GraphMatching::GraphMatching(const Graph &_graph) : graph(_graph) {}
GraphMatching::~GraphMatching() {}

vector<int> GraphMatching::initialMatching() const { return vector<int>(graph.nrVertices, 0); }

static const int32_t *ranges_of(const Graph &g)
{
    return g.neighbourRanges.empty() ? nullptr : reinterpret_cast<const int32_t *>(g.neighbourRanges.data());
}

void GraphMatching::getWeight(double &weight, long &size, const vector<int> &match, const Graph &g)
{
    assert((int)match.size() == g.nrVertices);
    const bool has_w = g.neighbourWeights.size() == g.neighbours.size() && !g.neighbours.empty();
    int64_t s = 0;
    if (mm_get_weight(g.nrVertices, ranges_of(g), g.neighbours.data(), has_w ? g.neighbourWeights.data() : nullptr,
                      match.data(), &weight, &s) != MM_OK) {
        cerr << "Unable to weigh matching: " << mm_last_error() << endl;
        throw exception();
    }
    size = (long)s;
}

bool GraphMatching::testMatching(const vector<int> &match, const Graph &g)
{
    if ((int)match.size() != g.nrVertices) {
        cerr << "Incompatible matching size!" << endl;
        return false;
    }
    int64_t c[4] = {0, 0, 0, 0};
    const int32_t rc = mm_test_matching(g.nrVertices, ranges_of(g), g.neighbours.data(), match.data(), c);
    cout << "Verified a matching of " << match.size() << " vertices, of these" << endl;
    cout << "    " << c[0] << " were matched to themselves," << endl;
    cout << "    " << c[1] << " were matched to more than one vertex," << endl;
    cout << "    " << c[2] << " were not matched to a neighbour," << endl;
    cout << "    " << c[3] << " can still be matched." << endl;
    return rc == MM_OK;
}

vector<unsigned int> peekRand(int count)
{
    // glibc keeps rand()'s state in the buffer last given to initstate/setstate;
    // switching away stores the read position into that buffer.  Copy it, draw
    // from the copy, then switch back: the process stream does not move.
    static char scratch[256];
    static char copy[256];
    char *prev = initstate(1u, scratch, sizeof(scratch));
    const int32_t hdr = *reinterpret_cast<int32_t *>(prev);
    static const int degrees[5] = {0, 7, 15, 31, 63};
    const int type = hdr % 5;
    const size_t bytes = sizeof(int32_t) * (size_t)(degrees[type < 0 ? 0 : type] + 1);
    std::memcpy(copy, prev, bytes);
    setstate(copy);
    vector<unsigned int> out((size_t)count);
    for (int i = 0; i < count; ++i) out[(size_t)i] = (unsigned int)random();
    setstate(prev);
    return out;
}

// Copies the state on entry and puts it back on exit, so whatever the runtime
// draws in between, the caller's stream continues exactly where it was.
namespace
{
size_t stateBytes(const char *state)
{
    static const int degrees[5] = {0, 7, 15, 31, 63};
    const int32_t hdr = *reinterpret_cast<const int32_t *>(state);
    const int type = hdr % 5;
    return sizeof(int32_t) * (size_t)(degrees[type < 0 ? 0 : type] + 1);
}
}  // namespace

RandStreamGuard::RandStreamGuard()
{
    static char scratch[256];
    char *prev = initstate(1u, scratch, sizeof(scratch));
    bytes = stateBytes(prev);
    std::memcpy(saved, prev, bytes);
    setstate(prev);
}

RandStreamGuard::~RandStreamGuard()
{
    // Restore into the spare of two buffers: setstate() writes the read
    // position of the buffer it leaves, which must not be the copy.
    static char live[2][256];
    static int which = 0;
    which ^= 1;
    std::memcpy(live[which], saved, bytes);
    setstate(live[which]);
}

GraphMatchingGPU::GraphMatchingGPU(const Graph &_graph, const int &_threadsPerBlock, const unsigned int &_selectBarrier)
    : GraphMatching(_graph), threadsPerBlock(_threadsPerBlock), selectBarrier(_selectBarrier), dgraph(nullptr),
      rounds(0)
{
    RandStreamGuard guard;
    int device = 0;
    hipGetDevice(&device);
    const bool weighted = graph.neighbourWeights.size() == graph.neighbours.size() && !graph.neighbours.empty();
    if (mm_graph_create(device, graph.nrVertices, (int64_t)graph.neighbours.size(), ranges_of(graph),
                        graph.neighbours.data(), weighted ? graph.neighbourWeights.data() : nullptr,
                        &dgraph) != MM_OK) {
        cerr << "Not enough memory on device to store this graph! (" << mm_last_error() << ")" << endl;
        throw exception();
    }
}

GraphMatchingGPU::~GraphMatchingGPU()
{
    RandStreamGuard guard;
    mm_graph_destroy(dgraph);
}

void GraphMatchingGPU::run(vector<int> &match, hipEvent_t &t1, hipEvent_t &t2, int kind, bool maximal) const
{
    assert((int)match.size() == graph.nrVertices);
    vector<unsigned int> seeds;
    if (maximal) {
        seeds = peekRand(MM_NR_MAX_MATCH_ROUNDS);
    } else {
        seeds.resize(MM_NR_MATCH_ROUNDS);
        for (unsigned int &s : seeds) s = (unsigned int)rand();
    }
    mm_opts o;
    std::memset(&o, 0, sizeof(o));
    o.kind = kind;
    o.maximal = maximal ? 1 : 0;
    o.max_rounds = maximal ? MM_NR_MAX_MATCH_ROUNDS : MM_NR_MATCH_ROUNDS;
    o.barrier = selectBarrier;
    o.seeds = seeds.data();
    o.n_seeds = (int32_t)seeds.size();
    o.block_size = (threadsPerBlock >= 64 && threadsPerBlock <= 1024 && threadsPerBlock % 64 == 0) ? threadsPerBlock : 0;
    mm_stats st;
    int32_t rc_match, rc_copy = MM_OK;
    {
        RandStreamGuard guard;
        hipEventRecord(t1, 0);
        hipEventSynchronize(t1);
        rc_match = mm_match(dgraph, &o, nullptr, &st);
        hipEventRecord(t2, 0);
        hipEventSynchronize(t2);
        if (rc_match == MM_OK && graph.nrVertices > 0) rc_copy = mm_graph_copy_match(dgraph, match.data());
    }
    if (rc_match != MM_OK) {
        cerr << "A HIP error occurred during the matching process: " << mm_last_error() << endl;
        throw exception();
    }
    if (rc_copy != MM_OK) {
        cerr << "Unable to retrieve data!" << endl;
        throw exception();
    }
    rounds = st.rounds;
    if (maximal)
        for (int i = 0; i < st.rounds; ++i) (void)rand();
}

GraphMatchingGPURandom::GraphMatchingGPURandom(const Graph &g, const int &t, const unsigned int &b)
    : GraphMatchingGPU(g, t, b) {}
GraphMatchingGPURandom::~GraphMatchingGPURandom() {}
void GraphMatchingGPURandom::performMatching(vector<int> &m, hipEvent_t &t1, hipEvent_t &t2) const
{
    run(m, t1, t2, MM_RANDOM, false);
}

GraphMatchingGPURandomMaximal::GraphMatchingGPURandomMaximal(const Graph &g, const int &t, const unsigned int &b)
    : GraphMatchingGPU(g, t, b) {}
GraphMatchingGPURandomMaximal::~GraphMatchingGPURandomMaximal() {}
void GraphMatchingGPURandomMaximal::performMatching(vector<int> &m, hipEvent_t &t1, hipEvent_t &t2) const
{
    run(m, t1, t2, MM_RANDOM, true);
}

GraphMatchingGPUWeighted::GraphMatchingGPUWeighted(const Graph &g, const int &t, const unsigned int &b)
    : GraphMatchingGPU(g, t, b)
{
    assert(graph.neighbourWeights.size() == graph.neighbours.size());
}
GraphMatchingGPUWeighted::~GraphMatchingGPUWeighted() {}
void GraphMatchingGPUWeighted::performMatching(vector<int> &m, hipEvent_t &t1, hipEvent_t &t2) const
{
    run(m, t1, t2, MM_WEIGHTED, false);
}

GraphMatchingGPUWeightedMaximal::GraphMatchingGPUWeightedMaximal(const Graph &g, const int &t, const unsigned int &b)
    : GraphMatchingGPU(g, t, b)
{
    assert(graph.neighbourWeights.size() == graph.neighbours.size());
}
GraphMatchingGPUWeightedMaximal::~GraphMatchingGPUWeightedMaximal() {}
void GraphMatchingGPUWeightedMaximal::performMatching(vector<int> &m, hipEvent_t &t1, hipEvent_t &t2) const
{
    run(m, t1, t2, MM_WEIGHTED, true);
}

}  // namespace mtc
