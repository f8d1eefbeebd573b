// Device-side result metrics: getWeight (src/matchcpu.cu:64-97) and
// testMatching (src/matchcpu.cu:127-246) of a handle's current result, for
// results too large to check on the host quickly (a 4096-graph batch holds
// ~46 M vertices).  A batch is checked graph by graph, as if testMatching
// ran on each graph alone: graph-local values, groups within one graph.
//
// testMatching sorts the vertices by match value (unmatched first) and walks
// the sorted order: it takes i1 and the next vertex i2; if i2 holds another
// value, i1 counts as "matched to itself" and the walk goes on after i2
// (which is never looked at as the start of its own group); otherwise the
// pair is checked for adjacency and the rest of the group counts as
// "matched to more than one vertex".  So the walk enters every group of
// equal values either at its first member or, after a single-member step,
// at its second: entry e in {0, 1}, with e' = [size == 1 + e] for the next
// group.  Here:
//   - groups are found without sorting: a value inside its graph's id range
//     (4 + local id) has a slot of its own, the graph's first caller id plus
//     the local id, so slot order is value order; per slot a counter and
//     three atomic minima give the size and the three smallest caller ids
//     (the oracle's order within a group; the reference leaves it to an
//     unstable sort);
//   - the entries e follow from a scan over the slots: each thread composes
//     the maps e -> e' of its run of slots (resetting to 0 at every graph's
//     first slot), each block composes its threads' maps, one thread chains
//     the blocks, and a second pass walks every run from its true entry;
//   - values outside their graph's range, which only invalid arrays hold,
//     come after all in-range values of their graph: they are listed as
//     (value, caller id) keys, sorted (bitonic network), and walked by one
//     thread from the state the scan leaves at each graph's end.
// A separate per-vertex pass counts the unmatched vertices with an unmatched
// neighbour and sums getWeight's size and weight (the first neighbour in CSR
// order that shares the value); its per-graph sums are kept per lane while
// the lane stays in one graph (graphs are contiguous caller-id ranges).
//
// Every buffer the kernels use lies in one scratch area laid out by
// check_layout(); the host side allocates nothing while it runs the kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <mutex>

#include "../../include/mm_match.h"
#include "mm_graph_impl.h"
#include "mm_internal.h"

namespace {

constexpr int CBLOCK = 256;
constexpr int SLOTS_PER_THREAD = 64;
constexpr int NACC = 6;  // under, multi, invalid, matchable, size, weight (double bits)
constexpr unsigned long long NO_KEY = ~0ull;

struct CDev {
    int32_t n;
    const int2 *ranges;   // internal rows
    const int32_t *nbrs;  // CSR order, internal ids
    const int2 *adj;      // weight-sorted rows (weighted graphs) or null
    const int32_t *res;   // result in the caller's order
    const int32_t *mint;  // the same result in the internal order
    const int32_t *perm;  // caller id -> internal id (null: the same)
    const int32_t *voff;  // batch: vertex offsets of the graphs (caller ids); null: one graph
    int32_t n_graphs;     // graphs (1 without voff)
    int32_t *count, *first, *second, *third;  // [n] per slot of in-range values
    int32_t *ctr;              // [0] vertices whose value is outside their graph's range, [1] keys listed
    unsigned long long *keys;  // (value << 32 | caller id) of those vertices
    int32_t ipb;               // caller ids per block of the per-vertex pass
    int32_t sgrid;             // blocks of the scan
    int32_t rows;              // rows of part: blocks of the scan / vertex passes, + 1 for the out-of-range walk
    uint8_t *tmap;             // per thread of the scan: bit e = exit entry for entry e
    uint8_t *bmap;             // per block of the scan
    uint8_t *bentry;           // per block of the scan: entry of its first slot
    int32_t *gexit;            // per graph: entry state after its last in-range slot
    unsigned long long *part;  // [rows][NACC]
    unsigned long long *tot;   // [NACC] totals
    double *gweight;           // per graph: weight summed over both endpoints, then halved
    unsigned long long *gsize;
    int32_t *gbad;             // per graph: 1 when invalid, then turned into "valid"
};

__device__ __forceinline__ int32_t graph_of(const CDev &d, int32_t c)
{
    if (!d.voff) return 0;
    int32_t lo = 0, hi = d.n_graphs - 1;
    while (lo < hi) {  // the last graph whose first id is <= c (empty graphs share their first id)
        const int32_t mid = (lo + hi + 1) >> 1;
        if (d.voff[mid] <= c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int32_t graph_begin(const CDev &d, int32_t gi) { return d.voff ? d.voff[gi] : 0; }
__device__ __forceinline__ int32_t graph_end(const CDev &d, int32_t gi) { return d.voff ? d.voff[gi + 1] : d.n; }

__device__ __forceinline__ int32_t internal_of(const CDev &d, int32_t c) { return d.perm ? d.perm[c] : c; }

// Slot of value m (>= 4) held by a vertex of graph gi, or -1 outside the
// graph's id range.
__device__ __forceinline__ int64_t slot_of(const CDev &d, int32_t gi, int32_t m)
{
    const int64_t base = graph_begin(d, gi);
    const int64_t local = (int64_t)m - 4;
    return local < (int64_t)graph_end(d, gi) - base ? base + local : -1;
}

__device__ __forceinline__ bool has_neighbour(const CDev &d, int32_t v, int32_t u)
{
    const int2 r = d.ranges[v];
    for (int32_t j = r.x; j < r.y; ++j)
        if (d.nbrs[j] == u) return true;
    return false;
}

__device__ __forceinline__ bool adjacent_callers(const CDev &d, int32_t a, int32_t b)
{
    const int32_t va = internal_of(d, a), vb = internal_of(d, b);
    return has_neighbour(d, va, vb) && has_neighbour(d, vb, va);
}

__device__ __forceinline__ float weight_to(const CDev &d, int32_t v, int32_t u)
{
    if (!d.adj) return 1.0f;
    const int2 r = d.ranges[v];
    for (int32_t j = r.x; j < r.y; ++j) {
        const int2 e = d.adj[j];
        if ((e.x >= 0 ? e.x : ~e.x) == u) return __int_as_float(e.y);
    }
    return 0.0f;
}

// Entry of the next group after a group of size s entered at e.
__device__ __forceinline__ int next_entry(int32_t s, int e) { return s == 1 + e ? 1 : 0; }
// Maps {0,1} -> {0,1} as two bits (bit e = image of e); compose(f, g) = g after f.
__device__ __forceinline__ uint8_t compose(uint8_t f, uint8_t g)
{
    return (uint8_t)(((g >> (f & 1)) & 1) | (((g >> ((f >> 1) & 1)) & 1) << 1));
}
constexpr uint8_t MAP_ID = 2;  // e -> e

// perm[p.x] = p.y for every {caller id, internal id} pair.
__global__ void k_c_perm(const int2 *pairs, int32_t n, int32_t *perm)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int2 p = pairs[k];
        perm[p.x] = p.y;
    }
}

// match[v] = in[o] for every (o, v) pair: a result in the caller's order into
// the internal order.
__global__ void k_c_permute(const int2 *pairs, int32_t n, const int32_t *in, int32_t *match)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int2 p = pairs[k];
        match[p.y] = in[p.x];
    }
}

// Clears the per-graph sums, the counters and the partial sums (the scan and
// the per-vertex pass fill different columns of different numbers of rows).
__global__ void k_c_zero(CDev d)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n_graphs; i += gridDim.x * blockDim.x) {
        d.gweight[i] = 0.0;
        d.gsize[i] = 0ull;
        d.gbad[i] = 0;
        d.gexit[i] = 0;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.rows * NACC; i += gridDim.x * blockDim.x) d.part[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x < 4) d.ctr[threadIdx.x] = 0;
}

// Empties the slots and counts the vertices whose value lies outside their
// graph's id range.
__global__ void k_c_init(CDev d)
{
    int32_t oob = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.n; i += (int64_t)gridDim.x * blockDim.x) {
        d.count[i] = 0;
        d.first[i] = INT_MAX;
        d.second[i] = INT_MAX;
        d.third[i] = INT_MAX;
        const int32_t m = d.res[i];
        if (m >= 4) oob += slot_of(d, graph_of(d, (int32_t)i), m) < 0;
    }
    for (int o = 32; o > 0; o >>= 1) oob += __shfl_xor(oob, o);
    if ((threadIdx.x & 63) == 0 && oob) atomicAdd(d.ctr, oob);
}

// pass 0: size and smallest member of every group; lists the out-of-range
// vertices.  pass 1: second smallest.  pass 2: third smallest.
__global__ void k_c_group(CDev d, int pass)
{
    for (int64_t cc = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; cc < d.n; cc += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = (int32_t)cc;
        const int32_t m = d.res[c];
        if (m < 4) continue;
        const int64_t s = slot_of(d, graph_of(d, c), m);
        if (s < 0) {
            if (pass == 0 && d.keys)
                d.keys[atomicAdd(d.ctr + 1, 1)] = ((unsigned long long)(uint32_t)m << 32) | (uint32_t)c;
            continue;
        }
        if (pass == 0) {
            atomicAdd(d.count + s, 1);
            atomicMin(d.first + s, c);
        } else if (pass == 1) {
            if (d.first[s] != c) atomicMin(d.second + s, c);
        } else if (d.first[s] != c && d.second[s] != c) {
            atomicMin(d.third + s, c);
        }
    }
}

// Keys [from, to) get the largest key (padding of the sort).
__global__ void k_c_pad(unsigned long long *keys, int64_t from, int64_t to)
{
    for (int64_t i = from + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < to; i += (int64_t)gridDim.x * blockDim.x)
        keys[i] = NO_KEY;
}

// One compare-exchange step (k, j) of a bitonic sorting network over the
// power-of-two count of keys.
__global__ void k_c_bitonic(unsigned long long *keys, int64_t count, int64_t k, int64_t j)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t l = i ^ j;
        if (l <= i) continue;
        const unsigned long long a = keys[i], b = keys[l];
        const bool up = (i & k) == 0;
        if (up ? a > b : a < b) {
            keys[i] = b;
            keys[l] = a;
        }
    }
}

// Slots [s0, s1) of scan thread t.
__device__ __forceinline__ void thread_slots(const CDev &d, int64_t &s0, int64_t &s1)
{
    s0 = ((int64_t)blockIdx.x * CBLOCK + threadIdx.x) * SLOTS_PER_THREAD;
    s1 = s0 + SLOTS_PER_THREAD < d.n ? s0 + SLOTS_PER_THREAD : d.n;
}

// The map e -> e' of every thread's run and of every block.
__global__ __launch_bounds__(CBLOCK) void k_c_scan_maps(CDev d)
{
    __shared__ uint8_t maps[CBLOCK];
    int64_t s0, s1;
    thread_slots(d, s0, s1);
    int a0 = 0, a1 = 1;
    if (s0 < s1) {
        int32_t gi = graph_of(d, (int32_t)s0);
        int32_t gstart = graph_begin(d, gi), gnext = graph_end(d, gi);
        for (int64_t s = s0; s < s1; ++s) {
            while (s >= gnext) {
                ++gi;
                gstart = graph_begin(d, gi);
                gnext = graph_end(d, gi);
            }
            if (s == gstart) a0 = a1 = 0;
            const int32_t cnt = d.count[s];
            if (cnt) {
                a0 = next_entry(cnt, a0);
                a1 = next_entry(cnt, a1);
            }
        }
    }
    const uint8_t f = (uint8_t)(a0 | (a1 << 1));
    d.tmap[(size_t)blockIdx.x * CBLOCK + threadIdx.x] = f;
    maps[threadIdx.x] = f;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint8_t b = MAP_ID;
        for (int t = 0; t < CBLOCK; ++t) b = compose(b, maps[t]);
        d.bmap[blockIdx.x] = b;
    }
}

// One thread chains the blocks' maps: the entry of every block's first slot.
__global__ void k_c_chain(CDev d)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int e = 0;
    for (int b = 0; b < d.sgrid; ++b) {
        d.bentry[b] = (uint8_t)e;
        e = (d.bmap[b] >> e) & 1;
    }
}

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long x)
{
    __shared__ unsigned long long ws[CBLOCK / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = x;
    __syncthreads();
    unsigned long long s = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < CBLOCK / 64; ++k) s += ws[k];
    return s;
}

__device__ __forceinline__ double block_sum_f64(double x)
{
    __shared__ double ws[CBLOCK / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = x;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < CBLOCK / 64; ++k) s += ws[k];
    return s;
}

// Walks every thread's run from its true entry: under / multi / invalid of
// each group, per-graph validity and each graph's exit state.
__global__ __launch_bounds__(CBLOCK) void k_c_scan_walk(CDev d)
{
    __shared__ uint8_t entry[CBLOCK];
    if (threadIdx.x == 0) {
        int e = d.bentry[blockIdx.x];
        for (int t = 0; t < CBLOCK; ++t) {
            entry[t] = (uint8_t)e;
            e = (d.tmap[(size_t)blockIdx.x * CBLOCK + t] >> e) & 1;
        }
    }
    __syncthreads();
    unsigned long long under = 0, multi = 0, invalid = 0;
    int64_t s0, s1;
    thread_slots(d, s0, s1);
    if (s0 < s1) {
        int e = entry[threadIdx.x];
        int32_t gi = graph_of(d, (int32_t)s0);
        int32_t gstart = graph_begin(d, gi), gnext = graph_end(d, gi);
        for (int64_t s = s0; s < s1; ++s) {
            while (s >= gnext) {
                ++gi;
                gstart = graph_begin(d, gi);
                gnext = graph_end(d, gi);
            }
            if (s == gstart) e = 0;
            const int32_t cnt = d.count[s];
            if (cnt) {
                int bad = 0;
                if (e < cnt) {
                    if (e == cnt - 1) {
                        ++under;
                        bad = 1;
                    } else {
                        const int32_t a = e == 0 ? d.first[s] : d.second[s];
                        const int32_t b = e == 0 ? d.second[s] : d.third[s];
                        if (!adjacent_callers(d, a, b)) {
                            ++invalid;
                            bad = 1;
                        }
                        if (cnt - e - 2 > 0) {
                            multi += (unsigned long long)(cnt - e - 2);
                            bad = 1;
                        }
                    }
                }
                if (bad) atomicMax(d.gbad + gi, 1);
                e = next_entry(cnt, e);
            }
            if (s + 1 == gnext) d.gexit[gi] = e;
        }
    }
    const unsigned long long tot[3] = {under, multi, invalid};
    for (int k = 0; k < 3; ++k) {
        const unsigned long long v = block_sum_u64(tot[k]);
        if (threadIdx.x == 0) d.part[(size_t)blockIdx.x * NACC + k] = v;
    }
}

// One thread walks the sorted out-of-range keys, graph by graph from the
// state each graph's in-range groups left (gexit, advanced in place).  Keys
// of one value are ordered by caller id, so the keys of one graph and value
// are consecutive, and each graph's values come in ascending order.
__global__ void k_c_oob_walk(CDev d, int64_t count)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    unsigned long long under = 0, multi = 0, invalid = 0;
    int64_t a = 0;
    while (a < count) {
        const uint32_t m = (uint32_t)(d.keys[a] >> 32);
        const int32_t gi = graph_of(d, (int32_t)(uint32_t)d.keys[a]);
        int64_t b = a + 1;
        while (b < count && (uint32_t)(d.keys[b] >> 32) == m && graph_of(d, (int32_t)(uint32_t)d.keys[b]) == gi) ++b;
        const int64_t cnt = b - a;
        const int e = d.gexit[gi];
        int bad = 0;
        if (e < cnt) {
            if (e == cnt - 1) {
                ++under;
                bad = 1;
            } else {
                if (!adjacent_callers(d, (int32_t)(uint32_t)d.keys[a + e], (int32_t)(uint32_t)d.keys[a + e + 1])) {
                    ++invalid;
                    bad = 1;
                }
                if (cnt - e - 2 > 0) {
                    multi += (unsigned long long)(cnt - e - 2);
                    bad = 1;
                }
            }
        }
        if (bad) d.gbad[gi] = 1;
        d.gexit[gi] = cnt == 1 + e ? 1 : 0;
        a = b;
    }
    unsigned long long *q = d.part + (size_t)(d.rows - 1) * NACC;
    q[0] = under;
    q[1] = multi;
    q[2] = invalid;
}

// One lane's sums for the graph it is in.
struct GraphAcc {
    int32_t g = -1;
    unsigned long long size = 0;
    double weight = 0.0;
};

__device__ __forceinline__ void post(const CDev &d, const GraphAcc &a)
{
    if (a.g < 0) return;
    if (a.size) atomicAdd(d.gsize + a.g, a.size);
    if (a.weight != 0.0) atomicAdd(d.gweight + a.g, a.weight);
}

// Per vertex: still matchable, getWeight's size and weight.  Block b takes the
// caller ids [b * ipb, (b + 1) * ipb), CBLOCK at a time.
__global__ __launch_bounds__(CBLOCK) void k_c_vertices(CDev d)
{
    unsigned long long matchable = 0, size = 0;
    double weight = 0.0;
    GraphAcc acc;
    const int64_t c0 = (int64_t)blockIdx.x * d.ipb;
    const int64_t c1 = c0 + d.ipb < d.n ? c0 + d.ipb : d.n;
    for (int64_t cc = c0 + threadIdx.x; cc < c1; cc += CBLOCK) {
        const int32_t c = (int32_t)cc;
        const int32_t m = d.res[c];
        const int32_t v = internal_of(d, c);
        const int2 r = d.ranges[v];
        if (m < 4) {
            for (int32_t j = r.x; j < r.y; ++j) {
                const int32_t u = d.nbrs[j];
                if (u != v && d.mint[u] < 4) {
                    ++matchable;
                    break;
                }
            }
            continue;
        }
        float w = 0.0f;
        bool found = false;
        for (int32_t j = r.x; j < r.y; ++j) {
            const int32_t u = d.nbrs[j];
            if (d.mint[u] == m) {
                w = weight_to(d, v, u);
                found = true;
                break;
            }
        }
        if (!found) continue;
        ++size;
        weight += (double)w;
        if (d.voff) {
            const int32_t gi = graph_of(d, c);
            if (gi != acc.g) {
                post(d, acc);
                acc = GraphAcc();
                acc.g = gi;
            }
            acc.size += 1;
            acc.weight += (double)w;
        }
    }
    if (d.voff) {
        // Lanes that end in the same graph as the whole wavefront post one sum.
        const int g0 = __shfl(acc.g, 0);
        if (__ballot(acc.g == g0) == ~0ull) {
            unsigned long long s = acc.size;
            double wsum = acc.weight;
            for (int o = 32; o > 0; o >>= 1) {
                s += __shfl_xor(s, o);
                wsum += __shfl_xor(wsum, o);
            }
            if ((threadIdx.x & 63) == 0) {
                GraphAcc t;
                t.g = g0;
                t.size = s;
                t.weight = wsum;
                post(d, t);
            }
        } else {
            post(d, acc);
        }
    }
    const unsigned long long m_sum = block_sum_u64(matchable);
    const unsigned long long s_sum = block_sum_u64(size);
    const double w_sum = block_sum_f64(weight);
    if (threadIdx.x == 0) {
        unsigned long long *q = d.part + (size_t)blockIdx.x * NACC;
        q[3] = m_sum;
        q[4] = s_sum;
        unsigned long long bits;
        std::memcpy(&bits, &w_sum, sizeof(bits));
        q[5] = bits;
    }
}

// Totals over the rows of partial sums (one block; fixed order), and the
// per-graph outputs: weight halved, "bad" turned into "valid".
__global__ __launch_bounds__(CBLOCK) void k_c_total(CDev d)
{
    unsigned long long acc[5] = {0, 0, 0, 0, 0};
    double w = 0.0;
    for (int r = threadIdx.x; r < d.rows; r += CBLOCK) {
        const unsigned long long *q = d.part + (size_t)r * NACC;
        for (int k = 0; k < 5; ++k) acc[k] += q[k];
        double x;
        std::memcpy(&x, &q[5], sizeof(x));
        w += x;
    }
    for (int k = 0; k < 5; ++k) {
        const unsigned long long s = block_sum_u64(acc[k]);
        if (threadIdx.x == 0) d.tot[k] = s;
    }
    const double ws = block_sum_f64(w);
    if (threadIdx.x == 0) {
        const double half = ws / 2.0;
        std::memcpy(&d.tot[5], &half, sizeof(half));
    }
    for (int i = threadIdx.x; i < d.n_graphs; i += CBLOCK) {
        d.gweight[i] = d.gweight[i] / 2.0;
        d.gbad[i] = d.gbad[i] ? 0 : 1;
    }
}

#define CK_TRY(expr, what)                                                                         \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return mm::set_errorf(e_ == hipErrorOutOfMemory ? MM_ERR_NOMEM : MM_ERR_HIP, "%s: %s", \
                                  what, hipGetErrorString(e_));                                    \
    } while (0)

// Scratch of one check: byte offsets into one area, each 256-byte aligned.
struct Layout {
    int grid, sgrid, vgrid, rows;
    int32_t ipb;
    size_t slots, perm, ctr, tot, part, maps, gexit, gweight, gsize, gbad, keys, bytes;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// key_cap: out-of-range keys to make room for at the end (0: none).
Layout check_layout(const mm_graph *g, int32_t ng, size_t key_cap)
{
    Layout L;
    const int64_t n = g->n;
    L.grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + CBLOCK - 1) / CBLOCK, (int64_t)g->num_cu * 8));
    L.sgrid = (int)std::max<int64_t>(1, (n + (int64_t)CBLOCK * SLOTS_PER_THREAD - 1) / ((int64_t)CBLOCK * SLOTS_PER_THREAD));
    L.ipb = (int32_t)(((n + L.grid - 1) / L.grid + CBLOCK - 1) / CBLOCK * CBLOCK);
    if (L.ipb < CBLOCK) L.ipb = CBLOCK;
    L.vgrid = (int)std::max<int64_t>(1, (n + L.ipb - 1) / L.ipb);
    L.rows = std::max(L.sgrid, L.vgrid) + 1;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = align256(off + bytes);
        return at;
    };
    L.slots = take(sizeof(int32_t) * 4 * (size_t)n);
    L.perm = take(sizeof(int32_t) * (size_t)n);
    L.ctr = take(sizeof(int32_t) * 4);
    L.tot = take(sizeof(unsigned long long) * NACC);
    L.part = take(sizeof(unsigned long long) * NACC * (size_t)L.rows);
    L.maps = take((size_t)L.sgrid * (CBLOCK + 2));
    L.gexit = take(sizeof(int32_t) * (size_t)ng);
    L.gweight = take(sizeof(double) * (size_t)ng);
    L.gsize = take(sizeof(unsigned long long) * (size_t)ng);
    L.gbad = take(sizeof(int32_t) * (size_t)ng);
    L.keys = take(sizeof(unsigned long long) * key_cap);
    L.bytes = off;
    return L;
}

size_t pow2_at_least(size_t x)
{
    size_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Where a check's scratch comes from.  base: the area of check_layout(g, ng,
// 0) (or, when key_cap > 0, of check_layout(g, ng, key_cap) with room for
// that many keys at L.keys).  perm_ready: base + L.perm already holds the
// caller -> internal map of g.  grow_keys: called with the out-of-range count
// when key_cap is too small; returns the keys' address, or null after setting
// an error.
struct Scratch {
    char *base = nullptr;
    size_t key_cap = 0;
    bool *perm_ready = nullptr;
    mm_graph *g = nullptr;
    unsigned long long *(*grow_keys)(mm_graph *, size_t) = nullptr;
};

// Runs the check with its scratch in s; per-graph outputs (may be NULL) go
// straight from the device to the caller.
int32_t run_check(mm_graph *g, const int32_t *voff, int32_t n_graphs, const Scratch &s, mm_check *out,
                  double *gw_out, int64_t *gs_out, int32_t *gv_out)
{
    const bool batch = voff && n_graphs > 0;
    const int32_t ng = batch ? n_graphs : 1;
    const Layout L = check_layout(g, ng, 0);
    const int64_t n = g->n;
    char *base = s.base;
    int32_t *perm = (int32_t *)(base + L.perm);
    const int pgrid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)g->num_cu * 8);
    if (g->relabelled && !*s.perm_ready) {
        hipLaunchKernelGGL(k_c_perm, dim3(pgrid), dim3(256), 0, g->stream, (const int2 *)g->d_unperm, g->n, perm);
        CK_TRY(hipGetLastError(), "k_c_perm");
        *s.perm_ready = true;
    }
    if (g->internal_stale) {
        // The last plain run wrote only the caller-order result: bring the
        // internal-order array the neighbour tests read up to date.
        hipLaunchKernelGGL(k_c_permute, dim3(pgrid), dim3(256), 0, g->stream, (const int2 *)g->d_unperm, g->n,
                           (const int32_t *)g->d_match_out, g->d_match);
        CK_TRY(hipGetLastError(), "k_c_permute");
        g->internal_stale = false;
    }
    CDev d;
    std::memset(&d, 0, sizeof(d));
    d.n = g->n;
    d.ranges = g->d_ranges;
    d.nbrs = g->d_nbrs;
    d.adj = g->weighted ? g->d_adj : nullptr;
    d.res = g->relabelled ? g->d_match_out : g->d_match;
    d.mint = g->d_match;
    d.perm = g->relabelled ? perm : nullptr;
    d.voff = batch ? g->d_voff : nullptr;
    d.n_graphs = ng;
    int32_t *slots = (int32_t *)(base + L.slots);
    d.count = slots;
    d.first = slots + n;
    d.second = slots + 2 * (size_t)n;
    d.third = slots + 3 * (size_t)n;
    d.ctr = (int32_t *)(base + L.ctr);
    d.ipb = L.ipb;
    d.sgrid = L.sgrid;
    d.rows = L.rows;
    d.tmap = (uint8_t *)(base + L.maps);
    d.bmap = d.tmap + (size_t)L.sgrid * CBLOCK;
    d.bentry = d.bmap + L.sgrid;
    d.gexit = (int32_t *)(base + L.gexit);
    d.part = (unsigned long long *)(base + L.part);
    d.tot = (unsigned long long *)(base + L.tot);
    d.gweight = (double *)(base + L.gweight);
    d.gsize = (unsigned long long *)(base + L.gsize);
    d.gbad = (int32_t *)(base + L.gbad);
    hipLaunchKernelGGL(k_c_zero, dim3(std::max(1, std::min(1024, (ng + 255) / 256))), dim3(256), 0, g->stream, d);
    hipLaunchKernelGGL(k_c_init, dim3(L.grid), dim3(CBLOCK), 0, g->stream, d);
    CK_TRY(hipGetLastError(), "k_c_init");
    int32_t oob = 0;
    CK_TRY(hipMemcpyAsync(&oob, d.ctr, sizeof(oob), hipMemcpyDeviceToHost, g->stream), "check counters");
    CK_TRY(hipStreamSynchronize(g->stream), "check counters");
    size_t sorted = 0;
    if (oob > 0) {
        sorted = pow2_at_least((size_t)oob);
        if (sorted <= s.key_cap) {
            d.keys = (unsigned long long *)(base + L.keys);
        } else {
            if (!s.grow_keys)
                return mm::set_error(MM_ERR_INVALID, "mm_graph_check: no room for the out-of-range values' keys");
            d.keys = s.grow_keys(s.g, sorted);
            if (!d.keys) return MM_ERR_NOMEM;
        }
    }
    for (int pass = 0; pass < 3; ++pass)
        hipLaunchKernelGGL(k_c_group, dim3(L.grid), dim3(CBLOCK), 0, g->stream, d, pass);
    hipLaunchKernelGGL(k_c_scan_maps, dim3(L.sgrid), dim3(CBLOCK), 0, g->stream, d);
    hipLaunchKernelGGL(k_c_chain, dim3(1), dim3(64), 0, g->stream, d);
    hipLaunchKernelGGL(k_c_scan_walk, dim3(L.sgrid), dim3(CBLOCK), 0, g->stream, d);
    hipLaunchKernelGGL(k_c_vertices, dim3(L.vgrid), dim3(CBLOCK), 0, g->stream, d);
    if (oob > 0) {
        const int sgrid_keys = (int)std::min<size_t>(1024, (sorted + 255) / 256);
        hipLaunchKernelGGL(k_c_pad, dim3(sgrid_keys), dim3(256), 0, g->stream, d.keys, (int64_t)oob, (int64_t)sorted);
        for (int64_t k = 2; k <= (int64_t)sorted; k <<= 1)
            for (int64_t j = k >> 1; j > 0; j >>= 1)
                hipLaunchKernelGGL(k_c_bitonic, dim3(sgrid_keys), dim3(256), 0, g->stream, d.keys, (int64_t)sorted, k, j);
        hipLaunchKernelGGL(k_c_oob_walk, dim3(1), dim3(64), 0, g->stream, d, (int64_t)oob);
    }
    hipLaunchKernelGGL(k_c_total, dim3(1), dim3(CBLOCK), 0, g->stream, d);
    CK_TRY(hipGetLastError(), "result check kernels");
    unsigned long long tot[NACC];
    CK_TRY(hipMemcpyAsync(tot, d.tot, sizeof(tot), hipMemcpyDeviceToHost, g->stream), "Unable to read check sums");
    if (batch) {
        if (gw_out)
            CK_TRY(hipMemcpyAsync(gw_out, d.gweight, sizeof(double) * (size_t)ng, hipMemcpyDeviceToHost, g->stream),
                   "Unable to read check sums");
        if (gs_out)
            CK_TRY(hipMemcpyAsync(gs_out, d.gsize, sizeof(int64_t) * (size_t)ng, hipMemcpyDeviceToHost, g->stream),
                   "Unable to read check sums");
        if (gv_out)
            CK_TRY(hipMemcpyAsync(gv_out, d.gbad, sizeof(int32_t) * (size_t)ng, hipMemcpyDeviceToHost, g->stream),
                   "Unable to read check sums");
    }
    CK_TRY(hipStreamSynchronize(g->stream), "result check");
    out->under = (int64_t)tot[0];
    out->multi = (int64_t)tot[1];
    out->invalid = (int64_t)tot[2];
    out->matchable = (int64_t)tot[3];
    out->size = (int64_t)tot[4];
    std::memcpy(&out->weight, &tot[5], sizeof(double));
    out->valid = out->under == 0 && out->multi == 0 && out->invalid == 0;
    return MM_OK;
}

// Handle-owned scratch: grown on demand, kept for the next check.
unsigned long long *grow_handle_keys(mm_graph *g, size_t cap)
{
    if (g->chk_key_cap >= cap) return g->d_chk_keys;
    if (g->d_chk_keys) {
        (void)hipFree(g->d_chk_keys);
        g->device_bytes -= (int64_t)(sizeof(unsigned long long) * g->chk_key_cap);
        g->d_chk_keys = nullptr;
        g->chk_key_cap = 0;
    }
    void *p = nullptr;
    if (mm_graph_alloc(g, &p, sizeof(unsigned long long) * cap, "the check's sort keys") != MM_OK) return nullptr;
    g->d_chk_keys = (unsigned long long *)p;
    g->chk_key_cap = cap;
    return g->d_chk_keys;
}

int32_t check_with_handle_scratch(mm_graph *g, const int32_t *voff, int32_t n_graphs, mm_check *out,
                                  double *gw_out, int64_t *gs_out, int32_t *gv_out)
{
    if (!out) return mm::set_error(MM_ERR_INVALID, "mm_graph_check: out is NULL");
    std::memset(out, 0, sizeof(*out));
    out->valid = 1;
    for (int32_t k = 0; k < n_graphs; ++k) {
        if (gw_out) gw_out[k] = 0.0;
        if (gs_out) gs_out[k] = 0;
        if (gv_out) gv_out[k] = 1;
    }
    if (g->n == 0) return MM_OK;
    CK_TRY(hipSetDevice(g->device), "hipSetDevice");
    const Layout L = check_layout(g, voff && n_graphs > 0 ? n_graphs : 1, 0);
    if (g->chk_bytes < L.bytes) {
        if (g->d_chk) {
            (void)hipFree(g->d_chk);
            g->device_bytes -= (int64_t)g->chk_bytes;
            g->d_chk = nullptr;
            g->chk_bytes = 0;
        }
        void *p = nullptr;
        const int32_t rc = mm_graph_alloc(g, &p, L.bytes, "check scratch");
        if (rc != MM_OK) return rc;
        g->d_chk = p;
        g->chk_bytes = L.bytes;
        g->chk_perm = false;
    }
    Scratch s;
    s.base = (char *)g->d_chk;
    s.perm_ready = &g->chk_perm;
    s.g = g;
    s.grow_keys = grow_handle_keys;
    return run_check(g, voff, n_graphs, s, out, gw_out, gs_out, gv_out);
}

// Scratch of mm_graph_check_noalloc: a file-scope device array, one copy per
// GPU the library runs on, held until the process ends.  Host code must not
// pass `scratch` itself to a kernel (on the host it names a stand-in, not a
// device address): it uses the address hipGetSymbolAddress returns for the
// current device, fetched once per device.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int32_t noalloc_scratch(int device, char **out)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return mm::set_error(MM_ERR_INVALID, "mm_graph_check_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        CK_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "mm_graph_check_noalloc: hipGetSymbolAddress");
        if (!p) return mm::set_error(MM_ERR_HIP, "mm_graph_check_noalloc: no device address for the scratch array");
        g_scratch_ptr[device] = p;
    }
    *out = (char *)g_scratch_ptr[device];
    return MM_OK;
}

}  // namespace

extern "C" {

int32_t mm_graph_check_noalloc(mm_graph *g, mm_check *out)
{
    mm::clear_error();
    if (!g) return mm::set_error(MM_ERR_INVALID, "mm_graph_check_noalloc: graph is NULL");
    if (!out) return mm::set_error(MM_ERR_INVALID, "mm_graph_check_noalloc: out is NULL");
    std::memset(out, 0, sizeof(*out));
    out->valid = 1;
    if (g->n == 0) return MM_OK;
    if (g->n > MM_CHECK_NOALLOC_MAX_N)
        return mm::set_errorf(MM_ERR_INVALID, "mm_graph_check_noalloc: %d vertices exceed the limit of %d",
                              (int)g->n, (int)MM_CHECK_NOALLOC_MAX_N);
    CK_TRY(hipSetDevice(g->device), "hipSetDevice");
    const size_t key_cap = pow2_at_least((size_t)g->n);
    const Layout L = check_layout(g, 1, key_cap);
    if (L.bytes > NOALLOC_SCRATCH_BYTES)
        return mm::set_errorf(MM_ERR_INVALID, "mm_graph_check_noalloc: the check needs %zu bytes of scratch, %zu fit",
                              L.bytes, NOALLOC_SCRATCH_BYTES);
    char *base = nullptr;
    const int32_t rc = noalloc_scratch(g->device, &base);
    if (rc != MM_OK) return rc;
    bool perm_ready = false;  // the array is shared by every handle on the device
    Scratch s;
    s.base = base;
    s.key_cap = key_cap;
    s.perm_ready = &perm_ready;
    s.g = g;
    return run_check(g, nullptr, 0, s, out, nullptr, nullptr, nullptr);
}

int32_t mm_graph_check(mm_graph *g, mm_check *out)
{
    mm::clear_error();
    if (!g) return mm::set_error(MM_ERR_INVALID, "mm_graph_check: graph is NULL");
    return check_with_handle_scratch(g, nullptr, 0, out, nullptr, nullptr, nullptr);
}

int32_t mm_graph_set_match(mm_graph *g, const int32_t *match)
{
    mm::clear_error();
    if (!g || (!match && g->n > 0)) return mm::set_error(MM_ERR_INVALID, "mm_graph_set_match: NULL argument");
    if (g->n == 0) return MM_OK;
    CK_TRY(hipSetDevice(g->device), "hipSetDevice");
    const size_t bytes = sizeof(int32_t) * (size_t)g->n;
    if (!g->relabelled) {
        CK_TRY(hipMemcpy(g->d_match, match, bytes, hipMemcpyHostToDevice), "Unable to set match");
        return MM_OK;
    }
    g->internal_stale = false;
    CK_TRY(hipMemcpy(g->d_match_out, match, bytes, hipMemcpyHostToDevice), "Unable to set match");
    const int grid = (int)std::min<int64_t>(((int64_t)g->n + 255) / 256, (int64_t)g->num_cu * 8);
    hipLaunchKernelGGL(k_c_permute, dim3(grid), dim3(256), 0, g->stream, (const int2 *)g->d_unperm, g->n,
                       (const int32_t *)g->d_match_out, g->d_match);
    CK_TRY(hipGetLastError(), "k_c_permute");
    CK_TRY(hipStreamSynchronize(g->stream), "Unable to set match");
    return MM_OK;
}

int32_t mm_batch_set_match(mm_batch *b, const int32_t *match)
{
    mm::clear_error();
    if (!b || !b->g) return mm::set_error(MM_ERR_INVALID, "mm_batch_set_match: batch is NULL");
    return mm_graph_set_match(b->g, match);
}

int32_t mm_batch_check(mm_batch *b, mm_check *total, double *graph_weight, int64_t *graph_size, int32_t *graph_valid)
{
    mm::clear_error();
    if (!b || !b->g) return mm::set_error(MM_ERR_INVALID, "mm_batch_check: batch is NULL");
    const int32_t ng = (int32_t)b->vertex_offsets.size() - 1;
    return check_with_handle_scratch(b->g, b->vertex_offsets.data(), ng, total, graph_weight, graph_size, graph_valid);
}

}  // extern "C"
