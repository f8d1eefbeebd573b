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
//     first slot), each block composes its threads' maps, the host chains the
//     blocks, and a second pass walks every run from its true entry;
//   - values outside their graph's range, which only invalid arrays hold,
//     come after all in-range values of their graph; they are listed, sorted
//     and walked on the host from the state the scan leaves at the graph's
//     end, and their pairs are tested for adjacency on the device.
// A separate per-vertex pass counts the unmatched vertices with an unmatched
// neighbour and sums getWeight's size and weight (the first neighbour in CSR
// order that shares the value); its per-graph sums are kept per lane while
// the lane stays in one graph (graphs are contiguous caller-id ranges).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <tuple>
#include <vector>

#include "../../include/mm_match.h"
#include "mm_graph_impl.h"
#include "mm_internal.h"

namespace {

constexpr int CBLOCK = 256;
constexpr int SLOTS_PER_THREAD = 64;
constexpr int NACC = 6;  // under, multi, invalid, matchable, size, weight (double bits)

struct CDev {
    int32_t n;
    const int2 *ranges;   // internal rows
    const int32_t *nbrs;  // CSR order, internal ids
    const int2 *adj;      // weight-sorted rows (weighted graphs) or null
    const int32_t *res;   // result in the caller's order
    const int32_t *mint;  // the same result in the internal order
    const int32_t *perm;  // caller id -> internal id (null: the same)
    const int32_t *voff;  // batch: vertex offsets of the graphs (caller ids); null: one graph
    int32_t n_graphs;
    int32_t *count, *first, *second, *third;  // [n] per slot of in-range values
    int32_t *oob;              // [0] vertices whose value is outside their graph's range, [1] list length
    int2 *oob_list;            // {caller id, value} of those vertices
    int32_t ipb;               // caller ids per block of the per-vertex pass
    uint8_t *tmap;             // per thread of the scan: bit e = exit entry for entry e
    uint8_t *bmap;             // per block of the scan
    const uint8_t *bentry;     // per block of the scan: entry of its first slot
    int32_t *gexit;            // per graph: entry state after its last in-range slot
    unsigned long long *part;  // [blocks][NACC]
    double *gweight;           // batch: per graph (weight summed over both endpoints)
    unsigned long long *gsize;
    int32_t *gbad;
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
__device__ __host__ __forceinline__ int next_entry(int32_t s, int e) { return s == 1 + e ? 1 : 0; }
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
    if ((threadIdx.x & 63) == 0 && oob) atomicAdd(d.oob, oob);
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
            if (pass == 0 && d.oob_list) d.oob_list[atomicAdd(d.oob + 1, 1)] = make_int2(c, m);
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
                if (bad && d.gbad) atomicMax(d.gbad + gi, 1);
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

// Adjacency of caller-id pairs (out-of-range groups): ok[k] = 1 when the two
// are each other's neighbours.
__global__ void k_c_pairs(CDev d, const int2 *pairs, int32_t count, int32_t *ok)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x)
        ok[k] = adjacent_callers(d, pairs[k].x, pairs[k].y) ? 1 : 0;
}

__global__ void k_c_zero(double *gw, unsigned long long *gs, int32_t *gb, int32_t *ge, int32_t count)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        if (gw) gw[i] = 0.0;
        if (gs) gs[i] = 0ull;
        if (gb) gb[i] = 0;
        ge[i] = 0;
    }
}

#define CK_TRY(expr, what)                                                                         \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return mm::set_errorf(e_ == hipErrorOutOfMemory ? MM_ERR_NOMEM : MM_ERR_HIP, "%s: %s", \
                                  what, hipGetErrorString(e_));                                    \
    } while (0)

// Device memory freed when the check returns.
struct Tmp {
    std::vector<void *> ptrs;
    ~Tmp()
    {
        for (void *q : ptrs) (void)hipFree(q);
    }
    hipError_t get(void **p, size_t bytes)
    {
        *p = nullptr;
        const hipError_t e = hipMalloc(p, bytes > 0 ? bytes : 1);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

int32_t run_check(mm_graph *g, const int32_t *voff, int32_t n_graphs, mm_check *out, double *gw_out,
                  int64_t *gs_out, int32_t *gv_out)
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
    const int64_t n = g->n;
    const int grid = (int)std::min<int64_t>((n + CBLOCK - 1) / CBLOCK, (int64_t)g->num_cu * 8);
    const int sgrid = (int)((n + (int64_t)CBLOCK * SLOTS_PER_THREAD - 1) / ((int64_t)CBLOCK * SLOTS_PER_THREAD));
    const int pblocks = std::max(grid, sgrid);
    int32_t rc;
    void *p = nullptr;
    if (!g->d_chk) {
        if ((rc = mm_graph_alloc(g, &p, sizeof(int32_t) * 5 * (size_t)n + 2 * sizeof(int32_t), "check scratch")) != MM_OK)
            return rc;
        g->d_chk = (int32_t *)p;
        g->chk_perm = false;
    }
    Tmp tmp;
    CK_TRY(tmp.get(&p, sizeof(unsigned long long) * NACC * (size_t)pblocks), "Not enough memory for check sums");
    unsigned long long *part = (unsigned long long *)p;
    CK_TRY(tmp.get(&p, (size_t)sgrid * (CBLOCK + 2)), "Not enough memory for check sums");
    uint8_t *maps = (uint8_t *)p;
    int32_t *perm = g->d_chk + 4 * (size_t)n;
    int32_t *oob = g->d_chk + 5 * (size_t)n;
    const int pgrid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)g->num_cu * 8);
    if (g->relabelled && !g->chk_perm) {
        hipLaunchKernelGGL(k_c_perm, dim3(pgrid), dim3(256), 0, g->stream, (const int2 *)g->d_unperm, g->n, perm);
        CK_TRY(hipGetLastError(), "k_c_perm");
        g->chk_perm = true;
    }
    if (g->internal_stale) {
        // The last plain run wrote only the caller-order result: bring the
        // internal-order array the neighbour tests read up to date.
        hipLaunchKernelGGL(k_c_permute, dim3(pgrid), dim3(256), 0, g->stream, (const int2 *)g->d_unperm, g->n,
                           (const int32_t *)g->d_match_out, g->d_match);
        CK_TRY(hipGetLastError(), "k_c_permute");
        g->internal_stale = false;
    }
    const int32_t ng = voff && n_graphs > 0 ? n_graphs : 1;
    CDev d;
    std::memset(&d, 0, sizeof(d));
    d.n = g->n;
    d.ranges = g->d_ranges;
    d.nbrs = g->d_nbrs;
    d.adj = g->weighted ? g->d_adj : nullptr;
    d.res = g->relabelled ? g->d_match_out : g->d_match;
    d.mint = g->d_match;
    d.perm = g->relabelled ? perm : nullptr;
    d.count = g->d_chk;
    d.first = g->d_chk + n;
    d.second = g->d_chk + 2 * (size_t)n;
    d.third = g->d_chk + 3 * (size_t)n;
    d.oob = oob;
    d.part = part;
    d.tmap = maps;
    d.bmap = maps + (size_t)sgrid * CBLOCK;
    d.bentry = maps + (size_t)sgrid * (CBLOCK + 1);
    CK_TRY(tmp.get(&p, sizeof(int32_t) * (size_t)ng), "Not enough memory for per-graph state");
    d.gexit = (int32_t *)p;
    if (voff && n_graphs > 0) {
        d.voff = g->d_voff;
        d.n_graphs = n_graphs;
        CK_TRY(tmp.get(&p, sizeof(double) * (size_t)ng), "Not enough memory for per-graph sums");
        d.gweight = (double *)p;
        CK_TRY(tmp.get(&p, sizeof(unsigned long long) * (size_t)ng), "Not enough memory for per-graph sums");
        d.gsize = (unsigned long long *)p;
        CK_TRY(tmp.get(&p, sizeof(int32_t) * (size_t)ng), "Not enough memory for per-graph sums");
        d.gbad = (int32_t *)p;
    }
    hipLaunchKernelGGL(k_c_zero, dim3(std::max(1, std::min(1024, (ng + 255) / 256))), dim3(256), 0, g->stream,
                       d.gweight, d.gsize, d.gbad, d.gexit, ng);
    CK_TRY(hipMemsetAsync(oob, 0, 2 * sizeof(int32_t), g->stream), "check counters");
    hipLaunchKernelGGL(k_c_init, dim3(grid), dim3(CBLOCK), 0, g->stream, d);
    CK_TRY(hipGetLastError(), "k_c_init");
    int32_t h_oob[2] = {0, 0};
    CK_TRY(hipMemcpyAsync(h_oob, oob, sizeof(h_oob), hipMemcpyDeviceToHost, g->stream), "check counters");
    CK_TRY(hipStreamSynchronize(g->stream), "check counters");
    if (h_oob[0] > 0) {
        CK_TRY(tmp.get(&p, sizeof(int2) * (size_t)h_oob[0]), "Not enough memory for the check's value list");
        d.oob_list = (int2 *)p;
    }
    for (int pass = 0; pass < 3; ++pass)
        hipLaunchKernelGGL(k_c_group, dim3(grid), dim3(CBLOCK), 0, g->stream, d, pass);
    hipLaunchKernelGGL(k_c_scan_maps, dim3(sgrid), dim3(CBLOCK), 0, g->stream, d);
    CK_TRY(hipGetLastError(), "result check kernels");
    // Chain the blocks' maps: the entry state of every block's first slot.
    std::vector<uint8_t> bmap((size_t)sgrid), bentry((size_t)sgrid);
    CK_TRY(hipMemcpyAsync(bmap.data(), d.bmap, bmap.size(), hipMemcpyDeviceToHost, g->stream), "check maps");
    CK_TRY(hipStreamSynchronize(g->stream), "check maps");
    int e = 0;
    for (int b = 0; b < sgrid; ++b) {
        bentry[(size_t)b] = (uint8_t)e;
        e = (bmap[(size_t)b] >> e) & 1;
    }
    CK_TRY(hipMemcpyAsync(maps + (size_t)sgrid * (CBLOCK + 1), bentry.data(), bentry.size(), hipMemcpyHostToDevice,
                          g->stream),
           "check maps");
    // k_c_scan_walk fills columns 0..2 of the first sgrid rows of the partial
    // sums, k_c_vertices columns 3..5 of the first vgrid rows.
    d.ipb = (int32_t)(((n + grid - 1) / grid + CBLOCK - 1) / CBLOCK * CBLOCK);
    const int vgrid = (int)((n + d.ipb - 1) / d.ipb);
    CK_TRY(hipMemsetAsync(part, 0, sizeof(unsigned long long) * NACC * (size_t)pblocks, g->stream), "check sums");
    hipLaunchKernelGGL(k_c_scan_walk, dim3(sgrid), dim3(CBLOCK), 0, g->stream, d);
    hipLaunchKernelGGL(k_c_vertices, dim3(vgrid), dim3(CBLOCK), 0, g->stream, d);
    CK_TRY(hipGetLastError(), "result check kernels");
    std::vector<unsigned long long> hpart((size_t)pblocks * NACC);
    CK_TRY(hipMemcpyAsync(hpart.data(), part, sizeof(unsigned long long) * hpart.size(), hipMemcpyDeviceToHost,
                          g->stream),
           "Unable to read check sums");
    std::vector<int32_t> gexit((size_t)ng);
    CK_TRY(hipMemcpyAsync(gexit.data(), d.gexit, sizeof(int32_t) * gexit.size(), hipMemcpyDeviceToHost, g->stream),
           "Unable to read check sums");
    std::vector<double> gw;
    std::vector<unsigned long long> gs;
    std::vector<int32_t> gb((size_t)ng, 0);
    if (d.voff) {
        gw.resize((size_t)ng);
        gs.resize((size_t)ng);
        CK_TRY(hipMemcpyAsync(gw.data(), d.gweight, sizeof(double) * gw.size(), hipMemcpyDeviceToHost, g->stream),
               "Unable to read check sums");
        CK_TRY(hipMemcpyAsync(gs.data(), d.gsize, sizeof(unsigned long long) * gs.size(), hipMemcpyDeviceToHost,
                              g->stream),
               "Unable to read check sums");
        CK_TRY(hipMemcpyAsync(gb.data(), d.gbad, sizeof(int32_t) * gb.size(), hipMemcpyDeviceToHost, g->stream),
               "Unable to read check sums");
    }
    std::vector<int2> olist((size_t)h_oob[0]);
    if (h_oob[0] > 0)
        CK_TRY(hipMemcpyAsync(olist.data(), d.oob_list, sizeof(int2) * olist.size(), hipMemcpyDeviceToHost, g->stream),
               "Unable to read the check's value list");
    CK_TRY(hipStreamSynchronize(g->stream), "result check");
    double weight = 0.0;
    for (int b = 0; b < pblocks; ++b) {
        const unsigned long long *q = hpart.data() + (size_t)b * NACC;
        out->under += (int64_t)q[0];
        out->multi += (int64_t)q[1];
        out->invalid += (int64_t)q[2];
        out->matchable += (int64_t)q[3];
        out->size += (int64_t)q[4];
        double w;
        std::memcpy(&w, &q[5], sizeof(w));
        weight += w;
    }
    if (!olist.empty()) {
        // Values outside their graph's range: after every in-range value of the
        // graph, in value order; walked from the state the scan left there.
        std::vector<std::tuple<int32_t, int32_t, int32_t>> items;  // (graph, value, caller id)
        items.reserve(olist.size());
        for (const int2 &q : olist) {
            int32_t gi = 0;
            if (d.voff) gi = (int32_t)(std::upper_bound(voff, voff + n_graphs + 1, q.x) - voff) - 1;
            items.emplace_back(gi, q.y, q.x);
        }
        std::sort(items.begin(), items.end());
        std::vector<int2> pairs;
        std::vector<int32_t> pair_graph;
        size_t a = 0;
        while (a < items.size()) {
            const int32_t gi = std::get<0>(items[a]);
            int ent = gexit[(size_t)gi];
            while (a < items.size() && std::get<0>(items[a]) == gi) {
                size_t b = a;
                while (b < items.size() && std::get<0>(items[b]) == gi && std::get<1>(items[b]) == std::get<1>(items[a]))
                    ++b;
                const int32_t cnt = (int32_t)(b - a);
                if (ent < cnt) {
                    if (ent == cnt - 1) {
                        ++out->under;
                        gb[(size_t)gi] = 1;
                    } else {
                        pairs.push_back(make_int2(std::get<2>(items[a + ent]), std::get<2>(items[a + ent + 1])));
                        pair_graph.push_back(gi);
                        if (cnt - ent - 2 > 0) {
                            out->multi += cnt - ent - 2;
                            gb[(size_t)gi] = 1;
                        }
                    }
                }
                ent = next_entry(cnt, ent);
                a = b;
            }
        }
        if (!pairs.empty()) {
            CK_TRY(tmp.get(&p, sizeof(int2) * pairs.size() + sizeof(int32_t) * pairs.size()),
                   "Not enough memory for the check's pair list");
            int2 *d_pairs = (int2 *)p;
            int32_t *d_ok = (int32_t *)(d_pairs + pairs.size());
            CK_TRY(hipMemcpyAsync(d_pairs, pairs.data(), sizeof(int2) * pairs.size(), hipMemcpyHostToDevice, g->stream),
                   "check pairs");
            const int cnt = (int)pairs.size();
            hipLaunchKernelGGL(k_c_pairs, dim3(std::min(1024, (cnt + 255) / 256)), dim3(256), 0, g->stream, d,
                               (const int2 *)d_pairs, cnt, d_ok);
            CK_TRY(hipGetLastError(), "k_c_pairs");
            std::vector<int32_t> ok(pairs.size());
            CK_TRY(hipMemcpyAsync(ok.data(), d_ok, sizeof(int32_t) * ok.size(), hipMemcpyDeviceToHost, g->stream),
                   "check pairs");
            CK_TRY(hipStreamSynchronize(g->stream), "check pairs");
            for (size_t k = 0; k < ok.size(); ++k)
                if (!ok[k]) {
                    ++out->invalid;
                    gb[(size_t)pair_graph[k]] = 1;
                }
        }
    }
    out->weight = weight / 2.0;
    out->valid = out->under == 0 && out->multi == 0 && out->invalid == 0;
    for (int32_t k = 0; k < n_graphs && d.voff; ++k) {
        if (gw_out) gw_out[k] = gw[(size_t)k] / 2.0;
        if (gs_out) gs_out[k] = (int64_t)gs[(size_t)k];
        if (gv_out) gv_out[k] = gb[(size_t)k] ? 0 : 1;
    }
    return MM_OK;
}

}  // namespace

extern "C" {

int32_t mm_graph_check(mm_graph *g, mm_check *out)
{
    mm::clear_error();
    if (!g) return mm::set_error(MM_ERR_INVALID, "mm_graph_check: graph is NULL");
    return run_check(g, nullptr, 0, out, nullptr, nullptr, nullptr);
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
    return run_check(b->g, b->vertex_offsets.data(), ng, total, graph_weight, graph_size, graph_valid);
}

}  // extern "C"
