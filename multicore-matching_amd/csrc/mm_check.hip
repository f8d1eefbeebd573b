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
// group.  Here, without sorting the vertices:
//   - one pass over the vertices in the internal order (rows read in place)
//     finds getWeight's partner (the first neighbour in CSR order that shares
//     the value), the still-matchable vertices, and whether a vertex has a
//     neighbour other than itself with its value;
//   - a value inside its graph's id range (4 + local id) has a slot of its
//     own, the graph's first caller id plus the local id, so slot order is
//     value order; the same pass counts every slot's members and, in the high
//     half of the same 64-bit word, the members without such a neighbour.  A
//     group of two entered at its first member is then valid exactly when
//     both members have one (each is the other's only same-valued vertex);
//   - the entries e follow from a scan over the slots (maps e -> e' composed
//     per thread, per block, and chained over the blocks by one block);
//   - members of groups of three or more, and values outside their graph's
//     range (invalid arrays only), are listed as (value, caller id) keys and
//     sorted (bitonic network): the walk tests the pair at its entry of a
//     large group by looking its members up there, and one thread walks the
//     out-of-range values from the state the scan leaves at each graph's end.
//
// Every buffer the kernels use lies in one scratch area laid out by
// check_layout(); the host side allocates nothing while it runs the kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>

#include "../../include/mm_match.h"
#include "mm_graph_impl.h"
#include "mm_internal.h"

namespace {

constexpr int CBLOCK = 256;
constexpr int SPT = 8;                          // slots per thread of the scan
constexpr int SLOTS_PER_BLOCK = CBLOCK * SPT;   // slots per block of the scan
constexpr int CHAIN_THREADS = 1024;             // one block chains the scan blocks
constexpr int NACC = 6;  // under, multi, invalid, matchable, size, weight (double bits)
constexpr unsigned long long NO_KEY = ~0ull;
constexpr unsigned long long LONE = 1ull << 32;  // a member without a same-valued neighbour

struct CDev {
    int32_t n;
    const int2 *ranges;   // internal rows
    const int32_t *nbrs;  // CSR order, internal ids
    const int2 *adj;      // weight-sorted rows (weighted graphs) or null
    const float *wcsr;    // CSR-order weights (weighted rows that may repeat an id) or null
    const int32_t *res;   // result in the caller's order
    const int32_t *mint;  // the same result in the internal order
    const int32_t *perm;  // caller id -> internal id (null: the same)
    const int32_t *loc;   // internal id -> caller id minus its graph's first id
    const int32_t *vbase; // internal id -> its graph's first caller id (batches; else null)
    const int32_t *voff;  // batch: vertex offsets of the graphs (caller ids); null: one graph
    int32_t n_graphs;     // graphs (1 without voff)
    unsigned long long *cnt;   // [n] per slot: members | members without a same-valued neighbour << 32
    int32_t *ctr;              // [0] members with a value outside their graph's range, [1] members of
                               // groups of three or more, [2] keys listed
    unsigned long long *keys;  // (value << 32 | caller id) of those members, sorted
    int64_t nkeys;             // keys after sorting (a power of two, padded with NO_KEY)
    int32_t ipb;               // internal ids per block of the vertex pass
    int32_t sblocks;           // blocks of the scan
    int32_t rows;              // rows of part: blocks of the scan / vertex pass, + 1 for the out-of-range walk
    uint8_t *bmap;             // per block of the scan: its map e -> e'
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

// Weight of entry j (v's row, neighbour u): the CSR-order copy when rows may
// repeat an id, else the one entry for u in v's weight-sorted row.
__device__ __forceinline__ float weight_at(const CDev &d, int32_t v, int32_t j, int32_t u)
{
    if (d.wcsr) return d.wcsr[j];
    if (!d.adj) return 1.0f;
    const int2 r = d.ranges[v];
    for (int32_t k0 = r.x; k0 < r.y; k0 += 4) {
        int2 e[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) e[q] = k0 + q < r.y ? d.adj[k0 + q] : make_int2(-1, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (k0 + q < r.y && (e[q].x >= 0 ? e[q].x : ~e[q].x) == u) return __int_as_float(e[q].y);
    }
    return 0.0f;
}

// Entry of the next group after a group of size s entered at e.
__device__ __forceinline__ int next_entry(uint32_t s, int e) { return s == 1u + (uint32_t)e ? 1 : 0; }
// Maps {0,1} -> {0,1} as two bits (bit e = image of e); compose(f, g) = g after f.
__device__ __forceinline__ uint8_t compose(uint8_t f, uint8_t g)
{
    return (uint8_t)(((g >> (f & 1)) & 1) | (((g >> ((f >> 1) & 1)) & 1) << 1));
}
__device__ __forceinline__ int apply(uint8_t f, int e) { return (f >> e) & 1; }
constexpr uint8_t MAP_ID = 2;  // e -> e

// The map of one slot: back to entry 0 at a graph's first slot, then the
// transition of the slot's group (none for an empty slot).
__device__ __forceinline__ uint8_t slot_map(uint32_t members, bool first_of_graph)
{
    uint8_t f = members ? (uint8_t)((members == 1u) | ((members == 2u) << 1)) : MAP_ID;
    return first_of_graph ? compose(0, f) : f;
}

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

// Empties the slots, the per-graph sums, the counters and the partial sums
// (the scan and the vertex pass fill different columns of different numbers
// of rows).
__global__ void k_c_zero(CDev d)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x, t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    ulonglong2 *c2 = reinterpret_cast<ulonglong2 *>(d.cnt);  // (the slots start 256-byte aligned)
    for (int64_t i = t0; i < (int64_t)d.n / 2; i += stride) c2[i] = make_ulonglong2(0ull, 0ull);
    if ((d.n & 1) && t0 == 0) d.cnt[d.n - 1] = 0ull;
    for (int64_t i = t0; i < d.n_graphs; i += stride) {
        d.gweight[i] = 0.0;
        d.gsize[i] = 0ull;
        d.gbad[i] = 0;
        d.gexit[i] = 0;
    }
    for (int64_t i = t0; i < (int64_t)d.rows * NACC; i += stride) d.part[i] = 0ull;
    if (t0 < 4) d.ctr[t0] = 0;
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

template <typename T>
__device__ __forceinline__ T wave_sum(T x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long x)
{
    __shared__ unsigned long long ws[CBLOCK / 64];
    x = wave_sum(x);
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
    x = wave_sum(x);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = x;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < CBLOCK / 64; ++k) s += ws[k];
    return s;
}

constexpr int RU = 4;  // row entries per vertex and step

// Per vertex, in the internal order (block b takes the ids [b * ipb, (b + 1) *
// ipb), CBLOCK at a time): still matchable; getWeight's partner (the first
// same-valued entry in CSR order), size and weight; the vertex's slot, counted
// with whether the vertex has a neighbour other than itself with its value.
// Counts the members whose value lies outside their graph's range.  Row
// entries are read RU at a time, all loads of a step issued before any is
// used; the first RU entries of the weight-sorted row are read beside the
// first step, as the partner's weight is usually among them; the slot counts
// are atomics whose old value is not waited for.
__global__ __launch_bounds__(CBLOCK) void k_c_vertices(CDev d)
{
    unsigned long long matchable = 0, size = 0;
    double weight = 0.0;
    int32_t oob = 0;
    GraphAcc acc;
    int32_t acc_base = -1, gbeg = 0, gend = d.n;
    const int64_t v0 = (int64_t)blockIdx.x * d.ipb;
    const int64_t v1 = v0 + d.ipb < d.n ? v0 + d.ipb : d.n;
    const bool pre = d.adj && !d.wcsr;
    for (int64_t vv = v0 + threadIdx.x; vv < v1; vv += CBLOCK) {
        const int32_t v = (int32_t)vv;
        const int32_t m = d.mint[v];
        const int2 r = d.ranges[v];
        const int32_t base = d.vbase ? d.vbase[v] : 0;
        const int32_t lc = d.loc[v];
        int2 head[RU];
#pragma unroll
        for (int k = 0; k < RU; ++k) head[k] = pre && m >= 4 && r.x + k < r.y ? d.adj[r.x + k] : make_int2(-1, 0);
        int32_t jf = -1, uf = -1;
        // m < 4: an unmatched neighbour other than v; m >= 4: a same-valued one
        bool hit = false;
        for (int32_t j = r.x; j < r.y && !hit; j += RU) {
            int32_t u[RU], mu[RU];
#pragma unroll
            for (int k = 0; k < RU; ++k) u[k] = j + k < r.y ? d.nbrs[j + k] : -1;
#pragma unroll
            for (int k = 0; k < RU; ++k) mu[k] = u[k] >= 0 ? d.mint[u[k]] : -1;
#pragma unroll
            for (int k = 0; k < RU; ++k) {
                if (hit || u[k] < 0) continue;
                if (m < 4) {
                    hit = u[k] != v && mu[k] < 4;
                } else if (mu[k] == m) {
                    if (jf < 0) {
                        jf = j + k;
                        uf = u[k];
                    }
                    hit = u[k] != v;
                }
            }
        }
        if (m < 4) {
            matchable += hit;
            continue;
        }
        if (d.voff && base != acc_base) {
            const int32_t gi = graph_of(d, base + lc);
            post(d, acc);
            acc = GraphAcc();
            acc.g = gi;
            acc_base = base;
            gbeg = graph_begin(d, gi);
            gend = graph_end(d, gi);
        }
        if (jf >= 0) {
            float w = 0.0f;
            bool have = false;
            if (pre) {
#pragma unroll
                for (int k = 0; k < RU; ++k)
                    if (!have && r.x + k < r.y && (head[k].x >= 0 ? head[k].x : ~head[k].x) == uf) {
                        w = __int_as_float(head[k].y);
                        have = true;
                    }
            }
            if (!have) w = weight_at(d, v, jf, uf);
            ++size;
            weight += (double)w;
            acc.size += 1;
            acc.weight += (double)w;
        }
        const int64_t local = (int64_t)m - 4;
        if (local >= (int64_t)gend - gbeg) ++oob;
        else atomicAdd(d.cnt + gbeg + local, 1ull + (hit ? 0ull : LONE));
    }
    if (d.voff) {
        // Lanes that end in the same graph as the whole wavefront post one sum.
        const int g0 = __shfl(acc.g, 0);
        if (__ballot(acc.g == g0) == ~0ull) {
            GraphAcc t;
            t.g = g0;
            t.size = wave_sum(acc.size);
            t.weight = wave_sum(acc.weight);
            if ((threadIdx.x & 63) == 0) post(d, t);
        } else {
            post(d, acc);
        }
    }
    oob = wave_sum(oob);
    if ((threadIdx.x & 63) == 0 && oob) atomicAdd(d.ctr + 0, oob);
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

// Lists (value << 32 | caller id) for every member whose value lies outside
// its graph's range or whose group has three or more members (invalid arrays
// only; the host launches this when the vertex pass counted any).
__global__ void k_c_list(CDev d)
{
    for (int64_t cc = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; cc < d.n; cc += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = (int32_t)cc;
        const int32_t m = d.res[c];
        if (m < 4) continue;
        const int32_t gi = graph_of(d, c);
        const int64_t gbeg = graph_begin(d, gi), local = (int64_t)m - 4;
        if (local >= (int64_t)graph_end(d, gi) - gbeg || (uint32_t)d.cnt[gbeg + local] >= 3u) {
            const int32_t k = atomicAdd(d.ctr + 2, 1);
            if (k < d.nkeys) d.keys[k] = ((unsigned long long)(uint32_t)m << 32) | (uint32_t)c;
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

// A scan thread's slots [s0, s0 + SPT): their counters (0 past the end), the
// graph of each slot and whether it is its graph's first.  Slot s0 is SPT-aligned.
struct ThreadSlots {
    unsigned long long c[SPT];
    int32_t gi[SPT];
    bool first[SPT];
};

__device__ __forceinline__ void load_slots(const CDev &d, int64_t s0, ThreadSlots &t)
{
    if (s0 + SPT <= d.n) {
        const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(d.cnt + s0);
#pragma unroll
        for (int k = 0; k < SPT / 2; ++k) {
            const ulonglong2 x = p[k];
            t.c[2 * k] = x.x;
            t.c[2 * k + 1] = x.y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < SPT; ++k) t.c[k] = s0 + k < d.n ? d.cnt[s0 + k] : 0ull;
    }
    int32_t gi = s0 < d.n ? graph_of(d, (int32_t)s0) : 0;
    int32_t gnext = graph_end(d, gi);
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
        const int64_t s = s0 + k;
        while (s < d.n && s >= gnext) {
            ++gi;
            gnext = graph_end(d, gi);
        }
        t.gi[k] = gi;
        t.first[k] = s < d.n && s == graph_begin(d, gi);
    }
}

// Inclusive scan of the lanes' maps in lane order (lane 0's applied first).
__device__ __forceinline__ uint8_t wave_scan_maps(uint8_t f)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint8_t y = (uint8_t)__shfl_up((int)f, off, 64);
        if (lane >= off) f = compose(y, f);
    }
    return f;
}

// Exclusive scan of the threads' maps over the block; *s_total gets the
// block's map.  Reached by all threads.
__device__ __forceinline__ uint8_t block_exclusive_maps(uint8_t f, uint8_t *s_wave, uint8_t *s_total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint8_t incl = wave_scan_maps(f);
    uint8_t excl = (uint8_t)__shfl_up((int)incl, 1, 64);
    if (lane == 0) excl = MAP_ID;
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint8_t acc = MAP_ID;
        for (int w = 0; w < nw; ++w) {
            const uint8_t t = s_wave[w];
            s_wave[w] = acc;
            acc = compose(acc, t);
        }
        *s_total = acc;
    }
    __syncthreads();
    return compose(s_wave[wave], excl);
}

// The map e -> e' of every scan block's slots; counts the members of groups
// of three or more.
__global__ __launch_bounds__(CBLOCK) void k_c_scan_maps(CDev d)
{
    __shared__ uint8_t s_wave[CBLOCK / 64], s_total;
    ThreadSlots t;
    load_slots(d, ((int64_t)blockIdx.x * CBLOCK + threadIdx.x) * SPT, t);
    uint8_t f = MAP_ID;
    int32_t big = 0;
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
        const uint32_t c = (uint32_t)t.c[k];
        f = compose(f, slot_map(c, t.first[k]));
        if (c >= 3u) big += (int32_t)c;
    }
    (void)block_exclusive_maps(f, s_wave, &s_total);
    if (threadIdx.x == 0) d.bmap[blockIdx.x] = s_total;
    big = wave_sum(big);
    if ((threadIdx.x & 63) == 0 && big) atomicAdd(d.ctr + 1, big);
}

// One block chains the scan blocks' maps: the entry of every block's first
// slot (entry 0 before the first slot).
__global__ __launch_bounds__(CHAIN_THREADS) void k_c_chain(CDev d)
{
    __shared__ uint8_t s_wave[CHAIN_THREADS / 64], s_total;
    const int per = (d.sblocks + CHAIN_THREADS - 1) / CHAIN_THREADS;
    const int b0 = threadIdx.x * per, b1 = min(d.sblocks, b0 + per);
    uint8_t f = MAP_ID;
    for (int b = b0; b < b1; ++b) f = compose(f, d.bmap[b]);
    int e = apply(block_exclusive_maps(f, s_wave, &s_total), 0);
    for (int b = b0; b < b1; ++b) {
        d.bentry[b] = (uint8_t)e;
        e = apply(d.bmap[b], e);
    }
}

// Caller id of member k (in caller-id order) of the group of value m in the
// graph whose ids start at gbeg, from the sorted keys (groups of three or more
// are listed whole).
__device__ __forceinline__ int32_t listed_member(const CDev &d, uint32_t m, int32_t gbeg, int k)
{
    if (d.nkeys <= 0) return 0;
    const unsigned long long want = ((unsigned long long)m << 32) | (uint32_t)gbeg;
    int64_t lo = 0, hi = d.nkeys;
    while (lo < hi) {  // the first key >= want
        const int64_t mid = (lo + hi) >> 1;
        if (d.keys[mid] < want) lo = mid + 1;
        else hi = mid;
    }
    lo += k;
    return lo < d.nkeys ? (int32_t)(uint32_t)d.keys[lo] : 0;  // (always listed: the guard keeps reads in bounds)
}

// Walks every slot from its true entry: under / multi / invalid of each
// group, per-graph validity and each graph's exit state.
__global__ __launch_bounds__(CBLOCK) void k_c_scan_walk(CDev d)
{
    __shared__ uint8_t s_wave[CBLOCK / 64], s_total;
    ThreadSlots t;
    const int64_t s0 = ((int64_t)blockIdx.x * CBLOCK + threadIdx.x) * SPT;
    load_slots(d, s0, t);
    uint8_t f = MAP_ID;
#pragma unroll
    for (int k = 0; k < SPT; ++k) f = compose(f, slot_map((uint32_t)t.c[k], t.first[k]));
    int e = apply(block_exclusive_maps(f, s_wave, &s_total), d.bentry[blockIdx.x]);
    unsigned long long under = 0, multi = 0, invalid = 0;
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
        const int64_t s = s0 + k;
        if (s >= d.n) break;
        if (t.first[k]) e = 0;
        const uint32_t cnt = (uint32_t)t.c[k];
        const uint32_t lone = (uint32_t)(t.c[k] >> 32);
        const int32_t gi = t.gi[k];
        if (cnt) {
            int bad = 0;
            if ((uint32_t)e < cnt) {
                if ((uint32_t)e == cnt - 1u) {
                    ++under;
                    bad = 1;
                } else if (cnt == 2u) {
                    if (lone) {  // entered at 0: the pair is the whole group
                        ++invalid;
                        bad = 1;
                    }
                } else {
                    const int32_t gbeg = graph_begin(d, gi);
                    const uint32_t m = 4u + (uint32_t)(s - gbeg);
                    if (!adjacent_callers(d, listed_member(d, m, gbeg, e), listed_member(d, m, gbeg, e + 1))) {
                        ++invalid;
                        bad = 1;
                    }
                    if (cnt - (uint32_t)e - 2u > 0u) {
                        multi += (unsigned long long)(cnt - (uint32_t)e - 2u);
                        bad = 1;
                    }
                }
            }
            if (bad) atomicMax(d.gbad + gi, 1);
            e = next_entry(cnt, e);
        }
        if (s + 1 == graph_end(d, gi)) d.gexit[gi] = e;
    }
    const unsigned long long tot[3] = {under, multi, invalid};
    for (int k = 0; k < 3; ++k) {
        const unsigned long long v = block_sum_u64(tot[k]);
        if (threadIdx.x == 0) d.part[(size_t)blockIdx.x * NACC + k] = v;
    }
}

// One thread walks the sorted keys of out-of-range values, graph by graph
// from the state each graph's in-range groups left (gexit, advanced in
// place); listed groups of in-range values were walked by the scan.  Keys of
// one value are ordered by caller id, so the keys of one graph and value are
// consecutive, and each graph's values come in ascending order.
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
        if ((int64_t)m - 4 < (int64_t)graph_end(d, gi) - graph_begin(d, gi)) {
            a = b;  // an in-range group
            continue;
        }
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
    int vgrid, sgrid, rows;
    int32_t ipb;
    size_t slots, perm, ctr, tot, part, maps, gexit, gweight, gsize, gbad, keys, bytes;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// key_cap: listed keys to make room for at the end (0: none).
Layout check_layout(const mm_graph *g, int32_t ng, size_t key_cap)
{
    Layout L;
    const int64_t n = g->n;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + CBLOCK - 1) / CBLOCK, (int64_t)g->num_cu * 8));
    L.sgrid = (int)std::max<int64_t>(1, (n + SLOTS_PER_BLOCK - 1) / SLOTS_PER_BLOCK);
    L.ipb = (int32_t)(((n + grid - 1) / grid + CBLOCK - 1) / CBLOCK * CBLOCK);
    if (L.ipb < CBLOCK) L.ipb = CBLOCK;
    L.vgrid = (int)std::max<int64_t>(1, (n + L.ipb - 1) / L.ipb);
    L.rows = std::max(L.sgrid, L.vgrid) + 1;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = align256(off + bytes);
        return at;
    };
    L.slots = take(sizeof(unsigned long long) * (size_t)n);
    L.perm = take(sizeof(int32_t) * (size_t)n);
    L.ctr = take(sizeof(int32_t) * 4);
    L.tot = take(sizeof(unsigned long long) * NACC);
    L.part = take(sizeof(unsigned long long) * NACC * (size_t)L.rows);
    L.maps = take((size_t)L.sgrid * 2);
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
// caller -> internal map of g.  grow_keys: called with the listed count when
// key_cap is too small; returns the keys' address, or null after setting an
// error.
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
        // internal-order array the vertex pass reads up to date.
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
    d.wcsr = g->weighted ? g->d_wcsr : nullptr;
    d.res = g->relabelled ? g->d_match_out : g->d_match;
    d.mint = g->d_match;
    d.perm = g->relabelled ? perm : nullptr;
    d.loc = g->d_loc;
    d.vbase = batch ? g->d_vbase : nullptr;
    d.voff = batch ? g->d_voff : nullptr;
    d.n_graphs = ng;
    d.cnt = (unsigned long long *)(base + L.slots);
    d.ctr = (int32_t *)(base + L.ctr);
    d.ipb = L.ipb;
    d.sblocks = L.sgrid;
    d.rows = L.rows;
    d.bmap = (uint8_t *)(base + L.maps);
    d.bentry = d.bmap + L.sgrid;
    d.gexit = (int32_t *)(base + L.gexit);
    d.part = (unsigned long long *)(base + L.part);
    d.tot = (unsigned long long *)(base + L.tot);
    d.gweight = (double *)(base + L.gweight);
    d.gsize = (unsigned long long *)(base + L.gsize);
    d.gbad = (int32_t *)(base + L.gbad);
    if (!d.loc || (batch && !d.vbase)) return mm::set_error(MM_ERR_INVALID, "mm_graph_check: handle has no vertex ids");
    const int zgrid = (int)std::max<int64_t>(1, std::min<int64_t>((n / 2 + 255) / 256, (int64_t)g->num_cu * 8));
    hipLaunchKernelGGL(k_c_zero, dim3(zgrid), dim3(256), 0, g->stream, d);
    hipLaunchKernelGGL(k_c_vertices, dim3(L.vgrid), dim3(CBLOCK), 0, g->stream, d);
    hipLaunchKernelGGL(k_c_scan_maps, dim3(L.sgrid), dim3(CBLOCK), 0, g->stream, d);
    hipLaunchKernelGGL(k_c_chain, dim3(1), dim3(CHAIN_THREADS), 0, g->stream, d);
    CK_TRY(hipGetLastError(), "k_c_vertices");
    int32_t h_ctr[2] = {0, 0};
    CK_TRY(hipMemcpyAsync(h_ctr, d.ctr, sizeof(h_ctr), hipMemcpyDeviceToHost, g->stream), "check counters");
    CK_TRY(hipStreamSynchronize(g->stream), "check counters");
    const int64_t listed = (int64_t)h_ctr[0] + h_ctr[1];
    if (listed > 0) {
        const size_t sorted = pow2_at_least((size_t)listed);
        if (sorted <= s.key_cap) {
            d.keys = (unsigned long long *)(base + L.keys);
        } else {
            if (!s.grow_keys) return mm::set_error(MM_ERR_INVALID, "mm_graph_check: no room for the listed values' keys");
            d.keys = s.grow_keys(s.g, sorted);
            if (!d.keys) return MM_ERR_NOMEM;
        }
        d.nkeys = (int64_t)sorted;
        const int kgrid = (int)std::min<size_t>(1024, (sorted + 255) / 256);
        hipLaunchKernelGGL(k_c_list, dim3(pgrid), dim3(256), 0, g->stream, d);
        hipLaunchKernelGGL(k_c_pad, dim3(kgrid), dim3(256), 0, g->stream, d.keys, listed, (int64_t)sorted);
        for (int64_t k = 2; k <= (int64_t)sorted; k <<= 1)
            for (int64_t j = k >> 1; j > 0; j >>= 1)
                hipLaunchKernelGGL(k_c_bitonic, dim3(kgrid), dim3(256), 0, g->stream, d.keys, (int64_t)sorted, k, j);
        CK_TRY(hipGetLastError(), "listed keys");
    }
    hipLaunchKernelGGL(k_c_scan_walk, dim3(L.sgrid), dim3(CBLOCK), 0, g->stream, d);
    if (h_ctr[0] > 0) hipLaunchKernelGGL(k_c_oob_walk, dim3(1), dim3(64), 0, g->stream, d, listed);
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
    char *base = nullptr;
    size_t have = 0;
    const int32_t rc = mm_noalloc_scratch(g->device, &base, &have);
    if (rc != MM_OK) return rc;
    if (L.bytes > have)
        return mm::set_errorf(MM_ERR_INVALID, "mm_graph_check_noalloc: the check needs %zu bytes of scratch, %zu fit",
                              L.bytes, have);
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
