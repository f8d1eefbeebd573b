// Device-side construction of a graph handle's arrays from the caller's CSR
// (the work behind mm_graph_create / mm_batch_create; the reference's
// constructor is a plain host-to-device copy, src/matchgpu.cu:39-66).
//
// The caller's ranges, neighbours and weights are copied to the device once;
// everything else is derived there:
//   - the symmetry test (every entry (i, u) has its reverse, with the same
//     weight) that makes proposal atomics exact, by binary search in the
//     neighbour's ascending row;
//   - the locality order of the vertices: seeds spread by a hash (one per
//     REGION vertices on average) grow regions by a breadth-first search from
//     all of them at once (a vertex joins the nearest seed, the smallest seed
//     id on ties); vertices the search does not reach get seeds of their own
//     (the smallest id among their unreached neighbours), until none is left.
//     The host orders the regions by a depth-first walk of the region graph
//     (graph by graph for a batch), so consecutive regions touch; vertices
//     follow their region's rank, then their distance from its seed, then
//     their id.  Vertices without neighbours are spread evenly over each
//     graph's order, as buckets of the order get equal work that way;
//   - the internal rows: offsets (a row whose first 8 entries would straddle
//     a 128-byte line starts at the next line, by a scan over the 16 states
//     of offset mod 16), renumbered neighbours, weight-sorted rows (rank
//     sort per lane group for rows of up to 64 entries, a radix sort for
//     longer ones; ties keep CSR order), the reverse positions of the 32-bit
//     proposal keys, the head copy and the caller-order pairs.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "../../include/mm_match.h"
#include "mm_graph_impl.h"
#include "mm_internal.h"

namespace {

constexpr unsigned long long UNSET = ~0ull;       // not reached yet
constexpr unsigned long long ISOLATED = ~1ull;    // no row entries: placed apart
constexpr uint32_t REGION = 2048;                 // mean vertices per hash seed
constexpr int UB = 256;                           // block size of the upload kernels
constexpr int ROWS_PER_THREAD = 64;               // offset scan

__device__ __forceinline__ uint32_t mix_id(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// Slot of this lane's item in a list shared by the whole grid, for lanes with
// pred set: one atomic per wavefront.  Every lane of the wavefront must call it.
__device__ __forceinline__ int32_t wave_slot(bool pred, int32_t *counter)
{
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return -1;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    int32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (int32_t)__popcll(m));
    base = __shfl(base, leader);
    return pred ? base + (int32_t)__popcll(m & ((1ull << lane) - 1ull)) : -1;
}

// Block-wide sum / max / or of one value per thread, posted to *dst with one
// global atomic per block (kernels of many small blocks otherwise queue
// hundreds of thousands of atomics on a single address).  Reached by all
// threads of the block.
template <typename T>
__device__ __forceinline__ T block_reduce(T x, int op)  // op 0: sum, 1: max, 2: or; result in thread 0
{
    __shared__ T s_part[UB / 64];
    for (int o = 32; o > 0; o >>= 1) {
        const T y = __shfl_xor(x, o);
        x = op == 0 ? x + y : (op == 1 ? (x > y ? x : y) : (x | y));
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            const T y = s_part[k];
            x = op == 0 ? x + y : (op == 1 ? (x > y ? x : y) : (x | y));
        }
    return x;
}

__device__ __forceinline__ int32_t graph_of(const int32_t *voff, int32_t ng, int32_t v)
{
    int32_t lo = 0, hi = ng - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (voff[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------- batch

// A batch's rows and ids are given per graph: row ranges shift by the
// graph's first entry, neighbour ids by its first vertex.
__global__ void k_u_rebase_rows(int2 *rg, int32_t n, const int32_t *voff, const long long *eoff, int32_t ng)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k = graph_of(voff, ng, (int32_t)v);
        const int2 r = rg[v];
        rg[v] = make_int2((int32_t)(eoff[k] + r.x), (int32_t)(eoff[k] + r.y));
    }
}

__global__ void k_u_rebase_ids(int32_t *nb, int64_t nnz, const int32_t *voff, const long long *eoff, int32_t ng)
{
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz; j += (int64_t)gridDim.x * blockDim.x) {
        int32_t lo = 0, hi = ng - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (eoff[mid] <= j) lo = mid;
            else hi = mid - 1;
        }
        nb[j] += voff[lo];
    }
}

// ------------------------------------------------------ checks, symmetry

// Rows longer than this are worked on by a whole block each (listed by the
// check); shorter rows by 8 lanes each.
constexpr int32_t LONG_ROW = 256;

// Per entry of row i (graph of ids [0, ids_end)): the id is inside the graph
// and larger than the previous entry's.
__device__ __forceinline__ void check_entry(const int32_t *nb, int64_t lo, int64_t j, int32_t ids_end, int &bad,
                                            int &unsorted)
{
    const int32_t u = nb[j];
    if (u < 0 || u >= ids_end) bad = 1;
    if (j > lo && u <= nb[j - 1]) unsorted = 1;
}

// Row i's entries [lo, hi) in the copied arrays and its graph's id count, or
// false when its range leaves the entries (a batch: its graph's block; the
// ranges are still local to the graph then).
__device__ __forceinline__ bool row_span(const int2 *rg, int64_t i, int32_t n, int64_t nnz, const int32_t *voff,
                                         const long long *eoff, int32_t ng, int64_t &lo, int64_t &hi, int32_t &ids_end)
{
    const int2 r = rg[i];
    int64_t e0 = 0, e1 = nnz;
    ids_end = n;
    if (voff) {
        const int32_t k = graph_of(voff, ng, (int32_t)i);
        e0 = eoff[k];
        e1 = eoff[k + 1];
        ids_end = voff[k + 1] - voff[k];
    }
    lo = e0 + r.x;
    hi = e0 + r.y;
    return r.x >= 0 && r.y >= r.x && hi <= e1;
}

// The CSR validation, before any kernel reads by the caller's values: row
// ranges inside the entries, neighbour ids inside the graph (local to their
// graph in a batch), and whether every row ascends.  Also the largest degree
// (flags[3]) and the degree histogram; rows longer than LONG_ROW are listed
// for k_u_check_long.
__global__ __launch_bounds__(UB) void k_u_check(const int2 *rg, const int32_t *nb, int32_t n, int64_t nnz,
                                                const int32_t *voff, const long long *eoff, int32_t ng, int32_t *flags,
                                                int32_t *longs, unsigned long long *hist)
{
    __shared__ unsigned long long s_hist[33];
    if (threadIdx.x < 33) s_hist[threadIdx.x] = 0;
    __syncthreads();
    const int gl = threadIdx.x & 7;
    const int64_t groups = (int64_t)gridDim.x * (blockDim.x >> 3);
    int bad = 0, unsorted = 0;
    int32_t maxdeg = 0;
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 3) + (threadIdx.x >> 3); i < n; i += groups) {
        int64_t lo, hi;
        int32_t ids_end;
        if (!row_span(rg, i, n, nnz, voff, eoff, ng, lo, hi, ids_end)) {
            bad = 1;
            continue;
        }
        const int32_t deg = (int32_t)(hi - lo);
        if (gl == 0) {
            maxdeg = max(maxdeg, deg);
            if (deg > 0) atomicAdd(s_hist + (31 - __clz((unsigned)deg)), 1ull);
        }
        if (deg > LONG_ROW) {
            if (gl == 0) longs[atomicAdd(flags + 2, 1)] = (int32_t)i;
            continue;
        }
        for (int64_t j = lo + gl; j < hi; j += 8) check_entry(nb, lo, j, ids_end, bad, unsorted);
    }
    maxdeg = block_reduce(maxdeg, 1);
    const int32_t anybad = block_reduce(bad | (unsorted << 1), 2);
    if (threadIdx.x == 0) {
        if (anybad & 1) atomicOr(flags, 1);
        if (anybad & 2) atomicOr(flags + 1, 1);
        if (maxdeg) atomicMax(flags + 3, maxdeg);
    }
    __syncthreads();
    if (threadIdx.x < 33 && s_hist[threadIdx.x]) atomicAdd(hist + threadIdx.x, s_hist[threadIdx.x]);
}

__global__ void k_u_check_long(const int2 *rg, const int32_t *nb, int32_t n, int64_t nnz, const int32_t *voff,
                               const long long *eoff, int32_t ng, int32_t *flags, const int32_t *longs, int32_t nlong)
{
    int bad = 0, unsorted = 0;
    for (int32_t k = blockIdx.x; k < nlong; k += gridDim.x) {
        int64_t lo, hi;
        int32_t ids_end;
        if (!row_span(rg, longs[k], n, nnz, voff, eoff, ng, lo, hi, ids_end)) continue;  // (flagged already)
        for (int64_t j = lo + threadIdx.x; j < hi; j += blockDim.x) check_entry(nb, lo, j, ids_end, bad, unsorted);
    }
    const unsigned long long a = __ballot(bad), b = __ballot(unsorted);
    if ((threadIdx.x & 63) == 0) {
        if (a) atomicOr(flags, 1);
        if (b) atomicOr(flags + 1, 1);
    }
}

// An entry i -> u (u != i) "points down" when u's row is shorter than i's,
// or as long and u > i; of the two entries of a symmetric pair exactly one
// points down.  Every down entry must find i in u's ascending row (the
// shorter one: hub rows are never searched) with an equal weight (both NaN
// count as equal); the reverses found are distinct up entries, so equal down
// and up counts make the pairing a bijection.  rev (may be null) receives
// each found pair's positions: rev[j] = position of the reverse of entry j.
__device__ __forceinline__ void sym_entry(const int2 *rg, const int32_t *nb, const float *w, int32_t i, int32_t deg_i,
                                          int32_t j, int32_t *rev, unsigned long long &down, unsigned long long &up,
                                          int &asym, int &wasym)
{
    const int32_t u = nb[j];
    if (u == i) {
        if (rev) rev[j] = j;  // a self-loop is its own reverse
        return;
    }
    const int2 ru = rg[u];
    const int32_t deg_u = ru.y - ru.x;
    if (!(deg_u < deg_i || (deg_u == deg_i && u > i))) {
        ++up;
        return;
    }
    ++down;
    int32_t lo = ru.x, hi = ru.y;
    while (lo < hi) {
        const int32_t mid = lo + ((hi - lo) >> 1);
        if (nb[mid] < i) lo = mid + 1;
        else hi = mid;
    }
    if (lo == ru.y || nb[lo] != i) {
        asym = 1;
        return;
    }
    if (w) {
        const float a = w[j], c = w[lo];
        if (!(a == c || (a != a && c != c))) wasym = 1;
    }
    if (rev) {
        rev[j] = lo;
        rev[lo] = j;
    }
}

__device__ __forceinline__ void sym_flush(unsigned long long down, unsigned long long up, int asym, int wasym,
                                          unsigned long long *counts, int32_t *flags)
{
    down = block_reduce(down, 0);
    up = block_reduce(up, 0);
    const int32_t f = block_reduce(asym | (wasym << 1), 2);
    if (threadIdx.x == 0) {
        if (down) atomicAdd(counts, down);
        if (up) atomicAdd(counts + 1, up);
        if (f & 1) atomicOr(flags, 1);
        if (f & 2) atomicOr(flags + 1, 1);
    }
}

// Rows up to LONG_ROW entries, 8 lanes per row.
__global__ void k_u_sym(const int2 *rg, const int32_t *nb, const float *w, int32_t n, int32_t *rev,
                        unsigned long long *counts, int32_t *flags)
{
    const int gl = threadIdx.x & 7;
    const int64_t groups = (int64_t)gridDim.x * (blockDim.x >> 3);
    unsigned long long down = 0, up = 0;
    int asym = 0, wasym = 0;
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 3) + (threadIdx.x >> 3); i < n; i += groups) {
        const int2 r = rg[i];
        if (r.y - r.x > LONG_ROW) continue;
        for (int32_t j = r.x + gl; j < r.y; j += 8)
            sym_entry(rg, nb, w, (int32_t)i, r.y - r.x, j, rev, down, up, asym, wasym);
    }
    sym_flush(down, up, asym, wasym, counts, flags);
}

// Longer rows, a block per row.
__global__ void k_u_sym_long(const int2 *rg, const int32_t *nb, const float *w, const int32_t *longs, int32_t nlong,
                             int32_t *rev, unsigned long long *counts, int32_t *flags)
{
    unsigned long long down = 0, up = 0;
    int asym = 0, wasym = 0;
    for (int32_t k = blockIdx.x; k < nlong; k += gridDim.x) {
        const int32_t i = longs[k];
        const int2 r = rg[i];
        for (int32_t j = r.x + threadIdx.x; j < r.y; j += blockDim.x)
            sym_entry(rg, nb, w, i, r.y - r.x, j, rev, down, up, asym, wasym);
    }
    sym_flush(down, up, asym, wasym, counts, flags);
}

// ---------------------------------------------------------- locality order

// label[v] = (distance << 32) | seed id of its region; UNSET / ISOLATED.
__global__ void k_u_label_init(const int2 *rg, int32_t n, unsigned long long *label, int32_t *front, int32_t *fcount)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int2 r = rg[v];
        if (r.y == r.x) {
            label[v] = ISOLATED;
            continue;
        }
        if (mix_id((uint32_t)v) % REGION == 0) {
            label[v] = (unsigned long long)(uint32_t)v;
            front[atomicAdd(fcount, 1)] = (int32_t)v;
        } else {
            label[v] = UNSET;
        }
    }
}

// One level of the search from every seed at once, 8 lanes per frontier
// vertex; a vertex reached for the first time is appended to the next
// frontier exactly once (the first atomicMin sees UNSET), and keeps the
// smallest label offered in this level.  The level's frontier size is read
// from the device (*nin), so several levels run between host read-backs;
// frontier vertices with more than LONG_ROW entries go to lfront for
// k_u_bfs_long.
__device__ __forceinline__ bool bfs_claims(unsigned long long *label, int32_t u, unsigned long long offer)
{
    const unsigned long long cur = label[u];
    if (cur == ISOLATED || cur <= offer) return false;
    return atomicMin(label + u, offer) == UNSET;
}

constexpr int BFS_BUF = 4096;  // a block's next-frontier entries gathered in LDS before one global append

__global__ __launch_bounds__(UB) void k_u_bfs(const int2 *rg, const int32_t *nb, unsigned long long *label,
                                              const int32_t *fin, const int32_t *nin, int32_t *fout, int32_t *nout,
                                              int32_t *lfront, int32_t *nlfront, uint32_t dist)
{
    __shared__ int32_t s_buf[BFS_BUF];
    __shared__ int32_t s_n, s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int gl = threadIdx.x & 7;
    const int32_t count = *nin;
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x >> 3);
    // every lane of a wavefront runs the same iterations (appends are one
    // atomic per wavefront)
    for (int64_t k0 = (int64_t)blockIdx.x * (blockDim.x >> 3) + ((threadIdx.x & ~63) >> 3); k0 < count; k0 += wstride) {
        const int64_t k = k0 + ((threadIdx.x & 63) >> 3);
        int32_t v = -1;
        int2 r = make_int2(0, 0);
        if (k < count) {
            v = fin[k];
            r = rg[v];
        }
        const bool longrow = r.y - r.x > LONG_ROW;
        const int32_t at = wave_slot(longrow && gl == 0, nlfront);
        if (at >= 0) lfront[at] = v;
        if (longrow) r = make_int2(0, 0);
        const unsigned long long offer =
            v >= 0 ? ((unsigned long long)(dist + 1) << 32) | (label[v] & 0xFFFFFFFFull) : 0ull;
        int32_t steps = (r.y - r.x + 7) >> 3;
        for (int o = 32; o > 0; o >>= 1) steps = max(steps, __shfl_xor(steps, o));
        for (int32_t t = 0; t < steps; ++t) {
            const int32_t j = r.x + t * 8 + gl;
            bool claim = false;
            int32_t u = -1;
            if (j < r.y) {
                u = nb[j];
                claim = u != v && bfs_claims(label, u, offer);
            }
            // the block's buffer first; the global list once it is full
            const int32_t at = wave_slot(claim, &s_n);
            if (claim && at < BFS_BUF) s_buf[at] = u;
            const int32_t slot = wave_slot(claim && at >= BFS_BUF, nout);
            if (slot >= 0) fout[slot] = u;
        }
    }
    __syncthreads();
    const int32_t kept = min(s_n, BFS_BUF);
    if (threadIdx.x == 0) s_base = kept ? atomicAdd(nout, kept) : 0;
    __syncthreads();
    for (int32_t k = threadIdx.x; k < kept; k += blockDim.x) fout[s_base + k] = s_buf[k];
}

// The same for the level's frontier vertices with more than LONG_ROW entries,
// a block each.
__global__ void k_u_bfs_long(const int2 *rg, const int32_t *nb, unsigned long long *label, const int32_t *lfront,
                             const int32_t *nlfront, int32_t *fout, int32_t *nout, uint32_t dist)
{
    const int32_t count = *nlfront;
    for (int32_t k = blockIdx.x; k < count; k += gridDim.x) {
        const int32_t v = lfront[k];
        const int2 r = rg[v];
        const unsigned long long offer = ((unsigned long long)(dist + 1) << 32) | (label[v] & 0xFFFFFFFFull);
        for (int32_t j0 = r.x; j0 < r.y; j0 += blockDim.x) {
            const int32_t j = j0 + threadIdx.x;
            bool claim = false;
            int32_t u = -1;
            if (j < r.y) {
                u = nb[j];
                claim = u != v && bfs_claims(label, u, offer);
            }
            const int32_t slot = wave_slot(claim, nout);
            if (claim) fout[slot] = u;
        }
    }
}

// Unreached vertices that are smaller than every unreached neighbour become
// seeds (decided from the labels alone, then applied).
__global__ void k_u_seed_mark(const int2 *rg, const int32_t *nb, const unsigned long long *label, int32_t n,
                              uint8_t *mark, int32_t *count)
{
    int32_t c = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        uint8_t m = 0;
        if (label[v] == UNSET) {
            m = 1;
            const int2 r = rg[v];
            for (int32_t j = r.x; j < r.y; ++j) {
                const int32_t u = nb[j];
                if (u < (int32_t)v && label[u] == UNSET) {
                    m = 0;
                    break;
                }
            }
        }
        mark[v] = m;
        c += m;
    }
    c = block_reduce(c, 0);
    if (threadIdx.x == 0 && c) atomicAdd(count, c);
}

__global__ void k_u_seed_apply(const uint8_t *mark, int32_t n, unsigned long long *label, int32_t *front, int32_t *fcount)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
        if (mark[v]) {
            label[v] = (unsigned long long)(uint32_t)v;
            front[atomicAdd(fcount, 1)] = (int32_t)v;
        }
}

// Region pairs: out[v] = (smaller seed << 32 | larger seed) of v's first
// neighbour in another region, or the largest key when there is none.  Every
// pair of touching regions shares many boundary vertices, so the region graph
// keeps its edges with one pair per vertex.
__global__ void k_u_pairs(const int2 *rg, const int32_t *nb, const unsigned long long *label, int32_t n,
                          unsigned long long *out)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long lv = label[v];
        unsigned long long p = UNSET;
        if (lv != ISOLATED) {
            const uint32_t rv = (uint32_t)lv;
            const int2 r = rg[v];
            for (int32_t j = r.x; j < r.y; ++j) {
                const unsigned long long lu = label[nb[j]];
                if (lu == ISOLATED || (uint32_t)lu == rv) continue;
                const uint32_t ru = (uint32_t)lu;
                p = rv < ru ? ((unsigned long long)rv << 32) | ru : ((unsigned long long)ru << 32) | rv;
                break;
            }
        }
        out[v] = p;
    }
}

// Seeds (in any order) and their count.
__global__ void k_u_seeds(const unsigned long long *label, int32_t n, int32_t *out, int32_t *count)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long l = label[v];
        if (l != ISOLATED && (l >> 32) == 0 && (uint32_t)l == (uint32_t)v) out[atomicAdd(count, 1)] = (int32_t)v;
    }
}

// Sort keys of the final order: regions by rank, then distance, then id;
// isolated vertices after all others, by id.  rank_of_seed[s] is the rank of
// the region whose seed is s.  iso_count[g]: isolated vertices of graph g.
__global__ void k_u_order_keys(const unsigned long long *label, int32_t n, const int32_t *rank_of_seed, int with_dist,
                               const int32_t *voff, int32_t ng, unsigned long long *keys, int32_t *iso_count)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); v0 < n; v0 += stride) {
        const int64_t v = v0 + (threadIdx.x & 63);
        bool iso = false;
        int32_t gi = 0;
        if (v < n) {
            const unsigned long long l = label[v];
            if (l == ISOLATED) {
                keys[v] = (1ull << 63) | (unsigned long long)v;
                iso = true;
                gi = voff ? graph_of(voff, ng, (int32_t)v) : 0;
            } else {
                const unsigned long long rank = (unsigned long long)rank_of_seed[(uint32_t)l];
                const unsigned long long dist = (l >> 32) < 255 ? (l >> 32) : 255;
                keys[v] = with_dist ? (rank << 40) | (dist << 32) | (unsigned long long)v
                                    : (rank << 32) | (unsigned long long)v;
            }
        }
        // isolated vertices of one graph: one atomic per wavefront
        const unsigned long long m = __ballot(iso);
        if (m == 0ull) continue;
        const int leader = __ffsll((long long)m) - 1;
        const int32_t g0 = __shfl(gi, leader);
        if (__ballot(iso && gi != g0) == 0ull) {
            if ((threadIdx.x & 63) == leader) atomicAdd(iso_count + g0, (int32_t)__popcll(m));
        } else if (iso) {
            atomicAdd(iso_count + gi, 1);
        }
    }
}

// out[seed[i]] = rank[i].
__global__ void k_u_scatter_rank(const int32_t *seed, const int32_t *rank, int32_t k, int32_t *out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) out[seed[i]] = rank[i];
}

// Per graph: first position, length, isolated vertices, where its isolated
// and its other vertices start in the sorted keys.  tail: each graph's
// isolated vertices follow its other ones (graphs where they are many: the
// first round's select settles them, and the first round's other kernels
// then scan only the ids before them); else they are spread evenly.
struct Segs {
    const int32_t *lo, *len, *k, *iso_at, *conn_at;
    int32_t count;
    int32_t tail;
};

// order[p] for every position: position q of a graph of len positions and k
// isolated vertices holds an isolated one when floor((q+1)k/len) steps up
// (spread), or when q >= len - k (tail).
__global__ void k_u_merge(const unsigned long long *sorted, Segs s, int32_t n, int32_t *order, int32_t *inv)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        int32_t lo = 0, hi = s.count - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (s.lo[mid] <= p) lo = mid;
            else hi = mid - 1;
        }
        const int64_t q = p - s.lo[lo], len = s.len[lo], k = s.k[lo];
        int64_t at;
        if (s.tail) {
            at = q < len - k ? s.conn_at[lo] + q : s.iso_at[lo] + (q - (len - k));
        } else {
            const int64_t before = (q * k) / len, upto = ((q + 1) * k) / len;
            at = upto > before ? s.iso_at[lo] + before : s.conn_at[lo] + (q - before);
        }
        const int32_t v = (int32_t)(sorted[at] & 0xFFFFFFFFull);
        order[p] = v;
        inv[v] = (int32_t)p;
    }
}

__global__ void k_u_identity(int32_t n, int32_t *order, int32_t *inv)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        order[v] = (int32_t)v;
        inv[v] = (int32_t)v;
    }
}

// ------------------------------------------------------------- row offsets

// Offset scan over the 16 states of (offset mod 16): a row of deg entries
// entered at offset o starts at o, or at the next multiple of 16 when
// align is set and its first min(deg, 8) entries would cross one.
struct OffMap {
    uint8_t exit[16];
    int64_t add[16];
};

__device__ __forceinline__ void row_step(int align, int32_t deg, int &s, int64_t &add)
{
    if (align && s + (deg < 8 ? deg : 8) > 16) {
        add += 16 - s;
        s = 0;
    }
    add += deg;
    s = (int)((s + deg) & 15);
}

__device__ __forceinline__ void compose_off(OffMap &a, const OffMap &b)
{
    OffMap r;
    for (int s = 0; s < 16; ++s) {
        r.exit[s] = b.exit[a.exit[s]];
        r.add[s] = a.add[s] + b.add[a.exit[s]];
    }
    a = r;
}

__global__ __launch_bounds__(UB) void k_u_off_maps(const int2 *rg, const int32_t *order, int32_t n, int align,
                                                   OffMap *tmaps, OffMap *bmaps)
{
    __shared__ OffMap maps[UB];
    const int64_t r0 = ((int64_t)blockIdx.x * UB + threadIdx.x) * ROWS_PER_THREAD;
    int st[16];
    int64_t add[16];
    for (int s = 0; s < 16; ++s) {
        st[s] = s;
        add[s] = 0;
    }
    for (int64_t i = r0; i < r0 + ROWS_PER_THREAD && i < n; ++i) {
        const int2 r = rg[order[i]];
        for (int s = 0; s < 16; ++s) row_step(align, r.y - r.x, st[s], add[s]);
    }
    OffMap m;
    for (int s = 0; s < 16; ++s) {
        m.exit[s] = (uint8_t)st[s];
        m.add[s] = add[s];
    }
    tmaps[(size_t)blockIdx.x * UB + threadIdx.x] = m;
    maps[threadIdx.x] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        OffMap b;
        for (int s = 0; s < 16; ++s) {
            b.exit[s] = (uint8_t)s;
            b.add[s] = 0;
        }
        for (int t = 0; t < UB; ++t) compose_off(b, maps[t]);
        bmaps[blockIdx.x] = b;
    }
}

// One thread chains the blocks: entry state and offset of every block, and
// the total in *total.
__global__ void k_u_off_chain(const OffMap *bmaps, int32_t blocks, int2 *bentry, long long *total)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int s = 0;
    int64_t off = 0;
    for (int b = 0; b < blocks; ++b) {
        bentry[b] = make_int2(s, 0);
        reinterpret_cast<long long *>(bentry + blocks)[b] = off;
        off += bmaps[b].add[s];
        s = bmaps[b].exit[s];
    }
    *total = off;
}

__global__ __launch_bounds__(UB) void k_u_off_walk(const int2 *rg, const int32_t *order, int32_t n, int align,
                                                   const OffMap *tmaps, const int2 *bentry, int32_t blocks, int2 *out)
{
    __shared__ int st_in[UB];
    __shared__ long long off_in[UB];
    if (threadIdx.x == 0) {
        int s = bentry[blockIdx.x].x;
        long long off = reinterpret_cast<const long long *>(bentry + blocks)[blockIdx.x];
        for (int t = 0; t < UB; ++t) {
            st_in[t] = s;
            off_in[t] = off;
            const OffMap &m = tmaps[(size_t)blockIdx.x * UB + t];
            off += m.add[s];
            s = m.exit[s];
        }
    }
    __syncthreads();
    const int64_t r0 = ((int64_t)blockIdx.x * UB + threadIdx.x) * ROWS_PER_THREAD;
    int s = st_in[threadIdx.x];
    long long off = off_in[threadIdx.x];
    for (int64_t i = r0; i < r0 + ROWS_PER_THREAD && i < n; ++i) {
        const int2 r = rg[order[i]];
        const int32_t deg = r.y - r.x;
        int64_t add = 0;
        row_step(align, deg, s, add);
        const long long start = off + (add - deg);
        out[i] = make_int2((int32_t)start, (int32_t)(start + deg));
        off += add;
    }
}

// ------------------------------------------------------- per-vertex arrays

__global__ void k_u_loc(const int32_t *order, int32_t n, const int32_t *voff, int32_t ng, int32_t *loc, int32_t *vbase)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t o = order[i];
        const int32_t b = voff ? voff[graph_of(voff, ng, o)] : 0;
        loc[i] = o - b;
        if (vbase) vbase[i] = b;
    }
}

// CSR-order weights at the internal positions (rows that may repeat an id), 8 lanes per row.
__global__ void k_u_wcsr(const int2 *rg, const float *w, const int32_t *order, const int2 *rn, int32_t n, float *out)
{
    const int gl = threadIdx.x & 7;
    const int64_t groups = (int64_t)gridDim.x * (blockDim.x >> 3);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 3) + (threadIdx.x >> 3); i < n; i += groups) {
        const int2 r = rg[order[i]];
        const int32_t dst = rn[i].x;
        for (int32_t k = gl; k < r.y - r.x; k += 8) out[dst + k] = w[r.x + k];
    }
}

// Renumbered CSR-order rows, 8 lanes per row.
__global__ void k_u_nbrs(const int2 *rg, const int32_t *nb, const int32_t *order, const int32_t *inv, const int2 *rn,
                         int32_t n, int32_t *out)
{
    const int gl = threadIdx.x & 7;
    const int64_t groups = (int64_t)gridDim.x * (blockDim.x >> 3);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 3) + (threadIdx.x >> 3); i < n; i += groups) {
        const int2 r = rg[order[i]];
        const int32_t dst = rn[i].x;
        for (int32_t k = gl; k < r.y - r.x; k += 8) out[dst + k] = inv[nb[r.x + k]];
    }
}

// ---------------------------------------------------------- sorted rows

// Sort key of a row entry: heavier first, -0 equal to +0, weights that can
// never win (NaN, <= -1) last; equal keys keep CSR order.
__device__ __forceinline__ uint32_t entry_key(float w)
{
    if (!(w > -1.0f)) return 0xFFFFFFFFu;
    const uint32_t bits = __float_as_uint(w == 0.0f ? 0.0f : w);
    const uint32_t ord = (bits & 0x80000000u) ? ~bits : (bits | 0x80000000u);
    return ~ord;
}

// out[i] = rn[list[i]].
__global__ void k_u_gather_ranges(const int32_t *list, int32_t k, const int2 *rn, int2 *out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) out[i] = rn[list[i]];
}

// Rows classified by length: [0] up to 16 entries, [1] up to 64, [2] longer.
// A block takes a contiguous chunk of rows: counts them per class in LDS,
// reserves its room with one global atomic per class, then fills it.
__global__ __launch_bounds__(UB) void k_u_classify(const int2 *rn, int32_t n, int32_t *lists, int32_t cap,
                                                   int32_t *counts)
{
    __shared__ int32_t s_cnt[3], s_base[3];
    if (threadIdx.x < 3) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t per = ((int64_t)n + gridDim.x - 1) / gridDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * per, i1 = min((int64_t)n, i0 + per);
    auto class_of = [&](int64_t i) {
        const int32_t deg = rn[i].y - rn[i].x;
        return deg == 0 ? -1 : (deg <= 16 ? 0 : (deg <= 64 ? 1 : 2));
    };
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const int c = class_of(i);
        if (c >= 0) atomicAdd(s_cnt + c, 1);
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        s_base[threadIdx.x] = s_cnt[threadIdx.x] ? atomicAdd(counts + threadIdx.x, s_cnt[threadIdx.x]) : 0;
        s_cnt[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int64_t b0 = i0; b0 < i1; b0 += blockDim.x) {
        const int64_t i = b0 + threadIdx.x;
        const int c = i < i1 ? class_of(i) : -1;
        for (int k = 0; k < 3; ++k) {
            const int32_t at = wave_slot(c == k, s_cnt + k);
            if (c == k) lists[(size_t)k * cap + s_base[k] + at] = (int32_t)i;
        }
    }
}

// Rank sort of rows of at most G entries, G lanes per row: the entry of lane k
// goes to position #{j : key_j < key_k, or key_j == key_k and j < k}.
template <int G>
__global__ void k_u_sort_small(const int32_t *list, int32_t count, const int2 *rg, const float *w, const int32_t *order,
                               const int2 *rn, const int32_t *nbrs_new, int2 *adj, int32_t *spos, int32_t *sperm)
{
    const int gl = threadIdx.x & (G - 1);
    const int64_t groups = (int64_t)gridDim.x * (blockDim.x / G);
    const int64_t start = (int64_t)blockIdx.x * (blockDim.x / G) + threadIdx.x / G;
    // every lane of a wavefront runs the same number of iterations (shuffles)
    const int64_t iters = (count + groups - 1) / groups;
    for (int64_t it = 0; it < iters; ++it) {
        const int64_t k = start + it * groups;
        const bool row = k < count;
        const int32_t i = row ? list[k] : 0;
        const int32_t src = row ? rg[order[i]].x : 0;
        const int32_t dst = row ? rn[i].x : 0;
        const int32_t deg = row ? rn[i].y - dst : 0;
        const bool mine = gl < deg;
        const float wk = mine ? w[src + gl] : 0.0f;
        const uint32_t key = mine ? entry_key(wk) : 0xFFFFFFFFu;
        int rank = 0;
        for (int j = 0; j < G; ++j) {
            const uint32_t kj = __shfl(key, j, G);
            rank += j < deg && (kj < key || (kj == key && j < gl));
        }
        if (mine) {
            const int32_t id = nbrs_new[dst + gl];
            adj[dst + rank] = make_int2(wk > -1.0f ? id : ~id, (int32_t)__float_as_uint(wk));
            if (spos) {
                spos[dst + gl] = rank;
                sperm[dst + rank] = gl;
            }
        }
    }
}

// Long rows: keys (long-row index << 32 | entry key) and values (position in
// the row) for one radix sort; lstart[r] is row r's first slot.
__global__ void k_u_long_keys(const int32_t *list, int32_t count, const long long *lstart, const int2 *rg,
                              const float *w, const int32_t *order, unsigned long long *keys, int32_t *vals)
{
    for (int32_t r = blockIdx.x; r < count; r += gridDim.x) {
        const int32_t i = list[r];
        const int2 src = rg[order[i]];
        const long long at = lstart[r];
        for (int32_t k = threadIdx.x; k < src.y - src.x; k += blockDim.x) {
            keys[at + k] = ((unsigned long long)r << 32) | entry_key(w[src.x + k]);
            vals[at + k] = k;
        }
    }
}

__global__ void k_u_long_place(const int32_t *list, int32_t count, const long long *lstart, const int32_t *vals,
                               const int2 *rg, const float *w, const int32_t *order, const int2 *rn,
                               const int32_t *nbrs_new, int2 *adj, int32_t *spos, int32_t *sperm)
{
    for (int32_t r = blockIdx.x; r < count; r += gridDim.x) {
        const int32_t i = list[r];
        const int32_t src = rg[order[i]].x, dst = rn[i].x, deg = rn[i].y - rn[i].x;
        const long long at = lstart[r];
        for (int32_t p = threadIdx.x; p < deg; p += blockDim.x) {
            const int32_t k = vals[at + p];
            const float wk = w[src + k];
            const int32_t id = nbrs_new[dst + k];
            adj[dst + p] = make_int2(wk > -1.0f ? id : ~id, (int32_t)__float_as_uint(wk));
            if (spos) {
                spos[dst + k] = p;
                sperm[dst + p] = k;
            }
        }
    }
}

// Reverse positions of the 32-bit keys: entry p of u's sorted row came from
// caller entry j of u's row (sperm), whose reverse entry rev[j] lies in the
// caller row of v; v's sorted entry for u gets position p.
__global__ void k_u_adjp(const int2 *rg, const int32_t *order, const int2 *rn, int32_t n, const int2 *adj,
                         const int32_t *spos, const int32_t *sperm, const int32_t *rev, int2 *adjp)
{
    const int gl = threadIdx.x & 7;
    const int64_t groups = (int64_t)gridDim.x * (blockDim.x >> 3);
    for (int64_t u = (int64_t)blockIdx.x * (blockDim.x >> 3) + (threadIdx.x >> 3); u < n; u += groups) {
        const int2 ru = rn[u];
        const int32_t cu = rg[order[u]].x;
        for (int32_t p = gl; p < ru.y - ru.x; p += 8) {
            const int32_t x = adj[ru.x + p].x, v = x >= 0 ? x : ~x;
            const int32_t jr = rev[cu + sperm[ru.x + p]];
            const int2 cv = rg[order[v]];
            if (jr < cv.x || jr >= cv.y) continue;  // (every entry of a symmetric CSR has its reverse)
            const int32_t base = rn[v].x;
            const int32_t at = base + spos[base + (jr - cv.x)];
            reinterpret_cast<int32_t *>(adjp + at)[1] = p;  // (the id half is adj's, copied before)
        }
    }
}

// Every entry of the sorted rows with key positions packed in 4 bytes
// (invalid << 31 | id << pb | position), as the head copy packs its entries.
__global__ void k_u_pack(const int2 *scan, int64_t nnz, uint32_t *packed, int pb)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        const int2 x = scan[i];
        const uint32_t id = (uint32_t)(x.x >= 0 ? x.x : ~x.x);
        packed[i] = (x.x < 0 ? 0x80000000u : 0u) | (id << pb) | ((uint32_t)x.y & ((1u << pb) - 1u));
    }
}

// The first 4 entries of every scanned row: 8-byte entries, or packed in 4
// bytes (invalid << 31 | id << pb | position) when pb > 0.
__global__ void k_u_heads(const int2 *rn, int32_t n, const int2 *scan, int2 *head, uint32_t *headc, int pb)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int2 r = rn[i];
        for (int k = 0; k < 4; ++k) {
            const int2 x = r.x + k < r.y ? scan[r.x + k] : make_int2(0, 0);
            if (headc) {
                const uint32_t id = (uint32_t)(x.x >= 0 ? x.x : ~x.x);
                headc[4 * i + k] = (x.x < 0 ? 0x80000000u : 0u) | (id << pb) | ((uint32_t)x.y & ((1u << pb) - 1u));
            } else {
                head[4 * i + k] = x;
            }
        }
    }
}

// Caller-order pairs grouped by the eighth of the internal id: keys and
// values for a stable 3-bit sort, and the group sizes.
__global__ void k_u_eighths(const int32_t *inv, int32_t n, uint32_t *keys, int32_t *vals, int32_t *counts)
{
    __shared__ int32_t hist[8];
    if (threadIdx.x < 8) hist[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(((int64_t)inv[o] * 8) / n);
        keys[o] = x;
        vals[o] = (int32_t)o;
        atomicAdd(hist + x, 1);
    }
    __syncthreads();
    if (threadIdx.x < 8 && hist[threadIdx.x]) atomicAdd(counts + threadIdx.x, hist[threadIdx.x]);
}

__global__ void k_u_pairs_out(const int32_t *vals, const int32_t *inv, int32_t n, int2 *pairs)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t o = vals[k];
        pairs[k] = make_int2(o, inv[o]);
    }
}

}  // namespace

// ------------------------------------------------------------------ host

struct UploadState {
    mm_graph *g = nullptr;
    int32_t n = 0;
    int64_t nnz = 0;
    int2 *rg = nullptr;
    int32_t *nb = nullptr;
    float *w = nullptr;
    int32_t *voff = nullptr;  // batch: vertex offsets on the device
    std::vector<int32_t> h_voff;
    int32_t *longs = nullptr;  // caller rows with more than LONG_ROW entries
    int32_t nlong = 0;
    bool strict = true;        // every row strictly ascending (no repeated id)
    bool want_rev = false;     // the graph could take 32-bit keys: the symmetry test keeps the reverses
    int32_t *rev = nullptr;    // rev[j]: position of the reverse of entry j (symmetric CSRs, 32-bit keys)
    int32_t *order = nullptr, *inv = nullptr;  // internal order: order[new] = old
    hipStream_t wstream = nullptr;  // the weights' copy runs on its own stream and host thread
    std::thread wthread;
    hipError_t wresult = hipSuccess;
    std::vector<void *> temps;
    // hipFree waits for the whole device, the weights' copy included: while
    // that copy runs, freed scratch is kept here and released by the join.
    std::vector<void *> deferred;
    hipError_t join_weights()
    {
        if (wthread.joinable()) wthread.join();
        for (void *p : deferred) (void)hipFree(p);
        deferred.clear();
        return wresult;
    }
    ~UploadState()
    {
        join_weights();
        if (wstream) (void)hipStreamDestroy(wstream);
        for (void *p : temps) (void)hipFree(p);
    }
};

namespace {

#define UP_TRY(expr, what)                                                                        \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return mm::set_errorf(e_ == hipErrorOutOfMemory ? MM_ERR_NOMEM : MM_ERR_HIP, "%s: %s", \
                                  what, hipGetErrorString(e_));                                   \
    } while (0)

// A copy ordered after the work already queued on stream s (the handle's
// stream does not synchronise with the null stream), then waited for.
hipError_t copy_on(hipStream_t s, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
}

template <typename T>
int32_t tmp_alloc(UploadState *u, T **p, size_t count, const char *what)
{
    void *q = nullptr;
    const hipError_t e = hipMalloc(&q, sizeof(T) * (count > 0 ? count : 1));
    if (e != hipSuccess)
        return mm::set_errorf(MM_ERR_NOMEM, "Not enough memory on device for %s (%zu bytes): %s", what,
                              sizeof(T) * count, hipGetErrorString(e));
    u->temps.push_back(q);
    *p = (T *)q;
    return MM_OK;
}

void tmp_free(UploadState *u, void *p)
{
    for (size_t k = 0; k < u->temps.size(); ++k)
        if (u->temps[k] == p) {
            if (u->wthread.joinable())
                u->deferred.push_back(p);
            else
                (void)hipFree(p);
            u->temps.erase(u->temps.begin() + (long)k);
            return;
        }
}

int grid_for(int64_t items, int per_block = UB)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((items + per_block - 1) / per_block, 65536));
}

// Grid of a grid-stride kernel that posts block totals with global atomics:
// about 8 blocks per CU, enough to fill the chip.
int grid_red(const UploadState *u, int64_t items)
{
    return std::min(grid_for(items), std::max(8, u->g->num_cu * 8));
}

int32_t read_i32(const int32_t *d, hipStream_t s, int32_t *out)
{
    UP_TRY(hipMemcpyAsync(out, d, sizeof(int32_t), hipMemcpyDeviceToHost, s), "upload counters");
    UP_TRY(hipStreamSynchronize(s), "upload counters");
    return MM_OK;
}

// Sorts count 64-bit keys (bits [0, end_bit)) from in to out.
int32_t sort_keys64(UploadState *u, const unsigned long long *in, unsigned long long *out, int64_t count, int end_bit)
{
    hipStream_t s = u->g->stream;
    size_t bytes = 0;
    UP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, in, out, (int)count, 0, end_bit, s), "radix sort");
    char *tmp = nullptr;
    int32_t rc = tmp_alloc(u, &tmp, bytes, "sort storage");
    if (rc != MM_OK) return rc;
    UP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp, bytes, in, out, (int)count, 0, end_bit, s), "radix sort");
    UP_TRY(hipStreamSynchronize(s), "radix sort");
    tmp_free(u, tmp);
    return MM_OK;
}

template <typename K>
int32_t sort_pairs(UploadState *u, const K *kin, K *kout, const int32_t *vin, int32_t *vout, int64_t count, int end_bit)
{
    hipStream_t s = u->g->stream;
    size_t bytes = 0;
    UP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)count, 0, end_bit, s),
           "radix sort");
    char *tmp = nullptr;
    int32_t rc = tmp_alloc(u, &tmp, bytes, "sort storage");
    if (rc != MM_OK) return rc;
    UP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, (int)count, 0, end_bit, s),
           "radix sort");
    UP_TRY(hipStreamSynchronize(s), "radix sort");
    tmp_free(u, tmp);
    return MM_OK;
}

int bits_for(uint64_t x)
{
    int b = 1;
    while (b < 64 && (1ull << b) <= x) ++b;
    return b;
}

// Grows regions from every seed in front[0, count) at once, LEVELS levels
// per host read-back of the frontier size (an even number: the frontier is
// back in fa after each batch).  lc / ll: the frontier and long-frontier sizes
// of each level of a batch; lf: the long frontier.
constexpr int LEVELS = 4;
int32_t grow_regions(UploadState *u, unsigned long long *label, int32_t *fa, int32_t *fb, int32_t *lf, int32_t *lc,
                     int32_t count)
{
    hipStream_t s = u->g->stream;
    int32_t *ll = lc + LEVELS + 1;
    uint32_t dist = 0;
    while (count > 0) {
        UP_TRY(hipMemsetAsync(lc, 0, sizeof(int32_t) * 2 * (LEVELS + 1), s), "upload counters");
        UP_TRY(hipMemcpyAsync(lc, &count, sizeof(int32_t), hipMemcpyHostToDevice, s), "upload counters");
        for (int l = 0; l < LEVELS; ++l, ++dist) {
            int32_t *fin = (l & 1) ? fb : fa, *fout = (l & 1) ? fa : fb;
            hipLaunchKernelGGL(k_u_bfs, dim3(2048), dim3(UB), 0, s, u->rg, u->nb, label, fin, lc + l, fout, lc + l + 1,
                               lf, ll + l, dist);
            if (u->nlong > 0)
                hipLaunchKernelGGL(k_u_bfs_long, dim3(1024), dim3(UB), 0, s, u->rg, u->nb, label, lf, ll + l, fout,
                                   lc + l + 1, dist);
        }
        UP_TRY(hipGetLastError(), "k_u_bfs");
        UP_TRY(copy_on(s, &count, lc + LEVELS, sizeof(int32_t), hipMemcpyDeviceToHost), "upload counters");
    }
    return MM_OK;
}

// order[new] = old and its inverse, on the device.
int32_t locality_order(UploadState *u, int32_t *order, int32_t *inv)
{
    mm_graph *g = u->g;
    hipStream_t s = g->stream;
    const int32_t n = u->n;
    const int32_t ng = u->voff ? (int32_t)u->h_voff.size() - 1 : 1;
    unsigned long long *label = nullptr;
    int32_t *fa = nullptr, *fb = nullptr, *ctr = nullptr;
    int32_t rc;
    int32_t *lf = nullptr, *lc = nullptr;
    if ((rc = tmp_alloc(u, &label, (size_t)n, "region labels")) != MM_OK ||
        (rc = tmp_alloc(u, &fa, (size_t)n, "search frontier")) != MM_OK ||
        (rc = tmp_alloc(u, &fb, (size_t)n, "search frontier")) != MM_OK ||
        (rc = tmp_alloc(u, &lf, (size_t)std::max(u->nlong, 1), "search frontier")) != MM_OK ||
        (rc = tmp_alloc(u, &lc, 2 * (LEVELS + 1), "search counters")) != MM_OK ||
        (rc = tmp_alloc(u, &ctr, 4, "upload counters")) != MM_OK)
        return rc;
    UP_TRY(hipMemsetAsync(ctr, 0, 4 * sizeof(int32_t), s), "upload counters");
    hipLaunchKernelGGL(k_u_label_init, dim3(grid_for(n)), dim3(UB), 0, s, u->rg, n, label, fa, ctr);
    UP_TRY(hipGetLastError(), "k_u_label_init");
    int32_t count = 0;
    if ((rc = read_i32(ctr, s, &count)) != MM_OK) return rc;
    uint8_t *mark = nullptr;
    if ((rc = tmp_alloc(u, &mark, (size_t)n, "seed marks")) != MM_OK) return rc;
    for (;;) {
        if ((rc = grow_regions(u, label, fa, fb, lf, lc, count)) != MM_OK) return rc;
        // vertices no search reached seed regions of their own
        UP_TRY(hipMemsetAsync(ctr, 0, 2 * sizeof(int32_t), s), "upload counters");
        hipLaunchKernelGGL(k_u_seed_mark, dim3(grid_red(u, n)), dim3(UB), 0, s, u->rg, u->nb, label, n, mark, ctr);
        UP_TRY(hipGetLastError(), "k_u_seed_mark");
        int32_t marked = 0;
        if ((rc = read_i32(ctr, s, &marked)) != MM_OK) return rc;
        if (marked == 0) break;
        hipLaunchKernelGGL(k_u_seed_apply, dim3(grid_for(n)), dim3(UB), 0, s, mark, n, label, fa, ctr + 1);
        UP_TRY(hipGetLastError(), "k_u_seed_apply");
        count = marked;
    }
    tmp_free(u, mark);
    tmp_free(u, fb);
    tmp_free(u, lf);
    tmp_free(u, lc);
    mm_upload_mark("order: regions grown", s);
    // Seeds, sorted: region r is the r-th smallest seed.
    int32_t *seeds = fa;
    UP_TRY(hipMemsetAsync(ctr, 0, sizeof(int32_t), s), "upload counters");
    hipLaunchKernelGGL(k_u_seeds, dim3(grid_for(n)), dim3(UB), 0, s, label, n, seeds, ctr);
    UP_TRY(hipGetLastError(), "k_u_seeds");
    int32_t nseeds = 0;
    if ((rc = read_i32(ctr, s, &nseeds)) != MM_OK) return rc;
    std::vector<int32_t> h_seeds((size_t)nseeds);
    UP_TRY(copy_on(s, h_seeds.data(), seeds, sizeof(int32_t) * (size_t)nseeds, hipMemcpyDeviceToHost), "region seeds");
    std::sort(h_seeds.begin(), h_seeds.end());
    // Region pairs, deduplicated on the device.
    unsigned long long *pcount = nullptr, *pa = nullptr, *pb = nullptr;
    if ((rc = tmp_alloc(u, &pcount, 1, "upload counters")) != MM_OK ||
        (rc = tmp_alloc(u, &pa, (size_t)n, "region pairs")) != MM_OK ||
        (rc = tmp_alloc(u, &pb, (size_t)n, "region pairs")) != MM_OK)
        return rc;
    hipLaunchKernelGGL(k_u_pairs, dim3(grid_for(n)), dim3(UB), 0, s, u->rg, u->nb, label, n, pa);
    UP_TRY(hipGetLastError(), "k_u_pairs");
    if ((rc = sort_keys64(u, pa, pb, n, 64)) != MM_OK) return rc;
    std::vector<unsigned long long> h_pairs;
    {
        size_t bytes = 0;
        UP_TRY(hipcub::DeviceSelect::Unique(nullptr, bytes, pb, pa, pcount, (int)n, s), "unique pairs");
        char *tmp = nullptr;
        if ((rc = tmp_alloc(u, &tmp, bytes, "unique storage")) != MM_OK) return rc;
        UP_TRY(hipcub::DeviceSelect::Unique(tmp, bytes, pb, pa, pcount, (int)n, s), "unique pairs");
        unsigned long long nu = 0;
        UP_TRY(copy_on(s, &nu, pcount, sizeof(nu), hipMemcpyDeviceToHost), "unique pairs");
        h_pairs.resize((size_t)nu);
        if (nu) UP_TRY(copy_on(s, h_pairs.data(), pa, sizeof(unsigned long long) * (size_t)nu, hipMemcpyDeviceToHost),
                       "unique pairs");
        if (!h_pairs.empty() && h_pairs.back() == UNSET) h_pairs.pop_back();  // vertices with no such neighbour
        tmp_free(u, tmp);
        tmp_free(u, pa);
        tmp_free(u, pb);
    }
    mm_upload_mark("order: region pairs", s);
    // Region graph (host): adjacency by region index, depth-first ranks, graph
    // by graph (regions never cross graphs), each walk from the smallest
    // unranked region and to the smallest unranked neighbour.
    const int32_t K = nseeds;
    auto region_of = [&](uint32_t seed) {
        return (int32_t)(std::lower_bound(h_seeds.begin(), h_seeds.end(), (int32_t)seed) - h_seeds.begin());
    };
    std::vector<int64_t> adj_off((size_t)K + 1, 0);
    std::vector<int32_t> pa_idx(h_pairs.size()), pb_idx(h_pairs.size());
    for (size_t k = 0; k < h_pairs.size(); ++k) {
        pa_idx[k] = region_of((uint32_t)(h_pairs[k] >> 32));
        pb_idx[k] = region_of((uint32_t)h_pairs[k]);
        ++adj_off[(size_t)pa_idx[k] + 1];
        ++adj_off[(size_t)pb_idx[k] + 1];
    }
    for (int32_t r = 0; r < K; ++r) adj_off[(size_t)r + 1] += adj_off[(size_t)r];
    std::vector<int32_t> adj_list((size_t)adj_off[(size_t)K]);
    {
        std::vector<int64_t> fill(adj_off.begin(), adj_off.end() - 1);
        for (size_t k = 0; k < h_pairs.size(); ++k) {
            adj_list[(size_t)fill[(size_t)pa_idx[k]]++] = pb_idx[k];
            adj_list[(size_t)fill[(size_t)pb_idx[k]]++] = pa_idx[k];
        }
    }
    for (int32_t r = 0; r < K; ++r) std::sort(adj_list.begin() + adj_off[(size_t)r], adj_list.begin() + adj_off[(size_t)r + 1]);
    std::vector<int32_t> rank((size_t)K, -1);
    std::vector<int64_t> cursor(adj_off.begin(), adj_off.end() - 1);
    std::vector<int32_t> stack;
    int32_t next_rank = 0;
    // regions of a graph are contiguous in seed order (seeds are vertex ids)
    for (int32_t r0 = 0; r0 < K; ++r0) {
        if (rank[(size_t)r0] >= 0) continue;
        rank[(size_t)r0] = next_rank++;
        stack.assign(1, r0);
        while (!stack.empty()) {
            const int32_t r = stack.back();
            int32_t nxt = -1;
            while (cursor[(size_t)r] < adj_off[(size_t)r + 1]) {
                const int32_t c = adj_list[(size_t)cursor[(size_t)r]++];
                if (rank[(size_t)c] < 0) {
                    nxt = c;
                    break;
                }
            }
            if (nxt < 0) {
                stack.pop_back();
                continue;
            }
            rank[(size_t)nxt] = next_rank++;
            stack.push_back(nxt);
        }
    }
    mm_upload_mark("order: region walk (host)", nullptr);
    // rank of each seed vertex, for the sort keys
    int32_t *rank_of_seed = nullptr, *d_rank = nullptr;
    if ((rc = tmp_alloc(u, &rank_of_seed, (size_t)n, "region ranks")) != MM_OK ||
        (rc = tmp_alloc(u, &d_rank, (size_t)std::max(K, 1), "region ranks")) != MM_OK)
        return rc;
    if (K > 0) {
        UP_TRY(copy_on(s, seeds, h_seeds.data(), sizeof(int32_t) * (size_t)K, hipMemcpyHostToDevice), "region ranks");
        UP_TRY(copy_on(s, d_rank, rank.data(), sizeof(int32_t) * (size_t)K, hipMemcpyHostToDevice), "region ranks");
        hipLaunchKernelGGL(k_u_scatter_rank, dim3(grid_for(K)), dim3(UB), 0, s, seeds, d_rank, K, rank_of_seed);
        UP_TRY(hipGetLastError(), "k_u_scatter_rank");
    }
    // Final order: sort keys, then spread each graph's isolated vertices.
    unsigned long long *keys = nullptr, *sorted = nullptr;
    int32_t *iso = nullptr;
    if ((rc = tmp_alloc(u, &keys, (size_t)n, "order keys")) != MM_OK ||
        (rc = tmp_alloc(u, &sorted, (size_t)n, "order keys")) != MM_OK ||
        (rc = tmp_alloc(u, &iso, (size_t)ng, "isolated counts")) != MM_OK)
        return rc;
    UP_TRY(hipMemsetAsync(iso, 0, sizeof(int32_t) * (size_t)ng, s), "isolated counts");
    const int with_dist = K < (1 << 23);
    hipLaunchKernelGGL(k_u_order_keys, dim3(grid_for(n)), dim3(UB), 0, s, label, n, rank_of_seed, with_dist, u->voff,
                       ng, keys, iso);
    UP_TRY(hipGetLastError(), "k_u_order_keys");
    if ((rc = sort_keys64(u, keys, sorted, n, 64)) != MM_OK) return rc;
    std::vector<int32_t> h_iso((size_t)ng);
    UP_TRY(copy_on(s, h_iso.data(), iso, sizeof(int32_t) * (size_t)ng, hipMemcpyDeviceToHost), "isolated counts");
    int64_t iso_total = 0;
    for (int32_t x : h_iso) iso_total += x;
    std::vector<int32_t> seg((size_t)ng * 5);
    int64_t iso_before = 0;
    for (int32_t k = 0; k < ng; ++k) {
        const int32_t lo = u->voff ? u->h_voff[(size_t)k] : 0;
        const int32_t len = u->voff ? u->h_voff[(size_t)k + 1] - lo : n;
        seg[(size_t)k] = lo;
        seg[(size_t)ng + k] = len > 0 ? len : 1;
        seg[2 * (size_t)ng + k] = h_iso[(size_t)k];
        seg[3 * (size_t)ng + k] = (int32_t)((n - iso_total) + iso_before);
        seg[4 * (size_t)ng + k] = (int32_t)(lo - iso_before);
        iso_before += h_iso[(size_t)k];
    }
    int32_t *d_seg = nullptr;
    if ((rc = tmp_alloc(u, &d_seg, seg.size(), "graph segments")) != MM_OK) return rc;
    UP_TRY(copy_on(s, d_seg, seg.data(), sizeof(int32_t) * seg.size(), hipMemcpyHostToDevice), "graph segments");
    Segs sg;
    sg.lo = d_seg;
    sg.len = d_seg + ng;
    sg.k = d_seg + 2 * (size_t)ng;
    sg.iso_at = d_seg + 3 * (size_t)ng;
    sg.conn_at = d_seg + 4 * (size_t)ng;
    sg.count = ng;
    // (the same rule as the matcher's settling of isolated vertices: at least 1% of the vertices)
    sg.tail = iso_total > 0 && iso_total * 100 >= (int64_t)n ? 1 : 0;
    if (sg.tail && !u->voff) u->g->iso_lo = (int32_t)(n - iso_total);
    hipLaunchKernelGGL(k_u_merge, dim3(grid_for(n)), dim3(UB), 0, s, sorted, sg, n, order, inv);
    UP_TRY(hipGetLastError(), "k_u_merge");
    UP_TRY(hipStreamSynchronize(s), "locality order");
    mm_upload_mark("order: sort and merge", s);
    tmp_free(u, d_seg);
    tmp_free(u, keys);
    tmp_free(u, sorted);
    tmp_free(u, iso);
    tmp_free(u, rank_of_seed);
    tmp_free(u, d_rank);
    tmp_free(u, label);
    tmp_free(u, fa);
    tmp_free(u, ctr);
    tmp_free(u, pcount);
    return MM_OK;
}

}  // namespace

namespace {

// The symmetry test (when every row ascends): push_ok_struct / push_ok_weights,
// and each entry's reverse when the graph can take 32-bit keys.  Needs the
// weights, so it joins their copy.
int32_t symmetry_test(UploadState *u)
{
    mm_graph *g = u->g;
    hipStream_t s = g->stream;
    const size_t nn = (size_t)g->n;
    int32_t rc;
    UP_TRY(u->join_weights(), "Unable to transfer graph weights to device");
    mm_upload_mark("weights copy (overlapped)", nullptr);
    g->push_ok_struct = 0;
    g->push_ok_weights = 0;
    if (!u->strict || !nn) return MM_OK;
    unsigned long long *counts = nullptr;
    int32_t *flags = nullptr;
    if ((rc = tmp_alloc(u, &counts, 2, "symmetry counts")) != MM_OK || (rc = tmp_alloc(u, &flags, 2, "symmetry flags")) != MM_OK)
        return rc;
    if (u->want_rev && g->nnz) {
        if ((rc = tmp_alloc(u, &u->rev, (size_t)g->nnz, "reverse positions")) != MM_OK) return rc;
        UP_TRY(hipMemsetAsync(u->rev, 0xFF, sizeof(int32_t) * (size_t)g->nnz, s), "reverse positions");  // -1: none
    }
    UP_TRY(hipMemsetAsync(counts, 0, 2 * sizeof(unsigned long long), s), "symmetry counts");
    UP_TRY(hipMemsetAsync(flags, 0, 2 * sizeof(int32_t), s), "symmetry flags");
    hipLaunchKernelGGL(k_u_sym, dim3(grid_red(u, (int64_t)nn * 8)), dim3(UB), 0, s, u->rg, u->nb, u->w, g->n, u->rev,
                       counts, flags);
    if (u->nlong > 0)
        hipLaunchKernelGGL(k_u_sym_long, dim3(std::min(u->nlong, 4096)), dim3(UB), 0, s, u->rg, u->nb, u->w, u->longs,
                           u->nlong, u->rev, counts, flags);
    UP_TRY(hipGetLastError(), "k_u_sym");
    unsigned long long h_counts[2];
    int32_t h_flags[2];
    UP_TRY(hipMemcpyAsync(h_counts, counts, sizeof(h_counts), hipMemcpyDeviceToHost, s), "symmetry counts");
    UP_TRY(copy_on(s, h_flags, flags, sizeof(h_flags), hipMemcpyDeviceToHost), "symmetry flags");
    const bool sym = !h_flags[0] && h_counts[0] == h_counts[1];
    g->push_ok_struct = sym;
    g->push_ok_weights = sym && !h_flags[1];
    tmp_free(u, counts);
    tmp_free(u, flags);
    if (!g->push_ok_weights && u->rev) {
        tmp_free(u, u->rev);
        u->rev = nullptr;
    }
    mm_upload_mark("symmetry test", s);
    return MM_OK;
}

}  // namespace

int32_t mm_upload_begin(mm_graph *g, const int32_t *ranges, const int32_t *nbrs, const float *w, const int32_t *voff,
                        const int64_t *eoff, int32_t n_graphs, int relabel, UploadState **out)
{
    *out = nullptr;
    UploadState *u = new (std::nothrow) UploadState();
    if (!u) return mm::set_error(MM_ERR_NOMEM, "mm_graph_create: host allocation failed");
    u->g = g;
    u->n = g->n;
    u->nnz = g->nnz;
    struct Guard {
        UploadState *&p;
        ~Guard() { delete p; }
    } guard{u};
    hipStream_t s = g->stream;
    const size_t nn = (size_t)g->n;
    const int32_t ng = voff && n_graphs > 0 ? n_graphs : 1;
    int32_t rc;
    if ((rc = tmp_alloc(u, &u->rg, nn, "caller ranges")) != MM_OK || (rc = tmp_alloc(u, &u->nb, (size_t)g->nnz, "caller rows")) != MM_OK)
        return rc;
    if (w && (rc = tmp_alloc(u, &u->w, (size_t)g->nnz, "caller weights")) != MM_OK) return rc;
    if (nn) UP_TRY(hipMemcpyAsync(u->rg, ranges, sizeof(int2) * nn, hipMemcpyHostToDevice, s), "Unable to transfer graph data to device");
    if (g->nnz)
        UP_TRY(hipMemcpyAsync(u->nb, nbrs, sizeof(int32_t) * (size_t)g->nnz, hipMemcpyHostToDevice, s),
               "Unable to transfer graph data to device");
    if (w && g->nnz) {
        // The weights are needed first by the symmetry test: their copy runs
        // beside the checks and the locality order, on a host thread and
        // stream of its own, after the rows' copy (two host copies at once
        // share the bus and take longer than one after the other).
        UP_TRY(hipStreamSynchronize(s), "Unable to transfer graph data to device");
        UP_TRY(hipStreamCreateWithFlags(&u->wstream, hipStreamNonBlocking), "hipStreamCreate");
        const int dev = g->device;
        const size_t bytes = sizeof(float) * (size_t)g->nnz;
        UploadState *up = u;
        u->wthread = std::thread([up, dev, w, bytes] {
            hipError_t e = hipSetDevice(dev);
            if (e == hipSuccess) e = hipMemcpyAsync(up->w, w, bytes, hipMemcpyHostToDevice, up->wstream);
            if (e == hipSuccess) e = hipStreamSynchronize(up->wstream);
            up->wresult = e;
        });
    }
    long long *d_eoff = nullptr;
    if (voff && n_graphs > 0) {
        u->h_voff.assign(voff, voff + n_graphs + 1);
        if ((rc = tmp_alloc(u, &u->voff, (size_t)n_graphs + 1, "graph offsets")) != MM_OK ||
            (rc = tmp_alloc(u, &d_eoff, (size_t)n_graphs + 1, "graph offsets")) != MM_OK)
            return rc;
        UP_TRY(hipMemcpyAsync(u->voff, voff, sizeof(int32_t) * ((size_t)n_graphs + 1), hipMemcpyHostToDevice, s),
               "Unable to transfer graph data to device");
        UP_TRY(hipMemcpyAsync(d_eoff, eoff, sizeof(long long) * ((size_t)n_graphs + 1), hipMemcpyHostToDevice, s),
               "Unable to transfer graph data to device");
    }
    mm_upload_mark("copy to device", s);
    // Ranges, ids and ascending rows (a batch's ranges and ids still local).
    int32_t *flags = nullptr;
    unsigned long long *hist = nullptr;
    if ((rc = tmp_alloc(u, &flags, 4, "check flags")) != MM_OK || (rc = tmp_alloc(u, &u->longs, nn, "long rows")) != MM_OK ||
        (rc = tmp_alloc(u, &hist, 33, "degree histogram")) != MM_OK)
        return rc;
    UP_TRY(hipMemsetAsync(flags, 0, 4 * sizeof(int32_t), s), "check flags");
    UP_TRY(hipMemsetAsync(hist, 0, 33 * sizeof(unsigned long long), s), "degree histogram");
    bool sorted = true;
    if (nn) {
        hipLaunchKernelGGL(k_u_check, dim3(grid_red(u, (int64_t)nn * 8)), dim3(UB), 0, s, u->rg, u->nb, g->n, g->nnz, u->voff,
                           d_eoff, ng, flags, u->longs, hist);
        UP_TRY(hipGetLastError(), "k_u_check");
        int32_t h_flags[4];
        unsigned long long h_hist[33];
        UP_TRY(hipMemcpyAsync(h_hist, hist, sizeof(h_hist), hipMemcpyDeviceToHost, s), "degree histogram");
        UP_TRY(copy_on(s, h_flags, flags, sizeof(h_flags), hipMemcpyDeviceToHost), "check flags");
        u->nlong = h_flags[2];
        g->max_degree = h_flags[3];
        for (int k = 0; k < 33; ++k) g->deg_hist[k] = (int64_t)h_hist[k];
        if (u->nlong > 0 && !h_flags[0]) {
            hipLaunchKernelGGL(k_u_check_long, dim3(std::min(u->nlong, 4096)), dim3(UB), 0, s, u->rg, u->nb, g->n, g->nnz,
                               u->voff, d_eoff, ng, flags, u->longs, u->nlong);
            UP_TRY(hipGetLastError(), "k_u_check_long");
            UP_TRY(copy_on(s, h_flags, flags, sizeof(h_flags), hipMemcpyDeviceToHost), "check flags");
        }
        if (h_flags[0]) return mm::set_error(MM_ERR_INVALID, "malformed CSR: range or neighbour id out of bounds");
        sorted = !h_flags[1];
        u->strict = sorted;
    }
    g->mean_degree = g->n > 0 ? (double)g->nnz / (double)g->n : 0.0;
    tmp_free(u, hist);
    if (d_eoff) {
        // one CSR: row ranges shift by their graph's first entry, ids by its first vertex
        if (nn)
            hipLaunchKernelGGL(k_u_rebase_rows, dim3(grid_for((int64_t)nn)), dim3(UB), 0, s, u->rg, g->n, u->voff, d_eoff,
                               n_graphs);
        if (g->nnz)
            hipLaunchKernelGGL(k_u_rebase_ids, dim3(grid_for(g->nnz)), dim3(UB), 0, s, u->nb, g->nnz, u->voff, d_eoff,
                               n_graphs);
        UP_TRY(hipGetLastError(), "batch rebase");
        UP_TRY(hipStreamSynchronize(s), "batch rebase");
        tmp_free(u, d_eoff);
    }
    // 32-bit proposal keys hold (position << L) | local id, L = bits of the
    // largest local id: possible only while the largest degree leaves room
    int32_t max_loc = g->n > 0 ? g->n - 1 : 0;
    if (voff && n_graphs > 0) {
        max_loc = 0;
        for (int32_t k = 0; k < n_graphs; ++k) max_loc = std::max(max_loc, voff[k + 1] - voff[k] - 1);
    }
    int locbits = 1;
    while (locbits < 31 && ((int64_t)1 << locbits) <= (int64_t)max_loc) ++locbits;
    const bool want_rev = w && locbits < 31 && (int64_t)g->max_degree < ((int64_t)1 << (32 - locbits));
    mm_upload_mark("id checks", s);
    // The internal order (reads rows and ids only).
    if ((rc = tmp_alloc(u, &u->order, nn, "vertex order")) != MM_OK || (rc = tmp_alloc(u, &u->inv, nn, "vertex order")) != MM_OK)
        return rc;
    g->relabelled = relabel != 0 && g->n > 1;
    if (g->relabelled) {
        if ((rc = locality_order(u, u->order, u->inv)) != MM_OK) return rc;
    } else if (nn) {
        hipLaunchKernelGGL(k_u_identity, dim3(grid_for((int64_t)nn)), dim3(UB), 0, s, g->n, u->order, u->inv);
        UP_TRY(hipGetLastError(), "k_u_identity");
    }
    mm_upload_mark("order", s);
    u->want_rev = want_rev;
    tmp_free(u, flags);
    *out = u;
    u = nullptr;  // (the guard deletes nothing now)
    return MM_OK;
}

void mm_upload_free(UploadState *u) { delete u; }

void mm_upload_mark(const char *phase, hipStream_t s)
{
    static const bool on = std::getenv("MM_UPLOAD_TIMING") != nullptr;
    if (!on) return;
    static thread_local std::chrono::steady_clock::time_point last;
    if (s) (void)hipStreamSynchronize(s);
    const auto now = std::chrono::steady_clock::now();
    if (phase) std::fprintf(stderr, "[mm upload] %-28s %9.2f ms\n", phase,
                            std::chrono::duration<double, std::milli>(now - last).count());
    last = now;
}

int32_t mm_upload_finish(mm_graph *g, UploadState *u, bool align_rows, bool heads)
{
    hipStream_t s = g->stream;
    const int32_t n = g->n;
    const size_t nn = (size_t)n;
    const int32_t ng = u->voff ? (int32_t)u->h_voff.size() - 1 : 1;
    int32_t rc;
    int32_t *order = u->order, *inv = u->inv;
    // Row offsets in the internal order.
    if ((rc = mm_graph_alloc(g, (void **)&g->d_ranges, sizeof(int2) * nn, "neighbour ranges")) != MM_OK) return rc;
    int64_t nnz_dev = 0;
    if (n > 0) {
        const int64_t threads = (n + ROWS_PER_THREAD - 1) / ROWS_PER_THREAD;
        const int32_t blocks = (int32_t)((threads + UB - 1) / UB);
        OffMap *tmaps = nullptr, *bmaps = nullptr;
        int2 *bentry = nullptr;
        long long *total = nullptr;
        if ((rc = tmp_alloc(u, &tmaps, (size_t)blocks * UB, "offset scan")) != MM_OK ||
            (rc = tmp_alloc(u, &bmaps, (size_t)blocks, "offset scan")) != MM_OK ||
            (rc = tmp_alloc(u, &bentry, (size_t)blocks * 2, "offset scan")) != MM_OK ||
            (rc = tmp_alloc(u, &total, 1, "offset scan")) != MM_OK)
            return rc;
        for (int align = align_rows ? 1 : 0; align >= 0; --align) {
            hipLaunchKernelGGL(k_u_off_maps, dim3(blocks), dim3(UB), 0, s, u->rg, order, n, align, tmaps, bmaps);
            hipLaunchKernelGGL(k_u_off_chain, dim3(1), dim3(64), 0, s, bmaps, blocks, bentry, total);
            UP_TRY(hipGetLastError(), "offset scan");
            long long h_total = 0;
            UP_TRY(hipMemcpyAsync(&h_total, total, sizeof(h_total), hipMemcpyDeviceToHost, s), "offset scan");
            UP_TRY(hipStreamSynchronize(s), "offset scan");
            if (h_total > INT32_MAX && align) continue;  // padded rows would overflow int32: pack them
            if (h_total > INT32_MAX) return mm::set_error(MM_ERR_INVALID, "mm_graph_create: too many row entries");
            hipLaunchKernelGGL(k_u_off_walk, dim3(blocks), dim3(UB), 0, s, u->rg, order, n, align, tmaps, bentry, blocks,
                               g->d_ranges);
            UP_TRY(hipGetLastError(), "k_u_off_walk");
            nnz_dev = h_total;
            break;
        }
        tmp_free(u, tmaps);
        tmp_free(u, bmaps);
        tmp_free(u, bentry);
        tmp_free(u, total);
    }
    mm_upload_mark("row offsets", s);
    // Per-vertex ids.
    if ((rc = mm_graph_alloc(g, (void **)&g->d_loc, sizeof(int32_t) * nn, "original ids")) != MM_OK) return rc;
    if (u->voff && (rc = mm_graph_alloc(g, (void **)&g->d_vbase, sizeof(int32_t) * nn, "graph bases")) != MM_OK)
        return rc;
    if (n > 0) {
        hipLaunchKernelGGL(k_u_loc, dim3(grid_for(n)), dim3(UB), 0, s, order, n, u->voff, ng, g->d_loc, g->d_vbase);
        UP_TRY(hipGetLastError(), "k_u_loc");
    }
    // Rows.
    if ((rc = mm_graph_alloc(g, (void **)&g->d_nbrs, sizeof(int32_t) * (size_t)std::max<int64_t>(nnz_dev, 1),
                             "neighbours")) != MM_OK)
        return rc;
    if (n > 0) {
        hipLaunchKernelGGL(k_u_nbrs, dim3(grid_for((int64_t)n * 8)), dim3(UB), 0, s, u->rg, u->nb, order, inv,
                           g->d_ranges, n, g->d_nbrs);
        UP_TRY(hipGetLastError(), "k_u_nbrs");
    }
    if (g->weighted && !u->strict && n > 0 && nnz_dev > 0) {
        // A row may hold one id twice with different weights: the result
        // check's getWeight reads the weight at the CSR position it stops at.
        if ((rc = mm_graph_alloc(g, (void **)&g->d_wcsr, sizeof(float) * (size_t)nnz_dev, "CSR-order weights")) != MM_OK)
            return rc;
        hipLaunchKernelGGL(k_u_wcsr, dim3(grid_for((int64_t)n * 8)), dim3(UB), 0, s, u->rg, u->w, order, g->d_ranges, n,
                           g->d_wcsr);
        UP_TRY(hipGetLastError(), "k_u_wcsr");
    }
    mm_upload_mark("ids and rows", s);
    if ((rc = symmetry_test(u)) != MM_OK) return rc;
    int32_t max_loc = n > 0 ? n - 1 : 0;
    if (u->voff) {
        max_loc = 0;
        for (int32_t k = 0; k < ng; ++k) max_loc = std::max(max_loc, u->h_voff[(size_t)k + 1] - u->h_voff[(size_t)k] - 1);
    }
    int locbits = 1;
    while (locbits < 31 && ((int64_t)1 << locbits) <= (int64_t)max_loc) ++locbits;
    const bool key32 = g->weighted && g->push_ok_weights && n > 0 && nnz_dev > 0 && locbits < 31 &&
                       (int64_t)g->max_degree < ((int64_t)1 << (32 - locbits)) && u->rev;
    int32_t *spos = nullptr, *sperm = nullptr;
    if (g->weighted) {
        if ((rc = mm_graph_alloc(g, (void **)&g->d_adj, sizeof(int2) * (size_t)std::max<int64_t>(nnz_dev, 1),
                                 "sorted weighted rows")) != MM_OK)
            return rc;
        if (key32 && ((rc = tmp_alloc(u, &spos, (size_t)nnz_dev, "sorted positions")) != MM_OK ||
                      (rc = tmp_alloc(u, &sperm, (size_t)nnz_dev, "sorted positions")) != MM_OK))
            return rc;
        int32_t *lists = nullptr, *counts = nullptr;
        if ((rc = tmp_alloc(u, &lists, 3 * nn, "row classes")) != MM_OK || (rc = tmp_alloc(u, &counts, 3, "row classes")) != MM_OK)
            return rc;
        UP_TRY(hipMemsetAsync(counts, 0, 3 * sizeof(int32_t), s), "row classes");
        hipLaunchKernelGGL(k_u_classify, dim3(grid_red(u, n)), dim3(UB), 0, s, g->d_ranges, n, lists, n, counts);
        UP_TRY(hipGetLastError(), "k_u_classify");
        int32_t h_counts[3];
        UP_TRY(hipMemcpyAsync(h_counts, counts, sizeof(h_counts), hipMemcpyDeviceToHost, s), "row classes");
        UP_TRY(hipStreamSynchronize(s), "row classes");
        if (h_counts[0] > 0)
            hipLaunchKernelGGL(k_u_sort_small<16>, dim3(grid_for((int64_t)h_counts[0] * 16)), dim3(UB), 0, s, lists,
                               h_counts[0], u->rg, u->w, order, g->d_ranges, g->d_nbrs, g->d_adj, spos, sperm);
        if (h_counts[1] > 0)
            hipLaunchKernelGGL(k_u_sort_small<64>, dim3(grid_for((int64_t)h_counts[1] * 64)), dim3(UB), 0, s, lists + nn,
                               h_counts[1], u->rg, u->w, order, g->d_ranges, g->d_nbrs, g->d_adj, spos, sperm);
        UP_TRY(hipGetLastError(), "k_u_sort_small");
        const int32_t nlong = h_counts[2];
        if (nlong > 0) {
            // slots of the long rows, in list order, from their lengths
            std::vector<long long> h_start((size_t)nlong + 1, 0);
            std::vector<int2> h_rr((size_t)nlong);
            int2 *d_rr = nullptr;
            if ((rc = tmp_alloc(u, &d_rr, (size_t)nlong, "long rows")) != MM_OK) return rc;
            hipLaunchKernelGGL(k_u_gather_ranges, dim3(grid_for(nlong)), dim3(UB), 0, s, lists + 2 * nn, nlong, g->d_ranges,
                               d_rr);
            UP_TRY(hipGetLastError(), "k_u_gather_ranges");
            UP_TRY(copy_on(s, h_rr.data(), d_rr, sizeof(int2) * (size_t)nlong, hipMemcpyDeviceToHost), "long rows");
            for (int32_t r = 0; r < nlong; ++r) h_start[(size_t)r + 1] = h_start[(size_t)r] + (h_rr[(size_t)r].y - h_rr[(size_t)r].x);
            const long long total = h_start[(size_t)nlong];
            long long *d_start = nullptr;
            unsigned long long *ka = nullptr, *kb = nullptr;
            int32_t *va = nullptr, *vb = nullptr;
            if ((rc = tmp_alloc(u, &d_start, (size_t)nlong + 1, "long rows")) != MM_OK ||
                (rc = tmp_alloc(u, &ka, (size_t)total, "long-row keys")) != MM_OK ||
                (rc = tmp_alloc(u, &kb, (size_t)total, "long-row keys")) != MM_OK ||
                (rc = tmp_alloc(u, &va, (size_t)total, "long-row keys")) != MM_OK ||
                (rc = tmp_alloc(u, &vb, (size_t)total, "long-row keys")) != MM_OK)
                return rc;
            UP_TRY(copy_on(s, d_start, h_start.data(), sizeof(long long) * h_start.size(), hipMemcpyHostToDevice),
                   "long rows");
            const int lgrid = std::min(nlong, 65536);
            hipLaunchKernelGGL(k_u_long_keys, dim3(lgrid), dim3(UB), 0, s, lists + 2 * nn, nlong, d_start, u->rg, u->w,
                               order, ka, va);
            UP_TRY(hipGetLastError(), "k_u_long_keys");
            if (total > INT32_MAX) return mm::set_error(MM_ERR_INVALID, "mm_graph_create: long rows too large to sort");
            if ((rc = sort_pairs(u, ka, kb, va, vb, total, 32 + bits_for((uint64_t)nlong))) != MM_OK) return rc;
            hipLaunchKernelGGL(k_u_long_place, dim3(lgrid), dim3(UB), 0, s, lists + 2 * nn, nlong, d_start, vb, u->rg,
                               u->w, order, g->d_ranges, g->d_nbrs, g->d_adj, spos, sperm);
            UP_TRY(hipGetLastError(), "k_u_long_place");
            UP_TRY(hipStreamSynchronize(s), "long rows");
            tmp_free(u, ka);
            tmp_free(u, kb);
            tmp_free(u, va);
            tmp_free(u, vb);
            tmp_free(u, d_start);
            tmp_free(u, d_rr);
        }
        tmp_free(u, lists);
        tmp_free(u, counts);
    }
    mm_upload_mark("sorted rows", s);
    if (key32) {
        if ((rc = mm_graph_alloc(g, (void **)&g->d_adjp, sizeof(int2) * (size_t)nnz_dev,
                                 "sorted rows with key positions")) != MM_OK)
            return rc;
        // (gaps between padded rows are never read; copy them too for a defined array)
        UP_TRY(hipMemcpyAsync(g->d_adjp, g->d_adj, sizeof(int2) * (size_t)nnz_dev, hipMemcpyDeviceToDevice, s),
               "key positions");
        hipLaunchKernelGGL(k_u_adjp, dim3(grid_for((int64_t)n * 8)), dim3(UB), 0, s, u->rg, order, g->d_ranges, n,
                           g->d_adj, spos, sperm, u->rev, g->d_adjp);
        UP_TRY(hipGetLastError(), "k_u_adjp");
        g->locbits = locbits;
    }
    mm_upload_mark("key positions", s);
    if (g->weighted && heads && n > 0) {
        const int2 *scan = key32 ? g->d_adjp : g->d_adj;
        int pb = 1, ib = 1;
        while (pb < 31 && ((int64_t)1 << pb) < (int64_t)g->max_degree) ++pb;
        while (ib < 31 && ((int64_t)1 << ib) < (int64_t)n) ++ib;
        if (key32 && ib + pb <= 31) {
            if ((rc = mm_graph_alloc(g, (void **)&g->d_headc, sizeof(uint32_t) * 4 * nn, "row heads")) != MM_OK) return rc;
            hipLaunchKernelGGL(k_u_heads, dim3(grid_for(n)), dim3(UB), 0, s, g->d_ranges, n, scan, nullptr, g->d_headc, pb);
            g->head_pb = pb;
            // (at least one entry: scans of an empty row read entry 0)
            const int64_t np = std::max<int64_t>(1, (int64_t)nnz_dev);
            if ((rc = mm_graph_alloc(g, (void **)&g->d_adjc, sizeof(uint32_t) * (size_t)np, "packed rows")) != MM_OK)
                return rc;
            if (nnz_dev == 0) UP_TRY(hipMemsetAsync(g->d_adjc, 0, sizeof(uint32_t), s), "packed rows");
            hipLaunchKernelGGL(k_u_pack, dim3(grid_for((int64_t)nnz_dev)), dim3(UB), 0, s, scan, (int64_t)nnz_dev,
                               g->d_adjc, pb);
        } else {
            if ((rc = mm_graph_alloc(g, (void **)&g->d_head, sizeof(int2) * 4 * nn, "row heads")) != MM_OK) return rc;
            hipLaunchKernelGGL(k_u_heads, dim3(grid_for(n)), dim3(UB), 0, s, g->d_ranges, n, scan, g->d_head, nullptr, 0);
        }
        UP_TRY(hipGetLastError(), "k_u_heads");
    }
    if (g->relabelled) {
        // (caller position, internal id) pairs grouped by the eighth of the
        // internal id, in caller order inside a group.
        uint32_t *ka = nullptr, *kb = nullptr;
        int32_t *va = nullptr, *vb = nullptr, *cnt = nullptr;
        if ((rc = tmp_alloc(u, &ka, nn, "vertex order")) != MM_OK || (rc = tmp_alloc(u, &kb, nn, "vertex order")) != MM_OK ||
            (rc = tmp_alloc(u, &va, nn, "vertex order")) != MM_OK || (rc = tmp_alloc(u, &vb, nn, "vertex order")) != MM_OK ||
            (rc = tmp_alloc(u, &cnt, 8, "vertex order")) != MM_OK)
            return rc;
        UP_TRY(hipMemsetAsync(cnt, 0, 8 * sizeof(int32_t), s), "vertex order");
        hipLaunchKernelGGL(k_u_eighths, dim3(grid_red(u, n)), dim3(UB), 0, s, inv, n, ka, va, cnt);
        UP_TRY(hipGetLastError(), "k_u_eighths");
        if ((rc = sort_pairs(u, ka, kb, va, vb, n, 3)) != MM_OK) return rc;
        if ((rc = mm_graph_alloc(g, (void **)&g->d_unperm, sizeof(int2) * nn, "vertex order")) != MM_OK ||
            (rc = mm_graph_alloc(g, (void **)&g->d_match_out, sizeof(int32_t) * nn, "match")) != MM_OK)
            return rc;
        hipLaunchKernelGGL(k_u_pairs_out, dim3(grid_for(n)), dim3(UB), 0, s, vb, inv, n, g->d_unperm);
        UP_TRY(hipGetLastError(), "k_u_pairs_out");
        int32_t h_cnt[8];
        UP_TRY(hipMemcpyAsync(h_cnt, cnt, sizeof(h_cnt), hipMemcpyDeviceToHost, s), "vertex order");
        UP_TRY(hipStreamSynchronize(s), "vertex order");
        g->unperm_off[0] = 0;
        for (int x = 0; x < 8; ++x) g->unperm_off[x + 1] = g->unperm_off[x] + h_cnt[x];
    }
    UP_TRY(hipStreamSynchronize(s), "Unable to prepare the graph on the device");
    mm_upload_mark("heads and caller pairs", s);
    return MM_OK;
}
