// Device-side result metrics: getWeight (src/matchcpu.cu:64-97) and
// testMatching (src/matchcpu.cu:127-246) of a handle's current result, for
// results too large to check on the host quickly (a 4096-graph batch holds
// ~46 M vertices).
//
// testMatching sorts the vertices by match value and inspects each group of
// equal values: a single member is "matched to itself", members beyond two
// are "matched to more than one vertex", and a first pair that are not each
// other's neighbours is "not matched to a neighbour".  Here the groups are
// found without sorting: the key of a matched vertex is its graph's first
// global id plus (value - 4), a per-key counter gives the group size and two
// atomic minima give the group's two smallest members, the pair that is
// checked.  The same pass counts the unmatched vertices with an unmatched
// neighbour and sums getWeight's size and weight (the first neighbour in CSR
// order that shares the value).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/mm_match.h"
#include "mm_graph_impl.h"
#include "mm_internal.h"

namespace {

constexpr int CBLOCK = 256;
constexpr int NACC = 6;  // under, multi, invalid, matchable, size, weight (double bits)

struct CDev {
    int32_t n;
    const int2 *ranges;
    const int32_t *nbrs;  // CSR order
    const int2 *adj;      // weight-sorted rows (weighted graphs) or null
    const int32_t *match;
    const int32_t *vbase;  // batch: first global id of the vertex's graph, else null
    const int32_t *voff;   // batch: graph vertex offsets (global ids)
    int32_t n_graphs;
    int32_t *count, *first, *second;
    unsigned long long *part;  // [blocks][NACC]
    double *gweight;           // batch: per graph
    unsigned long long *gsize;
    int32_t *gbad;
};

// Group key of a matched vertex, or -1 when the value lies outside its graph.
__device__ __forceinline__ int64_t group_key(const CDev &d, int32_t v, int32_t m)
{
    const int64_t base = d.vbase ? d.vbase[v] : 0;
    const int64_t key = base + (int64_t)m - 4;
    return key < d.n ? key : -1;
}

__device__ __forceinline__ bool has_neighbour(const CDev &d, int32_t v, int32_t u)
{
    const int2 r = d.ranges[v];
    for (int32_t j = r.x; j < r.y; ++j)
        if (d.nbrs[j] == u) return true;
    return false;
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

__device__ __forceinline__ int32_t graph_of(const CDev &d, int32_t v)
{
    const int32_t base = d.vbase[v];
    int32_t lo = 0, hi = d.n_graphs - 1;
    while (lo < hi) {  // the last graph whose first id is base (empty graphs share it)
        const int32_t mid = (lo + hi + 1) >> 1;
        if (d.voff[mid] <= base) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void k_c_init(CDev d)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.n; i += (int64_t)gridDim.x * blockDim.x) {
        d.count[i] = 0;
        d.first[i] = INT_MAX;
        d.second[i] = INT_MAX;
    }
}

__global__ void k_c_group(CDev d)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < d.n; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t m = d.match[v];
        if (m < 4) continue;
        const int64_t key = group_key(d, (int32_t)v, m);
        if (key < 0) continue;
        atomicAdd(d.count + key, 1);
        atomicMin(d.first + key, (int32_t)v);
    }
}

__global__ void k_c_second(CDev d)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < d.n; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t m = d.match[v];
        if (m < 4) continue;
        const int64_t key = group_key(d, (int32_t)v, m);
        if (key >= 0 && d.first[key] != (int32_t)v) atomicMin(d.second + key, (int32_t)v);
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

__global__ __launch_bounds__(CBLOCK) void k_c_eval(CDev d)
{
    unsigned long long under = 0, multi = 0, invalid = 0, matchable = 0, size = 0;
    double weight = 0.0;
    for (int64_t vv = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; vv < d.n; vv += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = (int32_t)vv;
        const int32_t m = d.match[v];
        const int2 r = d.ranges[v];
        if (m < 4) {
            for (int32_t j = r.x; j < r.y; ++j) {
                const int32_t u = d.nbrs[j];
                if (u != v && d.match[u] < 4) {
                    ++matchable;
                    break;
                }
            }
            continue;
        }
        int32_t bad = 0;
        const int64_t key = group_key(d, v, m);
        if (key < 0) {
            ++under;
            bad = 1;
        } else if (d.first[key] == v) {
            const int32_t s = d.count[key];
            if (s == 1) {
                ++under;
                bad = 1;
            } else {
                if (s > 2) {
                    multi += (unsigned long long)(s - 2);
                    bad = 1;
                }
                const int32_t u = d.second[key];
                if (!has_neighbour(d, v, u) || !has_neighbour(d, u, v)) {
                    ++invalid;
                    bad = 1;
                }
            }
        }
        // getWeight: the first neighbour in CSR order with the same value
        float w = 0.0f;
        bool found = false;
        for (int32_t j = r.x; j < r.y; ++j) {
            const int32_t u = d.nbrs[j];
            if (d.match[u] == m) {
                w = weight_to(d, v, u);
                found = true;
                break;
            }
        }
        if (found) {
            ++size;
            weight += (double)w;
        }
        if (d.vbase) {
            const int32_t gi = graph_of(d, v);
            if (found) {
                atomicAdd(d.gsize + gi, 1ull);
                atomicAdd(d.gweight + gi, 0.5 * (double)w);
            }
            if (bad) atomicMax(d.gbad + gi, 1);
        }
    }
    const unsigned long long acc[5] = {under, multi, invalid, matchable, size};
    for (int k = 0; k < 5; ++k) {
        const unsigned long long s = block_sum_u64(acc[k]);
        if (threadIdx.x == 0) d.part[(size_t)blockIdx.x * NACC + k] = s;
    }
    const double ws = block_sum_f64(weight);
    if (threadIdx.x == 0) {
        unsigned long long bits;
        std::memcpy(&bits, &ws, sizeof(bits));
        d.part[(size_t)blockIdx.x * NACC + 5] = bits;
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

__global__ void k_c_zero(double *gw, unsigned long long *gs, int32_t *gb, int32_t count)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        gw[i] = 0.0;
        gs[i] = 0ull;
        gb[i] = 0;
    }
}

#define CK_TRY(expr, what)                                                                         \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return mm::set_errorf(e_ == hipErrorOutOfMemory ? MM_ERR_NOMEM : MM_ERR_HIP, "%s: %s", \
                                  what, hipGetErrorString(e_));                                    \
    } while (0)

int32_t run_check(mm_graph *g, const int32_t *voff, int32_t n_graphs, mm_check *out, double *gw_out,
                  int64_t *gs_out, int32_t *gv_out)
{
    if (!out) return mm::set_error(MM_ERR_INVALID, "mm_graph_check: out is NULL");
    std::memset(out, 0, sizeof(*out));
    out->valid = 1;
    if (g->n == 0) return MM_OK;
    CK_TRY(hipSetDevice(g->device), "hipSetDevice");
    const int grid = (int)std::min<int64_t>(((int64_t)g->n + CBLOCK - 1) / CBLOCK, (int64_t)g->num_cu * 8);
    int32_t rc;
    void *p = nullptr;
    if (!g->d_chk) {
        if ((rc = mm_graph_alloc(g, &p, sizeof(int32_t) * 3 * (size_t)g->n, "check scratch")) != MM_OK) return rc;
        g->d_chk = (int32_t *)p;
    }
    if (!g->d_cpart) {
        if ((rc = mm_graph_alloc(g, &p, sizeof(unsigned long long) * NACC * (size_t)g->num_cu * 8, "check sums")) != MM_OK)
            return rc;
        g->d_cpart = (unsigned long long *)p;
    }
    struct Tmp {
        std::vector<void *> ptrs;
        ~Tmp()
        {
            for (void *q : ptrs) (void)hipFree(q);
        }
    } tmp;
    if (g->internal_stale) {
        // The last plain run wrote only the caller-order result: bring the
        // internal-order array the check reads up to date.
        const int pgrid = (int)std::min<int64_t>(((int64_t)g->n + 255) / 256, (int64_t)g->num_cu * 8);
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
    d.match = g->d_match;
    d.count = g->d_chk;
    d.first = g->d_chk + g->n;
    d.second = g->d_chk + 2 * (size_t)g->n;
    d.part = g->d_cpart;
    if (voff && n_graphs > 0) {
        d.vbase = g->d_vbase;
        d.voff = g->d_voff;
        d.n_graphs = n_graphs;
        for (size_t bytes : {sizeof(double), sizeof(unsigned long long), sizeof(int32_t)}) {
            void *q = nullptr;
            CK_TRY(hipMalloc(&q, bytes * (size_t)n_graphs), "Not enough memory on device for per-graph sums");
            tmp.ptrs.push_back(q);
        }
        d.gweight = (double *)tmp.ptrs[0];
        d.gsize = (unsigned long long *)tmp.ptrs[1];
        d.gbad = (int32_t *)tmp.ptrs[2];
        hipLaunchKernelGGL(k_c_zero, dim3(std::max(1, std::min(1024, (n_graphs + 255) / 256))), dim3(256), 0, g->stream,
                           d.gweight, d.gsize, d.gbad, n_graphs);
    }
    hipLaunchKernelGGL(k_c_init, dim3(grid), dim3(CBLOCK), 0, g->stream, d);
    hipLaunchKernelGGL(k_c_group, dim3(grid), dim3(CBLOCK), 0, g->stream, d);
    hipLaunchKernelGGL(k_c_second, dim3(grid), dim3(CBLOCK), 0, g->stream, d);
    hipLaunchKernelGGL(k_c_eval, dim3(grid), dim3(CBLOCK), 0, g->stream, d);
    CK_TRY(hipGetLastError(), "result check kernels");
    std::vector<unsigned long long> part((size_t)grid * NACC);
    CK_TRY(hipMemcpyAsync(part.data(), d.part, sizeof(unsigned long long) * part.size(), hipMemcpyDeviceToHost,
                          g->stream),
           "Unable to read check sums");
    std::vector<double> gw;
    std::vector<unsigned long long> gs;
    std::vector<int32_t> gb;
    if (d.vbase) {
        gw.resize((size_t)n_graphs);
        gs.resize((size_t)n_graphs);
        gb.resize((size_t)n_graphs);
        CK_TRY(hipMemcpyAsync(gw.data(), d.gweight, sizeof(double) * gw.size(), hipMemcpyDeviceToHost, g->stream),
               "Unable to read check sums");
        CK_TRY(hipMemcpyAsync(gs.data(), d.gsize, sizeof(unsigned long long) * gs.size(), hipMemcpyDeviceToHost,
                              g->stream),
               "Unable to read check sums");
        CK_TRY(hipMemcpyAsync(gb.data(), d.gbad, sizeof(int32_t) * gb.size(), hipMemcpyDeviceToHost, g->stream),
               "Unable to read check sums");
    }
    CK_TRY(hipStreamSynchronize(g->stream), "result check");
    double weight = 0.0;
    for (int b = 0; b < grid; ++b) {
        const unsigned long long *q = part.data() + (size_t)b * NACC;
        out->under += (int64_t)q[0];
        out->multi += (int64_t)q[1];
        out->invalid += (int64_t)q[2];
        out->matchable += (int64_t)q[3];
        out->size += (int64_t)q[4];
        double w;
        std::memcpy(&w, &q[5], sizeof(w));
        weight += w;
    }
    out->weight = weight / 2.0;
    out->valid = out->under == 0 && out->multi == 0 && out->invalid == 0;
    for (int32_t k = 0; k < n_graphs && d.vbase; ++k) {
        if (gw_out) gw_out[k] = gw[(size_t)k];
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

int32_t mm_batch_check(mm_batch *b, mm_check *total, double *graph_weight, int64_t *graph_size, int32_t *graph_valid)
{
    mm::clear_error();
    if (!b || !b->g) return mm::set_error(MM_ERR_INVALID, "mm_batch_check: batch is NULL");
    const int32_t ng = (int32_t)b->vertex_offsets.size() - 1;
    return run_check(b->g, b->vertex_offsets.data(), ng, total, graph_weight, graph_size, graph_valid);
}

}  // extern "C"
