// MI355X (gfx950) kernels and C ABI of the fine-grained greedy matcher.
//
// One matching round of the reference (src/matchgpu.cu:184-451) has four
// phases: hash-seeded blue/red select, blue->red request, red response and the
// match write-back.  Here a round is two or three launches:
//
//   request   blue vertices scan their rows (G lanes per row, or a whole block
//             for rows longer than the hub degree) and pick the heaviest red
//             neighbour, first in CSR order on ties (gwRequest,
//             src/matchgpu.cu:360-409) or the first red one (grRequest, 286-327).
//             PUSH: the proposal is also posted to the red vertex with a 64-bit
//             atomic max on (weight, -proposer), which replaces the respond scan.
//   respond   (PULL only) red vertices scan their rows for proposers
//             (gwRespond/grRespond, src/matchgpu.cu:329-357, 411-451).
//   match+select  every vertex active in this round resolves its pair
//             (gMatch, src/matchgpu.cu:260-283), and the ones still active are
//             recoloured with the next round's seed (gSelect, 184-218) and
//             appended to the next round's work lists.
//
// Only active vertices (match < 2) are ever touched after the first select:
// they live in four compacted lists per round (blue/red x row length below or
// above the hub degree), filled with one atomic per wavefront.  Neighbour
// colours are read from a 1-byte state mirror (st[]) instead of the 4-byte
// match array, which keeps the random gathers of a 2^22-vertex graph inside
// L2/MALL.  Every phase writes only the processed vertex's own slots and reads
// only slots its phase does not write, so the result is independent of the
// schedule and equals the serial statement of the rounds bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/mm_match.h"
#include "mm_internal.h"

namespace {

constexpr uint8_t ST_BLUE = 0, ST_RED = 1, ST_DEAD = 2, ST_MATCHED = 3;
constexpr int L_BLUE_LO = 0, L_BLUE_HI = 1, L_RED_LO = 2, L_RED_HI = 3, NLISTS = 4;
constexpr int32_t REQ_NONE = -1, REQ_DEAD = -2;
// int32 counters per round slot: [0..3] list sizes, [4] newly matched, [5] newly dead.
constexpr int CTR_STRIDE = 8;
// 64-bit counters per round slot: [0] degree sum of active vertices, [1] of blue ones.
constexpr int CTR64_STRIDE = 2;
constexpr int ROUND_SLOTS = MM_MAX_ROUNDS + 2;
constexpr int K_INIT = 0, K_REQUEST = 1, K_RESPOND = 2, K_MATCH = 3, K_CLEAR = 4;

struct Dev {
    int32_t n;
    int32_t cap;          // capacity of each work list
    int32_t hub_deg;
    uint32_t barrier;
    int32_t count_nnz;
    const int2 *ranges;
    const int32_t *nbrs;  // unweighted layout
    const int2 *adj;      // weighted layout: {neighbour, weight bits}
    int32_t *match;
    int32_t *req;
    uint8_t *st;
    unsigned long long *resp;  // [2][n], PUSH only
    int32_t *lists;            // [2][NLISTS][cap]
    int32_t *ctr;              // [ROUND_SLOTS][CTR_STRIDE]
    unsigned long long *ctr64; // [ROUND_SLOTS][CTR64_STRIDE]
    const uint32_t *seeds;     // seeds[r - 1] drives the select of round r
    const int32_t *vbase;      // batched: first id of the vertex's graph (else null)
    const int32_t *vgid;       // batched: graph index of the vertex
    int32_t *glast;            // batched: last round in which the graph had an active vertex
};

__device__ __forceinline__ int32_t *list_ptr(const Dev &d, int parity, int k)
{
    return d.lists + (size_t)(parity * NLISTS + k) * (size_t)d.cap;
}

__device__ __forceinline__ uint32_t rotl32(uint32_t x, uint32_t s) { return (x << s) | (x >> (32u - s)); }

// The select hash of src/matchgpu.cu:196-217: MD5-shaped, 16 steps, SHA-256
// constants K[0..15], MD5 round-1 rotations, per-step accumulation into h0..h3
// and g multiplied by the seed after every step.
__device__ __forceinline__ uint32_t select_hash(uint32_t i, uint32_t seed)
{
    uint32_t h0 = 0x67452301u, h1 = 0xefcdab89u, h2 = 0x98badcfeu, h3 = 0x10325476u;
    uint32_t a = h0, b = h1, c = h2, d = h3, g = i;
#define MM_STEP(K, R)                                         \
    {                                                         \
        const uint32_t f = (b & c) | (~b & d);                \
        const uint32_t t = d;                                 \
        d = c;                                                \
        c = b;                                                \
        b = b + rotl32(a + f + (K) + g, (R));                 \
        a = t;                                                \
        h0 += a; h1 += b; h2 += c; h3 += d;                   \
        g *= seed;                                            \
    }
    MM_STEP(0x428a2f98u, 7) MM_STEP(0x71374491u, 12) MM_STEP(0xb5c0fbcfu, 17) MM_STEP(0xe9b5dba5u, 22)
    MM_STEP(0x3956c25bu, 7) MM_STEP(0x59f111f1u, 12) MM_STEP(0x923f82a4u, 17) MM_STEP(0xab1c5ed5u, 22)
    MM_STEP(0xd807aa98u, 7) MM_STEP(0x12835b01u, 12) MM_STEP(0x243185beu, 17) MM_STEP(0x550c7dc3u, 22)
    MM_STEP(0x72be5d74u, 7) MM_STEP(0x80deb1feu, 12) MM_STEP(0x9bdc06a7u, 17) MM_STEP(0xc19bf174u, 22)
#undef MM_STEP
    return h0 + h1 + h2 + h3;
}

__device__ __forceinline__ int colour_of(uint32_t local_id, uint32_t seed, uint32_t barrier)
{
    return select_hash(local_id, seed) < barrier ? ST_BLUE : ST_RED;
}

// Appends value to list for every lane with pred set: one atomic per wavefront.
// Must be reached by all lanes of the wavefront.
__device__ __forceinline__ void wave_append(bool pred, int32_t value, int32_t *list, int32_t *counter)
{
    const unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, (int)__popcll(mask));
    base = __shfl(base, leader);
    if (pred) list[base + (int)__popcll(mask & ((1ull << lane) - 1ull))] = value;
}

__device__ __forceinline__ void wave_count(bool pred, int32_t *counter)
{
    const unsigned long long mask = __ballot(pred);
    if (mask != 0ull && __lane_id() == __ffsll((long long)mask) - 1) atomicAdd(counter, (int)__popcll(mask));
}

__device__ __forceinline__ void wave_sum64(unsigned long long x, unsigned long long *counter)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
    if (__lane_id() == 0 && x != 0ull) atomicAdd(counter, x);
}

// Candidate of a row scan: best weight, its CSR position and neighbour, and
// whether any neighbour is still unmatched (the reference's "dead" test).
struct Cand {
    float w;
    int32_t j;
    int32_t u;
    int32_t alive;
};

__device__ __forceinline__ Cand cand_none() { return Cand{-1.0f, INT_MAX, -1, 0}; }

// Order of the reference's sequential scans: strict '>' from -1.0 keeps the
// first maximum in CSR order (weighted); the first hit wins (random).
template <bool W>
__device__ __forceinline__ bool cand_better(const Cand &o, const Cand &m)
{
    if (W) return (o.w > m.w) || (o.w == m.w && o.j < m.j);
    return o.j < m.j;
}

template <bool W>
__device__ __forceinline__ Cand cand_shfl_reduce(Cand c, int width)
{
    for (int off = width >> 1; off >= 1; off >>= 1) {
        Cand o;
        o.w = __shfl_xor(c.w, off, width);
        o.j = __shfl_xor(c.j, off, width);
        o.u = __shfl_xor(c.u, off, width);
        o.alive = __shfl_xor(c.alive, off, width);
        const int alive = c.alive | o.alive;
        if (cand_better<W>(o, c)) c = o;
        c.alive = alive;
    }
    return c;
}

// Block-wide reduction; the result is valid in thread 0.
template <bool W>
__device__ Cand cand_block_reduce(Cand c)
{
    __shared__ float s_w[16];
    __shared__ int32_t s_j[16], s_u[16], s_a[16];
    c = cand_shfl_reduce<W>(c, 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = (blockDim.x + 63) >> 6;
    if (lane == 0) {
        s_w[wave] = c.w;
        s_j[wave] = c.j;
        s_u[wave] = c.u;
        s_a[wave] = c.alive;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        c = lane < nw ? Cand{s_w[lane], s_j[lane], s_u[lane], s_a[lane]} : cand_none();
        c = cand_shfl_reduce<W>(c, 64);
    }
    __syncthreads();
    return c;
}

template <bool A8>
__device__ __forceinline__ void load_entry(const Dev &d, int32_t j, int32_t &u, float &w)
{
    if (A8) {
        const int2 e = d.adj[j];
        u = e.x;
        w = __int_as_float(e.y);
    } else {
        u = d.nbrs[j];
        w = 1.0f;
    }
}

// Request scan of one row by `stride` cooperating lanes starting at `lane`.
template <bool W, bool A8>
__device__ __forceinline__ Cand scan_request(const Dev &d, int2 rg, int lane, int stride)
{
    Cand c = cand_none();
    for (int32_t j = rg.x + lane; j < rg.y; j += stride) {
        int32_t u;
        float w;
        load_entry<A8>(d, j, u, w);
        const uint8_t s = d.st[u];
        c.alive |= (s != ST_MATCHED);
        if (s == ST_RED) {
            if (W) {
                if (w > c.w) { c.w = w; c.j = j; c.u = u; }
            } else {
                c.j = j;
                c.u = u;
                break;  // later positions of this lane cannot win
            }
        }
    }
    return c;
}

// Respond scan (PULL): proposers are blue neighbours whose request names v.
template <bool W, bool A8>
__device__ __forceinline__ Cand scan_respond(const Dev &d, int32_t v, int2 rg, int lane, int stride)
{
    Cand c = cand_none();
    for (int32_t j = rg.x + lane; j < rg.y; j += stride) {
        int32_t u;
        float w;
        load_entry<A8>(d, j, u, w);
        if (d.st[u] != ST_BLUE || d.req[u] != v) continue;
        if (W) {
            if (w > c.w) { c.w = w; c.j = j; c.u = u; }
        } else {
            c.j = j;
            c.u = u;
            break;
        }
    }
    return c;
}

// 64-bit proposal key: order-preserving image of the weight (−0 folded onto +0,
// so equal weights tie) above the complement of the proposer id (smaller id
// wins a tie, i.e. the first proposer of an ascending row).
template <bool W>
__device__ __forceinline__ unsigned long long proposal_key(float w, int32_t v)
{
    uint32_t hi = 1u;
    if (W) {
        const uint32_t bits = __float_as_uint(w == 0.0f ? 0.0f : w);
        hi = (bits & 0x80000000u) ? ~bits : (bits | 0x80000000u);
    }
    return ((unsigned long long)hi << 32) | (unsigned long long)(0xFFFFFFFFu - (uint32_t)v);
}

template <bool W, bool PUSH>
__device__ __forceinline__ void write_request(const Dev &d, int32_t v, const Cand &c, unsigned long long *resp)
{
    if (c.j != INT_MAX) {
        d.req[v] = c.u;
        if (PUSH) atomicMax(resp + c.u, proposal_key<W>(c.w, v));
    } else {
        d.req[v] = c.alive ? REQ_NONE : REQ_DEAD;
    }
}

__global__ void k_clear(int32_t *ctr, unsigned long long *ctr64, int n32, int n64)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n32; i += gridDim.x * blockDim.x) ctr[i] = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n64; i += gridDim.x * blockDim.x) ctr64[i] = 0ull;
}

// Select of round 1 over all vertices (match starts at 0 everywhere,
// src/matchgpu.cu:486), building round 1's work lists.
template <bool PUSH>
__global__ void k_init_select(Dev d)
{
    const int parity = 1;
    int32_t *cnt = d.ctr + 1 * CTR_STRIDE;
    const uint32_t seed = d.seeds[0];
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < d.n;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = (int32_t)(base + (threadIdx.x & 63));
        const bool valid = v < d.n;
        int cls = -1;
        unsigned long long deg = 0;
        if (valid) {
            const int32_t vb = d.vbase ? d.vbase[v] : 0;
            const int c = colour_of((uint32_t)(v - vb), seed, d.barrier);
            d.match[v] = c;
            d.st[v] = (uint8_t)c;
            const int2 rg = d.ranges[v];
            deg = (unsigned long long)(rg.y - rg.x);
            cls = 2 * c + ((rg.y - rg.x) > d.hub_deg ? 1 : 0);
            if (PUSH && c == ST_RED) d.resp[(size_t)parity * d.n + v] = 0ull;
        }
#pragma unroll
        for (int k = 0; k < NLISTS; ++k) wave_append(cls == k, v, list_ptr(d, parity, k), cnt + k);
        if (d.count_nnz) wave_sum64(deg, d.ctr64 + 1 * CTR64_STRIDE + 0);
    }
}

// Request phase of round r.  Blocks [0, hub_blocks) take the hub rows one
// block per row; the rest take the other blue rows, G lanes per row.
template <bool W, bool A8, bool PUSH, int G>
__global__ void k_request(Dev d, int r, int hub_blocks)
{
    const int parity = r & 1;
    const int32_t *cnt = d.ctr + r * CTR_STRIDE;
    unsigned long long *resp = PUSH ? d.resp + (size_t)parity * d.n : nullptr;
    unsigned long long *blue_nnz = d.ctr64 + r * CTR64_STRIDE + 1;
    if ((int)blockIdx.x < hub_blocks) {
        const int32_t count = cnt[L_BLUE_HI];
        const int32_t *list = list_ptr(d, parity, L_BLUE_HI);
        for (int32_t idx = blockIdx.x; idx < count; idx += hub_blocks) {
            const int32_t v = list[idx];
            const int2 rg = d.ranges[v];
            Cand c = scan_request<W, A8>(d, rg, threadIdx.x, blockDim.x);
            c = cand_block_reduce<W>(c);
            if (threadIdx.x == 0) {
                write_request<W, PUSH>(d, v, c, resp);
                if (d.count_nnz) atomicAdd(blue_nnz, (unsigned long long)(rg.y - rg.x));
            }
        }
        return;
    }
    const int32_t count = cnt[L_BLUE_LO];
    const int32_t *list = list_ptr(d, parity, L_BLUE_LO);
    constexpr int GPW = 64 / G;  // groups per wavefront
    const int lane = threadIdx.x & 63, gl = lane & (G - 1), gw = lane / G;
    const int waves_per_block = blockDim.x >> 6;
    const int64_t wave0 = (int64_t)(blockIdx.x - hub_blocks) * waves_per_block + (threadIdx.x >> 6);
    const int64_t stride = (int64_t)(gridDim.x - hub_blocks) * waves_per_block * GPW;
    for (int64_t first = wave0 * GPW; first < count; first += stride) {
        const int64_t idx = first + gw;
        const bool valid = idx < count;
        const int32_t v = valid ? list[idx] : 0;
        const int2 rg = valid ? d.ranges[v] : make_int2(0, 0);
        Cand c = scan_request<W, A8>(d, rg, gl, G);
        if (G > 1) c = cand_shfl_reduce<W>(c, G);
        if (valid && gl == 0) write_request<W, PUSH>(d, v, c, resp);
        if (d.count_nnz) wave_sum64(valid && gl == 0 ? (unsigned long long)(rg.y - rg.x) : 0ull, blue_nnz);
    }
}

// Respond phase of round r (PULL): red vertices pick their proposer.
template <bool W, bool A8, int G>
__global__ void k_respond(Dev d, int r, int hub_blocks)
{
    const int parity = r & 1;
    const int32_t *cnt = d.ctr + r * CTR_STRIDE;
    if ((int)blockIdx.x < hub_blocks) {
        const int32_t count = cnt[L_RED_HI];
        const int32_t *list = list_ptr(d, parity, L_RED_HI);
        for (int32_t idx = blockIdx.x; idx < count; idx += hub_blocks) {
            const int32_t v = list[idx];
            const int2 rg = d.ranges[v];
            Cand c = scan_respond<W, A8>(d, v, rg, threadIdx.x, blockDim.x);
            c = cand_block_reduce<W>(c);
            if (threadIdx.x == 0) d.req[v] = c.j != INT_MAX ? c.u : REQ_NONE;
        }
        return;
    }
    const int32_t count = cnt[L_RED_LO];
    const int32_t *list = list_ptr(d, parity, L_RED_LO);
    constexpr int GPW = 64 / G;
    const int lane = threadIdx.x & 63, gl = lane & (G - 1), gw = lane / G;
    const int waves_per_block = blockDim.x >> 6;
    const int64_t wave0 = (int64_t)(blockIdx.x - hub_blocks) * waves_per_block + (threadIdx.x >> 6);
    const int64_t stride = (int64_t)(gridDim.x - hub_blocks) * waves_per_block * GPW;
    for (int64_t first = wave0 * GPW; first < count; first += stride) {
        const int64_t idx = first + gw;
        const bool valid = idx < count;
        const int32_t v = valid ? list[idx] : 0;
        const int2 rg = valid ? d.ranges[v] : make_int2(0, 0);
        Cand c = scan_respond<W, A8>(d, v, rg, gl, G);
        if (G > 1) c = cand_shfl_reduce<W>(c, G);
        if (valid && gl == 0) d.req[v] = c.j != INT_MAX ? c.u : REQ_NONE;
    }
}

// Match write-back of round r fused with the select of round r+1 (when
// do_select): resolves every vertex of round r's lists, recolours the ones
// still active and appends them to round r+1's lists.
template <bool PUSH>
__global__ void k_match_select(Dev d, int r, int do_select)
{
    const int parity = r & 1, nparity = parity ^ 1;
    int32_t *cnt = d.ctr + r * CTR_STRIDE;
    int32_t *ncnt = d.ctr + (r + 1) * CTR_STRIDE;
    const int32_t c0 = cnt[0], c1 = cnt[1], c2 = cnt[2], c3 = cnt[3];
    const int64_t total = (int64_t)c0 + c1 + c2 + c3;
    const unsigned long long *resp = PUSH ? d.resp + (size_t)parity * d.n : nullptr;
    unsigned long long *nresp = PUSH ? d.resp + (size_t)nparity * d.n : nullptr;
    const uint32_t seed = do_select ? d.seeds[r] : 0u;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < total;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = base + (threadIdx.x & 63);
        const bool valid = t < total;
        int cls = -1;
        bool now_matched = false, now_dead = false;
        int32_t v = 0;
        unsigned long long deg = 0;
        if (valid) {
            int k;
            int64_t idx = t;
            if (idx < c0) k = 0;
            else if ((idx -= c0) < c1) k = 1;
            else if ((idx -= c1) < c2) k = 2;
            else { idx -= c2; k = 3; }
            v = list_ptr(d, parity, k)[idx];
            const int colour = k >> 1, hub = k & 1;
            int32_t partner = -1;
            if (colour == ST_BLUE) {
                const int32_t rq = d.req[v];
                if (rq == REQ_DEAD) now_dead = true;
                else if (rq >= 0) {
                    if (PUSH) {
                        if ((uint32_t)resp[rq] == 0xFFFFFFFFu - (uint32_t)v) partner = rq;
                    } else if (d.req[rq] == v) partner = rq;
                }
            } else {
                if (PUSH) {
                    const unsigned long long key = resp[v];
                    if (key != 0ull) partner = (int32_t)(0xFFFFFFFFu - (uint32_t)key);
                } else {
                    const int32_t rq = d.req[v];
                    if (rq >= 0 && d.req[rq] == v) partner = rq;
                }
            }
            const int32_t vb = d.vbase ? d.vbase[v] : 0;
            if (partner >= 0) {
                now_matched = true;
                d.match[v] = 4 + min(v, partner) - vb;
                d.st[v] = ST_MATCHED;
            } else if (now_dead) {
                d.match[v] = 2;
                d.st[v] = ST_DEAD;
            } else if (do_select) {
                const int c = colour_of((uint32_t)(v - vb), seed, d.barrier);
                if (c != colour) {
                    d.match[v] = c;
                    d.st[v] = (uint8_t)c;
                }
                if (PUSH && c == ST_RED) nresp[v] = 0ull;
                cls = 2 * c + hub;
                if (d.count_nnz) {
                    const int2 rg = d.ranges[v];
                    deg = (unsigned long long)(rg.y - rg.x);
                }
                if (d.glast) atomicMax(d.glast + d.vgid[v], r + 1);
            }
        }
        wave_count(now_matched, cnt + 4);
        wave_count(now_dead, cnt + 5);
        if (do_select) {
#pragma unroll
            for (int k = 0; k < NLISTS; ++k) wave_append(cls == k, v, list_ptr(d, nparity, k), ncnt + k);
            if (d.count_nnz) wave_sum64(deg, d.ctr64 + (r + 1) * CTR64_STRIDE + 0);
        }
    }
}

// Builds the interleaved {neighbour, weight} rows of the weighted layout.
__global__ void k_interleave(int2 *adj, const int32_t *nbrs, const float *w, int64_t nnz)
{
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz; j += (int64_t)gridDim.x * blockDim.x)
        adj[j] = make_int2(nbrs[j], __float_as_int(w[j]));
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct mm_graph {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 256;
    int32_t n = 0;
    int64_t nnz = 0;
    int weighted = 0;
    int push_ok_struct = 0;   // symmetric with strictly ascending rows
    int push_ok_weights = 0;  // and w(u,v) == w(v,u)
    int32_t max_degree = 0;
    double mean_degree = 0.0;
    int64_t deg_hist[33] = {0};  // deg_hist[b]: vertices with degree in [2^b, 2^(b+1)) (b=0: deg 1)
    int2 *d_ranges = nullptr;
    int32_t *d_nbrs = nullptr;
    int2 *d_adj = nullptr;
    int32_t *d_match = nullptr;
    int32_t *d_req = nullptr;
    uint8_t *d_st = nullptr;
    unsigned long long *d_resp = nullptr;
    int32_t *d_lists = nullptr;
    int32_t *d_ctr = nullptr;
    unsigned long long *d_ctr64 = nullptr;
    uint32_t *d_seeds = nullptr;
    int32_t *h_ctr = nullptr;               // pinned
    unsigned long long *h_ctr64 = nullptr;  // pinned
    // batched graphs
    int32_t n_graphs = 0;
    int32_t *d_vbase = nullptr;
    int32_t *d_vgid = nullptr;
    int32_t *d_glast = nullptr;
    std::vector<int32_t> h_glast_init;
    std::vector<uint32_t> last_seeds;  // seed table currently on the device
    hipEvent_t ev_t1 = nullptr, ev_t2 = nullptr;
    std::vector<hipEvent_t> prof_ev;
    int64_t device_bytes = 0;
};

struct mm_batch {
    mm_graph *g = nullptr;
    std::vector<int32_t> vertex_offsets;
};

namespace {

#define MM_HIP_TRY(expr, what)                                                                     \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return mm::set_errorf(e_ == hipErrorOutOfMemory ? MM_ERR_NOMEM : MM_ERR_HIP, "%s: %s", \
                                  what, hipGetErrorString(e_));                                    \
    } while (0)

template <typename T>
int32_t dev_alloc(mm_graph *g, T **p, size_t count, const char *what)
{
    const size_t bytes = sizeof(T) * (count > 0 ? count : 1);
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return mm::set_errorf(MM_ERR_NOMEM, "Not enough memory on device for %s (%zu bytes): %s", what, bytes,
                              hipGetErrorString(e));
    }
    g->device_bytes += (int64_t)bytes;
    return MM_OK;
}

void release(mm_graph *g)
{
    if (!g) return;
    hipSetDevice(g->device);
    if (g->stream) hipStreamSynchronize(g->stream);
    void *ptrs[] = {g->d_ranges, g->d_nbrs, g->d_adj,  g->d_match, g->d_req,   g->d_st,    g->d_resp,
                    g->d_lists,  g->d_ctr,  g->d_ctr64, g->d_seeds, g->d_vbase, g->d_vgid, g->d_glast};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (g->h_ctr) hipHostFree(g->h_ctr);
    if (g->h_ctr64) hipHostFree(g->h_ctr64);
    if (g->ev_t1) hipEventDestroy(g->ev_t1);
    if (g->ev_t2) hipEventDestroy(g->ev_t2);
    for (hipEvent_t e : g->prof_ev) hipEventDestroy(e);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
}

// Host validation of the reference CSR (src/graph.h:41-60) and detection of
// the properties that make PUSH exact.
int32_t analyse_csr(mm_graph *g, const int32_t *ranges, const int32_t *nbrs, const float *w)
{
    const int32_t n = g->n;
    const int64_t nnz = g->nnz;
    std::vector<int32_t> bad(64, 0), unsorted(64, 0), asym(64, 0), wasym(64, 0), maxdeg(64, 0);
    std::vector<std::vector<int64_t>> hist(64, std::vector<int64_t>(33, 0));
    mm::parallel_for(n, [&](int64_t b, int64_t e, unsigned t) {
        for (int64_t i = b; i < e; ++i) {
            const int32_t lo = ranges[2 * i], hi = ranges[2 * i + 1];
            if (lo < 0 || hi < lo || (int64_t)hi > nnz) { bad[t] = 1; continue; }
            const int32_t deg = hi - lo;
            maxdeg[t] = std::max(maxdeg[t], deg);
            if (deg > 0) ++hist[t][31 - __builtin_clz((unsigned)deg)];
            for (int32_t j = lo; j < hi; ++j)
                if (nbrs[j] < 0 || nbrs[j] >= n) bad[t] = 1;
        }
    });
    for (int t = 0; t < 64; ++t)
        if (bad[t]) return mm::set_error(MM_ERR_INVALID, "malformed CSR: range or neighbour id out of bounds");
    mm::parallel_for(n, [&](int64_t b, int64_t e, unsigned t) {
        for (int64_t i = b; i < e; ++i) {
            const int32_t lo = ranges[2 * i], hi = ranges[2 * i + 1];
            for (int32_t j = lo; j < hi; ++j) {
                if (j > lo && nbrs[j] <= nbrs[j - 1]) unsorted[t] = 1;
                const int32_t u = nbrs[j];
                const int32_t *rb = nbrs + ranges[2 * u], *re = nbrs + ranges[2 * u + 1];
                const int32_t *it = std::lower_bound(rb, re, (int32_t)i);
                if (it == re || *it != (int32_t)i) { asym[t] = 1; continue; }
                if (w) {
                    const float a = w[j], c = w[it - nbrs];
                    if (!(a == c || (a != a && c != c))) wasym[t] = 1;
                }
            }
        }
    });
    int us = 0, as = 0, ws = 0;
    for (int t = 0; t < 64; ++t) {
        us |= unsorted[t];
        as |= asym[t];
        ws |= wasym[t];
        g->max_degree = std::max(g->max_degree, maxdeg[t]);
        for (int k = 0; k < 33; ++k) g->deg_hist[k] += hist[t][k];
    }
    g->push_ok_struct = !us && !as;
    g->push_ok_weights = g->push_ok_struct && !ws;
    g->mean_degree = n > 0 ? (double)nnz / (double)n : 0.0;
    return MM_OK;
}

int32_t create_graph(int32_t device, int32_t n, int64_t nnz, const int32_t *ranges, const int32_t *nbrs,
                     const float *w, mm_graph **out)
{
    mm::clear_error();
    if (!out) return mm::set_error(MM_ERR_INVALID, "mm_graph_create: out is NULL");
    *out = nullptr;
    if (n < 0 || n > INT32_MAX - 2 || nnz < 0 || nnz > INT32_MAX || (n > 0 && !ranges) || (nnz > 0 && !nbrs))
        return mm::set_error(MM_ERR_INVALID, "mm_graph_create: bad sizes or NULL arrays");
    mm_graph *g = new (std::nothrow) mm_graph();
    if (!g) return mm::set_error(MM_ERR_NOMEM, "mm_graph_create: host allocation failed");
    g->device = device;
    g->n = n;
    g->nnz = nnz;
    g->weighted = w != nullptr;
    // The CSR is checked on the host before any device is touched.
    int32_t rc = analyse_csr(g, ranges, nbrs, w);
    if (rc != MM_OK) { delete g; return rc; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        delete g;
        return mm::set_error(MM_ERR_HIP, "mm_graph_create: no HIP device available");
    }
    if (device < 0 || device >= ndev) {
        delete g;
        return mm::set_error(MM_ERR_INVALID, "mm_graph_create: bad device index");
    }
    auto fail = [&](int32_t code) { release(g); return code; };
    if (hipSetDevice(device) != hipSuccess) return fail(mm::set_error(MM_ERR_HIP, "hipSetDevice failed"));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        g->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(mm::set_error(MM_ERR_HIP, "hipStreamCreate failed"));
    if (hipEventCreate(&g->ev_t1) != hipSuccess || hipEventCreate(&g->ev_t2) != hipSuccess)
        return fail(mm::set_error(MM_ERR_HIP, "hipEventCreate failed"));
    const size_t nn = (size_t)n;
    if ((rc = dev_alloc(g, &g->d_ranges, nn, "neighbour ranges")) != MM_OK) return fail(rc);
    if (g->weighted) {
        if ((rc = dev_alloc(g, &g->d_adj, (size_t)nnz, "weighted rows")) != MM_OK) return fail(rc);
    } else {
        if ((rc = dev_alloc(g, &g->d_nbrs, (size_t)nnz, "neighbours")) != MM_OK) return fail(rc);
    }
    if ((rc = dev_alloc(g, &g->d_match, nn, "match")) != MM_OK) return fail(rc);
    if ((rc = dev_alloc(g, &g->d_req, nn, "requests")) != MM_OK) return fail(rc);
    if ((rc = dev_alloc(g, &g->d_st, nn, "state")) != MM_OK) return fail(rc);
    if (g->push_ok_struct && (rc = dev_alloc(g, &g->d_resp, 2 * nn, "proposal keys")) != MM_OK) return fail(rc);
    if ((rc = dev_alloc(g, &g->d_lists, 2 * NLISTS * nn, "work lists")) != MM_OK) return fail(rc);
    if ((rc = dev_alloc(g, &g->d_ctr, (size_t)ROUND_SLOTS * CTR_STRIDE, "counters")) != MM_OK) return fail(rc);
    if ((rc = dev_alloc(g, &g->d_ctr64, (size_t)ROUND_SLOTS * CTR64_STRIDE, "counters")) != MM_OK) return fail(rc);
    if ((rc = dev_alloc(g, &g->d_seeds, (size_t)MM_MAX_ROUNDS + 1, "seeds")) != MM_OK) return fail(rc);
    if (hipHostMalloc((void **)&g->h_ctr, sizeof(int32_t) * ROUND_SLOTS * CTR_STRIDE) != hipSuccess ||
        hipHostMalloc((void **)&g->h_ctr64, sizeof(unsigned long long) * ROUND_SLOTS * CTR64_STRIDE) != hipSuccess)
        return fail(mm::set_error(MM_ERR_NOMEM, "pinned host allocation failed"));
    if (n > 0 && hipMemcpy(g->d_ranges, ranges, sizeof(int2) * nn, hipMemcpyHostToDevice) != hipSuccess)
        return fail(mm::set_error(MM_ERR_HIP, "Unable to transfer graph data to device"));
    if (nnz > 0) {
        if (g->weighted) {
            int32_t *tn = nullptr;
            float *tw = nullptr;
            if (hipMalloc((void **)&tn, sizeof(int32_t) * (size_t)nnz) != hipSuccess ||
                hipMalloc((void **)&tw, sizeof(float) * (size_t)nnz) != hipSuccess) {
                if (tn) hipFree(tn);
                return fail(mm::set_error(MM_ERR_NOMEM, "Not enough memory on device to stage graph weights"));
            }
            hipError_t e1 = hipMemcpy(tn, nbrs, sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice);
            hipError_t e2 = hipMemcpy(tw, w, sizeof(float) * (size_t)nnz, hipMemcpyHostToDevice);
            if (e1 == hipSuccess && e2 == hipSuccess) {
                const int blocks = (int)std::min<int64_t>((nnz + 255) / 256, (int64_t)g->num_cu * 16);
                hipLaunchKernelGGL(k_interleave, dim3(blocks), dim3(256), 0, g->stream, g->d_adj, tn, tw, nnz);
                e1 = hipStreamSynchronize(g->stream);
            }
            hipFree(tn);
            hipFree(tw);
            if (e1 != hipSuccess || e2 != hipSuccess)
                return fail(mm::set_error(MM_ERR_HIP, "Unable to transfer graph weights to device"));
        } else if (hipMemcpy(g->d_nbrs, nbrs, sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess) {
            return fail(mm::set_error(MM_ERR_HIP, "Unable to transfer graph data to device"));
        }
    }
    *out = g;
    return MM_OK;
}

int pick_lanes(const mm_graph *g, int requested)
{
    if (requested > 0) return requested;
    int G = 1;
    while (G < 32 && (double)G * 1.5 < g->mean_degree) G <<= 1;
    return G;
}

int64_t hubs_above(const mm_graph *g, int32_t thr)
{
    // Upper bound on the vertices with degree > thr, from the log2 histogram.
    int64_t s = 0;
    for (int b = 0; b < 33; ++b)
        if (b >= 31 || ((int64_t)1 << (b + 1)) > (int64_t)thr) s += g->deg_hist[b];
    return s;
}

struct RunCfg {
    bool W, A8, PUSH;
    int G, block, hub_deg, grid_scan, grid_vertex, hub_blocks;
};

template <bool W, bool A8, bool PUSH, int G>
void launch_request_t(const RunCfg &c, const Dev &d, int r, hipStream_t s)
{
    hipLaunchKernelGGL((k_request<W, A8, PUSH, G>), dim3(c.grid_scan + c.hub_blocks), dim3(c.block), 0, s, d, r,
                       c.hub_blocks);
}

template <bool W, bool A8, int G>
void launch_respond_t(const RunCfg &c, const Dev &d, int r, hipStream_t s)
{
    hipLaunchKernelGGL((k_respond<W, A8, G>), dim3(c.grid_scan + c.hub_blocks), dim3(c.block), 0, s, d, r,
                       c.hub_blocks);
}

template <bool W, bool A8, bool PUSH>
void launch_request_g(const RunCfg &c, const Dev &d, int r, hipStream_t s)
{
    switch (c.G) {
    case 1: launch_request_t<W, A8, PUSH, 1>(c, d, r, s); break;
    case 2: launch_request_t<W, A8, PUSH, 2>(c, d, r, s); break;
    case 4: launch_request_t<W, A8, PUSH, 4>(c, d, r, s); break;
    case 8: launch_request_t<W, A8, PUSH, 8>(c, d, r, s); break;
    case 16: launch_request_t<W, A8, PUSH, 16>(c, d, r, s); break;
    case 32: launch_request_t<W, A8, PUSH, 32>(c, d, r, s); break;
    default: launch_request_t<W, A8, PUSH, 64>(c, d, r, s); break;
    }
}

template <bool W, bool A8>
void launch_respond_g(const RunCfg &c, const Dev &d, int r, hipStream_t s)
{
    switch (c.G) {
    case 1: launch_respond_t<W, A8, 1>(c, d, r, s); break;
    case 2: launch_respond_t<W, A8, 2>(c, d, r, s); break;
    case 4: launch_respond_t<W, A8, 4>(c, d, r, s); break;
    case 8: launch_respond_t<W, A8, 8>(c, d, r, s); break;
    case 16: launch_respond_t<W, A8, 16>(c, d, r, s); break;
    case 32: launch_respond_t<W, A8, 32>(c, d, r, s); break;
    default: launch_respond_t<W, A8, 64>(c, d, r, s); break;
    }
}

void launch_request(const RunCfg &c, const Dev &d, int r, hipStream_t s)
{
    if (c.W) {
        if (c.A8) { if (c.PUSH) launch_request_g<true, true, true>(c, d, r, s); else launch_request_g<true, true, false>(c, d, r, s); }
        else { if (c.PUSH) launch_request_g<true, false, true>(c, d, r, s); else launch_request_g<true, false, false>(c, d, r, s); }
    } else {
        if (c.A8) { if (c.PUSH) launch_request_g<false, true, true>(c, d, r, s); else launch_request_g<false, true, false>(c, d, r, s); }
        else { if (c.PUSH) launch_request_g<false, false, true>(c, d, r, s); else launch_request_g<false, false, false>(c, d, r, s); }
    }
}

void launch_respond(const RunCfg &c, const Dev &d, int r, hipStream_t s)
{
    if (c.W) { if (c.A8) launch_respond_g<true, true>(c, d, r, s); else launch_respond_g<true, false>(c, d, r, s); }
    else { if (c.A8) launch_respond_g<false, true>(c, d, r, s); else launch_respond_g<false, false>(c, d, r, s); }
}

int32_t run_matching(mm_graph *g, const mm_opts *o, mm_stats *stats, std::vector<int32_t> *graph_rounds)
{
    if (!o) return mm::set_error(MM_ERR_INVALID, "mm_match: opts is NULL");
    if (o->kind != MM_RANDOM && o->kind != MM_WEIGHTED) return mm::set_error(MM_ERR_INVALID, "mm_match: bad kind");
    const int maximal = o->maximal ? 1 : 0;
    const int R_max = o->max_rounds > 0 ? o->max_rounds : (maximal ? MM_NR_MAX_MATCH_ROUNDS : MM_NR_MATCH_ROUNDS);
    if (R_max < 1 || R_max > MM_MAX_ROUNDS) return mm::set_error(MM_ERR_INVALID, "mm_match: max_rounds out of range");
    if (!o->seeds || o->n_seeds < 1) return mm::set_error(MM_ERR_SEEDS, "mm_match: no seeds");
    if (!maximal && o->n_seeds < R_max) return mm::set_error(MM_ERR_SEEDS, "mm_match: fewer seeds than rounds");
    const int block = o->block_size > 0 ? o->block_size : 256;
    if (block < 64 || block > 1024 || block % 64) return mm::set_error(MM_ERR_INVALID, "mm_match: bad block_size");
    const int G = pick_lanes(g, o->lanes_per_vertex);
    if (G < 1 || G > 64 || (G & (G - 1))) return mm::set_error(MM_ERR_INVALID, "mm_match: bad lanes_per_vertex");
    RunCfg c;
    c.W = o->kind == MM_WEIGHTED;
    c.A8 = g->weighted != 0;
    const bool push_ok = c.W && g->weighted ? g->push_ok_weights : g->push_ok_struct;
    if (o->respond == MM_RESPOND_PUSH && !push_ok)
        return mm::set_error(MM_ERR_INVALID, "mm_match: PUSH needs a symmetric CSR with ascending rows and symmetric weights");
    if (o->respond < 0 || o->respond > MM_RESPOND_PUSH) return mm::set_error(MM_ERR_INVALID, "mm_match: bad respond mode");
    c.PUSH = o->respond == MM_RESPOND_PUSH || (o->respond == MM_RESPOND_AUTO && push_ok);
    c.G = G;
    c.block = block;
    c.hub_deg = o->hub_degree > 0 ? o->hub_degree : 256;
    const int blocks_per_cu = std::max(1, 2048 / block);
    const int64_t groups_per_block = block / G;
    c.grid_scan = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)g->n + groups_per_block - 1) / groups_per_block,
                                                               (int64_t)g->num_cu * blocks_per_cu));
    c.grid_vertex = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)g->n + block - 1) / block,
                                                                 (int64_t)g->num_cu * blocks_per_cu));
    c.hub_blocks = (int)std::min<int64_t>(hubs_above(g, c.hub_deg), (int64_t)g->num_cu * blocks_per_cu);

    if (hipSetDevice(g->device) != hipSuccess) return mm::set_error(MM_ERR_HIP, "hipSetDevice failed");
    const int n_copy = std::min(o->n_seeds, MM_MAX_ROUNDS);
    // The device keeps the last seed table; it is only re-sent when it changes.
    if ((int)g->last_seeds.size() != n_copy ||
        std::memcmp(g->last_seeds.data(), o->seeds, sizeof(uint32_t) * (size_t)n_copy) != 0) {
        g->last_seeds.assign(o->seeds, o->seeds + n_copy);
        MM_HIP_TRY(hipMemcpyAsync(g->d_seeds, g->last_seeds.data(), sizeof(uint32_t) * (size_t)n_copy,
                                  hipMemcpyHostToDevice, g->stream),
                   "Unable to set seeds");
    }
    if (g->d_glast && !g->h_glast_init.empty())
        MM_HIP_TRY(hipMemcpyAsync(g->d_glast, g->h_glast_init.data(), sizeof(int32_t) * g->h_glast_init.size(),
                                  hipMemcpyHostToDevice, g->stream),
                   "Unable to reset batch state");

    Dev d;
    d.n = g->n;
    d.cap = g->n;
    d.hub_deg = c.hub_deg;
    d.barrier = o->barrier;
    d.count_nnz = o->profile ? 1 : 0;
    d.ranges = g->d_ranges;
    d.nbrs = g->d_nbrs;
    d.adj = g->d_adj;
    d.match = g->d_match;
    d.req = g->d_req;
    d.st = g->d_st;
    d.resp = g->d_resp;
    d.lists = g->d_lists;
    d.ctr = g->d_ctr;
    d.ctr64 = g->d_ctr64;
    d.seeds = g->d_seeds;
    d.vbase = g->d_vbase;
    d.vgid = g->d_vgid;
    d.glast = g->d_glast;

    const bool prof = o->profile != 0;
    size_t ev_used = 0;
    float kms[MM_KERNEL_CLASSES] = {0};
    int klaunch[MM_KERNEL_CLASSES] = {0};
    std::vector<int> ev_class;
    auto ev_begin = [&](int cls) -> int32_t {
        ++klaunch[cls];
        if (!prof) return MM_OK;
        while (g->prof_ev.size() < ev_used + 2) {
            hipEvent_t e;
            MM_HIP_TRY(hipEventCreate(&e), "hipEventCreate");
            g->prof_ev.push_back(e);
        }
        ev_class.push_back(cls);
        MM_HIP_TRY(hipEventRecord(g->prof_ev[ev_used], g->stream), "hipEventRecord");
        return MM_OK;
    };
    auto ev_end = [&]() -> int32_t {
        if (!prof) return MM_OK;
        MM_HIP_TRY(hipEventRecord(g->prof_ev[ev_used + 1], g->stream), "hipEventRecord");
        ev_used += 2;
        return MM_OK;
    };
    int32_t rc;
#define MM_LAUNCH(cls, stmt)                             \
    do {                                                 \
        if ((rc = ev_begin(cls)) != MM_OK) return rc;    \
        stmt;                                            \
        if ((rc = ev_end()) != MM_OK) return rc;         \
    } while (0)

    MM_LAUNCH(K_CLEAR, hipLaunchKernelGGL(k_clear, dim3(8), dim3(256), 0, g->stream, g->d_ctr, g->d_ctr64,
                                          ROUND_SLOTS * CTR_STRIDE, ROUND_SLOTS * CTR64_STRIDE));
    MM_HIP_TRY(hipEventRecord(g->ev_t1, g->stream), "hipEventRecord");
    if (g->n > 0) {
        if (c.PUSH) MM_LAUNCH(K_INIT, hipLaunchKernelGGL(k_init_select<true>, dim3(c.grid_vertex), dim3(block), 0, g->stream, d));
        else MM_LAUNCH(K_INIT, hipLaunchKernelGGL(k_init_select<false>, dim3(c.grid_vertex), dim3(block), 0, g->stream, d));
    }
    int rounds = -1;
    const int chunk = 4;
    // A maximal run may stop early; it can only go as far as the seeds given.
    const int R_run = maximal ? std::min(R_max, o->n_seeds) : R_max;
    int r = 1;
    while (r <= R_run && g->n > 0) {
        const int end = std::min(R_run, r + chunk - 1);
        for (int rr = r; rr <= end; ++rr) {
            const int do_select = rr < R_run;
            MM_LAUNCH(K_REQUEST, launch_request(c, d, rr, g->stream));
            if (!c.PUSH) MM_LAUNCH(K_RESPOND, launch_respond(c, d, rr, g->stream));
            if (c.PUSH) MM_LAUNCH(K_MATCH, hipLaunchKernelGGL(k_match_select<true>, dim3(c.grid_vertex), dim3(block), 0, g->stream, d, rr, do_select));
            else MM_LAUNCH(K_MATCH, hipLaunchKernelGGL(k_match_select<false>, dim3(c.grid_vertex), dim3(block), 0, g->stream, d, rr, do_select));
        }
        MM_HIP_TRY(hipGetLastError(), "A HIP error occurred during the matching process");
        if (maximal) {
            const int k_hi = std::min(end + 1, R_run);
            MM_HIP_TRY(hipMemcpyAsync(g->h_ctr + (size_t)r * CTR_STRIDE, g->d_ctr + (size_t)r * CTR_STRIDE,
                                      sizeof(int32_t) * (size_t)(k_hi - r + 1) * CTR_STRIDE, hipMemcpyDeviceToHost,
                                      g->stream),
                       "Unable to read round counters");
            MM_HIP_TRY(hipStreamSynchronize(g->stream), "A HIP error occurred during the matching process");
            for (int k = r; k <= k_hi; ++k) {
                const int32_t *q = g->h_ctr + (size_t)k * CTR_STRIDE;
                if ((int64_t)q[0] + q[1] + q[2] + q[3] == 0) { rounds = k; break; }
            }
            if (rounds > 0) break;
        }
        r = end + 1;
    }
    if (maximal && rounds < 0 && R_run < R_max && g->n > 0)
        return mm::set_error(MM_ERR_SEEDS, "mm_match: ran out of seeds before the matching converged");
    if (g->n == 0) rounds = maximal ? 1 : R_max;
    if (rounds < 0) rounds = R_max;
    MM_HIP_TRY(hipEventRecord(g->ev_t2, g->stream), "hipEventRecord");
    if (stats) {
        MM_HIP_TRY(hipMemcpyAsync(g->h_ctr, g->d_ctr, sizeof(int32_t) * ROUND_SLOTS * CTR_STRIDE,
                                  hipMemcpyDeviceToHost, g->stream),
                   "Unable to read round counters");
        MM_HIP_TRY(hipMemcpyAsync(g->h_ctr64, g->d_ctr64, sizeof(unsigned long long) * ROUND_SLOTS * CTR64_STRIDE,
                                  hipMemcpyDeviceToHost, g->stream),
                   "Unable to read round counters");
    }
    if (graph_rounds && g->d_glast) {
        graph_rounds->assign((size_t)g->n_graphs, 0);
        MM_HIP_TRY(hipMemcpyAsync(graph_rounds->data(), g->d_glast, sizeof(int32_t) * (size_t)g->n_graphs,
                                  hipMemcpyDeviceToHost, g->stream),
                   "Unable to read batch rounds");
    }
    MM_HIP_TRY(hipStreamSynchronize(g->stream), "A HIP error occurred during the matching process");
    MM_HIP_TRY(hipGetLastError(), "A HIP error occurred during the matching process");
    if (graph_rounds && g->d_glast) {
        for (auto &x : *graph_rounds) x = maximal ? std::min(x + 1, R_max) : R_max;
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->rounds = rounds;
        stats->respond_used = c.PUSH ? MM_RESPOND_PUSH : MM_RESPOND_PULL;
        stats->lanes_used = c.G;
        float ms = 0.0f;
        hipEventElapsedTime(&ms, g->ev_t1, g->ev_t2);
        stats->match_ms = ms;
        int64_t cum = 0;
        for (int k = 1; k <= rounds && k <= MM_MAX_ROUNDS; ++k) {
            const int32_t *q = g->h_ctr + (size_t)k * CTR_STRIDE;
            const unsigned long long *q64 = g->h_ctr64 + (size_t)k * CTR64_STRIDE;
            cum += q[4];
            stats->active[k - 1] = (int64_t)q[0] + q[1] + q[2] + q[3];
            stats->blue[k - 1] = (int64_t)q[0] + q[1];
            stats->matched[k - 1] = cum;
            stats->active_nnz[k - 1] = (int64_t)q64[0];
            stats->blue_nnz[k - 1] = (int64_t)q64[1];
        }
        for (int k = 0; k < MM_KERNEL_CLASSES; ++k) stats->kernel_launches[k] = klaunch[k];
        if (prof) {
            for (size_t e = 0; e < ev_class.size(); ++e) {
                float t = 0.0f;
                hipEventElapsedTime(&t, g->prof_ev[2 * e], g->prof_ev[2 * e + 1]);
                kms[ev_class[e]] += t;
            }
            for (int k = 0; k < MM_KERNEL_CLASSES; ++k) stats->kernel_ms[k] = kms[k];
        }
    }
#undef MM_LAUNCH
    return MM_OK;
}

}  // namespace

extern "C" {

int32_t mm_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int32_t mm_graph_create(int32_t device, int32_t n, int64_t nnz, const int32_t *ranges, const int32_t *neighbours,
                        const float *weights, mm_graph **out)
{
    return create_graph(device, n, nnz, ranges, neighbours, weights, out);
}

int32_t mm_graph_destroy(mm_graph *g)
{
    release(g);
    return MM_OK;
}

int32_t mm_graph_get_info(const mm_graph *g, mm_graph_info *info)
{
    if (!g || !info) return mm::set_error(MM_ERR_INVALID, "mm_graph_get_info: NULL argument");
    info->n = g->n;
    info->weighted = g->weighted;
    info->nnz = g->nnz;
    info->max_degree = g->max_degree;
    info->push_ok = g->weighted ? g->push_ok_weights : g->push_ok_struct;
    info->device_bytes = g->device_bytes;
    return MM_OK;
}

int32_t mm_match(mm_graph *g, const mm_opts *opts, int32_t *match_out, mm_stats *stats)
{
    mm::clear_error();
    if (!g) return mm::set_error(MM_ERR_INVALID, "mm_match: graph is NULL");
    int32_t rc = run_matching(g, opts, stats, nullptr);
    if (rc != MM_OK) return rc;
    if (match_out) return mm_graph_copy_match(g, match_out);
    return MM_OK;
}

int32_t mm_graph_copy_match(mm_graph *g, int32_t *match_out)
{
    if (!g || !match_out) return mm::set_error(MM_ERR_INVALID, "mm_graph_copy_match: NULL argument");
    if (g->n == 0) return MM_OK;
    if (hipSetDevice(g->device) != hipSuccess) return mm::set_error(MM_ERR_HIP, "hipSetDevice failed");
    MM_HIP_TRY(hipMemcpyAsync(match_out, g->d_match, sizeof(int32_t) * (size_t)g->n, hipMemcpyDeviceToHost, g->stream),
               "Unable to retrieve data");
    MM_HIP_TRY(hipStreamSynchronize(g->stream), "Unable to retrieve data");
    return MM_OK;
}

int32_t mm_batch_create(int32_t device, int32_t n_graphs, const int32_t *vertex_offsets, const int64_t *edge_offsets,
                        const int32_t *ranges, const int32_t *neighbours, const float *weights, mm_batch **out)
{
    mm::clear_error();
    if (!out || n_graphs < 0 || !vertex_offsets || !edge_offsets)
        return mm::set_error(MM_ERR_INVALID, "mm_batch_create: bad arguments");
    *out = nullptr;
    const int64_t total_n = vertex_offsets[n_graphs];
    const int64_t total_nnz = edge_offsets[n_graphs];
    if (vertex_offsets[0] != 0 || edge_offsets[0] != 0 || total_n > INT32_MAX - 2 || total_nnz > INT32_MAX)
        return mm::set_error(MM_ERR_INVALID, "mm_batch_create: offsets must start at 0 and fit int32");
    for (int32_t k = 0; k < n_graphs; ++k)
        if (vertex_offsets[k + 1] < vertex_offsets[k] || edge_offsets[k + 1] < edge_offsets[k])
            return mm::set_error(MM_ERR_INVALID, "mm_batch_create: offsets must be non-decreasing");
    // Rebase every graph into one global CSR.
    std::vector<int32_t> granges((size_t)total_n * 2), gnbrs((size_t)total_nnz);
    std::vector<int32_t> vbase((size_t)total_n), vgid((size_t)total_n);
    int bad = 0;
    mm::parallel_for(n_graphs, [&](int64_t b, int64_t e, unsigned) {
        for (int64_t k = b; k < e; ++k) {
            const int32_t v0 = vertex_offsets[k], v1 = vertex_offsets[k + 1];
            const int64_t e0 = edge_offsets[k], e1 = edge_offsets[k + 1];
            const int32_t nk = v1 - v0;
            for (int32_t v = v0; v < v1; ++v) {
                const int64_t lo = ranges[2 * (size_t)v], hi = ranges[2 * (size_t)v + 1];
                if (lo < 0 || hi < lo || e0 + hi > e1) { bad = 1; continue; }
                granges[2 * (size_t)v] = (int32_t)(e0 + lo);
                granges[2 * (size_t)v + 1] = (int32_t)(e0 + hi);
                vbase[(size_t)v] = v0;
                vgid[(size_t)v] = (int32_t)k;
            }
            for (int64_t j = e0; j < e1; ++j) {
                const int32_t u = neighbours[j];
                if (u < 0 || u >= nk) { bad = 1; continue; }
                gnbrs[(size_t)j] = u + v0;
            }
        }
    });
    if (bad) return mm::set_error(MM_ERR_INVALID, "mm_batch_create: malformed per-graph CSR");
    mm_graph *g = nullptr;
    int32_t rc = create_graph(device, (int32_t)total_n, total_nnz, granges.data(), gnbrs.data(), weights, &g);
    if (rc != MM_OK) return rc;
    g->n_graphs = n_graphs;
    if ((rc = dev_alloc(g, &g->d_vbase, (size_t)total_n, "batch bases")) != MM_OK ||
        (rc = dev_alloc(g, &g->d_vgid, (size_t)total_n, "batch ids")) != MM_OK ||
        (rc = dev_alloc(g, &g->d_glast, (size_t)n_graphs, "batch rounds")) != MM_OK) {
        release(g);
        return rc;
    }
    g->h_glast_init.resize((size_t)n_graphs);
    for (int32_t k = 0; k < n_graphs; ++k) g->h_glast_init[(size_t)k] = vertex_offsets[k + 1] > vertex_offsets[k] ? 1 : 0;
    if (total_n > 0 &&
        (hipMemcpy(g->d_vbase, vbase.data(), sizeof(int32_t) * (size_t)total_n, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(g->d_vgid, vgid.data(), sizeof(int32_t) * (size_t)total_n, hipMemcpyHostToDevice) != hipSuccess)) {
        release(g);
        return mm::set_error(MM_ERR_HIP, "mm_batch_create: upload failed");
    }
    mm_batch *b = new (std::nothrow) mm_batch();
    if (!b) {
        release(g);
        return mm::set_error(MM_ERR_NOMEM, "mm_batch_create: host allocation failed");
    }
    b->g = g;
    b->vertex_offsets.assign(vertex_offsets, vertex_offsets + n_graphs + 1);
    *out = b;
    return MM_OK;
}

int32_t mm_batch_destroy(mm_batch *b)
{
    if (!b) return MM_OK;
    release(b->g);
    delete b;
    return MM_OK;
}

int32_t mm_batch_match(mm_batch *b, const mm_opts *opts, int32_t *match_out, int32_t *rounds_out, mm_stats *stats)
{
    mm::clear_error();
    if (!b || !b->g) return mm::set_error(MM_ERR_INVALID, "mm_batch_match: batch is NULL");
    std::vector<int32_t> gr;
    int32_t rc = run_matching(b->g, opts, stats, rounds_out ? &gr : nullptr);
    if (rc != MM_OK) return rc;
    if (rounds_out) std::copy(gr.begin(), gr.end(), rounds_out);
    if (match_out) return mm_graph_copy_match(b->g, match_out);
    return MM_OK;
}

}  // extern "C"
