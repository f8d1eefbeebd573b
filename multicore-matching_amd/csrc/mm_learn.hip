// Learned vertex selection (gfx950): a float32 8 -> 32 -> 32 -> 1 network
// picks each active vertex's blue probability at every select; request,
// respond and match keep the reference rules (src/matchgpu.cu:260-451).
// Also the network's forward/backward and Adam kernels that train it on round
// traces of the plain rounds.
//
// The rounds here are full-n kernels (one thread per vertex, rows scanned in
// the handle's matching order: weight-sorted for weighted rows, CSR order
// otherwise), four launches per round:
//   select   features of the vertex from its row, network, colour
//   request  blue: first red neighbour in matching order (heaviest, first in
//            CSR order on ties), else "none" or "dead"
//   respond  red: first blue neighbour in matching order that proposed to it
//   match    mutual pairs become 4 + min(id), dead vertices 2, the rest keep
//            this round's colour
// Each phase writes only the processed vertex's slots and reads slots its
// phase does not write, so the result is schedule-independent.
//
// All network arithmetic (features, dot products, the blue threshold) is
// float32/float64 with explicit fused multiply-adds and no contraction
// elsewhere (this file is compiled with -ffp-contract=off), so a host
// statement of the same steps reproduces every colour bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/mm_learn.h"
#include "../../include/mm_match.h"
#include "mm_graph_impl.h"
#include "mm_hash.h"
#include "mm_internal.h"

namespace {

constexpr int NF = MM_MLP_FEATURES, NH = MM_MLP_HIDDEN, NP = MM_MLP_PARAMS;
constexpr int LCTR = 4;  // per round: active at select, blue, newly matched, newly dead
constexpr int LPART = 4096;  // most blocks of a counting launch
constexpr int LSLOTS = MM_MAX_ROUNDS + 2;
constexpr int32_t LREQ_NONE = -1, LREQ_DEAD = -2;
// col[v]: this round's colour of an active vertex, LC_DONE once it left the rounds.
constexpr uint8_t LC_BLUE = 0, LC_RED = 1, LC_RETIRE = 2, LC_DONE = 3;
// st[v]: state after the previous round: 0 blue, 1 red, 2 dead, 3 matched
// (match[v] in one byte, the form neighbour gathers read).
constexpr int SEL_G = 8;   // lanes per row in the select's feature scans
constexpr int SCAN_G = 4;  // lanes per row in the request / respond scans
constexpr int SCAN_ROWS = 4;  // rows each select group scans at once

static_assert(MM_MLP_OFF_B3 + 1 == NP, "parameter layout");

// ---------------------------------------------------------------------------
// Exact network arithmetic

__host__ __device__ inline float relu(float a) { return a > 0.0f ? a : 0.0f; }

// z = W3 . relu(W2 . relu(W1 . x + b1) + b2) + b3, fused multiply-adds in index order.
__host__ __device__ inline float mlp_z(const float *__restrict__ P, const float *x)
{
    float h1[NH];
#pragma unroll
    for (int j = 0; j < NH; ++j) {
        float a = P[MM_MLP_OFF_B1 + j];
#pragma unroll
        for (int k = 0; k < NF; ++k) a = fmaf(P[MM_MLP_OFF_W1 + j * NF + k], x[k], a);
        h1[j] = relu(a);
    }
    float z = P[MM_MLP_OFF_B3];
#pragma unroll 4
    for (int j = 0; j < NH; ++j) {
        float a = P[MM_MLP_OFF_B2 + j];
#pragma unroll
        for (int k = 0; k < NH; ++k) a = fmaf(P[MM_MLP_OFF_W2 + j * NH + k], h1[k], a);
        z = fmaf(P[MM_MLP_OFF_W3 + j], relu(a), z);
    }
    return z;
}

// mlp_z for two feature rows at once: every weight is loaded once for both,
// and each row's arithmetic is exactly mlp_z's.
__device__ inline void mlp_z2(const float *__restrict__ P, const float *x0, const float *x1, float &z0, float &z1)
{
    float h0[NH], h1[NH];
#pragma unroll
    for (int j = 0; j < NH; ++j) {
        float a0 = P[MM_MLP_OFF_B1 + j], a1 = a0;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const float wjk = P[MM_MLP_OFF_W1 + j * NF + k];
            a0 = fmaf(wjk, x0[k], a0);
            a1 = fmaf(wjk, x1[k], a1);
        }
        h0[j] = relu(a0);
        h1[j] = relu(a1);
    }
    z0 = P[MM_MLP_OFF_B3];
    z1 = z0;
#pragma unroll 2
    for (int j = 0; j < NH; ++j) {
        float a0 = P[MM_MLP_OFF_B2 + j], a1 = a0;
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            const float wjk = P[MM_MLP_OFF_W2 + j * NH + k];
            a0 = fmaf(wjk, h0[k], a0);
            a1 = fmaf(wjk, h1[k], a1);
        }
        z0 = fmaf(P[MM_MLP_OFF_W3 + j], relu(a0), z0);
        z1 = fmaf(P[MM_MLP_OFF_W3 + j], relu(a1), z1);
    }
}

// floor(65536 * s(z)), s(z) = 1 / (1 + 2^(-z log2 e)); 2^f by the degree-7
// Taylor polynomial of exp(f ln 2).  The floor is settled with exact double
// products (a 17-bit integer times a float fits a double's mantissa).
__host__ __device__ inline uint32_t blue_threshold(float z)
{
    const float t = -z * 0x1.715476p+0f;
    if (!(t < 24.0f)) return 0u;
    if (t < -24.0f) return 65536u;
    const float fi = floorf(t);
    const float f = t - fi;
    float p = 0x1.ffcbfcp-17f;
    p = fmaf(p, f, 0x1.430912p-13f);
    p = fmaf(p, f, 0x1.5d87fep-10f);
    p = fmaf(p, f, 0x1.3b2ab6p-7f);
    p = fmaf(p, f, 0x1.c6b08ep-5f);
    p = fmaf(p, f, 0x1.ebfbe0p-3f);
    p = fmaf(p, f, 0x1.62e430p-1f);
    p = fmaf(p, f, 1.0f);
    const float den = 1.0f + ldexpf(p, (int)fi);
    const double dd = (double)den;
    uint32_t T = (uint32_t)(65536.0f / den);
    if (T > 65536u) T = 65536u;
    while (T < 65536u && (double)(T + 1u) * dd <= 65536.0) ++T;
    while (T > 0u && (double)T * dd > 65536.0) --T;
    return T;
}

// ---------------------------------------------------------------------------
// Learned rounds

struct LDev {
    int32_t n;
    uint32_t barrier;
    int32_t use_mlp;
    int32_t trace;
    uint32_t trace_keep;
    int64_t trace_cap;
    const int2 *ranges;
    const int32_t *nbrs;
    const int2 *adj;
    const int32_t *loc;
    int32_t *match;
    int32_t *req;
    uint8_t *col;
    uint8_t *st;
    float *reqw;
    float *feat;
    int32_t *ctr;  // [LSLOTS][LCTR]
    int32_t *part; // [LCTR][LPART] per-block partial counts of the current round
    float *sx, *sy, *sw;
    int32_t *sv, *sr;
    unsigned long long *scount;
    unsigned long long *resp;  // PUSH proposal keys, one per vertex
    int32_t *lact, *lblue;     // work lists of the round: block b owns [b * seg, (b + 1) * seg)
    int32_t *lcnt;             // [2][LPART] per-block list lengths (active, blue)
    const int32_t *lprev;      // the previous round's active list (same segments) ...
    const int32_t *lcnt_prev;  // ... and its per-block lengths
    int64_t seg;
};

// Row entry j in matching order: weight-sorted pairs (id complemented when the
// weight can never win) or the CSR-order ids with unit weights.
template <bool A8>
__device__ __forceinline__ void l_entry(const LDev &d, int32_t j, int32_t &u, bool &valid, float &w)
{
    if (A8) {
        const int2 e = d.adj[j];
        valid = e.x >= 0;
        u = valid ? e.x : ~e.x;
        w = __int_as_float(e.y);
    } else {
        u = d.nbrs[j];
        valid = true;
        w = 1.0f;
    }
}

// Sum of x over the block (every thread must call it; blockDim.x a multiple
// of 64, at most 1024); the result is valid in thread 0.
__device__ __forceinline__ int block_sum(int x)
{
    __shared__ int ws[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    const int wid = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) ws[wid] = x;
    __syncthreads();
    int s = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += ws[k];
    return s;
}

__global__ void k_l_start(int32_t *match, uint8_t *st, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        match[i] = 0;
        st[i] = 0;
    }
}

__global__ void k_l_fill(int32_t *p, int64_t count, int32_t value)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

constexpr int LTILE = 256;  // vertices per tile = threads per block of every round kernel

// Sum over the G aligned lanes of a group (G a power of two <= 64).
template <int G>
__device__ __forceinline__ int group_sum(int x)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// The two smallest of the group's (p1 < p2) position pairs.
template <int G>
__device__ __forceinline__ void group_min2(int &p1, int &p2)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        const int q1 = __shfl_xor(p1, o), q2 = __shfl_xor(p2, o);
        const int n1 = min(p1, q1);
        p2 = min(max(p1, q1), min(p2, q2));
        p1 = n1;
    }
}

// Exclusive prefix of flag over the LTILE threads of the block: this thread's
// position among the flagged ones; total receives their number.  Every thread
// must call it.
__device__ __forceinline__ int block_prefix(bool flag, int &total)
{
    __shared__ int wsum[LTILE / 64];
    const unsigned long long m = __ballot(flag);
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) wsum[wid] = (int)__popcll(m);
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < LTILE / 64; ++k) {
        const int c = wsum[k];
        if (k < wid) before += c;
        total += c;
    }
    return before + (int)__popcll(m & ((1ull << lane) - 1ull));
}

// Select of round r.  Block b takes tiles b, b + gridDim, ... of STILE
// vertices and keeps the active ones of each tile in LDS, packed.  Phase 1:
// groups of G lanes scan the packed vertices' rows (coalesced chunks of G
// entries) for the unmatched / previously-red neighbour counts and the row
// positions of the first two candidates.  Phase 2: each thread takes two
// packed vertices, reads those entries, hashes, evaluates the network for both
// at once and draws the colours (PUSH rounds also clear the proposal keys).
// The active and the blue vertices go to block b's segments of the round's
// work lists, so the other phases touch only them.  The network parameters
// arrive as their own read-only, non-aliased argument, so their wave-uniform
// loads are scalar.
constexpr int STILE = 2 * LTILE;  // vertices per select tile

struct SelOut {
    int c;
    float x[NF];
    bool need_net;
};

// Features and plain colour of packed vertex k of the tile (phase 2).
template <bool A8>
__device__ __forceinline__ void sel_features(const LDev &d, int r, uint32_t seed, bool scan, int32_t v,
                                             const int32_t *s_p1, const int32_t *s_p2, const int32_t *s_alive,
                                             const int32_t *s_redp, int k, uint32_t &vl, SelOut &o)
{
    vl = (uint32_t)d.loc[v];
    const uint32_t hv = mm_select_hash(vl, seed);
    o.c = hv < d.barrier ? LC_BLUE : LC_RED;
    o.need_net = false;
    if (!scan) return;
    const int2 rg = d.ranges[v];
    const int32_t p1 = s_p1[k], p2 = s_p2[k], alive = s_alive[k];
    float w1 = 0.0f, w2 = 0.0f, bit = 0.0f;
    if (p1 != INT_MAX) {
        int32_t u;
        bool valid;
        l_entry<A8>(d, p1, u, valid, w1);
        bit = hv < mm_select_hash((uint32_t)d.loc[u], seed) ? 1.0f : 0.0f;
    }
    if (p2 != INT_MAX) {
        int32_t u;
        bool valid;
        l_entry<A8>(d, p2, u, valid, w2);
    }
    o.x[0] = (float)min(rg.y - rg.x, 1024) * 0.015625f;
    o.x[1] = w1;
    o.x[2] = w2;
    o.x[3] = w1 > 0.0f ? (float)((double)w2 / (double)w1) : 0.0f;
    o.x[4] = (float)min(alive, 1024) * 0.015625f;
    o.x[5] = (float)min(s_redp[k], 1024) * 0.015625f;
    o.x[6] = (float)(r - 1) * 0.03125f;
    o.x[7] = bit;
    if (d.trace) {
#pragma unroll
        for (int q = 0; q < NF; ++q) d.feat[(size_t)v * NF + q] = o.x[q];
    }
    if (d.use_mlp && r <= MM_LEARN_ROUNDS) {
        if (alive == 0) o.c = LC_RETIRE;
        else o.need_net = true;
    }
}

__device__ __forceinline__ int sel_draw(uint32_t vl, uint32_t seed, float z, int plain)
{
    if (!isfinite(z)) return plain;
    const uint32_t T = blue_threshold(z);
    return (mm_select_hash(vl, seed ^ MM_LEARN_DRAW_TWEAK) >> 16) < T ? LC_BLUE : LC_RED;
}

template <bool A8, bool PUSH, int G>
__global__ __launch_bounds__(LTILE) void k_l_select(LDev d, int r, uint32_t seed, const float *__restrict__ params)
{
    __shared__ int32_t s_idx[STILE], s_p1[STILE], s_p2[STILE], s_alive[STILE], s_redp[STILE];
    const int t = threadIdx.x;
    const int gl = t & (G - 1);
    const int grp = t / G;
    constexpr int GROUPS = LTILE / G;
    const bool scan = (d.use_mlp && r <= MM_LEARN_ROUNDS) || d.trace;
    int32_t *act_seg = d.lact + (size_t)blockIdx.x * d.seg;
    int32_t *blue_seg = d.lblue + (size_t)blockIdx.x * d.seg;
    int n_act = 0, n_blue = 0;
    // Round 1 takes tiles b, b + gridDim, ... of the ids; later rounds take the
    // vertices block b listed as active in the round before (a superset of the
    // active ones, in ascending order), so a round costs what its survivors cost.
    const bool from_list = r > 1;
    const int32_t *prev = d.lprev + (size_t)blockIdx.x * d.seg;
    const int64_t limit = from_list ? (int64_t)d.lcnt_prev[blockIdx.x] : (int64_t)d.n;
    const int64_t stride = from_list ? (int64_t)STILE : (int64_t)gridDim.x * STILE;
    for (int64_t base = from_list ? 0 : (int64_t)blockIdx.x * STILE; base < limit; base += stride) {
        int32_t c0 = -1, c1 = -1;
        if (base + t < limit) c0 = from_list ? prev[base + t] : (int32_t)(base + t);
        if (base + LTILE + t < limit) c1 = from_list ? prev[base + LTILE + t] : (int32_t)(base + LTILE + t);
        int A0, A1;
        const bool act0 = c0 >= 0 && d.match[c0] < 2;
        const bool act1 = c1 >= 0 && d.match[c1] < 2;
        const int pos0 = block_prefix(act0, A0);
        if (act0) s_idx[pos0] = c0;
        const int pos1 = block_prefix(act1, A1);
        if (act1) s_idx[A0 + pos1] = c1;
        const int A = A0 + A1;
        __syncthreads();
        if (scan) {
            if (r == 1) {
                // Round 1: every vertex is unmatched, so a row's unmatched
                // neighbours are all of it, none was red before, and its first
                // two candidates are its first two entries that can win (rows
                // keep those first).
                for (int k = t; k < A; k += LTILE) {
                    const int2 rg = d.ranges[s_idx[k]];
                    int32_t p1 = INT_MAX, p2 = INT_MAX;
                    for (int32_t j = rg.x; j < rg.y && j < rg.x + 2; ++j) {
                        int32_t u;
                        bool valid;
                        float w;
                        l_entry<A8>(d, j, u, valid, w);
                        if (!valid) break;
                        if (p1 == INT_MAX) p1 = j;
                        else p2 = j;
                    }
                    s_p1[k] = p1;
                    s_p2[k] = p2;
                    s_alive[k] = rg.y - rg.x;
                    s_redp[k] = 0;
                }
            } else
            // SCAN_ROWS rows per group at once: their entry loads, then their
            // state gathers, are in flight together.
            for (int k0 = grp; k0 < A; k0 += GROUPS * SCAN_ROWS) {
                int2 rg[SCAN_ROWS];
                int alive[SCAN_ROWS], redp[SCAN_ROWS], p1[SCAN_ROWS], p2[SCAN_ROWS];
                int32_t maxlen = 0;
#pragma unroll
                for (int q = 0; q < SCAN_ROWS; ++q) {
                    const int k = k0 + q * GROUPS;
                    rg[q] = k < A ? d.ranges[s_idx[k]] : make_int2(0, 0);
                    maxlen = max(maxlen, rg[q].y - rg[q].x);
                    alive[q] = redp[q] = 0;
                    p1[q] = p2[q] = INT_MAX;
                }
                for (int32_t off = gl; off < maxlen; off += G) {
                    int32_t u[SCAN_ROWS];
                    bool valid[SCAN_ROWS], in[SCAN_ROWS];
#pragma unroll
                    for (int q = 0; q < SCAN_ROWS; ++q) {
                        in[q] = off < rg[q].y - rg[q].x;
                        float w;
                        u[q] = 0;
                        valid[q] = false;
                        if (in[q]) l_entry<A8>(d, rg[q].x + off, u[q], valid[q], w);
                    }
                    int su[SCAN_ROWS];
#pragma unroll
                    for (int q = 0; q < SCAN_ROWS; ++q) su[q] = in[q] ? (int)d.st[u[q]] : 3;
#pragma unroll
                    for (int q = 0; q < SCAN_ROWS; ++q) {
                        if (su[q] >= 3) continue;
                        ++alive[q];
                        redp[q] += su[q] == 1;
                        if (valid[q]) {
                            const int32_t j = rg[q].x + off;
                            if (p1[q] == INT_MAX) p1[q] = j;
                            else if (p2[q] == INT_MAX) p2[q] = j;
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < SCAN_ROWS; ++q) {
                    const int k = k0 + q * GROUPS;
                    const int a = group_sum<G>(alive[q]);
                    const int rp = group_sum<G>(redp[q]);
                    group_min2<G>(p1[q], p2[q]);
                    if (gl == 0 && k < A) {
                        s_p1[k] = p1[q];
                        s_p2[k] = p2[q];
                        s_alive[k] = a;
                        s_redp[k] = rp;
                    }
                }
            }
            __syncthreads();
        }
        // phase 2: packed vertices t and t + LTILE
        const bool has0 = t < A, has1 = t + LTILE < A;
        int32_t v0 = -1, v1 = -1;
        uint32_t vl0 = 0, vl1 = 0;
        SelOut o0, o1;
        o0.c = o1.c = LC_RED;
        o0.need_net = o1.need_net = false;
        if (has0) {
            v0 = s_idx[t];
            sel_features<A8>(d, r, seed, scan, v0, s_p1, s_p2, s_alive, s_redp, t, vl0, o0);
        }
        if (has1) {
            v1 = s_idx[t + LTILE];
            sel_features<A8>(d, r, seed, scan, v1, s_p1, s_p2, s_alive, s_redp, t + LTILE, vl1, o1);
        }
        if (o0.need_net && o1.need_net) {
            float z0, z1;
            mlp_z2(params, o0.x, o1.x, z0, z1);
            o0.c = sel_draw(vl0, seed, z0, o0.c);
            o1.c = sel_draw(vl1, seed, z1, o1.c);
        } else if (o0.need_net) {
            o0.c = sel_draw(vl0, seed, mlp_z(params, o0.x), o0.c);
        } else if (o1.need_net) {
            o1.c = sel_draw(vl1, seed, mlp_z(params, o1.x), o1.c);
        }
        if (has0) {
            d.col[v0] = (uint8_t)o0.c;
            if (PUSH) d.resp[v0] = 0ull;
            act_seg[n_act + t] = v0;
        }
        if (has1) {
            d.col[v1] = (uint8_t)o1.c;
            if (PUSH) d.resp[v1] = 0ull;
            act_seg[n_act + LTILE + t] = v1;
        }
        int B0, B1;
        const bool blue0 = has0 && o0.c == LC_BLUE, blue1 = has1 && o1.c == LC_BLUE;
        const int bpos0 = block_prefix(blue0, B0);
        if (blue0) blue_seg[n_blue + bpos0] = v0;
        const int bpos1 = block_prefix(blue1, B1);
        if (blue1) blue_seg[n_blue + B0 + bpos1] = v1;
        n_act += A;
        n_blue += B0 + B1;
    }
    if (t == 0) {
        d.lcnt[blockIdx.x] = n_act;
        d.lcnt[LPART + blockIdx.x] = n_blue;
    }
}

// Proposal key of a PUSH request: order-preserving image of the weight (-0
// folded onto +0) above the complement of the proposer's graph-local id, so
// the largest key is the heaviest proposer, the smallest id on ties -- the
// first in an ascending row, as the reference's respond scan picks it.
__device__ __forceinline__ unsigned long long l_key(float w, uint32_t vl)
{
    const uint32_t bits = __float_as_uint(w == 0.0f ? 0.0f : w);
    const uint32_t hi = (bits & 0x80000000u) ? ~bits : (bits | 0x80000000u);
    return ((unsigned long long)hi << 32) | (unsigned long long)(0xFFFFFFFFu - vl);
}

__device__ __forceinline__ float l_key_weight(unsigned long long k)
{
    const uint32_t hi = (uint32_t)(k >> 32);
    return __uint_as_float((hi & 0x80000000u) ? (hi & 0x7FFFFFFFu) : ~hi);
}

// Bits of this lane's group in a wavefront ballot, as a G-bit mask.
template <int G>
__device__ __forceinline__ unsigned long long group_bits(unsigned long long ballot)
{
    const int shift = (__lane_id() / G) * G;
    return G == 64 ? ballot : (ballot >> shift) & ((1ull << G) - 1ull);
}

// First candidate of v's row in matching order, G lanes per row, G entries
// per step; RESPOND looks for a blue unmatched proposer of v, else for a red
// unmatched neighbour (alive: v has an unmatched neighbour).  Every lane of
// the wavefront must call it, has = false for lanes without a vertex.
template <bool A8, bool RESPOND, int G>
__device__ __forceinline__ bool l_first(const LDev &d, bool has, int32_t v, int32_t &best, float &bw, bool &alive)
{
    const int gl = __lane_id() & (G - 1);
    int2 rg = make_int2(0, 0);
    if (has) rg = d.ranges[v];
    bool found = false;
    alive = false;
    for (int32_t jb = rg.x; __ballot(has && !found && jb < rg.y) != 0ull; jb += G) {
        const bool live = has && !found && jb < rg.y;
        const int32_t j = jb + gl;
        bool cand = false;
        int32_t u = 0;
        float w = 0.0f;
        if (live && j < rg.y) {
            bool valid;
            l_entry<A8>(d, j, u, valid, w);
            // col: this round's colour of an active neighbour, LC_DONE once
            // it left (an active vertex never neighbours a dead one)
            const int cu = d.col[u];
            if (RESPOND) {
                cand = valid && cu == LC_BLUE && d.req[u] == v;
            } else {
                alive |= cu != LC_DONE;
                cand = valid && cu == LC_RED;
            }
        }
        const unsigned long long gm = group_bits<G>(__ballot(cand));
        const int src = (__lane_id() & ~(G - 1)) + (gm ? __ffsll((long long)gm) - 1 : 0);
        const int32_t su = __shfl(u, src);
        const float sw = __shfl(w, src);
        if (live && gm) {
            best = su;
            bw = sw;
            found = true;
        }
    }
    if (!RESPOND) alive = group_bits<G>(__ballot(alive)) != 0ull;
    return found;
}

// Request: block b's blue list, G lanes per vertex.
template <bool A8, bool PUSH, int G>
__global__ __launch_bounds__(LTILE) void k_l_request(LDev d)
{
    const int32_t cnt = d.lcnt[LPART + blockIdx.x];
    const int32_t *list = d.lblue + (size_t)blockIdx.x * d.seg;
    constexpr int GPB = LTILE / G;
    const int grp = threadIdx.x / G, gl = threadIdx.x & (G - 1);
    const int32_t rounded = (cnt + GPB - 1) / GPB * GPB;
    for (int32_t k = grp; k < rounded; k += GPB) {
        const bool has = k < cnt;
        const int32_t v = has ? list[k] : 0;
        int32_t best = -1;
        float bw = 0.0f;
        bool alive;
        const bool found = l_first<A8, false, G>(d, has, v, best, bw, alive);
        if (has && gl == 0) {
            d.req[v] = found ? best : (alive ? LREQ_NONE : LREQ_DEAD);
            d.reqw[v] = found ? bw : 0.0f;
            if (PUSH && found) atomicMax(d.resp + best, l_key(bw, (uint32_t)d.loc[v]));
        }
    }
}

// Respond (PULL): block b's active list; each red vertex takes its first blue
// proposer in matching order.
template <bool A8, int G>
__global__ __launch_bounds__(LTILE) void k_l_respond(LDev d)
{
    const int32_t cnt = d.lcnt[blockIdx.x];
    const int32_t *list = d.lact + (size_t)blockIdx.x * d.seg;
    constexpr int GPB = LTILE / G;
    const int grp = threadIdx.x / G, gl = threadIdx.x & (G - 1);
    const int32_t rounded = (cnt + GPB - 1) / GPB * GPB;
    for (int32_t k = grp; k < rounded; k += GPB) {
        const int32_t v = k < cnt ? list[k] : 0;
        const bool red = k < cnt && d.col[v] == LC_RED;
        int32_t best = -1;
        float bw = 0.0f;
        bool alive;
        const bool found = l_first<A8, true, G>(d, red, v, best, bw, alive);
        if (red && gl == 0) {
            d.req[v] = found ? best : LREQ_NONE;
            if (found) d.reqw[v] = bw;
        }
    }
}

// Match: block b's active list, one thread per vertex.
template <bool PUSH>
__global__ __launch_bounds__(LTILE) void k_l_match(LDev d, int r, uint32_t seed)
{
    const int32_t cnt = d.lcnt[blockIdx.x];
    const int32_t *list = d.lact + (size_t)blockIdx.x * d.seg;
    int n_mat = 0, n_dead = 0;
    const int32_t rounded = (cnt + LTILE - 1) / LTILE * LTILE;
    for (int32_t k = threadIdx.x; k < rounded; k += LTILE) {
        const bool act = k < cnt;
        int32_t nm = 0, v = 0;
        int c = LC_RED;
        uint32_t vl = 0;
        if (act) {
            v = list[k];
            c = d.col[v];
            const int32_t q = d.req[v];
            vl = (uint32_t)d.loc[v];
            if (c == LC_RETIRE || (c == LC_BLUE && q == LREQ_DEAD)) {
                nm = 2;
            } else if (PUSH) {
                // red: the winning key names its proposer; blue: it won iff its target's key names it
                if (c == LC_RED) {
                    const unsigned long long key = d.resp[v];
                    if (key != 0ull) {
                        nm = 4 + (int32_t)min(vl, 0xFFFFFFFFu - (uint32_t)key);
                        d.reqw[v] = l_key_weight(key);
                    } else {
                        nm = c;
                    }
                } else if (q >= 0 && (uint32_t)d.resp[q] == 0xFFFFFFFFu - vl) {
                    nm = 4 + (int32_t)min(vl, (uint32_t)d.loc[q]);
                } else {
                    nm = c;
                }
            } else if (q >= 0 && d.req[q] == v) {
                nm = 4 + (int32_t)min(vl, (uint32_t)d.loc[q]);
            } else {
                nm = c;
            }
            d.match[v] = nm;
            d.st[v] = (uint8_t)(nm >= 4 ? 3 : nm);
            if (nm >= 2) d.col[v] = LC_DONE;
        }
        n_mat += act && nm >= 4;
        n_dead += act && nm == 2;
        if (d.trace) {
            const bool keep = act && nm >= 4 && (mm_select_hash(vl, seed ^ MM_LEARN_TRACE_TWEAK) >> 16) < d.trace_keep;
            const unsigned long long mask = __ballot(keep);
            if (mask) {
                const int lane = __lane_id();
                const int leader = __ffsll((long long)mask) - 1;
                unsigned long long slot0 = 0;
                if (lane == leader) slot0 = atomicAdd(d.scount, (unsigned long long)__popcll(mask));
                slot0 = __shfl(slot0, leader);
                const unsigned long long slot = slot0 + __popcll(mask & ((1ull << lane) - 1ull));
                if (keep && slot < (unsigned long long)d.trace_cap) {
#pragma unroll
                    for (int kk = 0; kk < NF; ++kk) d.sx[slot * NF + kk] = d.feat[(size_t)v * NF + kk];
                    d.sy[slot] = c == LC_BLUE ? 1.0f : 0.0f;
                    d.sw[slot] = d.reqw[v];
                    d.sv[slot] = (int32_t)vl;
                    d.sr[slot] = r;
                }
            }
        }
    }
    n_mat = block_sum(n_mat);
    n_dead = block_sum(n_dead);
    if (threadIdx.x == 0) {
        d.part[2 * LPART + blockIdx.x] = n_mat;
        d.part[3 * LPART + blockIdx.x] = n_dead;
    }
}

// ctr[r] = {active, blue, newly matched, newly dead} from the round's per-block counts.
__global__ __launch_bounds__(256 * LCTR) void k_l_reduce(const int32_t *lcnt, const int32_t *part, int nblk,
                                                         int32_t *ctr_r)
{
    // 256 threads per counter, four loads in flight per thread.
    __shared__ int ws[4 * LCTR];
    const int c = threadIdx.x >> 8, t = threadIdx.x & 255;
    const int32_t *src = c < 2 ? lcnt + c * LPART : part + c * LPART;
    int x = 0;
#pragma unroll 4
    for (int b = t; b < nblk; b += 256) x += src[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = x;
    __syncthreads();
    if (t == 0) ctr_r[c] = ws[4 * c] + ws[4 * c + 1] + ws[4 * c + 2] + ws[4 * c + 3];
}

// ---------------------------------------------------------------------------
// Training kernels

constexpr int GB = 64;         // samples per chunk = threads per block (one wavefront)
constexpr int HS = NH + 1;     // padded LDS row stride
constexpr int OWN = (NP + GB - 1) / GB;  // gradient entries owned by each thread

__global__ void k_mlp_forward(const float *__restrict__ P, const float *__restrict__ X, int64_t n, float *z)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float x[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) x[k] = X[i * NF + k];
        z[i] = mlp_z(P, x);
    }
}

// Per-block partial sums of the gradient of (1/n_batch) sum w (softplus(z) - y z)
// over chunks of GB samples (samples idx[i], or i when idx is NULL).
__global__ __launch_bounds__(GB) void k_mlp_grad(const float *__restrict__ P, const float *__restrict__ X,
                                                  const float *__restrict__ Y, const float *__restrict__ Wt,
                                                  const int32_t *__restrict__ idx, int64_t n, int64_t chunks,
                                                  float inv_n, float *partial, double *lossp)
{
    __shared__ float xs[GB][NF + 1];
    __shared__ float h1s[GB][HS];
    __shared__ float d1s[GB][HS];
    __shared__ float h2s[GB][HS];
    __shared__ float d2s[GB][HS];
    __shared__ float dzs[GB];
    __shared__ double ls[GB];
    const int t = threadIdx.x;
    float acc[OWN];
#pragma unroll
    for (int q = 0; q < OWN; ++q) acc[q] = 0.0f;
    double lacc = 0.0;
    for (int64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
        const int64_t i = ch * GB + t;
        const bool valid = i < n;
        const int64_t s = valid ? (idx ? (int64_t)idx[i] : i) : 0;
        float x[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) x[k] = valid ? X[s * NF + k] : 0.0f;
        float h1[NH];
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            float a = P[MM_MLP_OFF_B1 + j];
#pragma unroll
            for (int k = 0; k < NF; ++k) a = fmaf(P[MM_MLP_OFF_W1 + j * NF + k], x[k], a);
            h1[j] = relu(a);
            h1s[t][j] = h1[j];
        }
        float z = P[MM_MLP_OFF_B3];
#pragma unroll 2
        for (int j = 0; j < NH; ++j) {
            float a = P[MM_MLP_OFF_B2 + j];
#pragma unroll
            for (int k = 0; k < NH; ++k) a = fmaf(P[MM_MLP_OFF_W2 + j * NH + k], h1[k], a);
            h2s[t][j] = relu(a);
            z = fmaf(P[MM_MLP_OFF_W3 + j], relu(a), z);
        }
        float dz = 0.0f;
        if (valid) {
            const float y = Y[s], w = Wt[s];
            const float sp = fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z)));
            lacc += (double)(w * (sp - y * z));
            const float sig = 1.0f / (1.0f + expf(-z));
            dz = w * (sig - y) * inv_n;
        }
        dzs[t] = dz;
        // Backward: relu'(a) = [relu(a) > 0].
        float d2[NH];
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            d2[j] = h2s[t][j] > 0.0f ? dz * P[MM_MLP_OFF_W3 + j] : 0.0f;
            d2s[t][j] = d2[j];
        }
#pragma unroll 2
        for (int k = 0; k < NH; ++k) {
            float g = 0.0f;
#pragma unroll
            for (int j = 0; j < NH; ++j) g = fmaf(P[MM_MLP_OFF_W2 + j * NH + k], d2[j], g);
            d1s[t][k] = h1[k] > 0.0f ? g : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < NF; ++k) xs[t][k] = x[k];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < OWN; ++q) {
            const int e = t + GB * q;
            if (e >= NP) break;
            float sacc = 0.0f;
            if (e < MM_MLP_OFF_B1) {
                const int j = e / NF, k = e % NF;
                for (int ii = 0; ii < GB; ++ii) sacc = fmaf(d1s[ii][j], xs[ii][k], sacc);
            } else if (e < MM_MLP_OFF_W2) {
                const int j = e - MM_MLP_OFF_B1;
                for (int ii = 0; ii < GB; ++ii) sacc += d1s[ii][j];
            } else if (e < MM_MLP_OFF_B2) {
                const int j = (e - MM_MLP_OFF_W2) / NH, k = (e - MM_MLP_OFF_W2) % NH;
                for (int ii = 0; ii < GB; ++ii) sacc = fmaf(d2s[ii][j], h1s[ii][k], sacc);
            } else if (e < MM_MLP_OFF_W3) {
                const int j = e - MM_MLP_OFF_B2;
                for (int ii = 0; ii < GB; ++ii) sacc += d2s[ii][j];
            } else if (e < MM_MLP_OFF_B3) {
                const int j = e - MM_MLP_OFF_W3;
                for (int ii = 0; ii < GB; ++ii) sacc = fmaf(dzs[ii], h2s[ii][j], sacc);
            } else {
                for (int ii = 0; ii < GB; ++ii) sacc += dzs[ii];
            }
            acc[q] += sacc;
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < OWN; ++q) {
        const int e = t + GB * q;
        if (e < NP) partial[(size_t)blockIdx.x * NP + e] = acc[q];
    }
    ls[t] = lacc;
    __syncthreads();
    for (int off = GB / 2; off > 0; off >>= 1) {
        if (t < off) ls[t] += ls[t + off];
        __syncthreads();
    }
    if (t == 0) lossp[blockIdx.x] = ls[0];
}

// grad[e] = sum over blocks of partial[b][e] in block order; *loss_acc += sum of lossp.
__global__ void k_mlp_reduce(const float *partial, const double *lossp, int nblk, float *grad, double *loss_acc)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < NP) {
        float s = 0.0f;
        for (int b = 0; b < nblk; ++b) s += partial[(size_t)b * NP + e];
        grad[e] = s;
    }
    if (e == 0) {
        double l = 0.0;
        for (int b = 0; b < nblk; ++b) l += lossp[b];
        *loss_acc += l;
    }
}

__global__ void k_adam(float *P, const float *G, float *M, float *V, float lr_t, float b1, float b2, float eps)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NP) return;
    const float g = G[e];
    const float m = b1 * M[e] + (1.0f - b1) * g;
    const float v = b2 * V[e] + (1.0f - b2) * g * g;
    M[e] = m;
    V[e] = v;
    P[e] -= lr_t * m / (sqrtf(v) + eps);
}

__global__ void k_zero_d(double *p, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Host side

#define ML_TRY(expr, what)                                                                         \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return mm::set_errorf(e_ == hipErrorOutOfMemory ? MM_ERR_NOMEM : MM_ERR_HIP, "%s: %s", \
                                  what, hipGetErrorString(e_));                                    \
    } while (0)

// Device buffers freed on scope exit (entry points that are not tied to a graph).
struct DevBufs {
    std::vector<void *> ptrs;
    ~DevBufs()
    {
        for (void *p : ptrs) hipFree(p);
    }
    template <typename T>
    int32_t alloc(T **p, size_t count, const char *what)
    {
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, sizeof(T) * (count > 0 ? count : 1));
        if (e != hipSuccess) {
            *p = nullptr;
            return mm::set_errorf(MM_ERR_NOMEM, "Not enough memory on device for %s: %s", what, hipGetErrorString(e));
        }
        ptrs.push_back(q);
        *p = (T *)q;
        return MM_OK;
    }
};

template <typename T>
int32_t graph_buf(mm_graph *g, T **p, size_t count, const char *what)
{
    if (*p) return MM_OK;
    void *q = nullptr;
    const int32_t rc = mm_graph_alloc(g, &q, sizeof(T) * (count > 0 ? count : 1), what);
    *p = (T *)q;
    return rc;
}

int32_t set_device(int32_t device)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return mm::set_error(MM_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return mm::set_error(MM_ERR_INVALID, "bad device index");
    ML_TRY(hipSetDevice(device), "hipSetDevice");
    return MM_OK;
}

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int32_t run_learned(mm_graph *g, const mm_opts *o, mm_learn_opts *lo, mm_stats *stats)
{
    if (!o) return mm::set_error(MM_ERR_INVALID, "mm_match_learned: opts is NULL");
    if (o->kind != MM_RANDOM && o->kind != MM_WEIGHTED)
        return mm::set_error(MM_ERR_INVALID, "mm_match_learned: bad kind");
    const int maximal = o->maximal ? 1 : 0;
    const int R_max = o->max_rounds > 0 ? o->max_rounds : (maximal ? MM_NR_MAX_MATCH_ROUNDS : MM_NR_MATCH_ROUNDS);
    if (R_max < 1 || R_max > MM_MAX_ROUNDS)
        return mm::set_error(MM_ERR_INVALID, "mm_match_learned: max_rounds out of range");
    if (!o->seeds || o->n_seeds < 1) return mm::set_error(MM_ERR_SEEDS, "mm_match_learned: no seeds");
    if (!maximal && o->n_seeds < R_max) return mm::set_error(MM_ERR_SEEDS, "mm_match_learned: fewer seeds than rounds");
    const int block = o->block_size > 0 ? o->block_size : 256;
    if (block < 64 || block > 1024 || block % 64)
        return mm::set_error(MM_ERR_INVALID, "mm_match_learned: bad block_size");
    const bool trace = lo && lo->trace;
    if (trace && (lo->trace_cap < 0 || (lo->trace_cap > 0 && (!lo->trace_x || !lo->trace_y || !lo->trace_w))))
        return mm::set_error(MM_ERR_INVALID, "mm_match_learned: trace arrays missing");
    if (trace && lo->trace_keep > 65536u)
        return mm::set_error(MM_ERR_INVALID, "mm_match_learned: trace_keep above 65536");
    const bool use_mlp = lo && lo->params;
    const bool A8 = o->kind == MM_WEIGHTED && g->weighted;
    // Proposals by atomic max where that equals the reference's respond scan
    // (symmetric CSR, ascending rows, symmetric weights), as in mm_match.
    const bool push = (A8 ? g->push_ok_weights : g->push_ok_struct) && g->d_resp != nullptr;
    const size_t nn = (size_t)g->n;

    ML_TRY(hipSetDevice(g->device), "hipSetDevice");
    int32_t rc;
    if ((rc = graph_buf(g, &g->d_lcol, nn, "learned colours")) != MM_OK ||
        (rc = graph_buf(g, &g->d_lreqw, nn, "proposal weights")) != MM_OK ||
        (rc = graph_buf(g, &g->d_lctr, (size_t)LSLOTS * LCTR, "learned counters")) != MM_OK ||
        (rc = graph_buf(g, &g->d_params, (size_t)NP, "network parameters")) != MM_OK ||
        (rc = graph_buf(g, &g->d_lsamp, 2, "trace counters")) != MM_OK)
        return rc;
    if (trace && (rc = graph_buf(g, &g->d_lfeat, nn * NF, "features")) != MM_OK) return rc;
    if ((rc = graph_buf(g, &g->d_lpart, (size_t)LCTR * LPART, "partial counts")) != MM_OK) return rc;
    if ((rc = graph_buf(g, &g->d_lcnt, (size_t)4 * LPART, "list lengths")) != MM_OK) return rc;
    if (!g->h_lctr && hipHostMalloc((void **)&g->h_lctr, sizeof(int32_t) * LSLOTS * LCTR) != hipSuccess)
        return mm::set_error(MM_ERR_NOMEM, "pinned host allocation failed");

    DevBufs tb;
    LDev d;
    std::memset(&d, 0, sizeof(d));
    d.n = g->n;
    d.barrier = o->barrier;
    d.use_mlp = use_mlp ? 1 : 0;
    d.trace = trace ? 1 : 0;
    d.trace_keep = trace ? lo->trace_keep : 0u;
    d.trace_cap = trace ? lo->trace_cap : 0;
    d.ranges = g->d_ranges;
    d.nbrs = g->d_nbrs;
    d.adj = g->d_adj;
    d.loc = g->d_loc;
    d.match = g->d_match;
    d.req = g->d_req;
    d.col = g->d_lcol;
    d.st = g->d_st;
    d.reqw = g->d_lreqw;
    d.feat = g->d_lfeat;
    d.ctr = g->d_lctr;
    d.part = g->d_lpart;
    d.scount = g->d_lsamp;
    d.resp = g->d_resp;
    if (trace) {
        const size_t cap = (size_t)lo->trace_cap;
        if ((rc = tb.alloc(&d.sx, cap * NF, "trace features")) != MM_OK || (rc = tb.alloc(&d.sy, cap, "trace")) != MM_OK ||
            (rc = tb.alloc(&d.sw, cap, "trace")) != MM_OK || (rc = tb.alloc(&d.sv, cap, "trace")) != MM_OK ||
            (rc = tb.alloc(&d.sr, cap, "trace")) != MM_OK)
            return rc;
    }
    hipStream_t st = g->stream;
    if (use_mlp)
        ML_TRY(hipMemcpyAsync(g->d_params, lo->params, sizeof(float) * NP, hipMemcpyHostToDevice, st),
               "Unable to set network parameters");

    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>({((int64_t)g->n + block - 1) / block,
                                                                   (int64_t)g->num_cu * (2048 / block), LPART}));
    // Every round kernel runs sgrid blocks; block b's tiles feed its own list segments.
    const int64_t tiles = ((int64_t)g->n + STILE - 1) / STILE;
    const int sgrid = (int)std::max<int64_t>(1, std::min<int64_t>({tiles, (int64_t)g->num_cu * 8, LPART}));
    d.seg = (tiles + sgrid - 1) / sgrid * STILE;
    // Lists: active (by round parity, the select reads the previous round's)
    // and blue; list lengths likewise by round parity.
    if ((size_t)(3 * sgrid * d.seg) > g->llist_cap) {
        if (g->d_llist) (void)hipFree(g->d_llist);
        g->d_llist = nullptr;
        g->llist_cap = 0;
        void *q = nullptr;
        if ((rc = mm_graph_alloc(g, &q, sizeof(int32_t) * 3 * (size_t)sgrid * (size_t)d.seg, "learned work lists")) !=
            MM_OK)
            return rc;
        g->d_llist = (int32_t *)q;
        g->llist_cap = (size_t)(3 * sgrid * d.seg);
    }
    int32_t *act_list[2] = {g->d_llist, g->d_llist + (size_t)sgrid * d.seg};
    int32_t *cnt_list[2] = {g->d_lcnt, g->d_lcnt + 2 * LPART};
    d.lblue = g->d_llist + 2 * (size_t)sgrid * d.seg;
    ML_TRY(hipEventRecord(g->ev_t1, st), "hipEventRecord");
    hipLaunchKernelGGL(k_l_fill, dim3(1), dim3(256), 0, st, d.ctr, (int64_t)LSLOTS * LCTR, 0);
    hipLaunchKernelGGL(k_l_fill, dim3(1), dim3(64), 0, st, (int32_t *)d.scount, (int64_t)4, 0);
    if (g->n > 0) hipLaunchKernelGGL(k_l_start, dim3(grid), dim3(block), 0, st, d.match, d.st, (int64_t)g->n);
    const int R_run = maximal ? std::min(R_max, o->n_seeds) : R_max;
    // Maximal runs read the counters back every 4 rounds: an empty learned round
    // still sweeps the select (about as costly as a read-back), so the plain
    // rounds' longer predicted chunks do not pay off here.
    const int chunk = maximal ? 4 : R_run;
    int rounds = -1;
    for (int r = 1; r <= R_run && g->n > 0;) {
        const int end = std::min(R_run, r + chunk - 1);
        for (int rr = r; rr <= end; ++rr) {
            const uint32_t seed = o->seeds[rr - 1];
            d.lact = act_list[rr & 1];
            d.lcnt = cnt_list[rr & 1];
            d.lprev = act_list[(rr - 1) & 1];
            d.lcnt_prev = cnt_list[(rr - 1) & 1];
            const float *P = (const float *)g->d_params;
#define ML_ROUND(A8_, PUSH_)                                                                                   \
    do {                                                                                                       \
        hipLaunchKernelGGL((k_l_select<A8_, PUSH_, SEL_G>), dim3(sgrid), dim3(LTILE), 0, st, d, rr, seed, P); \
        hipLaunchKernelGGL((k_l_request<A8_, PUSH_, SCAN_G>), dim3(sgrid), dim3(LTILE), 0, st, d);            \
        if (!PUSH_) hipLaunchKernelGGL((k_l_respond<A8_, SCAN_G>), dim3(sgrid), dim3(LTILE), 0, st, d);        \
        hipLaunchKernelGGL((k_l_match<PUSH_>), dim3(sgrid), dim3(LTILE), 0, st, d, rr, seed);                 \
    } while (0)
            if (A8 && push) ML_ROUND(true, true);
            else if (A8) ML_ROUND(true, false);
            else if (push) ML_ROUND(false, true);
            else ML_ROUND(false, false);
#undef ML_ROUND
            hipLaunchKernelGGL(k_l_reduce, dim3(1), dim3(256 * LCTR), 0, st, (const int32_t *)d.lcnt,
                               (const int32_t *)d.part, sgrid, d.ctr + (size_t)rr * LCTR);
        }
        ML_TRY(hipGetLastError(), "A HIP error occurred during the matching process");
        if (maximal) {
            ML_TRY(hipMemcpyAsync(g->h_lctr + (size_t)r * LCTR, d.ctr + (size_t)r * LCTR,
                                  sizeof(int32_t) * (size_t)(end - r + 1) * LCTR, hipMemcpyDeviceToHost, st),
                   "Unable to read round counters");
            ML_TRY(hipStreamSynchronize(st), "A HIP error occurred during the matching process");
            for (int k = r; k <= end; ++k)
                if (g->h_lctr[(size_t)k * LCTR] == 0) {
                    rounds = k;
                    break;
                }
            if (rounds > 0) break;
        }
        r = end + 1;
    }
    if (maximal && rounds < 0 && R_run < R_max && g->n > 0)
        return mm::set_error(MM_ERR_SEEDS, "mm_match_learned: ran out of seeds before the matching converged");
    if (g->n == 0) rounds = maximal ? 1 : R_max;
    if (rounds < 0) rounds = R_max;
    mm_graph_enqueue_unpermute(g);
    g->internal_stale = false;
    ML_TRY(hipEventRecord(g->ev_t2, st), "hipEventRecord");
    ML_TRY(hipMemcpyAsync(g->h_lctr, d.ctr, sizeof(int32_t) * LSLOTS * LCTR, hipMemcpyDeviceToHost, st),
           "Unable to read round counters");
    unsigned long long offered = 0;
    ML_TRY(hipMemcpyAsync(&offered, d.scount, sizeof(offered), hipMemcpyDeviceToHost, st), "Unable to read trace");
    ML_TRY(hipStreamSynchronize(st), "A HIP error occurred during the matching process");
    ML_TRY(hipGetLastError(), "A HIP error occurred during the matching process");

    if (trace) {
        const int64_t count = std::min<int64_t>((int64_t)offered, lo->trace_cap);
        std::vector<float> x((size_t)count * NF), y((size_t)count), w((size_t)count);
        std::vector<int32_t> v((size_t)count), rr((size_t)count);
        if (count > 0) {
            ML_TRY(hipMemcpy(x.data(), d.sx, sizeof(float) * x.size(), hipMemcpyDeviceToHost), "Unable to read trace");
            ML_TRY(hipMemcpy(y.data(), d.sy, sizeof(float) * y.size(), hipMemcpyDeviceToHost), "Unable to read trace");
            ML_TRY(hipMemcpy(w.data(), d.sw, sizeof(float) * w.size(), hipMemcpyDeviceToHost), "Unable to read trace");
            ML_TRY(hipMemcpy(v.data(), d.sv, sizeof(int32_t) * v.size(), hipMemcpyDeviceToHost), "Unable to read trace");
            ML_TRY(hipMemcpy(rr.data(), d.sr, sizeof(int32_t) * rr.size(), hipMemcpyDeviceToHost), "Unable to read trace");
        }
        // (round, graph-local id, position) order: samples of a batch's graphs share ids, the position keeps it total.
        std::vector<int64_t> ord((size_t)count);
        std::iota(ord.begin(), ord.end(), (int64_t)0);
        std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
            if (rr[(size_t)a] != rr[(size_t)b]) return rr[(size_t)a] < rr[(size_t)b];
            if (v[(size_t)a] != v[(size_t)b]) return v[(size_t)a] < v[(size_t)b];
            return a < b;
        });
        for (int64_t k = 0; k < count; ++k) {
            const size_t s = (size_t)ord[(size_t)k];
            std::memcpy(lo->trace_x + k * NF, &x[s * NF], sizeof(float) * NF);
            lo->trace_y[k] = y[s];
            lo->trace_w[k] = w[s];
            if (lo->trace_v) lo->trace_v[k] = v[s];
            if (lo->trace_round) lo->trace_round[k] = rr[s];
        }
        lo->trace_count = count;
        lo->trace_offered = (int64_t)offered;
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->rounds = rounds;
        stats->respond_used = MM_RESPOND_PULL;
        stats->lanes_used = 1;
        float ms = 0.0f;
        hipEventElapsedTime(&ms, g->ev_t1, g->ev_t2);
        stats->match_ms = ms;
        int64_t cum = 0;
        for (int k = 1; k <= rounds && k <= MM_MAX_ROUNDS; ++k) {
            const int32_t *q = g->h_lctr + (size_t)k * LCTR;
            cum += q[2];
            stats->active[k - 1] = q[0];
            stats->blue[k - 1] = q[1];
            stats->matched[k - 1] = cum;
        }
    }
    return MM_OK;
}

}  // namespace

extern "C" {

int32_t mm_match_learned(mm_graph *g, const mm_opts *opts, mm_learn_opts *lopts, int32_t *match_out,
                         mm_stats *stats)
{
    mm::clear_error();
    if (!g) return mm::set_error(MM_ERR_INVALID, "mm_match_learned: graph is NULL");
    const int32_t rc = run_learned(g, opts, lopts, stats);
    if (rc != MM_OK) return rc;
    if (match_out) return mm_graph_copy_match(g, match_out);
    return MM_OK;
}

int32_t mm_mlp_init(uint64_t seed, float *params_out)
{
    if (!params_out) return mm::set_error(MM_ERR_INVALID, "mm_mlp_init: NULL output");
    uint64_t s = seed;
    auto uni = [&](float bound) {
        const double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
        return (float)((2.0 * u - 1.0) * bound);
    };
    std::memset(params_out, 0, sizeof(float) * NP);
    const float a1 = std::sqrt(6.0f / NF), a2 = std::sqrt(6.0f / NH);
    for (int e = MM_MLP_OFF_W1; e < MM_MLP_OFF_B1; ++e) params_out[e] = uni(a1);
    for (int e = MM_MLP_OFF_W2; e < MM_MLP_OFF_B2; ++e) params_out[e] = uni(a2);
    for (int e = MM_MLP_OFF_W3; e < MM_MLP_OFF_B3; ++e) params_out[e] = uni(a2);
    return MM_OK;
}

int32_t mm_mlp_forward(int32_t device, const float *params, const float *x, int64_t n, float *z_out)
{
    mm::clear_error();
    if (!params || n < 0 || (n > 0 && (!x || !z_out))) return mm::set_error(MM_ERR_INVALID, "mm_mlp_forward: bad arguments");
    int32_t rc = set_device(device);
    if (rc != MM_OK) return rc;
    if (n == 0) return MM_OK;
    DevBufs b;
    float *dp, *dx, *dz;
    if ((rc = b.alloc(&dp, NP, "parameters")) != MM_OK || (rc = b.alloc(&dx, (size_t)n * NF, "features")) != MM_OK ||
        (rc = b.alloc(&dz, (size_t)n, "outputs")) != MM_OK)
        return rc;
    ML_TRY(hipMemcpy(dp, params, sizeof(float) * NP, hipMemcpyHostToDevice), "upload");
    ML_TRY(hipMemcpy(dx, x, sizeof(float) * (size_t)n * NF, hipMemcpyHostToDevice), "upload");
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(256), 0, 0, dp, dx, n, dz);
    ML_TRY(hipGetLastError(), "k_mlp_forward");
    ML_TRY(hipMemcpy(z_out, dz, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost), "download");
    return MM_OK;
}

int32_t mm_mlp_loss_grad(int32_t device, const float *params, const float *x, const float *y, const float *w,
                         int64_t n, double *loss_out, float *grad_out)
{
    mm::clear_error();
    if (!params || n <= 0 || !x || !y || !w || !grad_out)
        return mm::set_error(MM_ERR_INVALID, "mm_mlp_loss_grad: bad arguments");
    int32_t rc = set_device(device);
    if (rc != MM_OK) return rc;
    const int64_t chunks = (n + GB - 1) / GB;
    const int nblk = (int)std::min<int64_t>(chunks, 1024);
    DevBufs b;
    float *dp, *dx, *dy, *dw, *part, *dg;
    double *lossp, *dl;
    if ((rc = b.alloc(&dp, NP, "parameters")) != MM_OK || (rc = b.alloc(&dx, (size_t)n * NF, "features")) != MM_OK ||
        (rc = b.alloc(&dy, (size_t)n, "labels")) != MM_OK || (rc = b.alloc(&dw, (size_t)n, "weights")) != MM_OK ||
        (rc = b.alloc(&part, (size_t)nblk * NP, "partial gradients")) != MM_OK ||
        (rc = b.alloc(&dg, NP, "gradient")) != MM_OK || (rc = b.alloc(&lossp, (size_t)nblk, "loss")) != MM_OK ||
        (rc = b.alloc(&dl, 1, "loss")) != MM_OK)
        return rc;
    ML_TRY(hipMemcpy(dp, params, sizeof(float) * NP, hipMemcpyHostToDevice), "upload");
    ML_TRY(hipMemcpy(dx, x, sizeof(float) * (size_t)n * NF, hipMemcpyHostToDevice), "upload");
    ML_TRY(hipMemcpy(dy, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice), "upload");
    ML_TRY(hipMemcpy(dw, w, sizeof(float) * (size_t)n, hipMemcpyHostToDevice), "upload");
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, 0, dl, 1);
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(GB), 0, 0, dp, dx, dy, dw, (const int32_t *)nullptr, n, chunks,
                       (float)(1.0 / (double)n), part, lossp);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NP + 255) / 256), dim3(256), 0, 0, part, lossp, nblk, dg, dl);
    ML_TRY(hipGetLastError(), "k_mlp_grad");
    double lsum = 0.0;
    ML_TRY(hipMemcpy(grad_out, dg, sizeof(float) * NP, hipMemcpyDeviceToHost), "download");
    ML_TRY(hipMemcpy(&lsum, dl, sizeof(double), hipMemcpyDeviceToHost), "download");
    if (loss_out) *loss_out = lsum / (double)n;
    return MM_OK;
}

int32_t mm_mlp_train(int32_t device, const float *x, const float *y, const float *w, int64_t n,
                     const mm_train_opts *t, float *params_inout, float *epoch_loss_out)
{
    mm::clear_error();
    if (!t || !params_inout || n <= 0 || !x || !y || !w || t->epochs < 1 || t->batch < 1 || n > INT32_MAX)
        return mm::set_error(MM_ERR_INVALID, "mm_mlp_train: bad arguments");
    int32_t rc = set_device(device);
    if (rc != MM_OK) return rc;
    const int64_t B = t->batch;
    const int nblk_max = (int)std::min<int64_t>((B + GB - 1) / GB, 1024);
    DevBufs b;
    float *dp, *dx, *dy, *dw, *part, *dg, *dm, *dv;
    double *lossp, *dl;
    int32_t *didx;
    if ((rc = b.alloc(&dp, NP, "parameters")) != MM_OK || (rc = b.alloc(&dx, (size_t)n * NF, "features")) != MM_OK ||
        (rc = b.alloc(&dy, (size_t)n, "labels")) != MM_OK || (rc = b.alloc(&dw, (size_t)n, "weights")) != MM_OK ||
        (rc = b.alloc(&part, (size_t)nblk_max * NP, "partial gradients")) != MM_OK ||
        (rc = b.alloc(&dg, NP, "gradient")) != MM_OK || (rc = b.alloc(&dm, NP, "moments")) != MM_OK ||
        (rc = b.alloc(&dv, NP, "moments")) != MM_OK || (rc = b.alloc(&lossp, (size_t)nblk_max, "loss")) != MM_OK ||
        (rc = b.alloc(&dl, (size_t)t->epochs, "loss")) != MM_OK || (rc = b.alloc(&didx, (size_t)n, "order")) != MM_OK)
        return rc;
    ML_TRY(hipMemcpy(dp, params_inout, sizeof(float) * NP, hipMemcpyHostToDevice), "upload");
    ML_TRY(hipMemcpy(dx, x, sizeof(float) * (size_t)n * NF, hipMemcpyHostToDevice), "upload");
    ML_TRY(hipMemcpy(dy, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice), "upload");
    ML_TRY(hipMemcpy(dw, w, sizeof(float) * (size_t)n, hipMemcpyHostToDevice), "upload");
    ML_TRY(hipMemset(dm, 0, sizeof(float) * NP), "clear");
    ML_TRY(hipMemset(dv, 0, sizeof(float) * NP), "clear");
    hipLaunchKernelGGL(k_zero_d, dim3((t->epochs + 63) / 64), dim3(64), 0, 0, dl, t->epochs);
    const float b1 = t->beta1 > 0.0f ? t->beta1 : 0.9f, b2 = t->beta2 > 0.0f ? t->beta2 : 0.999f;
    const float eps = t->eps > 0.0f ? t->eps : 1e-8f, lr = t->lr > 0.0f ? t->lr : 1e-3f;
    std::vector<int32_t> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    uint64_t s = t->seed;
    int64_t step = 0;
    for (int ep = 0; ep < t->epochs; ++ep) {
        for (int64_t i = n - 1; i > 0; --i) std::swap(order[(size_t)i], order[(size_t)(splitmix64(s) % (uint64_t)(i + 1))]);
        ML_TRY(hipMemcpy(didx, order.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice), "upload");
        for (int64_t lo = 0; lo < n; lo += B) {
            const int64_t cnt = std::min(B, n - lo);
            const int64_t chunks = (cnt + GB - 1) / GB;
            const int nblk = (int)std::min<int64_t>(chunks, nblk_max);
            hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(GB), 0, 0, dp, dx, dy, dw, didx + lo, cnt, chunks,
                               (float)(1.0 / (double)cnt), part, lossp);
            hipLaunchKernelGGL(k_mlp_reduce, dim3((NP + 255) / 256), dim3(256), 0, 0, part, lossp, nblk, dg, dl + ep);
            ++step;
            const double c1 = 1.0 - std::pow((double)b1, (double)step), c2 = 1.0 - std::pow((double)b2, (double)step);
            hipLaunchKernelGGL(k_adam, dim3((NP + 255) / 256), dim3(256), 0, 0, dp, dg, dm, dv,
                               (float)(lr * std::sqrt(c2) / c1), b1, b2, eps);
        }
        ML_TRY(hipGetLastError(), "training kernels");
    }
    std::vector<double> el((size_t)t->epochs);
    ML_TRY(hipMemcpy(params_inout, dp, sizeof(float) * NP, hipMemcpyDeviceToHost), "download");
    ML_TRY(hipMemcpy(el.data(), dl, sizeof(double) * el.size(), hipMemcpyDeviceToHost), "download");
    if (epoch_loss_out)
        for (int ep = 0; ep < t->epochs; ++ep) epoch_loss_out[ep] = (float)(el[(size_t)ep] / (double)n);
    return MM_OK;
}

}  // extern "C"
