// MI355X (gfx950) kernels and C ABI of the fine-grained greedy matcher.
//
// One matching round of the reference (src/matchgpu.cu:184-451) has four
// phases: hash-seeded blue/red select, blue->red request, red response and the
// match write-back.  Here a round is two or three launches:
//
//   request   blue vertices scan their weight-sorted rows and pick the first
//             red neighbour in row order: the heaviest, first in CSR order on
//             ties (gwRequest, src/matchgpu.cu:360-409), or the first red one
//             of a CSR-order row (grRequest, 286-327).  Round 1 gives each lane
//             one vertex of its bucket's id range and reads a dense copy of the
//             rows' first 4 entries; later rounds give G lanes to each entry of
//             a work list (reading a packed 4-byte copy of the rows when the
//             graph has one), and rows longer than the hub degree a whole
//             block.  The list rounds' requests are stored by list slot, so
//             the match loads them together with the list entries.
//             PUSH: the proposal is also posted to the red vertex with one
//             atomic (32-bit min on the proposer's position in the red
//             vertex's sorted row, or 64-bit max on (weight, -proposer)),
//             which replaces the respond scan.  In round 1 a proposal to a
//             vertex of the proposer's own bucket meets the others in the
//             block's LDS copy of the bucket's keys, stored once at the end
//             (most rows' heaviest red neighbour lies in their own bucket).
//   respond   (PULL only) red vertices scan their rows for proposers
//             (gwRespond/grRespond, src/matchgpu.cu:329-357, 411-451).
//   match+select  every vertex active in this round resolves its pair
//             (gMatch, src/matchgpu.cu:260-283); settled vertices store their
//             final value straight into the result in the caller's order, and
//             the ones still active are recoloured with the next round's seed
//             (gSelect, 184-218) and appended to the next round's work lists.
//
// Vertices without neighbours, when they are at least 1% of a graph, are
// settled by the round-1 select itself (their whole run follows from their
// ids and the seeds) and never enter a work list.
//
// Only active vertices are touched after round 1: they live in four compacted
// lists per round (blue/red x row length below or above the hub degree),
// filled with one LDS atomic per wavefront.  Neighbour colours are read from a
// 1-byte state mirror (st[]), which keeps the random gathers of a 2^22-vertex
// graph inside L2/MALL.  Every phase writes only slots that no thread of the
// same phase reads (the processed vertex's own state, request and list slots;
// proposal keys only through atomics, global or on the block's LDS copy;
// result slots, which no phase reads,
// once per settled vertex), so the result is independent of the schedule and
// equals the serial statement of the rounds bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <climits>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/mm_match.h"
#include "mm_graph_impl.h"
#include "mm_hash.h"
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
    int32_t nb;           // buckets = blocks of the per-bucket kernels
    int32_t bcap;         // capacity of one bucket; bucket b owns ids [b*bcap, (b+1)*bcap) in round 1
    const int2 *ranges;
    const int32_t *nbrs;  // unweighted layout
    const int2 *adj;      // weighted rows sorted for matching: {neighbour id (~id when the
                          // weight can never win), weight bits}
    const int2 *head;     // entries 0..3 of every adj row, 4 per vertex (null: none kept)
    const uint32_t *headc;// the same in 4 bytes per entry: invalid << 31 | id << head_pb | position
                          // (graphs with 32-bit keys whose ids and positions fit; null otherwise)
    int32_t head_pb;
    const uint32_t *adjc; // every entry of the sorted rows packed as headc packs them (or null)
    int32_t *reqs;        // PUSH list rounds: the bucket lists' requests by list slot (bucket b's at
                          // b * bcap), so the match loads them with the entries; null: by vertex (req)
    int32_t *out;         // result in the caller's order, written as vertices settle
    const int32_t *vbase; // batches: first caller id of each vertex's graph (null: single graph)
    int32_t *req;
    uint8_t *st;
    unsigned long long *resp;  // [2][n], PUSH only; with key32 the same memory holds uint32 [2][n]
    int32_t key32;             // PUSH keys are 32-bit: (position of the proposer in the red vertex's
                               // sorted row) << locbits | proposer's local id; the scanned rows then
                               // carry those positions where the weights would be
    int32_t locbits;
    int32_t listless1;         // round 1 runs over the buckets' id ranges: no round-1 lists of short rows
    int32_t no_hubs;           // no row is longer than the hub degree
    int32_t split1;            // round 1 (32-bit keys, head scans): a proposal to a vertex of the
                               // proposer's own bucket goes to the block's LDS copy of the bucket's
                               // keys, stored afterwards to bucket_keys(); the others take atomics
                               // on the round's keys; a key is the smaller of the two
    int4 *lists;               // [2][NLISTS][cap] entries {v, row begin, row end, loc[v]};
                               // classes 0 and 2 split into nb buckets of bcap
    int32_t *ctr;              // [ROUND_SLOTS][CTR_STRIDE] totals; [1],[3] = hub list sizes (atomics)
    int32_t *ctrb;             // [ROUND_SLOTS][4][nb] per-bucket: blue-lo, red-lo, matched, dead
    unsigned long long *ctr64; // [ROUND_SLOTS][CTR64_STRIDE] degree sums (profile)
    unsigned long long *ctrb64;// [ROUND_SLOTS][2][nb] per-bucket degree sums (profile)
    unsigned long long *hub64; // [ROUND_SLOTS][2] degree sums of hub rows (profile)
    int32_t *ctr_next;         // the other counter set, cleared here for the next run
    unsigned long long *hub64_next;
    const uint32_t *seeds;     // seeds[r - 1] drives the select of round r
    const int32_t *loc;        // graph-local ORIGINAL id of each (internally renumbered) vertex:
                               // the hash input, the proposal tie-break id and the match value base
    uint8_t *vround;           // batched: round in which the vertex stopped being active (255: never)
    int32_t *iso;              // vertices without neighbours settled by the round-1 select (null: they
                               // go through the rounds): per round r and bucket b, iso[(2r + c) nb + b]
                               // of them were blue (c = 0) / red (c = 1) in round r; iso[2 ROUND_SLOTS nb
                               // + b] = the last such round of bucket b
    int32_t R_run;             // rounds this run may execute
    int32_t n1, bcap1;         // the id range [0, n1) that round 1's request and match scan (without
                               // lists), bcap1 ids per bucket: all ids, or those before the isolated
                               // vertices the select settles (Dev::iso, a single graph that keeps
                               // them last)
};

// Work-list entry: the vertex, its row range and its original id travel
// together, so the scans start without dependent gathers.
__device__ __forceinline__ int4 load_ent(const int4 *p)
{
    typedef int v4i __attribute__((ext_vector_type(4)));
    const v4i e = __builtin_nontemporal_load(reinterpret_cast<const v4i *>(p));
    return make_int4(e.x, e.y, e.z, e.w);
}

__device__ __forceinline__ int4 *list_ptr(const Dev &d, int parity, int k)
{
    return d.lists + (size_t)(parity * NLISTS + k) * (size_t)d.cap;
}

// Bucket owned by this block.  Workgroups are dealt to the 8 XCDs round
// robin (workgroup i runs on XCD i % 8), so block i takes bucket
// (i % 8) * (nb / 8) + i / 8: every XCD owns one contiguous eighth of the
// vertex ids, and the neighbour-state gathers of its blocks stay inside its
// own L2.  nb is a multiple of 8.
__device__ __forceinline__ int bucket_of_block(const Dev &d)
{
    const int i = blockIdx.x;
    return (i & 7) * (d.nb >> 3) + (i >> 3);
}

__device__ __forceinline__ uint32_t select_hash(uint32_t i, uint32_t seed) { return mm_select_hash(i, seed); }

// Result index base of internal vertex v: its graph's first caller id
// (batches; 0 for a single graph).  v's value goes to out[out_base + loc[v]].
__device__ __forceinline__ int32_t out_base(const Dev &d, int32_t v) { return d.vbase ? d.vbase[v] : 0; }

__device__ __forceinline__ int colour_of(uint32_t local_id, uint32_t seed, uint32_t barrier)
{
    return select_hash(local_id, seed) < barrier ? ST_BLUE : ST_RED;
}

// Appends value to list for every lane with pred set: one atomic per wavefront.
// Must be reached by all lanes of the wavefront.
__device__ __forceinline__ void wave_append(bool pred, int4 value, int4 *list, int32_t *counter)
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

// Row entry formats (template parameter A8): 0 ids only (d.nbrs, unweighted
// rows), 1 {id, weight bits or key position} pairs (d.adj), 2 the pairs with
// key positions packed in 4 bytes, invalid << 31 | id << head_pb | position
// (d.adjc).
__device__ __forceinline__ void unpack_entry(const Dev &d, uint32_t x, int32_t &u, bool &valid, float &w)
{
    valid = (x >> 31) == 0u;
    u = (int32_t)((x >> d.head_pb) & ((1u << (31 - d.head_pb)) - 1u));
    w = __int_as_float((int32_t)(x & ((1u << d.head_pb) - 1u)));  // (the key position)
}

// Neighbour id of row entry j; valid is false for a weight that can never
// win (sorted weighted rows store those ids complemented).  Rows are streamed
// once per round: non-temporal loads, so they do not evict the gathered
// state arrays from L2.
template <int A8>
__device__ __forceinline__ void load_id(const Dev &d, int32_t j, int32_t &u, bool &valid, float &w)
{
    if (A8 == 2) {
        unpack_entry(d, __builtin_nontemporal_load(d.adjc + j), u, valid, w);
    } else if (A8) {
        typedef int v2i __attribute__((ext_vector_type(2)));
        const v2i e = __builtin_nontemporal_load(reinterpret_cast<const v2i *>(d.adj + j));
        valid = e.x >= 0;
        u = valid ? e.x : ~e.x;
        w = valid ? __int_as_float(e.y) : -2.0f;
    } else {
        u = __builtin_nontemporal_load(d.nbrs + j);
        valid = true;
        w = 1.0f;
    }
}

// Raw row entry j ({id, weight bits} weighted, {id, 0} otherwise), decoded
// only where it is used: a load whose value is not read yet costs no wait.
template <int A8>
__device__ __forceinline__ int2 load_raw(const Dev &d, int32_t j)
{
    typedef int v2i __attribute__((ext_vector_type(2)));
    if (A8 == 2) return make_int2((int32_t)__builtin_nontemporal_load(d.adjc + j), 0);
    if (A8) {
        const v2i e = __builtin_nontemporal_load(reinterpret_cast<const v2i *>(d.adj + j));
        return make_int2(e.x, e.y);
    }
    return make_int2(__builtin_nontemporal_load(d.nbrs + j), 0);
}

template <int A8>
__device__ __forceinline__ void decode_raw(const Dev &d, int2 e, int32_t &u, bool &valid, float &w)
{
    if (A8 == 2) {
        unpack_entry(d, (uint32_t)e.x, u, valid, w);
    } else if (A8) {
        valid = e.x >= 0;
        u = valid ? e.x : ~e.x;
        w = valid ? __int_as_float(e.y) : -2.0f;
    } else {
        u = e.x;
        valid = true;
        w = 1.0f;
    }
}

// First row step of list entry e for lane gl: entry e.y + gl of the row.
// Always a valid address -- clamped into the row, or entry 0 for an empty row
// (the caller masks lanes past the row end) -- so the load can be issued
// unconditionally: a load under a branch makes the compiler wait for every
// load in flight at the first use after it.
template <int A8>
__device__ __forceinline__ int2 first_raw(const Dev &d, const int4 &e, int gl)
{
    const int32_t j = e.y + gl < e.z ? e.y + gl : (e.z > e.y ? e.z - 1 : 0);
    return load_raw<A8>(d, j);
}

template <int A8>
__device__ __forceinline__ void load_entry(const Dev &d, int32_t j, int32_t &u, float &w)
{
    bool valid;
    load_id<A8>(d, j, u, valid, w);
}

// ---------------------------------------------------------------------------
// First-candidate scans.  Rows are stored in PRIORITY ORDER: for weighted
// matching sorted by (weight descending, CSR position ascending) with the
// weights that can never win (NaN, <= -1) last; for random matching in CSR
// order.  The reference's choice -- the heaviest red neighbour, first in CSR
// order on ties (gwRequest, src/matchgpu.cu:360-409), or the first red one
// (grRequest, 286-327); likewise for proposers in gwRespond/grRespond -- is
// then the FIRST candidate in row order, and a scan stops there.  In round 1
// that is about two entries into a row instead of all of it.
// ---------------------------------------------------------------------------

// Candidate test of entry (u, w) of v's row; alive is set for the request
// scan's dead test (an unmatched neighbour).
template <bool W, bool RESPOND>
__device__ __forceinline__ bool first_candidate(const Dev &d, int32_t v, int32_t u, bool valid, bool &alive)
{
    const uint8_t s = d.st[u];
    bool cand;
    if (RESPOND) {
        cand = s == ST_BLUE && d.req[u] == v;
    } else {
        alive = alive || s != ST_MATCHED;
        cand = s == ST_RED;
    }
    if (W) cand = cand && valid;
    return cand;
}

// first_candidate on an already gathered state s = st[u] (and, RESPOND,
// q = req[u]).
template <bool W, bool RESPOND>
__device__ __forceinline__ bool candidate_of(int32_t v, uint8_t s, int32_t q, bool valid, bool &alive)
{
    bool cand;
    if (RESPOND) {
        cand = s == ST_BLUE && q == v;
    } else {
        alive = alive || s != ST_MATCHED;
        cand = s == ST_RED;
    }
    if (W) cand = cand && valid;
    return cand;
}

// 64-bit proposal key: order-preserving image of the weight (-0 folded onto +0,
// so equal weights tie) above the complement of the proposer's id (smaller id
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

// Proposal keys of round parity p (64-bit, or 32-bit in the same memory).
__device__ __forceinline__ unsigned long long *resp_slot(const Dev &d, int p)
{
    if (d.key32) return reinterpret_cast<unsigned long long *>(reinterpret_cast<uint32_t *>(d.resp) + (size_t)p * d.n);
    return d.resp + (size_t)p * d.n;
}
// Empty key: 0 for the 64-bit max, all ones for the 32-bit min.
__device__ __forceinline__ void clear_key(const Dev &d, unsigned long long *slot, int32_t v)
{
    if (d.key32) reinterpret_cast<uint32_t *>(slot)[v] = 0xFFFFFFFFu;
    else slot[v] = 0ull;
}
// Round 1's keys of proposals made inside a bucket (split1): the third n-entry
// quarter of the key memory, which 32-bit keys leave unused.
__device__ __forceinline__ uint32_t *bucket_keys(const Dev &d)
{
    return reinterpret_cast<uint32_t *>(d.resp) + 2 * (size_t)d.n;
}
// Key at vertex v, widened: 32-bit keys come back with the empty key as 0 and
// the proposer's local id, complemented, in the low half -- the form of the
// 64-bit keys, which partner_of decodes.
// split: round 1 with split1, whose key is also in bucket_keys().
__device__ __forceinline__ unsigned long long key_at(const Dev &d, const unsigned long long *slot, int32_t v,
                                                     bool split = false)
{
    if (!d.key32) return slot[v];
    uint32_t k = reinterpret_cast<const uint32_t *>(slot)[v];
    if (split) k = min(k, bucket_keys(d)[v]);
    if (k == 0xFFFFFFFFu) return 0ull;
    return (1ull << 32) | (unsigned long long)(0xFFFFFFFFu - (k & ((1u << d.locbits) - 1u)));
}

// Writes v's request (or, RESPOND, its accepted proposer) for winner entry
// (u, w) or none (u < 0).  s_key (32-bit keys only): the block's LDS keys of
// the ids [klo, khi), which take the proposals to those ids.
template <bool W, bool PUSH, bool RESPOND>
__device__ __forceinline__ void first_write(const Dev &d, int32_t v, int32_t vloc, int32_t u, float w, bool alive,
                                            unsigned long long *resp, uint32_t *s_key = nullptr, int64_t klo = 0,
                                            int64_t khi = 0, int32_t *rqa = nullptr)
{
    int32_t *const rp = rqa ? rqa : d.req + v;  // (rqa: the request's list slot)
    if (u >= 0) {
        *rp = u;
        if (!RESPOND && PUSH) {
            if (d.key32) {  // (w holds the position bits of the winning entry)
                const uint32_t key = ((uint32_t)__float_as_int(w) << d.locbits) | (uint32_t)vloc;
                if (s_key && u >= klo && u < khi) atomicMin(s_key + (u - klo), key);
                else atomicMin(reinterpret_cast<uint32_t *>(resp) + u, key);
            } else
                atomicMax(resp + u, proposal_key<W>(w, vloc));
        }
    } else {
        *rp = (RESPOND || alive) ? REQ_NONE : REQ_DEAD;
    }
}

// Steps 1, 2, ... of a row scan (entries from base on), G lanes per row; the
// group stops at its first candidate.  Wave-uniform control flow: every lane
// of the wavefront must call it (done lanes idle).
template <bool W, int A8, bool RESPOND, int G>
__device__ __forceinline__ void scan_rest(const Dev &d, int32_t v, int2 rg, int32_t base, bool done, bool &alive,
                                          int32_t &win_u, float &win_w)
{
    const int lane = threadIdx.x & 63, gl = lane & (G - 1), gw = lane / G;
    const unsigned long long gmask = G == 64 ? ~0ull : (((1ull << G) - 1ull) << (gw * G));
    if (base >= rg.y) done = true;
    while (__ballot(!done) != 0ull) {
        const int32_t j = base + gl;
        int32_t u = 0;
        float w = 1.0f;
        bool cand = false;
        if (!done && j < rg.y) {
            bool ok;
            load_id<A8>(d, j, u, ok, w);
            const uint8_t su8 = d.st[u];
            const int32_t sq = RESPOND ? d.req[u] : 0;
            cand = candidate_of<W, RESPOND>(v, su8, sq, ok, alive);
        }
        const unsigned long long m = __ballot(cand) & gmask;
        const int src = m ? __ffsll((long long)m) - 1 : lane;
        const int32_t su = __shfl(u, src);
        const float sw = __shfl(w, src);
        if (!done && m) {
            win_u = su;
            win_w = sw;
            done = true;
        }
        base += G;
        if (base >= rg.y) done = true;
    }
}

// Rows deferred to a block's second pass: a first step that settles no row
// keeps the whole wavefront waiting for the next dependent step, so rows that
// need more than one step are set aside (in dynamic LDS, one 16-byte entry
// per thread of the block) and scanned together after the first steps of all
// rows; a row that finds no room is finished in place.
extern __shared__ __align__(16) int4 s_deferred[];

// Second pass over the nd rows deferred to s_def: {v (top bit: alive so
// far), next entry, row end, loc}, G lanes per row.  With rq the first
// field is the row's list slot instead of v: requests go to rq[slot].
template <bool W, int A8, bool PUSH, bool RESPOND, int G>
__device__ void deferred_pass(const Dev &d, const int4 *s_def, int nd, unsigned long long *resp,
                              uint32_t *s_key = nullptr, int64_t klo = 0, int64_t khi = 0, int32_t *rq = nullptr)
{
    constexpr int GPW = 64 / G;
    const int lane = threadIdx.x & 63, gl = lane & (G - 1), gw = lane / G;
    const unsigned long long gmask = G == 64 ? ~0ull : (((1ull << G) - 1ull) << (gw * G));
    const int step = (int)(blockDim.x >> 6) * GPW;
    for (int i0 = (threadIdx.x >> 6) * GPW; i0 < nd; i0 += step) {
        const bool valid = i0 + gw < nd;
        const int4 e = valid ? s_def[i0 + gw] : make_int4(0, 0, 0, 0);
        const int32_t v = e.x & INT_MAX;
        bool alive = e.x < 0;
        const int2 rg = make_int2(e.y, e.z);
        int32_t win_u = -1;
        float win_w = 1.0f;
        scan_rest<W, A8, RESPOND, G>(d, v, rg, rg.x, !valid, alive, win_u, win_w);
        alive = (__ballot(alive) & gmask) != 0ull;
        if (valid && gl == 0)
            first_write<W, PUSH, RESPOND>(d, v, e.w, win_u, win_w, alive, resp, s_key, klo, khi, rq ? rq + v : nullptr);
    }
}

// G lanes per row, G entries per step; a group stops at its first candidate
// (lowest lane of the first step that has one).  Wave-uniform control flow.
// Software-pipelined: while a group scans its row, the list entry two steps
// ahead and the first row chunk of the next entry are already in flight, so
// a row that settles in its first chunk (most of them) costs one dependent
// memory trip (the neighbour-state gathers) instead of two.  Vector loads
// return in issue order, so the gathers of the first step are issued before
// the prefetches: waiting for them then leaves the prefetches (HBM) in flight
// instead of waiting for those too.
// rq (request scans of PUSH rounds): the requests go to rq[list slot].
template <bool W, int A8, bool PUSH, bool RESPOND, int G>
__device__ void first_scan_bucket(const Dev &d, const int4 *list, int32_t count, unsigned long long *resp,
                                  unsigned long long &nnz, int32_t *rq = nullptr)
{
    constexpr int GPW = 64 / G;
    int4 *s_def = s_deferred;
    const int DEF_CAP = (int)blockDim.x;
    __shared__ int32_t s_ndef;
    if (count <= 0) return;
    if (threadIdx.x == 0) s_ndef = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, gl = lane & (G - 1), gw = lane / G;
    const unsigned long long gmask = G == 64 ? ~0ull : (((1ull << G) - 1ull) << (gw * G));
    const int step = (int)(blockDim.x >> 6) * GPW;
    // list entries are read at clamped indices (always in bounds); validity
    // is decided by the unclamped index.  With 16 or fewer rows per
    // wavefront the step's entries are one coalesced dword load (lane q
    // reads word q & 3 of entry q >> 2) and every group takes its entry's
    // words by shuffles, instead of every lane loading its group's 16 bytes.
    constexpr bool WORDS = GPW <= 16;
    auto ent_at = [&](int32_t i0) -> int4 {
        if (WORDS) {
            const int q = lane & (4 * GPW - 1);
            const int32_t w = __builtin_nontemporal_load(
                reinterpret_cast<const int32_t *>(list + min(i0 + (q >> 2), count - 1)) + (q & 3));
            return make_int4(w, 0, 0, 0);
        }
        return load_ent(list + min(i0 + gw, count - 1));
    };
    auto ent_of = [&](const int4 &x) -> int4 {
        if (!WORDS) return x;
        return make_int4(__shfl(x.x, gw * 4), __shfl(x.x, gw * 4 + 1), __shfl(x.x, gw * 4 + 2), __shfl(x.x, gw * 4 + 3));
    };
    int first = (threadIdx.x >> 6) * GPW;
    int4 cur = ent_of(ent_at(first));
    int4 nxt = ent_at(first + step);
    int2 praw = first_raw<A8>(d, cur, gl);  // first step of cur's row
    for (; first < count; first += step) {
        const bool valid = first + gw < count;
        const int4 e = cur;
        const int32_t v = e.x;
        const int2 rg = valid ? make_int2(e.y, e.z) : make_int2(0, 0);
        int32_t pu;
        bool pok;
        float pw;
        decode_raw<A8>(d, praw, pu, pok, pw);
        // (lanes past the row end gather vertex 0's slot: an edgeless graph's
        // row storage holds no ids at all)
        const bool in0 = rg.x + gl < rg.y;
        const int32_t g0 = in0 ? pu : 0;
        const uint8_t s0 = d.st[g0];
        const int32_t q0 = RESPOND ? d.req[g0] : 0;
        // keep the scheduler from hoisting the prefetches above the gathers
        __builtin_amdgcn_sched_barrier(0);
        // (the next entry's words are shuffled out before the load two steps
        // ahead is issued: the load can then take their register, and the
        // loop-carried copy that would wait for it goes away)
        const int4 nxe = ent_of(nxt);
        const int4 nn = ent_at(first + 2 * step);
        const int2 nraw = first_raw<A8>(d, nxe, gl);
        // step 0
        int32_t win_u = -1;
        float win_w = 1.0f;
        bool alive = false, done = rg.x >= rg.y;
        const bool cand = !done && in0 && candidate_of<W, RESPOND>(v, s0, q0, pok, alive);
        const unsigned long long m = __ballot(cand) & gmask;
        const int src = m ? __ffsll((long long)m) - 1 : lane;
        const int32_t su = __shfl(pu, src);
        const float sw = __shfl(pw, src);
        if (!done && m) {
            win_u = su;
            win_w = sw;
            done = true;
        }
        if (rg.x + G >= rg.y) done = true;
        alive = (__ballot(alive) & gmask) != 0ull;
        // unsettled rows go to the second pass while LDS room lasts
        int slot = -1;
        if (!done && gl == 0) slot = atomicAdd(&s_ndef, 1);
        slot = __shfl(slot, lane & ~(G - 1));
        const bool deferred = slot >= 0 && slot < DEF_CAP;
        if (deferred && gl == 0)
            s_def[slot] = make_int4((rq ? first + gw : v) | (alive ? INT_MIN : 0), rg.x + G, rg.y, e.w);
        scan_rest<W, A8, RESPOND, G>(d, v, rg, rg.x + G, done || deferred, alive, win_u, win_w);
        alive = (__ballot(alive) & gmask) != 0ull;
        if (valid && gl == 0) {
            if (!deferred)
                first_write<W, PUSH, RESPOND>(d, v, e.w, win_u, win_w, alive, resp, nullptr, 0, 0,
                                              rq ? rq + first + gw : nullptr);
            nnz += (unsigned long long)(rg.y - rg.x);
        }
        cur = nxe;
        nxt = nn;
        praw = nraw;
    }
    __syncthreads();
    deferred_pass<W, A8, PUSH, RESPOND, G>(d, s_def, min(s_ndef, DEF_CAP), resp, nullptr, 0, 0, rq);
}

// Round-1 scans over the head copy, one lane per row, straight over the
// bucket's id range [lo, hi) (every vertex is active in round 1): a lane reads
// its vertex's state, row range, id and the row's 4 leading entries (two
// 16-byte loads) at once, all coalesced, then gathers the 4 neighbour states
// at once -- two dependent memory steps for 64 rows per wavefront.  Blue rows
// with no candidate among those 4 and more entries go to the second pass (G2
// lanes per row); rows above the hub degree are the hub blocks' work.
// Weighted rows only (the head copy exists for those).
template <bool PUSH, int G2, bool HC>
__device__ void head_scan_range(const Dev &d, int64_t lo, int64_t hi, unsigned long long *resp,
                                unsigned long long &nnz)
{
    int4 *s_def = s_deferred;
    const int DEF_CAP = (int)blockDim.x;
    // split1: the bucket's keys, after the deferred rows
    uint32_t *s_key = d.split1 ? reinterpret_cast<uint32_t *>(s_deferred + blockDim.x) : nullptr;
    __shared__ int32_t s_ndef;
    if (hi <= lo) return;
    if (threadIdx.x == 0) s_ndef = 0;
    if (s_key)
        for (int64_t i = threadIdx.x; i < hi - lo; i += blockDim.x) s_key[i] = 0xFFFFFFFFu;
    __syncthreads();
    typedef int v4i __attribute__((ext_vector_type(4)));
    for (int64_t base = lo + (threadIdx.x & ~63u); base < hi; base += blockDim.x) {
        const int64_t vi = base + (threadIdx.x & 63);
        const int32_t v = (int32_t)min(vi, hi - 1);
        const uint8_t sv0 = d.st[v];
        const int2 rg = d.ranges[v];
        const int32_t vloc = d.loc[v];
        int32_t u[4];
        bool ok[4];
        float w[4];
        uint8_t sv[4];
        if (HC) {
            const v4i h = __builtin_nontemporal_load(reinterpret_cast<const v4i *>(d.headc + (size_t)v * 4));
            const uint32_t hw[4] = {(uint32_t)h.x, (uint32_t)h.y, (uint32_t)h.z, (uint32_t)h.w};
            const uint32_t idmask = (1u << (31 - d.head_pb)) - 1u, pmask = (1u << d.head_pb) - 1u;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ok[k] = (hw[k] >> 31) == 0u;
                u[k] = (int32_t)((hw[k] >> d.head_pb) & idmask);
                w[k] = __int_as_float((int32_t)(hw[k] & pmask));  // (the key position)
            }
        } else {
            const v4i *hp = reinterpret_cast<const v4i *>(d.head + (size_t)v * 4);
            const v4i h01 = __builtin_nontemporal_load(hp), h23 = __builtin_nontemporal_load(hp + 1);
            const int2 raw[4] = {make_int2(h01.x, h01.y), make_int2(h01.z, h01.w), make_int2(h23.x, h23.y),
                                 make_int2(h23.z, h23.w)};
#pragma unroll
            for (int k = 0; k < 4; ++k) decode_raw<1>(d, raw[k], u[k], ok[k], w[k]);
        }
        const int32_t deg0 = rg.y - rg.x;
        const bool valid = vi < hi && sv0 == ST_BLUE && deg0 <= d.hub_deg;
        const int32_t deg = valid ? deg0 : 0;
        // (entries past the row end gather vertex 0's slot)
#pragma unroll
        for (int k = 0; k < 4; ++k) sv[k] = d.st[k < deg ? u[k] : 0];
        bool alive = false;
        int32_t win_u = -1;
        float win_w = 1.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < deg && win_u < 0 && candidate_of<true, false>(0, sv[k], 0, ok[k], alive)) {
                win_u = u[k];
                win_w = w[k];
            }
        bool deferred = false;
        if (valid && win_u < 0 && deg > 4) {
            const int slot = atomicAdd(&s_ndef, 1);
            if (slot < DEF_CAP) {
                s_def[slot] = make_int4(v | (alive ? INT_MIN : 0), rg.x + 4, rg.y, vloc);
                deferred = true;
            } else {
                // no room left: finish the row here
                for (int32_t j = rg.x + 4; j < rg.y && win_u < 0; ++j) {
                    int32_t uj;
                    bool okj;
                    float wj;
                    load_id<HC ? 2 : 1>(d, j, uj, okj, wj);
                    if (candidate_of<true, false>(0, d.st[uj], 0, okj, alive)) {
                        win_u = uj;
                        win_w = wj;
                    }
                }
            }
        }
        if (valid && !deferred) first_write<true, PUSH, false>(d, v, vloc, win_u, win_w, alive, resp, s_key, lo, hi);
        if (valid) nnz += (unsigned long long)deg;
    }
    __syncthreads();
    deferred_pass<true, HC ? 2 : 1, PUSH, false, G2>(d, s_def, min(s_ndef, DEF_CAP), resp, s_key, lo, hi);
    if (s_key) {
        __syncthreads();
        uint32_t *bk = bucket_keys(d);
        for (int64_t i = threadIdx.x; i < hi - lo; i += blockDim.x) bk[lo + i] = s_key[i];
    }
}

// Rows longer than the hub degree: one block per row, blockDim entries per
// step, stopping at the step with the first candidate.
template <bool W, int A8, bool PUSH, bool RESPOND>
__device__ void first_scan_hubs(const Dev &d, const int4 *list, int32_t count, int hb, int hub_blocks,
                                unsigned long long *resp, unsigned long long &nnz)
{
    __shared__ int32_t s_first, s_alive;
    for (int32_t idx = hb; idx < count; idx += hub_blocks) {
        const int4 ent = list[idx];
        const int32_t v = ent.x;
        const int2 rg = make_int2(ent.y, ent.z);
        if (threadIdx.x == 0) {
            s_first = INT_MAX;
            s_alive = 0;
        }
        __syncthreads();
        for (int32_t base = rg.x; base < rg.y; base += blockDim.x) {
            const int32_t j = base + threadIdx.x;
            if (j < rg.y) {
                int32_t u;
                float w;
                bool ok, alive = false;
                load_id<A8>(d, j, u, ok, w);
                if (first_candidate<W, RESPOND>(d, v, u, ok, alive)) atomicMin(&s_first, j);
                if (alive) s_alive = 1;
            }
            __syncthreads();
            const int32_t f = s_first;
            __syncthreads();
            if (f != INT_MAX) break;
        }
        if (threadIdx.x == 0) {
            int32_t u = -1;
            float w = 0.0f;
            if (s_first != INT_MAX) load_entry<A8>(d, s_first, u, w);
            first_write<W, PUSH, RESPOND>(d, v, ent.w, u, w, s_alive != 0, resp);
            nnz += (unsigned long long)(rg.y - rg.x);
        }
        __syncthreads();
    }
}



__device__ __forceinline__ int32_t *ctrb_ptr(const Dev &d, int r, int j) { return d.ctrb + ((size_t)r * 4 + j) * d.nb; }
__device__ __forceinline__ unsigned long long *ctrb64_ptr(const Dev &d, int r, int j)
{
    return d.ctrb64 + ((size_t)r * 2 + j) * d.nb;
}

// Appends value to this block's bucket for every lane with pred set; the
// bucket counter lives in LDS, so no global atomic is involved.  Must be
// reached by all lanes of the wavefront.
__device__ __forceinline__ void bucket_append(bool pred, int4 value, int4 *bucket, int32_t *s_count)
{
    const unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(s_count, (int)__popcll(mask));
    base = __shfl(base, leader);
    if (pred) bucket[base + (int)__popcll(mask & ((1ull << lane) - 1ull))] = value;
}

// bucket_append of an 8-byte entry.
__device__ __forceinline__ void bucket_append2(bool pred, int2 value, int2 *bucket, int32_t *s_count)
{
    const unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(s_count, (int)__popcll(mask));
    base = __shfl(base, leader);
    if (pred) bucket[base + (int)__popcll(mask & ((1ull << lane) - 1ull))] = value;
}

// bucket_append without the entry: counts the lanes with pred set.
__device__ __forceinline__ void bucket_count(bool pred, int32_t *s_count)
{
    const unsigned long long mask = __ballot(pred);
    if (mask != 0ull && __lane_id() == __ffsll((long long)mask) - 1) atomicAdd(s_count, (int)__popcll(mask));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// Block sum of per-thread values; the result is valid in thread 0.
template <typename T>
__device__ T block_sum(T x)
{
    __shared__ T s_part[16];
    x = wave_sum(x);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = x;
    __syncthreads();
    T t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) t += s_part[w];
    __syncthreads();
    return t;
}

// Select of round 1 (match starts at 0 everywhere, src/matchgpu.cu:486).
// Block b colours the ids of bucket b and files them into round 1's lists.
// Held to 64 VGPRs (8 wavefronts per SIMD): two blocks of 1024 threads share
// a CU, so all the buckets' blocks of a 2^22-vertex graph run at once.
template <bool PUSH>
__global__ __attribute__((amdgpu_flat_work_group_size(1, 1024), amdgpu_waves_per_eu(8))) void k_init_select(Dev d)
{
    __shared__ int32_t s_cnt[2];
    const int parity = 1, b = bucket_of_block(d);
    if (threadIdx.x < 2) s_cnt[threadIdx.x] = 0;
    // Runs alternate between two counter sets; the hub-list sizes and hub
    // degree sums are accumulated with atomics, so this run clears the other
    // set for the next one (this run's set was cleared by the previous run).
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < ROUND_SLOTS * CTR_STRIDE; i += blockDim.x) d.ctr_next[i] = 0;
        for (int i = threadIdx.x; i < ROUND_SLOTS * 2; i += blockDim.x) d.hub64_next[i] = 0ull;
    }
    __syncthreads();
    const uint32_t seed = d.seeds[0];
    const int64_t lo = (int64_t)b * d.bcap, hi = min((int64_t)d.n, lo + d.bcap);
    int4 *blue = list_ptr(d, parity, L_BLUE_LO) + (size_t)b * d.bcap;
    int4 *red = list_ptr(d, parity, L_RED_LO) + (size_t)b * d.bcap;
    unsigned long long deg_sum = 0, hub_deg_sum = 0;
    // IU vertices per thread and step, all loads issued before any hash.  The
    // row ranges serve the hub test, the round's lists and the degree sums
    // only: without lists, hubs or sums they are not read.
    const bool need_rows = !d.listless1 || !d.no_hubs || d.count_nnz;
    // Without lists or hubs, a lane takes 4 consecutive ids: one 16-byte
    // load of their original ids, one 4-byte store of their states, one
    // 16-byte store (32 with 64-bit keys) of their empty keys.  The address
    // unit's work grows with the lanes a memory instruction serves, so 4 ids
    // per lane cost a fraction of 4 single-id accesses.
    // (the degree sums of a profiled run read the ranges as two more 16-byte
    // loads per lane)
    const bool quads = d.listless1 && d.no_hubs && (lo & 3) == 0;
    if (quads) {
        typedef int v4i __attribute__((ext_vector_type(4)));
        int32_t nblue = 0, nred = 0;
        const int64_t nq = (hi - lo) >> 2;
        for (int64_t k = threadIdx.x; k < nq; k += blockDim.x) {
            const int64_t v0 = lo + 4 * k;
            const v4i l4 = *reinterpret_cast<const v4i *>(d.loc + v0);
            if (d.count_nnz) {
                const v4i *rp = reinterpret_cast<const v4i *>(d.ranges + v0);
                const v4i r01 = rp[0], r23 = rp[1];
                deg_sum += (unsigned long long)((r01.y - r01.x) + (r01.w - r01.z) + (r23.y - r23.x) + (r23.w - r23.z));
            }
            const int32_t lv[4] = {l4.x, l4.y, l4.z, l4.w};
            uint32_t packed = 0u;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int c = colour_of((uint32_t)lv[q], seed, d.barrier);
                packed |= (uint32_t)c << (8 * q);
                nblue += c == ST_BLUE;
                nred += c == ST_RED;
            }
            *reinterpret_cast<uint32_t *>(d.st + v0) = packed;
            if (d.vround) *reinterpret_cast<uint32_t *>(d.vround + v0) = 0xFFFFFFFFu;
            if (PUSH) {
                if (d.key32) {
                    *reinterpret_cast<v4i *>(reinterpret_cast<uint32_t *>(resp_slot(d, parity)) + v0) =
                        v4i{-1, -1, -1, -1};
                } else {
                    v4i *kp = reinterpret_cast<v4i *>(resp_slot(d, parity) + v0);
                    kp[0] = v4i{0, 0, 0, 0};
                    kp[1] = v4i{0, 0, 0, 0};
                }
            }
        }
        // the last (hi - lo) % 4 ids, one per thread
        for (int64_t v = lo + 4 * nq + threadIdx.x; v < hi; v += blockDim.x) {
            if (d.count_nnz) deg_sum += (unsigned long long)(d.ranges[v].y - d.ranges[v].x);
            const int c = colour_of((uint32_t)d.loc[v], seed, d.barrier);
            d.st[v] = (uint8_t)c;
            if (d.vround) d.vround[v] = 255;
            nblue += c == ST_BLUE;
            nred += c == ST_RED;
            if (PUSH) clear_key(d, resp_slot(d, parity), (int32_t)v);
        }
        nblue = wave_sum(nblue);
        nred = wave_sum(nred);
        if ((threadIdx.x & 63) == 0) {
            if (nblue) atomicAdd(s_cnt + 0, nblue);
            if (nred) atomicAdd(s_cnt + 1, nred);
        }
    }
    constexpr int IU = 2;
    for (int64_t base = lo + (threadIdx.x & ~63); !quads && base < hi; base += (int64_t)blockDim.x * IU) {
        int32_t vl[IU];
        int2 rg[IU];
#pragma unroll
        for (int q = 0; q < IU; ++q) {
            const int64_t v = base + (int64_t)q * blockDim.x + (threadIdx.x & 63);
            vl[q] = 0;
            rg[q] = make_int2(0, 0);
            if (v < hi) {
                vl[q] = d.loc[v];
                if (need_rows) rg[q] = d.ranges[v];
            }
        }
#pragma unroll
        for (int q = 0; q < IU; ++q) {
            const int32_t v = (int32_t)(base + (int64_t)q * blockDim.x + (threadIdx.x & 63));
            int cls = -1;
            const int4 ent = make_int4(v, rg[q].x, rg[q].y, vl[q]);
            if (v < hi) {
                const int c = colour_of((uint32_t)vl[q], seed, d.barrier);
                d.st[v] = (uint8_t)c;
                if (d.vround) d.vround[v] = 255;
                const int32_t deg = rg[q].y - rg[q].x;
                const int hub = deg > d.hub_deg ? 1 : 0;
                cls = 2 * c + hub;
                if (hub) hub_deg_sum += (unsigned long long)deg;
                else deg_sum += (unsigned long long)deg;
                // every vertex's key is emptied, not only the red ones, so the
                // stores cover whole lines (a blue vertex's key is never read)
                if (PUSH) clear_key(d, resp_slot(d, parity), v);
            }
            if (d.listless1) {
                bucket_count(cls == L_BLUE_LO, s_cnt + 0);
                bucket_count(cls == L_RED_LO, s_cnt + 1);
            } else {
                bucket_append(cls == L_BLUE_LO, ent, blue, s_cnt + 0);
                bucket_append(cls == L_RED_LO, ent, red, s_cnt + 1);
            }
            wave_append(cls == L_BLUE_HI, ent, list_ptr(d, parity, L_BLUE_HI), d.ctr + 1 * CTR_STRIDE + L_BLUE_HI);
            wave_append(cls == L_RED_HI, ent, list_ptr(d, parity, L_RED_HI), d.ctr + 1 * CTR_STRIDE + L_RED_HI);
        }
    }
    if (d.count_nnz) {
        deg_sum = block_sum(deg_sum);
        hub_deg_sum = block_sum(hub_deg_sum);
        if (threadIdx.x == 0) {
            ctrb64_ptr(d, 1, 0)[b] = deg_sum;
            if (hub_deg_sum) atomicAdd(d.hub64 + 1 * 2 + 0, hub_deg_sum);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ctrb_ptr(d, 1, 0)[b] = s_cnt[0];
        ctrb_ptr(d, 1, 1)[b] = s_cnt[1];
    }
}

// Vertices without neighbours (select of round 1, graphs with many of them).
// Such a vertex is never proposed to and never proposes: in every round it is
// red and stays, or blue and dies (its request is n + 1, src/matchgpu.cu:
// 320-325).  Its whole run depends only on its id and the round seeds, so the
// round-1 select settles it at once instead of carrying it through every
// round's lists: it is marked dead for the kernels that scan by id, its final
// value is written (2 in the round it is first blue; its colour, red, if it
// is red to the last round), and the rounds' counters get its share later
// (k_reduce_counts), so round counts and statistics are unchanged.
//
// The red ones are queued, per bucket, in the bucket's segments of the next
// round's work lists (round 2's lists are written only by round 1's match),
// as {id, original id} pairs; after the bucket's scan the block takes the
// queue through rounds 2, 3, ... together, each round hashing only the ones
// still red and compacting them into the other segment.
__device__ void iso_rounds(const Dev &d, int2 *qa, int2 *qb, int32_t *s_q, int32_t *s_iso, int32_t *s_last)
{
    const int lane = threadIdx.x & 63;
    for (int r = 2; r <= d.R_run; ++r) {
        __syncthreads();
        const int32_t qn = s_q[0];
        if (qn == 0) break;
        if (threadIdx.x == 0) {
            s_q[1] = 0;
            if (r > *s_last) *s_last = r;
        }
        __syncthreads();
        const uint32_t seed = d.seeds[r - 1];
        int32_t nb = 0, nr = 0;
        // (every lane of a wavefront reaches the append: the loop bound is wave-uniform)
        for (int32_t i0 = threadIdx.x & ~63; i0 < qn; i0 += blockDim.x) {
            const int32_t i = i0 + lane;
            bool keep = false;
            int2 e = make_int2(0, 0);
            if (i < qn) {
                e = qa[i];
                if (colour_of((uint32_t)e.y, seed, d.barrier) == ST_BLUE) {
                    d.out[out_base(d, e.x) + e.y] = 2;
                    if (d.vround) d.vround[e.x] = (uint8_t)r;
                    ++nb;
                } else {
                    keep = true;
                    ++nr;
                }
            }
            bucket_append2(keep, e, qb, s_q + 1);
        }
        nb = wave_sum(nb);
        nr = wave_sum(nr);
        if (lane == 0) {
            if (nb) atomicAdd(s_iso + 2 * r, nb);
            if (nr) atomicAdd(s_iso + 2 * r + 1, nr);
        }
        __syncthreads();
        if (threadIdx.x == 0) s_q[0] = s_q[1];
        int2 *t = qa;
        qa = qb;
        qb = t;
    }
    // red to the last round: the run ends with them active and red
    __syncthreads();
    for (int32_t i = threadIdx.x; i < s_q[0]; i += blockDim.x) {
        const int2 e = qa[i];
        d.out[out_base(d, e.x) + e.y] = ST_RED;
    }
}

// k_init_select's general path (row ranges read, lists or hubs) for graphs
// whose vertices without neighbours are settled here (Dev::iso).
template <bool PUSH>
__global__ __attribute__((amdgpu_flat_work_group_size(1, 1024), amdgpu_waves_per_eu(8))) void k_init_select_iso(Dev d)
{
    constexpr int IU = 2;
    __shared__ int32_t s_cnt[2], s_q[2], s_last;
    __shared__ int32_t s_iso[2 * ROUND_SLOTS];
    const int parity = 1, b = bucket_of_block(d);
    int2 *qa = reinterpret_cast<int2 *>(list_ptr(d, 0, L_BLUE_LO) + (size_t)b * d.bcap);
    int2 *qb = reinterpret_cast<int2 *>(list_ptr(d, 0, L_RED_LO) + (size_t)b * d.bcap);
    if (threadIdx.x < 2) {
        s_cnt[threadIdx.x] = 0;
        s_q[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0) s_last = 0;
    for (int i = threadIdx.x; i < 2 * ROUND_SLOTS; i += blockDim.x) s_iso[i] = 0;
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < ROUND_SLOTS * CTR_STRIDE; i += blockDim.x) d.ctr_next[i] = 0;
        for (int i = threadIdx.x; i < ROUND_SLOTS * 2; i += blockDim.x) d.hub64_next[i] = 0ull;
    }
    __syncthreads();
    const uint32_t seed = d.seeds[0];
    const int64_t lo = (int64_t)b * d.bcap, hi = min((int64_t)d.n, lo + d.bcap);
    int4 *blue = list_ptr(d, parity, L_BLUE_LO) + (size_t)b * d.bcap;
    int4 *red = list_ptr(d, parity, L_RED_LO) + (size_t)b * d.bcap;
    unsigned long long deg_sum = 0, hub_deg_sum = 0;
    int32_t iso_b = 0, iso_r = 0;
    for (int64_t base = lo + (threadIdx.x & ~63); base < hi; base += (int64_t)blockDim.x * IU) {
        int32_t vl[IU];
        int2 rg[IU];
#pragma unroll
        for (int q = 0; q < IU; ++q) {
            const int64_t v = base + (int64_t)q * blockDim.x + (threadIdx.x & 63);
            vl[q] = 0;
            rg[q] = make_int2(0, 0);
            if (v < hi) {
                vl[q] = d.loc[v];
                // (ids from n1 on have no neighbours: their row ranges are not needed)
                if (v < d.n1) rg[q] = d.ranges[v];
            }
        }
#pragma unroll
        for (int q = 0; q < IU; ++q) {
            const int32_t v = (int32_t)(base + (int64_t)q * blockDim.x + (threadIdx.x & 63));
            int cls = -1;
            bool queue = false;
            const int4 ent = make_int4(v, rg[q].x, rg[q].y, vl[q]);
            if (v < hi) {
                const int c = colour_of((uint32_t)vl[q], seed, d.barrier);
                if (d.vround) d.vround[v] = 255;
                const int32_t deg = rg[q].y - rg[q].x;
                if (deg == 0) {
                    // (no kernel reads the state of an id from n1 on, nor the key of a
                    // vertex without neighbours: nobody proposes to it)
                    if (v < d.n1) d.st[v] = ST_DEAD;
                    if (c == ST_BLUE) {
                        d.out[out_base(d, v) + vl[q]] = 2;
                        if (d.vround) d.vround[v] = 1;
                        ++iso_b;
                    } else {
                        queue = true;
                        ++iso_r;
                    }
                } else {
                    d.st[v] = (uint8_t)c;
                    const int hub = deg > d.hub_deg ? 1 : 0;
                    cls = 2 * c + hub;
                    if (hub) hub_deg_sum += (unsigned long long)deg;
                    else deg_sum += (unsigned long long)deg;
                    if (PUSH) clear_key(d, resp_slot(d, parity), v);
                }
            }
            bucket_append2(queue, make_int2(v, vl[q]), qa, s_q);
            if (d.listless1) {
                bucket_count(cls == L_BLUE_LO, s_cnt + 0);
                bucket_count(cls == L_RED_LO, s_cnt + 1);
            } else {
                bucket_append(cls == L_BLUE_LO, ent, blue, s_cnt + 0);
                bucket_append(cls == L_RED_LO, ent, red, s_cnt + 1);
            }
            wave_append(cls == L_BLUE_HI, ent, list_ptr(d, parity, L_BLUE_HI), d.ctr + 1 * CTR_STRIDE + L_BLUE_HI);
            wave_append(cls == L_RED_HI, ent, list_ptr(d, parity, L_RED_HI), d.ctr + 1 * CTR_STRIDE + L_RED_HI);
        }
    }
    iso_rounds(d, qa, qb, s_q, s_iso, &s_last);
    iso_b = wave_sum(iso_b);
    iso_r = wave_sum(iso_r);
    if ((threadIdx.x & 63) == 0) {
        if (iso_b) atomicAdd(s_iso + 2, iso_b);
        if (iso_r) atomicAdd(s_iso + 3, iso_r);
    }
    if (d.count_nnz) {
        deg_sum = block_sum(deg_sum);
        hub_deg_sum = block_sum(hub_deg_sum);
        if (threadIdx.x == 0) {
            ctrb64_ptr(d, 1, 0)[b] = deg_sum;
            if (hub_deg_sum) atomicAdd(d.hub64 + 1 * 2 + 0, hub_deg_sum);
        }
    }
    __syncthreads();
    // this run's isolated counts of the bucket, round by round, and its last round
    const int32_t last = s_iso[2] + s_iso[3] > 0 ? max(s_last, 1) : 0;
    for (int i = threadIdx.x; i < 2 * (last + 1); i += blockDim.x)
        if (i >= 2) d.iso[(size_t)i * d.nb + b] = s_iso[i];
    if (threadIdx.x == 0) {
        d.iso[(size_t)2 * ROUND_SLOTS * d.nb + b] = last;
        ctrb_ptr(d, 1, 0)[b] = s_cnt[0];
        ctrb_ptr(d, 1, 1)[b] = s_cnt[1];
    }
}

// Request phase of round r.  Blocks [0, nb) take bucket b of the blue rows,
// G lanes per row; blocks [nb, nb + hub_blocks) take the hub rows, one block
// per row.
template <bool W, int A8, bool PUSH, int G, bool HEAD = false, bool HC = false>
__global__ void k_request(Dev d, int r, int hub_blocks)
{
    const int parity = r & 1;
    unsigned long long *resp = PUSH ? resp_slot(d, parity) : nullptr;
    unsigned long long nnz = 0;
    if ((int)blockIdx.x >= d.nb) {
        first_scan_hubs<W, A8, PUSH, false>(d, list_ptr(d, parity, L_BLUE_HI), d.ctr[r * CTR_STRIDE + L_BLUE_HI],
                                            blockIdx.x - d.nb, hub_blocks, resp, nnz);
        if (d.count_nnz && threadIdx.x == 0 && nnz) atomicAdd(d.hub64 + (size_t)r * 2 + 1, nnz);
        return;
    }
    const int b = bucket_of_block(d);
    if (HEAD)
        head_scan_range<PUSH, 8, HC>(d, (int64_t)b * d.bcap1, min((int64_t)d.n1, (int64_t)(b + 1) * d.bcap1),
                                     resp, nnz);
    else
        first_scan_bucket<W, A8, PUSH, false, G>(d, list_ptr(d, parity, L_BLUE_LO) + (size_t)b * d.bcap,
                                                 ctrb_ptr(d, r, 0)[b], resp, nnz,
                                                 PUSH && d.reqs ? d.reqs + (size_t)b * d.bcap : nullptr);
    if (d.count_nnz) {
        nnz = block_sum(nnz);
        if (threadIdx.x == 0) ctrb64_ptr(d, r, 1)[b] = nnz;
    }
}

// ---------------------------------------------------------------------------
// Entry-parallel tiles.  A block takes up to blockDim vertices of its bucket,
// stages their ids and row starts in LDS with an exclusive scan of the row
// lengths, and then hands out ROW ENTRIES, not rows, to its threads: each
// thread finds the row of entry e by binary search over the LDS prefix and
// issues U independent row loads before using any.  The per-row winner is an
// LDS 64-bit atomic max on (order-preserving weight, ~position in row), i.e.
// largest weight, first in CSR order on ties -- the reference's strict '>'
// scan from -1.0.  Only two dependent memory steps remain per entry (row
// entry, neighbour state), whatever the degree distribution.
// ---------------------------------------------------------------------------

// Exclusive block scan of x; total in *s_total.  Reached by all threads.
__device__ int32_t block_exclusive_scan(int32_t x, int32_t *s_wave, int32_t *s_total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    int32_t incl = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t acc = 0;
        for (int w = 0; w < nw; ++w) {
            const int32_t t = s_wave[w];
            s_wave[w] = acc;
            acc += t;
        }
        *s_total = acc;
    }
    __syncthreads();
    return s_wave[wave] + incl - x;
}

// Largest k in [0, tn) with off[k] <= e (off[0] == 0 <= e).
__device__ __forceinline__ int tile_row_of(const int32_t *off, int tn, int32_t e)
{
    int lo = 0, hi = tn - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint32_t order_bits(float w)
{
    const uint32_t bits = __float_as_uint(w == 0.0f ? 0.0f : w);
    return (bits & 0x80000000u) ? ~bits : (bits | 0x80000000u);
}

// RESPOND == false: request scan of blue rows; true: respond scan of red rows.
// U: row entries each thread has in flight.
template <bool W, int A8, bool PUSH, bool RESPOND, int U>
__device__ void tile_scan_bucket(const Dev &d, const int4 *list, int32_t count, unsigned long long *resp)
{
    extern __shared__ unsigned long long s_dyn[];
    const int T = blockDim.x;
    unsigned long long *s_key = s_dyn;                          // [T]
    int32_t *s_v = reinterpret_cast<int32_t *>(s_key + T);      // [T]
    int32_t *s_beg = s_v + T;                                   // [T]
    int32_t *s_off = s_beg + T;                                 // [T]
    int32_t *s_alive = s_off + T;                               // [T]
    int32_t *s_wave = s_alive + T;                              // [16]
    int32_t *s_total = s_wave + 16;                             // [1]
    const int tid = threadIdx.x;
    for (int t0 = 0; t0 < count; t0 += T) {
        const int tn = min(T, count - t0);
        int32_t deg = 0;
        if (tid < tn) {
            const int4 ent = load_ent(list + t0 + tid);
            const int32_t v = ent.x;
            const int2 rg = make_int2(ent.y, ent.z);
            s_v[tid] = v;
            s_beg[tid] = rg.x;
            deg = rg.y - rg.x;
            s_key[tid] = 0ull;
            s_alive[tid] = 0;
        }
        s_off[tid] = block_exclusive_scan(deg, s_wave, s_total);
        __syncthreads();
        const int32_t E = *s_total;
        for (int32_t e0 = 0; e0 < E; e0 += T * U) {
            int32_t u[U], k[U], pos[U];
            float w[U];
#pragma unroll
            for (int q = 0; q < U; ++q) {
                const int32_t e = e0 + q * T + tid;
                k[q] = -1;
                if (e < E) {
                    k[q] = tile_row_of(s_off, tn, e);
                    pos[q] = e - s_off[k[q]];
                    load_entry<A8>(d, s_beg[k[q]] + pos[q], u[q], w[q]);
                }
            }
            uint8_t s[U];
            int32_t rq[U];
#pragma unroll
            for (int q = 0; q < U; ++q) {
                s[q] = k[q] >= 0 ? d.st[u[q]] : ST_MATCHED;
                if (RESPOND) rq[q] = (k[q] >= 0 && s[q] == ST_BLUE) ? d.req[u[q]] : -1;
            }
#pragma unroll
            for (int q = 0; q < U; ++q) {
                if (k[q] < 0) continue;
                bool cand;
                if (RESPOND) {
                    cand = s[q] == ST_BLUE && rq[q] == s_v[k[q]];
                } else {
                    if (s[q] != ST_MATCHED) s_alive[k[q]] = 1;
                    cand = s[q] == ST_RED;
                }
                if (W) cand = cand && w[q] > -1.0f;
                if (cand) {
                    const unsigned long long key = ((unsigned long long)(W ? order_bits(w[q]) : 1u) << 32) |
                                                   (unsigned long long)(0xFFFFFFFFu - (uint32_t)pos[q]);
                    atomicMax(s_key + k[q], key);
                }
            }
        }
        __syncthreads();
        if (tid < tn) {
            const int32_t v = s_v[tid];
            const unsigned long long key = s_key[tid];
            if (key != 0ull) {
                const int32_t j = s_beg[tid] + (int32_t)(0xFFFFFFFFu - (uint32_t)key);
                int32_t un;
                float wn;
                load_entry<A8>(d, j, un, wn);
                d.req[v] = un;
                if (!RESPOND && PUSH)
                    atomicMax(resp + un,
                              (key & 0xFFFFFFFF00000000ull) | (unsigned long long)(0xFFFFFFFFu - (uint32_t)d.loc[v]));
            } else {
                d.req[v] = RESPOND ? REQ_NONE : (s_alive[tid] ? REQ_NONE : REQ_DEAD);
            }
        }
        __syncthreads();
    }
}

template <bool W, int A8, bool PUSH, int U>
__global__ void k_request_tile(Dev d, int r, int hub_blocks)
{
    const int parity = r & 1;
    unsigned long long *resp = PUSH ? d.resp + (size_t)parity * d.n : nullptr;
    if ((int)blockIdx.x >= d.nb) {
        unsigned long long nnz = 0;
        first_scan_hubs<W, A8, PUSH, false>(d, list_ptr(d, parity, L_BLUE_HI), d.ctr[r * CTR_STRIDE + L_BLUE_HI],
                                            blockIdx.x - d.nb, hub_blocks, resp, nnz);
        if (d.count_nnz && threadIdx.x == 0 && nnz) atomicAdd(d.hub64 + (size_t)r * 2 + 1, nnz);
        return;
    }
    const int b = bucket_of_block(d);
    const int32_t count = ctrb_ptr(d, r, 0)[b];
    const int4 *list = list_ptr(d, parity, L_BLUE_LO) + (size_t)b * d.bcap;
    tile_scan_bucket<W, A8, PUSH, false, U>(d, list, count, resp);
    if (d.count_nnz) {
        unsigned long long nnz = 0;
        for (int i = threadIdx.x; i < count; i += blockDim.x) {
            const int4 e = list[i];
            nnz += (unsigned long long)(e.z - e.y);
        }
        nnz = block_sum(nnz);
        if (threadIdx.x == 0) ctrb64_ptr(d, r, 1)[b] = nnz;
    }
}

template <bool W, int A8, int U>
__global__ void k_respond_tile(Dev d, int r, int hub_blocks)
{
    const int parity = r & 1;
    if ((int)blockIdx.x >= d.nb) {
        unsigned long long nnz = 0;
        first_scan_hubs<W, A8, false, true>(d, list_ptr(d, parity, L_RED_HI), d.ctr[r * CTR_STRIDE + L_RED_HI],
                                            blockIdx.x - d.nb, hub_blocks, nullptr, nnz);
        return;
    }
    const int b = bucket_of_block(d);
    tile_scan_bucket<W, A8, false, true, U>(d, list_ptr(d, parity, L_RED_LO) + (size_t)b * d.bcap,
                                         ctrb_ptr(d, r, 1)[b], nullptr);
}

__host__ __device__ constexpr size_t tile_smem_bytes(int T) { return (size_t)T * (8 + 4 * 4) + 17 * 4; }

// Respond phase of round r (PULL): red vertices pick their proposer.
template <bool W, int A8, int G>
__global__ void k_respond(Dev d, int r, int hub_blocks)
{
    const int parity = r & 1;
    unsigned long long nnz = 0;
    if ((int)blockIdx.x >= d.nb) {
        first_scan_hubs<W, A8, false, true>(d, list_ptr(d, parity, L_RED_HI), d.ctr[r * CTR_STRIDE + L_RED_HI],
                                            blockIdx.x - d.nb, hub_blocks, nullptr, nnz);
        return;
    }
    const int b = bucket_of_block(d);
    first_scan_bucket<W, A8, false, true, G>(d, list_ptr(d, parity, L_RED_LO) + (size_t)b * d.bcap,
                                             ctrb_ptr(d, r, 1)[b], nullptr, nnz);
}

// Writes the outcome of round r for vertex v (gMatch, src/matchgpu.cu:260-283)
// and, when do_select, its colour for round r+1 (gSelect).  Returns the class
// of v in round r+1, or -1 when it left the active set.
// partner is the partner's graph-local original id (-1: none); vloc is v's.
// Final values go straight into the caller-order result as vertices settle
// (a 4-byte scattered store per vertex over the whole run, instead of a
// separate permutation pass at its end).  out index of a vertex with
// graph-local id loc in the graph of internal vertex v: vbase[v] + loc.

// partner_of's value for a blue vertex of a PUSH round that won its proposal:
// its result is written by the red end, which knows both ids.
constexpr int32_t PARTNER_WRITES = INT_MAX;

template <bool PUSH>
__device__ __forceinline__ int settle_vertex(const Dev &d, int32_t v, int32_t vloc, int colour, int hub,
                                             int32_t partner, bool dead, int r, int do_select, uint32_t seed,
                                             unsigned long long *nresp)
{
    if (partner >= 0) {
        if (partner != PARTNER_WRITES) {
            const int32_t m = 4 + min(vloc, partner), ob = out_base(d, v);
            d.out[ob + vloc] = m;
            // PUSH: the red end also writes the blue end's value (same graph)
            if (PUSH && colour == ST_RED) d.out[ob + partner] = m;
        }
        d.st[v] = ST_MATCHED;
        if (d.vround) d.vround[v] = (uint8_t)r;
        return -1;
    }
    if (dead) {
        d.out[out_base(d, v) + vloc] = 2;
        d.st[v] = ST_DEAD;
        if (d.vround) d.vround[v] = (uint8_t)r;
        return -1;
    }
    if (!do_select) {
        // the run ends with v still active: its value is this round's colour
        d.out[out_base(d, v) + vloc] = colour;
        return -1;
    }
    const int c = colour_of((uint32_t)vloc, seed, d.barrier);
    if (c != colour) d.st[v] = (uint8_t)c;
    if (PUSH && c == ST_RED) clear_key(d, nresp, v);
    return 2 * c + hub;
}

// Partner of v after round r's responses, as a graph-local original id, from
// its request a = req[v], the word k read for it (PUSH: the proposal key at the
// red end, whose low half is the winner's local id; PULL: req[] of the other
// end) and pl = loc[a] (PULL only).  -1 when unmatched; PARTNER_WRITES for a
// blue vertex whose proposal won (PUSH); dead flags a blue vertex without
// unmatched neighbours.
template <bool PUSH>
__device__ __forceinline__ int32_t partner_of(int32_t v, int32_t vloc, int colour, int32_t a, unsigned long long k,
                                              int32_t pl, bool &dead)
{
    if (colour == ST_BLUE) {
        if (a == REQ_DEAD) {
            dead = true;
            return -1;
        }
        if (a < 0) return -1;
        if (PUSH) return (uint32_t)k == 0xFFFFFFFFu - (uint32_t)vloc ? PARTNER_WRITES : -1;
        return (int32_t)k == v ? pl : -1;
    }
    if (PUSH) return k != 0ull ? (int32_t)(0xFFFFFFFFu - (uint32_t)k) : -1;
    return (a >= 0 && (int32_t)k == v) ? pl : -1;
}

// Loads of partner_of's inputs for vertex v.
// split: the keys are split between resp and bucket_keys() (round 1, split1).
template <bool PUSH>
__device__ __forceinline__ void partner_words(const Dev &d, int32_t v, int colour, const unsigned long long *resp,
                                              bool split, int32_t &a, unsigned long long &k)
{
    a = -1;
    k = 0ull;
    if (PUSH && colour == ST_RED) k = key_at(d, resp, v, split);
    else a = d.req[v];
}
template <bool PUSH>
__device__ __forceinline__ void partner_words2(const Dev &d, int colour, const unsigned long long *resp, bool split,
                                               int32_t a, unsigned long long &k, int32_t &pl)
{
    pl = -1;
    if (a < 0) return;
    if (PUSH) {
        if (colour == ST_BLUE) k = key_at(d, resp, a, split);
    } else {
        k = (unsigned long long)(uint32_t)d.req[a];
        pl = d.loc[a];
    }
}

// Match write-back of round r fused with the select of round r+1: block b
// resolves bucket b of round r's lists (and a share of the hub lists),
// recolours the vertices still active and files them into bucket b of round
// r+1's lists.
// TIGHT: held to 64 VGPRs (8 wavefronts per SIMD, a few spilled values), so
// two blocks of 1024 threads share a CU and all the buckets' blocks of a
// 2^22-vertex graph run at once: its list rounds 2-5 took 30/20/14/9 instead
// of 35/23/17/12 us.  Batches of small graphs keep 72 registers in rounds 2-4
// (held there too, a batch of 4096 took 4.62 instead of 4.50 ms) and are held
// from round 5 on, when few vertices are left.  Round 1's pass over every id takes
// 2 vertices per thread instead of 4 (56 registers, no spills) and also runs
// 8 wavefronts per SIMD (61 instead of 64 us).
template <bool PUSH, bool LL = false, bool TIGHT = false>
__global__ __attribute__((amdgpu_flat_work_group_size(1, 1024), amdgpu_waves_per_eu((TIGHT || LL) ? 8 : 1)))
void k_match_select(Dev d, int r, int do_select)
{
    __shared__ int32_t s_cnt[2];
    const int parity = r & 1, nparity = parity ^ 1, b = bucket_of_block(d);
    if (threadIdx.x < 2) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    // LL (round 1 without lists): the items are the bucket's ids, every one
    // active; rows above the hub degree are left to the hub part below.
    const int64_t id0 = (int64_t)b * (LL ? d.bcap1 : d.bcap);
    const int32_t c0 = LL ? 0 : ctrb_ptr(d, r, 0)[b], c1 = LL ? 0 : ctrb_ptr(d, r, 1)[b];
    const int32_t total = LL ? (int32_t)max((int64_t)0, min((int64_t)d.n1, id0 + d.bcap1) - id0) : c0 + c1;
    const int4 *blue = list_ptr(d, parity, L_BLUE_LO) + (size_t)b * d.bcap;
    const int4 *red = list_ptr(d, parity, L_RED_LO) + (size_t)b * d.bcap;
    int4 *nblue = list_ptr(d, nparity, L_BLUE_LO) + (size_t)b * d.bcap;
    int4 *nred = list_ptr(d, nparity, L_RED_LO) + (size_t)b * d.bcap;
    const unsigned long long *resp = PUSH ? resp_slot(d, parity) : nullptr;
    unsigned long long *nresp = PUSH ? resp_slot(d, nparity) : nullptr;
    const bool split = PUSH && LL && d.split1;
    const uint32_t seed = do_select ? d.seeds[r] : 0u;
    int32_t n_matched = 0, n_dead = 0;
    unsigned long long deg_sum = 0, hub_deg_sum = 0;
    // MU vertices per thread: all loads of a step are issued before any store.
    constexpr int MU = LL ? 2 : 4;
    // PUSH list rounds: blue vertices' requests by list slot, loaded with the entries
    const int32_t *rq = PUSH && !LL && d.reqs ? d.reqs + (size_t)b * d.bcap : nullptr;
    for (int base = threadIdx.x & ~63; base < total; base += blockDim.x * MU) {
        int4 e[MU];
        int32_t v[MU], a[MU], vl[MU], pl[MU], ra[MU];
        unsigned long long k[MU];
        int col[MU];
#pragma unroll
        for (int q = 0; q < MU; ++q) {
            const int t = base + q * blockDim.x + (threadIdx.x & 63);
            col[q] = -1;
            ra[q] = -1;
            e[q] = make_int4(0, 0, 0, 0);
            if (LL) {
                if (t < total) {
                    const int32_t vv = (int32_t)(id0 + t);
                    const int2 rg = d.ranges[vv];
                    e[q] = make_int4(vv, rg.x, rg.y, d.loc[vv]);
                    col[q] = d.st[vv];
                    // (settled by the select: a vertex without neighbours)
                    if (col[q] >= ST_DEAD) col[q] = -1;
                }
            } else if (t < total) {
                col[q] = t < c0 ? ST_BLUE : ST_RED;
                e[q] = load_ent(t < c0 ? blue + t : red + (t - c0));
                if (rq && t < c0) ra[q] = rq[t];
            }
            v[q] = e[q].x;
            vl[q] = e[q].w;
        }
        if (LL) {
#pragma unroll
            for (int q = 0; q < MU; ++q)
                if (col[q] >= 0 && e[q].z - e[q].y > d.hub_deg) col[q] = -1;
        }
        // (with rq a blue vertex's key load goes out with the red ones')
#pragma unroll
        for (int q = 0; q < MU; ++q) {
            a[q] = -1;
            k[q] = 0ull;
            if (rq && col[q] == ST_BLUE) {
                a[q] = ra[q];
                if (a[q] >= 0) k[q] = key_at(d, resp, a[q], split);
            } else if (col[q] >= 0) {
                partner_words<PUSH>(d, v[q], col[q], resp, split, a[q], k[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < MU; ++q) {
            pl[q] = -1;
            if (col[q] >= 0 && !(rq && col[q] == ST_BLUE)) partner_words2<PUSH>(d, col[q], resp, split, a[q], k[q], pl[q]);
        }
#pragma unroll
        for (int q = 0; q < MU; ++q) {
            int cls = -1;
            if (col[q] >= 0) {
                bool dd = false;
                const int32_t partner = partner_of<PUSH>(v[q], vl[q], col[q], a[q], k[q], pl[q], dd);
                n_matched += partner >= 0;
                n_dead += dd;
                cls = settle_vertex<PUSH>(d, v[q], vl[q], col[q], 0, partner, dd, r, do_select, seed, nresp);
                if (cls >= 0 && d.count_nnz) deg_sum += (unsigned long long)(e[q].z - e[q].y);
            }
            bucket_append(cls == L_BLUE_LO, e[q], nblue, s_cnt + 0);
            bucket_append(cls == L_RED_LO, e[q], nred, s_cnt + 1);
        }
    }
    // hub rows: a share of the global hub lists
    const int32_t h0 = d.ctr[r * CTR_STRIDE + L_BLUE_HI], h1 = d.ctr[r * CTR_STRIDE + L_RED_HI];
    const int4 *hblue = list_ptr(d, parity, L_BLUE_HI), *hred = list_ptr(d, parity, L_RED_HI);
    for (int64_t base = (int64_t)b * blockDim.x + (threadIdx.x & ~63); base < (int64_t)h0 + h1;
         base += (int64_t)d.nb * blockDim.x) {
        const int64_t t = base + (threadIdx.x & 63);
        int cls = -1;
        int4 ent = make_int4(0, 0, 0, 0);
        if (t < (int64_t)h0 + h1) {
            const int colour = t < h0 ? ST_BLUE : ST_RED;
            ent = t < h0 ? hblue[t] : hred[t - h0];
            const int32_t v = ent.x, vloc = ent.w;
            int32_t a, pl;
            unsigned long long k;
            partner_words<PUSH>(d, v, colour, resp, split, a, k);
            partner_words2<PUSH>(d, colour, resp, split, a, k, pl);
            bool dd = false;
            const int32_t partner = partner_of<PUSH>(v, vloc, colour, a, k, pl, dd);
            n_matched += partner >= 0;
            n_dead += dd;
            cls = settle_vertex<PUSH>(d, v, vloc, colour, 1, partner, dd, r, do_select, seed, nresp);
            if (cls >= 0 && d.count_nnz) hub_deg_sum += (unsigned long long)(ent.z - ent.y);
        }
        wave_append(cls == L_BLUE_HI, ent, list_ptr(d, nparity, L_BLUE_HI), d.ctr + (r + 1) * CTR_STRIDE + L_BLUE_HI);
        wave_append(cls == L_RED_HI, ent, list_ptr(d, nparity, L_RED_HI), d.ctr + (r + 1) * CTR_STRIDE + L_RED_HI);
    }
    // the matched and dead counts share one reduction (low and high halves)
    const unsigned long long md = block_sum((unsigned long long)(uint32_t)n_matched |
                                            ((unsigned long long)(uint32_t)n_dead << 32));
    n_matched = (int32_t)(uint32_t)md;
    n_dead = (int32_t)(md >> 32);
    if (d.count_nnz) {
        deg_sum = block_sum(deg_sum);
        hub_deg_sum = block_sum(hub_deg_sum);
        if (threadIdx.x == 0 && do_select) {
            ctrb64_ptr(d, r + 1, 0)[b] = deg_sum;
            if (hub_deg_sum) atomicAdd(d.hub64 + (size_t)(r + 1) * 2 + 0, hub_deg_sum);
        }
    }
    // (every wavefront made its appends to s_cnt before entering block_sum,
    // whose barriers order them before thread 0's reads)
    if (threadIdx.x == 0) {
        ctrb_ptr(d, r, 2)[b] = n_matched;
        ctrb_ptr(d, r, 3)[b] = n_dead;
        if (do_select) {
            ctrb_ptr(d, r + 1, 0)[b] = s_cnt[0];
            ctrb_ptr(d, r + 1, 1)[b] = s_cnt[1];
        }
    }
}

// Sums the per-bucket counters of rounds [r0, r0 + gridDim.x) into the
// per-round totals (one block of up to 1024 threads per round; every thread
// issues all its bucket loads before adding, and the eight sums share one
// wavefront reduction and one LDS pass).  Assigns, so it may be re-run.
__global__ __attribute__((amdgpu_flat_work_group_size(1, 1024))) void k_reduce_counts(Dev d, int r0, int with_nnz)
{
    const int r = r0 + blockIdx.x;
    int32_t a[6] = {0, 0, 0, 0, 0, 0};
    unsigned long long z[2] = {0ull, 0ull};
    const int32_t *iso_last = d.iso ? d.iso + (size_t)2 * ROUND_SLOTS * d.nb : nullptr;
    for (int b = threadIdx.x; b < d.nb; b += blockDim.x) {
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] += ctrb_ptr(d, r, j)[b];
        if (with_nnz) {
            z[0] += ctrb64_ptr(d, r, 0)[b];
            z[1] += ctrb64_ptr(d, r, 1)[b];
        }
        // vertices without neighbours settled by the round-1 select (Dev::iso);
        // a bucket's slots past its last round hold stale values
        if (iso_last) {
            const int32_t last = iso_last[b];
            const int32_t ib = d.iso[(size_t)(2 * r) * d.nb + b], ir = d.iso[(size_t)(2 * r + 1) * d.nb + b];
            if (r <= last) {
                a[4] += ib;
                a[5] += ir;
            }
        }
    }
    __shared__ int32_t s_a[16][6];
    __shared__ unsigned long long s_z[16][2];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 6; ++j) a[j] = wave_sum(a[j]);
    if (with_nnz) {
        z[0] = wave_sum(z[0]);
        z[1] = wave_sum(z[1]);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int j = 0; j < 6; ++j) s_a[w][j] = a[j];
        s_z[w][0] = z[0];
        s_z[w][1] = z[1];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int x = 1; x < (int)((blockDim.x + 63) >> 6); ++x) {
#pragma unroll
            for (int j = 0; j < 6; ++j) a[j] += s_a[x][j];
            z[0] += s_z[x][0];
            z[1] += s_z[x][1];
        }
        int32_t *c = d.ctr + (size_t)r * CTR_STRIDE;
        c[L_BLUE_LO] = a[0] + a[4];
        c[L_RED_LO] = a[1] + a[5];
        c[4] = a[2];
        c[5] = a[3] + a[4];
        if (with_nnz) {
            d.ctr64[(size_t)r * CTR64_STRIDE + 0] = z[0] + d.hub64[(size_t)r * 2 + 0];
            d.ctr64[(size_t)r * CTR64_STRIDE + 1] = z[1] + d.hub64[(size_t)r * 2 + 1];
        }
    }
}

// Writes the state array back in the caller's vertex numbering: coalesced
// writes, reads gathered from the (L2/MALL-resident) state array.
// Offsets of the eight groups of (output position, internal id) pairs.
struct UnpermOff {
    int32_t off[9];
};

// out[o] = match[v] for every pair {o, v}.  The pairs are grouped by the
// eighth of the internal ids that v falls in (ordered by o within a group), and
// workgroup i works on group i % 8: workgroups are dealt to the 8 XCDs round
// robin, so each XCD gathers from one 1/8 slice of match[], which stays in its
// own L2, instead of every XCD missing on random lines of the whole array.
__global__ void k_unpermute(const int2 *__restrict__ pairs, UnpermOff g, const int32_t *__restrict__ match,
                            int32_t *__restrict__ out)
{
    const int x = blockIdx.x & 7;
    const int64_t step = (int64_t)(gridDim.x >> 3) * blockDim.x;
    for (int64_t k = g.off[x] + (int64_t)(blockIdx.x >> 3) * blockDim.x + threadIdx.x; k < g.off[x + 1]; k += step) {
        typedef int v2i __attribute__((ext_vector_type(2)));
        const v2i p = __builtin_nontemporal_load(reinterpret_cast<const v2i *>(pairs + k));
        out[p.x] = match[p.y];
    }
}

// Rounds each batched graph would execute alone: one plus the last round in
// which one of its vertices was still active (src/matchgpu.cu:604 loop).
// Three launches: clear the per-graph maxima, fold the vertices' last rounds
// into them, turn them into round counts.  The fold gives each thread a run
// of 64 consecutive vertices (graphs are contiguous id ranges, found by a
// binary search over the graph offsets for the run's first vertex) and posts
// one atomic max per graph the run touches.
__global__ void k_graph_clear(int32_t *gmax, int32_t n_graphs)
{
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_graphs; i += gridDim.x * blockDim.x) gmax[i] = 0;
}

__global__ void k_graph_fold(const uint8_t *vround, const int32_t *voff, int32_t n_graphs, int32_t *gmax)
{
    constexpr int RUN = 64;
    const int32_t n = voff[n_graphs];
    const int64_t nruns = ((int64_t)n + RUN - 1) / RUN;
    // wave-uniform loop: the wave's maxima are combined before the atomics,
    // which all go to the same few addresses
    for (int64_t t0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); t0 < nruns;
         t0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = t0 + (threadIdx.x & 63);
        const bool valid = t < nruns;
        const int32_t v0 = valid ? (int32_t)(t * RUN) : 0, v1 = valid ? min(n, v0 + RUN) : 0;
        int32_t lo = 0, hi = n_graphs - 1;  // last graph with voff[g] <= v0
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (voff[mid] <= v0) lo = mid;
            else hi = mid - 1;
        }
        // the run's bytes: four 16-byte loads (a full run lies inside the array)
        uint32_t wds[RUN / 4];
        if (v1 - v0 == RUN) {
            const uint4 *p = reinterpret_cast<const uint4 *>(vround + v0);
#pragma unroll
            for (int q = 0; q < RUN / 16; ++q) {
                const uint4 c = p[q];
                wds[4 * q] = c.x;
                wds[4 * q + 1] = c.y;
                wds[4 * q + 2] = c.z;
                wds[4 * q + 3] = c.w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < RUN / 4; ++q) wds[q] = 0;
            for (int32_t v = v0; v < v1; ++v) wds[(v - v0) >> 2] |= (uint32_t)vround[v] << (8 * ((v - v0) & 3));
        }
        int32_t g = lo, end = voff[g + 1], x = 0;
        if (end >= v1) {
            // one graph (the usual case): the byte maximum of the run
#pragma unroll
            for (int q = 0; q < RUN / 4; ++q) {
                const uint32_t w = wds[q];
                x = max(x, (int32_t)max(max(w & 0xFFu, (w >> 8) & 0xFFu), max((w >> 16) & 0xFFu, w >> 24)));
            }
        } else {
            for (int32_t v = v0; v < v1; ++v) {
                while (v >= end) {  // (empty graphs are skipped)
                    if (x) atomicMax(gmax + g, x);
                    x = 0;
                    ++g;
                    end = voff[g + 1];
                }
                x = max(x, (int32_t)((wds[(v - v0) >> 2] >> (8 * ((v - v0) & 3))) & 0xFFu));
            }
        }
        // (g, x): the last graph this run touched and its maximum so far
        const int32_t g0 = __shfl(g, 0);
        if (__ballot(valid && g != g0) == 0ull) {
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) x = max(x, __shfl_xor(x, off, 64));
            if ((threadIdx.x & 63) == 0 && x) atomicMax(gmax + g0, x);
        } else if (valid && x) {
            atomicMax(gmax + g, x);
        }
    }
}

__global__ void k_graph_rounds(const int32_t *gmax, const int32_t *voff, int32_t n_graphs, int32_t R_run,
                               int32_t R_max, int32_t maximal, int32_t *out)
{
    for (int32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n_graphs; g += gridDim.x * blockDim.x) {
        const int32_t x = gmax[g];
        int32_t rounds;
        if (!maximal) rounds = R_max;
        else if (voff[g + 1] == voff[g]) rounds = 1;
        else if (x == 255) rounds = R_run;
        else rounds = min(R_max, x + 1);
        out[g] = rounds;
    }
}


}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------


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
    void *ptrs[] = {g->d_ranges, g->d_nbrs,  g->d_adj,    g->d_adjp,   g->d_head,   g->d_headc,  g->d_adjc,  g->d_match, g->d_req,   g->d_st,   g->d_reqs,
                    g->d_resp,   g->d_lists, g->d_ctr,    g->d_ctrb,  g->d_ctr64, g->d_ctrb64,
                    g->d_hub64,  g->d_iso,   g->d_seeds, g->d_loc,    g->d_voff,  g->d_vround, g->d_grounds,
                    g->d_unperm, g->d_match_out, g->d_lcol, g->d_lreqw, g->d_lfeat, g->d_lctr, g->d_lpart, g->d_lcnt, g->d_llist, g->d_params, g->d_vbase, g->d_chk, g->d_chk_keys, g->d_wcsr,
                    g->d_lsamp};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (g->h_ctr) hipHostFree(g->h_ctr);
    if (g->h_ctr64) hipHostFree(g->h_ctr64);
    if (g->h_lctr) hipHostFree(g->h_lctr);
    if (g->ev_t1) hipEventDestroy(g->ev_t1);
    if (g->ev_t2) hipEventDestroy(g->ev_t2);
    for (hipEvent_t e : g->prof_ev) hipEventDestroy(e);
    for (CachedExec &ce : g->execs) hipGraphExecDestroy(ce.exec);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
}

// A graph handle from the caller's CSR.  Batches (voff / eoff / n_graphs):
// ranges and neighbour ids local to each graph.  relabel: renumber the
// vertices internally in a locality order; align_rows: start a row at the
// next 128-byte line when its first 8 entries would straddle one; heads: keep
// the sorted rows' first 4 entries in a dense array.  The host checks the CSR;
// the device builds everything else (mm_upload.hip).
int32_t create_graph(int32_t device, int32_t n, int64_t nnz, const int32_t *ranges, const int32_t *nbrs,
                     const float *w, const int32_t *voff, const int64_t *eoff, int32_t n_graphs, int relabel,
                     bool align_rows, bool heads, mm_graph **out)
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
    mm_upload_mark(nullptr, nullptr);
    int32_t rc;
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
    mm_upload_mark("device setup", nullptr);
    UploadState *up = nullptr;
    if ((rc = mm_upload_begin(g, ranges, nbrs, w, voff, eoff, n_graphs, relabel, &up)) != MM_OK) return fail(rc);
    struct UpGuard {
        UploadState *p;
        ~UpGuard() { mm_upload_free(p); }
    } up_guard{up};
    const size_t nn = (size_t)n;
    if ((rc = dev_alloc(g, &g->d_match, nn, "match")) != MM_OK) return fail(rc);
    if ((rc = dev_alloc(g, &g->d_req, nn, "requests")) != MM_OK) return fail(rc);
    if ((rc = dev_alloc(g, &g->d_st, nn, "state")) != MM_OK) return fail(rc);
    // Buckets: at most one per block of the smallest block size (64 threads).
    g->nb_max = std::max(1, g->num_cu * 32);
    g->list_cap = (int64_t)n + g->nb_max;
    if ((rc = dev_alloc(g, &g->d_lists, 2 * NLISTS * (size_t)g->list_cap, "work lists")) != MM_OK) return fail(rc);
    if ((rc = dev_alloc(g, &g->d_reqs, (size_t)g->list_cap, "requests by list slot")) != MM_OK) return fail(rc);
    static_assert(sizeof(*g->d_lists) == 16, "work-list entries are 16 bytes");
    if ((rc = dev_alloc(g, &g->d_ctr, (size_t)2 * ROUND_SLOTS * CTR_STRIDE, "counters")) != MM_OK) return fail(rc);
    if ((rc = dev_alloc(g, &g->d_ctrb, (size_t)ROUND_SLOTS * 4 * g->nb_max, "bucket counters")) != MM_OK)
        return fail(rc);
    if ((rc = dev_alloc(g, &g->d_ctr64, (size_t)ROUND_SLOTS * CTR64_STRIDE, "counters")) != MM_OK) return fail(rc);
    if ((rc = dev_alloc(g, &g->d_ctrb64, (size_t)ROUND_SLOTS * 2 * g->nb_max, "bucket counters")) != MM_OK)
        return fail(rc);
    if ((rc = dev_alloc(g, &g->d_hub64, (size_t)2 * ROUND_SLOTS * 2, "counters")) != MM_OK) return fail(rc);
    int64_t with_nbrs = 0;
    for (int k = 0; k < 33; ++k) with_nbrs += g->deg_hist[k];
    g->n_isolated = (int64_t)n - with_nbrs;
    if (g->n_isolated * 100 >= (int64_t)n && g->n_isolated > 0 &&
        (rc = dev_alloc(g, &g->d_iso, ((size_t)ROUND_SLOTS * 2 + 1) * g->nb_max, "isolated counters")) != MM_OK)
        return fail(rc);
    if (hipMemset(g->d_ctr, 0, sizeof(int32_t) * 2 * ROUND_SLOTS * CTR_STRIDE) != hipSuccess ||
        hipMemset(g->d_hub64, 0, sizeof(unsigned long long) * 2 * ROUND_SLOTS * 2) != hipSuccess)
        return fail(mm::set_error(MM_ERR_HIP, "Unable to clear counters"));
    if ((rc = dev_alloc(g, &g->d_seeds, (size_t)MM_MAX_ROUNDS + 1, "seeds")) != MM_OK) return fail(rc);
    if (hipHostMalloc((void **)&g->h_ctr, sizeof(int32_t) * ROUND_SLOTS * CTR_STRIDE) != hipSuccess ||
        hipHostMalloc((void **)&g->h_ctr64, sizeof(unsigned long long) * ROUND_SLOTS * CTR64_STRIDE) != hipSuccess)
        return fail(mm::set_error(MM_ERR_NOMEM, "pinned host allocation failed"));
    mm_upload_mark("matching buffers", nullptr);
    if ((rc = mm_upload_finish(g, up, align_rows, heads)) != MM_OK) return fail(rc);
    // (the symmetry test, inside the finish, decides whether proposals are pushed)
    if (g->push_ok_struct && (rc = dev_alloc(g, &g->d_resp, 2 * nn, "proposal keys")) != MM_OK) return fail(rc);
    *out = g;
    return MM_OK;
}

// Lanes per row of the first-candidate scans (0: auto = 4); MM_LANES_TILES
// selects the entry-parallel tile kernels (G = 0 internally).
int pick_lanes(const mm_graph *, int requested)
{
    if (requested == MM_LANES_TILES) return 0;
    if (requested < 0) return -1;
    return requested > 0 ? requested : 4;
}

int64_t hubs_above(const mm_graph *g, int32_t thr)
{
    // Upper bound on the vertices with degree > thr, from the log2 histogram.
    int64_t s = 0;
    for (int b = 0; b < 33; ++b)
        if (b >= 31 || ((int64_t)1 << (b + 1)) > (int64_t)thr) s += g->deg_hist[b];
    return s;
}

// Dynamic LDS of the first-candidate scan kernels: the deferred-row buffer.
size_t deferred_smem_bytes(int block) { return sizeof(int4) * (size_t)block; }
// LDS of a round-1 head-scan block: the deferred rows, then (split1) the bucket's keys.
size_t head_smem_bytes(int block, int bcap, bool split) { return deferred_smem_bytes(block) + (split ? 4 * (size_t)bcap : 0); }
constexpr size_t HEAD_SMEM_MAX = 60 * 1024;  // (static LDS of the kernel comes on top)

struct RunCfg {
    bool W, A8, PUSH;
    bool HEAD;  // round 1 runs without lists, one lane per row over the head copy
    bool PK;    // the list rounds' first-candidate scans read the packed 4-byte rows
    int G, U, block, hub_deg, nb, bcap, hub_blocks;
};

template <bool W, int A8, bool PUSH, int G>
void launch_request_t(const RunCfg &c, const Dev &d, int r, hipStream_t s)
{
    hipLaunchKernelGGL((k_request<W, A8, PUSH, G>), dim3(c.nb + c.hub_blocks), dim3(c.block),
                       deferred_smem_bytes(c.block), s, d, r, c.hub_blocks);
}

template <bool W, int A8, int G>
void launch_respond_t(const RunCfg &c, const Dev &d, int r, hipStream_t s)
{
    hipLaunchKernelGGL((k_respond<W, A8, G>), dim3(c.nb + c.hub_blocks), dim3(c.block),
                       deferred_smem_bytes(c.block), s, d, r, c.hub_blocks);
}

template <bool W, int A8, bool PUSH>
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

template <bool W, int A8>
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

template <bool W, int A8, bool PUSH>
void launch_request_tile(const RunCfg &c, const Dev &d, int r, hipStream_t s)
{
    const dim3 grid(c.nb + c.hub_blocks), blk(c.block);
    const size_t sm = tile_smem_bytes(c.block);
    if (c.U == 4) hipLaunchKernelGGL((k_request_tile<W, A8, PUSH, 4>), grid, blk, sm, s, d, r, c.hub_blocks);
    else if (c.U == 8) hipLaunchKernelGGL((k_request_tile<W, A8, PUSH, 8>), grid, blk, sm, s, d, r, c.hub_blocks);
    else hipLaunchKernelGGL((k_request_tile<W, A8, PUSH, 16>), grid, blk, sm, s, d, r, c.hub_blocks);
}

void launch_request(const RunCfg &c, const Dev &d, int r, hipStream_t s)
{
    if (r == 1 && c.HEAD) {
        const dim3 grid(c.nb + c.hub_blocks), blk(c.block);
        const size_t sm = head_smem_bytes(c.block, c.bcap, d.split1 != 0);
        // (HEAD implies PUSH)
        if (d.headc) hipLaunchKernelGGL((k_request<true, true, true, 4, true, true>), grid, blk, sm, s, d, r, c.hub_blocks);
        else hipLaunchKernelGGL((k_request<true, true, true, 4, true, false>), grid, blk, sm, s, d, r, c.hub_blocks);
        return;
    }
    if (c.G == 0) {
        if (c.W) {
            if (c.A8) { if (c.PUSH) launch_request_tile<true, true, true>(c, d, r, s); else launch_request_tile<true, true, false>(c, d, r, s); }
            else { if (c.PUSH) launch_request_tile<true, false, true>(c, d, r, s); else launch_request_tile<true, false, false>(c, d, r, s); }
        } else {
            if (c.A8) { if (c.PUSH) launch_request_tile<false, true, true>(c, d, r, s); else launch_request_tile<false, true, false>(c, d, r, s); }
            else { if (c.PUSH) launch_request_tile<false, false, true>(c, d, r, s); else launch_request_tile<false, false, false>(c, d, r, s); }
        }
        return;
    }
    if (c.PK) {  // (implies weighted PUSH scans with 32-bit keys)
        launch_request_g<true, 2, true>(c, d, r, s);
        return;
    }
    if (c.W) {
        if (c.A8) { if (c.PUSH) launch_request_g<true, true, true>(c, d, r, s); else launch_request_g<true, true, false>(c, d, r, s); }
        else { if (c.PUSH) launch_request_g<true, false, true>(c, d, r, s); else launch_request_g<true, false, false>(c, d, r, s); }
    } else {
        if (c.A8) { if (c.PUSH) launch_request_g<false, true, true>(c, d, r, s); else launch_request_g<false, true, false>(c, d, r, s); }
        else { if (c.PUSH) launch_request_g<false, false, true>(c, d, r, s); else launch_request_g<false, false, false>(c, d, r, s); }
    }
}

template <bool W, int A8>
void launch_respond_tile(const RunCfg &c, const Dev &d, int r, hipStream_t s)
{
    const dim3 grid(c.nb + c.hub_blocks), blk(c.block);
    const size_t sm = tile_smem_bytes(c.block);
    if (c.U == 4) hipLaunchKernelGGL((k_respond_tile<W, A8, 4>), grid, blk, sm, s, d, r, c.hub_blocks);
    else if (c.U == 8) hipLaunchKernelGGL((k_respond_tile<W, A8, 8>), grid, blk, sm, s, d, r, c.hub_blocks);
    else hipLaunchKernelGGL((k_respond_tile<W, A8, 16>), grid, blk, sm, s, d, r, c.hub_blocks);
}

void launch_respond(const RunCfg &c, const Dev &d, int r, hipStream_t s)
{
    if (c.G == 0) {
        if (c.W) { if (c.A8) launch_respond_tile<true, true>(c, d, r, s); else launch_respond_tile<true, false>(c, d, r, s); }
        else { if (c.A8) launch_respond_tile<false, true>(c, d, r, s); else launch_respond_tile<false, false>(c, d, r, s); }
        return;
    }
    if (c.W) { if (c.A8) launch_respond_g<true, true>(c, d, r, s); else launch_respond_g<true, false>(c, d, r, s); }
    else { if (c.A8) launch_respond_g<false, true>(c, d, r, s); else launch_respond_g<false, false>(c, d, r, s); }
}

// Waits for the stream: polls for up to 2 ms (a step of a large graph ends
// within that, and polling wakes the caller about 2 us sooner than a blocking
// wait), then blocks.
hipError_t wait_stream(hipStream_t s)
{
    const auto t0 = std::chrono::steady_clock::now();
    hipError_t q;
    while ((q = hipStreamQuery(s)) == hipErrorNotReady)
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) return hipStreamSynchronize(s);
    return q;
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
    if (G < 0 || G > 64 || (G & (G - 1))) return mm::set_error(MM_ERR_INVALID, "mm_match: bad lanes_per_vertex");
    RunCfg c;
    c.W = o->kind == MM_WEIGHTED;
    // Weighted matching reads the weight-sorted rows; random matching the CSR-order ids.
    c.A8 = c.W && g->weighted != 0;
    const bool push_ok = c.W && g->weighted ? g->push_ok_weights : g->push_ok_struct;
    if (o->respond == MM_RESPOND_PUSH && !push_ok)
        return mm::set_error(MM_ERR_INVALID, "mm_match: PUSH needs a symmetric CSR with ascending rows and symmetric weights");
    if (o->respond < 0 || o->respond > MM_RESPOND_PUSH) return mm::set_error(MM_ERR_INVALID, "mm_match: bad respond mode");
    c.PUSH = o->respond == MM_RESPOND_PUSH || (o->respond == MM_RESPOND_AUTO && push_ok);
    c.G = G;
    // 32-bit keys with the first-candidate scans, when the graph has the rows for them
    const bool key32 = c.PUSH && c.A8 && G != 0 && g->d_adjp != nullptr;
    // The head copy (4-byte entries only with 32-bit keys) serves the first
    // round of weighted PUSH scans (not the tiles); that round then runs
    // without lists.
    c.HEAD = c.A8 && c.PUSH && G != 0 && ((key32 && g->d_headc != nullptr) || g->d_head != nullptr);
    c.PK = key32 && g->d_adjc != nullptr;
    c.U = o->entries_per_thread > 0 ? o->entries_per_thread : 4;
    if (c.U != 4 && c.U != 8 && c.U != 16)
        return mm::set_error(MM_ERR_INVALID, "mm_match: entries_per_thread must be 4, 8 or 16");
    c.block = block;
    c.hub_deg = o->hub_degree > 0 ? o->hub_degree : 256;
    const int blocks_per_cu = std::max(1, 2048 / block);
    // One bucket per block of the per-bucket kernels, about 2 * block ids each.
    c.nb = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)g->n + 2 * block - 1) / (2 * block),
                                                        std::min<int64_t>(g->nb_max, (int64_t)g->num_cu * blocks_per_cu)));
    c.nb = (c.nb + 7) & ~7;  // bucket_of_block deals buckets to XCDs in eighths
    c.bcap = (int)(((int64_t)g->n + c.nb - 1) / c.nb);
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

    Dev d;
    d.n = g->n;
    d.cap = (int32_t)g->list_cap;
    d.hub_deg = c.hub_deg;
    d.barrier = o->barrier;
    d.count_nnz = o->profile ? 1 : 0;
    d.nb = c.nb;
    d.bcap = c.bcap;
    d.ranges = g->d_ranges;
    d.nbrs = g->d_nbrs;
    d.adj = g->d_adj;
    d.head = c.HEAD ? g->d_head : nullptr;
    d.headc = c.HEAD && key32 ? g->d_headc : nullptr;
    if (c.HEAD && !d.head && !d.headc) return mm::set_error(MM_ERR_INVALID, "mm_match: no head copy for round 1");
    // (the packed rows exist wherever the packed head copy does)
    d.adjc = key32 ? g->d_adjc : nullptr;
    if (d.headc && !d.adjc) return mm::set_error(MM_ERR_INVALID, "mm_match: no packed rows beside the packed head copy");
    d.key32 = key32 ? 1 : 0;
    d.listless1 = c.HEAD ? 1 : 0;
    d.no_hubs = g->max_degree <= c.hub_deg ? 1 : 0;
    // proposals inside a bucket meet in LDS in round 1 when the bucket's keys fit
    d.split1 = c.HEAD && key32 && head_smem_bytes(c.block, c.bcap, true) <= HEAD_SMEM_MAX ? 1 : 0;
    d.head_pb = g->head_pb;
    d.locbits = g->locbits;
    if (d.key32) d.adj = g->d_adjp;
    // Vertices write their final values straight into the caller-order result
    // as they settle, so no permutation pass follows the last round.
    d.out = g->relabelled ? g->d_match_out : g->d_match;
    d.vbase = g->d_vbase;
    d.req = g->d_req;
    // the first-candidate request scans of PUSH rounds leave the bucket lists'
    // requests by list slot (the entry-parallel tiles write them by vertex)
    d.reqs = c.PUSH && c.G != 0 ? g->d_reqs : nullptr;
    d.st = g->d_st;
    d.resp = g->d_resp;
    d.lists = g->d_lists;
    // This run's counter set; its init clears the other set for the next run,
    // so the next run switches sets once this run's init has been submitted
    // (whatever happens after that).
    const int cset = g->ctr_set;
    struct SetFlip {
        mm_graph *g;
        bool init_submitted = false;
        ~SetFlip()
        {
            if (init_submitted) g->ctr_set ^= 1;
        }
    } flip{g};
    d.ctr = g->d_ctr + (size_t)cset * ROUND_SLOTS * CTR_STRIDE;
    d.ctr_next = g->d_ctr + (size_t)(cset ^ 1) * ROUND_SLOTS * CTR_STRIDE;
    d.ctrb = g->d_ctrb;
    d.ctr64 = g->d_ctr64;
    d.ctrb64 = g->d_ctrb64;
    d.hub64 = g->d_hub64 + (size_t)cset * ROUND_SLOTS * 2;
    d.hub64_next = g->d_hub64 + (size_t)(cset ^ 1) * ROUND_SLOTS * 2;
    d.seeds = g->d_seeds;
    d.loc = g->d_loc;
    d.vround = g->d_vround;
    d.iso = g->d_iso;
    // round 1 scans by id only the ids before the isolated vertices the select settles
    d.n1 = g->d_iso && g->iso_lo < g->n ? g->iso_lo : g->n;
    d.bcap1 = (int32_t)(((int64_t)d.n1 + c.nb - 1) / c.nb);

    const bool prof = o->profile != 0;
    const bool use_graph = !prof && o->graph_mode == 0;
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

    // A maximal run may stop early; it can only go as far as the seeds given.
    const int R_run = maximal ? std::min(R_max, o->n_seeds) : R_max;
    d.R_run = R_run;
    // Fixed runs are one launch sequence; maximal runs check the round
    // counters after every chunk of rounds (rounds past convergence are no-ops).
    // The first chunk is 4 rounds; later ones last as long as the decay of the
    // active count predicts the run still needs, so a typical run reads the
    // counters back twice.
    int chunk = maximal ? 4 : R_run;
    auto enq_reduce = [&](int r0, int r1) -> int32_t {
        MM_LAUNCH(K_CLEAR, hipLaunchKernelGGL(k_reduce_counts, dim3(r1 - r0 + 1), dim3(1024), 0, g->stream, d, r0,
                                              d.count_nnz));
        return MM_OK;
    };
    // Enqueues rounds [r0, r1] (and the clear + round-1 select when r0 == 1);
    // maximal runs also total the counters of rounds [r0, r1 + 1].
    auto enqueue = [&](int r0, int r1) -> int32_t {
        if (r0 == 1) {
            if (d.iso && c.PUSH) MM_LAUNCH(K_INIT, hipLaunchKernelGGL(k_init_select_iso<true>, dim3(c.nb), dim3(block), 0, g->stream, d));
            else if (d.iso) MM_LAUNCH(K_INIT, hipLaunchKernelGGL(k_init_select_iso<false>, dim3(c.nb), dim3(block), 0, g->stream, d));
            else if (c.PUSH) MM_LAUNCH(K_INIT, hipLaunchKernelGGL(k_init_select<true>, dim3(c.nb), dim3(block), 0, g->stream, d));
            else MM_LAUNCH(K_INIT, hipLaunchKernelGGL(k_init_select<false>, dim3(c.nb), dim3(block), 0, g->stream, d));
            if (!use_graph) flip.init_submitted = true;
        }
        for (int rr = r0; rr <= r1; ++rr) {
            const int do_select = rr < R_run;
            // From round 3 on the rows left are few and their first entries
            // mostly matched: twice the lanes per row (fewer dependent steps).
            RunCfg cr = c;
            if (rr >= 3 && c.G > 0) cr.G = std::min(64, 2 * c.G);
            MM_LAUNCH(K_REQUEST, launch_request(cr, d, rr, g->stream));
            if (!c.PUSH) MM_LAUNCH(K_RESPOND, launch_respond(cr, d, rr, g->stream));
            const bool ll = rr == 1 && d.listless1;
            // the 64-register match build: a single graph's list rounds, and a
            // batch's from round 5, when its buckets hold few vertices and two
            // blocks per CU halve the launch's block waves (batch of 4096:
            // 4.47 -> 4.42 ms; from round 3 on it is slower)
            bool tight = g->n_graphs == 0 || rr >= 5;
            // From round 3 on a single graph's buckets hold a few hundred
            // vertices: 512-thread blocks of the unbounded build (no spills,
            // three blocks per CU) beat 1024-thread blocks of the held one
            // (RGG-22: 0.315 -> 0.308 ms per step; 256 threads, or from round 2
            // on, are slower).
            int mblock = block;
            if (g->n_graphs == 0 && rr >= 3 && block > 512) {
                mblock = 512;
                tight = false;
            }
            if (c.PUSH && ll) MM_LAUNCH(K_MATCH, hipLaunchKernelGGL((k_match_select<true, true>), dim3(c.nb), dim3(block), 0, g->stream, d, rr, do_select));
            else if (c.PUSH && tight) MM_LAUNCH(K_MATCH, hipLaunchKernelGGL((k_match_select<true, false, true>), dim3(c.nb), dim3(mblock), 0, g->stream, d, rr, do_select));
            else if (c.PUSH) MM_LAUNCH(K_MATCH, hipLaunchKernelGGL((k_match_select<true, false>), dim3(c.nb), dim3(mblock), 0, g->stream, d, rr, do_select));
            else if (ll) MM_LAUNCH(K_MATCH, hipLaunchKernelGGL((k_match_select<false, true>), dim3(c.nb), dim3(block), 0, g->stream, d, rr, do_select));
            else if (tight) MM_LAUNCH(K_MATCH, hipLaunchKernelGGL((k_match_select<false, false, true>), dim3(c.nb), dim3(mblock), 0, g->stream, d, rr, do_select));
            else MM_LAUNCH(K_MATCH, hipLaunchKernelGGL((k_match_select<false, false>), dim3(c.nb), dim3(mblock), 0, g->stream, d, rr, do_select));
        }
        if (maximal) return enq_reduce(r0, std::min(r1 + 1, R_run));
        return MM_OK;
    };
    // Replays the captured launch sequence of rounds [r0, r1], capturing it on first use.
    ExecKey key;
    std::memset(&key, 0, sizeof(key));
    key.R_run = R_run; key.W = c.W; key.A8 = c.A8; key.PUSH = c.PUSH; key.G = c.G; key.U = c.U; key.block = c.block;
    key.hub_deg = c.hub_deg; key.nb = c.nb; key.bcap = c.bcap; key.maximal = maximal;
    key.hub_blocks = c.hub_blocks; key.barrier = o->barrier; key.count_nnz = d.count_nnz; key.cset = cset;
    auto run_chunk = [&](int r0, int r1) -> int32_t {
        if (!use_graph) return enqueue(r0, r1);
        key.r0 = r0;
        key.r1 = r1;
        for (const CachedExec &ce : g->execs)
            if (std::memcmp(&ce.key, &key, sizeof(key)) == 0) {
                MM_HIP_TRY(hipGraphLaunch(ce.exec, g->stream), "hipGraphLaunch");
                if (r0 == 1) flip.init_submitted = true;
                return MM_OK;
            }
        hipGraph_t graph = nullptr;
        MM_HIP_TRY(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        const int32_t erc = enqueue(r0, r1);
        const hipError_t ec = hipStreamEndCapture(g->stream, &graph);
        if (erc != MM_OK) { if (graph) hipGraphDestroy(graph); return erc; }
        MM_HIP_TRY(ec, "hipStreamEndCapture");
        CachedExec ce;
        ce.key = key;
        const hipError_t ei = hipGraphInstantiate(&ce.exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        MM_HIP_TRY(ei, "hipGraphInstantiate");
        if (g->execs.size() >= 256) {
            hipGraphExecDestroy(g->execs.front().exec);
            g->execs.erase(g->execs.begin());
        }
        g->execs.push_back(ce);
        MM_HIP_TRY(hipGraphLaunch(ce.exec, g->stream), "hipGraphLaunch");
        if (r0 == 1) flip.init_submitted = true;
        return MM_OK;
    };

    // (the events time the rounds for stats->match_ms only)
    if (stats) MM_HIP_TRY(hipEventRecord(g->ev_t1, g->stream), "hipEventRecord");
    int rounds = -1;
    int r = 1;
    if (g->n > 0) {
        while (r <= R_run) {
            const int end = std::min(R_run, r + chunk - 1);
            if ((rc = run_chunk(r, end)) != MM_OK) return rc;
            MM_HIP_TRY(hipGetLastError(), "A HIP error occurred during the matching process");
            if (maximal) {
                const int k_hi = std::min(end + 1, R_run);
                MM_HIP_TRY(hipMemcpyAsync(g->h_ctr + (size_t)r * CTR_STRIDE, d.ctr + (size_t)r * CTR_STRIDE,
                                          sizeof(int32_t) * (size_t)(k_hi - r + 1) * CTR_STRIDE,
                                          hipMemcpyDeviceToHost, g->stream),
                           "Unable to read round counters");
                MM_HIP_TRY(wait_stream(g->stream), "A HIP error occurred during the matching process");
                auto active = [&](int k) {
                    const int32_t *q = g->h_ctr + (size_t)k * CTR_STRIDE;
                    return (int64_t)q[0] + q[1] + q[2] + q[3];
                };
                for (int k = r; k <= k_hi; ++k)
                    if (active(k) == 0) { rounds = k; break; }
                if (rounds > 0) break;
                chunk = predicted_rounds(k_hi > r ? active(k_hi - 1) : 0, active(k_hi));
            }
            r = end + 1;
        }
    }
    if (maximal && rounds < 0 && R_run < R_max && g->n > 0)
        return mm::set_error(MM_ERR_SEEDS, "mm_match: ran out of seeds before the matching converged");
    if (g->n == 0) rounds = maximal ? 1 : R_max;
    if (rounds < 0) rounds = R_max;
    // The result is already in the caller's numbering; the internal-order
    // array is rebuilt from it only if a device check asks for it.
    g->internal_stale = g->relabelled;
    if (stats) MM_HIP_TRY(hipEventRecord(g->ev_t2, g->stream), "hipEventRecord");
    // Totals of every executed round (a maximal run's last, empty round has none).
    const int r_stats = g->n > 0 ? std::min(rounds, R_run) : 0;
    if (stats && r_stats > 0) {
        if ((rc = enq_reduce(1, r_stats)) != MM_OK) return rc;
        MM_HIP_TRY(hipMemcpyAsync(g->h_ctr, d.ctr, sizeof(int32_t) * ROUND_SLOTS * CTR_STRIDE,
                                  hipMemcpyDeviceToHost, g->stream),
                   "Unable to read round counters");
        MM_HIP_TRY(hipMemcpyAsync(g->h_ctr64, g->d_ctr64, sizeof(unsigned long long) * ROUND_SLOTS * CTR64_STRIDE,
                                  hipMemcpyDeviceToHost, g->stream),
                   "Unable to read round counters");
    }
    if (graph_rounds && g->d_vround) {
        graph_rounds->assign((size_t)g->n_graphs, 0);
        if (g->n_graphs > 0) {
            const int gb = (int)std::min<int64_t>(((int64_t)g->n_graphs + 255) / 256, 1024);
            const int fb = (int)std::min<int64_t>(((int64_t)g->n / 64 + 255) / 256 + 1, (int64_t)g->num_cu * 8);
            int32_t *gmax = g->d_grounds + g->n_graphs;
            hipLaunchKernelGGL(k_graph_clear, dim3(gb), dim3(256), 0, g->stream, gmax, g->n_graphs);
            hipLaunchKernelGGL(k_graph_fold, dim3(fb), dim3(256), 0, g->stream, (const uint8_t *)g->d_vround,
                               (const int32_t *)g->d_voff, g->n_graphs, gmax);
            hipLaunchKernelGGL(k_graph_rounds, dim3(gb), dim3(256), 0, g->stream, (const int32_t *)gmax,
                               (const int32_t *)g->d_voff, g->n_graphs, R_run, R_max, maximal, g->d_grounds);
            MM_HIP_TRY(hipMemcpyAsync(graph_rounds->data(), g->d_grounds, sizeof(int32_t) * (size_t)g->n_graphs,
                                      hipMemcpyDeviceToHost, g->stream),
                       "Unable to read batch rounds");
        }
    }
    MM_HIP_TRY(wait_stream(g->stream), "A HIP error occurred during the matching process");
    MM_HIP_TRY(hipGetLastError(), "A HIP error occurred during the matching process");
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->rounds = rounds;
        stats->respond_used = c.PUSH ? MM_RESPOND_PUSH : MM_RESPOND_PULL;
        stats->lanes_used = c.G;
        float ms = 0.0f;
        hipEventElapsedTime(&ms, g->ev_t1, g->ev_t2);
        stats->match_ms = ms;
        int64_t cum = 0;
        for (int k = 1; k <= r_stats && k <= MM_MAX_ROUNDS; ++k) {
            const int32_t *q = g->h_ctr + (size_t)k * CTR_STRIDE;
            const unsigned long long *q64 = g->h_ctr64 + (size_t)k * CTR64_STRIDE;
            const int64_t active = (int64_t)q[0] + q[1] + q[2] + q[3];
            if (active > 0) cum += q[4];
            stats->active[k - 1] = active;
            stats->blue[k - 1] = (int64_t)q[0] + q[1];
            stats->matched[k - 1] = cum;
            stats->active_nnz[k - 1] = active > 0 ? (int64_t)q64[0] : 0;
            stats->blue_nnz[k - 1] = active > 0 ? (int64_t)q64[1] : 0;
        }
        for (int k = r_stats + 1; k <= rounds && k <= MM_MAX_ROUNDS; ++k) stats->matched[k - 1] = cum;
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

namespace {

// Coordinate search over lanes per row, block size and hub degree (see
// mm_autotune in mm_match.h).  Every candidate runs once untimed (it captures
// its launch sequence), then reps times; the median match_ms counts.
int32_t autotune(mm_graph *g, const mm_opts *base, const mm_tune_opts *t, mm_tune_result *out)
{
    if (!g || !base || !out) return mm::set_error(MM_ERR_INVALID, "mm_autotune: NULL argument");
    const int reps = t && t->reps > 0 ? t->reps : 3;
    const bool verify = t && t->verify;
    std::memset(out, 0, sizeof(*out));
    std::vector<int32_t> ref, cur;
    if (verify) {
        ref.resize((size_t)g->n);
        cur.resize((size_t)g->n);
    }
    struct Cand {
        int32_t block, lanes, hub, unroll;
    };
    Cand best{base->block_size, base->lanes_per_vertex, base->hub_degree, base->entries_per_thread};
    float best_ms = 0.0f;
    bool have = false;
    auto time_cand = [&](const Cand &c, float *ms_out) -> int32_t {
        mm_opts o = *base;
        o.block_size = c.block;
        o.lanes_per_vertex = c.lanes;
        o.hub_degree = c.hub;
        o.entries_per_thread = c.unroll;
        o.profile = 0;
        mm_stats st;
        std::vector<float> ms;
        for (int k = 0; k <= reps; ++k) {
            const int32_t rc = run_matching(g, &o, &st, nullptr);
            if (rc != MM_OK) return rc;
            if (k > 0) ms.push_back(st.match_ms);
        }
        std::sort(ms.begin(), ms.end());
        *ms_out = ms[ms.size() / 2];
        ++out->candidates;
        if (verify && g->n > 0) {
            const int32_t rc = mm_graph_copy_match(g, have ? cur.data() : ref.data());
            if (rc != MM_OK) return rc;
            if (have && cur != ref) ++out->mismatches;
        }
        return MM_OK;
    };
    auto consider = [&](const Cand &c) -> int32_t {
        float ms = 0.0f;
        const int32_t rc = time_cand(c, &ms);
        if (rc != MM_OK) return rc;
        if (!have) out->default_ms = ms;
        if (!have || ms < best_ms) {
            best = c;
            best_ms = ms;
        }
        have = true;
        return MM_OK;
    };
    int32_t rc = consider(best);
    if (rc != MM_OK) return rc;
    // lanes per row (entry-parallel tiles with 4, 8, 16 entries per thread)
    {
        const Cand b0 = best;
        for (int lanes : {1, 2, 4, 8, 16, 32, 64}) {
            Cand c = b0;
            c.lanes = lanes;
            c.unroll = 0;
            if ((rc = consider(c)) != MM_OK) return rc;
        }
        for (int u : {4, 8, 16}) {
            Cand c = b0;
            c.lanes = MM_LANES_TILES;
            c.unroll = u;
            if ((rc = consider(c)) != MM_OK) return rc;
        }
    }
    {
        const Cand b0 = best;
        for (int block : {64, 128, 256, 512, 1024}) {
            Cand c = b0;
            c.block = block;
            if ((rc = consider(c)) != MM_OK) return rc;
        }
    }
    if (g->max_degree > 32) {
        const Cand b0 = best;
        for (int hub : {32, 64, 128, 256, 512, 1024, 4096}) {
            if (hub >= g->max_degree && hub != 256) continue;
            Cand c = b0;
            c.hub = hub;
            if ((rc = consider(c)) != MM_OK) return rc;
        }
    }
    out->block_size = best.block > 0 ? best.block : 256;
    out->lanes_per_vertex = best.lanes != 0 ? best.lanes : 4;
    out->hub_degree = best.hub > 0 ? best.hub : 256;
    out->entries_per_thread = best.lanes == MM_LANES_TILES ? (best.unroll > 0 ? best.unroll : 4) : 0;
    out->best_ms = best_ms;
    return MM_OK;
}

}  // namespace

int32_t mm_graph_alloc(mm_graph *g, void **p, size_t bytes, const char *what)
{
    char *q = nullptr;
    const int32_t rc = dev_alloc(g, &q, bytes, what);
    *p = q;
    return rc;
}

// Rounds a maximal run is expected to need after a round with `now` active
// vertices, when the round before had `before`: the active count shrinks by a
// roughly constant factor per round, so log(now) / log(before / now) rounds
// reach zero.  Clamped to [2, 32]; a short guess costs one more read-back, a
// long one a few empty rounds (a few microseconds each).
int predicted_rounds(int64_t before, int64_t now)
{
    if (now <= 0) return 2;
    if (before <= now) return 8;
    const double shrink = std::log((double)before / (double)now);
    const double need = std::log((double)now + 1.0) / std::max(shrink, 1e-3) + 2.0;
    return (int)std::min(32.0, std::max(2.0, std::ceil(need)));
}

void mm_graph_enqueue_unpermute(mm_graph *g)
{
    if (g->relabelled && g->n > 0) {
        UnpermOff off;
        std::memcpy(off.off, g->unperm_off, sizeof(off.off));
        const int per_xcd = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)g->n / 8 + 255) / 256,
                                                                          (int64_t)std::max(1, g->num_cu / 8) * 4));
        hipLaunchKernelGGL(k_unpermute, dim3(8 * per_xcd), dim3(256), 0, g->stream, (const int2 *)g->d_unperm, off,
                           (const int32_t *)g->d_match, g->d_match_out);
    }
}

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
    return create_graph(device, n, nnz, ranges, neighbours, weights, nullptr, nullptr, 0, 1, true, true, out);
}

int32_t mm_graph_create_ex(int32_t device, int32_t n, int64_t nnz, const int32_t *ranges, const int32_t *neighbours,
                           const float *weights, uint32_t flags, mm_graph **out)
{
    if (flags & ~(uint32_t)(MM_CREATE_KEEP_ORDER | MM_CREATE_PACKED_ROWS | MM_CREATE_NO_HEADS))
        return mm::set_error(MM_ERR_INVALID, "mm_graph_create_ex: bad flags");
    return create_graph(device, n, nnz, ranges, neighbours, weights, nullptr, nullptr, 0, (flags & MM_CREATE_KEEP_ORDER) ? 0 : 1,
                        (flags & MM_CREATE_PACKED_ROWS) == 0, (flags & MM_CREATE_NO_HEADS) == 0, out);
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
    info->key_bits = !info->push_ok ? 0 : (g->d_adjp ? 32 : 64);
    info->iso_settled = g->d_iso ? 1 : 0;
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
    MM_HIP_TRY(hipMemcpyAsync(match_out, g->relabelled ? g->d_match_out : g->d_match, sizeof(int32_t) * (size_t)g->n,
                              hipMemcpyDeviceToHost, g->stream),
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
    // The graphs' rows and ids stay local to each graph; the device rebases
    // them into one CSR (mm_upload.hip).
    mm_graph *g = nullptr;
    int32_t rc = create_graph(device, (int32_t)total_n, total_nnz, ranges, neighbours, weights, vertex_offsets,
                              edge_offsets, n_graphs, 1, true, true, &g);
    if (rc != MM_OK) {
        if (rc == MM_ERR_INVALID && std::strstr(mm_last_error(), "malformed CSR"))
            return mm::set_error(MM_ERR_INVALID, "mm_batch_create: malformed per-graph CSR");
        return rc;
    }
    g->n_graphs = n_graphs;
    if ((rc = dev_alloc(g, &g->d_voff, (size_t)n_graphs + 1, "batch offsets")) != MM_OK ||
        (rc = dev_alloc(g, &g->d_vround, (size_t)total_n, "batch rounds")) != MM_OK ||
        (rc = dev_alloc(g, &g->d_grounds, 2 * (size_t)n_graphs, "batch rounds")) != MM_OK) {
        release(g);
        return rc;
    }
    if (hipMemcpy(g->d_voff, vertex_offsets, sizeof(int32_t) * ((size_t)n_graphs + 1), hipMemcpyHostToDevice) !=
        hipSuccess) {
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

int32_t mm_batch_get_info(const mm_batch *b, mm_graph_info *info)
{
    if (!b || !b->g) return mm::set_error(MM_ERR_INVALID, "mm_batch_get_info: NULL argument");
    return mm_graph_get_info(b->g, info);
}

int32_t mm_batch_destroy(mm_batch *b)
{
    if (!b) return MM_OK;
    release(b->g);
    delete b;
    return MM_OK;
}

int32_t mm_autotune(mm_graph *g, const mm_opts *opts, const mm_tune_opts *t, mm_tune_result *out)
{
    mm::clear_error();
    return autotune(g, opts, t, out);
}

int32_t mm_batch_autotune(mm_batch *b, const mm_opts *opts, const mm_tune_opts *t, mm_tune_result *out)
{
    mm::clear_error();
    if (!b || !b->g) return mm::set_error(MM_ERR_INVALID, "mm_batch_autotune: batch is NULL");
    return autotune(b->g, opts, t, out);
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
