/*
 * CPU oracle for fine-grained greedy matching.  TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the round semantics of the reference's GPU
 * matcher and its CPU twin so that the HIP kernels can be checked against it
 * bit for bit.  It is the checker, never the thing measured or shipped: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Reference anchors (BasFaggingerAuer/Multicore-Matching):
 *   select hash ............ src/matchgpu.cu:166-218, twin src/matchtbb.cu:121-156
 *   match write-back ....... src/matchgpu.cu:260-283, twin src/matchtbb.cu:171-194
 *   random request/respond . src/matchgpu.cu:286-357, twin src/matchtbb.cu:244-331
 *   weighted req/respond ... src/matchgpu.cu:360-451, twin src/matchtbb.cu:384-485
 *   round loops ............ src/matchgpu.cu:507-524 (fixed), 604-615 (maximal),
 *                            src/matchtbb.cu:525-550, 594-619
 *   state encoding ......... src/matchcpu.h:24-25, 40-41; src/matchgpu.cu:157-164
 *   getWeight .............. src/matchcpu.cu:64-97
 *   testMatching ........... src/matchcpu.cu:127-246
 *   CPU greedy types 0/4/5 . src/matchcpu.cu:259-303, 695-745, 777-810 (type 5
 *                            lives in mm_oracle_sort.cpp, it needs std::sort)
 *   CPU types 2/3 .......... src/matchcpu.cu:518-682 (Karp-Sipser), 304-426
 *                            (dynamic minimum degree); type 1 (static minimum
 *                            degree, 428-516) is in mm_oracle_sort.cpp
 *
 * Every phase of a round writes only the slot of the vertex it processes and
 * reads only slots the phase does not write, so the OpenMP loops below give
 * the same result as the serial ones for any thread count.
 *
 * Parity status of each function (how it is pinned, see tests/test_oracle.py):
 *   orc_hash / orc_colour: constants re-derived from mathematics (SHA-256 K[j]
 *     from cube roots of primes, MD5 IV and shifts), a second independent
 *     restatement in numpy over ~1e6 (i, seed) pairs, special cases (vertex 0
 *     is seed-independent, seed 1 keeps g = i).  The reference itself cannot be
 *     executed here, so the absolute bit output of the hash is pinned only by
 *     these checks.
 *   orc_match: brute force over all colourings of tiny graphs against a
 *     literal pure-Python statement of the phases, closed forms (single edge,
 *     stars, ties), the unit-weight identity weighted == random, validity and
 *     maximality invariants, serial == OpenMP.
 *   orc_get_weight / orc_test_matching: hand-computed known answers.
 *   orc_cpu_random / orc_cpu_weighted: hand-computed known answers and the
 *     textbook 1/2-approximation bound against brute-force maximum matchings.
 *   orc_cpu_karp_sipser: maximum matchings on forests (Karp and Sipser's
 *     degree-1 rule is optimal there; checked against a tree DP), hand-computed
 *     known answers, validity and maximality.
 *   orc_cpu_min_deg: closed forms (paths and cycles floor(n/2), stars 1,
 *     complete bipartite min(a, b)), a hand-computed known answer that
 *     exercises the recency order among equal degrees, validity, maximality.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* State encoding of match[] (src/matchgpu.cu:157-164). */
enum { ORC_BLUE = 0, ORC_RED = 1, ORC_DEAD = 2 };
#define ORC_IS_MATCHED(m) ((m) >= 4)
#define ORC_MATCH_VAL(i, j) (4 + ((i) < (j) ? (i) : (j)))

/* First 16 SHA-256 round constants, as read by the select hash
 * (src/matchgpu.cu:167-169; only K[0..15] are used). */
static const uint32_t ORC_K[16] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u,
    0x923f82a4u, 0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u,
    0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u};
/* MD5 round-1 rotations (src/matchgpu.cu:177). */
static const uint32_t ORC_R[4] = {7u, 12u, 17u, 22u};

static inline uint32_t orc_rotl(uint32_t x, uint32_t s) { return (x << s) | (x >> (32u - s)); }

/* Sum h0+h1+h2+h3 of the 16-step MD5-shaped hash of (i, seed)
 * (src/matchgpu.cu:196-217). */
uint32_t orc_hash(uint32_t i, uint32_t seed)
{
    uint32_t h0 = 0x67452301u, h1 = 0xefcdab89u, h2 = 0x98badcfeu, h3 = 0x10325476u;
    uint32_t a = h0, b = h1, c = h2, d = h3, g = i;
    for (int j = 0; j < 16; ++j) {
        const uint32_t f = (b & c) | (~b & d);
        const uint32_t na = d, nc = b, nd = c;
        const uint32_t nb = b + orc_rotl(a + f + ORC_K[j] + g, ORC_R[j & 3]);
        a = na; b = nb; c = nc; d = nd;
        h0 += a; h1 += b; h2 += c; h3 += d;
        g *= seed;
    }
    return h0 + h1 + h2 + h3;
}

/* 0 = blue, 1 = red (src/matchgpu.cu:217). */
int32_t orc_colour(uint32_t i, uint32_t seed, uint32_t barrier)
{
    return orc_hash(i, seed) < barrier ? ORC_BLUE : ORC_RED;
}

void orc_hash_many(int64_t count, const uint32_t *ids, const uint32_t *seeds, uint32_t *out)
{
    for (int64_t k = 0; k < count; ++k) out[k] = orc_hash(ids[k], seeds[k]);
}

static int orc_threads(int nthreads)
{
#ifdef _OPENMP
    return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

/*
 * One full matching run: Appendix-A semantics of the reference rounds.
 *   weighted: 0 random (grRequest/grRespond), 1 weighted (gwRequest/gwRespond)
 *   maximal:  0 -> exactly max_rounds rounds (types 6/8, NR_MATCH_ROUNDS);
 *             1 -> while (keep && ++count < max_rounds + 1) (types 7/9/10/11)
 *   seeds:    one per executed round, in order (the reference draws rand()).
 * Outputs: match[n]; round_active[r] = vertices with match<2 seen by the
 * select of round r+1; round_matched[r] = matched vertices after round r+1.
 * Returns the number of rounds executed (= seeds consumed), or -1 when
 * n_seeds is too small.
 */
int32_t orc_match(int32_t n, const int32_t *ranges, const int32_t *nbrs, const float *w,
                  int32_t weighted, int32_t maximal, int32_t max_rounds, uint32_t barrier,
                  const uint32_t *seeds, int32_t n_seeds, int32_t *match,
                  int64_t *round_active, int64_t *round_matched, int32_t nthreads)
{
    const int nt = orc_threads(nthreads);
    int32_t *req = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!req) return -1;
    for (int32_t i = 0; i < n; ++i) { match[i] = 0; req[i] = 0; }

    int32_t rounds = 0;
    int keep = 1;
    /* src/matchgpu.cu:604 is `while (keep == 1 && ++count < NR_MAX_MATCH_ROUNDS)`
     * with NR_MAX_MATCH_ROUNDS = 256, i.e. at most 255 rounds; the fixed loop
     * (src/matchgpu.cu:507) runs exactly NR_MATCH_ROUNDS.  Both are expressed
     * here by max_rounds. */
    while (rounds < max_rounds && (!maximal || keep)) {
        if (rounds >= n_seeds) { free(req); return -1; }
        const uint32_t seed = seeds[rounds];
        int64_t active = 0;

        /* select: src/matchgpu.cu:184-258 */
#pragma omp parallel for num_threads(nt) reduction(+ : active) schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            if (match[i] >= 2) continue;
            ++active;
            match[i] = orc_colour((uint32_t)i, seed, barrier);
        }
        keep = active > 0;

        /* request: src/matchgpu.cu:286-327 (random), 360-409 (weighted) */
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1024)
        for (int32_t i = 0; i < n; ++i) {
            if (match[i] != ORC_BLUE) { req[i] = n; continue; }
            int32_t dead = 1, cand = n;
            float best = -1.0f;
            for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1]; ++j) {
                const int32_t u = nbrs[j];
                const int32_t mu = match[u];
                if (mu >= 4) continue;
                dead = 0;
                if (mu != ORC_RED) continue;
                if (!weighted) { cand = u; break; }
                if (w[j] > best) { best = w[j]; cand = u; }
            }
            /* random: a found red neighbour ends the scan with dead == 0 */
            req[i] = cand < n ? cand : n + dead;
        }

        /* respond: src/matchgpu.cu:329-357 (random), 411-451 (weighted) */
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1024)
        for (int32_t i = 0; i < n; ++i) {
            if (match[i] != ORC_RED) continue;
            int32_t cand = n;
            float best = -1.0f;
            for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1]; ++j) {
                const int32_t u = nbrs[j];
                if (match[u] != ORC_BLUE || req[u] != i) continue;
                if (!weighted) { cand = u; break; }
                if (w[j] > best) { best = w[j]; cand = u; }
            }
            if (cand < n) req[i] = cand;
        }

        /* match: src/matchgpu.cu:260-283 */
        int64_t matched = 0;
#pragma omp parallel for num_threads(nt) reduction(+ : matched) schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            const int32_t r = req[i];
            if (r == n + 1) match[i] = ORC_DEAD;
            else if (r < n && req[r] == i) match[i] = ORC_MATCH_VAL(i, r);
            matched += ORC_IS_MATCHED(match[i]);
        }

        if (round_active) round_active[rounds] = active;
        if (round_matched) round_matched[rounds] = matched;
        ++rounds;
    }
    free(req);
    return rounds;
}

/* getWeight: src/matchcpu.cu:64-97.  size counts matched VERTICES; the weight
 * is summed in double over both endpoints and halved. */
void orc_get_weight(int32_t n, const int32_t *ranges, const int32_t *nbrs, const float *w,
                    const int32_t *match, double *weight_out, int64_t *size_out)
{
    double weight = 0.0;
    int64_t size = 0;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t m = match[i];
        if (!ORC_IS_MATCHED(m)) continue;
        for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1]; ++j) {
            if (match[nbrs[j]] == m) {
                weight += w ? (double)w[j] : 1.0;
                size += 1;
                break;
            }
        }
    }
    *weight_out = weight / 2.0;
    *size_out = size;
}

static _Thread_local const int32_t *orc_sort_key; /* per thread: callers may check in parallel */
static int orc_cmp_match(const void *pa, const void *pb)
{
    /* SortVerticesMatch (src/matchcpu.cu:99-125): unmatched first, then by
     * match value.  Ties are broken by vertex id here (the reference's
     * std::sort leaves their order open); the order within a group decides
     * only which pair of a group of three or more is tested for adjacency. */
    const int32_t a = *(const int32_t *)pa, b = *(const int32_t *)pb;
    const int32_t va = orc_sort_key[a], vb = orc_sort_key[b];
    const int ma = ORC_IS_MATCHED(va), mb = ORC_IS_MATCHED(vb);
    if (ma != mb) return ma ? 1 : -1;
    if (ma && va != vb) return va < vb ? -1 : 1;
    return a < b ? -1 : (a > b);
}

/* testMatching: src/matchcpu.cu:127-246.  counts[0..3] = under-matched,
 * multi-matched, invalidly matched, still matchable.  Returns 1 when valid. */
int32_t orc_test_matching(int32_t n, const int32_t *ranges, const int32_t *nbrs,
                          const int32_t *match, int64_t *counts)
{
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!order) return -1;
    for (int32_t i = 0; i < n; ++i) order[i] = i;
    orc_sort_key = match;
    qsort(order, (size_t)n, sizeof(int32_t), orc_cmp_match);

    int32_t k = 0;
    while (k < n && !ORC_IS_MATCHED(match[order[k]])) ++k;
    int64_t under = 0, multi = 0, invalid = 0, matchable = 0;
    while (k < n) {
        const int32_t i1 = order[k];
        const int32_t cur = match[i1];
        ++k;
        if (k == n) { ++under; continue; }
        const int32_t i2 = order[k];
        ++k;
        if (match[i2] != cur) { ++under; continue; }
        int f1 = 0, f2 = 0;
        for (int32_t j = ranges[2 * i1]; j < ranges[2 * i1 + 1] && !f1; ++j) f1 = nbrs[j] == i2;
        for (int32_t j = ranges[2 * i2]; j < ranges[2 * i2 + 1] && !f2; ++j) f2 = nbrs[j] == i1;
        if (!f1 || !f2) ++invalid;
        while (k < n && match[order[k]] == cur) { ++multi; ++k; }
    }
    for (int32_t i = 0; i < n; ++i) {
        if (ORC_IS_MATCHED(match[i])) continue;
        int32_t valid = 0;
        for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1]; ++j) {
            const int32_t u = nbrs[j];
            if (u == i) continue;
            if (!ORC_IS_MATCHED(match[u])) ++valid;
        }
        if (valid > 0) ++matchable;
    }
    free(order);
    if (counts) { counts[0] = under; counts[1] = multi; counts[2] = invalid; counts[3] = matchable; }
    return under == 0 && multi == 0 && invalid == 0;
}

/* Type 0, random greedy: src/matchcpu.cu:259-303. */
void orc_cpu_random(int32_t n, const int32_t *ranges, const int32_t *nbrs, int32_t *match)
{
    for (int32_t i = 0; i < n; ++i) match[i] = 0;
    for (int32_t i = 0; i < n; ++i) {
        if (ORC_IS_MATCHED(match[i])) continue;
        int32_t cand = n;
        for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1]; ++j) {
            const int32_t u = nbrs[j];
            if (!ORC_IS_MATCHED(match[u]) && u != i) { cand = u; break; }
        }
        if (cand < n) match[i] = match[cand] = ORC_MATCH_VAL(i, cand);
    }
}

/* Type 4, weighted vertex-order greedy: src/matchcpu.cu:695-745. */
void orc_cpu_weighted(int32_t n, const int32_t *ranges, const int32_t *nbrs, const float *w,
                      int32_t *match)
{
    for (int32_t i = 0; i < n; ++i) match[i] = 0;
    for (int32_t i = 0; i < n; ++i) {
        if (ORC_IS_MATCHED(match[i])) continue;
        float best = -1.0f;
        int32_t cand = n;
        for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1]; ++j) {
            const int32_t u = nbrs[j];
            if (!ORC_IS_MATCHED(match[u]) && u != i && w[j] > best) { best = w[j]; cand = u; }
        }
        if (cand < n) match[i] = match[cand] = ORC_MATCH_VAL(i, cand);
    }
}

/* Type 2, one-sided Karp-Sipser: src/matchcpu.cu:518-682.  deg[] starts at
 * the row lengths.  Singletons (degree exactly 1) are served first-in
 * first-out: each is matched to its first unmatched neighbour, and every row
 * entry of that partner lowers its target's degree.  With no singleton left,
 * one pass over the remaining degree > 1 vertices, in vertex order (vertices
 * never re-enter that set, so "in order" is a scan of ids with a membership
 * flag), matches each still-unmatched vertex of nonzero degree to its first
 * unmatched neighbour and lowers the targets of both rows; every vertex the
 * pass reaches leaves the set.  Whenever an unmatched vertex's degree becomes
 * exactly 1 it leaves the set and joins the singletons. */
void orc_cpu_karp_sipser(int32_t n, const int32_t *ranges, const int32_t *nbrs, int32_t *match)
{
    int32_t *deg = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * n + 1));
    char *in_set = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    int64_t q_head = 0, q_tail = 0, in_set_count = 0;
    for (int32_t i = 0; i < n; ++i) {
        match[i] = 0;
        deg[i] = ranges[2 * i + 1] - ranges[2 * i];
        if (deg[i] == 1) queue[q_tail++] = i;
        else if (deg[i] > 1) { in_set[i] = 1; ++in_set_count; }
    }
#define ORC_KS_LOWER(c)                                                              \
    for (int32_t jj = ranges[2 * (c)]; jj < ranges[2 * (c) + 1]; ++jj) {             \
        const int32_t t = nbrs[jj];                                                  \
        deg[t] -= 1;                                                                 \
        if (deg[t] == 1 && !ORC_IS_MATCHED(match[t])) {                              \
            if (in_set[t]) { in_set[t] = 0; --in_set_count; }                        \
            queue[q_tail++] = t;                                                     \
        }                                                                            \
    }
    while (q_head < q_tail || in_set_count > 0) {
        while (q_head < q_tail) {
            const int32_t i = queue[q_head++];
            if (ORC_IS_MATCHED(match[i]) || deg[i] < 1) continue;
            int32_t cand = -1;
            for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1] && cand < 0; ++j)
                if (nbrs[j] != i && !ORC_IS_MATCHED(match[nbrs[j]])) cand = nbrs[j];
            if (cand < 0) continue;
            match[i] = match[cand] = ORC_MATCH_VAL(i, cand);
            ORC_KS_LOWER(cand)
        }
        for (int32_t i = 0; i < n && in_set_count > 0; ++i) {
            if (!in_set[i]) continue;
            if (!ORC_IS_MATCHED(match[i]) && deg[i] != 0) {
                int32_t cand = -1;
                for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1] && cand < 0; ++j)
                    if (nbrs[j] != i && !ORC_IS_MATCHED(match[nbrs[j]])) cand = nbrs[j];
                if (cand >= 0) {
                    match[i] = match[cand] = ORC_MATCH_VAL(i, cand);
                    ORC_KS_LOWER(i)
                    ORC_KS_LOWER(cand)
                }
            }
            if (in_set[i]) { in_set[i] = 0; --in_set_count; }
        }
    }
#undef ORC_KS_LOWER
    free(deg);
    free(queue);
    free(in_set);
}

/* Type 3, two-sided dynamic minimum degree: src/matchcpu.cu:304-426.  The
 * reference keeps (degree, vertex) in a std::multimap, where equal degrees
 * come out in insertion order; here each queued vertex carries the time of
 * its last (re)insertion and every step scans for the smallest (degree, time)
 * -- quadratic, for small graphs.  Step: take that vertex i; its partner is
 * the unmatched, queued neighbour of least degree (first in row order on
 * ties); both leave the queue; then every row entry of i, and then of the
 * partner, that names an unmatched queued vertex other than i lowers its
 * degree by one (re-inserting it now, or dropping it at degree 0). */
void orc_cpu_min_deg(int32_t n, const int32_t *ranges, const int32_t *nbrs, int32_t *match)
{
    int32_t *deg = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int64_t *stamp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    char *queued = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    int64_t clock = 0, left = 0;
    for (int32_t i = 0; i < n; ++i) {
        match[i] = 0;
        deg[i] = ranges[2 * i + 1] - ranges[2 * i];
        if (deg[i] > 0) { queued[i] = 1; stamp[i] = clock++; ++left; }
    }
    while (left > 0) {
        int32_t i = -1;
        for (int32_t v = 0; v < n; ++v)
            if (queued[v] && (i < 0 || deg[v] < deg[i] || (deg[v] == deg[i] && stamp[v] < stamp[i]))) i = v;
        queued[i] = 0;
        --left;
        int32_t cand = -1;
        for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1]; ++j) {
            const int32_t u = nbrs[j];
            if (u == i || ORC_IS_MATCHED(match[u]) || !queued[u]) continue;
            if (cand < 0 || deg[u] < deg[cand]) cand = u;
        }
        if (cand < 0) continue;
        match[i] = match[cand] = ORC_MATCH_VAL(i, cand);
        queued[cand] = 0;
        --left;
        for (int side = 0; side < 2; ++side) {
            const int32_t c = side == 0 ? i : cand;
            for (int32_t j = ranges[2 * c]; j < ranges[2 * c + 1]; ++j) {
                const int32_t u = nbrs[j];
                if (u == i || ORC_IS_MATCHED(match[u]) || !queued[u]) continue;
                if (deg[u] > 1) { deg[u] -= 1; stamp[u] = clock++; }
                else { queued[u] = 0; --left; }
            }
        }
    }
    free(deg);
    free(stamp);
    free(queued);
}

