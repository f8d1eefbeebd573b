/*
 * mm_match.h — C ABI of the MI355X (gfx950) fine-grained greedy matcher.
 *
 * Plain pointers and sizes only; no C++ or torch types cross this boundary.
 * Every entry point returns MM_OK (0) or a positive MM_ERR_* code and leaves a
 * message for mm_last_error().  The C++ classes in
 * multicore-matching_amd/csrc/matchgpu.h (GraphMatchingGPU*) sit on top of it.
 *
 * Reference interfaces replaced (BasFaggingerAuer/Multicore-Matching):
 *   mm_graph_create   <- GraphMatchingGPU ctor, device upload of the Graph CSR
 *                        (src/matchgpu.cu:39-66, weights 97-141; src/matchgpu.h:29-42)
 *   mm_graph_destroy  <- GraphMatchingGPU* dtors (src/matchgpu.cu:68-73, 117-121, 143-147)
 *   mm_match          <- GraphMatchingGPU{Random,RandomMaximal,Weighted,WeightedMaximal}
 *                        ::performMatching (src/matchgpu.cu:453-851) and its kernels
 *                        gSelect/gaSelect/grRequest/grRespond/gwRequest/gwRespond/gMatch
 *                        (src/matchgpu.cu:184-451)
 *   mm_batch_*        <- no counterpart: many independent graphs per launch
 *   mm_gen_*          <- no counterpart: seeded synthetic workload generators
 */
#ifndef MM_MATCH_H
#define MM_MATCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MM_OK 0
#define MM_ERR_INVALID 1   /* bad argument or malformed CSR */
#define MM_ERR_NOMEM 2     /* device or host allocation failed */
#define MM_ERR_HIP 3       /* HIP runtime / kernel error */
#define MM_ERR_SEEDS 4     /* fewer seeds than rounds to run */

/* Matching kinds: src/matchgpu.cu:286-357 (random), 360-451 (weighted). */
#define MM_RANDOM 0
#define MM_WEIGHTED 1

/* Round counts of the reference loops (src/matchcpu.h:24-25). */
#define MM_NR_MATCH_ROUNDS 20       /* fixed variants, types 6 and 8 */
#define MM_NR_MAX_MATCH_ROUNDS 255  /* maximal variants: while (keep && ++count < 256) */
#define MM_MAX_ROUNDS 255
/* Default select barrier, p = 0.534059 (bin/match.cpp:182-183). */
#define MM_DEFAULT_BARRIER 0x88B81733u

/* Respond strategy.  PULL scans every red vertex's row exactly as the
 * reference does and is exact for any CSR.  PUSH lets each blue vertex post
 * its proposal with one 64-bit atomic max; it returns the same match array
 * when the CSR is structurally symmetric, every row is strictly ascending and
 * w(u,v) == w(v,u) — mm_graph_create checks this and AUTO picks PUSH only
 * then. */
#define MM_LANES_TILES (-1)

#define MM_RESPOND_AUTO 0
#define MM_RESPOND_PULL 1
#define MM_RESPOND_PUSH 2

typedef struct mm_graph mm_graph;

typedef struct mm_opts {
    int32_t kind;             /* MM_RANDOM or MM_WEIGHTED */
    int32_t maximal;          /* 0: exactly max_rounds rounds; 1: stop after the first round whose
                                 select saw no active vertex, at most max_rounds rounds */
    int32_t max_rounds;       /* 1..MM_MAX_ROUNDS; 0 = 20 (fixed) or 255 (maximal) */
    uint32_t barrier;         /* blue iff hash < barrier */
    const uint32_t *seeds;    /* host array, one seed per round, >= rounds that can run */
    int32_t n_seeds;
    int32_t block_size;       /* threads per block, multiple of 64 in [64, 1024]; 0 = 256 */
    int32_t lanes_per_vertex; /* 1,2,4,...,64 lanes scan one row and stop at its first candidate
                                 (0 = 4); MM_LANES_TILES: entry-parallel tiles (rows staged in LDS,
                                 one row entry per thread, full rows) */
    int32_t hub_degree;       /* rows longer than this are scanned by a whole block; 0 = 256 */
    int32_t entries_per_thread; /* tiles: row entries each thread keeps in flight (4, 8, 16); 0 = 4 */
    int32_t respond;          /* MM_RESPOND_* */
    int32_t profile;          /* 1: time every kernel with HIP events (slower; fills kernel_ms and *_nnz) */
    int32_t graph_mode;       /* 0: capture each launch sequence once as a hipGraph and replay it;
                                 1: plain stream launches (profile == 1 implies 1) */
} mm_opts;

#define MM_KERNEL_CLASSES 5   /* 0 init-select, 1 request, 2 respond, 3 match-select, 4 counters */

typedef struct mm_stats {
    int32_t rounds;                    /* rounds executed = seeds consumed */
    int32_t respond_used;              /* MM_RESPOND_PULL or MM_RESPOND_PUSH */
    int32_t lanes_used;                /* lanes per row that ran (0: entry-parallel tiles) */
    int32_t reserved;
    float match_ms;                    /* device time from before the first round to after the last
                                          (the reference's t1 -> t2, src/matchgpu.cu:500-527) */
    float kernel_ms[MM_KERNEL_CLASSES];/* profile == 1: summed event time per kernel class */
    int32_t kernel_launches[MM_KERNEL_CLASSES];
    int64_t active[MM_MAX_ROUNDS + 1]; /* [r] vertices with match < 2 at the select of round r+1 */
    int64_t blue[MM_MAX_ROUNDS + 1];   /* [r] of those, coloured blue */
    int64_t matched[MM_MAX_ROUNDS + 1];/* [r] matched vertices after round r+1 (cumulative) */
    int64_t active_nnz[MM_MAX_ROUNDS + 1]; /* [r] sum of degrees of the active vertices (profile == 1) */
    int64_t blue_nnz[MM_MAX_ROUNDS + 1];   /* [r] sum of degrees of the blue vertices (profile == 1) */
} mm_stats;

typedef struct mm_graph_info {
    int32_t n;
    int32_t weighted;
    int64_t nnz;
    int32_t max_degree;
    int32_t push_ok;          /* symmetric, strictly ascending rows, symmetric weights */
    int64_t device_bytes;     /* device memory held by the handle */
    int32_t key_bits;         /* 32: PUSH proposals of the lane scans are 32-bit keys (position of the
                                 edge in the red vertex's sorted row, then local id: needs
                                 max_degree < 2^(32 - bits of the largest local id));
                                 64: (weight, local id) keys; 0: no PUSH */
    int32_t iso_settled;      /* 1: at least 1% of the vertices have no neighbour, and the first round's
                                 select settles those (their colours in every round follow from the
                                 seeds alone); the work lists then carry only the others */
} mm_graph_info;

/* Last error message of the calling thread ("" when none). */
const char *mm_last_error(void);
/* Number of visible HIP devices (0 when there is no GPU or no runtime). */
int32_t mm_device_count(void);

/* Upload a CSR graph in the reference layout (src/graph.h:41-60):
 *   ranges      int32[2n]: ranges[2i] = begin, ranges[2i+1] = end of row i (the Graph's int2)
 *   neighbours  int32[nnz]
 *   weights     float[nnz] or NULL (unit weights; MM_WEIGHTED then sees all ties)
 * The device copy is owned by the handle.  device = HIP device ordinal. */
int32_t mm_graph_create(int32_t device, int32_t n, int64_t nnz, const int32_t *ranges,
                        const int32_t *neighbours, const float *weights, mm_graph **out);
/* As mm_graph_create, with flags.  By default the handle renumbers the vertices
 * internally in a locality order (compact regions of up to 2048 vertices,
 * each grown breadth-first), so that neighbours sit close together in memory; results, hash inputs and tie-breaks still use the caller's ids.
 * MM_CREATE_KEEP_ORDER keeps the caller's numbering on the device. */
#define MM_CREATE_KEEP_ORDER 1u
/* MM_CREATE_PACKED_ROWS stores the device rows back to back.  By default a row
 * whose first 8 entries would straddle a 128-byte line starts at the next line
 * (about 1.75 entries of padding per row on average), so scans of scattered
 * rows in the later rounds fetch one line per row step.  Results are the same
 * either way. */
#define MM_CREATE_PACKED_ROWS 2u
/* MM_CREATE_NO_HEADS skips the head copy of a weighted graph.  By default the
 * handle also keeps the first 4 entries of every weight-sorted row in a dense
 * array (32 bytes per vertex): the first round's scans, which visit about half
 * of all rows and usually stop within 4 entries, read that array (one lane per
 * row, all coalesced) instead of the rows themselves.  Results are the same
 * either way. */
#define MM_CREATE_NO_HEADS 4u
int32_t mm_graph_create_ex(int32_t device, int32_t n, int64_t nnz, const int32_t *ranges,
                           const int32_t *neighbours, const float *weights, uint32_t flags, mm_graph **out);
int32_t mm_graph_destroy(mm_graph *g);
int32_t mm_graph_get_info(const mm_graph *g, mm_graph_info *info);

/* Run one matching.  match_out (host, n entries) receives the final state array
 * in the reference encoding (0 blue, 1 red, 2 dead, >=4 matched: 4 + min(i, partner));
 * pass NULL to leave the result on the device (mm_graph_copy_match fetches it).
 * stats may be NULL. */
int32_t mm_match(mm_graph *g, const mm_opts *opts, int32_t *match_out, mm_stats *stats);
int32_t mm_graph_copy_match(mm_graph *g, int32_t *match_out);

/* Batched matching: many independent graphs concatenated into one CSR.
 *   vertex_offsets int32[n_graphs+1], edge_offsets int64[n_graphs+1]
 *   ranges int32[2 * total_n] with positions LOCAL to each graph's edge block,
 *   neighbours int32[total_nnz] with ids LOCAL to each graph, weights or NULL.
 * One launch sequence matches every graph; graph g's result equals mm_match on
 * graph g alone with the same seeds (hash input, sentinels and match values
 * use graph-local ids). */
typedef struct mm_batch mm_batch;
int32_t mm_batch_create(int32_t device, int32_t n_graphs, const int32_t *vertex_offsets,
                        const int64_t *edge_offsets, const int32_t *ranges,
                        const int32_t *neighbours, const float *weights, mm_batch **out);
int32_t mm_batch_destroy(mm_batch *b);
/* Properties of the batch's concatenated device graph (as mm_graph_get_info). */
int32_t mm_batch_get_info(const mm_batch *b, mm_graph_info *info);
/* rounds_out int32[n_graphs] (may be NULL): rounds graph g executed alone. */
int32_t mm_batch_match(mm_batch *b, const mm_opts *opts, int32_t *match_out,
                       int32_t *rounds_out, mm_stats *stats);

/* Autotuner (no reference counterpart; the reference fixes 256 threads per
 * block, bin/match.cpp:186).  Searches, one parameter at a time, the lanes per
 * row (a row is scanned by 1..64 lanes, or by entry-parallel tiles with 4, 8
 * or 16 entries per thread), the block size (64..1024) and the hub degree
 * (rows longer than it are scanned by a whole block), timing each candidate
 * with device events on this graph and these seeds, and returns the fastest.
 * Tiles never change results: with verify = 1 every candidate's match array
 * is compared with the first one's and differences are counted. */
typedef struct mm_tune_opts {
    int32_t reps;     /* timed runs per candidate, the median is kept; 0 = 3 */
    int32_t verify;   /* 1: compare every candidate's match array with the first */
} mm_tune_opts;
typedef struct mm_tune_result {
    int32_t block_size, lanes_per_vertex, hub_degree, entries_per_thread;  /* the fastest candidate */
    float best_ms;        /* its median match_ms */
    float default_ms;     /* median match_ms of the first candidate (the opts as given) */
    int32_t candidates;   /* candidates timed */
    int32_t mismatches;   /* candidates whose match array differed (verify = 1) */
} mm_tune_result;
/* opts: the run to tune (kind, maximal, max_rounds, seeds, barrier, respond);
 * its tile fields are the search's starting point (0 = defaults). */
int32_t mm_autotune(mm_graph *g, const mm_opts *opts, const mm_tune_opts *t, mm_tune_result *out);
int32_t mm_batch_autotune(mm_batch *b, const mm_opts *opts, const mm_tune_opts *t, mm_tune_result *out);

/* Result metrics on the device (src/matchcpu.cu:64-97 getWeight, 127-246
 * testMatching) of the handle's current result, in graph-local values for a
 * batch: size = matched vertices, weight summed in double over both endpoints
 * and halved, and the testMatching counts.  Groups of equal match values are
 * formed per graph, for any value >= 4 (values outside the graph's id range
 * included, as testMatching's sort groups them); a group's "first two"
 * members, the pair whose adjacency is checked, are its two smallest caller
 * ids (the reference leaves that order to an unstable sort).  The weight of a
 * matched vertex is that of its first CSR neighbour with the same value,
 * exact for CSRs without repeated neighbours. */
typedef struct mm_check {
    int64_t under;      /* matched to themselves */
    int64_t multi;      /* matched to more than one vertex */
    int64_t invalid;    /* not matched to a neighbour */
    int64_t matchable;  /* unmatched vertices with an unmatched neighbour */
    int64_t size;       /* matched vertices (getWeight's size) */
    double weight;      /* getWeight's weight */
    int32_t valid;      /* under == multi == invalid == 0 */
} mm_check;
int32_t mm_graph_check(mm_graph *g, mm_check *out);
/* mm_graph_check without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture
 * aside -- the call synchronises its stream -- tight serving loops, harnesses
 * that forbid hipMalloc).  Same results as mm_graph_check, for graphs of up to
 * MM_CHECK_NOALLOC_MAX_N vertices (larger ones: MM_ERR_INVALID).  Its scratch
 * is one 128 MiB device array per GPU, held until the process ends and shared
 * by every call on that GPU: calls that overlap on one GPU race. */
#define MM_CHECK_NOALLOC_MAX_N 4194304
int32_t mm_graph_check_noalloc(mm_graph *g, mm_check *out);
/* Replaces the handle's current result with match (host, n entries, caller's
 * order) -- to check an array computed elsewhere. */
int32_t mm_graph_set_match(mm_graph *g, const int32_t *match);
/* Batch: replaces the batch's current result with match (host, total_n
 * entries: the graphs' arrays concatenated, graph-local values). */
int32_t mm_batch_set_match(mm_batch *b, const int32_t *match);
/* Batch: totals, and per graph (arrays of n_graphs, each may be NULL) the
 * weight, the matched vertices and whether the graph's result is valid. */
int32_t mm_batch_check(mm_batch *b, mm_check *total, double *graph_weight, int64_t *graph_size,
                       int32_t *graph_valid);

/* Sequential host matchers, types 0-5 of the reference's getMatcher
 * (bin/match.cpp:136-143; src/matchcpu.cu:259-810): 0 random greedy,
 * 1 static minimum degree, 2 Karp-Sipser, 3 dynamic minimum degree,
 * 4 weighted vertex-order greedy, 5 edge-sorted greedy (edges from the rows,
 * i < j, row order).  Host only; match_out gets n entries in the reference
 * encoding.  weights may be NULL (unit weights). */
int32_t mm_host_match(int32_t type, int32_t n, int64_t nnz, const int32_t *ranges, const int32_t *neighbours,
                      const float *weights, int32_t *match_out);

/* Result metrics of a match array (host; src/matchcpu.cu:64-97 getWeight and
 * 127-246 testMatching).  mm_get_weight: size = matched VERTICES, weight summed
 * in double over both endpoints and halved.  mm_test_matching fills
 * counts[4] = {matched to themselves, matched to more than one vertex, not
 * matched to a neighbour, still matchable} and returns MM_OK when the first
 * three are zero (MM_ERR_INVALID otherwise). */
int32_t mm_get_weight(int32_t n, const int32_t *ranges, const int32_t *neighbours, const float *weights,
                      const int32_t *match, double *weight_out, int64_t *size_out);
int32_t mm_test_matching(int32_t n, const int32_t *ranges, const int32_t *neighbours, const int32_t *match,
                         int64_t *counts);

/* Seeded synthetic CSR generators (host).  Arrays are malloc'd by the library
 * and released with mm_csr_free.  Rows are strictly ascending; weights are
 * w = (k + 1) / 2^24 with k a hash of (seed, min(u,v), max(u,v)), mirrored. */
typedef struct mm_csr {
    int32_t n;
    int64_t nnz;
    int32_t *ranges;      /* int32[2n] */
    int32_t *neighbours;  /* int32[nnz] */
    float *weights;       /* float[nnz] or NULL */
} mm_csr;
/* Random geometric graph: n points uniform in [0,1)^2, edge iff distance < r,
 * r = radius_scale * sqrt(ln n / n) (radius_scale 0.55 for the standard workload). */
int32_t mm_gen_rgg(int32_t n, double radius_scale, uint64_t seed, int32_t weighted, mm_csr *out);
/* R-MAT: 2^scale vertices, edge_factor * 2^scale samples, quadrant
 * probabilities a, b, c (d = 1 - a - b - c); symmetrised, deduplicated,
 * self-loops removed. */
int32_t mm_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                    uint64_t seed, int32_t weighted, mm_csr *out);
/* rows x cols 2-D grid, 4-neighbour stencil, vertex id = r * cols + c. */
int32_t mm_gen_grid(int32_t rows, int32_t cols, uint64_t seed, int32_t weighted, mm_csr *out);
void mm_csr_free(mm_csr *csr);

#ifdef __cplusplus
}
#endif

#endif /* MM_MATCH_H */
