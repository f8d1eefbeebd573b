/*
 * mm_learn.h — learned vertex selection and its training kernels (gfx950).
 *
 * No counterpart in the reference (BasFaggingerAuer/Multicore-Matching).  The
 * reference colours every active vertex blue with one fixed probability taken
 * from a hash of (vertex, seed) (gSelect, src/matchgpu.cu:184-218).  Here a
 * small network chooses each vertex's blue probability from local features at
 * every select; request, respond and match keep the reference rules
 * (gwRequest/gwRespond/gMatch, src/matchgpu.cu:260-451; grRequest/grRespond
 * for MM_RANDOM).  With params == NULL the rounds are the plain ones and the
 * match array equals mm_match's bit for bit.
 *
 * Network: float32, 8 features -> 32 -> 32 -> 1, ReLU, MM_MLP_PARAMS values
 * in this order (row-major):
 *   [0, 256)     W1[32][8]     [256, 288)   b1[32]
 *   [288, 1312)  W2[32][32]    [1312, 1344) b2[32]
 *   [1344, 1376) W3[32]        [1376]       b3
 *   z = W3 . relu(W2 . relu(W1 . x + b1) + b2) + b3,  blue probability sigmoid(z).
 * Every dot product is a chain of fused multiply-adds in index order starting
 * from the bias, so host and device agree bit for bit.
 *
 * Features of an active vertex v (match < 2) at the select of round r
 * (1-based), over the unmatched neighbours u (match[u] < 4):
 *   x0 = min(deg v, 1024) / 64
 *   x1 = w1, the heaviest weight to an unmatched neighbour among the weights
 *        that can win a request (> -1); 0 when there is none
 *   x2 = w2, the second heaviest (w1 again when it repeats); 0 when fewer than two
 *   x3 = w2 / w1 (divided in double, rounded to float) when w1 > 0, else 0
 *   x4 = min(unmatched neighbours, 1024) / 64
 *   x5 = min(neighbours that were red after the previous round (match == 1), 1024) / 64
 *   x6 = (r - 1) / 32
 *   x7 = 1 when hash(v, seed_r) < hash(u*, seed_r) for the heaviest unmatched
 *        neighbour u* (first in CSR order on ties), else 0
 * MM_RANDOM reads every weight as 1 and u* is the first unmatched neighbour.
 * hash is the reference select hash of the vertex's graph-local id.
 *
 * Colouring (params != NULL, rounds 1..MM_LEARN_ROUNDS; later rounds use the
 * plain colouring, so a network that keeps two neighbours red forever cannot
 * stop the rounds from converging):
 *   - a vertex with no unmatched neighbour leaves the rounds at once (state 2);
 *   - a non-finite z falls back to the plain colour (hash(v, seed_r) < barrier);
 *   - otherwise T = floor(65536 * s(z)) with s(z) = 1 / (1 + 2^(-z * log2 e)),
 *     2^f evaluated by the degree-7 Taylor polynomial of exp(f ln 2) in Horner
 *     form with fused multiply-adds, and v is blue iff
 *     (hash(v, seed_r ^ 0x5BD1E995) >> 16) < T.
 */
#ifndef MM_LEARN_H
#define MM_LEARN_H

#include <stdint.h>

#include "mm_match.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MM_MLP_FEATURES 8
#define MM_MLP_HIDDEN 32
#define MM_MLP_PARAMS 1377
#define MM_MLP_OFF_W1 0
#define MM_MLP_OFF_B1 256
#define MM_MLP_OFF_W2 288
#define MM_MLP_OFF_B2 1312
#define MM_MLP_OFF_W3 1344
#define MM_MLP_OFF_B3 1376
/* Seed tweaks of the colour draw and of the trace sampling hash. */
#define MM_LEARN_DRAW_TWEAK 0x5BD1E995u
#define MM_LEARN_TRACE_TWEAK 0x2545F491u
/* Rounds whose select uses the network; later rounds are plain. */
#define MM_LEARN_ROUNDS 32

typedef struct mm_learn_opts {
    const float *params;    /* host, MM_MLP_PARAMS floats; NULL: plain hash colouring */
    int32_t trace;          /* 1: record one sample per vertex matched in a round */
    uint32_t trace_keep;    /* record a matched vertex iff (hash(v, seed_r ^ MM_LEARN_TRACE_TWEAK) >> 16)
                               < trace_keep; 65536 records all */
    int64_t trace_cap;      /* capacity of the arrays below */
    float *trace_x;         /* host [trace_cap][8]: features at the select of that round */
    float *trace_y;         /* host [trace_cap]: 1 if the vertex was blue, 0 if red */
    float *trace_w;         /* host [trace_cap]: weight of the matched edge */
    int32_t *trace_v;       /* host [trace_cap] or NULL: graph-local id */
    int32_t *trace_round;   /* host [trace_cap] or NULL: round (1-based) */
    int64_t trace_count;    /* out: samples stored, ordered by (round, id) */
    int64_t trace_offered;  /* out: samples that passed the sampling hash (may exceed trace_cap) */
} mm_learn_opts;

/* One matching with learned selection.  opts: kind, maximal, max_rounds,
 * barrier, seeds, n_seeds and block_size are used as in mm_match; stats gets
 * rounds, match_ms, active, blue and matched.  lopts may be NULL (plain). */
int32_t mm_match_learned(mm_graph *g, const mm_opts *opts, mm_learn_opts *lopts, int32_t *match_out,
                         mm_stats *stats);

/* He-uniform weights and zero biases from a SplitMix64 stream (host). */
int32_t mm_mlp_init(uint64_t seed, float *params_out);
/* z for n feature rows x[n][8] on the device (same arithmetic as the select). */
int32_t mm_mlp_forward(int32_t device, const float *params, const float *x, int64_t n, float *z_out);
/* Weighted logistic loss L = (1/n) sum_i w_i (softplus(z_i) - y_i z_i) and its
 * gradient in the parameter layout, by the device forward/backward kernels. */
int32_t mm_mlp_loss_grad(int32_t device, const float *params, const float *x, const float *y, const float *w,
                         int64_t n, double *loss_out, float *grad_out);

typedef struct mm_train_opts {
    int32_t epochs;         /* passes over the samples */
    int32_t batch;          /* samples per Adam step */
    float lr;               /* Adam step size */
    float beta1, beta2, eps;
    uint64_t seed;          /* sample order */
} mm_train_opts;
/* Minibatch Adam on the device over the samples (x[n][8], y[n], w[n]);
 * params_inout is the start point and receives the result;
 * epoch_loss_out[epochs] (may be NULL) the mean loss of each epoch.  The
 * reductions run in a fixed order, so a run is reproducible. */
int32_t mm_mlp_train(int32_t device, const float *x, const float *y, const float *w, int64_t n,
                     const mm_train_opts *t, float *params_inout, float *epoch_loss_out);

#ifdef __cplusplus
}
#endif

#endif /* MM_LEARN_H */
