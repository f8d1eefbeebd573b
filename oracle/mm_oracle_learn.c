/*
 * CPU oracle of the learned selection.  TEST INFRASTRUCTURE ONLY: only
 * tests/ load it, as the checker of the device kernels.
 *
 * The learned selection has no counterpart in the reference
 * (BasFaggingerAuer/Multicore-Matching); this file restates the specification
 * in include/mm_learn.h: the eight features, the 8-32-32-1 network evaluated
 * with fused multiply-adds in index order, the blue threshold, and rounds whose
 * request / respond / match phases follow the reference rules
 * (src/matchgpu.cu:260-451, the same phases as orc_match in mm_oracle.c) with
 * colours held apart from the match array.  Compiled with -ffp-contract=off,
 * so every float operation is the one written.
 *
 * Parity status (tests/test_oracle_learn.py):
 *   orc_blue_threshold: against 65536 * sigmoid(z) from numpy (within one
 *     step of 2^-16), monotone, saturates at both ends.
 *   orc_mlp_z: against a numpy float64 forward pass (within float rounding).
 *   orc_match_learned: with params == NULL equals orc_match bit for bit;
 *     a network that keeps vertices red still converges (plain rounds after
 *     round 32), to a valid maximal matching;
 *     closed forms (a perfect matching of disjoint heavy edges is matched in
 *     round 1 by a network that follows the hash-order bit; a network that
 *     always says "red" matches nothing and retires nothing while neighbours
 *     remain); validity of every result.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

uint32_t orc_hash(uint32_t i, uint32_t seed);

#define NF 8
#define NH 32
#define OFF_W1 0
#define OFF_B1 256
#define OFF_W2 288
#define OFF_B2 1312
#define OFF_W3 1344
#define OFF_B3 1376
#define DRAW_TWEAK 0x5BD1E995u
#define TRACE_TWEAK 0x2545F491u
#define LEARN_ROUNDS 32 /* rounds whose select uses the network */

static float orc_relu(float a) { return a > 0.0f ? a : 0.0f; }

float orc_mlp_z(const float *P, const float *x)
{
    float h1[NH];
    for (int j = 0; j < NH; ++j) {
        float a = P[OFF_B1 + j];
        for (int k = 0; k < NF; ++k) a = fmaf(P[OFF_W1 + j * NF + k], x[k], a);
        h1[j] = orc_relu(a);
    }
    float z = P[OFF_B3];
    for (int j = 0; j < NH; ++j) {
        float a = P[OFF_B2 + j];
        for (int k = 0; k < NH; ++k) a = fmaf(P[OFF_W2 + j * NH + k], h1[k], a);
        z = fmaf(P[OFF_W3 + j], orc_relu(a), z);
    }
    return z;
}

void orc_mlp_z_many(const float *P, const float *x, int64_t n, float *z)
{
    for (int64_t i = 0; i < n; ++i) z[i] = orc_mlp_z(P, x + i * NF);
}

/* floor(65536 / (1 + 2^(-z log2 e))), 2^f by the degree-7 Taylor polynomial
 * of exp(f ln 2) (coefficients (ln 2)^k / k! rounded to float). */
uint32_t orc_blue_threshold(float z)
{
    const float t = -z * 1.44269502162933349609375f;
    if (!(t < 24.0f)) return 0u;
    if (t < -24.0f) return 65536u;
    const float fi = floorf(t);
    const float f = t - fi;
    static const float c[8] = {1.0f, 0.693147182464599609375f, 0.2402265071868896484375f,
                               0.0555041097104549407958984375f, 0.00961812864989042282104492188f,
                               0.00133335578721016645431518555f, 0.000154035296873189508914947510f,
                               1.52527336467755958437919616699e-05f};
    float p = c[7];
    for (int k = 6; k >= 0; --k) p = fmaf(p, f, c[k]);
    const float den = 1.0f + ldexpf(p, (int)fi);
    /* the largest T with T * den <= 65536, decided in exact double products */
    const double q = floor(65536.0 / (double)den);
    uint32_t T = q > 65536.0 ? 65536u : (uint32_t)q;
    while (T < 65536u && (double)(T + 1u) * (double)den <= 65536.0) ++T;
    while (T > 0u && (double)T * (double)den > 65536.0) --T;
    return T;
}

void orc_blue_threshold_many(const float *z, int64_t n, uint32_t *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = orc_blue_threshold(z[i]);
}

enum { L_BLUE = 0, L_RED = 1, L_RETIRE = 2 };

/*
 * Learned rounds (see include/mm_learn.h).  weighted: 1 heaviest red
 * neighbour / heaviest proposer, 0 first in CSR order (unit weights).  params
 * NULL: plain colouring.  trace_*: when trace_cap > 0, one sample per vertex
 * matched in a round whose sampling hash passes trace_keep, in (round, id)
 * order; *trace_count receives the number offered.  Returns the rounds
 * executed, or -1 when the seeds run out.
 */
int32_t orc_match_learned(int32_t n, const int32_t *ranges, const int32_t *nbrs, const float *w,
                          int32_t weighted, int32_t maximal, int32_t max_rounds, uint32_t barrier,
                          const uint32_t *seeds, int32_t n_seeds, const float *params, int32_t *match,
                          int64_t *round_active, int64_t *round_matched, uint32_t trace_keep, int64_t trace_cap,
                          float *trace_x, float *trace_y, float *trace_w, int32_t *trace_v, int32_t *trace_round,
                          int64_t *trace_count)
{
    const size_t nn = (size_t)(n > 0 ? n : 1);
    int32_t *req = (int32_t *)malloc(sizeof(int32_t) * nn);
    uint8_t *col = (uint8_t *)malloc(nn);
    float *reqw = (float *)malloc(sizeof(float) * nn);
    float *feat = (float *)malloc(sizeof(float) * NF * nn);
    if (!req || !col || !reqw || !feat) {
        free(req); free(col); free(reqw); free(feat);
        return -1;
    }
    for (int32_t i = 0; i < n; ++i) { match[i] = 0; req[i] = -1; col[i] = L_RED; reqw[i] = 0.0f; }
    const int tracing = trace_cap > 0;
    int64_t offered = 0, matched_total = 0;
    int32_t rounds = 0;
    int keep = 1;
    while (rounds < max_rounds && (!maximal || keep)) {
        if (rounds >= n_seeds) { rounds = -1; break; }
        const uint32_t seed = seeds[rounds];
        const int32_t r = rounds + 1;
        int64_t active = 0;

        /* select: every vertex reads match[] and writes only its own colour
         * and features, so the loop runs on all host threads */
#pragma omp parallel for reduction(+ : active) schedule(dynamic, 4096)
        for (int32_t i = 0; i < n; ++i) {
            if (match[i] >= 2) continue;
            ++active;
            const uint32_t hv = orc_hash((uint32_t)i, seed);
            const int plain = hv < barrier ? L_BLUE : L_RED;
            const int net = params && r <= LEARN_ROUNDS;
            if (!net && !tracing) { col[i] = (uint8_t)plain; continue; }
            int32_t alive = 0, redp = 0, ustar = -1;
            float w1 = 0.0f, w2 = 0.0f;
            int have2 = 0;
            for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1]; ++j) {
                const int32_t u = nbrs[j];
                const int32_t mu = match[u];
                if (mu >= 4) continue;
                ++alive;
                redp += mu == 1;
                const float wj = (weighted && w) ? w[j] : 1.0f;
                if (!(wj > -1.0f)) continue;
                if (ustar < 0) { ustar = u; w1 = wj; }
                else if (wj > w1) { w2 = w1; have2 = 1; w1 = wj; ustar = u; }
                else if (!have2 || wj > w2) { w2 = wj; have2 = 1; }
            }
            float *x = feat + (size_t)i * NF;
            const int32_t deg = ranges[2 * i + 1] - ranges[2 * i];
            x[0] = (float)(deg < 1024 ? deg : 1024) * 0.015625f;
            x[1] = w1;
            x[2] = w2;
            x[3] = w1 > 0.0f ? (float)((double)w2 / (double)w1) : 0.0f;
            x[4] = (float)(alive < 1024 ? alive : 1024) * 0.015625f;
            x[5] = (float)(redp < 1024 ? redp : 1024) * 0.015625f;
            x[6] = (float)(r - 1) * 0.03125f;
            x[7] = (ustar >= 0 && hv < orc_hash((uint32_t)ustar, seed)) ? 1.0f : 0.0f;
            if (!net) { col[i] = (uint8_t)plain; continue; }
            if (alive == 0) { col[i] = L_RETIRE; continue; }
            const float z = orc_mlp_z(params, x);
            if (!isfinite(z)) { col[i] = (uint8_t)plain; continue; }
            const uint32_t T = orc_blue_threshold(z);
            col[i] = (orc_hash((uint32_t)i, seed ^ DRAW_TWEAK) >> 16) < T ? L_BLUE : L_RED;
        }
        keep = active > 0;

        /* request (src/matchgpu.cu:286-327 random, 360-409 weighted) */
#pragma omp parallel for schedule(dynamic, 4096)
        for (int32_t i = 0; i < n; ++i) {
            if (match[i] >= 2) continue;
            if (col[i] != L_BLUE) { req[i] = -1; continue; }
            int alive = 0;
            int32_t cand = -1;
            float best = -1.0f;
            for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1]; ++j) {
                const int32_t u = nbrs[j];
                const int32_t mu = match[u];
                if (mu >= 4) continue;
                alive = 1;
                if (mu >= 2 || col[u] != L_RED) continue;
                const float wj = (weighted && w) ? w[j] : 1.0f;
                if (!weighted) { cand = u; best = 1.0f; break; }
                if (wj > best) { best = wj; cand = u; }
            }
            req[i] = cand >= 0 ? cand : (alive ? -1 : -2);
            reqw[i] = cand >= 0 ? best : 0.0f;
        }

        /* respond (src/matchgpu.cu:329-357 random, 411-451 weighted): red
         * vertices write their own req[] and read only blue vertices' req[] */
#pragma omp parallel for schedule(dynamic, 4096)
        for (int32_t i = 0; i < n; ++i) {
            if (match[i] >= 2 || col[i] != L_RED) continue;
            int32_t cand = -1;
            float best = -1.0f;
            for (int32_t j = ranges[2 * i]; j < ranges[2 * i + 1]; ++j) {
                const int32_t u = nbrs[j];
                if (match[u] >= 2 || col[u] != L_BLUE || req[u] != i) continue;
                const float wj = (weighted && w) ? w[j] : 1.0f;
                if (!weighted) { cand = u; best = 1.0f; break; }
                if (wj > best) { best = wj; cand = u; }
            }
            if (cand >= 0) { req[i] = cand; reqw[i] = best; }
        }

        /* match (src/matchgpu.cu:260-283); the others keep this round's colour */
        int64_t newly = 0;
        for (int32_t i = 0; i < n; ++i) {
            if (match[i] >= 2) continue;
            const int32_t q = req[i];
            int32_t nm;
            if (col[i] == L_RETIRE || q == -2) nm = 2;
            else if (q >= 0 && req[q] == i) nm = 4 + (i < q ? i : q);
            else nm = col[i];
            if (nm >= 4) {
                ++newly;
                if (tracing && (orc_hash((uint32_t)i, seed ^ TRACE_TWEAK) >> 16) < trace_keep) {
                    if (offered < trace_cap) {
                        memcpy(trace_x + offered * NF, feat + (size_t)i * NF, sizeof(float) * NF);
                        trace_y[offered] = col[i] == L_BLUE ? 1.0f : 0.0f;
                        trace_w[offered] = reqw[i];
                        if (trace_v) trace_v[offered] = i;
                        if (trace_round) trace_round[offered] = r;
                    }
                    ++offered;
                }
            }
            /* the phase reads req[] and col[] only, so writing in place is safe */
            match[i] = nm;
        }
        matched_total += newly;
        if (round_active && rounds < 256) round_active[rounds] = active;
        if (round_matched && rounds < 256) round_matched[rounds] = matched_total;
        ++rounds;
    }
    if (trace_count) *trace_count = offered;
    free(req); free(col); free(reqw); free(feat);
    return rounds;
}
