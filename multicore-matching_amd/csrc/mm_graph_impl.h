// The device-side state of one uploaded graph (or batch of graphs), shared by
// the kernel sources of the matcher library.
#ifndef MM_GRAPH_IMPL_H
#define MM_GRAPH_IMPL_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/mm_match.h"

// Key of a captured launch sequence: every value baked into its kernel arguments.
struct ExecKey {
    int32_t R_run, W, A8, PUSH, G, U, block, hub_deg, nb, bcap, maximal, hub_blocks, count_nnz, r0, r1, cset;
    uint32_t barrier;
};
struct CachedExec {
    ExecKey key;
    hipGraphExec_t exec;
};

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
    float *d_wcsr = nullptr;  // weighted graphs whose rows may repeat an id: weights in CSR order at the internal
                              // positions (the result check's getWeight); null otherwise
    int2 *d_adjp = nullptr;   // adj with each weight replaced by the edge's position in the other end's
                              // sorted row (32-bit proposal keys; PUSH graphs that fit them)
    int locbits = 0;          // bits of local id in a 32-bit proposal key
    int2 *d_head = nullptr;   // weighted: entries 0..3 of each sorted row (adjp when present, else adj),
                              // 4 per vertex (round-1 scans)
    uint32_t *d_headc = nullptr;  // d_head packed to 4 bytes per entry (invalid << 31 | id << head_pb | position)
    int head_pb = 0;
    uint32_t *d_adjc = nullptr;   // allocated with d_headc: every entry of d_adjp packed the same way
                                  // (the list rounds' row scans)
    int32_t *d_match = nullptr;
    int32_t *d_req = nullptr;
    uint8_t *d_st = nullptr;
    int32_t *d_reqs = nullptr;    // requests of the bucket lists by list slot (PUSH list rounds)
    unsigned long long *d_resp = nullptr;
    int4 *d_lists = nullptr;
    int32_t *d_ctr = nullptr;
    int32_t *d_ctrb = nullptr;
    unsigned long long *d_ctr64 = nullptr;
    unsigned long long *d_ctrb64 = nullptr;
    unsigned long long *d_hub64 = nullptr;
    int32_t nb_max = 1;       // bucket slots allocated per counter row
    int64_t n_isolated = 0;   // vertices without neighbours
    int32_t iso_lo = INT32_MAX; // a single relabelled graph with many of them: the first internal id of
                              // those, which the internal order puts after all others
    int32_t *d_iso = nullptr; // graphs with many of them: per round slot and bucket, the isolated vertices
                              // blue and red in that round ([ROUND_SLOTS][2][nb_max]), then per bucket the
                              // last round with one ([nb_max])
    int64_t list_cap = 0;     // capacity of one work list
    uint32_t *d_seeds = nullptr;
    int32_t *h_ctr = nullptr;               // pinned
    unsigned long long *h_ctr64 = nullptr;  // pinned
    // batched graphs
    int32_t n_graphs = 0;
    int32_t *d_loc = nullptr;        // graph-local original id per internal vertex
    int2 *d_unperm = nullptr;        // {original position, internal id} pairs, grouped by eighth of internal ids
    int32_t unperm_off[9] = {0};     // group offsets in d_unperm
    int32_t *d_match_out = nullptr;  // result in the caller's numbering, when relabelled
    int32_t *d_vbase = nullptr;      // batches: first global id of each internal vertex's graph
    void *d_chk = nullptr;           // scratch of the device result check (mm_check.hip lays it out)
    size_t chk_bytes = 0;
    bool chk_perm = false;           // d_chk holds the caller -> internal id map
    unsigned long long *d_chk_keys = nullptr;  // the check's sort keys of out-of-range values
    size_t chk_key_cap = 0;
    bool relabelled = false;
    bool internal_stale = false;     // d_match is behind d_match_out (the plain rounds write only the latter)
    int32_t *d_voff = nullptr;
    uint8_t *d_vround = nullptr;
    int32_t *d_grounds = nullptr;
    std::vector<uint32_t> last_seeds;  // seed table currently on the device
    int ctr_set = 0;                   // counter set of the next run (runs alternate between two)
    std::vector<CachedExec> execs;     // captured launch sequences
    hipEvent_t ev_t1 = nullptr, ev_t2 = nullptr;
    std::vector<hipEvent_t> prof_ev;
    int64_t device_bytes = 0;
    // learned selection (allocated on first use)
    uint8_t *d_lcol = nullptr;      // colour chosen by the learned select: 0 blue, 1 red, 2 no unmatched neighbour
    float *d_lreqw = nullptr;       // weight of the edge a vertex proposed along or accepted
    float *d_lfeat = nullptr;       // [n][8] features of the current round (traces only)
    int32_t *d_lctr = nullptr;      // [MM_MAX_ROUNDS + 2][4] per-round counters of the learned rounds
    int32_t *d_lpart = nullptr;     // per-block partial counts of the current learned round
    int32_t *d_lcnt = nullptr;      // per-block work-list lengths of the current learned round
    int32_t *d_llist = nullptr;     // work lists of the learned rounds (active, blue)
    size_t llist_cap = 0;
    float *d_params = nullptr;      // network parameters
    unsigned long long *d_lsamp = nullptr;  // [2] samples written / offered
    int32_t *h_lctr = nullptr;      // pinned
};

struct mm_batch {
    mm_graph *g = nullptr;
    std::vector<int32_t> vertex_offsets;
};

// Device allocation charged to the graph; MM_OK or MM_ERR_NOMEM (message set).
int32_t mm_graph_alloc(mm_graph *g, void **p, size_t bytes, const char *what);

// Device-side construction of a handle's arrays (mm_upload.hip).  begin copies
// the caller's CSR to the device, checks its row ranges and neighbour ids
// (MM_ERR_INVALID when one leaves its graph) and records the degree
// statistics, rebases a batch's per-graph rows into one CSR by voff / eoff,
// computes the internal order (relabel: the locality order, else the
// caller's); finish builds the rows and, when every row is strictly
// ascending, tests the symmetry that proposal atomics need (sets push_ok_*,
// so d_resp is allocated after it; also keeps every entry's reverse when the
// graph can take 32-bit keys), then the sorted rows, key positions, heads and
// caller-order pairs.  The weights' copy runs beside begin and the rows, up to
// the symmetry test.  mm_upload_free releases the temporaries.
struct UploadState;
// Upload phase times on stderr when the environment sets MM_UPLOAD_TIMING:
// the time since the previous mark (stream s drained first; null: none).
void mm_upload_mark(const char *phase, hipStream_t s);
int32_t mm_upload_begin(mm_graph *g, const int32_t *ranges, const int32_t *nbrs, const float *w, const int32_t *voff,
                        const int64_t *eoff, int32_t n_graphs, int relabel, UploadState **out);
int32_t mm_upload_finish(mm_graph *g, UploadState *u, bool align_rows, bool heads);
void mm_upload_free(UploadState *u);
// Enqueues the copy of the internal match array into the caller's vertex order
// (relabelled graphs only; a no-op otherwise).
void mm_graph_enqueue_unpermute(mm_graph *g);
// The static scratch of the no-allocation check on `device` (the current
// device): its device address and size (mm_scratch.hip).
int32_t mm_noalloc_scratch(int device, char **out, size_t *bytes);
// Rounds a maximal run should run next, from the active counts of its last
// two rounds (before, now); the chunk length between counter read-backs.
int predicted_rounds(int64_t before, int64_t now);

#endif
