/* C ABI of the MI355X (gfx950) MNI support-evaluation library, libtfsm_gpu.so.
 *
 * The reference (lyuheng/T-FSM, parallel build) has no plugin or FFI layer:
 * the seam it exposes is two C++ calls inside one translation unit, and each
 * entry point below replaces one of them:
 *
 *   per pattern, at activation (parallel/comper.h:1655-1700):
 *       GMatchEngine::set(&grami.pruned_graph, pattern);
 *       DPisoFilter(false, freq);                 // gmatch.h:417-472
 *       buildTable(tc->edge_matrix);              // gmatch.h:995-1130
 *       generateGQLQueryPlan(u, order[u]);        // gmatch.h:692-727
 *       generateBN(...);                          // gmatch.h:803-828
 *   per task (comper.h:431-533, 1181-1300):
 *       int GMatchEngine::execute(u, v_idx, embedding, ms);   // gmatch.h:205-222
 *       followed by the MNI bookkeeping of Comper::compute / postprocess.
 *
 * Here one call evaluates a whole pattern (tfsm_eval_pattern: filter, table,
 * plans and every (u, v) existence check with the early-frequent and
 * early-infrequent exits) or every pattern of one extension level
 * (tfsm_eval_level).  Return codes replace the reference's magic ints
 * MATCH_FOUND -1 / MATCH_NOT_FOUND -2 / TIMEOUT -3 (comper.h:30-32) and its
 * exit(-1) on I/O errors: every call returns a tfsm_status, there is no
 * timeout status (every (u, v) gets a final verdict), and nothing exits or
 * throws across the boundary.
 *
 * Threading: a context is not thread-safe; one host thread drives it.
 * Ownership: input arrays are caller-owned and only read during the call;
 * device memory (graph, scratch, retained domains) is context-owned.
 */
#ifndef TFSM_GPU_H
#define TFSM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TFSM_MAX_PATTERN_VERTICES 16
#define TFSM_MAX_PATTERN_EDGES 64

typedef enum {
    TFSM_OK = 0,
    TFSM_ERR_ARGUMENT = -1,     /* malformed input (pattern, sizes, null pointers) */
    TFSM_ERR_HIP = -2,          /* a HIP runtime call failed */
    TFSM_ERR_NO_DEVICE = -3,    /* no usable GPU */
    TFSM_ERR_LIMIT = -4,        /* input beyond a compiled limit */
    TFSM_ERR_HANDLE = -5        /* unknown or released domain handle */
} tfsm_status;

/* Pruned data graph (the output of the reference's construct_freq_graph,
 * graph.h:685-803): dense vertex ids, per-vertex adjacency ascending by
 * neighbour id, labels given as dense indices. */
typedef struct {
    uint32_t num_vertices;
    uint64_t num_adj;               /* directed entries = 2 |E| */
    const uint32_t *offsets;        /* num_vertices + 1 */
    const uint32_t *nbrs;           /* num_adj */
    const uint32_t *elabels;        /* num_adj, dense edge label of each entry */
    const uint32_t *vlabels;        /* num_vertices, dense vertex label */
    uint32_t num_vlabels;
    uint32_t num_elabels;
} tfsm_csr;

/* A connected pattern (the reference's Pattern, graph.h:419-550). */
typedef struct {
    uint32_t num_vertices;
    uint32_t num_edges;
    uint32_t vlabel[TFSM_MAX_PATTERN_VERTICES];     /* dense labels */
    uint32_t edge_a[TFSM_MAX_PATTERN_EDGES];
    uint32_t edge_b[TFSM_MAX_PATTERN_EDGES];
    uint32_t edge_label[TFSM_MAX_PATTERN_EDGES];    /* dense labels */
} tfsm_pattern;

enum {
    TFSM_MODE_CAPPED = 0,   /* reference semantics: stop at the frequent / infrequent verdict */
    TFSM_MODE_EXACT = 1,    /* decide every (u, v): exact MNI support */
    TFSM_MODE_FRACTION = 2  /* Fraction-Score support (the reference's fraction-score build,
                               fraction-score/graph.h:895-922, comper.h:478-508, 1082-1106):
                               u's support is the sum over its valid candidates v of
                               min over u's pattern neighbours' labels l of score(v, l)
                               (tfsm_fraction_scores), the pattern's the minimum over u;
                               frequent when it is at least freq.  Every (u, v) of a pattern
                               is decided unless a vertex can no longer reach freq valid
                               candidates (then the pattern is infrequent: a score is at most 1) */
};

enum {
    TFSM_ORDER_REFERENCE = 0,   /* u ascending, candidates in domain order (comper.h:1181-1239) */
    TFSM_ORDER_LEARNED = 1      /* ordered by the ranker installed with tfsm_ranker_set */
};

/* Matching order inside one existence search.  The verdicts do not depend on
 * it; the work does. */
enum {
    TFSM_MATCH_ADAPTIVE = 0,    /* next = the unmatched neighbour whose shortest backward row under the
                                   partial embedding is shortest (patterns with at most 12 edges;
                                   larger ones use the reference order) */
    TFSM_MATCH_REFERENCE = 1    /* GQL order fixed per start vertex (gmatch.h:692-727, 803-828) */
};

/* Push-down pruning (the reference's sub-pattern cache, decompose.h:19-94 and
 * comper.h:1582-1612): vertex `vertex` of pattern `pattern` of the call keeps
 * only candidates that lie in the retained child domain (C minus proven-invalid
 * candidates) of vertex `source_vertex` of the retained handle `handle`, a
 * frequent sub-pattern that embeds into the pattern with source_vertex on
 * vertex.  Sound: a candidate outside it has no embedding of the sub-pattern,
 * so none of the pattern.  The reference keeps only the proven-invalid part. */
typedef struct {
    uint32_t pattern;           /* index in the call's pattern array */
    uint32_t vertex;
    uint64_t handle;            /* a retained domain handle (tfsm_domain_*) */
    uint32_t source_vertex;
    uint32_t reserved;          /* must be 0 */
} tfsm_pushdown;

typedef struct {
    uint32_t mode;              /* TFSM_MODE_* */
    uint32_t retain_domains;    /* keep a frequent pattern's child domains on the device */
    uint32_t task_order;        /* TFSM_ORDER_* */
    uint32_t trace;             /* record per-task verdict and step count (tfsm_read_trace) */
    uint32_t grid_waves_per_cu; /* resident waves per CU of the search kernel (0: tuned default) */
    uint32_t split_steps;       /* DFS steps after which a search donates its untried
                                   candidates to idle lanes (0: tuned default,
                                   0xffffffff: never); verdicts do not depend on it */
    uint32_t split_fanout;      /* records one donation splits its candidates into
                                   (0: tuned default, at most 64) */
    uint32_t long_steps;        /* DFS steps after which a root search moves to the
                                   wave-cooperative search (0: default, which is
                                   never; 0xffffffff: never); verdicts do not depend on it */
    uint32_t long_waves_per_cu; /* waves per CU of the wave-cooperative search (0: tuned default) */
    uint32_t long_donate_every; /* its rounds between offers of work to idle waves
                                   (0: tuned default, 0xffffffff: never) */
    uint32_t long_fanout;       /* records per offer (0: tuned default, at most 64) */
    uint32_t match_order;       /* TFSM_MATCH_* (0: adaptive) */
    uint32_t group_lanes;       /* most lanes that share one root task when a search launch has
                                   fewer tasks than resident lanes (0: tuned default, 1: one lane
                                   per task, else a power of two <= 64); verdicts do not depend on it */
    uint32_t reserved;          /* must be 0 */
    uint32_t num_pushdown;      /* entries of pushdown (0: none) */
    const tfsm_pushdown *pushdown;
    /* more of the search kernel's shape; none changes a verdict */
    uint32_t steps_per_round;   /* DFS steps a lane takes between bookkeeping rounds (0: 8; at most 64) */
    uint32_t refill_min;        /* fewest and most lanes' worth of root tasks a wave takes from the */
    uint32_t refill_max;        /*   launch's counter at once (0: 128 and 256; 1 <= min <= max <= 4096) */
    uint32_t prune_batch;       /* DP-iso passes launched between convergence checks (0: 6; at most 64);
                                   a device-driven call runs prune_batch passes (0: 8)
                                   and is redone host-driven when they do not reach the fixpoint */
    uint32_t level_mode;        /* TFSM_LEVEL_* (0: device-driven with a captured hipGraph for
                                   tfsm_eval_level, host-driven for tfsm_eval_pattern) */
    uint32_t hot_group_lanes;   /* most lanes per task of the learned order's hot tier (the deciding
                                   vertex's tasks of the patterns expected infrequent; 0: group_lanes,
                                   else a power of two <= 64); verdicts do not depend on it */
    uint32_t ldf_push;          /* a pattern vertex with no parent domain takes its LDF candidates
                                   from the adjacency of a neighbouring vertex's parent domain,
                                   instead of its whole label bucket, when a bound on the edges
                                   walked is at most ldf_push percent of the bucket (0: 100,
                                   in calls whose pushed buckets hold at least 64 items per
                                   thread of one block per CU; 0xffffffff: never); the
                                   filtered domains do not depend on it */
} tfsm_eval_options;

/* How one call's launch sequence is driven.  None changes a result.
 * Host-driven: the host reads the domain sizes back after the filter, makes
 * the decisions and layout and launches the tables and the search.
 * Device-driven: one launch sequence with no host round trip inside the
 * call; the decisions, layout and task counts are computed on the device from
 * regions sized by the candidates' sources, every kernel reads its arguments
 * from a per-context descriptor, and with TFSM_LEVEL_GRAPH the sequence is
 * captured once per context and option set in a hipGraph and replayed.  A
 * call the device-driven path does not take (traced, Fraction-Score, patterns
 * over 8 vertices or in the reference matching order, the wave-cooperative
 * phase, more than 4096 patterns, regions over the device budget) runs
 * host-driven; so does one whose filter needs more passes than the sequence
 * holds (redone from the start). */
enum {
    TFSM_LEVEL_DEFAULT = 0,
    TFSM_LEVEL_HOST = 1,
    TFSM_LEVEL_DEVICE = 2,      /* device-driven, kernels launched one by one */
    TFSM_LEVEL_GRAPH = 3        /* device-driven, one hipGraph launch */
};

/* decided_by */
enum { TFSM_DECIDED_FILTER = 0, TFSM_DECIDED_EXPRESS = 1, TFSM_DECIDED_SEARCH = 2 };

typedef struct {
    int32_t frequent;
    int32_t decided_by;
    uint32_t support_lower;     /* min over u of proven-valid candidates */
    uint32_t support_upper;     /* min over u of candidates not proven invalid */
    uint32_t domain_size[TFSM_MAX_PATTERN_VERTICES];    /* |C(u)| after filtering */
    uint32_t valid_count[TFSM_MAX_PATTERN_VERTICES];
    uint32_t invalid_count[TFSM_MAX_PATTERN_VERTICES];
    uint64_t searches;          /* existence searches started */
    uint64_t verdicts;          /* final (u, v) verdicts, by search or by propagation */
    uint64_t domain_handle;     /* retained child domains (0: none) */
    /* TFSM_MODE_FRACTION only (else 0): per u, the sum of the scores of its
     * proven-valid candidates, exact for every pattern not rejected by the
     * valid-count bound (so for every frequent one); fraction_support is
     * their minimum.  Each sum is the double nearest the exact sum of the
     * scores (ties to even): it does not depend on the order of the
     * additions, which in the reference follows its threads. */
    double fraction[TFSM_MAX_PATTERN_VERTICES];
    double fraction_support;
} tfsm_pattern_result;

typedef struct tfsm_ctx tfsm_ctx;

/* Upload the graph to device 0 (current HIP device).
 * Replaces Worker::load_data's hand-off of grami.pruned_graph (worker.h:47-55). */
int tfsm_ctx_create(const tfsm_csr *g, tfsm_ctx **out);
int tfsm_ctx_destroy(tfsm_ctx *ctx);
/* Capacity of the device pool that holds the auxiliary tables (entries of 4
 * bytes; the default is max(8 x adjacency entries, 2^22)).  It still grows on
 * demand: a call whose tables do not fit is redone with the pool enlarged.
 * No counterpart in the reference, which allocates its tables per pattern
 * (gmatch.h:995-1130). */
int tfsm_ctx_set_pool_capacity(tfsm_ctx *ctx, uint64_t entries);

/* One pattern, one launch sequence (comper.h:1655-1700 + 431-533 + 1181-1300).
 * parent_handle: retained domains of the DFS-code parent (0: label buckets, as
 * for the children of size-2 patterns, worker.h:57-106); the first
 * `parent vertices` candidate lists start from it (gmatch.h:335-359). */
int tfsm_eval_pattern(tfsm_ctx *ctx, const tfsm_pattern *p, uint64_t parent_handle, uint32_t freq,
                      const tfsm_eval_options *opt, tfsm_pattern_result *out);

/* Every pattern of one extension level in one fused launch sequence.
 * Both eval calls return with the results on the host and the copies of the
 * retained child domains still queued on the context's stream; every later
 * call that reads device memory from the host waits for the stream first. */
int tfsm_eval_level(tfsm_ctx *ctx, const tfsm_pattern *ps, const uint64_t *parent_handles, uint32_t n,
                    uint32_t freq, const tfsm_eval_options *opt, tfsm_pattern_result *outs);

/* Fraction scores of the graph's vertices (fraction-score/graph.h:895-922):
 * out[v * num_vlabels + l] = score(v, l), from s = 0, for each neighbour w of
 * v with label l in ascending id order (each adjacency entry once):
 * s = min(1, s + 1 / NLF(w, label(v))), in IEEE double arithmetic in that
 * order, so each value equals the reference's bit for bit; 0 where v has no
 * neighbour with label l.  Computed on the device at the first call of this
 * function or of a TFSM_MODE_FRACTION evaluation and kept with the context
 * (8 * num_vertices * num_vlabels bytes).  Copies up to cap values and
 * returns num_vertices * num_vlabels (negative: a tfsm_status).  The graph
 * may have at most 64 edge labels (TFSM_ERR_LIMIT beyond). */
int64_t tfsm_fraction_scores(tfsm_ctx *ctx, double *out, uint64_t cap);

/* Child domain of pattern vertex u (C(u) minus proven-invalid candidates,
 * comper.h:1001-1018) of a retained handle; *size gets the full size. */
int tfsm_domain_read(tfsm_ctx *ctx, uint64_t handle, uint32_t u, uint32_t *out, uint32_t cap, uint32_t *size);
int tfsm_domain_release(tfsm_ctx *ctx, uint64_t handle);

/* Inspection of the last tfsm_eval_pattern call (tests and tracing):
 *   tfsm_last_domain  : C(u) after filtering (sorted ids), returns size
 *   tfsm_last_states  : per candidate of u: 0 undecided, 1 valid, 2 invalid
 *   tfsm_last_aux     : auxiliary table a->b (rows over C(a), ascending
 *                       indices into C(b)), the reference's edge_matrix[a][b];
 *                       returns entries or -1 when (a, b) is not a pattern edge
 *   tfsm_last_order   : matching order starting at u */
int64_t tfsm_last_domain(tfsm_ctx *ctx, uint32_t u, uint32_t *out, uint32_t cap);
int64_t tfsm_last_states(tfsm_ctx *ctx, uint32_t u, uint8_t *out, uint32_t cap);
int64_t tfsm_last_aux(tfsm_ctx *ctx, uint32_t a, uint32_t b, uint32_t *off, uint32_t off_cap, uint32_t *idx,
                      uint32_t idx_cap);
int tfsm_last_order(tfsm_ctx *ctx, uint32_t u, uint32_t *order);

/* Per-task trace of the last call with opt.trace set, in task order of the
 * reference (pattern, u, candidate index): verdict (0 skipped, 1 valid,
 * 2 invalid, 3 aborted) in bits 30-31 and DFS steps in bits 0-29. */
int64_t tfsm_read_trace(tfsm_ctx *ctx, uint32_t *out, uint64_t cap);

/* Timing and traffic of the last eval call, measured with HIP events on the
 * library's stream.  Algorithmic bytes count each input byte the phase must
 * read and each output byte it must write once:
 *   table  : per directed pattern edge a->b, 8 B of segment bounds and 4 B per
 *            scanned neighbour id for every v in C(a), plus 8 B per row
 *            (start, end) and 4 B per table entry written;
 *   search : per (u, v) verdict, sum over pattern neighbours u' of u of
 *            8 B of row bounds + 4 B per entry of the depth-1 row A[u->u'][v]
 *            (taken at the mean row length of u). */
typedef struct {
    float filter_ms;            /* LDF + DP-iso passes + compaction (incl. host round trips) */
    float table_ms;             /* one-pass table build */
    float search_ms;            /* existence-search kernel launch alone */
    uint32_t search_launches;
    uint64_t table_bytes;
    uint64_t search_bytes;
    uint32_t search_splits;     /* sub-tree pieces donated between lanes */
    uint32_t search_watchdog;   /* 1 when a drained wave left on its idle limit */
    float rank_ms;              /* features, scores and priority sort before the search */
    uint32_t level_mode;        /* TFSM_LEVEL_* that ran the call (the last slice's) */
    uint32_t filter_passes;     /* DP-iso passes that did work */
} tfsm_call_stats;
int tfsm_last_stats(tfsm_ctx *ctx, tfsm_call_stats *out);

/* ---- Learned task ordering (no counterpart in the reference, which always
 * refills tasks in u-ascending, domain order, comper.h:1181-1239).
 *
 * A 24 -> 64 (ReLU) -> 1 perceptron scores every (pattern, u, candidate)
 * task from 24 integer-derived features; its output is a logit of
 * P(candidate valid).  Weights are stored as BF16 on the device and every
 * sum is accumulated in FP32.  With task_order = TFSM_ORDER_LEARNED the
 * scores only reorder and abandon work: the existence search still decides
 * every verdict, so results equal those of the reference order. */
#define TFSM_RANKER_FEATURES 24
#define TFSM_RANKER_HIDDEN 64

typedef struct {
    float w1[TFSM_RANKER_HIDDEN * TFSM_RANKER_FEATURES];   /* row-major [hidden][feature] */
    float b1[TFSM_RANKER_HIDDEN];
    float w2[TFSM_RANKER_HIDDEN];
    float b2;
} tfsm_ranker_weights;

/* Install weights (rounded to BF16) / read back the installed BF16 weights. */
int tfsm_ranker_set(tfsm_ctx *ctx, const tfsm_ranker_weights *w);
int tfsm_ranker_get(tfsm_ctx *ctx, tfsm_ranker_weights *w);
/* Logits of n feature rows (row-major n x 24) with the installed weights. */
int tfsm_ranker_forward(tfsm_ctx *ctx, const float *x, uint64_t n, float *logits);
/* Mean binary cross-entropy over n rows (labels 0/1) and its gradient with
 * respect to the installed weights. */
int tfsm_ranker_gradient(tfsm_ctx *ctx, const float *x, const float *y, uint64_t n, tfsm_ranker_weights *grad,
                         float *loss);
/* tfsm_ranker_gradient for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Inputs and outputs are DEVICE pointers
 * owned by the caller: d_x (n x 24 rows), d_y (n labels), d_grad (the
 * mean gradient, laid out as tfsm_ranker_weights: 1665 floats), d_loss (the
 * mean loss; may be NULL).  It allocates nothing: the sums live in a fixed
 * device array of TFSM_NOALLOC_SCRATCH_FLOATS floats (one per GPU, held
 * until the process ends), so calls that overlap on one GPU race; each
 * call clears the array first.  n is at most TFSM_NOALLOC_MAX_ROWS
 * (TFSM_ERR_LIMIT beyond).  Returns when the results are written. */
#define TFSM_NOALLOC_SCRATCH_FLOATS 1666
#define TFSM_NOALLOC_MAX_ROWS (1ull << 31)
int tfsm_ranker_gradient_noalloc(tfsm_ctx *ctx, const float *d_x, const float *d_y, uint64_t n, float *d_grad,
                                 float *d_loss);
/* Adam on the GPU from the installed weights (FP32 master copy, BF16 device
 * weights refreshed after each step); mini-batches drawn with a seeded
 * shuffle.  Installs the result; *loss gets the last epoch's mean loss. */
int tfsm_ranker_train(tfsm_ctx *ctx, const float *x, const float *y, uint64_t n, uint32_t epochs, uint32_t batch,
                      float lr, uint64_t seed, float *loss);
/* Features of the tasks of the last eval call made with opt.trace set, one
 * row of 24 per task in reference task order (as tfsm_read_trace). */
int64_t tfsm_read_features(tfsm_ctx *ctx, float *out, uint64_t cap_rows);

const char *tfsm_status_string(int status);
const char *tfsm_last_error(tfsm_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif
