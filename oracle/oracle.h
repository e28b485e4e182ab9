/* CPU oracle for the T-FSM MNI support-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain CPU restatement of the reference
 * algorithm (lyuheng/T-FSM, parallel build), used as the checker in tests/,
 * in the smoke check and as a timed CPU restatement in bench.py.  Nothing in
 * the product (t-fsm_amd/) may include, link or call it.
 *
 * Every function cites the reference file:line it restates in oracle.cpp.
 * Pinning (what checks each function against something other than itself):
 *   orc_load                 - golden pattern sets of the reference binary,
 *                              hand-computed sample-graph fixture
 *   orc_edge_stats           - closed form min(|A|,|B|) on fixtures
 *   orc_eval (filter)        - soundness vs brute force (domains contain every
 *                              true MNI vertex) and arc-consistency invariant
 *   orc_eval (aux table)     - recomputed from the definition by a second,
 *                              direct adjacency test in tests/
 *   orc_eval (verdict, MNI)  - brute-force enumeration on tiny graphs, closed
 *                              forms (cliques, cycles, stars), reference
 *                              pattern sets through orc_mine
 *   orc_mine                 - the reference binaries' pattern sets
 *   orc_fraction_scores      - bit for bit against the reference's own
 *                              compute_fraction_score (a harness around
 *                              fraction-score/graph.h), closed forms on stars
 *   orc_edge_fraction,
 *   orc_exact_sum            - Python's math.fsum on the same values; the
 *                              reference harness's sums to within rounding
 *   orc_eval / orc_mine (mode 2) - the reference fraction-score build's
 *                              pattern sets
 */
#ifndef TFSM_ORACLE_H
#define TFSM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_graph orc_graph;
typedef struct orc_eval orc_eval;

/* Load and prune a `.lg` graph at support `freq`; NULL on error. */
orc_graph *orc_load(const char *path, int freq);
void orc_graph_free(orc_graph *g);
uint32_t orc_num_vertices(const orc_graph *g);
uint64_t orc_num_adj(const orc_graph *g);
/* offsets[n+1], nbrs[num_adj], elabels[num_adj] (original values), vlabels[n] (original values) */
void orc_csr(const orc_graph *g, uint32_t *offsets, uint32_t *nbrs, int32_t *elabels, int32_t *vlabels);
/* Surviving size-2 statistics: 5 int32 per entry (la, el, lb, |A|, |B|); returns count. */
int orc_edge_stats(const orc_graph *g, int32_t *out, int cap);

/* Fraction-Score table (fraction-score/graph.h:895-922): out[v * L + l], l the
 * rank of the label among the pruned graph's labels (ascending); returns n * L. */
int64_t orc_fraction_scores(const orc_graph *g, double *out, uint64_t cap);
/* Size-2 Fraction-Score sums (fraction-score/graph.h:803-824) per surviving
 * size-2 entry, in orc_edge_stats order: (sum over candA, sum over candB). */
int orc_edge_fraction(const orc_graph *g, double *out, int cap);
/* The double nearest the exact sum of x[0..n) (Shewchuk partials, ties to even). */
double orc_exact_sum(const double *x, uint64_t n);

/* Evaluate one pattern.
 *   nv, vlab[nv]           pattern vertex labels (original values)
 *   ne, edges[3*ne]        (a, b, edge label) per pattern edge
 *   parent_nv              number of leading pattern vertices whose candidates
 *                          start from the parent domains (0: none)
 *   parent_off[parent_nv+1], parent_dom[...]   parent domains (sorted ids)
 *   mode                   0 capped (reference early exits), 1 exact,
 *                          2 Fraction-Score (fraction-score/comper.h)
 */
orc_eval *orc_eval_pattern(const orc_graph *g, int nv, const int32_t *vlab, int ne, const int32_t *edges,
                           int parent_nv, const uint32_t *parent_off, const uint32_t *parent_dom,
                           int freq, int mode);
void orc_eval_free(orc_eval *e);
int orc_eval_frequent(const orc_eval *e);
/* 0: rejected by the filter (a domain below freq), 1: express path, 2: search */
int orc_eval_decided_by(const orc_eval *e);
uint32_t orc_eval_support_lower(const orc_eval *e);
uint32_t orc_eval_support_upper(const orc_eval *e);
uint64_t orc_eval_searches(const orc_eval *e);
uint64_t orc_eval_verdicts(const orc_eval *e);
/* Mode 2: per u, the summed scores of the candidates found valid (all of
 * them when the search ran to the end); returns the minimum over u. */
double orc_eval_fraction(const orc_eval *e, double *per_vertex);
/* Filtered candidate domain of u (after DP-iso); returns its size, copies up to cap ids. */
uint32_t orc_eval_domain(const orc_eval *e, int u, uint32_t *out, uint32_t cap);
/* Per-candidate state of u: 0 undecided, 1 valid, 2 invalid; returns |C(u)|. */
uint32_t orc_eval_states(const orc_eval *e, int u, uint8_t *out, uint32_t cap);
/* Domain of u handed to children: C(u) minus proven-invalid candidates. */
uint32_t orc_eval_child_domain(const orc_eval *e, int u, uint32_t *out, uint32_t cap);
/* Auxiliary table rows over C(a) listing indices into C(b); -1 if (a,b) is not a pattern edge.
 * Returns the number of entries; fills off[|C(a)|+1] and idx[...] when non-NULL. */
int64_t orc_eval_aux(const orc_eval *e, int a, int b, uint32_t *off, uint32_t *idx);
/* GQL matching order starting at u (nv entries). */
void orc_eval_order(const orc_eval *e, int u, uint32_t *order);
/* Backward neighbours of depth d in the order starting at u (up to nv entries, in
 * the pattern's adjacency order); returns their number, -1 without plans. */
int orc_eval_bn(const orc_eval *e, int u, int d, uint32_t *out);

/* Brute-force MNI: enumerates every injective label- and edge-label-preserving
 * map of the pattern into the pruned graph.  Fills images[u] counts; returns MNI.
 * Only for tiny graphs. */
uint32_t orc_bruteforce_mni(const orc_graph *g, int nv, const int32_t *vlab, int ne, const int32_t *edges,
                            uint32_t *image_sizes);

/* Full CPU mining run (pattern growth from the product host code, evaluation
 * from this oracle) over `threads` threads.  Writes one line per frequent
 * pattern to out_path: "<dfs code key>\t<support lower>\t<support upper>".
 * Returns the number of frequent patterns (size-2 included), or -1. */
int64_t orc_mine(const orc_graph *g, int freq, int max_nodes, int threads, int mode, const char *out_path,
                 double *mining_seconds, uint64_t *patterns_evaluated, uint64_t *verdicts);

#ifdef __cplusplus
}
#endif

#endif
