// Device-side data layout and kernels of the MNI evaluation library (gfx950).
//
// Layout in HBM:
//   graph  : label-partitioned CSR.  Adjacency of v is sorted by
//            (label(w), edge label, w); seg[v*(K+1)+k] is the absolute start
//            of key k = label*EL + elabel and seg[v*(K+1)+K] the end of v's
//            list.  NLF(v, l) (graph.h:850-866 of the reference) is
//            seg[v][(l+1)EL] - seg[v][l EL]; a label/edge-label restricted
//            neighbour scan is one contiguous range.
//   domain : per pattern vertex, a membership bitmap over graph vertices
//            (W = ceil(n/32) words), a rank array (prefix popcounts per word)
//            and the sorted candidate list C(u).  rank replaces the
//            reference's |V|-sized flag[] map (gmatch.h:1020-1030).
//   table  : per directed pattern edge a->b ("slot"), row offsets over C(a)
//            and ascending indices into C(b) in one shared pool: the
//            reference's edge_matrix[a][b] (types.h:67-90, gmatch.h:995-1130).
//   verdict: valid / invalid bitmaps over candidate indices of each u plus
//            counters, replacing domain_matches / non_candidates
//            (comper.h:104-298).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tfsm_gpu.h"

namespace tfsm_dev {

constexpr int kMaxV = TFSM_MAX_PATTERN_VERTICES;
constexpr int kMaxE = TFSM_MAX_PATTERN_EDGES;
constexpr int kMaxSlots = 2 * kMaxE;
constexpr int kBlock = 256;            // threads per block of every kernel
constexpr int kCutEdges = 12;          // pattern edges of the adaptive matching order
constexpr uint8_t kNoPush = 0xff;
constexpr uint32_t kPushItem = 0x80000000u;   // LDF work entry x: a pushed vertex (its items index src[push_x])

struct DevGraph {
    uint32_t n;        // vertices
    uint32_t W;        // bitmap words
    uint32_t K;        // keys = num_vlabels * num_elabels
    uint32_t EL;       // edge labels
    const uint32_t *seg;     // n * (K + 1)
    const uint32_t *nbr;     // partitioned adjacency
    // per vertex, bit (k mod 64) set for every key k with a neighbour: a
    // necessary condition for v to have a neighbour over key k, tested with
    // one load before the segment offsets
    const unsigned long long *keymask;
};

// The early exits a support mode allows the existence search: the infrequent
// verdict (some vertex can no longer reach freq valid candidates) in capped
// and Fraction-Score mode (a score is at most 1, so a vertex's summed score is
// at most its valid count), the frequent verdict (every vertex reached freq
// valid candidates) in capped mode only.
__host__ __device__ inline bool mode_stops_infrequent(uint32_t mode) { return mode != TFSM_MODE_EXACT; }
__host__ __device__ inline bool mode_stops_frequent(uint32_t mode) { return mode == TFSM_MODE_CAPPED; }

// Fraction-Score sum of one (pattern, vertex): the exact sum of the scores
// of its valid candidates in units of 2^-96, as a 128-bit integer in two
// words (every score is a double in [2^-32, 1], a multiple of 2^-84).
struct FracSum {
    unsigned long long lo, hi;
};

// Host-written description of one pattern in one eval call.
struct PatStatic {
    uint32_t nv, nslots, freq, mode;
    uint32_t vlab[kMaxV];
    uint32_t deg[kMaxV];
    uint32_t nbr_n[kMaxV];
    uint8_t nbr_v[kMaxV][kMaxV];        // pattern neighbours of u
    uint32_t nbr_key[kMaxV][kMaxV];     // segment key of the edge u - nbr_v
    unsigned long long need_mask[kMaxV];   // keymask bits every candidate of u must have
    uint32_t nlf_n[kMaxV];
    uint32_t nlf_lab[kMaxV][kMaxV];
    uint32_t nlf_cnt[kMaxV][kMaxV];
    // LDF sources: parent domain (with its invalid bits) or label bucket
    const uint32_t *src[kMaxV];
    const uint32_t *src_invalid[kMaxV];
    // label bucket sources: the key masks of the bucket's vertices in bucket
    // order (a coalesced read beside src); null for parent domains
    const unsigned long long *src_mask[kMaxV];
    uint32_t src_n[kMaxV];
    // a label-bucket vertex u whose pattern neighbour push_x[u] has a parent
    // domain takes its LDF candidates from that domain's neighbours over the
    // key push_key[u] (label(u) x edge label of x - u) instead of the whole
    // bucket (kNoPush: from the bucket)
    uint8_t push_x[kMaxV];
    uint32_t push_key[kMaxV];
    // domains
    // per 32 data vertices, a pair of words: the membership bits of C(u),
    // then the rank (index in C(u)) of the group's first member, so a table
    // build reads both with one 8-byte load
    uint32_t *bits[kMaxV];
    uint32_t *dom[kMaxV];
    uint32_t dom_size[kMaxV];           // filled by the host once known
    uint32_t *valid_bits[kMaxV];
    uint32_t *invalid_bits[kMaxV];
    // a device-driven call's copies kept after it (retained domains): the
    // domain and its invalid bits, written by the call's last kernel (null:
    // the host copies what it keeps after the call)
    uint32_t *keep_dom[kMaxV];
    uint32_t *keep_inv[kMaxV];
    // directed pattern edges
    uint8_t slot_a[kMaxSlots], slot_b[kMaxSlots];
    uint32_t slot_key[kMaxSlots];
    uint32_t *aux_off[kMaxSlots];       // per row of C(a): (start, end) absolute pool positions
    // matching plans per start vertex u: order[u][d], bn_slot[u][d][k]
    uint8_t order[kMaxV][kMaxV];
    uint8_t bn_n[kMaxV][kMaxV];
    uint8_t bn_slot[kMaxV][kMaxV][kMaxV];
    uint8_t sec_base[kMaxV][kMaxV];     // first cursor slot of depth d's non-primary rows
    uint16_t same_before[kMaxV][kMaxV]; // earlier depths whose vertex has d's label
    uint8_t same_after[kMaxV][kMaxV];   // 1 when a later depth has d's label
    uint32_t task_prefix[kMaxV + 1];    // reference task order: u-major prefix of |C(u)|
    // adaptive matching order (lane_dfs.h): on for this pattern, and the
    // other pattern vertices with each vertex's label
    uint32_t adaptive;
    uint16_t same_lab[kMaxV];
    uint32_t edge_pack[kCutEdges / 4];  // edges a | b << 4, one byte each, 0xff past the last
    uint32_t cut_ao[2 * kCutEdges];     // slot's row pairs at aux_region + cut_ao[slot]
    // device-driven level calls (plan_pattern): a distinct-label acyclic
    // pattern (decided by the filter's arc consistency), and where the
    // pattern's table offsets and verdict bitmaps start in the call's regions
    // (each sized by the pattern's sources, upper bounds of its domains)
    uint32_t express;
    uint32_t aux_base;                  // words into the table-offset region
    uint32_t verdict_base;              // words into the verdict region
    uint32_t aux_m0;                    // its first (pattern, slot) entry of the table work list
};

// Device-written state of one pattern.
struct PatState {
    uint32_t dom_size[kMaxV];
    uint32_t valid_cnt[kMaxV];
    uint32_t invalid_cnt[kMaxV];
    uint32_t ndone;                     // vertices that reached freq valid candidates
    uint32_t status;                    // 0 undecided, 1 frequent, 2 infrequent
    unsigned long long searches;
    uint32_t aux_total[kMaxSlots];      // entries of each slot's table
    uint32_t decided;                   // TFSM_DECIDED_* (device-driven level calls)
    uint32_t pad;
};

// Layout of one pattern once its domain sizes are known (reference
// isCandidateSetValid and the express rule of comper.h:1619-1653, then the
// sizes of buildTable's tables and of the verdict bitmaps): the decision,
// and for a search-decided pattern its task prefix, each slot's row-pair
// array at aux_region + aux_base (slot after slot, two words per row of
// C(slot_a)) and its valid / invalid bitmaps at verdict_region +
// verdict_base (vertex after vertex).  Device-driven level calls run it on
// the device (plan_level, in the last compaction block) and repeat it on the
// host after the call.
// Returns the decision; `dom` holds the sizes.
__host__ __device__ inline uint32_t plan_pattern(PatStatic &P, const uint32_t *dom, uint32_t *aux_region,
                                                 uint32_t *verdict_region) {
    uint32_t mn = 0xffffffffu;
    for (uint32_t u = 0; u < P.nv; ++u) {
        P.dom_size[u] = dom[u];
        mn = dom[u] < mn ? dom[u] : mn;
    }
    if (P.express) return TFSM_DECIDED_EXPRESS;
    if (mn < P.freq) return TFSM_DECIDED_FILTER;
    P.task_prefix[0] = 0;
    for (uint32_t u = 0; u < P.nv; ++u) P.task_prefix[u + 1] = P.task_prefix[u] + P.dom_size[u];
    uint32_t ao = P.aux_base;
    for (uint32_t sl = 0; sl < P.nslots; ++sl) {
        if (sl < 2u * kCutEdges) P.cut_ao[sl] = ao;
        P.aux_off[sl] = aux_region + ao;
        ao += 2u * P.dom_size[P.slot_a[sl]];
    }
    uint32_t vo = P.verdict_base;
    for (uint32_t u = 0; u < P.nv; ++u) {
        const uint32_t wds = (P.dom_size[u] + 31u) / 32u;
        P.valid_bits[u] = verdict_region + vo;
        vo += wds;
        P.invalid_bits[u] = verdict_region + vo;
        vo += wds;
    }
    return TFSM_DECIDED_SEARCH;
}

// A push-down source: a retained child domain (sorted ids) and its bitmap
// of proven-invalid entries.
struct PushSrc {
    const uint32_t *dom;
    const uint32_t *invalid;
    uint32_t n;
    uint32_t pad;
};

// One device-to-device copy of `words` 32-bit words (a zero fill when src is null).
struct CopyDesc {
    const uint32_t *src;
    uint32_t *dst;
    unsigned long long words;
};

// Flattened work over (pattern, x, item) ranges: item t belongs to range m
// with base[m] <= t < base[m+1].
struct WorkTable {
    const unsigned long long *base;     // M + 1
    const uint32_t *pat;                // M
    const uint32_t *x;                  // M (pattern vertex or slot)
    uint32_t M;
    unsigned long long total;
};

// A donated piece of one (u, v) search: the candidates [lo, hi) at depth
// `floor` of the matching order of u, below the fixed prefix of candidate
// indices chosen at depths 0 .. floor-1 (adaptive matching order: of the
// pattern vertices pord[0 .. floor]).
struct SplitRecord {
    uint32_t seq;                          // the launch epoch once the record is written
    uint32_t p, u, ci, floor, cnt_id, lo, hi;
    uint32_t prim;                         // which backward-neighbour row [lo, hi) indexes
    uint32_t prefix[kMaxV];
    uint32_t cpos;                         // wave-cooperative pieces: tested candidates at
    uint32_t cmask_lo, cmask_hi;           //   positions cpos + i, i a set bit of the mask
    uint8_t pord[kMaxV];                   // pattern vertex per depth (adaptive order)
};
static_assert(sizeof(SplitRecord) == 128, "one record per 128-byte line");

// Global bookkeeping of one search launch (zeroed before each launch).
// Each hot counter sits on its own 128-byte line so that polling one does not
// slow the atomics on another.
struct SearchCounters {
    unsigned long long main_next;          // next root task
    uint32_t pad0[30];
    uint32_t busy_waves;                   // waves with running pieces or claimed records
    uint32_t pad1[31];
    uint32_t split_tail;                   // records reserved by producers
    uint32_t pad2[31];
    uint32_t split_head;                   // records claimed by consumers (<= split_tail)
    uint32_t pad3[31];
    uint32_t cnt_alloc;                    // piece counters handed out
    uint32_t watchdog;                     // set when a wave left on its idle limit
    uint32_t pad4[30];
    uint32_t long_cnt;                     // root tasks handed to the wave-cooperative search
    uint32_t pad5[31];
    uint32_t long_next;                    // next of those taken by a wave
    uint32_t pad6[31];
    uint32_t long_busy;                    // second-phase waves with a task or piece
    uint32_t pad7[31];
    unsigned long long hot_next;           // next hot-tier entry
    uint32_t pad8[30];
};

struct VerifyArgs {
    const PatStatic *ps;
    PatState *st;
    const unsigned long long *task_base;   // P + 1; task_base[P] is the number of tasks
    uint32_t P;
    uint32_t epoch;                        // record epoch (device-driven level calls)
    SearchCounters *ctr;
    const uint32_t *pool;
    const uint32_t *aux_region;            // every slot's row offsets of the call
    uint32_t max_edges;                    // the most edges of any searched pattern of the call
    const uint32_t *overflow;              // nonzero: the tables did not fit the pool (do nothing)
    uint32_t *trace;                       // nullable, indexed by reference task order
    const uint32_t *order;                 // nullable: search order -> reference task index
    const uint32_t *order_pu;              // with order: (pattern << 4) | pattern vertex of each entry
    const uint32_t *split;                 // with order: entries before *split form the first part
    uint32_t part;                         // 0: every entry, 1: the first part, 2: the rest
    uint32_t abort_stopped;                // capped mode: abandon tasks of a u that has freq valid
    uint32_t steps_per_round;
    uint32_t refill_min, refill_max;     // lanes' worth of root tasks per wave refill
    uint32_t split_steps;                  // donate work after this many steps (0: never)
    uint32_t split_fanout;                 // records per donation (>= 1)
    SplitRecord *records;                  // split_cap records
    uint32_t split_cap;
    uint32_t *piece_cnt;                   // split_cap counters
    uint32_t *piece_incomplete;            // split_cap flags
    uint32_t long_steps;                   // root tasks still running after this many steps move on
    uint32_t long_cap;                     // capacity of long_list (0: no wave-cooperative search)
    uint32_t *long_list;                   // reference task indices for the wave-cooperative search
    uint32_t long_donate_every;            // its rounds between offers of work (0: never)
    uint32_t long_fanout;                  // records per offer (>= 1)
    uint32_t group_max;                    // most lanes per root task (power of two; <= 1: one)
    uint32_t hot_group_max;                // the same for the hot tier's tasks
    uint32_t group_base;                   // piece counter of this launch's entry 0 in a lane group
    uint32_t group_cap;                    // entries a launch may run in lane groups
    const uint32_t *hot;                   // with order: entries before *hot (of the first part)
                                           // form the hot tier, searched with the largest groups
};

constexpr uint32_t kClockWays = 16;

// Everything the kernels of one device-driven level call read, in one
// descriptor per context slot (c_level in kernels.inc).  The kernels of such
// a call take only the slot (and a pass or part number), so their launch
// sequence is the same for every level with the same options and can be
// captured once in a hipGraph and replayed; the host writes the descriptor
// before each launch.  Sizes the host cannot know before the filter (domain
// sizes, task counts, table work) are computed on the device by
// the last compaction block (plan_level) and read by the later kernels from
// device memory.
struct LevelArgs {
    const CopyDesc *fills;                 // zero fills before the filter
    uint32_t nfills;
    uint32_t P;                            // patterns
    DevGraph g;
    PatStatic *ps;
    PatState *st;                          // P + 1: the last entry's first word is the filter's change flag
    // filter
    WorkTable wldf;                        // LDF: (pattern, u, source item), pushed vertices flagged
    WorkTable wsrc;                        // DP-iso: (pattern, u, source candidate)
    const uint32_t *ex_off;                // push-down sources (null: none)
    const PushSrc *ex;
    uint32_t *pflags;                      // per DP-iso pass: 1 when it removed a candidate
    uint32_t *chg;                         // three change-flag arrays of nchg words
    uint32_t nchg;
    uint32_t passes;                       // DP-iso passes of the call
    const uint32_t *pairs_p, *pairs_u;     // compaction: (pattern, u) pairs
    uint32_t npairs;
    uint32_t *passes_run;                  // passes that did work (counted by the passes themselves)
    uint32_t *blocks_done;                 // compaction blocks finished (the last one plans the level)
    // layout (plan_level)
    uint32_t *aux_region;
    uint32_t *verdict_region;
    unsigned long long *task_base;         // P + 1
    unsigned long long *aux_base;          // aux_M + 1: the table work list's ranges
    const uint32_t *aux_pat, *aux_x;       // aux_M: (pattern, slot) of every slot of every pattern
    uint32_t aux_M;
    // tables
    uint32_t *pool, *pool_used, *overflow;
    uint32_t pool_cap;
    // learned order
    const uint16_t *W;
    float *logits, *sums, *qsums, *ssums;
    uint32_t *keys, *pu, *hist, *tiles, *order, *order_pu;
    uint32_t nhist, ntiles;
    // existence search: the first launch and the deferred one
    VerifyArgs va[2];
    // phase clock (the device's constant-rate wall clock), five stamps of
    // kClockWays words each, a block using word (block mod kClockWays) so
    // that the atomics spread over lines: [0] the first block's start (min),
    // [1] the end of the filter and the plan, [2] of the tables (max), [3]
    // the first search block's start (min), [4] the last search block's end
    // (max); the host uploads the starting values with the call and reduces
    // each stamp's words
    unsigned long long *clk;
};

}  // namespace tfsm_dev
