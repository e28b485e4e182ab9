// Level-synchronous MNI miner on top of the GPU evaluation library.
//
// Replaces the reference's Worker / Comper task loop (worker.h:4-166,
// comper.h:1302-1715) for this build: the size-2 level comes from the edge
// statistics (grami.h:83-169, 459-481), every later level is the set of
// canonical rightmost extensions of the previous level's frequent patterns
// (grami.h:172-294), evaluated on the GPU either one pattern per launch
// sequence (plain) or the whole level in one fused sequence (batched).  A
// child's first candidate lists start from its parent's domains minus the
// parent's proven-invalid candidates (comper.h:1069-1152); children of
// size-2 patterns start from the label buckets (worker.h:57-106).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "graph.h"
#include "pattern.h"
#include "../../include/tfsm_gpu.h"

namespace tfsm {

struct MinerOptions {
    int32_t freq = 0;
    int max_nodes = -1;
    int threads = 1;           // host threads for pattern growth
    bool batched = false;      // one eval call per level instead of per pattern
    int mode = TFSM_MODE_CAPPED;
    bool learned = false;      // learned task order (requires a ranker on the context)
    bool trace = false;        // keep per-task traces for ranker training
    bool pushdown = false;     // push-down pruning from cached frequent sub-patterns (pushdown.h)
    // search-kernel shape (grid_waves_per_cu, split_*, long_*); mode, order,
    // trace and retention are set from the fields above
    tfsm_eval_options tuning{};
};

struct FoundPattern {
    Pattern p;
    uint32_t support_lower = 0, support_upper = 0;
    double fraction = 0;       // Fraction-Score support (TFSM_MODE_FRACTION)
};

struct LevelStats {
    uint32_t size = 0;          // pattern vertex count of the level
    uint64_t evaluated = 0, frequent = 0;
    double wall_ms = 0;
    double prep_ms = 0;         // host: device patterns and growth start, before the eval calls
    double eval_ms = 0;         // host wall time inside the level's eval calls
    double post_ms = 0;         // host: verdicts, retained handles and the next level's candidates
};

struct MinerStats {
    double mining_s = 0;
    uint64_t evaluated = 0;         // candidate patterns evaluated (size >= 3)
    uint64_t frequent = 0;          // frequent patterns, size-2 included
    uint64_t verdicts = 0;          // final (u, v) verdicts
    uint64_t searches = 0;          // existence searches
    uint64_t eval_calls = 0;        // tfsm_eval_* calls
    double gpu_ms[3] = {0, 0, 0};   // filter, table, search (device time)
    uint64_t search_launches = 0;
    double search_kernel_ms = 0;    // sum of search-kernel launch durations
    uint64_t bytes_table = 0, bytes_search = 0;   // algorithmic bytes
    std::vector<LevelStats> levels;
    int capped_vertices = 0;        // nonzero: growth stopped at this many vertices (the device limit)
    uint64_t dropped_large = 0;     // children over the device's edge limit, not evaluated
    uint64_t pushdown_sources = 0;  // push-down entries handed to the evaluation
    uint64_t domain_sum = 0;        // filtered domain sizes summed over the evaluated patterns
    double eval_wall_ms = 0;        // host wall time inside the eval calls
    double growth_wait_ms = 0;      // host wall time waiting for the next level's growth
    uint64_t filter_passes = 0;     // DP-iso passes that did work
    uint64_t device_calls = 0;      // eval calls run device-driven (TFSM_LEVEL_DEVICE or _GRAPH)
    uint64_t graph_calls = 0;       // of those, launched as one captured hipGraph
};

// Runs the whole mining job; `found` receives the frequent patterns in a
// deterministic order (by size, then DFS-code key).  Returns 0 or a tfsm_status.
int mine(const DataGraph &g, tfsm_ctx *ctx, const MinerOptions &opt, std::vector<FoundPattern> &found,
         MinerStats &stats, std::string &err);

// Host CSR view of the pruned graph for tfsm_ctx_create.
tfsm_csr csr_view(const DataGraph &g);

// grami_results.txt in the reference serial build's format (serial/grami.h:622-635).
void write_results(const std::vector<FoundPattern> &found, const std::string &path);

}  // namespace tfsm
