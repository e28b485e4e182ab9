// Data graph for the MNI miner: `.lg` ingest, frequent-label pruning and
// renumbering, and the size-2 edge statistics.
//
// Behaviour follows the reference loader, parallel/graph.h:557-803:
//   * vertex labels with fewer than `freq` vertices are dropped, and edge
//     labels with fewer than `freq` edges (graph.h:662-683);
//   * the kept vertices are renumbered densely in `v`-line order, kept edges
//     keep `e`-line order, and each adjacency list is sorted by neighbour id
//     (graph.h:685-769);
//   * the size-2 statistics ("hashedEdges", graph.h:733-753) are taken over
//     every input edge, with an endpoint whose label was dropped mapped to new
//     id 0 (the reference's `id_map[...]` default), and entries with either
//     side below `freq` removed (graph.h:789-800).
#pragma once

#include <cstdint>
#include <string>
#include <vector>
#include <map>
#include <set>
#include <tuple>

namespace tfsm {

struct InputGraph {
    std::vector<uint32_t> vid;      // ids in `v`-line order
    std::vector<int32_t> vlab;      // label of the i-th `v` line
    std::vector<int32_t> label_of;  // label indexed by vertex id (size max id + 1)
    std::vector<uint32_t> eu, ev;   // edges in `e`-line order
    std::vector<int32_t> el;
};

// Reads a `t/v/e` file.  Throws std::runtime_error on I/O or format errors.
InputGraph read_lg(const std::string &path);

struct EdgeStat {                  // one surviving size-2 pattern (la <= lb)
    int32_t la, el, lb;
    uint32_t na, nb;               // |candA|, |candB|; MNI = min(na, nb)
};

struct DataGraph {
    int32_t freq = 0;
    uint32_t n = 0;                          // vertices after pruning
    std::vector<int32_t> vlabel_values;      // dense vertex label -> original label (ascending)
    std::vector<int32_t> elabel_values;      // dense edge label -> original label (ascending)
    std::vector<uint32_t> vlabel;            // dense label per vertex
    std::vector<uint32_t> offsets;           // n + 1
    std::vector<uint32_t> nbrs;              // per vertex ascending by neighbour id
    std::vector<uint32_t> elab;              // dense edge label per adjacency entry
    std::vector<EdgeStat> edge_stats;        // sorted by (la, el, lb)

    int dense_vlabel(int32_t orig) const;    // -1 if not frequent
    int dense_elabel(int32_t orig) const;
    uint64_t num_adj() const { return nbrs.size(); }
};

DataGraph prune_graph(const InputGraph &in, int32_t freq);

}  // namespace tfsm
