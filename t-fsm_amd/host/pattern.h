// Patterns as gSpan DFS codes, rightmost-path extension and the minimum-DFS-
// code test.  Host-side pattern growth that produces the candidate patterns
// whose MNI support the GPU evaluates.
//
//   * DFS-code order and `Pattern::add_edge` semantics: parallel/graph.h:245-339
//     (dfs_code_t::compareTo) and graph.h:934-979 (add_edge/extend);
//   * rightmost path: graph.h:919-932 (buildRMPath);
//   * extension families and their label pruning: parallel/grami.h:172-294;
//   * canonical test: parallel/grami.h:840-872 and canonical.h:157-346.
#pragma once

#include <cstdint>
#include <map>
#include <set>
#include <string>
#include <vector>

namespace tfsm {

constexpr int kMaxPatternVertices = 16;
constexpr int kMaxPatternEdges = 64;

struct DfsCode {
    int from = 0, to = 0;
    int32_t fl = 0, el = 0, tl = 0;   // from label, edge label, to label (original values)
    bool operator==(const DfsCode &o) const {
        return from == o.from && to == o.to && fl == o.fl && el == o.el && tl == o.tl;
    }
    bool operator!=(const DfsCode &o) const { return !(*this == o); }
    bool forward() const { return from < to; }
};

struct Pattern {
    std::vector<DfsCode> code;
    std::vector<int32_t> vlab;                                  // per pattern vertex
    std::vector<std::vector<std::pair<int, int32_t>>> adj;      // (neighbour, edge label), ascending neighbour
    std::vector<int> rmpath;                                    // code indices, last edge first

    int nv() const { return (int)vlab.size(); }
    int ne() const { return (int)code.size(); }
    void push(const DfsCode &c);            // add the edge and rebuild the rightmost path
    bool adjacent(int a, int b) const;
    int32_t edge_label(int a, int b) const; // requires adjacent(a, b)
    int degree(int a) const { return (int)adj[a].size(); }
    bool distinct_labels() const;
    bool acyclic() const { return ne() == nv() - 1; }   // connected by construction
    std::string key() const;                 // "fl,el,tl,from,to;..." over the code, for sets and maps
    std::string to_lg_block() const;         // `v id label` lines then `e from to elabel` lines
};

// Label adjacency learnt from the frequent size-2 patterns (grami.h:153-162).
struct GrowthContext {
    std::map<int32_t, std::set<int32_t>> neighbor_labels;
    std::set<int32_t> freq_edge_labels;
    int max_nodes = -1;                      // -1: unlimited
};

// Minimum DFS code test on the pattern's own graph.
bool is_min_code(const Pattern &p);

// Rightmost extensions of `p` that pass `is_min_code`, in a deterministic order.
std::vector<Pattern> extend(const Pattern &p, const GrowthContext &ctx);

// Single-edge pattern (0, 1, la, el, lb).
Pattern single_edge(int32_t la, int32_t el, int32_t lb);

}  // namespace tfsm
