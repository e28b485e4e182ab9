// Push-down pruning on the host side (the reference's sub-pattern cache,
// parallel/decompose.h:19-94 and parallel/comper.h:1582-1612, 1035-1041,
// 1113-1119).
//
// Every frequent pattern evaluated on the GPU leaves a retained child domain
// per vertex (C(u) minus proven-invalid candidates).  A new candidate pattern
// P is decomposed as the reference does: for each edge, the connected part of
// P without that edge that holds P's last vertex, when it has more than one
// vertex.  Each such sub-pattern S that is itself a cached frequent pattern
// Q (up to isomorphism) limits P's candidates: vertex x of S (a vertex of P)
// keeps only candidates in Q's child domain of the vertex x maps to.  The
// reference keys its cache by the exact vertex and edge layout; here the key
// is a canonical form, so every isomorphic sub-pattern is found.
#pragma once

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "pattern.h"
#include "../../include/tfsm_gpu.h"

namespace tfsm {

// Canonical form of a small labelled graph: `key` is the same for exactly the
// isomorphic graphs; order[k] is the vertex at canonical position k.
struct Canon {
    std::string key;
    std::vector<int> order;
};
// vertices 0 .. n-1 with labels lab; el[a * n + b] is the edge label, or -1.
Canon canonical_form(int n, const int32_t *lab, const int32_t *el);
Canon canonical_form(const Pattern &p);

// A sub-pattern of P: its vertices (ids in P) and its edges' removal index.
struct SubPattern {
    std::vector<int> verts;            // P's vertex of each sub-pattern vertex
    std::vector<int32_t> lab;
    std::vector<int32_t> el;           // n x n, -1 for no edge
};
// decompose.h:19-94: per edge of P, the part of P without it that holds P's
// last vertex, when it has more than one vertex.
std::vector<SubPattern> decompose(const Pattern &p);

struct PushdownCache {
    struct Entry {
        uint64_t handle;
        std::vector<int> order;        // the cached pattern's vertex at each canonical position
    };
    std::map<std::string, Entry> entries;
    mutable std::mutex mu;
    void insert(const Pattern &p, uint64_t handle);
    // Push-down entries for candidate `pattern_index` (vertices of P), skipping
    // sources that are `skip_handle` (the parent, already the LDF source).
    std::vector<tfsm_pushdown> lookup(const Pattern &p, uint32_t pattern_index, uint64_t skip_handle) const;
};

}  // namespace tfsm
