// Reference-side binding: the header a maintainer of lyuheng/T-FSM adds to
// its parallel/ tree to evaluate patterns with this library instead of
// GMatchEngine.  It is written against the reference's own types:
//   Graph::vertices_ (a Vertices = vector<vertex_t>, parallel/graph.h:62, 104-145),
//   vertex_t::{id, label, edges} and edge_t::{from, label, to} (graph.h:27-60),
//   Pattern::{vertices_, dfscodes} (graph.h:419-550) and
//   dfs_code_t::{from, to, edge_label} (graph.h:245-339).
// Include it after the reference's graph.h; link with -ltfsm_gpu and put this
// repository's include/ on the include path.
//
// It replaces two seams of the reference:
//   activation (comper.h:1655-1700): gmatch_engine.set, DPisoFilter,
//     buildTable, generateGQLQueryPlan / generateBN;
//   the per-task execute calls (comper.h:431-533) and the MNI bookkeeping
//     around them (comper.h:1181-1300).
// One gpu_eval call does both for one pattern; gpu_eval_level for a level.
#ifndef TFSM_REF_GPU_EVAL_H
#define TFSM_REF_GPU_EVAL_H

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "tfsm_gpu.h"

// The pruned data graph with dense labels, uploaded once.
struct GpuGraph {
    std::vector<uint32_t> off, nbr, el, vl;
    std::map<vLabel, uint32_t> vmap;   // original -> dense vertex label
    std::map<eLabel, uint32_t> emap;   // original -> dense edge label
    tfsm_ctx *ctx = nullptr;
    std::string error;
};

// `pruned` is GraMi::pruned_graph after construct_freq_graph (graph.h:685-803):
// vertex ids are dense and each adjacency list is sorted by neighbour id.
// Returns a tfsm_status.
inline int gpu_graph_init(GpuGraph &G, const Graph &pruned) {
    const uint32_t n = (uint32_t)pruned.vertices_.size();
    G.off.assign(n + 1, 0);
    G.vl.resize(n);
    for (uint32_t v = 0; v < n; ++v) {
        const vertex_t &vx = pruned.vertices_[v];
        G.vmap.emplace(vx.label, (uint32_t)G.vmap.size());
        for (const edge_t &e : vx.edges) G.emap.emplace(e.label, (uint32_t)G.emap.size());
    }
    for (uint32_t v = 0; v < n; ++v) {
        const vertex_t &vx = pruned.vertices_[v];
        for (const edge_t &e : vx.edges) {
            G.nbr.push_back((uint32_t)e.to);
            G.el.push_back(G.emap.at(e.label));
        }
        G.off[v + 1] = (uint32_t)G.nbr.size();
        G.vl[v] = G.vmap.at(vx.label);
    }
    tfsm_csr c;
    c.num_vertices = n;
    c.num_adj = G.nbr.size();
    c.offsets = G.off.data();
    c.nbrs = G.nbr.data();
    c.elabels = G.el.data();
    c.vlabels = G.vl.data();
    c.num_vlabels = (uint32_t)G.vmap.size();
    c.num_elabels = (uint32_t)G.emap.size();
    const int rc = tfsm_ctx_create(&c, &G.ctx);
    if (rc != TFSM_OK) G.error = tfsm_status_string(rc);
    return rc;
}

inline void gpu_graph_destroy(GpuGraph &G) {
    if (G.ctx) tfsm_ctx_destroy(G.ctx);
    G.ctx = nullptr;
}

// Dense-label form of a reference Pattern.  False when a label does not occur
// in the pruned graph: the pattern then has no embedding and is infrequent.
inline bool gpu_pattern(const GpuGraph &G, const Pattern &p, tfsm_pattern &tp) {
    tp = tfsm_pattern{};
    if (p.vertices_.size() > TFSM_MAX_PATTERN_VERTICES || p.dfscodes.size() > TFSM_MAX_PATTERN_EDGES) return false;
    tp.num_vertices = (uint32_t)p.vertices_.size();
    for (uint32_t i = 0; i < tp.num_vertices; ++i) {
        auto it = G.vmap.find(p.vertices_[i].label);
        if (it == G.vmap.end()) return false;
        tp.vlabel[i] = it->second;
    }
    uint32_t k = 0;
    for (const dfs_code_t &d : p.dfscodes) {   // <from, to, from_label, edge_label, to_label>
        auto it = G.emap.find(d.edge_label);
        if (it == G.emap.end()) return false;
        tp.edge_a[k] = (uint32_t)d.from;
        tp.edge_b[k] = (uint32_t)d.to;
        tp.edge_label[k] = it->second;
        ++k;
    }
    tp.num_edges = k;
    return true;
}

// Replaces DPisoFilter / buildTable / the plans and every execute task of one
// pattern.  parent_handle: the parent's retained domains (0 for the children
// of size-2 patterns), the role of Pattern::parent_prog (graph.h:402-414,
// comper.h:1001-1018).  *child_handle gets this pattern's retained domains
// when it is frequent (0 otherwise): pass it as its children's parent_handle
// and release it with tfsm_domain_release once they have all been evaluated.
// Returns a tfsm_status; *frequent gets the verdict.
inline int gpu_eval(GpuGraph &G, const Pattern &p, uint64_t parent_handle, int freq, bool *frequent,
                    uint64_t *child_handle) {
    *frequent = false;
    *child_handle = 0;
    tfsm_pattern tp;
    if (!gpu_pattern(G, p, tp)) return TFSM_OK;
    tfsm_eval_options opt = tfsm_eval_options();   // zeros: capped mode, reference task order,
                                                   // adaptive matching order, tuned search shape
    opt.mode = TFSM_MODE_CAPPED;
    opt.retain_domains = 1;
    tfsm_pattern_result r;
    const int rc = tfsm_eval_pattern(G.ctx, &tp, parent_handle, (uint32_t)freq, &opt, &r);
    if (rc != TFSM_OK) {
        G.error = tfsm_last_error(G.ctx);
        return rc;
    }
    *frequent = r.frequent != 0;
    *child_handle = r.domain_handle;
    return TFSM_OK;
}

// Every pattern of one level in one call; frequent[i] and child_handles[i]
// as for gpu_eval.
inline int gpu_eval_level(GpuGraph &G, const std::vector<Pattern *> &ps, const std::vector<uint64_t> &parents,
                          int freq, std::vector<char> &frequent, std::vector<uint64_t> &child_handles) {
    frequent.assign(ps.size(), 0);
    child_handles.assign(ps.size(), 0);
    std::vector<tfsm_pattern> tps;
    std::vector<uint64_t> par;
    std::vector<size_t> where;
    for (size_t i = 0; i < ps.size(); ++i) {
        tfsm_pattern tp;
        if (!gpu_pattern(G, *ps[i], tp)) continue;
        tps.push_back(tp);
        par.push_back(parents[i]);
        where.push_back(i);
    }
    if (tps.empty()) return TFSM_OK;
    tfsm_eval_options opt = tfsm_eval_options();
    opt.mode = TFSM_MODE_CAPPED;
    opt.retain_domains = 1;
    std::vector<tfsm_pattern_result> res(tps.size());
    const int rc = tfsm_eval_level(G.ctx, tps.data(), par.data(), (uint32_t)tps.size(), (uint32_t)freq, &opt,
                                   res.data());
    if (rc != TFSM_OK) {
        G.error = tfsm_last_error(G.ctx);
        return rc;
    }
    for (size_t k = 0; k < where.size(); ++k) {
        frequent[where[k]] = res[k].frequent != 0;
        child_handles[where[k]] = res[k].domain_handle;
    }
    return TFSM_OK;
}

#endif
