// gSpan growth and the minimum-DFS-code test.  Behaviour cited in pattern.h.
#include "pattern.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <stdexcept>

namespace tfsm {

void Pattern::push(const DfsCode &c) {
    if (code.empty()) {
        vlab = {c.fl, c.tl};
        adj.assign(2, {});
    } else if (c.to >= nv()) {
        if (c.to != nv()) throw std::invalid_argument("forward edge must add vertex nv()");
        vlab.push_back(c.tl);
        adj.emplace_back();
    }
    if (nv() > kMaxPatternVertices || ne() + 1 > kMaxPatternEdges)
        throw std::length_error("pattern exceeds the supported size");
    auto ins = [&](int a, int b, int32_t l) {
        auto &v = adj[a];
        auto it = std::upper_bound(v.begin(), v.end(), b,
                                   [](int x, const std::pair<int, int32_t> &e) { return x < e.first; });
        v.insert(it, {b, l});
    };
    ins(c.from, c.to, c.el);
    ins(c.to, c.from, c.el);
    code.push_back(c);
    // rightmost path: forward edges chained back from the last forward edge
    rmpath.clear();
    int prev = -1;
    for (int i = ne() - 1; i >= 0; --i) {
        if (code[i].forward() && (rmpath.empty() || prev == code[i].to)) {
            prev = code[i].from;
            rmpath.push_back(i);
        }
    }
}

bool Pattern::adjacent(int a, int b) const {
    for (auto &e : adj[a]) if (e.first == b) return true;
    return false;
}

int32_t Pattern::edge_label(int a, int b) const {
    for (auto &e : adj[a]) if (e.first == b) return e.second;
    throw std::invalid_argument("no such pattern edge");
}

bool Pattern::distinct_labels() const {
    std::vector<int32_t> l = vlab;
    std::sort(l.begin(), l.end());
    return std::adjacent_find(l.begin(), l.end()) == l.end();
}

std::string Pattern::key() const {
    std::string s;
    for (auto &c : code) {
        s += std::to_string(c.from) + "," + std::to_string(c.to) + "," + std::to_string(c.fl) + "," +
             std::to_string(c.el) + "," + std::to_string(c.tl) + ";";
    }
    return s;
}

std::string Pattern::to_lg_block() const {
    std::string s;
    for (int i = 0; i < nv(); ++i) s += "v " + std::to_string(i) + " " + std::to_string(vlab[i]) + "\n";
    for (auto &c : code)
        s += "e " + std::to_string(c.from) + " " + std::to_string(c.to) + " " + std::to_string(c.el) + "\n";
    return s;
}

Pattern single_edge(int32_t la, int32_t el, int32_t lb) {
    Pattern p;
    p.push(DfsCode{0, 1, la, el, lb});
    return p;
}

namespace {

// One embedding of the growing minimum code into the pattern's own graph.
struct Proj {
    int8_t vmap[kMaxPatternVertices];   // DFS id -> pattern vertex
    uint64_t eused;                      // pattern edge ids already matched
    uint32_t vused;                      // pattern vertices already matched
};

struct SelfGraph {
    const Pattern &p;
    int eid[kMaxPatternVertices][kMaxPatternVertices];
    explicit SelfGraph(const Pattern &pp) : p(pp) {
        for (auto &r : eid) for (int &x : r) x = -1;
        for (int i = 0; i < p.ne(); ++i) {
            eid[p.code[i].from][p.code[i].to] = i;
            eid[p.code[i].to][p.code[i].from] = i;
        }
    }
};

std::vector<int> rightmost_path(const std::vector<DfsCode> &mc) {
    std::vector<int> rm;
    int prev = -1;
    for (int i = (int)mc.size() - 1; i >= 0; --i) {
        if (mc[i].forward() && (rm.empty() || prev == mc[i].to)) {
            prev = mc[i].from;
            rm.push_back(i);
        }
    }
    return rm;
}

}  // namespace

// Builds the minimum DFS code edge by edge over all self-embeddings of the
// prefix, and stops at the first position where the pattern's code differs.
// Next-edge preference, as in gSpan: backward edges from the rightmost vertex
// (shallowest target first, then smaller edge label), then forward edges from
// the rightmost vertex, then forward edges from rightmost-path vertices,
// deepest first; forward keys compare (edge label, target label).  The label
// side conditions are those of canonical.h:157-253.  Like grami.h:840-872 the
// first edge is taken as the minimum (from label, edge label, to label) triple
// and the comparison starts at the second edge.
bool is_min_code(const Pattern &p) {
    if (p.ne() <= 1) return true;
    SelfGraph g(p);
    const int n = p.nv();

    // first edge: minimum (la, el, lb) over oriented edges with la <= lb
    int32_t b0 = INT_MAX, b1 = INT_MAX, b2 = INT_MAX;
    for (int a = 0; a < n; ++a)
        for (auto &e : p.adj[a]) {
            int32_t la = p.vlab[a], lb = p.vlab[e.first];
            if (la > lb) continue;
            if (std::make_tuple(la, e.second, lb) < std::make_tuple(b0, b1, b2)) { b0 = la; b1 = e.second; b2 = lb; }
        }
    std::vector<DfsCode> mc{DfsCode{0, 1, b0, b1, b2}};
    std::vector<Proj> projs;
    for (int a = 0; a < n; ++a)
        for (auto &e : p.adj[a]) {
            if (p.vlab[a] == b0 && e.second == b1 && p.vlab[e.first] == b2) {
                Proj pr;
                std::memset(pr.vmap, -1, sizeof(pr.vmap));
                pr.vmap[0] = (int8_t)a; pr.vmap[1] = (int8_t)e.first;
                pr.eused = 1ull << g.eid[a][e.first];
                pr.vused = (1u << a) | (1u << e.first);
                projs.push_back(pr);
            }
        }

    std::vector<Proj> next;
    for (int k = 1; k < p.ne(); ++k) {
        std::vector<int> rm = rightmost_path(mc);
        const int maxtoc = mc[rm[0]].to;
        const int32_t minlabel = mc[0].fl;
        bool found = false;
        DfsCode nc;

        // backward edges from the rightmost vertex
        for (int i = (int)rm.size() - 1; i >= 1 && !found; --i) {
            const int t = mc[rm[i]].from;
            int32_t best = INT_MAX;
            for (auto &pr : projs) {
                int gr = pr.vmap[maxtoc], gt = pr.vmap[t];
                int id = g.eid[gr][gt];
                if (id < 0 || (pr.eused >> id & 1)) continue;
                int32_t le = p.edge_label(gr, gt);
                int32_t e1l = p.edge_label(pr.vmap[mc[rm[i]].from], pr.vmap[mc[rm[i]].to]);
                int32_t e1to = p.vlab[pr.vmap[mc[rm[i]].to]];
                if (e1l < le || (e1l == le && e1to <= p.vlab[gr])) best = std::min(best, le);
            }
            if (best == INT_MAX) continue;
            found = true;
            nc = DfsCode{maxtoc, t, mc[rm[0]].tl, best, mc[rm[i]].fl};
            if (nc != p.code[k]) return false;
            next.clear();
            for (auto &pr : projs) {
                int gr = pr.vmap[maxtoc], gt = pr.vmap[t];
                int id = g.eid[gr][gt];
                if (id < 0 || (pr.eused >> id & 1)) continue;
                int32_t le = p.edge_label(gr, gt);
                if (le != best) continue;
                int32_t e1l = p.edge_label(pr.vmap[mc[rm[i]].from], pr.vmap[mc[rm[i]].to]);
                int32_t e1to = p.vlab[pr.vmap[mc[rm[i]].to]];
                if (!(e1l < le || (e1l == le && e1to <= p.vlab[gr]))) continue;
                Proj q = pr;
                q.eused |= 1ull << id;
                next.push_back(q);
            }
        }

        // forward edges from the rightmost vertex
        if (!found) {
            std::pair<int32_t, int32_t> best{INT_MAX, INT_MAX};
            for (auto &pr : projs) {
                int gr = pr.vmap[maxtoc];
                for (auto &e : p.adj[gr]) {
                    int w = e.first;
                    if ((pr.vused >> w & 1) || p.vlab[w] < minlabel) continue;
                    best = std::min(best, std::make_pair(e.second, p.vlab[w]));
                }
            }
            if (best.first != INT_MAX) {
                found = true;
                nc = DfsCode{maxtoc, maxtoc + 1, p.vlab[projs[0].vmap[maxtoc]], best.first, best.second};
                if (nc != p.code[k]) return false;
                next.clear();
                for (auto &pr : projs) {
                    int gr = pr.vmap[maxtoc];
                    for (auto &e : p.adj[gr]) {
                        int w = e.first;
                        if ((pr.vused >> w & 1) || p.vlab[w] < minlabel) continue;
                        if (e.second != best.first || p.vlab[w] != best.second) continue;
                        Proj q = pr;
                        q.vmap[maxtoc + 1] = (int8_t)w;
                        q.eused |= 1ull << g.eid[gr][w];
                        q.vused |= 1u << w;
                        next.push_back(q);
                    }
                }
            }
        }

        // forward edges from the rightmost path, deepest source first
        for (int i = 0; i < (int)rm.size() && !found; ++i) {
            const int f = mc[rm[i]].from, to = mc[rm[i]].to;
            std::pair<int32_t, int32_t> best{INT_MAX, INT_MAX};
            auto ok = [&](const Proj &pr, int w, int32_t le) {
                int gf = pr.vmap[f], gto = pr.vmap[to];
                if (w == gto || p.vlab[w] < minlabel || (pr.vused >> w & 1)) return false;
                int32_t el0 = p.edge_label(gf, gto);
                int32_t tol = p.vlab[gto];
                return el0 < le || (el0 == le && tol <= p.vlab[w]);
            };
            for (auto &pr : projs)
                for (auto &e : p.adj[pr.vmap[f]])
                    if (ok(pr, e.first, e.second)) best = std::min(best, std::make_pair(e.second, p.vlab[e.first]));
            if (best.first == INT_MAX) continue;
            found = true;
            nc = DfsCode{f, maxtoc + 1, mc[rm[i]].fl, best.first, best.second};
            if (nc != p.code[k]) return false;
            next.clear();
            for (auto &pr : projs) {
                int gf = pr.vmap[f];
                for (auto &e : p.adj[gf]) {
                    int w = e.first;
                    if (!ok(pr, w, e.second)) continue;
                    if (e.second != best.first || p.vlab[w] != best.second) continue;
                    Proj q = pr;
                    q.vmap[maxtoc + 1] = (int8_t)w;
                    q.eused |= 1ull << g.eid[gf][w];
                    q.vused |= 1u << w;
                    next.push_back(q);
                }
            }
        }

        if (!found) return true;   // nothing extends the prefix (canonical.h:344)
        mc.push_back(nc);
        projs.swap(next);
    }
    return true;
}

std::vector<Pattern> extend(const Pattern &p, const GrowthContext &ctx) {
    std::vector<Pattern> out;
    const int last = p.nv() - 1;
    const int32_t last_label = p.vlab[last];
    const int32_t min_label = p.code[0].fl;
    const bool may_grow = ctx.max_nodes == -1 || p.nv() < ctx.max_nodes;
    static const std::set<int32_t> kEmpty;
    auto nbr_labels = [&](int32_t l) -> const std::set<int32_t> & {
        auto it = ctx.neighbor_labels.find(l);
        return it == ctx.neighbor_labels.end() ? kEmpty : it->second;
    };
    auto emit = [&](const DfsCode &c) {
        Pattern q = p;
        q.push(c);
        if (is_min_code(q)) out.push_back(std::move(q));
    };

    // forward from the rightmost vertex (grami.h:182-210)
    if (may_grow) {
        for (int32_t nl : nbr_labels(last_label))
            for (int32_t el : ctx.freq_edge_labels)
                if (nl >= min_label) emit(DfsCode{last, last + 1, last_label, el, nl});
    }
    // backward from the rightmost vertex to rightmost-path vertices (grami.h:212-253)
    for (int i = (int)p.rmpath.size(); i > 1; --i) {
        const DfsCode &rc = p.code[p.rmpath[i - 1]];
        const int cur = rc.from;
        const int32_t cur_label = p.vlab[cur];
        if (!nbr_labels(last_label).count(cur_label)) continue;
        if (p.adjacent(last, cur)) continue;
        for (int32_t el : ctx.freq_edge_labels)
            if (el > rc.el || (el == rc.el && last_label >= rc.tl))
                emit(DfsCode{last, cur, last_label, el, cur_label});
    }
    // forward from rightmost-path vertices (grami.h:255-293)
    if (may_grow) {
        for (int i = 0; i < (int)p.rmpath.size(); ++i) {
            const DfsCode &rc = p.code[p.rmpath[i]];
            const int cur = rc.from;
            const int32_t cur_label = p.vlab[cur];
            for (int32_t nl : nbr_labels(cur_label))
                for (int32_t el : ctx.freq_edge_labels)
                    if (rc.el < el || (rc.el == el && rc.tl <= nl))
                        emit(DfsCode{cur, last + 1, cur_label, el, nl});
        }
    }
    return out;
}

}  // namespace tfsm
