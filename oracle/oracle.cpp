// CPU oracle for the T-FSM MNI support-evaluation path.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  A plain, deliberately direct
// restatement of the reference's algorithm, written for checking rather than
// speed.  Reference citations are to lyuheng/T-FSM, parallel/ build.
#include "oracle.h"

#include "../t-fsm_amd/host/pattern.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

using std::vector;

struct orc_graph {
    int freq = 0;
    vector<int32_t> vlabel;                                    // original label per pruned vertex
    vector<vector<std::pair<uint32_t, int32_t>>> adj;          // (to, edge label) ascending `to`
    std::map<int32_t, vector<uint32_t>> nodes_by_label;
    vector<std::unordered_map<int32_t, int>> nlf;
    std::set<int32_t> freq_edge_labels;
    std::map<std::tuple<int32_t, int32_t, int32_t>, std::pair<std::set<uint32_t>, std::set<uint32_t>>> hashed_edges;
    vector<std::map<int32_t, double>> vtx_frac;                // fraction-score/graph.h:895-922
    uint32_t n() const { return (uint32_t)vlabel.size(); }
    double frac(uint32_t v, int32_t l) const {                // operator[] of the reference: 0 when absent
        auto it = vtx_frac[v].find(l);
        return it == vtx_frac[v].end() ? 0.0 : it->second;
    }
};

namespace {

// Correctly rounded sum of doubles: Shewchuk's exact partials ("Adaptive
// Precision Floating-Point Arithmetic and Fast Robust Geometric Predicates",
// 1997, GROW-EXPANSION with zero elimination), rounded once to nearest with
// ties to even (the round-half-even step of Python's math.fsum).  The
// Fraction-Score support of this build is the double nearest the exact sum.
struct ExactSum {
    vector<double> p;                                          // non-overlapping, increasing magnitude
    void add(double x) {
        size_t k = 0;
        for (double y : p) {
            if (std::fabs(x) < std::fabs(y)) std::swap(x, y);
            const double hi = x + y, lo = y - (hi - x);
            if (lo != 0.0) p[k++] = lo;
            x = hi;
        }
        p.resize(k);
        p.push_back(x);
    }
    double rounded() const {
        if (p.empty()) return 0.0;
        size_t n = p.size();
        double hi = p[--n], lo = 0.0;
        while (n > 0) {
            const double x = hi, y = p[--n];
            hi = x + y;
            lo = y - (hi - x);
            if (lo != 0.0) break;
        }
        // hi + lo is exact; a tie (lo exactly half an ulp) with more of the
        // same sign below it rounds away from hi
        if (n > 0 && ((lo < 0 && p[n - 1] < 0) || (lo > 0 && p[n - 1] > 0))) {
            const double y = lo * 2, x = hi + y;
            if (y == x - hi) hi = x;
        }
        return hi;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// Loading: read_input / construct_graph / find_frequent_labels /
// construct_freq_graph / build_nlf, parallel/graph.h:557-803, 850-866.
// ---------------------------------------------------------------------------
extern "C" orc_graph *orc_load(const char *path, int freq) {
    std::ifstream fin(path);
    if (!fin.is_open()) return nullptr;
    vector<vector<std::string>> input;                         // graph.h:572-601
    std::string line;
    uint32_t vertex_num = 0;
    while (std::getline(fin, line)) {
        std::istringstream ss(line);
        vector<std::string> tok;
        std::string t;
        while (ss >> t) {
            if (t[0] == 'v') vertex_num++;
            tok.push_back(t);
        }
        if (!tok.empty()) input.push_back(tok);
    }
    auto *g = new orc_graph;
    g->freq = freq;
    // full-graph label counts (graph.h:603-683)
    std::map<int32_t, int> vcount;
    std::map<int32_t, int> ecount2;                           // both directions, as edgesByLabel
    for (auto &tok : input) {
        if (tok[0] == "v") vcount[std::stoi(tok[2])]++;
        else if (tok[0] == "e") ecount2[std::stoi(tok[3])] += 2;
    }
    std::set<int32_t> freq_node_labels;
    for (auto &kv : vcount) if (kv.second >= freq) freq_node_labels.insert(kv.first);
    for (auto &kv : ecount2) if (kv.second / 2 >= freq) g->freq_edge_labels.insert(kv.first);

    // construct_freq_graph (graph.h:685-803)
    vector<int32_t> labels;                                    // by v-line order
    std::unordered_map<uint32_t, uint32_t> id_map;
    uint32_t vertex_id = 0;
    for (auto &tok : input) {
        if (tok[0] == "t") continue;
        if (tok[0] == "v") {
            uint32_t id = (uint32_t)std::stoul(tok[1]);
            int32_t label = std::stoi(tok[2]);
            labels.push_back(label);
            if (freq_node_labels.count(label)) {
                g->vlabel.push_back(label);
                g->adj.emplace_back();
                id_map[id] = vertex_id++;
            }
        } else if (tok[0] == "e") {
            uint32_t from = (uint32_t)std::stoul(tok[1]), to = (uint32_t)std::stoul(tok[2]);
            int32_t label = std::stoi(tok[3]);
            int32_t lf = labels[from], lt = labels[to];
            if (freq_node_labels.count(lf) && freq_node_labels.count(lt) && g->freq_edge_labels.count(label)) {
                g->adj[id_map[from]].push_back({id_map[to], label});
                g->adj[id_map[to]].push_back({id_map[from], label});
            }
            uint32_t nf = id_map[from], nt = id_map[to];      // operator[]: 0 for dropped vertices
            if (lf < lt) {
                auto &e = g->hashed_edges[std::make_tuple(lf, label, lt)];
                e.first.insert(nf); e.second.insert(nt);
            } else if (lf == lt) {
                auto &e = g->hashed_edges[std::make_tuple(lf, label, lt)];
                e.first.insert(nf); e.first.insert(nt); e.second.insert(nt); e.second.insert(nf);
            } else {
                auto &e = g->hashed_edges[std::make_tuple(lt, label, lf)];
                e.first.insert(nt); e.second.insert(nf);
            }
        }
    }
    for (auto &a : g->adj)
        std::stable_sort(a.begin(), a.end(), [](const std::pair<uint32_t, int32_t> &x, const std::pair<uint32_t, int32_t> &y) {
            return x.first < y.first;
        });
    for (uint32_t i = 0; i < g->n(); ++i) g->nodes_by_label[g->vlabel[i]].push_back(i);
    for (auto it = g->hashed_edges.begin(); it != g->hashed_edges.end();) {
        if ((int)it->second.first.size() < freq || (int)it->second.second.size() < freq) it = g->hashed_edges.erase(it);
        else ++it;
    }
    g->nlf.resize(g->n());                                     // build_nlf, graph.h:850-866
    for (uint32_t i = 0; i < g->n(); ++i)
        for (auto &e : g->adj[i]) g->nlf[i][g->vlabel[e.first]] += 1;
    // compute_fraction_score, fraction-score/graph.h:895-922: every edge
    // entry (i, nbr) adds 1 / NLF(i, label(nbr)) to nbr's score for label(i),
    // capped at 1 after each add, in the order i ascending, nbr ascending
    g->vtx_frac.resize(g->n());
    for (uint32_t i = 0; i < g->n(); ++i)
        for (auto &e : g->adj[i]) {
            double &s = g->vtx_frac[e.first][g->vlabel[i]];
            s += 1 / (double)g->nlf[i][g->vlabel[e.first]];
            if (s > 1) s = 1.0;
        }
    return g;
}

extern "C" int64_t orc_fraction_scores(const orc_graph *g, double *out, uint64_t cap) {
    const uint64_t L = g->nodes_by_label.size();
    uint64_t k = 0;
    for (uint32_t v = 0; v < g->n(); ++v)
        for (auto &kv : g->nodes_by_label) {
            if (out && k < cap) out[k] = g->frac(v, kv.first);
            ++k;
        }
    return (int64_t)(L * g->n());
}

// The size-2 Fraction-Score filter, fraction-score/graph.h:803-824: over
// each surviving endpoint set, sum_a = sum of score(v, to label) over candA,
// sum_b = sum of score(v, from label) over candB.
extern "C" int orc_edge_fraction(const orc_graph *g, double *out, int cap) {
    int k = 0;
    for (auto &kv : g->hashed_edges) {
        if (k < cap) {
            ExactSum a, b;
            for (uint32_t v : kv.second.first) a.add(g->frac(v, std::get<2>(kv.first)));
            for (uint32_t v : kv.second.second) b.add(g->frac(v, std::get<0>(kv.first)));
            out[2 * k] = a.rounded();
            out[2 * k + 1] = b.rounded();
        }
        ++k;
    }
    return k;
}

extern "C" double orc_exact_sum(const double *x, uint64_t n) {
    ExactSum s;
    for (uint64_t i = 0; i < n; ++i) s.add(x[i]);
    return s.rounded();
}

extern "C" void orc_graph_free(orc_graph *g) { delete g; }
extern "C" uint32_t orc_num_vertices(const orc_graph *g) { return g->n(); }
extern "C" uint64_t orc_num_adj(const orc_graph *g) {
    uint64_t s = 0;
    for (auto &a : g->adj) s += a.size();
    return s;
}
extern "C" void orc_csr(const orc_graph *g, uint32_t *offsets, uint32_t *nbrs, int32_t *elabels, int32_t *vlabels) {
    uint64_t k = 0;
    for (uint32_t v = 0; v < g->n(); ++v) {
        offsets[v] = (uint32_t)k;
        vlabels[v] = g->vlabel[v];
        for (auto &e : g->adj[v]) { nbrs[k] = e.first; elabels[k] = e.second; ++k; }
    }
    offsets[g->n()] = (uint32_t)k;
}
extern "C" int orc_edge_stats(const orc_graph *g, int32_t *out, int cap) {
    int k = 0;
    for (auto &kv : g->hashed_edges) {
        if (k < cap) {
            out[5 * k + 0] = std::get<0>(kv.first);
            out[5 * k + 1] = std::get<1>(kv.first);
            out[5 * k + 2] = std::get<2>(kv.first);
            out[5 * k + 3] = (int32_t)kv.second.first.size();
            out[5 * k + 4] = (int32_t)kv.second.second.size();
        }
        ++k;
    }
    return k;
}

// ---------------------------------------------------------------------------
// Pattern evaluation.
// ---------------------------------------------------------------------------
namespace {

const uint32_t kInvalid = 1000000000u;                         // INVALID_VERTEX_ID, gmatch.h:12

struct PGraph {                                                // query graph as the engine sees it
    int nv = 0;
    vector<int32_t> vlab;
    vector<vector<std::pair<int, int32_t>>> adj;               // ascending neighbour (Pattern::insert)
    int deg(int u) const { return (int)adj[u].size(); }
    int32_t elabel(int a, int b) const {
        for (auto &e : adj[a]) if (e.first == b) return e.second;
        return -1;
    }
};

struct Aux {                                                   // Edges, types.h:67-90
    bool present = false;
    vector<uint32_t> off, idx;
};

}  // namespace

struct orc_eval {
    int nv = 0;
    vector<vector<uint32_t>> C;
    vector<vector<uint8_t>> state;
    vector<vector<Aux>> aux;
    vector<vector<uint32_t>> order;
    vector<vector<vector<int>>> bn;
    int frequent = 0, decided_by = 0;
    uint32_t lower = 0, upper = 0;
    uint64_t searches = 0, verdicts = 0;
    vector<double> fraction;                                   // mode 2: per u, summed scores of valid candidates
};

namespace {

// LDF_pruning, gmatch.h:301-415 (no push-down non-candidates).
bool ldf(const orc_graph &g, const PGraph &q, int parent_nv, const uint32_t *parent_off, const uint32_t *parent_dom,
         vector<vector<uint32_t>> &C) {
    C.assign(q.nv, {});
    for (int i = 0; i < q.nv; ++i) {
        std::unordered_map<int32_t, int> qnlf;
        for (auto &e : q.adj[i]) qnlf[q.vlab[e.first]] += 1;
        const int degree = q.deg(i);
        auto test = [&](uint32_t v) {
            if ((int)g.adj[v].size() < degree) return false;
            for (auto &kv : qnlf) {
                auto it = g.nlf[v].find(kv.first);
                if (it == g.nlf[v].end() || it->second < kv.second) return false;
            }
            return true;
        };
        if (i < parent_nv) {
            for (uint32_t k = parent_off[i]; k < parent_off[i + 1]; ++k)
                if (test(parent_dom[k])) C[i].push_back(parent_dom[k]);
        } else {
            auto it = g.nodes_by_label.find(q.vlab[i]);
            if (it != g.nodes_by_label.end())
                for (uint32_t v : it->second) if (test(v)) C[i].push_back(v);
        }
        if (C[i].empty()) return false;
    }
    return true;
}

// pruneCandidates, gmatch.h:601-658.
bool prune(const orc_graph &g, const PGraph &q, vector<vector<uint32_t>> &C, int u, const vector<int> &pivots,
           vector<uint32_t> &flag) {
    bool shrunk = false;
    const int32_t ql = q.vlab[u];
    const int qd = q.deg(u);
    uint32_t count = 0;
    vector<uint32_t> updated;
    for (int p : pivots) {
        const int32_t el = q.elabel(u, p);
        for (uint32_t v : C[p]) {
            if (v == kInvalid) continue;
            for (auto &e : g.adj[v]) {
                uint32_t w = e.first;
                if (flag[w] == count && el == e.second && g.vlabel[w] == ql && (int)g.adj[w].size() >= qd) {
                    flag[w] += 1;
                    if (count == 0) updated.push_back(w);
                }
            }
        }
        count += 1;
    }
    for (auto &v : C[u]) {
        if (v == kInvalid) continue;
        if (flag[v] != count) { v = kInvalid; shrunk = true; }
    }
    for (uint32_t w : updated) flag[w] = 0;
    return shrunk;
}

// DPisoFilter / filterToConsistency, gmatch.h:417-509, with the plan of
// generateDPisoFilterPlan / selectDPisoStartVertex / bfsTraversal, 511-599,
// compactCandidates 660-674.
void dpiso(const orc_graph &g, const PGraph &q, vector<vector<uint32_t>> &C) {
    int start = 0;
    double min_score = (double)g.n();
    for (int i = 0; i < q.nv; ++i) {
        if (q.deg(i) <= 1) continue;
        double s = C[i].size() / (double)q.deg(i);
        if (s < min_score) { min_score = s; start = i; }
    }
    vector<int> bfs;
    vector<bool> seen(q.nv, false);
    bfs.push_back(start); seen[start] = true;
    for (size_t h = 0; h < bfs.size(); ++h)
        for (auto &e : q.adj[bfs[h]])
            if (!seen[e.first]) { seen[e.first] = true; bfs.push_back(e.first); }
    vector<int> pos(q.nv);
    for (int i = 0; i < q.nv; ++i) pos[bfs[i]] = i;
    vector<vector<int>> bnl(q.nv), fnl(q.nv);
    for (int u = 0; u < q.nv; ++u)
        for (auto &e : q.adj[u]) (pos[e.first] < pos[u] ? bnl[u] : fnl[u]).push_back(e.first);
    vector<uint32_t> flag(g.n(), 0);
    bool shrunk = true;
    while (shrunk) {
        for (int i = 1; i < q.nv; ++i) shrunk = prune(g, q, C, bfs[i], bnl[bfs[i]], flag);
        for (int i = q.nv - 2; i >= 0; --i) shrunk = prune(g, q, C, bfs[i], fnl[bfs[i]], flag) | shrunk;
    }
    for (auto &c : C) c.erase(std::remove(c.begin(), c.end(), kInvalid), c.end());
}

// buildTable(Edges***&), gmatch.h:995-1130: aux[x][y] has rows over C(x)
// listing ascending indices into C(y).
void build_table(const orc_graph &g, const PGraph &q, const vector<vector<uint32_t>> &C, vector<vector<Aux>> &aux) {
    aux.assign(q.nv, vector<Aux>(q.nv));
    vector<int> order(q.nv);
    for (int i = 0; i < q.nv; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int l, int r) {
        if (q.deg(l) == q.deg(r)) return l < r;
        return q.deg(l) > q.deg(r);
    });
    vector<uint32_t> flag(g.n(), 0);
    for (int u : order) {
        bool recorded = false;
        for (auto &qe : q.adj[u]) {
            int un = qe.first;
            if (aux[u][un].present) continue;
            const int32_t el = q.elabel(un, u);
            if (!recorded) {
                for (uint32_t j = 0; j < C[u].size(); ++j) flag[C[u][j]] = j + 1;
                recorded = true;
            }
            Aux &fw = aux[un][u], &bw = aux[u][un];
            fw.present = bw.present = true;
            fw.off.assign(C[un].size() + 1, 0);
            bw.off.assign(C[u].size() + 1, 0);
            for (uint32_t j = 0; j < C[un].size(); ++j) {
                fw.off[j] = (uint32_t)fw.idx.size();
                uint32_t v = C[un][j];
                for (auto &e : g.adj[v]) {
                    if (flag[e.first] != 0 && e.second == el) {
                        uint32_t p = flag[e.first] - 1;
                        fw.idx.push_back(p);
                        bw.off[p + 1] += 1;
                    }
                }
            }
            fw.off[C[un].size()] = (uint32_t)fw.idx.size();
            for (size_t j = 1; j < bw.off.size(); ++j) bw.off[j] += bw.off[j - 1];
            bw.idx.assign(fw.idx.size(), 0);
            vector<uint32_t> cur(bw.off.begin(), bw.off.end() - 1);
            for (uint32_t j = 0; j < C[un].size(); ++j)
                for (uint32_t k = fw.off[j]; k < fw.off[j + 1]; ++k) bw.idx[cur[fw.idx[k]]++] = j;
        }
        if (recorded) for (uint32_t v : C[u]) flag[v] = 0;
    }
}

// generateGQLQueryPlan(int, ui*&) + generateBN(...), gmatch.h:692-727, 803-828.
void plan(const orc_graph &g, const PGraph &q, const vector<vector<uint32_t>> &C, int start, vector<uint32_t> &order,
          vector<vector<int>> &bn) {
    vector<bool> visited(q.nv, false), adjacent(q.nv, false);
    order.assign(q.nv, 0);
    auto update = [&](int u) {
        visited[u] = true;
        for (auto &e : q.adj[u]) adjacent[e.first] = true;
    };
    order[0] = start;
    update(start);
    for (int i = 1; i < q.nv; ++i) {
        int next = 0;
        size_t min_value = (size_t)g.n() + 1;
        for (int j = 0; j < q.nv; ++j) {
            if (visited[j] || !adjacent[j]) continue;
            if (C[j].size() < min_value) { min_value = C[j].size(); next = j; }
            else if (C[j].size() == min_value && q.deg(j) > q.deg(next)) next = j;
        }
        update(next);
        order[i] = next;
    }
    bn.assign(q.nv, {});
    vector<bool> vis(q.nv, false);
    vis[order[0]] = true;
    for (int i = 1; i < q.nv; ++i) {
        for (auto &e : q.adj[order[i]]) if (vis[e.first]) bn[i].push_back(e.first);
        vis[order[i]] = true;
    }
}

// LFTJ (gmatch.h:1198-1280) without the timeout, with the candidate sets of
// generateValidCandidateIndex (gmatch.h:1508-1546): the bn[0] row intersected
// with every other backward-neighbour row.  Returns true with emb_idx filled.
bool exists(const PGraph &q, const vector<vector<uint32_t>> &C, const vector<vector<Aux>> &aux,
            const vector<uint32_t> &order, const vector<vector<int>> &bn, uint32_t idx0, vector<uint32_t> &emb_idx) {
    const int n = q.nv;
    vector<vector<uint32_t>> cand(n);
    vector<size_t> counter(n, 0);
    vector<uint32_t> emb_v(n, 0);
    emb_idx.assign(n, 0);
    emb_idx[order[0]] = idx0;
    emb_v[order[0]] = C[order[0]][idx0];
    auto gen = [&](int d) {
        const int u = order[d];
        const Aux &a0 = aux[bn[d][0]][u];
        uint32_t r = emb_idx[bn[d][0]];
        vector<uint32_t> cur(a0.idx.begin() + a0.off[r], a0.idx.begin() + a0.off[r + 1]);
        for (size_t i = 1; i < bn[d].size(); ++i) {
            const Aux &a = aux[bn[d][i]][u];
            uint32_t rr = emb_idx[bn[d][i]];
            vector<uint32_t> out;
            std::set_intersection(cur.begin(), cur.end(), a.idx.begin() + a.off[rr], a.idx.begin() + a.off[rr + 1],
                                  std::back_inserter(out));
            cur.swap(out);
        }
        cand[d].swap(cur);
        counter[d] = 0;
    };
    if (n == 1) return true;
    int d = 1;
    gen(1);
    while (true) {
        while (counter[d] < cand[d].size()) {
            uint32_t c = cand[d][counter[d]++];
            const int u = order[d];
            uint32_t v = C[u][c];
            bool used = false;
            for (int k = 0; k < d; ++k) if (emb_v[order[k]] == v) { used = true; break; }
            if (used) continue;
            emb_idx[u] = c;
            emb_v[u] = v;
            if (d == n - 1) return true;
            ++d;
            gen(d);
        }
        --d;
        if (d < 1) return false;
    }
}

PGraph make_query(int nv, const int32_t *vlab, int ne, const int32_t *edges) {
    PGraph q;
    q.nv = nv;
    q.vlab.assign(vlab, vlab + nv);
    q.adj.assign(nv, {});
    for (int i = 0; i < ne; ++i) {
        int a = edges[3 * i], b = edges[3 * i + 1];
        int32_t l = edges[3 * i + 2];
        q.adj[a].push_back({b, l});
        q.adj[b].push_back({a, l});
    }
    for (auto &a : q.adj) std::sort(a.begin(), a.end());
    return q;
}

}  // namespace

extern "C" orc_eval *orc_eval_pattern(const orc_graph *gp, int nv, const int32_t *vlab, int ne, const int32_t *edges,
                                      int parent_nv, const uint32_t *parent_off, const uint32_t *parent_dom, int freq,
                                      int mode) {
    const orc_graph &g = *gp;
    PGraph q = make_query(nv, vlab, ne, edges);
    auto *r = new orc_eval;
    r->nv = nv;
    // activation, comper.h:1614-1700
    if (!ldf(g, q, parent_nv, parent_off, parent_dom, r->C)) {
        // an empty candidate set empties every domain of a connected pattern
        // under refinement; the reference stops here (gmatch.h:387-391)
        r->C.assign(nv, {});
        r->state.resize(nv);
        for (int u = 0; u < nv; ++u) r->state[u].assign(r->C[u].size(), 0);
        r->decided_by = 0;
        return r;
    }
    dpiso(g, q, r->C);
    r->state.resize(nv);
    for (int u = 0; u < nv; ++u) r->state[u].assign(r->C[u].size(), 0);
    uint32_t min_size = ~0u;
    for (auto &c : r->C) min_size = std::min<uint32_t>(min_size, (uint32_t)c.size());
    r->upper = min_size;
    bool distinct = true;
    for (int a = 0; a < nv; ++a)
        for (int b = a + 1; b < nv; ++b) if (q.vlab[a] == q.vlab[b]) distinct = false;
    // Fraction-Score term of candidate v of u (fraction-score/comper.h:486-499):
    // min over u's pattern neighbours of v's score for the neighbour's label
    auto frac_agg = [&](int u, uint32_t v) {
        double f = std::numeric_limits<double>::max();
        for (auto &e : q.adj[u]) f = std::min(f, g.frac(v, q.vlab[e.first]));
        return f;
    };
    r->fraction.assign(nv, 0.0);
    if (distinct && ne == nv - 1) {
        // express path, comper.h:1619-1653: arc-consistent domains are exact
        r->decided_by = 1;
        r->lower = r->upper = min_size;
        r->frequent = (int)min_size >= freq;
        for (auto &s : r->state) std::fill(s.begin(), s.end(), 1);
        for (auto &c : r->C) r->verdicts += c.size();
        if (mode == 2) {
            // fraction_score_is_freq, fraction-score/comper.h:1082-1106
            double mn = std::numeric_limits<double>::max();
            for (int u = 0; u < nv; ++u) {
                ExactSum s;
                for (uint32_t v : r->C[u]) s.add(frac_agg(u, v));
                r->fraction[u] = s.rounded();
                mn = std::min(mn, r->fraction[u]);
            }
            r->frequent = r->frequent && mn >= (double)freq;
        }
        return r;
    }
    if ((int)min_size < freq) { r->decided_by = 0; return r; }          // isCandidateSetValid(freq)
    r->decided_by = 2;
    build_table(g, q, r->C, r->aux);
    r->order.resize(nv);
    r->bn.resize(nv);
    for (int u = 0; u < nv; ++u) plan(g, q, r->C, u, r->order[u], r->bn[u]);

    // task processing: refill_Q order (u ascending, domain order), compute,
    // refill_check and postprocess, comper.h:431-533, 1181-1300.  Mode 2
    // (the fraction-score build) keeps u's summed score beside its valid
    // count: a vertex stops when the sum reaches freq (comper.h:478-508) and
    // the pattern fails when the sum plus the unfinished candidates cannot
    // (comper.h:299-301)
    const bool early = mode == 0 || mode == 2;
    vector<ExactSum> vsup(nv);
    vector<uint32_t> nvalid(nv, 0), ninvalid(nv, 0), finished(nv, 0);
    vector<bool> stops(nv, false);
    int done = 0;
    bool decided = false, frequent = false;
    vector<uint32_t> emb;
    for (int u = 0; u < nv && !decided; ++u) {
        for (uint32_t i = 0; i < r->C[u].size() && !decided; ++i) {
            if (early && stops[u]) break;
            if (r->state[u][i] == 1) {                              // already valid through propagation
                if (mode == 1) { r->verdicts++; continue; }
            }
            r->searches++;
            if (exists(q, r->C, r->aux, r->order[u], r->bn[u], i, emb)) {
                for (int j = 0; j < nv; ++j) {
                    if (r->state[j][emb[j]] != 1) {
                        r->state[j][emb[j]] = 1;
                        nvalid[j]++;
                        if (mode == 2) vsup[j].add(frac_agg(j, r->C[j][emb[j]]));
                        const bool reached = mode == 2 ? vsup[j].rounded() >= (double)freq : (int)nvalid[j] >= freq;
                        if (early && !stops[j] && reached) {
                            stops[j] = true;
                            if (++done == nv) { decided = true; frequent = true; }
                        }
                    }
                }
            } else {
                r->state[u][i] = 2;
                ninvalid[u]++;
            }
            r->verdicts++;
            finished[u]++;
            if (mode == 0 && !decided && (int64_t)nvalid[u] + (int64_t)r->C[u].size() - finished[u] < freq) {
                decided = true;                                     // domain_early_exit, comper.h:288-297
                frequent = false;
            }
            if (mode == 2 && !decided) {
                ExactSum bound = vsup[u];
                bound.add((double)((int64_t)r->C[u].size() - finished[u]));
                if (bound.rounded() < (double)freq) { decided = true; frequent = false; }
            }
        }
    }
    uint32_t lo = ~0u, hi = ~0u;
    for (int u = 0; u < nv; ++u) {
        lo = std::min(lo, nvalid[u]);
        hi = std::min(hi, (uint32_t)r->C[u].size() - ninvalid[u]);
    }
    r->lower = lo;
    r->upper = hi;
    r->frequent = mode == 1 ? (int)lo >= freq : (decided ? frequent : (int)lo >= freq);
    if (mode == 2) {
        bool all = true;
        for (int u = 0; u < nv; ++u) {
            r->fraction[u] = vsup[u].rounded();
            all = all && r->fraction[u] >= (double)freq;
        }
        r->frequent = decided ? frequent : all;
    }
    return r;
}

extern "C" double orc_eval_fraction(const orc_eval *e, double *per_vertex) {
    double mn = 0;
    for (int u = 0; u < e->nv; ++u) {
        const double f = u < (int)e->fraction.size() ? e->fraction[u] : 0.0;
        if (per_vertex) per_vertex[u] = f;
        mn = u == 0 ? f : std::min(mn, f);
    }
    return mn;
}

extern "C" void orc_eval_free(orc_eval *e) { delete e; }
extern "C" int orc_eval_frequent(const orc_eval *e) { return e->frequent; }
extern "C" int orc_eval_decided_by(const orc_eval *e) { return e->decided_by; }
extern "C" uint32_t orc_eval_support_lower(const orc_eval *e) { return e->lower; }
extern "C" uint32_t orc_eval_support_upper(const orc_eval *e) { return e->upper; }
extern "C" uint64_t orc_eval_searches(const orc_eval *e) { return e->searches; }
extern "C" uint64_t orc_eval_verdicts(const orc_eval *e) { return e->verdicts; }

extern "C" uint32_t orc_eval_domain(const orc_eval *e, int u, uint32_t *out, uint32_t cap) {
    const auto &c = e->C[u];
    for (uint32_t i = 0; i < c.size() && i < cap; ++i) out[i] = c[i];
    return (uint32_t)c.size();
}
extern "C" uint32_t orc_eval_states(const orc_eval *e, int u, uint8_t *out, uint32_t cap) {
    const auto &s = e->state[u];
    for (uint32_t i = 0; i < s.size() && i < cap; ++i) out[i] = s[i];
    return (uint32_t)s.size();
}
extern "C" uint32_t orc_eval_child_domain(const orc_eval *e, int u, uint32_t *out, uint32_t cap) {
    uint32_t k = 0;
    for (size_t i = 0; i < e->C[u].size(); ++i) {
        if (e->state[u][i] == 2) continue;                          // remove_non_cands, comper.h:1001-1018
        if (k < cap) out[k] = e->C[u][i];
        ++k;
    }
    return k;
}
extern "C" int64_t orc_eval_aux(const orc_eval *e, int a, int b, uint32_t *off, uint32_t *idx) {
    if (e->aux.empty() || !e->aux[a][b].present) return -1;
    const Aux &x = e->aux[a][b];
    if (off) std::copy(x.off.begin(), x.off.end(), off);
    if (idx) std::copy(x.idx.begin(), x.idx.end(), idx);
    return (int64_t)x.idx.size();
}
extern "C" void orc_eval_order(const orc_eval *e, int u, uint32_t *order) {
    for (int i = 0; i < e->nv; ++i) order[i] = e->order.empty() ? 0 : e->order[u][i];
}

extern "C" int orc_eval_bn(const orc_eval *e, int u, int d, uint32_t *out) {
    if (e->bn.empty() || u < 0 || u >= e->nv || d < 0 || d >= e->nv) return -1;
    const vector<int> &b = e->bn[u][d];
    for (size_t k = 0; k < b.size(); ++k) out[k] = (uint32_t)b[k];
    return (int)b.size();
}

// ---------------------------------------------------------------------------
// Brute force: every injective, label- and edge-label-preserving map.
// ---------------------------------------------------------------------------
extern "C" uint32_t orc_bruteforce_mni(const orc_graph *gp, int nv, const int32_t *vlab, int ne, const int32_t *edges,
                                       uint32_t *image_sizes) {
    const orc_graph &g = *gp;
    PGraph q = make_query(nv, vlab, ne, edges);
    vector<std::set<uint32_t>> images(nv);
    vector<uint32_t> m(nv);
    auto has_edge = [&](uint32_t a, uint32_t b, int32_t l) {
        for (auto &e : g.adj[a]) if (e.first == b && e.second == l) return true;
        return false;
    };
    std::function<void(int)> rec = [&](int d) {
        if (d == nv) {
            for (int u = 0; u < nv; ++u) images[u].insert(m[u]);
            return;
        }
        for (uint32_t v = 0; v < g.n(); ++v) {
            if (g.vlabel[v] != q.vlab[d]) continue;
            bool ok = true;
            for (int k = 0; k < d && ok; ++k) if (m[k] == v) ok = false;
            for (auto &e : q.adj[d])
                if (ok && e.first < d && !has_edge(v, m[e.first], e.second)) ok = false;
            if (!ok) continue;
            m[d] = v;
            rec(d + 1);
        }
    };
    rec(0);
    uint32_t mni = ~0u;
    for (int u = 0; u < nv; ++u) {
        if (image_sizes) image_sizes[u] = (uint32_t)images[u].size();
        mni = std::min(mni, (uint32_t)images[u].size());
    }
    return nv == 0 ? 0 : mni;
}

// ---------------------------------------------------------------------------
// Full mining run: size-2 level from the endpoint sets (grami.h:83-169,
// 459-481), then level by level; children of size-2 patterns start from the
// label buckets (worker.h:57-106 leaves their parent_prog unset), deeper
// children from the parent's domains minus proven-invalid vertices
// (comper.h:1069-1152).
// ---------------------------------------------------------------------------
namespace {

struct Item {
    tfsm::Pattern p;
    int parent = -1;                 // index into the previous level's results, -1: label buckets
};

struct Done {
    vector<uint32_t> off, dom;       // child domains over the pattern's vertices
};

}  // namespace

extern "C" int64_t orc_mine(const orc_graph *gp, int freq, int max_nodes, int threads, int mode, const char *out_path,
                            double *mining_seconds, uint64_t *patterns_evaluated, uint64_t *verdicts) {
    const orc_graph &g = *gp;
    auto t0 = std::chrono::steady_clock::now();
    FILE *out = std::fopen(out_path, "w");
    if (!out) return -1;
    tfsm::GrowthContext ctx;
    ctx.max_nodes = max_nodes;
    ctx.freq_edge_labels = g.freq_edge_labels;
    int64_t nfreq = 0;
    uint64_t evaluated = 0, nverdicts = 0;
    vector<Item> level;
    // mode 2: the size-2 Fraction-Score filter (fraction-score/graph.h:803-824)
    auto hashed = g.hashed_edges;
    if (mode == 2)
        for (auto it = hashed.begin(); it != hashed.end();) {
            ExactSum a, b;
            for (uint32_t v : it->second.first) a.add(g.frac(v, std::get<2>(it->first)));
            for (uint32_t v : it->second.second) b.add(g.frac(v, std::get<0>(it->first)));
            if (a.rounded() < (double)freq || b.rounded() < (double)freq) it = hashed.erase(it);
            else ++it;
        }
    for (auto &kv : hashed) {
        int32_t la = std::get<0>(kv.first), el = std::get<1>(kv.first), lb = std::get<2>(kv.first);
        uint32_t s = (uint32_t)std::min(kv.second.first.size(), kv.second.second.size());
        ctx.neighbor_labels[la].insert(lb);
        ctx.neighbor_labels[lb].insert(la);
        tfsm::Pattern p = tfsm::single_edge(la, el, lb);
        std::fprintf(out, "%s\t%u\t%u", p.key().c_str(), s, s);
        if (mode == 2) {
            ExactSum a, b;
            for (uint32_t v : kv.second.first) a.add(g.frac(v, lb));
            for (uint32_t v : kv.second.second) b.add(g.frac(v, la));
            std::fprintf(out, "\t%a", std::min(a.rounded(), b.rounded()));
        }
        std::fprintf(out, "\n");
        ++nfreq;
    }
    for (auto &kv : hashed) {
        tfsm::Pattern p = tfsm::single_edge(std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first));
        for (auto &c : tfsm::extend(p, ctx)) level.push_back(Item{std::move(c), -1});
    }
    vector<Done> prev;
    while (!level.empty()) {
        vector<Done> cur(level.size());
        vector<char> freq_flag(level.size(), 0);
        vector<uint32_t> lo(level.size()), hi(level.size());
        vector<uint64_t> vd(level.size());
        vector<double> fr(level.size());
        std::atomic<size_t> next{0};
        auto work = [&]() {
            size_t i;
            while ((i = next.fetch_add(1)) < level.size()) {
                const tfsm::Pattern &p = level[i].p;
                vector<int32_t> edges;
                for (auto &c : p.code) { edges.push_back(c.from); edges.push_back(c.to); edges.push_back(c.el); }
                const Done *par = level[i].parent >= 0 ? &prev[level[i].parent] : nullptr;
                int pnv = par ? (int)par->off.size() - 1 : 0;
                orc_eval *e = orc_eval_pattern(&g, p.nv(), p.vlab.data(), p.ne(), edges.data(), pnv,
                                               par ? par->off.data() : nullptr, par ? par->dom.data() : nullptr, freq, mode);
                freq_flag[i] = (char)e->frequent;
                lo[i] = e->lower; hi[i] = e->upper; vd[i] = e->verdicts;
                fr[i] = orc_eval_fraction(e, nullptr);
                if (e->frequent) {
                    cur[i].off.push_back(0);
                    for (int u = 0; u < p.nv(); ++u) {
                        for (size_t k = 0; k < e->C[u].size(); ++k)
                            if (e->state[u][k] != 2) cur[i].dom.push_back(e->C[u][k]);
                        cur[i].off.push_back((uint32_t)cur[i].dom.size());
                    }
                }
                orc_eval_free(e);
            }
        };
        vector<std::thread> pool;
        for (int t = 0; t < std::max(1, threads); ++t) pool.emplace_back(work);
        for (auto &t : pool) t.join();
        vector<Item> nextlevel;
        for (size_t i = 0; i < level.size(); ++i) {
            evaluated++;
            nverdicts += vd[i];
            if (!freq_flag[i]) continue;
            ++nfreq;
            std::fprintf(out, "%s\t%u\t%u", level[i].p.key().c_str(), lo[i], hi[i]);
            if (mode == 2) std::fprintf(out, "\t%a", fr[i]);
            std::fprintf(out, "\n");
            for (auto &c : tfsm::extend(level[i].p, ctx)) nextlevel.push_back(Item{std::move(c), (int)i});
        }
        prev.swap(cur);
        level.swap(nextlevel);
    }
    std::fclose(out);
    if (mining_seconds)
        *mining_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (patterns_evaluated) *patterns_evaluated = evaluated;
    if (verdicts) *verdicts = nverdicts;
    return nfreq;
}
