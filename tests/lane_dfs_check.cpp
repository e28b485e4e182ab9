// Host check of the lane-per-task existence search
// (t-fsm_amd/csrc/lane_dfs.h): the same source, run on the CPU over the CPU
// oracle's filtered domains, auxiliary tables and matching orders, must
// reproduce the oracle's exact-mode verdict of every candidate it is given,
// in the reference matching order and in the adaptive one.  Every memory
// access the search makes goes through a bounds-checked accessor; a
// violation exits with status 2.
//
// Pieces: every `donate_every` steps the search donates the untried
// candidates of its shallowest open depth, cut into up to `fanout` pieces
// (the records of the search kernel); pieces run one after another and may
// donate again.  The task's verdict is "found" when any piece finds an
// embedding, "exhausted" when every piece is exhausted.
//
//   lane_dfs_check SPEC DONATE_EVERY FANOUT   (SPEC: the pattern file written
//                                             by tests/test_lane_dfs.py)
//
// Random mode: synthetic patterns and tables, checked against a plain
// recursive search written here independently of the lane code (same
// verdicts; in the reference order without pieces, the same candidates
// entered in the same order):
//
//   lane_dfs_check --random SEED INSTANCES DONATE_EVERY FANOUT
//
// Level mode (every pattern of SPEC in one launch with the kernel's record
// queue; REFERENCE 1 searches every pattern in the reference order):
//
//   lane_dfs_check --level SPEC SPLIT_STEPS FANOUT LANES SEED REFERENCE
//
// An optional last argument GROUP (a power of two, default 1) runs every
// root task as GROUP interleaved lanes (phases 0 .. GROUP-1 of the depth-1
// candidates, as the search kernel's lane groups do), each of which may
// donate pieces; the task's verdict combines them all.  `crit_*` then sums,
// over tasks, the most steps any one lane of the group took.
#define TFSM_LANE_FN inline
#include "../t-fsm_amd/csrc/lane_dfs.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../oracle/oracle.h"

namespace {

[[noreturn]] void violation(const char *what, uint64_t a, uint64_t b) {
    std::fprintf(stderr, "bounds violation: %s (%llu vs %llu)\n", what, (unsigned long long)a,
                 (unsigned long long)b);
    std::exit(2);
}

constexpr int kMV = 8;
// the search instance's edge bound (the kernel's ME): 12, or 8 with -DLANE_ME=8
#ifndef LANE_ME
#define LANE_ME 12
#endif
constexpr int kLaneME = LANE_ME;

struct HostCols {
    uint32_t idx_[kMV], vid_[kMV], pos_[kMV], end_[kMV], cur_[kMV], cend_[kMV], curv_[kMV], meta_[kMV];
    uint32_t n = 0;    // pattern vertices: idx and vid by vertex, the rest by depth < n
    void poison() {
        for (int i = 0; i < kMV; ++i)
            idx_[i] = vid_[i] = pos_[i] = end_[i] = cur_[i] = cend_[i] = curv_[i] = meta_[i] = 0xdeadbeefu;
    }
    uint32_t &at(uint32_t *a, uint32_t i, uint32_t lim, const char *what) {
        if (i >= lim) violation(what, i, lim);
        return a[i];
    }
    uint32_t &idx(uint32_t x) { return at(idx_, x, n, "idx"); }
    uint32_t &vid(uint32_t j) { return at(vid_, j, n, "vid"); }
    uint32_t &pos(uint32_t d) { return at(pos_, d, n, "pos"); }
    uint32_t &end(uint32_t d) { return at(end_, d, n, "end"); }
    uint32_t &cur(uint32_t g) { return at(cur_, g, kMV, "cur"); }
    uint32_t &cend(uint32_t g) { return at(cend_, g, kMV, "cend"); }
    uint32_t &curv(uint32_t g) { return at(curv_, g, kMV, "curv"); }
    uint32_t &meta(uint32_t d) { return at(meta_, d, n, "meta"); }
};

struct HostAcc {
    uint32_t n = 0;
    bool adaptive_ = false;
    std::vector<int32_t> lab;
    std::vector<std::vector<uint32_t>> dom;                      // C(x), sorted data vertex ids
    std::vector<uint32_t> pool_;                                 // every table, back to back
    // level mode: the tables of every pattern of the launch in one pool (as
    // on the device, where positions are unique across a launch's patterns)
    const std::vector<uint32_t> *shared = nullptr;
    const std::vector<uint32_t> &pl() const { return shared ? *shared : pool_; }
    std::vector<std::vector<std::vector<uint32_t>>> off;         // off[b][x]: row starts over C(b), absolute
    std::vector<std::vector<uint32_t>> ord;                      // ord[u][d]
    std::vector<std::vector<std::vector<uint32_t>>> bn;          // bn[u][d]: backward neighbours of depth d
    std::vector<std::pair<uint32_t, uint32_t>> edges;            // pattern edges in pattern order

    void chk(uint32_t u, uint32_t d, const char *what) const {
        if (u >= n || d >= n) violation(what, u, d);
    }
    uint32_t nv() const { return n; }
    uint32_t edge_max() const { return (uint32_t)edges.size(); }
    bool adaptive() const { return adaptive_; }
    uint32_t pool(uint32_t pos) const {
        if (pos >= pl().size()) violation("pool", pos, pl().size());
        return pl()[pos];
    }
    uint32_t vid(uint32_t x, uint32_t c) const {
        if (x >= n || c >= dom[x].size()) violation("vid", c, x < n ? dom[x].size() : 0);
        return dom[x][c];
    }
    // slot s = b * n + x: rows over C(b) listing indices into C(x)
    void row(uint32_t s, uint32_t r, uint32_t &lo, uint32_t &hi) const {
        if (s >= n * n) violation("slot", s, n * n);
        const std::vector<uint32_t> &o = off[s / n][s % n];
        if (o.empty()) violation("row: no table", s / n, s % n);
        if ((size_t)r + 1 >= o.size()) violation("row index", r, o.size());
        lo = o[r];
        hi = o[r + 1];
        if (lo > hi || hi > pl().size()) violation("row range", hi, pl().size());
    }
    uint32_t order(uint32_t u, uint32_t d) const {
        chk(u, d, "order");
        return ord[u][d];
    }
    uint32_t bn_n(uint32_t u, uint32_t d) const {
        chk(u, d, "bn_n");
        return (uint32_t)bn[u][d].size();
    }
    uint32_t bn_slot(uint32_t u, uint32_t d, uint32_t k) const {
        if (k >= bn_n(u, d)) violation("bn_slot", k, bn[u][d].size());
        return bn[u][d][k] * n + ord[u][d];
    }
    uint32_t slot_src(uint32_t s) const {
        if (s >= n * n) violation("slot_src", s, n * n);
        return s / n;
    }
    uint32_t sec_base(uint32_t u, uint32_t d) const {
        chk(u, d, "sec_base");
        uint32_t g = 0;
        for (uint32_t j = 1; j < d; ++j) g += (uint32_t)bn[u][j].size() - 1;
        return g;
    }
    uint32_t same_before(uint32_t u, uint32_t d) const {
        chk(u, d, "same_before");
        uint32_t m = 0;
        for (uint32_t j = 0; j < d; ++j)
            if (lab[ord[u][j]] == lab[ord[u][d]]) m |= 1u << j;
        return m;
    }
    uint32_t same_after(uint32_t u, uint32_t d) const {
        chk(u, d, "same_after");
        for (uint32_t j = d + 1; j < n; ++j)
            if (lab[ord[u][j]] == lab[ord[u][d]]) return 1;
        return 0;
    }
    uint32_t edge_x(uint32_t e) const {
        if (e >= edges.size()) violation("edge_x", e, edges.size());
        return edges[e].first;
    }
    uint32_t edge_y(uint32_t e) const {
        if (e >= edges.size()) violation("edge_y", e, edges.size());
        return edges[e].second;
    }
    uint32_t edge_pack(uint32_t i) const {
        if (i >= tfsm_lane::kCutMax / 4) violation("edge_pack", i, tfsm_lane::kCutMax / 4);
        uint32_t w = 0;
        for (uint32_t k = 0; k < 4; ++k) {
            const uint32_t e = 4 * i + k;
            const uint32_t b = e < edges.size() ? (edges[e].first | edges[e].second << 4) : 0xffu;
            w |= b << (8 * k);
        }
        return w;
    }
    // the host keeps the slot itself as its "offset position"
    uint32_t cut_ao(uint32_t slot) const {
        if (slot >= 2 * edges.size()) violation("cut_ao", slot, 2 * edges.size());
        return slot;
    }
    void row_at(uint32_t slot, uint32_t r, uint32_t &lo, uint32_t &hi) const {
        const uint32_t e = slot / 2;
        row(slot % 2 == 0 ? edge_x(e) * n + edge_y(e) : edge_y(e) * n + edge_x(e), r, lo, hi);
    }
    uint32_t same_lab(uint32_t x) const {
        if (x >= n) violation("same_lab", x, n);
        uint32_t m = 0;
        for (uint32_t y = 0; y < n; ++y)
            if (y != x && lab[y] == lab[x]) m |= 1u << y;
        return m;
    }
    // the library's rule for the adaptive order
    bool adaptive_fits() const {
        return edges.size() <= (size_t)kLaneME && edges.size() - (n - 1) <= (size_t)std::min(kMV, 15);
    }
};

struct Piece {
    uint32_t floor, prim, lo, hi;
    uint32_t prefix[kMV];
    uint8_t pord[kMV];
};

struct TaskResult {
    bool found = false, all_exhausted = true;
    uint64_t steps = 0, pieces = 0;
    uint64_t crit = 0;         // most steps of one root lane
};

uint32_t g_group = 1;          // lanes per root task

// Runs task (u, ci) with pieces; `seq` (when set) collects the entered
// (depth, candidate) pairs of the root piece.  ORD: the lane instance (the
// adaptive-only one is what the search kernel runs when every pattern of a
// launch is in the adaptive order).
template <int ORD>
TaskResult run_task_in(const HostAcc &a, uint32_t u, uint32_t ci, uint32_t donate_every, uint32_t fanout,
                       const std::vector<std::pair<uint32_t, uint32_t>> &edges_pat,
                       std::vector<std::pair<uint32_t, uint32_t>> *seq) {
    TaskResult res;
    HostCols c;
    c.n = a.n;
    tfsm_lane::Lane<kMV, ORD, kLaneME> L;
    std::vector<Piece> queue;
    auto check_embedding = [&]() {
        if (c.idx(u) != ci) violation("embedding root", c.idx(u), ci);
        for (uint32_t i = 0; i < a.n; ++i)
            for (uint32_t j = i + 1; j < a.n; ++j)
                if (a.dom[i][c.idx(i)] == a.dom[j][c.idx(j)]) violation("not injective", i, j);
        for (auto &e : edges_pat) {
            const std::vector<uint32_t> &o = a.off[e.first][e.second];
            bool hit = false;
            for (uint32_t q = o[c.idx(e.first)]; q < o[c.idx(e.first) + 1]; ++q) hit |= a.pool_[q] == c.idx(e.second);
            if (!hit) violation("embedding edge missing", e.first, e.second);
        }
    };
    auto run = [&](bool root_piece) {
        uint32_t since = 0;
        while (true) {
            const uint32_t r = L.step(c, a);
            ++res.steps;
            // the candidate the step entered, for the visit sequence
            if (seq && root_piece && L.last_depth != tfsm_lane::kNone) seq->emplace_back(L.last_depth, L.last_cand);
            if (r == tfsm_lane::kLaneFound) {
                check_embedding();
                return r;
            }
            if (r == tfsm_lane::kLaneExhausted) return r;
            if (donate_every && ++since >= donate_every) {
                since = 0;
                const uint32_t k = L.open_depth(c);
                if (k <= L.d) {
                    const uint32_t lo = c.pos(k), rem = c.end(k) - lo;
                    const uint32_t m = std::min(rem, fanout);
                    for (uint32_t q = 0; q < m; ++q) {
                        Piece p;
                        p.floor = k;
                        p.prim = tfsm_lane::meta_prim(c.meta(k));
                        p.lo = lo + (uint32_t)((uint64_t)rem * q / m);
                        p.hi = lo + (uint32_t)((uint64_t)rem * (q + 1) / m);
                        for (uint32_t j = 0; j < kMV; ++j) {
                            p.prefix[j] = 0xdeadbeefu;
                            p.pord[j] = 0xee;
                        }
                        for (uint32_t j = 0; j <= k && j < a.n; ++j) {
                            const uint32_t w = L.vertex_at(c, a, j);
                            p.pord[j] = (uint8_t)w;
                            if (j < k) p.prefix[j] = c.idx(w);
                        }
                        queue.push_back(p);
                    }
                    c.end(k) = lo;
                }
            }
        }
    };
    uint32_t r;
    for (uint32_t phase = 0; phase < g_group; ++phase) {
        c.poison();
        L.start_root(c, a, u, ci, g_group, phase);
        const uint64_t s0 = res.steps;
        r = run(g_group == 1);
        res.crit = std::max(res.crit, res.steps - s0);
        res.found |= r == tfsm_lane::kLaneFound;
        res.all_exhausted &= r == tfsm_lane::kLaneExhausted;
    }
    for (size_t q = 0; q < queue.size(); ++q) {
        const Piece p = queue[q];
        if (p.prefix[0] != ci || p.pord[0] != u) violation("piece root", p.prefix[0], ci);
        c.poison();
        L.start_piece(c, a, u, p.floor, p.prefix, p.pord, p.prim, p.lo, p.hi);
        r = run(false);
        res.found |= r == tfsm_lane::kLaneFound;
        res.all_exhausted &= r == tfsm_lane::kLaneExhausted;
        ++res.pieces;
    }
    return res;
}

// Adaptive-order tasks alternate between the two instances that run them.
TaskResult run_task(const HostAcc &a, uint32_t u, uint32_t ci, uint32_t donate_every, uint32_t fanout,
                    const std::vector<std::pair<uint32_t, uint32_t>> &edges_pat,
                    std::vector<std::pair<uint32_t, uint32_t>> *seq) {
    if (a.adaptive_ && (ci & 1u))
        return run_task_in<tfsm_lane::kOrderAdaptive>(a, u, ci, donate_every, fanout, edges_pat, seq);
    return run_task_in<tfsm_lane::kOrderMixed>(a, u, ci, donate_every, fanout, edges_pat, seq);
}

// Random instances: vertices 0..n-1 with labels in {0, 1}; a connected
// pattern; domains of random size with random data vertex ids (collisions
// only within a label); a table for every pattern edge in both directions
// with random ascending rows (empty, short, long), the two directions
// transposes of each other.
struct Rng {
    uint64_t s;
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    uint32_t below(uint32_t n) { return (uint32_t)(next() % n); }
};

void random_instance(Rng &rng, HostAcc &a, std::vector<std::vector<int>> &adj) {
    const int n = 3 + (int)rng.below(4);
    a = HostAcc();
    a.n = (uint32_t)n;
    a.lab.resize(n);
    for (auto &l : a.lab) l = (int)rng.below(2);
    adj.assign(n, std::vector<int>(n, 0));
    for (int x = 1; x < n; ++x) {
        const int y = (int)rng.below((uint32_t)x);
        adj[x][y] = adj[y][x] = 1;
        a.edges.emplace_back((uint32_t)y, (uint32_t)x);
    }
    for (int t = 0; t < n; ++t) {
        const int x = (int)rng.below(n), y = (int)rng.below(n);
        if (x != y && !adj[x][y] && rng.below(2) == 0) {
            adj[x][y] = adj[y][x] = 1;
            a.edges.emplace_back((uint32_t)x, (uint32_t)y);
        }
    }
    a.dom.resize(n);
    for (int x = 0; x < n; ++x) {
        const uint32_t sz = 20 + rng.below(200);
        uint32_t v = (uint32_t)a.lab[x];
        for (uint32_t i = 0; i < sz; ++i) {
            v += 2 * (1 + rng.below(3));
            a.dom[x].push_back(v);
        }
    }
    // rows of b -> x for b < x at random; x -> b is their transpose (as
    // in a real table pair, where j is in row i of b -> x exactly when i is
    // in row j of x -> b)
    a.off.assign(n, std::vector<std::vector<uint32_t>>(n));
    for (int b = 0; b < n; ++b)
        for (int x = b + 1; x < n; ++x) {
            if (!adj[b][x]) continue;
            const uint32_t cb = (uint32_t)a.dom[b].size(), cx = (uint32_t)a.dom[x].size();
            std::vector<std::vector<uint32_t>> fwd(cb), bwd(cx);
            for (uint32_t r = 0; r < cb; ++r) {
                const uint32_t kind = rng.below(8);
                const uint32_t len = kind == 0 ? 0 : kind < 5 ? rng.below(8) : kind < 7 ? 20 + rng.below(60) : cx / 2;
                for (uint32_t c = 0; c < cx; ++c)
                    if (rng.below(cx) < len) {
                        fwd[r].push_back(c);
                        bwd[c].push_back(r);
                    }
            }
            for (int dir = 0; dir < 2; ++dir) {
                const std::vector<std::vector<uint32_t>> &rows = dir ? bwd : fwd;
                std::vector<uint32_t> o{(uint32_t)a.pool_.size()};
                for (auto &row : rows) {
                    a.pool_.insert(a.pool_.end(), row.begin(), row.end());
                    o.push_back((uint32_t)a.pool_.size());
                }
                (dir ? a.off[x][b] : a.off[b][x]) = std::move(o);
            }
        }
    a.ord.assign(n, std::vector<uint32_t>());
    a.bn.assign(n, std::vector<std::vector<uint32_t>>(n));
    for (int u = 0; u < n; ++u) {                 // BFS order from u
        std::vector<int> seen(n, 0);
        a.ord[u].push_back((uint32_t)u);
        seen[u] = 1;
        for (size_t q = 0; q < a.ord[u].size(); ++q)
            for (int y = 0; y < n; ++y)
                if (adj[a.ord[u][q]][y] && !seen[y]) {
                    seen[y] = 1;
                    a.ord[u].push_back((uint32_t)y);
                }
        for (int d = 0; d < n; ++d)
            for (int j = 0; j < d; ++j)
                if (adj[a.ord[u][j]][a.ord[u][d]]) a.bn[u][d].push_back(a.ord[u][j]);
    }
}

// Plain recursive existence search in the reference order: candidates of
// depth d are the entries of the primary (shortest, first on ties) backward
// row present in every other one, entered in ascending order.
bool plain_exists(const HostAcc &a, uint32_t u, std::vector<uint32_t> &idx, uint32_t d,
                  std::vector<std::pair<uint32_t, uint32_t>> &seq) {
    const uint32_t n = a.n;
    if (d == n) return true;
    const uint32_t x = a.ord[u][d];
    const auto &B = a.bn[u][d];
    uint32_t prim = 0, best = 0xffffffffu;
    for (uint32_t k = 0; k < B.size(); ++k) {
        const std::vector<uint32_t> &o = a.off[B[k]][x];
        const uint32_t len = o[idx[B[k]] + 1] - o[idx[B[k]]];
        if (len < best) { best = len; prim = k; }
    }
    const std::vector<uint32_t> &op = a.off[B[prim]][x];
    for (uint32_t p = op[idx[B[prim]]]; p < op[idx[B[prim]] + 1]; ++p) {
        const uint32_t c = a.pool_[p];
        bool ok = true;
        for (uint32_t k = 0; k < B.size() && ok; ++k) {
            if (k == prim) continue;
            const std::vector<uint32_t> &o = a.off[B[k]][x];
            bool in = false;
            for (uint32_t q = o[idx[B[k]]]; q < o[idx[B[k]] + 1]; ++q) in |= a.pool_[q] == c;
            ok = in;
        }
        for (uint32_t j = 0; j < d && ok; ++j)
            if (a.dom[a.ord[u][j]][idx[a.ord[u][j]]] == a.dom[x][c]) ok = false;
        if (!ok) continue;
        seq.emplace_back(d, c);
        idx[x] = c;
        if (plain_exists(a, u, idx, d + 1, seq)) return true;
    }
    return false;
}

// Accessor over the oracle's domains, tables (appended to one pool) and
// reference-order plans of one evaluated pattern.
void build_acc(orc_eval *e, int nv, const std::vector<int32_t> &vl, int ne, const std::vector<int32_t> &ed,
               HostAcc &a) {
    a.n = (uint32_t)nv;
    a.lab = vl;
    a.dom.resize(nv);
    for (int x = 0; x < nv; ++x) {
        const uint32_t sz = orc_eval_domain(e, x, nullptr, 0);
        a.dom[x].resize(sz);
        orc_eval_domain(e, x, a.dom[x].data(), sz);
    }
    a.off.assign(nv, std::vector<std::vector<uint32_t>>(nv));
    for (int b = 0; b < nv; ++b)
        for (int x = 0; x < nv; ++x) {
            const int64_t cnt = orc_eval_aux(e, b, x, nullptr, nullptr);
            if (cnt < 0) continue;
            std::vector<uint32_t> o(a.dom[b].size() + 1), ix((size_t)cnt);
            orc_eval_aux(e, b, x, o.data(), ix.data());
            const uint32_t base = (uint32_t)a.pool_.size();
            a.pool_.insert(a.pool_.end(), ix.begin(), ix.end());
            for (auto &v : o) v += base;
            a.off[b][x] = std::move(o);
        }
    std::vector<std::vector<int>> adj(nv, std::vector<int>(nv, 0));
    for (int i = 0; i < ne; ++i) {
        adj[ed[3 * i]][ed[3 * i + 1]] = adj[ed[3 * i + 1]][ed[3 * i]] = 1;
        a.edges.emplace_back((uint32_t)ed[3 * i], (uint32_t)ed[3 * i + 1]);
    }
    a.ord.assign(nv, std::vector<uint32_t>(nv));
    a.bn.assign(nv, std::vector<std::vector<uint32_t>>(nv));
    for (int u = 0; u < nv; ++u) {
        orc_eval_order(e, u, a.ord[u].data());
        for (int d = 0; d < nv; ++d)
            for (int j = 0; j < d; ++j)
                if (adj[a.ord[u][j]][a.ord[u][d]]) a.bn[u][d].push_back(a.ord[u][j]);
    }
}


// Level mode: every pattern of the spec evaluated together as one launch of
// the search kernel runs them, with its record queue: `lanes` virtual lanes
// take root tasks (pattern-major, u ascending, candidates in order) while
// any remain, then claim records from a FIFO; a piece donates the untried
// candidates of its shallowest open depth, in up to `fanout` records of
// consecutive ranges, once it has run split_steps steps since its last
// donation, checked every 64 steps as in the kernel, and only when every
// root task has been handed out and no record waits.  Each task has a piece
// counter; its verdict is "valid" when some piece found an embedding and
// "invalid" when the last piece ends without one.  The lanes run
// `round` steps per turn in an order shuffled per round (seeded), so pieces
// of different patterns interleave.  Verdicts must equal the oracle's.
struct LevelTask {
    uint32_t p, u, ci;
    uint32_t pieces = 0;
    bool found = false;
    int verdict = 0;        // 0 pending, 1 valid, 2 invalid
};

struct LevelRecord {
    uint32_t task, floor, prim, lo, hi;
    uint32_t prefix[kMV];
    uint8_t pord[kMV];
};

int level_mode(const char *spec_path, uint32_t split_steps, uint32_t fanout, uint32_t nlanes, uint64_t seed,
               bool reference_order) {
    std::ifstream in(spec_path);
    std::string graph_path;
    int freq = 0;
    uint32_t samples = 0;
    in >> graph_path >> freq >> samples;
    orc_graph *g = orc_load(graph_path.c_str(), freq);
    if (!g) return 64;
    std::string line;
    std::getline(in, line);
    std::vector<HostAcc> accs;
    std::vector<std::vector<std::vector<uint8_t>>> want;     // oracle states per pattern, vertex, candidate
    std::vector<LevelTask> tasks;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        int nv, ne;
        if (!(ls >> nv)) continue;
        std::vector<int32_t> vl(nv);
        for (auto &x : vl) ls >> x;
        ls >> ne;
        std::vector<int32_t> ed(3 * ne);
        for (auto &x : ed) ls >> x;
        if (nv > kMV) continue;
        orc_eval *e = orc_eval_pattern(g, nv, vl.data(), ne, ed.data(), 0, nullptr, nullptr, freq, 1);
        if (!e) continue;
        if (orc_eval_decided_by(e) != 2) {
            orc_eval_free(e);
            continue;
        }
        HostAcc a;
        build_acc(e, nv, vl, ne, ed, a);
        a.adaptive_ = !reference_order && a.adaptive_fits();
        const uint32_t p = (uint32_t)accs.size();
        std::vector<std::vector<uint8_t>> st(nv);
        for (int u = 0; u < nv; ++u) {
            st[u].resize(a.dom[u].size());
            orc_eval_states(e, u, st[u].data(), (uint32_t)st[u].size());
            for (uint32_t ci = 0; ci < st[u].size(); ++ci) tasks.push_back(LevelTask{p, (uint32_t)u, ci});
        }
        accs.push_back(std::move(a));
        want.push_back(std::move(st));
        orc_eval_free(e);
    }
    orc_graph_free(g);
    // one pool for the launch: each pattern's rows move to absolute positions
    std::vector<uint32_t> pool;
    for (HostAcc &a : accs) {
        const uint32_t base = (uint32_t)pool.size();
        pool.insert(pool.end(), a.pool_.begin(), a.pool_.end());
        for (auto &ob : a.off)
            for (auto &o : ob)
                for (auto &v : o) v += base;
        a.pool_.clear();
    }
    for (HostAcc &a : accs) a.shared = &pool;
    struct VLane {
        bool active = false;
        uint32_t task = 0, steps = 0, since = 0;
        HostCols c;
        tfsm_lane::Lane<kMV, tfsm_lane::kOrderMixed, kLaneME> L;
    };
    std::vector<VLane> lanes(nlanes);
    std::vector<LevelRecord> fifo;          // reserved records; head never passes the tail
    size_t head = 0, next_task = 0;
    uint64_t pieces = 0, records = 0, steps_total = 0;
    Rng rng{seed};
    std::vector<uint32_t> turn(nlanes);
    for (uint32_t i = 0; i < nlanes; ++i) turn[i] = i;
    auto end_piece = [&](VLane &ln, bool found) {
        LevelTask &t = tasks[ln.task];
        t.found |= found;
        if (--t.pieces == 0) t.verdict = t.found ? 1 : 2;
        ln.active = false;
    };
    const uint32_t round = 4;
    for (;;) {
        bool any = false;
        for (uint32_t i = nlanes - 1; i > 0; --i) std::swap(turn[i], turn[rng.below(i + 1)]);
        for (uint32_t li : turn) {
            VLane &ln = lanes[li];
            if (!ln.active) {
                if (next_task < tasks.size()) {
                    const LevelTask &t = tasks[next_task];
                    ln.task = (uint32_t)next_task++;
                    tasks[ln.task].pieces = 1;
                    ln.c.poison();
                    ln.c.n = accs[t.p].n;
                    ln.L.start_root(ln.c, accs[t.p], t.u, t.ci);
                } else if (head < fifo.size()) {
                    const LevelRecord r = fifo[head++];
                    const LevelTask &t = tasks[r.task];
                    if (r.prefix[0] != t.ci || r.pord[0] != t.u) violation("record root", r.prefix[0], t.ci);
                    ln.task = r.task;
                    ln.c.poison();
                    ln.c.n = accs[t.p].n;
                    ln.L.start_piece(ln.c, accs[t.p], t.u, r.floor, r.prefix, r.pord, r.prim, r.lo, r.hi);
                    ++pieces;
                } else {
                    continue;
                }
                ln.active = true;
                ln.steps = ln.since = 0;
            }
            any = true;
            const HostAcc &a = accs[tasks[ln.task].p];
            for (uint32_t k = 0; k < round && ln.active; ++k) {
                ++ln.steps;
                ++ln.since;
                if ((ln.steps & 63u) == 0 && split_steps && ln.since >= split_steps && next_task >= tasks.size() &&
                    head >= fifo.size()) {
                    ln.since = 0;
                    const uint32_t kd = ln.L.open_depth(ln.c);
                    if (kd <= ln.L.d) {
                        const uint32_t lo = ln.c.pos(kd), rem = ln.c.end(kd) - lo;
                        const uint32_t m = std::min(rem, fanout);
                        tasks[ln.task].pieces += m;
                        for (uint32_t q = 0; q < m; ++q) {
                            LevelRecord r;
                            r.task = ln.task;
                            r.floor = kd;
                            r.prim = tfsm_lane::meta_prim(ln.c.meta(kd));
                            r.lo = lo + (uint32_t)((uint64_t)rem * q / m);
                            r.hi = lo + (uint32_t)((uint64_t)rem * (q + 1) / m);
                            for (uint32_t j = 0; j < kMV; ++j) {
                                r.prefix[j] = 0xdeadbeefu;
                                r.pord[j] = 0xee;
                            }
                            for (uint32_t j = 0; j <= kd && j < a.n; ++j) {
                                const uint32_t w = ln.L.vertex_at(ln.c, a, j);
                                r.pord[j] = (uint8_t)w;
                                if (j < kd) r.prefix[j] = ln.c.idx(w);
                            }
                            fifo.push_back(r);
                            ++records;
                        }
                        ln.c.end(kd) = lo;
                    }
                }
                const uint32_t r = ln.L.step(ln.c, a);
                ++steps_total;
                if (r == tfsm_lane::kLaneFound) end_piece(ln, true);
                else if (r == tfsm_lane::kLaneExhausted) end_piece(ln, false);
            }
        }
        if (!any) break;
    }
    uint64_t mismatches = 0, valid = 0;
    for (const LevelTask &t : tasks) {
        const uint8_t w = want[t.p][t.u][t.ci];
        valid += t.verdict == 1;
        if (t.verdict != w && ++mismatches <= 10)
            std::fprintf(stderr, "mismatch: pattern %u u %u ci %u got %d want %d\n", t.p, t.u, t.ci, t.verdict, w);
    }
    std::printf("level_patterns %zu tasks %zu valid %llu mismatches %llu records %llu pieces %llu steps %llu\n",
                accs.size(), tasks.size(), (unsigned long long)valid, (unsigned long long)mismatches,
                (unsigned long long)records, (unsigned long long)pieces, (unsigned long long)steps_total);
    return mismatches || head != fifo.size() ? 1 : 0;
}

int random_mode(uint64_t seed, uint32_t instances, uint32_t donate_every, uint32_t fanout) {
    Rng rng{seed};
    uint64_t checked = 0, mismatches = 0, pieces = 0, found_total = 0, steps_ref = 0, steps_ad = 0, crit = 0;
    for (uint32_t it = 0; it < instances; ++it) {
        HostAcc a;
        std::vector<std::vector<int>> adj;
        random_instance(rng, a, adj);
        if (!a.adaptive_fits()) continue;
        for (uint32_t u = 0; u < a.n; ++u)
            for (uint32_t ci = 0; ci < a.dom[u].size(); ci += 1 + rng.below(3)) {
                std::vector<uint32_t> idx(a.n, 0);
                idx[u] = ci;
                std::vector<std::pair<uint32_t, uint32_t>> want_seq, got_seq;
                const bool want = plain_exists(a, u, idx, 1, want_seq);
                for (int adaptive = 0; adaptive < 2; ++adaptive) {
                    a.adaptive_ = adaptive != 0;
                    got_seq.clear();
                    const bool visit = !donate_every && !adaptive && g_group == 1;
                    TaskResult r = run_task(a, u, ci, donate_every, fanout, a.edges, visit ? &got_seq : nullptr);
                    (adaptive ? steps_ad : steps_ref) += r.steps;
                    crit += r.crit;
                    pieces += r.pieces;
                    found_total += r.found;
                    ++checked;
                    const bool bad = r.found != want || (!r.found && !r.all_exhausted) || (visit && got_seq != want_seq);
                    if (bad && ++mismatches <= 10)
                        std::fprintf(stderr, "mismatch: instance %u u %u ci %u adaptive %d got %d want %d (%zu vs %zu)\n",
                                     it, u, ci, adaptive, r.found, want, got_seq.size(), want_seq.size());
                }
            }
    }
    std::printf("random checked %llu mismatches %llu pieces %llu found %llu steps_reference %llu steps_adaptive %llu "
                "crit %llu\n",
                (unsigned long long)checked, (unsigned long long)mismatches, (unsigned long long)pieces,
                (unsigned long long)found_total, (unsigned long long)steps_ref, (unsigned long long)steps_ad,
                (unsigned long long)crit);
    return mismatches ? 1 : 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc == 8 && std::string(argv[1]) == "--level")
        return level_mode(argv[2], (uint32_t)std::atoi(argv[3]), (uint32_t)std::atoi(argv[4]),
                          (uint32_t)std::atoi(argv[5]), std::strtoull(argv[6], nullptr, 10), std::atoi(argv[7]) != 0);
    if ((argc == 6 || argc == 7) && std::string(argv[1]) == "--random") {
        if (argc == 7) g_group = (uint32_t)std::atoi(argv[6]);
        if (g_group < 1 || g_group > 64 || (g_group & (g_group - 1))) return 64;
        return random_mode(std::strtoull(argv[2], nullptr, 10), (uint32_t)std::atoi(argv[3]),
                           (uint32_t)std::atoi(argv[4]), (uint32_t)std::atoi(argv[5]));
    }
    if (argc != 4 && argc != 5) {
        std::fprintf(stderr, "usage: lane_dfs_check SPEC DONATE_EVERY FANOUT [GROUP]\n");
        return 64;
    }
    if (argc == 5) g_group = (uint32_t)std::atoi(argv[4]);
    if (g_group < 1 || g_group > 64 || (g_group & (g_group - 1))) return 64;
    const uint32_t donate_every = (uint32_t)std::atoi(argv[2]), fanout = (uint32_t)std::atoi(argv[3]);
    std::ifstream in(argv[1]);
    std::string graph_path;
    int freq = 0;
    uint32_t samples = 0;
    in >> graph_path >> freq >> samples;
    orc_graph *g = orc_load(graph_path.c_str(), freq);
    if (!g) {
        std::fprintf(stderr, "cannot load %s\n", graph_path.c_str());
        return 64;
    }
    std::string line;
    std::getline(in, line);
    uint64_t checked = 0, mismatches = 0, patterns = 0, pieces = 0, steps_ref = 0, steps_ad = 0, crit = 0;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        int nv, ne;
        if (!(ls >> nv)) continue;
        std::vector<int32_t> vl(nv);
        for (auto &x : vl) ls >> x;
        ls >> ne;
        std::vector<int32_t> ed(3 * ne);
        for (auto &x : ed) ls >> x;
        if (nv > kMV) continue;
        orc_eval *e = orc_eval_pattern(g, nv, vl.data(), ne, ed.data(), 0, nullptr, nullptr, freq, 1);
        if (!e) continue;
        if (orc_eval_decided_by(e) != 2) {
            orc_eval_free(e);
            continue;
        }
        ++patterns;
        HostAcc a;
        build_acc(e, nv, vl, ne, ed, a);
        const bool fits = a.adaptive_fits();
        for (int u = 0; u < nv; ++u) {
            std::vector<uint8_t> states(a.dom[u].size());
            orc_eval_states(e, u, states.data(), (uint32_t)states.size());
            const size_t sz = states.size();
            const size_t stride = samples && sz > samples ? sz / samples : 1;
            for (size_t ci = 0; ci < sz; ci += stride)
                for (int adaptive = 0; adaptive < (fits ? 2 : 1); ++adaptive) {
                    a.adaptive_ = adaptive != 0;
                    TaskResult r = run_task(a, (uint32_t)u, (uint32_t)ci, donate_every, fanout, a.edges, nullptr);
                    (adaptive ? steps_ad : steps_ref) += r.steps;
                    crit += r.crit;
                    pieces += r.pieces;
                    const uint8_t want = states[ci];
                    const uint8_t got = r.found ? 1 : r.all_exhausted ? 2 : 0;
                    ++checked;
                    if (got != want && ++mismatches <= 10)
                        std::fprintf(stderr, "mismatch: pattern %llu u %d ci %zu adaptive %d got %d want %d\n",
                                     (unsigned long long)patterns, u, ci, adaptive, got, want);
                }
        }
        orc_eval_free(e);
    }
    orc_graph_free(g);
    std::printf("patterns %llu checked %llu mismatches %llu pieces %llu steps_reference %llu steps_adaptive %llu "
                "crit %llu\n",
                (unsigned long long)patterns, (unsigned long long)checked, (unsigned long long)mismatches,
                (unsigned long long)pieces, (unsigned long long)steps_ref, (unsigned long long)steps_ad,
                (unsigned long long)crit);
    return mismatches ? 1 : 0;
}
