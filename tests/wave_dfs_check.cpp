// Host check of the wave-cooperative existence search
// (t-fsm_amd/csrc/wave_dfs.h): the same source, run with an emulated wave
// (a loop over 64 lanes) on the CPU oracle's filtered domains, auxiliary
// tables and matching orders, must reproduce the oracle's exact-mode verdict
// of every candidate it is given.  Every memory access the search makes goes
// through a bounds-checked accessor; a violation exits with status 2.
//
// Pieces: the search donates the pending work of its shallowest open depth
// every `donate_every` rounds (the untested range split into up to `fanout`
// pieces, the tested candidates with the first one); the pieces
// run one after another, and the task's verdict is "found" when any piece
// finds an embedding, "exhausted" when every piece is exhausted.
//
//   wave_dfs_check SPEC DONATE_EVERY FANOUT   (SPEC: see the pattern file
//                                             written by tests/test_wave_dfs.py)
//
// Random mode: synthetic patterns and tables with long rows (well past the
// 64-entry chunks and windows), checked against a plain recursive search
// written here independently of the wave code:
//
//   wave_dfs_check --random SEED INSTANCES DONATE_EVERY FANOUT
#define TFSM_WAVE_FN inline
#include "../t-fsm_amd/csrc/wave_dfs.h"

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

struct HostWave {
    template <class F>
    unsigned long long ballot(F f) const {
        unsigned long long m = 0;
        for (uint32_t l = 0; l < tfsm_wave::kLanes; ++l)
            if (f(l)) m |= 1ull << l;
        return m;
    }
    // dst[i] = f(i) for i < count (one lane each on the device); the rest of
    // dst is poisoned so that a read of an unstaged entry shows up
    template <class F>
    void stage(uint32_t *dst, uint32_t count, F f) const {
        if (count > tfsm_wave::kLanes) violation("stage count", count, tfsm_wave::kLanes);
        for (uint32_t i = 0; i < count; ++i) dst[i] = f(i);
        for (uint32_t i = count; i < tfsm_wave::kLanes; ++i) dst[i] = 0xdeadbeefu;
    }
};

struct HostAcc {
    uint32_t n = 0;
    std::vector<std::vector<uint32_t>> dom;                      // C(x), sorted data vertex ids
    std::vector<uint32_t> pool_;                                 // every table, back to back
    std::vector<std::vector<std::vector<uint32_t>>> off;         // off[b][x]: row starts over C(b), absolute
    std::vector<std::vector<uint32_t>> ord;                      // ord[u][d]
    std::vector<std::vector<std::vector<uint32_t>>> bn;          // bn[u][d]: backward neighbours of depth d
    std::vector<std::vector<uint32_t>> sb;                       // sb[u][d]: earlier depths with d's label

    uint32_t nv() const { return n; }
    uint32_t order(uint32_t u, uint32_t d) const {
        if (u >= n || d >= n) violation("order", u, d);
        return ord[u][d];
    }
    uint32_t bn_n(uint32_t u, uint32_t d) const {
        if (u >= n || d >= n) violation("bn_n", u, d);
        return (uint32_t)bn[u][d].size();
    }
    uint32_t bn_vertex(uint32_t u, uint32_t d, uint32_t k) const {
        if (k >= bn_n(u, d)) violation("bn_vertex", k, bn[u][d].size());
        return bn[u][d][k];
    }
    void row(uint32_t u, uint32_t d, uint32_t k, uint32_t idx_b, uint32_t &lo, uint32_t &hi) const {
        const uint32_t b = bn_vertex(u, d, k), x = order(u, d);
        const std::vector<uint32_t> &o = off[b][x];
        if (o.empty()) violation("row: no table", b, x);
        if ((size_t)idx_b + 1 >= o.size()) violation("row index", idx_b, o.size());
        lo = o[idx_b];
        hi = o[idx_b + 1];
        if (lo > hi || hi > pool_.size()) violation("row range", hi, pool_.size());
    }
    uint32_t pool(uint32_t pos) const {
        if (pos >= pool_.size()) violation("pool", pos, pool_.size());
        return pool_[pos];
    }
    uint32_t vid(uint32_t x, uint32_t c) const {
        if (x >= n || c >= dom[x].size()) violation("vid", c, x < n ? dom[x].size() : 0);
        return dom[x][c];
    }
    uint32_t same_before(uint32_t u, uint32_t d) const {
        if (u >= n || d >= n) violation("same_before", u, d);
        return sb[u][d];
    }
};

struct Piece {
    uint32_t floor, lo, hi, cpos;
    unsigned long long cmask;
    uint32_t prefix[16];
};

struct HostHooks {
    uint32_t abort_every = 0, donate_every = 0, fanout = 1, u = 0;
    const HostAcc *a = nullptr;
    std::vector<Piece> *queue = nullptr;
    uint64_t donations = 0;
    std::vector<std::pair<uint32_t, uint32_t>> *seq = nullptr;   // entered (depth, candidate), when set
    bool abort() const { return false; }
    void entered(uint32_t d, uint32_t c) {
        if (seq) seq->emplace_back(d, c);
    }
    template <int MV>
    bool donate(uint32_t k, uint32_t lo, uint32_t hi, uint32_t cpos, unsigned long long cmask,
                const tfsm_wave::WaveStack<MV> &s) {
        if (hi < lo || (hi == lo && !cmask) || k == 0 || k >= a->n) violation("donated work", lo, hi);
        const uint32_t m = std::max<uint32_t>(1, std::min<uint32_t>(fanout, hi - lo));
        for (uint32_t r = 0; r < m; ++r) {
            Piece p;
            p.floor = k;
            p.lo = lo + (uint32_t)((uint64_t)(hi - lo) * r / m);
            p.hi = lo + (uint32_t)((uint64_t)(hi - lo) * (r + 1) / m);
            p.cpos = cpos;
            p.cmask = r == 0 ? cmask : 0ull;
            for (uint32_t j = 0; j < k; ++j) p.prefix[j] = s.idx[a->order(u, j)];
            queue->push_back(p);
        }
        ++donations;
        return true;
    }
};

// Random instances.  Vertices 0..n-1 with labels in {0, 1}; a connected
// pattern; domains of random size with random data vertex ids (collisions
// only within a label, as in a real graph); a table for every pattern edge in
// both directions with random ascending rows, some far longer than 64.
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
    const int n = 3 + (int)rng.below(3);
    a = HostAcc();
    a.n = (uint32_t)n;
    std::vector<int> lab(n);
    for (auto &l : lab) l = (int)rng.below(2);
    adj.assign(n, std::vector<int>(n, 0));
    for (int x = 1; x < n; ++x) {                 // a spanning tree, then extra edges
        const int y = (int)rng.below((uint32_t)x);
        adj[x][y] = adj[y][x] = 1;
    }
    for (int t = 0; t < n; ++t) {
        const int x = (int)rng.below(n), y = (int)rng.below(n);
        if (x != y && rng.below(3) == 0) adj[x][y] = adj[y][x] = 1;
    }
    a.dom.resize(n);
    for (int x = 0; x < n; ++x) {
        const uint32_t sz = 40 + rng.below(400);
        std::vector<uint32_t> ids;
        uint32_t v = (uint32_t)lab[x];
        for (uint32_t i = 0; i < sz; ++i) {
            v += 2 * (1 + rng.below(3));          // ascending, parity = label
            ids.push_back(v);
        }
        a.dom[x] = ids;
    }
    a.off.assign(n, std::vector<std::vector<uint32_t>>(n));
    for (int b = 0; b < n; ++b)
        for (int x = 0; x < n; ++x) {
            if (!adj[b][x]) continue;
            std::vector<uint32_t> o{(uint32_t)a.pool_.size()};
            const uint32_t cx = (uint32_t)a.dom[x].size();
            for (size_t r = 0; r < a.dom[b].size(); ++r) {
                const uint32_t kind = rng.below(8);
                const uint32_t len = kind == 0 ? 0 : kind < 5 ? rng.below(12) : kind < 7 ? 40 + rng.below(120) : cx / 2;
                std::vector<uint32_t> row;
                for (uint32_t c = 0; c < cx; ++c)
                    if (rng.below(cx) < len) row.push_back(c);
                a.pool_.insert(a.pool_.end(), row.begin(), row.end());
                o.push_back((uint32_t)a.pool_.size());
            }
            a.off[b][x] = std::move(o);
        }
    a.ord.assign(n, std::vector<uint32_t>());
    a.bn.assign(n, std::vector<std::vector<uint32_t>>(n));
    a.sb.assign(n, std::vector<uint32_t>(n, 0));
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
            for (int j = 0; j < d; ++j) {
                if (adj[a.ord[u][j]][a.ord[u][d]]) a.bn[u][d].push_back(a.ord[u][j]);
                if (lab[a.ord[u][j]] == lab[a.ord[u][d]]) a.sb[u][d] |= 1u << j;
            }
    }
}

// Plain recursive existence search over the same tables: candidates of depth
// d are the entries of the first backward row present in every other one,
// entered in ascending order (recorded in seq).
bool plain_exists(const HostAcc &a, uint32_t u, std::vector<uint32_t> &idx, uint32_t d,
                  std::vector<std::pair<uint32_t, uint32_t>> &seq) {
    const uint32_t n = a.n;
    if (d == n) return true;
    const uint32_t x = a.ord[u][d];
    uint32_t lo0, hi0;
    a.row(u, d, 0, idx[a.bn[u][d][0]], lo0, hi0);
    for (uint32_t p = lo0; p < hi0; ++p) {
        const uint32_t c = a.pool_[p];
        bool ok = true;
        for (uint32_t k = 1; k < a.bn[u][d].size() && ok; ++k) {
            uint32_t lo, hi;
            a.row(u, d, k, idx[a.bn[u][d][k]], lo, hi);
            bool in = false;
            for (uint32_t q = lo; q < hi; ++q) in |= a.pool_[q] == c;
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

int random_mode(uint64_t seed, uint32_t instances, uint32_t donate_every, uint32_t fanout) {
    Rng rng{seed};
    uint64_t checked = 0, mismatches = 0, donations = 0, found_total = 0;
    tfsm_wave::WaveStack<16> st;
    HostWave w;
    for (uint32_t it = 0; it < instances; ++it) {
        HostAcc a;
        std::vector<std::vector<int>> adj;
        random_instance(rng, a, adj);
        for (uint32_t u = 0; u < a.n; ++u)
            for (uint32_t ci = 0; ci < a.dom[u].size(); ci += 1 + rng.below(4)) {
                std::vector<uint32_t> idx(a.n, 0);
                idx[u] = ci;
                std::vector<std::pair<uint32_t, uint32_t>> want_seq, got_seq;
                const bool want = plain_exists(a, u, idx, 1, want_seq);
                std::vector<Piece> queue;
                HostHooks h;
                if (!donate_every) h.seq = &got_seq;      // without pieces the visit order is the plain one
                h.donate_every = donate_every;
                h.fanout = fanout;
                h.u = u;
                h.a = &a;
                h.queue = &queue;
                tfsm_wave::wave_start_root<16>(a, u, ci, st);
                uint32_t r = tfsm_wave::wave_run<16>(w, a, u, 1, st, h);
                bool found = r == tfsm_wave::kWaveFound;
                for (size_t q = 0; q < queue.size(); ++q) {
                    const Piece p = queue[q];
                    tfsm_wave::wave_start_piece<16>(a, u, p.floor, p.prefix, p.lo, p.hi, p.cpos, p.cmask, st);
                    r = tfsm_wave::wave_run<16>(w, a, u, p.floor, st, h);
                    found |= r == tfsm_wave::kWaveFound;
                }
                donations += h.donations;
                found_total += found;
                ++checked;
                const bool bad = found != want || (!donate_every && got_seq != want_seq);
                if (bad && ++mismatches <= 10)
                    std::fprintf(stderr, "mismatch: instance %u u %u ci %u got %d want %d (entered %zu vs %zu)\n", it,
                                 u, ci, found, want, got_seq.size(), want_seq.size());
            }
    }
    std::printf("random checked %llu mismatches %llu donations %llu found %llu\n", (unsigned long long)checked,
                (unsigned long long)mismatches, (unsigned long long)donations, (unsigned long long)found_total);
    return mismatches ? 1 : 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc == 6 && std::string(argv[1]) == "--random")
        return random_mode(std::strtoull(argv[2], nullptr, 10), (uint32_t)std::atoi(argv[3]),
                           (uint32_t)std::atoi(argv[4]), (uint32_t)std::atoi(argv[5]));
    if (argc != 4) {
        std::fprintf(stderr, "usage: wave_dfs_check SPEC DONATE_EVERY FANOUT\n");
        return 64;
    }
    const uint32_t donate_every = (uint32_t)std::atoi(argv[2]), fanout = (uint32_t)std::atoi(argv[3]);
    uint64_t donations = 0;
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
    uint64_t checked = 0, mismatches = 0, patterns = 0;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        int nv, ne;
        if (!(ls >> nv)) continue;
        std::vector<int32_t> vl(nv);
        for (auto &x : vl) ls >> x;
        ls >> ne;
        std::vector<int32_t> ed(3 * ne);
        for (auto &x : ed) ls >> x;
        orc_eval *e = orc_eval_pattern(g, nv, vl.data(), ne, ed.data(), 0, nullptr, nullptr, freq, 1);
        if (!e) continue;
        if (orc_eval_decided_by(e) != 2 || nv > 16) {
            orc_eval_free(e);
            continue;
        }
        ++patterns;
        HostAcc a;
        a.n = (uint32_t)nv;
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
        for (int i = 0; i < ne; ++i) adj[ed[3 * i]][ed[3 * i + 1]] = adj[ed[3 * i + 1]][ed[3 * i]] = 1;
        a.ord.assign(nv, std::vector<uint32_t>(nv));
        a.bn.assign(nv, std::vector<std::vector<uint32_t>>(nv));
        a.sb.assign(nv, std::vector<uint32_t>(nv, 0));
        for (int u = 0; u < nv; ++u) {
            orc_eval_order(e, u, a.ord[u].data());
            for (int d = 0; d < nv; ++d)
                for (int j = 0; j < d; ++j) {
                    if (adj[a.ord[u][j]][a.ord[u][d]]) a.bn[u][d].push_back(a.ord[u][j]);
                    if (vl[a.ord[u][j]] == vl[a.ord[u][d]]) a.sb[u][d] |= 1u << j;
                }
        }
        tfsm_wave::WaveStack<16> st;
        HostWave w;
        for (int u = 0; u < nv; ++u) {
            std::vector<uint8_t> states(a.dom[u].size());
            orc_eval_states(e, u, states.data(), (uint32_t)states.size());
            const size_t sz = states.size();
            const size_t step = samples && sz > samples ? sz / samples : 1;
            for (size_t ci = 0; ci < sz; ci += step) {
                std::vector<Piece> queue;
                HostHooks h;
                h.donate_every = donate_every;
                h.fanout = fanout;
                h.u = (uint32_t)u;
                h.a = &a;
                h.queue = &queue;
                tfsm_wave::wave_start_root<16>(a, (uint32_t)u, (uint32_t)ci, st);
                uint32_t r = tfsm_wave::wave_run<16>(w, a, (uint32_t)u, 1, st, h);
                bool found = r == tfsm_wave::kWaveFound, all_exhausted = r == tfsm_wave::kWaveExhausted;
                auto check_embedding = [&]() {
                    // injective, edge-consistent and mapping u to ci
                    if (st.idx[u] != ci) violation("embedding root", st.idx[u], ci);
                    for (int i = 0; i < nv; ++i)
                        for (int j = i + 1; j < nv; ++j)
                            if (a.dom[i][st.idx[i]] == a.dom[j][st.idx[j]]) violation("not injective", i, j);
                    for (int k = 0; k < ne; ++k) {
                        const uint32_t pa = ed[3 * k], pb = ed[3 * k + 1];
                        const std::vector<uint32_t> &o = a.off[pa][pb];
                        bool hit = false;
                        for (uint32_t q = o[st.idx[pa]]; q < o[st.idx[pa] + 1]; ++q) hit |= a.pool_[q] == st.idx[pb];
                        if (!hit) violation("embedding edge missing", pa, pb);
                    }
                };
                if (found) check_embedding();
                for (size_t q = 0; q < queue.size(); ++q) {
                    const Piece p = queue[q];
                    if (p.prefix[0] != ci) violation("piece root", p.prefix[0], ci);
                    tfsm_wave::wave_start_piece<16>(a, (uint32_t)u, p.floor, p.prefix, p.lo, p.hi, p.cpos, p.cmask, st);
                    r = tfsm_wave::wave_run<16>(w, a, (uint32_t)u, p.floor, st, h);
                    if (r == tfsm_wave::kWaveFound) {
                        found = true;
                        check_embedding();
                    }
                    all_exhausted &= r == tfsm_wave::kWaveExhausted;
                }
                donations += h.donations;
                const uint8_t want = states[ci];
                const uint8_t got = found ? 1 : all_exhausted ? 2 : 0;
                ++checked;
                if (got != want) {
                    ++mismatches;
                    if (mismatches <= 10)
                        std::fprintf(stderr, "mismatch: pattern %llu u %d ci %zu got %d want %d\n",
                                     (unsigned long long)patterns, u, ci, got, want);
                }
            }
        }
        orc_eval_free(e);
    }
    orc_graph_free(g);
    std::printf("patterns %llu checked %llu mismatches %llu donations %llu\n", (unsigned long long)patterns,
                (unsigned long long)checked, (unsigned long long)mismatches, (unsigned long long)donations);
    return mismatches ? 1 : 0;
}
