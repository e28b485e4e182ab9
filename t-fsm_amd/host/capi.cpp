// libtfsm_host.so: C entry points over the host side (loader, pattern growth,
// miner) for Python callers (tests, bench).  The GPU evaluation itself is
// reached only through libtfsm_gpu.so's C ABI.
#include <cstring>
#include <string>
#include <vector>

#include "graph.h"
#include "csr_cache.h"
#include "miner.h"
#include "pushdown.h"
#include "pattern.h"
#include "ranker.h"
#include "tuning.h"

extern "C" {

struct tfsmh_graph {
    tfsm::DataGraph g;
    std::string err;
};

static std::string g_err;

const char *tfsmh_last_error() { return g_err.c_str(); }

tfsmh_graph *tfsmh_load(const char *path, int freq) {
    try {
        auto *h = new tfsmh_graph;
        h->g = tfsm::prune_graph(tfsm::read_lg(path), freq);
        return h;
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}

// As tfsmh_load, through the binary cache beside the file (csr_cache.h);
// *hit (when given) is 1 when the cache was used.
tfsmh_graph *tfsmh_load_cached(const char *path, int freq, int *hit) {
    try {
        auto *h = new tfsmh_graph;
        bool used = false;
        h->g = tfsm::load_graph_cached(path, freq, &used);
        if (hit) *hit = used ? 1 : 0;
        return h;
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}

void tfsmh_free(tfsmh_graph *h) { delete h; }

// n, num_adj, num_vlabels, num_elabels, num_edge_stats
void tfsmh_info(const tfsmh_graph *h, uint64_t *out5) {
    out5[0] = h->g.n;
    out5[1] = h->g.nbrs.size();
    out5[2] = h->g.vlabel_values.size();
    out5[3] = h->g.elabel_values.size();
    out5[4] = h->g.edge_stats.size();
}

// Dense CSR plus label value tables; arrays sized from tfsmh_info.
void tfsmh_csr(const tfsmh_graph *h, uint32_t *offsets, uint32_t *nbrs, uint32_t *elabels, uint32_t *vlabels,
               int32_t *vlabel_values, int32_t *elabel_values) {
    const auto &g = h->g;
    std::memcpy(offsets, g.offsets.data(), g.offsets.size() * 4);
    if (!g.nbrs.empty()) {
        std::memcpy(nbrs, g.nbrs.data(), g.nbrs.size() * 4);
        std::memcpy(elabels, g.elab.data(), g.elab.size() * 4);
    }
    if (g.n) std::memcpy(vlabels, g.vlabel.data(), g.n * 4);
    std::memcpy(vlabel_values, g.vlabel_values.data(), g.vlabel_values.size() * 4);
    std::memcpy(elabel_values, g.elabel_values.data(), g.elabel_values.size() * 4);
}

// 5 int32 per entry: la, el, lb, |A|, |B|
void tfsmh_edge_stats(const tfsmh_graph *h, int32_t *out) {
    size_t k = 0;
    for (auto &e : h->g.edge_stats) {
        out[k++] = e.la; out[k++] = e.el; out[k++] = e.lb;
        out[k++] = (int32_t)e.na; out[k++] = (int32_t)e.nb;
    }
}

int tfsmh_load_ranker(tfsm_ctx *ctx, const char *path) { return tfsm::load_ranker_file(ctx, path); }

// Search-kernel shape from a tuning file into *opt; 0 or -1 (message in tfsmh_last_error).
int tfsmh_load_tuning(const char *path, tfsm_eval_options *opt) {
    std::string err;
    if (!tfsm::load_tuning_file(path, *opt, err)) {
        g_err = err;
        return -1;
    }
    return 0;
}

int tfsmh_ctx_create(const tfsmh_graph *h, tfsm_ctx **out) {
    tfsm_csr c = tfsm::csr_view(h->g);
    return tfsm_ctx_create(&c, out);
}

// Runs the miner.  stats_out (16 doubles): mining_s, evaluated, frequent,
// verdicts, searches, eval_calls, filter_ms, table_ms, search_ms,
// search_launches, bytes_table, bytes_search, levels, max pattern size,
// the sum of filtered domain sizes over the evaluated patterns, push-down
// entries used.
// results_path: grami_results.txt format; keys_path: one line per frequent
// pattern "<dfs key>\t<support lower>\t<support upper>" (either may be NULL).
// tuning (nullable): search-kernel shape for every evaluation call.
int64_t tfsmh_mine(const tfsmh_graph *h, tfsm_ctx *ctx, int freq, int max_nodes, int threads, int batched, int mode,
                   int learned, int pushdown, const char *results_path, const char *keys_path, double *stats_out,
                   const tfsm_eval_options *tuning) {
    tfsm::MinerOptions opt;
    if (tuning) opt.tuning = *tuning;
    opt.freq = freq;
    opt.max_nodes = max_nodes;
    opt.threads = threads;
    opt.batched = batched != 0;
    opt.mode = mode;
    opt.learned = learned != 0;
    opt.pushdown = pushdown != 0;
    std::vector<tfsm::FoundPattern> found;
    tfsm::MinerStats st;
    std::string err;
    int rc = tfsm::mine(h->g, ctx, opt, found, st, err);
    if (rc != 0) {
        g_err = err;
        return rc;
    }
    if (results_path) tfsm::write_results(found, results_path);
    if (keys_path) {
        FILE *f = std::fopen(keys_path, "w");
        if (f) {
            // Fraction-Score runs add the support as a fourth column (hex float: every bit)
            for (auto &fp : found) {
                std::fprintf(f, "%s\t%u\t%u", fp.p.key().c_str(), fp.support_lower, fp.support_upper);
                if (mode == TFSM_MODE_FRACTION) std::fprintf(f, "\t%a", fp.fraction);
                std::fprintf(f, "\n");
            }
            std::fclose(f);
        }
    }
    if (stats_out) {
        double v[20] = {st.mining_s, (double)st.evaluated, (double)st.frequent, (double)st.verdicts, (double)st.searches,
                        (double)st.eval_calls, st.gpu_ms[0], st.gpu_ms[1], st.gpu_ms[2], (double)st.search_launches,
                        (double)st.bytes_table, (double)st.bytes_search, (double)st.levels.size(),
                        st.levels.empty() ? 0.0 : (double)st.levels.back().size, (double)st.domain_sum,
                        (double)st.pushdown_sources, (double)st.filter_passes, (double)st.device_calls,
                        (double)st.graph_calls, 0.0};
        std::memcpy(stats_out, v, sizeof(v));
    }
    return (int64_t)found.size();
}

// Canonical rightmost extensions of one pattern (codes as 5 ints per edge),
// with the label pruning learnt from the graph's frequent size-2 patterns.
// Writes children back to back (each ne + 1 codes) into out (cap ints);
// returns the number of children or -1 when out is too small.
int tfsmh_extend(const tfsmh_graph *h, int max_nodes, const int32_t *code, int ne, int32_t *out, int cap) {
    try {
        tfsm::GrowthContext gc;
        gc.max_nodes = max_nodes;
        for (int32_t l : h->g.elabel_values) gc.freq_edge_labels.insert(l);
        for (auto &e : h->g.edge_stats) {
            gc.neighbor_labels[e.la].insert(e.lb);
            gc.neighbor_labels[e.lb].insert(e.la);
        }
        tfsm::Pattern p;
        for (int i = 0; i < ne; ++i)
            p.push(tfsm::DfsCode{code[5 * i], code[5 * i + 1], code[5 * i + 2], code[5 * i + 3], code[5 * i + 4]});
        std::vector<tfsm::Pattern> kids = tfsm::extend(p, gc);
        int k = 0;
        for (auto &c : kids) {
            if (k + 5 * c.ne() > cap) return -1;
            for (auto &d : c.code) {
                out[k++] = d.from; out[k++] = d.to; out[k++] = d.fl; out[k++] = d.el; out[k++] = d.tl;
            }
        }
        return (int)kids.size();
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

// Canonical form of a small labelled graph (push-down cache keys): n
// vertices with labels lab, edge labels el[a * n + b] (-1: none).  Writes the
// key (NUL-terminated, cap bytes) and order[k] = the vertex at canonical
// position k; returns the key length or -1.
int tfsmh_canonical(int n, const int32_t *lab, const int32_t *el, char *key, int cap, int32_t *order) {
    try {
        if (n < 1 || n > tfsm::kMaxPatternVertices) return -1;
        tfsm::Canon c = tfsm::canonical_form(n, lab, el);
        if ((int)c.key.size() + 1 > cap) return -1;
        std::memcpy(key, c.key.c_str(), c.key.size() + 1);
        for (int k = 0; k < n; ++k) order[k] = c.order[k];
        return (int)c.key.size();
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

// The push-down decomposition of a pattern (5 ints per code edge): per
// sub-pattern, its vertex count m then its m pattern vertices, back to back in
// out (cap ints).  Returns the number of sub-patterns or -1.
int tfsmh_decompose(const int32_t *code, int ne, int32_t *out, int cap) {
    try {
        tfsm::Pattern p;
        for (int i = 0; i < ne; ++i)
            p.push(tfsm::DfsCode{code[5 * i], code[5 * i + 1], code[5 * i + 2], code[5 * i + 3], code[5 * i + 4]});
        int k = 0, cnt = 0;
        for (const tfsm::SubPattern &s : tfsm::decompose(p)) {
            if (k + 1 + (int)s.verts.size() > cap) return -1;
            out[k++] = (int32_t)s.verts.size();
            for (int v : s.verts) out[k++] = v;
            ++cnt;
        }
        return cnt;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

// Minimum-DFS-code test of a code given as 5 ints per edge (from, to, fl, el, tl).
int tfsmh_is_min_code(const int32_t *code, int ne) {
    try {
        tfsm::Pattern p;
        for (int i = 0; i < ne; ++i)
            p.push(tfsm::DfsCode{code[5 * i], code[5 * i + 1], code[5 * i + 2], code[5 * i + 3], code[5 * i + 4]});
        return tfsm::is_min_code(p) ? 1 : 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

}  // extern "C"
