// Level-synchronous miner; behaviour cited in miner.h.
#include "miner.h"
#include "fraction.h"
#include "pushdown.h"
#include "worker_pool.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace tfsm {

tfsm_csr csr_view(const DataGraph &g) {
    tfsm_csr c;
    c.num_vertices = g.n;
    c.num_adj = g.nbrs.size();
    c.offsets = g.offsets.data();
    c.nbrs = g.nbrs.data();
    c.elabels = g.elab.data();
    c.vlabels = g.vlabel.data();
    c.num_vlabels = (uint32_t)g.vlabel_values.size();
    c.num_elabels = (uint32_t)g.elabel_values.size();
    return c;
}

namespace {

// largest level whose children are grown while the GPU evaluates it
constexpr size_t kSpeculativeGrowth = 4096;

struct Candidate {
    Pattern p;
    uint64_t parent_handle = 0;
};

// Dense-label form of a pattern; false when a label is absent from the pruned
// graph (then the pattern has no embedding and is infrequent).
bool to_device_pattern(const DataGraph &g, const Pattern &p, tfsm_pattern &out) {
    std::memset(&out, 0, sizeof(out));
    out.num_vertices = (uint32_t)p.nv();
    out.num_edges = (uint32_t)p.ne();
    for (int i = 0; i < p.nv(); ++i) {
        int d = g.dense_vlabel(p.vlab[i]);
        if (d < 0) return false;
        out.vlabel[i] = (uint32_t)d;
    }
    for (int e = 0; e < p.ne(); ++e) {
        int d = g.dense_elabel(p.code[e].el);
        if (d < 0) return false;
        out.edge_a[e] = (uint32_t)p.code[e].from;
        out.edge_b[e] = (uint32_t)p.code[e].to;
        out.edge_label[e] = (uint32_t)d;
    }
    return true;
}


}  // namespace

int mine(const DataGraph &g, tfsm_ctx *ctx, const MinerOptions &opt, std::vector<FoundPattern> &found,
         MinerStats &stats, std::string &err) {
    auto t0 = std::chrono::steady_clock::now();
    stats = MinerStats{};
    found.clear();
    GrowthContext gc;
    // growth stops at the device's pattern limits rather than failing the run:
    // at most TFSM_MAX_PATTERN_VERTICES vertices (forward edges), and children
    // with more than TFSM_MAX_PATTERN_EDGES edges are dropped below
    gc.max_nodes = opt.max_nodes < 0 || opt.max_nodes > TFSM_MAX_PATTERN_VERTICES ? TFSM_MAX_PATTERN_VERTICES
                                                                                     : opt.max_nodes;
    if (opt.max_nodes < 0 || opt.max_nodes > TFSM_MAX_PATTERN_VERTICES)
        stats.capped_vertices = TFSM_MAX_PATTERN_VERTICES;
    for (int32_t l : g.elabel_values) gc.freq_edge_labels.insert(l);
    // size-2 level (grami.h:83-169): every surviving endpoint-set entry is
    // frequent; with Fraction-Score support, those whose two summed scores
    // reach freq (fraction-score/graph.h:803-824)
    std::vector<std::pair<double, double>> frac2;
    if (opt.mode == TFSM_MODE_FRACTION) {
        const int64_t cells = tfsm_fraction_scores(ctx, nullptr, 0);
        if (cells < 0) {
            err = std::string("fraction scores: ") + tfsm_last_error(ctx);
            return (int)cells;
        }
        std::vector<double> score((size_t)cells);
        tfsm_fraction_scores(ctx, score.data(), (uint64_t)cells);
        frac2 = edge_fraction(g, score);
    }
    std::vector<Pattern> size2;
    for (size_t i = 0; i < g.edge_stats.size(); ++i) {
        const EdgeStat &e = g.edge_stats[i];
        if (!frac2.empty() && (frac2[i].first < (double)opt.freq || frac2[i].second < (double)opt.freq)) continue;
        gc.neighbor_labels[e.la].insert(e.lb);
        gc.neighbor_labels[e.lb].insert(e.la);
        FoundPattern f;
        f.p = single_edge(e.la, e.el, e.lb);
        f.support_lower = f.support_upper = std::min(e.na, e.nb);
        if (!frac2.empty()) f.fraction = std::min(frac2[i].first, frac2[i].second);
        size2.push_back(f.p);
        found.push_back(std::move(f));
    }
    stats.frequent = found.size();
    stats.levels.push_back(LevelStats{2, 0, found.size(), 0});

    std::vector<Candidate> level;
    {
        std::vector<std::vector<Pattern>> kids(size2.size());
        parallel_for(size2.size(), opt.threads, [&](size_t i) { kids[i] = extend(size2[i], gc); });
        for (auto &k : kids)
            for (auto &c : k) level.push_back(Candidate{std::move(c), 0});
    }
    // the size-2 level's wall time includes growing its children
    stats.levels[0].wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    tfsm_eval_options eo = opt.tuning;
    eo.mode = (uint32_t)opt.mode;
    eo.retain_domains = 1;
    eo.task_order = opt.learned ? TFSM_ORDER_LEARNED : TFSM_ORDER_REFERENCE;
    eo.trace = opt.trace ? 1 : 0;

    // retained-domain handles of the previous level (parents of `level`) and
    // of the level being evaluated: released on every exit, errors included
    std::vector<uint64_t> live_handles, new_handles;
    struct Release {
        tfsm_ctx *ctx;
        std::vector<uint64_t> *a, *b;
        ~Release() {
            for (auto *v : {a, b}) {
                for (uint64_t h : *v) tfsm_domain_release(ctx, h);
                v->clear();
            }
        }
    } release{ctx, &live_handles, &new_handles};
    // push-down pruning: frequent patterns' retained domains stay in the cache
    // (and alive) until the run ends
    PushdownCache cache;
    std::vector<uint64_t> cached_handles;
    Release release_cached{ctx, &cached_handles, &cached_handles};
    while (!level.empty()) {
        auto tl = std::chrono::steady_clock::now();
        const size_t n = level.size();
        std::vector<tfsm_pattern> dp(n);
        std::vector<uint64_t> parents(n);
        std::vector<char> ok(n);
        for (size_t i = 0; i < n; ++i) {
            ok[i] = to_device_pattern(g, level[i].p, dp[i]);
            parents[i] = level[i].parent_handle;
        }
        std::vector<tfsm_pattern_result> res(n);
        std::memset(res.data(), 0, n * sizeof(tfsm_pattern_result));
        // push-down entries of each candidate from the cached frequent patterns
        std::vector<std::vector<tfsm_pushdown>> push(n);
        if (opt.pushdown && !cache.entries.empty()) {
            parallel_for(n, opt.threads, [&](size_t i) {
                if (ok[i]) push[i] = cache.lookup(level[i].p, 0, parents[i]);
            });
            for (auto &v : push) stats.pushdown_sources += v.size();
        }
        // the next level's candidates are grown for every pattern of this
        // level while the GPU evaluates it (growth is pure: a pattern's
        // children depend on the pattern alone); only the frequent ones'
        // children are kept, so the candidates are those of growing after
        // the verdicts.  Once the verdicts are in, the patterns not yet grown
        // that are infrequent are skipped.  Levels of more than
        // kSpeculativeGrowth patterns are grown after the verdicts only: on a
        // level of 10^5 patterns of which 1% are frequent, growing them all
        // makes millions of children and costs more than the GPU
        std::vector<std::vector<Pattern>> kids(n);
        std::unique_ptr<std::atomic<uint8_t>[]> skip(new std::atomic<uint8_t>[n]);
        const bool speculate = n <= kSpeculativeGrowth;
        for (size_t i = 0; i < n; ++i) skip[i].store(speculate ? 0 : 1, std::memory_order_relaxed);
        Background grower(speculate ? n : 0, opt.threads, [&](size_t i) {
            if (!skip[i].load(std::memory_order_relaxed)) kids[i] = extend(level[i].p, gc);
        });
        const auto te = std::chrono::steady_clock::now();
        const double prep_ms = std::chrono::duration<double, std::milli>(te - tl).count();
        auto account = [&]() {
            tfsm_call_stats cs;
            tfsm_last_stats(ctx, &cs);
            stats.eval_calls++;
            stats.gpu_ms[0] += cs.filter_ms;
            stats.gpu_ms[1] += cs.table_ms;
            stats.gpu_ms[2] += cs.search_ms;
            stats.search_launches += cs.search_launches;
            stats.search_kernel_ms += cs.search_ms;
            stats.bytes_table += cs.table_bytes;
            stats.bytes_search += cs.search_bytes;
            stats.filter_passes += cs.filter_passes;
            stats.device_calls += cs.level_mode == TFSM_LEVEL_DEVICE || cs.level_mode == TFSM_LEVEL_GRAPH;
            stats.graph_calls += cs.level_mode == TFSM_LEVEL_GRAPH;
        };
        if (opt.batched) {
            std::vector<tfsm_pattern> sub;
            std::vector<uint64_t> subpar;
            std::vector<size_t> where;
            std::vector<tfsm_pushdown> subpush;
            for (size_t i = 0; i < n; ++i)
                if (ok[i]) {
                    for (tfsm_pushdown d : push[i]) {
                        d.pattern = (uint32_t)sub.size();
                        subpush.push_back(d);
                    }
                    sub.push_back(dp[i]);
                    subpar.push_back(parents[i]);
                    where.push_back(i);
                }
            std::vector<tfsm_pattern_result> subres(sub.size());
            if (!sub.empty()) {
                tfsm_eval_options lo = eo;
                lo.num_pushdown = (uint32_t)subpush.size();
                lo.pushdown = subpush.data();
                int rc = tfsm_eval_level(ctx, sub.data(), subpar.data(), (uint32_t)sub.size(), (uint32_t)opt.freq, &lo,
                                         subres.data());
                if (rc != TFSM_OK) { err = tfsm_last_error(ctx); return rc; }
                account();
            }
            for (size_t k = 0; k < where.size(); ++k) res[where[k]] = subres[k];
            for (auto &r : subres)
                if (r.domain_handle) new_handles.push_back(r.domain_handle);
        } else {
            for (size_t i = 0; i < n; ++i) {
                if (!ok[i]) continue;
                tfsm_eval_options po = eo;
                po.num_pushdown = (uint32_t)push[i].size();
                po.pushdown = push[i].data();
                int rc = tfsm_eval_pattern(ctx, &dp[i], parents[i], (uint32_t)opt.freq, &po, &res[i]);
                if (rc != TFSM_OK) { err = tfsm_last_error(ctx); return rc; }
                if (res[i].domain_handle) new_handles.push_back(res[i].domain_handle);
                account();
            }
        }
        const auto tp = std::chrono::steady_clock::now();
        const double eval_ms = std::chrono::duration<double, std::milli>(tp - te).count();
        stats.eval_wall_ms += eval_ms;
        for (size_t i = 0; i < n; ++i)
            if (!res[i].frequent) skip[i].store(1, std::memory_order_relaxed);
        // the parents' domains are no longer needed once all their children ran
        // (with push-down pruning they stay cached until the run ends)
        if (!opt.pushdown) {
            for (uint64_t h : live_handles) tfsm_domain_release(ctx, h);
            live_handles.swap(new_handles);
        } else {
            for (size_t i = 0; i < n; ++i)
                if (res[i].frequent && res[i].domain_handle) cache.insert(level[i].p, res[i].domain_handle);
            cached_handles.insert(cached_handles.end(), new_handles.begin(), new_handles.end());
        }
        new_handles.clear();

        LevelStats ls;
        ls.size = level.empty() ? 0 : (uint32_t)level[0].p.nv();
        ls.evaluated = n;
        std::vector<size_t> freq_idx;
        for (size_t i = 0; i < n; ++i) {
            stats.evaluated++;
            for (int u = 0; u < level[i].p.nv(); ++u) stats.domain_sum += res[i].domain_size[u];
            stats.verdicts += res[i].verdicts;
            stats.searches += res[i].searches;
            if (!res[i].frequent) continue;
            freq_idx.push_back(i);
        }
        ls.frequent = freq_idx.size();
        const auto tw = std::chrono::steady_clock::now();
        grower.wait();
        if (!speculate)
            parallel_for(freq_idx.size(), opt.threads, [&](size_t k) { kids[freq_idx[k]] = extend(level[freq_idx[k]].p, gc); });
        stats.growth_wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tw).count();
        std::vector<Candidate> next;
        for (size_t k = 0; k < freq_idx.size(); ++k) {
            const size_t i = freq_idx[k];
            FoundPattern f;
            f.p = level[i].p;
            f.support_lower = res[i].support_lower;
            f.support_upper = res[i].support_upper;
            f.fraction = res[i].fraction_support;
            found.push_back(std::move(f));
            for (auto &c : kids[i]) {
                if (c.ne() > TFSM_MAX_PATTERN_EDGES) {
                    stats.dropped_large++;
                    continue;
                }
                next.push_back(Candidate{std::move(c), res[i].domain_handle});
            }
        }
        stats.frequent += freq_idx.size();
        const auto tn = std::chrono::steady_clock::now();
        ls.wall_ms = std::chrono::duration<double, std::milli>(tn - tl).count();
        ls.prep_ms = prep_ms;
        ls.eval_ms = eval_ms;
        ls.post_ms = std::chrono::duration<double, std::milli>(tn - tp).count();
        stats.levels.push_back(ls);
        level.swap(next);
    }
    std::stable_sort(found.begin(), found.end(), [](const FoundPattern &a, const FoundPattern &b) {
        if (a.p.nv() != b.p.nv()) return a.p.nv() < b.p.nv();
        return a.p.key() < b.p.key();
    });
    stats.mining_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

void write_results(const std::vector<FoundPattern> &found, const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return;
    for (auto &fp : found) {
        std::fputs("=====================================\n", f);
        std::fputs(fp.p.to_lg_block().c_str(), f);
    }
    std::fclose(f);
}

}  // namespace tfsm
