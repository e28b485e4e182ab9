// `run`: the miner's command line.
//
//   run -file G.lg -freq F [-thread T] [-maxNodes K]
//       [-gpu plain|batched] [-support capped|exact|fraction] [-learned RANKER]
//       [-tune TUNING.json] [-out grami_results.txt] [-stats stats.json]
//       [-csrCache 1] [-pushdown 1]
//
// Same flags and console summary as the reference (parallel/run.cpp:11-82);
// flags are read by name in any order (the reference reads argv[2] as the
// file).  -learned takes a ranker weight file written by the ranker trainer,
// -tune a search-kernel shape written by the autotuner.  -csrCache 1 loads
// the pruned graph from a binary cache beside the file (`<file>.f<freq>.csr`,
// written on the first run and checked against the file's bytes).
// -pushdown 1 turns on push-down pruning from cached frequent sub-patterns
// (the reference's decompose.h / comper.h:1582-1612 cache).  -support
// fraction mines with Fraction-Score support (the reference's fraction-score
// build, same flags and output).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "csr_cache.h"
#include "graph.h"
#include "miner.h"
#include "ranker.h"
#include "tuning.h"

namespace {

const char *flag(int argc, char **argv, const char *name) {
    for (int i = 1; i + 1 < argc; ++i)
        if (std::strcmp(argv[i], name) == 0) return argv[i + 1];
    return nullptr;
}

}  // namespace

int main(int argc, char **argv) {
    using clk = std::chrono::steady_clock;
    const char *file = flag(argc, argv, "-file");
    const char *freq_s = flag(argc, argv, "-freq");
    if (!file || !freq_s) {
        std::cerr << "usage: run -file G.lg -freq F [-thread T] [-maxNodes K] [-gpu plain|batched]"
                     " [-support capped|exact|fraction] [-learned RANKER] [-tune TUNING] [-out PATH] [-stats PATH]\n";
        return 2;
    }
    tfsm::MinerOptions opt;
    opt.freq = std::atoi(freq_s);
    if (opt.freq < 1) { std::cerr << "run: -freq must be at least 1\n"; return 2; }
    if (const char *s = flag(argc, argv, "-thread")) opt.threads = std::max(1, std::atoi(s));
    if (const char *s = flag(argc, argv, "-maxNodes")) opt.max_nodes = std::atoi(s);
    if (const char *s = flag(argc, argv, "-gpu")) opt.batched = std::strcmp(s, "batched") == 0;
    if (const char *s = flag(argc, argv, "-support")) {
        if (std::strcmp(s, "capped") == 0) opt.mode = TFSM_MODE_CAPPED;
        else if (std::strcmp(s, "exact") == 0) opt.mode = TFSM_MODE_EXACT;
        else if (std::strcmp(s, "fraction") == 0) opt.mode = TFSM_MODE_FRACTION;
        else { std::cerr << "run: -support must be capped, exact or fraction\n"; return 2; }
    }
    if (const char *s = flag(argc, argv, "-pushdown")) opt.pushdown = std::strcmp(s, "0") != 0;
    const char *ranker_path = flag(argc, argv, "-learned");
    if (const char *t = flag(argc, argv, "-tune")) {
        std::string err;
        if (!tfsm::load_tuning_file(t, opt.tuning, err)) {
            std::cerr << "run: " << err << "\n";
            return 2;
        }
    }
    std::string out = flag(argc, argv, "-out") ? flag(argc, argv, "-out") : "./grami_results.txt";
    const char *stats_path = flag(argc, argv, "-stats");

    auto t1 = clk::now();
    tfsm::DataGraph g;
    try {
        const char *cache = flag(argc, argv, "-csrCache");
        if (cache && std::strcmp(cache, "0") != 0) g = tfsm::load_graph_cached(file, opt.freq);
        else g = tfsm::prune_graph(tfsm::read_lg(file), opt.freq);
    } catch (const std::exception &e) {
        std::cerr << e.what() << "\n";
        return 1;
    }
    tfsm_csr csr = tfsm::csr_view(g);
    tfsm_ctx *ctx = nullptr;
    int rc = tfsm_ctx_create(&csr, &ctx);
    if (rc != TFSM_OK) {
        std::cerr << "run: GPU context: " << tfsm_status_string(rc) << "\n";
        return 1;
    }
    if (ranker_path) {
        rc = tfsm::load_ranker_file(ctx, ranker_path);
        if (rc != TFSM_OK) {
            std::cerr << "run: cannot install ranker " << ranker_path << ": " << tfsm_status_string(rc) << " "
                      << tfsm_last_error(ctx) << "\n";
            tfsm_ctx_destroy(ctx);
            return 1;
        }
        opt.learned = true;
    }
    auto t2 = clk::now();
    std::vector<tfsm::FoundPattern> found;
    tfsm::MinerStats st;
    std::string err;
    rc = tfsm::mine(g, ctx, opt, found, st, err);
    auto t3 = clk::now();
    if (rc != 0) {
        std::cerr << "run: mining failed: " << tfsm_status_string(rc) << " " << err << "\n";
        tfsm_ctx_destroy(ctx);
        return 1;
    }
    tfsm::write_results(found, out);
    if (st.capped_vertices)
        std::fprintf(stderr, "run: warning: pattern growth stops at %d vertices (the device's pattern limit)\n",
                     st.capped_vertices);
    if (st.dropped_large)
        std::fprintf(stderr, "run: warning: %llu candidate patterns over %d edges were not evaluated\n",
                     (unsigned long long)st.dropped_large, TFSM_MAX_PATTERN_EDGES);
    auto sec =[](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    std::printf("[TIME] Loading Graph Time: %.3f s\n", sec(t1, t2));
    std::printf("[TIME] Mining Time: %.3f s\n", sec(t2, t3));
    std::printf("[TIME] Total Elapsed Time: %.3f s\n", sec(t1, t3));
    std::printf("[INFO] # Frequent Patterns: %zu\n", found.size());
    if (stats_path) {
        FILE *f = std::fopen(stats_path, "w");
        if (f) {
            std::fprintf(f,
                         "{\"mining_s\": %.6f, \"evaluated\": %llu, \"frequent\": %llu, \"verdicts\": %llu, "
                         "\"searches\": %llu, \"eval_calls\": %llu, \"filter_ms\": %.4f, \"table_ms\": %.4f, "
                         "\"search_ms\": %.4f, \"search_launches\": %llu, \"bytes_table\": %llu, \"bytes_search\": %llu, "
                         "\"eval_wall_ms\": %.3f, \"growth_wait_ms\": %.3f, \"levels\": [",
                         st.mining_s, (unsigned long long)st.evaluated, (unsigned long long)st.frequent,
                         (unsigned long long)st.verdicts, (unsigned long long)st.searches,
                         (unsigned long long)st.eval_calls, st.gpu_ms[0], st.gpu_ms[1], st.gpu_ms[2],
                         (unsigned long long)st.search_launches, (unsigned long long)st.bytes_table,
                         (unsigned long long)st.bytes_search, st.eval_wall_ms, st.growth_wait_ms);
            for (size_t i = 0; i < st.levels.size(); ++i) {
                const tfsm::LevelStats &l = st.levels[i];
                std::fprintf(f,
                             "%s{\"vertices\": %u, \"evaluated\": %llu, \"frequent\": %llu, \"wall_ms\": %.3f, "
                             "\"prep_ms\": %.3f, \"eval_ms\": %.3f, \"post_ms\": %.3f}",
                             i ? ", " : "", l.size, (unsigned long long)l.evaluated, (unsigned long long)l.frequent,
                             l.wall_ms, l.prep_ms, l.eval_ms, l.post_ms);
            }
            std::fprintf(f, "]}\n");
            std::fclose(f);
        }
    }
    tfsm_ctx_destroy(ctx);
    return 0;
}
