// Dumps what the reference's Fraction-Score loader computes (lyuheng/T-FSM
// fraction-score/graph.h): the per-vertex scores of compute_fraction_score
// (graph.h:895-922) and the size-2 entries that survive its Fraction-Score
// filter (graph.h:789-826), with the two sums that filter compares.
// Built by oracle/ref.mk from the reference sources where they lie (this file
// is the only translation unit of ours); the output becomes the fixtures that
// pin the CPU oracle's Fraction-Score functions.
//
//   ref_fraction_dump GRAPH.lg FREQ OUT
// OUT lines (doubles as C99 hex floats, so every bit is kept):
//   N <vertices>
//   S <v> <label> <score>                  one per (pruned vertex, label) entry
//   E <la> <el> <lb> <|A|> <|B|> <sum_A> <sum_B>   surviving size-2 entries,
//     sums over candA of score(v, lb) and over candB of score(v, la), added in
//     the reference's own iteration order of the sets
#include "gmatch.h"

#include <algorithm>
#include <cstdio>
#include <map>
#include <tuple>
#include <vector>

// the static member the reference's growth code reads (defined by its
// fraction-score/run.cpp, which is not linked here; growth is not used)
int Settings::maxNumNodes = -1;

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: ref_fraction_dump GRAPH FREQ OUT\n");
        return 64;
    }
    const int freq = std::atoi(argv[2]);
    Graph graph(freq);
    grami.nsupport_ = freq;
    graph.load_graph(grami.pruned_graph, argv[1], " ", 1);
    Graph &g = grami.pruned_graph;
    FILE *out = std::fopen(argv[3], "w");
    if (!out) return 65;
    std::fprintf(out, "N %u\n", (unsigned)g.size());
    for (ui v = 0; v < g.size(); ++v) {
        std::map<vLabel, double> sorted(g.vtx_frac[v].begin(), g.vtx_frac[v].end());
        for (auto &kv : sorted) std::fprintf(out, "S %u %d %a\n", (unsigned)v, (int)kv.first, kv.second);
    }
    std::vector<std::tuple<int, int, int, size_t, size_t, double, double>> edges;
    for (auto &kv : g.hashedEdges) {
        double a = 0, b = 0;
        for (VertexID v : kv.second.candA) a += g.vtx_frac[v][kv.first.to_label];
        for (VertexID v : kv.second.candB) b += g.vtx_frac[v][kv.first.from_label];
        edges.emplace_back((int)kv.first.from_label, (int)kv.first.edge_label, (int)kv.first.to_label,
                           kv.second.candA.size(), kv.second.candB.size(), a, b);
    }
    std::sort(edges.begin(), edges.end());
    for (auto &e : edges)
        std::fprintf(out, "E %d %d %d %zu %zu %a %a\n", std::get<0>(e), std::get<1>(e), std::get<2>(e), std::get<3>(e),
                     std::get<4>(e), std::get<5>(e), std::get<6>(e));
    std::fclose(out);
    return 0;
}
