// Dumps, for a list of patterns, what the reference's own matching engine
// computes at pattern activation (lyuheng/T-FSM parallel/gmatch.h, driven as
// parallel/comper.h:1655-1685 drives it, without push-down pruning):
//   the filtered domains after DPisoFilter (gmatch.h:417-472),
//   the auxiliary tables of buildTable (gmatch.h:995-1130),
//   the GQL order and backward neighbours per start vertex
//   (generateGQLQueryPlan / generateBN, gmatch.h:692-727, 803-828).
// Built by tests/golden/make_gmatch_fixtures.py from the reference sources
// where they lie (this file is the only translation unit of ours); the
// output becomes the fixtures that pin the CPU oracle.
//
//   ref_gmatch_dump GRAPH.lg FREQ SPEC OUT
// SPEC: one pattern per line, as a DFS code key "a,b,la,el,lb;..." (original
// labels).  OUT gets, per pattern:
//   P <key>
//   F <0|1>                    DPisoFilter's result with support 1
//   D <u> <n> <ids...>         C(u), ascending pruned-graph vertex ids
//   T <a> <b> <rows> <entries> <offsets...> <indices...>   edge_matrix[a][b]
//   O <u> <order...>           matching order starting at u
//   B <u> <d> <bn...>          backward neighbours of depth d of that order
#include "gmatch.h"

#include <cstdio>
#include <fstream>
#include <sstream>

// the static member the reference's growth code reads (defined by its
// parallel/run.cpp, which is not linked here; growth is not used)
int Settings::maxNumNodes = -1;

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: ref_gmatch_dump GRAPH FREQ SPEC OUT\n");
        return 64;
    }
    const int freq = std::atoi(argv[2]);
    Graph graph(freq);
    grami.nsupport_ = freq;
    graph.load_graph(grami.pruned_graph, argv[1], " ", 1);
    std::ifstream spec(argv[3]);
    FILE *out = std::fopen(argv[4], "w");
    if (!out) return 65;
    GMatchEngine engine;
    std::string key;
    while (std::getline(spec, key)) {
        if (key.empty()) continue;
        Pattern p(true);
        std::stringstream ks(key);
        std::string code;
        while (std::getline(ks, code, ';')) {
            if (code.empty()) continue;
            int a, b, la, el, lb;
            if (std::sscanf(code.c_str(), "%d,%d,%d,%d,%d", &a, &b, &la, &el, &lb) != 5) return 66;
            dfs_code_t d((VertexID)a, (VertexID)b, la, el, lb);
            p.extend(d);
        }
        const ui nv = (ui)p.size();
        p.non_candidates.resize(nv);
        engine.set(&grami.pruned_graph, &p);
        const bool ok = engine.DPisoFilter(false, 1);
        std::fprintf(out, "P %s\nF %d\n", key.c_str(), ok ? 1 : 0);
        if (!ok) continue;
        for (ui u = 0; u < nv; ++u) {
            Domain &dm = p.get_cands()[u];
            std::fprintf(out, "D %u %d", u, dm.size());
            for (VertexID v : dm.candidate) std::fprintf(out, " %u", v);
            std::fprintf(out, "\n");
        }
        Edges ***em = new Edges **[nv];
        for (ui i = 0; i < nv; ++i) {
            em[i] = new Edges *[nv];
            for (ui j = 0; j < nv; ++j) em[i][j] = NULL;
        }
        engine.buildTable(em);
        for (ui a = 0; a < nv; ++a)
            for (ui b = 0; b < nv; ++b) {
                Edges *t = em[a][b];
                if (!t) continue;
                std::fprintf(out, "T %u %u %u %u", a, b, t->vertex_count_, t->edge_count_);
                for (ui i = 0; i <= t->vertex_count_; ++i) std::fprintf(out, " %u", t->offset_[i]);
                for (ui i = 0; i < t->edge_count_; ++i) std::fprintf(out, " %u", t->edge_[i]);
                std::fprintf(out, "\n");
                delete t;
            }
        for (ui i = 0; i < nv; ++i) delete[] em[i];
        delete[] em;
        for (ui u = 0; u < nv; ++u) {
            ui *order = NULL, *bn_count = NULL, **bn = NULL;
            engine.generateGQLQueryPlan((int)u, order);
            engine.generateBN(bn_count, bn, order);
            std::fprintf(out, "O %u", u);
            for (ui d = 0; d < nv; ++d) std::fprintf(out, " %u", order[d]);
            std::fprintf(out, "\n");
            for (ui d = 1; d < nv; ++d) {
                std::fprintf(out, "B %u %u", u, d);
                for (ui k = 0; k < bn_count[d]; ++k) std::fprintf(out, " %u", bn[d][k]);
                std::fprintf(out, "\n");
            }
            for (ui d = 0; d < nv; ++d) delete[] bn[d];
            delete[] bn;
            delete[] bn_count;
            delete[] order;
        }
        p.get_cands().clear();
    }
    std::fclose(out);
    return 0;
}
