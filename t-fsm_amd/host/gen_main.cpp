// tfsm_gen: write a named synthetic workload as an `.lg` file.
//
//   tfsm_gen <gnm2k|powerlaw100k|rmat20> <out.lg> [seed]
#include "gen.h"

#include <cstdlib>
#include <iostream>

int main(int argc, char **argv) {
    if (argc < 3) {
        std::cerr << "usage: tfsm_gen <gnm2k|powerlaw100k|rmat20> <out.lg> [seed]\n";
        return 2;
    }
    std::string name = argv[1];
    uint64_t seed = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : tfsm::default_seed(name);
    try {
        tfsm::LabeledGraph g = tfsm::gen_named(name, seed);
        tfsm::write_lg(g, argv[2]);
        std::cout << name << " seed=" << seed << " n=" << g.n << " m=" << g.edges.size() << "\n";
    } catch (const std::exception &e) {
        std::cerr << "tfsm_gen: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
