// Fraction-Score support of the size-2 level (the reference's fraction-score
// build filters its size-2 entries by summed scores, fraction-score/graph.h:
// 803-824, after computing the scores, graph.h:895-922).
#pragma once

#include <cstdint>
#include <vector>

#include "graph.h"

namespace tfsm {

// For each entry of g.edge_stats (la, el, lb): the sum over candA (vertices of
// label la with an el-edge to a label-lb vertex) of score(v, lb), and over
// candB (label lb, an el-edge to label la) of score(v, la).  score is the
// library's table (tfsm_fraction_scores: n x dense labels).  Each sum is the
// double nearest the exact sum (ties to even), as the library's pattern
// supports.
std::vector<std::pair<double, double>> edge_fraction(const DataGraph &g, const std::vector<double> &score);

}  // namespace tfsm
