// Size-2 Fraction-Score sums; behaviour cited in fraction.h.
#include "fraction.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>

namespace tfsm {

namespace {

// A score (a double in [2^-32, 1]) in units of 2^-96: exact.
unsigned __int128 fixed96(double f) {
    if (f == 0.0) return 0;
    int e = 0;
    const double m = std::frexp(f, &e);                 // f = m * 2^e, m in [0.5, 1)
    const unsigned long long mi = (unsigned long long)std::ldexp(m, 53);   // exact: 53 significant bits
    const int sh = e - 53 + 96;
    return sh >= 0 ? (unsigned __int128)mi << sh : (unsigned __int128)(mi >> -sh);
}

double rounded96(unsigned __int128 x) { return std::ldexp((double)x, -96); }

}  // namespace

std::vector<std::pair<double, double>> edge_fraction(const DataGraph &g, const std::vector<double> &score) {
    const size_t L = g.vlabel_values.size();
    std::map<std::tuple<int, int, int>, size_t> index;         // dense (la, el, lb) -> entry
    for (size_t i = 0; i < g.edge_stats.size(); ++i) {
        const EdgeStat &e = g.edge_stats[i];
        const int la = g.dense_vlabel(e.la), el = g.dense_elabel(e.el), lb = g.dense_vlabel(e.lb);
        if (la >= 0 && el >= 0 && lb >= 0) index[std::make_tuple(la, el, lb)] = i;
    }
    std::vector<unsigned __int128> sa(g.edge_stats.size(), 0), sb(g.edge_stats.size(), 0);
    std::vector<std::pair<uint32_t, uint32_t>> keys;            // (edge label, neighbour label) of v
    for (uint32_t v = 0; v < g.n; ++v) {
        keys.clear();
        for (uint32_t k = g.offsets[v]; k < g.offsets[v + 1]; ++k) keys.emplace_back(g.elab[k], g.vlabel[g.nbrs[k]]);
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        const int lv = (int)g.vlabel[v];
        for (auto &kw : keys) {
            const int el = (int)kw.first, lw = (int)kw.second;
            const double s = score[(size_t)v * L + lw];
            if (lv <= lw) {                                     // v in candA of (lv, el, lw)
                auto it = index.find(std::make_tuple(lv, el, lw));
                if (it != index.end()) sa[it->second] += fixed96(s);
            }
            if (lv >= lw) {                                     // v in candB of (lw, el, lv)
                auto it = index.find(std::make_tuple(lw, el, lv));
                if (it != index.end()) sb[it->second] += fixed96(s);
            }
        }
    }
    std::vector<std::pair<double, double>> out(g.edge_stats.size());
    for (size_t i = 0; i < out.size(); ++i) out[i] = {rounded96(sa[i]), rounded96(sb[i])};
    return out;
}

}  // namespace tfsm
