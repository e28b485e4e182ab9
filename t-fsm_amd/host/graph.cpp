// `.lg` ingest and frequent-label pruning.  Behaviour cited per function in
// graph.h; this file is the product's own fast implementation (one pass over a
// memory image of the file, sort-based endpoint sets instead of hash sets).
#include "graph.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>

namespace tfsm {

namespace {

struct Cursor {
    const char *p, *end;
    bool at_eol() const { return p >= end || *p == '\n'; }
    void skip_spaces() { while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p; }
    void next_line() { while (p < end && *p != '\n') ++p; if (p < end) ++p; }
    bool read_int(long long &out) {
        skip_spaces();
        if (at_eol()) return false;
        bool neg = false;
        if (*p == '-') { neg = true; ++p; }
        if (p >= end || *p < '0' || *p > '9') return false;
        long long v = 0;
        while (p < end && *p >= '0' && *p <= '9') { v = v * 10 + (*p - '0'); ++p; }
        out = neg ? -v : v;
        return true;
    }
};

}  // namespace

InputGraph read_lg(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("Can not open the graph file " + path + " .");
    std::fseek(f, 0, SEEK_END);
    long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<char> buf((size_t)std::max(size, 0L));
    if (size > 0 && std::fread(buf.data(), 1, (size_t)size, f) != (size_t)size) {
        std::fclose(f);
        throw std::runtime_error("short read on " + path);
    }
    std::fclose(f);

    InputGraph g;
    Cursor c{buf.data(), buf.data() + buf.size()};
    size_t line = 0;
    uint32_t max_id = 0;
    while (c.p < c.end) {
        ++line;
        c.skip_spaces();
        if (c.at_eol()) { c.next_line(); continue; }
        char kind = *c.p++;
        long long a, b, d;
        if (kind == 't') {
            // header: counts are informational only
        } else if (kind == 'v') {
            if (!c.read_int(a) || !c.read_int(b) || a < 0)
                throw std::runtime_error("bad v line " + std::to_string(line) + " in " + path);
            g.vid.push_back((uint32_t)a);
            g.vlab.push_back((int32_t)b);
            max_id = std::max(max_id, (uint32_t)a);
        } else if (kind == 'e') {
            if (!c.read_int(a) || !c.read_int(b) || !c.read_int(d) || a < 0 || b < 0)
                throw std::runtime_error("bad e line " + std::to_string(line) + " in " + path);
            g.eu.push_back((uint32_t)a);
            g.ev.push_back((uint32_t)b);
            g.el.push_back((int32_t)d);
        } else {
            throw std::runtime_error("Reading input error! line " + std::to_string(line) + " in " + path);
        }
        c.next_line();
    }
    g.label_of.assign(g.vid.empty() ? 0 : (size_t)max_id + 1, 0);
    std::vector<uint8_t> seen(g.label_of.size(), 0);
    for (size_t i = 0; i < g.vid.size(); ++i) {
        g.label_of[g.vid[i]] = g.vlab[i];
        seen[g.vid[i]] = 1;
    }
    for (size_t i = 0; i < g.eu.size(); ++i) {
        if (g.eu[i] >= seen.size() || g.ev[i] >= seen.size() || !seen[g.eu[i]] || !seen[g.ev[i]])
            throw std::runtime_error("edge " + std::to_string(i) + " names an undeclared vertex in " + path);
    }
    return g;
}

int DataGraph::dense_vlabel(int32_t orig) const {
    auto it = std::lower_bound(vlabel_values.begin(), vlabel_values.end(), orig);
    return (it != vlabel_values.end() && *it == orig) ? (int)(it - vlabel_values.begin()) : -1;
}

int DataGraph::dense_elabel(int32_t orig) const {
    auto it = std::lower_bound(elabel_values.begin(), elabel_values.end(), orig);
    return (it != elabel_values.end() && *it == orig) ? (int)(it - elabel_values.begin()) : -1;
}

DataGraph prune_graph(const InputGraph &in, int32_t freq) {
    DataGraph g;
    g.freq = freq;
    // label frequencies over the whole input (graph.h:662-683)
    std::map<int32_t, int64_t> vcount, ecount;
    for (int32_t l : in.vlab) vcount[l]++;
    for (int32_t l : in.el) ecount[l]++;
    for (auto &kv : vcount) if (kv.second >= freq) g.vlabel_values.push_back(kv.first);
    for (auto &kv : ecount) if (kv.second >= freq) g.elabel_values.push_back(kv.first);

    // renumber kept vertices in v-line order (graph.h:700-712)
    const uint32_t kNone = ~0u;
    std::vector<uint32_t> newid(in.label_of.size(), kNone);
    for (size_t i = 0; i < in.vid.size(); ++i) {
        int dl = g.dense_vlabel(in.vlab[i]);
        if (dl >= 0) {
            newid[in.vid[i]] = g.n++;
            g.vlabel.push_back((uint32_t)dl);
        }
    }

    // kept edges -> CSR sorted by neighbour id (graph.h:714-769)
    std::vector<uint32_t> deg(g.n, 0);
    std::vector<uint32_t> keep_edges;
    keep_edges.reserve(in.eu.size());
    for (size_t i = 0; i < in.eu.size(); ++i) {
        uint32_t a = newid[in.eu[i]], b = newid[in.ev[i]];
        if (a == kNone || b == kNone || g.dense_elabel(in.el[i]) < 0) continue;
        keep_edges.push_back((uint32_t)i);
        deg[a]++; deg[b]++;
    }
    g.offsets.assign(g.n + 1, 0);
    for (uint32_t v = 0; v < g.n; ++v) g.offsets[v + 1] = g.offsets[v] + deg[v];
    g.nbrs.resize(g.offsets[g.n]);
    g.elab.resize(g.offsets[g.n]);
    std::vector<uint32_t> fill(g.offsets.begin(), g.offsets.end() - 1);
    for (uint32_t i : keep_edges) {
        uint32_t a = newid[in.eu[i]], b = newid[in.ev[i]];
        uint32_t e = (uint32_t)g.dense_elabel(in.el[i]);
        g.nbrs[fill[a]] = b; g.elab[fill[a]++] = e;
        g.nbrs[fill[b]] = a; g.elab[fill[b]++] = e;
    }
    {
        std::vector<std::pair<uint32_t, uint32_t>> tmp;
        for (uint32_t v = 0; v < g.n; ++v) {
            uint32_t s = g.offsets[v], t = g.offsets[v + 1];
            tmp.clear();
            for (uint32_t k = s; k < t; ++k) tmp.emplace_back(g.nbrs[k], g.elab[k]);
            std::stable_sort(tmp.begin(), tmp.end(),
                             [](const std::pair<uint32_t, uint32_t> &x, const std::pair<uint32_t, uint32_t> &y) { return x.first < y.first; });
            for (uint32_t k = s; k < t; ++k) { g.nbrs[k] = tmp[k - s].first; g.elab[k] = tmp[k - s].second; }
        }
    }

    // size-2 endpoint sets over every input edge (graph.h:733-753, 789-800)
    std::map<std::tuple<int32_t, int32_t, int32_t>, std::pair<std::vector<uint32_t>, std::vector<uint32_t>>> sets;
    for (size_t i = 0; i < in.eu.size(); ++i) {
        int32_t lf = in.label_of[in.eu[i]], lt = in.label_of[in.ev[i]], le = in.el[i];
        uint32_t f = newid[in.eu[i]] == kNone ? 0 : newid[in.eu[i]];
        uint32_t t = newid[in.ev[i]] == kNone ? 0 : newid[in.ev[i]];
        if (lf < lt) {
            auto &s = sets[std::make_tuple(lf, le, lt)];
            s.first.push_back(f); s.second.push_back(t);
        } else if (lf == lt) {
            auto &s = sets[std::make_tuple(lf, le, lt)];
            s.first.push_back(f); s.first.push_back(t);
            s.second.push_back(t); s.second.push_back(f);
        } else {
            auto &s = sets[std::make_tuple(lt, le, lf)];
            s.first.push_back(t); s.second.push_back(f);
        }
    }
    for (auto &kv : sets) {
        auto &A = kv.second.first, &B = kv.second.second;
        std::sort(A.begin(), A.end()); A.erase(std::unique(A.begin(), A.end()), A.end());
        std::sort(B.begin(), B.end()); B.erase(std::unique(B.begin(), B.end()), B.end());
        if ((int64_t)A.size() < freq || (int64_t)B.size() < freq) continue;
        g.edge_stats.push_back({std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first),
                                (uint32_t)A.size(), (uint32_t)B.size()});
    }
    return g;
}

}  // namespace tfsm
