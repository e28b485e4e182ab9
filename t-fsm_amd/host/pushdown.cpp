// Push-down pruning: decomposition, canonical forms and the cache; see pushdown.h.
#include "pushdown.h"

#include <algorithm>
#include <functional>
#include <tuple>

namespace tfsm {

Canon canonical_form(int n, const int32_t *lab, const int32_t *el) {
    // vertex invariants: label, degree, sorted (neighbour label, edge label)
    std::vector<std::vector<int64_t>> inv(n);
    for (int v = 0; v < n; ++v) {
        std::vector<int64_t> nb;
        for (int w = 0; w < n; ++w)
            if (el[v * n + w] >= 0) nb.push_back(((int64_t)lab[w] << 32) | (uint32_t)el[v * n + w]);
        std::sort(nb.begin(), nb.end());
        inv[v].push_back(lab[v]);
        inv[v].push_back((int64_t)nb.size());
        inv[v].insert(inv[v].end(), nb.begin(), nb.end());
    }
    // positions are filled in invariant order; within equal invariants every
    // arrangement is tried and the lexicographically least code kept
    std::vector<int> byinv(n);
    for (int i = 0; i < n; ++i) byinv[i] = i;
    std::stable_sort(byinv.begin(), byinv.end(), [&](int a, int b) { return inv[a] < inv[b]; });
    // code of position k: label, then edge label + 1 (0: none) to each earlier position
    std::vector<int64_t> best, cur;
    std::vector<int> best_order, order;
    std::vector<char> used(n, 0);
    bool have = false;
    std::vector<size_t> mark;
    std::function<void(int)> rec;
    rec = [&](int k) {
        if (k == n) {
            if (!have || cur < best) {
                best = cur;
                best_order = order;
                have = true;
            }
            return;
        }
        // candidates: unused vertices with the invariant of position k's slot
        const std::vector<int64_t> &want = inv[byinv[k]];
        for (int v = 0; v < n; ++v) {
            if (used[v] || inv[v] != want) continue;
            const size_t before = cur.size();
            cur.push_back(lab[v]);
            for (int j = 0; j < k; ++j) cur.push_back((int64_t)el[order[j] * n + v] + 1);
            // prune: a prefix already above the best cannot win
            bool worse = false;
            if (have) {
                const size_t m = std::min(cur.size(), best.size());
                for (size_t i = 0; i < m; ++i) {
                    if (cur[i] != best[i]) {
                        worse = cur[i] > best[i];
                        break;
                    }
                }
            }
            if (!worse) {
                used[v] = 1;
                order.push_back(v);
                rec(k + 1);
                order.pop_back();
                used[v] = 0;
            }
            cur.resize(before);
        }
    };
    rec(0);
    Canon c;
    c.order = best_order;
    c.key = std::to_string(n) + ":";
    for (int64_t x : best) c.key += std::to_string(x) + ",";
    return c;
}

Canon canonical_form(const Pattern &p) {
    const int n = p.nv();
    std::vector<int32_t> el((size_t)n * n, -1);
    for (const DfsCode &d : p.code) el[d.from * n + d.to] = el[d.to * n + d.from] = d.el;
    return canonical_form(n, p.vlab.data(), el.data());
}

std::vector<SubPattern> decompose(const Pattern &p) {
    const int n = p.nv(), last = n - 1;
    std::vector<SubPattern> out;
    for (int e = 0; e < p.ne(); ++e) {
        const int a = p.code[e].from, b = p.code[e].to;
        // the part without edge e that holds the last vertex
        std::vector<char> seen(n, 0);
        std::vector<int> stack{last}, verts;
        seen[last] = 1;
        while (!stack.empty()) {
            const int x = stack.back();
            stack.pop_back();
            verts.push_back(x);
            for (auto &nb : p.adj[x]) {
                const int y = nb.first;
                if ((x == a && y == b) || (x == b && y == a) || seen[y]) continue;
                seen[y] = 1;
                stack.push_back(y);
            }
        }
        if (verts.size() < 2) continue;
        std::sort(verts.begin(), verts.end());
        SubPattern s;
        s.verts = verts;
        const int m = (int)verts.size();
        s.lab.resize(m);
        s.el.assign((size_t)m * m, -1);
        std::vector<int> at(n, -1);
        for (int i = 0; i < m; ++i) {
            at[verts[i]] = i;
            s.lab[i] = p.vlab[verts[i]];
        }
        for (int f = 0; f < p.ne(); ++f) {
            if (f == e) continue;
            const int x = at[p.code[f].from], y = at[p.code[f].to];
            if (x < 0 || y < 0) continue;
            s.el[x * m + y] = s.el[y * m + x] = p.code[f].el;
        }
        out.push_back(std::move(s));
    }
    return out;
}

void PushdownCache::insert(const Pattern &p, uint64_t handle) {
    Canon c = canonical_form(p);
    std::lock_guard<std::mutex> lk(mu);
    entries[c.key] = Entry{handle, c.order};
}

std::vector<tfsm_pushdown> PushdownCache::lookup(const Pattern &p, uint32_t pattern_index, uint64_t skip_handle) const {
    std::vector<tfsm_pushdown> out;
    for (const SubPattern &s : decompose(p)) {
        const Canon c = canonical_form((int)s.verts.size(), s.lab.data(), s.el.data());
        const Entry *hit = nullptr;
        {
            std::lock_guard<std::mutex> lk(mu);
            auto it = entries.find(c.key);
            if (it != entries.end()) hit = &it->second;
        }
        if (!hit || hit->handle == skip_handle) continue;
        for (size_t k = 0; k < c.order.size(); ++k) {
            tfsm_pushdown d{};
            d.pattern = pattern_index;
            d.vertex = (uint32_t)s.verts[c.order[k]];
            d.handle = hit->handle;
            d.source_vertex = (uint32_t)hit->order[k];
            out.push_back(d);
        }
    }
    // one entry per (vertex, handle, source vertex)
    std::sort(out.begin(), out.end(), [](const tfsm_pushdown &x, const tfsm_pushdown &y) {
        return std::tie(x.vertex, x.handle, x.source_vertex) < std::tie(y.vertex, y.handle, y.source_vertex);
    });
    out.erase(std::unique(out.begin(), out.end(),
                          [](const tfsm_pushdown &x, const tfsm_pushdown &y) {
                              return x.vertex == y.vertex && x.handle == y.handle &&
                                     x.source_vertex == y.source_vertex;
                          }),
              out.end());
    return out;
}

}  // namespace tfsm
