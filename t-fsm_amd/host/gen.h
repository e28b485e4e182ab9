// Seeded synthetic graph generators for the T-FSM workloads, writing the
// reference's `.lg` text format (README.md:73-94 of the reference: a `t N M`
// line, then `v id label degree` lines sorted by id, then `e u v elabel`).
//
// PRNG: splitmix64 seeds xoshiro256** (Blackman & Vigna).  Every generator is
// a pure function of its parameters and seed, so the same call reproduces the
// same file on any host.
#pragma once

#include <cstdint>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>
#include <algorithm>
#include <stdexcept>

namespace tfsm {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto &w : s) w = splitmix(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform in [0, 1) with 53 random bits
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    // uniform integer in [0, n) by 128-bit multiply-shift (Lemire, no rejection:
    // bias < n / 2^64, irrelevant at these sizes)
    uint64_t below(uint64_t n) {
        return (uint64_t)(((unsigned __int128)next() * n) >> 64);
    }
};

// Open-addressing set of undirected edges keyed (min << 32 | max); keeps the
// insertion order in `order` so the written edge list is the generation order.
class EdgeSet {
public:
    explicit EdgeSet(size_t expected) {
        size_t cap = 16;
        while (cap < expected * 2) cap <<= 1;
        slots_.assign(cap, kEmpty);
        mask_ = cap - 1;
        order.reserve(expected);
    }
    bool insert(uint32_t a, uint32_t b) {
        uint64_t key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
        uint64_t h = key * 0x9E3779B97F4A7C15ull;
        size_t i = (size_t)(h >> 20) & mask_;
        while (true) {
            if (slots_[i] == kEmpty) { slots_[i] = key; order.push_back(key); return true; }
            if (slots_[i] == key) return false;
            i = (i + 1) & mask_;
        }
    }
    size_t size() const { return order.size(); }
    std::vector<uint64_t> order;
private:
    static constexpr uint64_t kEmpty = ~0ull;
    std::vector<uint64_t> slots_;
    size_t mask_;
};

struct LabeledGraph {
    uint32_t n = 0;
    std::vector<uint32_t> vlabel;          // n
    std::vector<uint64_t> edges;           // (min << 32 | max), generation order
    std::vector<uint32_t> elabel;          // one per edge
};

// Walker alias table for sampling an index proportionally to weights.
struct Alias {
    std::vector<double> prob;
    std::vector<uint32_t> alias;
    explicit Alias(const std::vector<double> &w) {
        size_t n = w.size();
        prob.resize(n); alias.resize(n);
        double sum = 0; for (double x : w) sum += x;
        std::vector<double> p(n);
        std::vector<uint32_t> small, large;
        for (size_t i = 0; i < n; ++i) {
            p[i] = w[i] * n / sum;
            (p[i] < 1.0 ? small : large).push_back((uint32_t)i);
        }
        while (!small.empty() && !large.empty()) {
            uint32_t s = small.back(); small.pop_back();
            uint32_t l = large.back(); large.pop_back();
            prob[s] = p[s]; alias[s] = l;
            p[l] = (p[l] + p[s]) - 1.0;
            (p[l] < 1.0 ? small : large).push_back(l);
        }
        for (uint32_t l : large) { prob[l] = 1.0; alias[l] = l; }
        for (uint32_t s : small) { prob[s] = 1.0; alias[s] = s; }
    }
    uint32_t sample(Rng &r) const {
        uint32_t i = (uint32_t)r.below(prob.size());
        return r.uniform() < prob[i] ? i : alias[i];
    }
};

// G(n, m): m distinct non-loop undirected edges drawn uniformly; vertex labels
// uniform over [0, nlabels); every edge label is `elabel`.
inline LabeledGraph gen_gnm(uint32_t n, uint64_t m, uint32_t nlabels, uint32_t elabel, uint64_t seed) {
    if (m > (uint64_t)n * (n - 1) / 2) throw std::invalid_argument("gnm: too many edges");
    Rng r(seed);
    LabeledGraph g; g.n = n;
    g.vlabel.resize(n);
    for (auto &l : g.vlabel) l = (uint32_t)r.below(nlabels);
    EdgeSet es(m);
    while (es.size() < m) {
        uint32_t a = (uint32_t)r.below(n), b = (uint32_t)r.below(n);
        if (a != b) es.insert(a, b);
    }
    g.edges = std::move(es.order);
    g.elabel.assign(g.edges.size(), elabel);
    return g;
}

// Chung-Lu power law: w_i = c * (i+1)^(-1/(gamma-1)), scaled so the mean is
// `mean_deg`, then capped at `max_w`; endpoints drawn independently with
// probability proportional to w (alias sampling) until m distinct non-loop
// edges exist.  An edge whose endpoint already has degree `max_w` is rejected,
// so no realised degree exceeds the cap.  Vertex labels ~ Zipf(s) over
// [0, nlabels) (rank r -> label r-1).
inline LabeledGraph gen_chung_lu(uint32_t n, uint64_t m, double gamma, double mean_deg, double max_w,
                                 uint32_t nlabels, double zipf_s, uint32_t elabel, uint64_t seed) {
    Rng r(seed);
    LabeledGraph g; g.n = n;
    std::vector<double> w(n);
    double sum = 0;
    for (uint32_t i = 0; i < n; ++i) { w[i] = std::pow((double)(i + 1), -1.0 / (gamma - 1.0)); sum += w[i]; }
    double scale = mean_deg * n / sum;
    for (auto &x : w) x = std::min(x * scale, max_w);
    // labels first (independent of degree)
    std::vector<double> cdf(nlabels);
    double z = 0;
    for (uint32_t k = 0; k < nlabels; ++k) { z += 1.0 / std::pow((double)(k + 1), zipf_s); cdf[k] = z; }
    for (auto &c : cdf) c /= z;
    g.vlabel.resize(n);
    for (auto &l : g.vlabel) {
        double x = r.uniform();
        uint32_t k = (uint32_t)(std::lower_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
        l = std::min(k, nlabels - 1);
    }
    Alias al(w);
    EdgeSet es(m);
    std::vector<uint32_t> deg(n, 0);
    const uint32_t cap = (uint32_t)max_w;
    while (es.size() < m) {
        uint32_t a = al.sample(r), b = al.sample(r);
        if (a == b || deg[a] >= cap || deg[b] >= cap) continue;
        if (es.insert(a, b)) { deg[a]++; deg[b]++; }
    }
    g.edges = std::move(es.order);
    g.elabel.assign(g.edges.size(), elabel);
    return g;
}

// R-MAT (Chakrabarti et al.) on 2^scale vertices with quadrant probabilities
// a, b, c, d = 1-a-b-c; edges are symmetrised, loops and duplicates dropped,
// and sampling continues until m distinct edges exist.  Vertex labels uniform
// over [0, nlabels); edge labels uniform over [1, nelabels].
inline LabeledGraph gen_rmat(uint32_t scale, uint64_t m, double a, double b, double c,
                             uint32_t nlabels, uint32_t nelabels, uint64_t seed) {
    Rng r(seed);
    LabeledGraph g; g.n = 1u << scale;
    g.vlabel.resize(g.n);
    for (auto &l : g.vlabel) l = (uint32_t)r.below(nlabels);
    EdgeSet es(m);
    const double ab = a + b, abc = a + b + c;
    while (es.size() < m) {
        uint32_t u = 0, v = 0;
        for (uint32_t bit = 0; bit < scale; ++bit) {
            double x = r.uniform();
            uint32_t ub = 0, vb = 0;
            if (x < a) { }
            else if (x < ab) { vb = 1; }
            else if (x < abc) { ub = 1; }
            else { ub = 1; vb = 1; }
            u = (u << 1) | ub; v = (v << 1) | vb;
        }
        if (u != v) es.insert(u, v);
    }
    g.edges = std::move(es.order);
    g.elabel.resize(g.edges.size());
    for (auto &e : g.elabel) e = 1 + (uint32_t)r.below(nelabels);
    return g;
}

inline void write_lg(const LabeledGraph &g, const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<uint32_t> deg(g.n, 0);
    for (uint64_t e : g.edges) { deg[e >> 32]++; deg[(uint32_t)e]++; }
    std::vector<char> buf(1 << 20);
    setvbuf(f, buf.data(), _IOFBF, buf.size());
    std::fprintf(f, "t %u %zu\n", g.n, g.edges.size());
    for (uint32_t i = 0; i < g.n; ++i) std::fprintf(f, "v %u %u %u\n", i, g.vlabel[i], deg[i]);
    for (size_t i = 0; i < g.edges.size(); ++i)
        std::fprintf(f, "e %u %u %u\n", (uint32_t)(g.edges[i] >> 32), (uint32_t)g.edges[i], g.elabel[i]);
    std::fclose(f);
}

// The named workloads.  `seed` defaults to the workload's fixed seed; the
// learned-ordering runs use the same generators with other seeds.
inline LabeledGraph gen_named(const std::string &name, uint64_t seed) {
    if (name == "gnm2k")       // G(2000, 8000), 8 uniform labels, edge label 1
        return gen_gnm(2000, 8000, 8, 1, seed);
    if (name == "powerlaw100k") // Chung-Lu 100k / 1.1M, gamma 2.1, mean 22, cap 1500, Zipf(1) over 29 labels
        return gen_chung_lu(100000, 1100000, 2.1, 22.0, 1500.0, 29, 1.0, 1, seed);
    if (name == "rmat20")      // R-MAT scale 20, 16 * 2^20 edges, 16 labels, 4 edge labels
        return gen_rmat(20, 16ull << 20, 0.57, 0.19, 0.19, 16, 4, seed);
    throw std::invalid_argument("unknown workload " + name);
}

inline uint64_t default_seed(const std::string &name) {
    if (name == "gnm2k") return 1;
    if (name == "powerlaw100k") return 2;
    if (name == "rmat20") return 4;
    return 0;
}

}  // namespace tfsm
