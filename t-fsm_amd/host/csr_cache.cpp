// Binary cache of the pruned data graph (format in csr_cache.h).
#include "csr_cache.h"

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace tfsm {

namespace {

constexpr char kMagic[8] = {'T', 'F', 'S', 'M', 'C', 'S', 'R', '1'};

// Header, then the arrays in this order: vlabel_values, elabel_values,
// vlabel, offsets, nbrs, elab (4-byte entries), edge_stats (5 x 4 bytes).
struct Header {
    char magic[8];
    uint64_t src_hash, src_size;
    int32_t freq;
    uint32_t n;
    uint64_t n_vlv, n_elv, n_vlabel, n_offsets, n_nbrs, n_elab, n_stats;
};
static_assert(sizeof(EdgeStat) == 20, "edge statistics are five 4-byte fields");

inline uint64_t mix(uint64_t h, uint64_t w) {
    h ^= w;
    h *= 0x9e3779b97f4a7c15ull;
    return h ^ (h >> 31);
}

template <class T>
bool put(FILE *f, const std::vector<T> &v) {
    return v.empty() || std::fwrite(v.data(), sizeof(T), v.size(), f) == v.size();
}

template <class T>
bool get(FILE *f, std::vector<T> &v, uint64_t n) {
    v.resize((size_t)n);
    return n == 0 || std::fread(v.data(), sizeof(T), (size_t)n, f) == (size_t)n;
}

}  // namespace

uint64_t file_hash(const std::string &path, uint64_t *size_out) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("Can not open the graph file " + path + " .");
    std::vector<char> buf(1u << 22);
    uint64_t h = 0x243f6a8885a308d3ull, total = 0;
    size_t got;
    uint8_t tail[8];
    size_t ntail = 0;
    while ((got = std::fread(buf.data(), 1, buf.size(), f)) > 0) {
        size_t i = 0;
        // finish a word begun in the previous block
        while (ntail && ntail < 8 && i < got) tail[ntail++] = (uint8_t)buf[i++];
        if (ntail == 8) {
            uint64_t w;
            std::memcpy(&w, tail, 8);
            h = mix(h, w);
            ntail = 0;
        }
        for (; i + 8 <= got; i += 8) {
            uint64_t w;
            std::memcpy(&w, buf.data() + i, 8);
            h = mix(h, w);
        }
        while (i < got) tail[ntail++] = (uint8_t)buf[i++];
        total += got;
    }
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) throw std::runtime_error("read error on " + path);
    uint64_t w = 0;
    for (size_t i = 0; i < ntail; ++i) w |= (uint64_t)tail[i] << (8 * i);
    h = mix(mix(h, w), total);
    if (size_out) *size_out = total;
    return h;
}

std::string csr_cache_path(const std::string &lg, int32_t freq) { return lg + ".f" + std::to_string(freq) + ".csr"; }

bool read_csr_cache(const std::string &cache, uint64_t src_hash, uint64_t src_size, int32_t freq, DataGraph &g) {
    FILE *f = std::fopen(cache.c_str(), "rb");
    if (!f) return false;
    Header h;
    bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, kMagic, 8) == 0 &&
              h.src_hash == src_hash && h.src_size == src_size && h.freq == freq && h.n_vlabel == h.n &&
              h.n_offsets == (uint64_t)h.n + 1 && h.n_elab == h.n_nbrs;
    DataGraph d;
    if (ok) {
        d.freq = h.freq;
        d.n = h.n;
        ok = get(f, d.vlabel_values, h.n_vlv) && get(f, d.elabel_values, h.n_elv) && get(f, d.vlabel, h.n_vlabel) &&
             get(f, d.offsets, h.n_offsets) && get(f, d.nbrs, h.n_nbrs) && get(f, d.elab, h.n_elab) &&
             get(f, d.edge_stats, h.n_stats);
    }
    // nothing may follow the arrays, and the offsets must describe the lists
    if (ok) ok = std::fgetc(f) == EOF && d.offsets.front() == 0 && d.offsets.back() == d.nbrs.size();
    if (ok)
        for (uint32_t v = 0; v < d.n && ok; ++v) ok = d.offsets[v] <= d.offsets[v + 1];
    std::fclose(f);
    if (ok) g = std::move(d);
    return ok;
}

bool write_csr_cache(const std::string &cache, uint64_t src_hash, uint64_t src_size, const DataGraph &g) {
    // written under a temporary name and renamed, so a reader never sees a
    // half-written file
    const std::string tmp = cache + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return false;
    Header h;
    std::memcpy(h.magic, kMagic, 8);
    h.src_hash = src_hash;
    h.src_size = src_size;
    h.freq = g.freq;
    h.n = g.n;
    h.n_vlv = g.vlabel_values.size();
    h.n_elv = g.elabel_values.size();
    h.n_vlabel = g.vlabel.size();
    h.n_offsets = g.offsets.size();
    h.n_nbrs = g.nbrs.size();
    h.n_elab = g.elab.size();
    h.n_stats = g.edge_stats.size();
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && put(f, g.vlabel_values) && put(f, g.elabel_values) &&
              put(f, g.vlabel) && put(f, g.offsets) && put(f, g.nbrs) && put(f, g.elab) && put(f, g.edge_stats);
    ok = std::fclose(f) == 0 && ok;
    if (ok) ok = std::rename(tmp.c_str(), cache.c_str()) == 0;
    if (!ok) std::remove(tmp.c_str());
    return ok;
}

DataGraph load_graph_cached(const std::string &lg, int32_t freq, bool *hit) {
    uint64_t size = 0;
    const uint64_t hash = file_hash(lg, &size);
    const std::string cache = csr_cache_path(lg, freq);
    DataGraph g;
    if (read_csr_cache(cache, hash, size, freq, g)) {
        if (hit) *hit = true;
        return g;
    }
    g = prune_graph(read_lg(lg), freq);
    (void)write_csr_cache(cache, hash, size, g);
    if (hit) *hit = false;
    return g;
}

}  // namespace tfsm
