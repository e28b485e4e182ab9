// Binary cache of the pruned data graph, so that a `.lg` is parsed once per
// frequency threshold.  The cache file sits beside the input as
// `<file>.f<freq>.csr` and holds the DataGraph arrays of prune_graph(read_lg
// (file), freq) as flat little-endian arrays after a header carrying a 64-bit
// hash and the size of the `.lg` bytes it was made from.  A cache whose
// header, sizes or source hash do not match is ignored and rewritten.  The
// reference has no such cache: its loader parses the text on every run
// (graph.h:557-803).
#pragma once

#include <cstdint>
#include <string>

#include "graph.h"

namespace tfsm {

// 64-bit hash of a file's bytes (eight-byte words mixed by multiply and
// xor-shift, then the tail bytes and the length).  Throws on I/O errors.
uint64_t file_hash(const std::string &path, uint64_t *size_out = nullptr);

// The cache file's path for (lg, freq).
std::string csr_cache_path(const std::string &lg, int32_t freq);

// Loads the pruned graph for (lg, freq): from the cache when it is valid for
// the file's current bytes, else by parsing and pruning, after which the
// cache is written (best effort: a failed write is ignored).  `hit`, when
// given, tells which happened.
DataGraph load_graph_cached(const std::string &lg, int32_t freq, bool *hit = nullptr);

// The two halves, for tests: read a cache (false when absent or invalid for
// the given source hash and size) and write one.
bool read_csr_cache(const std::string &cache, uint64_t src_hash, uint64_t src_size, int32_t freq, DataGraph &g);
bool write_csr_cache(const std::string &cache, uint64_t src_hash, uint64_t src_size, const DataGraph &g);

}  // namespace tfsm
