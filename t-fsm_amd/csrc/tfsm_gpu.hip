// libtfsm_gpu.so: the C ABI of include/tfsm_gpu.h on MI355X (gfx950).
//
// One translation unit: the kernels (kernels.inc) and the host side that
// plans, launches and reads back an evaluation of one pattern
// (tfsm_eval_pattern) or one extension level (tfsm_eval_level).  Two drivers
// run the same kernels:
//   host-driven:   LDF -> DP-iso passes until a pass removes nothing ->
//                  compaction -> (host) sizes, express / filter decisions,
//                  matching plans, layout -> one-pass table build ->
//                  learned order -> existence search -> (host) verdicts and
//                  retained child domains;
//   device-driven: the same phases in one launch sequence with no host round
//                  trip inside it (the last compaction block makes the
//                  decisions and layout on the device), every kernel reading
//                  a per-context descriptor, so that the sequence is captured
//                  once in a hipGraph and replayed for every level of the
//                  same shape; the host reads the results back once.
#include "kernels.inc"

#include <algorithm>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

using namespace tfsm_dev;

// Scratch of tfsm_ranker_gradient_noalloc: the gradient and loss sums, one
// copy per GPU for the process's life.  Host code never passes this name to
// a kernel (in host code it is a stand-in, not a device address): each
// context looks up its device's copy with hipGetSymbolAddress at creation.
constexpr uint32_t kStaticGradWords = kParams + 1;
static __device__ __align__(256) float scratch[kStaticGradWords];
static_assert(kStaticGradWords == TFSM_NOALLOC_SCRATCH_FLOATS, "the header states the scratch size");

namespace {

__global__ __launch_bounds__(kBlock) void k_fill_f32(float *p, uint32_t n, float v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// out[i] = sums[i] * scale for the kParams gradient words, and the loss
__global__ __launch_bounds__(kBlock) void k_scale_grad(const float *sums, float scale, float *grad_out,
                                                       float *loss_out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)kParams; i += gridDim.x * blockDim.x)
        grad_out[i] = sums[i] * scale;
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss_out) *loss_out = sums[kParams] * scale;
}

}  // namespace

// Search-kernel work splitting: record capacity per launch, the default step
// budget after which a piece donates work and the default number of records
// a donation is cut into.
constexpr uint32_t kSplitCap = 1u << 21;
constexpr uint32_t kDefaultSplitSteps = 1024;
constexpr uint32_t kDefaultSplitFanout = 1;
// Lane groups: the default most lanes per root task, and the entries one
// launch may run in groups (each has a piece counter beyond the records'
// counters, two regions: one per search launch of a call).
constexpr uint32_t kDefaultGroupLanes = 8;
constexpr uint32_t kGroupCap = 1u << 16;
// Wave-cooperative second phase: default step budget of a root task in the
// first phase (off by default: on the power-law workloads its rounds, binary
// searches in global memory, cost more than the lane-per-task steps they
// replace), list capacity, waves per CU, offer interval and fanout.
constexpr uint32_t kDefaultLongSteps = 0xffffffffu;
constexpr unsigned long long kLongCap = 1ull << 20;
constexpr uint32_t kDefaultLongWavesPerCu = 8;
constexpr uint32_t kDefaultLongDonateEvery = 32;
constexpr uint32_t kDefaultLongFanout = 4;

namespace {

struct HipError {
    hipError_t err;
    const char *what;
};

#define HIPX(expr)                                              \
    do {                                                        \
        hipError_t e_ = (expr);                                 \
        if (e_ != hipSuccess) throw HipError{e_, #expr};        \
    } while (0)

struct ArgError {
    int code;
    std::string msg;
};

// Chunked device arena: bump allocation, chunks never move, reset() reuses
// every chunk.  Holds the scratch of one eval call.
class Arena {
public:
    ~Arena() {
        for (auto &c : chunks_) (void)hipFree(c.ptr);
    }
    void reset() {
        for (auto &c : chunks_) c.used = 0;
    }
    template <class T>
    T *alloc(size_t count) {
        size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
        if (bytes == 0) bytes = 256;
        for (auto &c : chunks_) {
            if (c.cap - c.used >= bytes) {
                T *p = reinterpret_cast<T *>(c.ptr + c.used);
                c.used += bytes;
                return p;
            }
        }
        size_t cap = std::max(bytes, size_t(256) << 20);
        char *ptr = nullptr;
        HIPX(hipMalloc(&ptr, cap));
        chunks_.push_back({ptr, cap, bytes});
        return reinterpret_cast<T *>(ptr);
    }

private:
    struct Chunk {
        char *ptr;
        size_t cap, used;
    };
    std::vector<Chunk> chunks_;
};

// Size-class cache of device buffers for retained domains.
class BufferCache {
public:
    ~BufferCache() {
        for (auto &kv : free_) (void)hipFree(kv.second);
    }
    void *get(size_t bytes, size_t *cls) {
        size_t c = 4096;
        while (c < bytes) c <<= 1;
        *cls = c;
        auto it = free_.find(c);
        if (it != free_.end()) {
            void *p = it->second;
            free_.erase(it);
            return p;
        }
        void *p = nullptr;
        HIPX(hipMalloc(&p, c));
        return p;
    }
    void put(void *p, size_t cls) { free_.emplace(cls, p); }

private:
    std::multimap<size_t, void *> free_;
};

struct Retained {
    void *buf = nullptr;
    size_t cls = 0;
    uint32_t nv = 0;
    uint32_t vlab[kMaxV];            // dense labels of the pattern's vertices
    uint32_t size[kMaxV];
    uint32_t *dom[kMaxV];           // device pointers into buf
    uint32_t *invalid[kMaxV];
};

template <class T>
struct Pinned {
    T *ptr = nullptr;
    size_t cap = 0;
    ~Pinned() { if (ptr) (void)hipHostFree(ptr); }
    T *ensure(size_t n) {
        if (n > cap) {
            if (ptr) HIPX(hipHostFree(ptr));
            cap = std::max(n, cap * 2);
            HIPX(hipHostMalloc(reinterpret_cast<void **>(&ptr), cap * sizeof(T), hipHostMallocDefault));
        }
        return ptr;
    }
};

// Pinned host staging for one call's uploads: each upload is copied here and
// sent with an asynchronous copy (from pageable memory the runtime would
// stage it synchronously).  Chunks are only added during a call and reused
// by the next one, which starts after the previous call's final sync.
struct PinnedStage {
    std::vector<std::pair<char *, size_t>> chunks;
    size_t cur = 0, off = 0;
    ~PinnedStage() {
        for (auto &ch : chunks) (void)hipHostFree(ch.first);
    }
    void reset() { cur = off = 0; }
    void *alloc(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        while (cur < chunks.size() && off + bytes > chunks[cur].second) {
            ++cur;
            off = 0;
        }
        if (cur == chunks.size()) {
            const size_t sz = std::max<size_t>(bytes, 4u << 20);
            char *p = nullptr;
            HIPX(hipHostMalloc(reinterpret_cast<void **>(&p), sz, hipHostMallocDefault));
            chunks.emplace_back(p, sz);
            off = 0;
        }
        void *r = chunks[cur].first + off;
        off += bytes;
        return r;
    }
};

constexpr uint32_t kMaxPrunePasses = 4096;   // filter passes of one call
constexpr uint32_t kPruneBatch = 6;          // passes launched between convergence checks
// search steps between bookkeeping rounds, and a wave's root-task refills in
// lanes' worth (the autotuner's choices on both power-law workloads: 8 or 16
// steps and refills of at least 128 were 2-10% faster than 4 and 64)
constexpr uint32_t kDefaultStepsPerRound = 8;
constexpr uint32_t kDefaultRefillMin = 128;
constexpr uint32_t kDefaultRefillMax = 256;
constexpr uint32_t kPrioBlocks = 512;        // most blocks of the priority sort (one histogram column each)
// device-driven calls: blocks of the priority sort (fixed, so that the launch
// sequence does not depend on the task count), most patterns, and the most
// device memory their regions (sized by the candidates' sources) may take
constexpr uint32_t kGraphPrioBlocks = 512;
constexpr uint32_t kDeviceMaxPatterns = 4096;
constexpr size_t kDeviceRegionBudget = size_t(4) << 30;
constexpr uint32_t kNoSlot = 0xffffffffu;
constexpr unsigned long long kDeviceKeepBudget = 256ull << 20;   // bytes of kept-domain buffers per call
// a device-driven call's read-back block: the phase clock, two search
// counters, four words (pool used, overflow, filter passes run, compaction
// blocks done), then the pattern states
constexpr size_t kOutCtr = 5 * kClockWays * sizeof(unsigned long long);
constexpr size_t kOutFlags = kOutCtr + 2 * sizeof(SearchCounters);
constexpr size_t kOutStates = (kOutFlags + 4 * sizeof(uint32_t) + 255) & ~(size_t)255;
constexpr uint32_t kFanBins = 32;
constexpr size_t kFanMaxCells = 1u << 16;     // label x key cells with fan-out bins
constexpr uint32_t kLdfPushPct = 100;
constexpr uint32_t kPushMinItemsPerThread = 64;   // bucket items per thread of a one-block-per-CU grid
// launch grids of a device-driven call, which with the options make a graph's key
struct LevelShape {
    uint32_t learned, passes, gfill, gldf, gsrc, gpair, gaux, gsample, grank, vgrid, me8, keep;
    bool operator==(const LevelShape &o) const { return std::memcmp(this, &o, sizeof(o)) == 0; }
};

}  // namespace

struct tfsm_ctx {
    int device = 0;
    int num_cus = 0;
    int verify_blocks8 = 1, verify_blocks8a = 1, verify_blocks16 = 1;   // resident search workgroups per CU
    hipStream_t stream = nullptr;
    // graph
    uint32_t n = 0, L = 0, EL = 0, K = 0, W = 0;
    uint64_t nadj = 0;
    uint32_t *d_seg = nullptr, *d_nbr = nullptr, *d_bucket = nullptr;
    unsigned long long *d_keymask = nullptr; // per vertex: keys (mod 64) with a neighbour
    unsigned long long *d_keymask_bucket = nullptr;  // the same in label-bucket order
    double *d_score = nullptr;               // Fraction-Score table, n x L (computed on first use)
    std::vector<uint32_t> bucket_off;        // host copy, L + 1
    // fan-out bins for LDF push decisions: for label l and segment key k,
    // bin b of (l * K + k) * kFanBins counts the vertices of label l with
    // 2^b .. 2^(b+1) - 1 neighbours over k (fan_n) and sums those counts
    // (fan_sum); empty when L * K is too large to keep
    std::vector<uint32_t> fan_n;
    std::vector<unsigned long long> fan_sum;
    DevGraph g{};
    // scratch
    Arena work;
    uint32_t *d_pool = nullptr;
    uint32_t pool_cap = 0;
    uint32_t *d_flags = nullptr;             // [0] pool_used, [1] overflow (after the two d_ctr entries)
    uint32_t *d_pflags = nullptr;            // per filter pass: 1 when it removed a candidate
    PinnedStage stage;                       // upload staging of the current call
    SearchCounters *d_ctr = nullptr;         // first and deferred search launch
    SplitRecord *d_records = nullptr;        // kSplitCap records, reused across launches by epoch
    uint32_t *d_piece_cnt = nullptr, *d_piece_inc = nullptr;
    uint32_t epoch = 0;
    Pinned<PatState> h_state;
    Pinned<SearchCounters> h_ctr;
    Pinned<FracSum> h_frac;
    // retained domains
    BufferCache cache;
    std::unordered_map<uint64_t, Retained> retained;
    uint64_t next_handle = 1;
    // last call
    std::vector<PatStatic> last_static;
    std::vector<PatState> last_state;
    uint32_t last_pool_used = 0;            // pool entries the last call's tables reserved
    std::vector<int> last_decided;
    uint32_t *d_trace = nullptr;
    uint64_t trace_len = 0, trace_cap = 0;
    // ranker
    uint16_t *d_rank_w = nullptr;            // installed BF16 parameters (kParams)
    bool ranker_installed = false;
    float *d_feat = nullptr;                 // features of the last traced call
    float *d_static_grad = nullptr;          // this device's copy of `scratch`
    uint64_t feat_len = 0, feat_cap = 0;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ev_sync = nullptr;           // end of the stream's work, polled by sync_spin
    tfsm_call_stats stats{};
    std::string err;
    // device-driven level calls
    uint32_t level_slot = kNoSlot;          // this context's descriptor in c_level
    bool stream_idle = false;               // no work queued since the last eval call's final sync
    LevelArgs *d_level = nullptr;           // its device address
    Pinned<LevelArgs> h_level;              // the descriptor's staging
    Pinned<unsigned char> h_out;            // a device-driven call's read-back block
    int wall_khz = 100000;                  // rate of the device's wall clock
    std::vector<std::pair<LevelShape, hipGraphExec_t>> graphs;   // captured launch sequences
};

namespace {
// c_level slots in use, per device
std::mutex g_slot_mutex;
std::map<int, uint32_t> g_slots_used;

uint32_t acquire_slot(int device) {
    std::lock_guard<std::mutex> lk(g_slot_mutex);
    uint32_t &m = g_slots_used[device];
    for (uint32_t i = 0; i < kLevelSlots; ++i)
        if (!(m >> i & 1u)) {
            m |= 1u << i;
            return i;
        }
    return kNoSlot;
}

void release_slot(int device, uint32_t slot) {
    if (slot == kNoSlot) return;
    std::lock_guard<std::mutex> lk(g_slot_mutex);
    g_slots_used[device] &= ~(1u << slot);
}
}  // namespace

namespace {

int fail(tfsm_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

// `eval`: an eval entry point, which reads the idle flag itself; every other
// entry point may queue work on the stream, so it clears the flag
template <class F>
int guarded(tfsm_ctx *c, F &&f, bool eval = false) {
    if (!eval) c->stream_idle = false;
    try {
        return f();
    } catch (const HipError &e) {
        return fail(c, TFSM_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.err));
    } catch (const ArgError &e) {
        return fail(c, e.code, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(c, TFSM_ERR_LIMIT, "host allocation failed");
    }
}

// Upper bound on the neighbours over segment key `key` of any `s` vertices of
// label `lab`: whole fan-out bins from the largest down, then the rest of `s`
// at the next bin's largest fan-out (at most twice the true sum of the s
// largest fan-outs).
unsigned long long fan_bound(const tfsm_ctx *c, uint32_t lab, uint32_t key, uint32_t s) {
    const size_t cell = ((size_t)lab * c->K + key) * kFanBins;
    unsigned long long est = 0;
    for (int b = kFanBins - 1; b >= 0 && s; --b) {
        const uint32_t n = c->fan_n[cell + b];
        if (n <= s) {
            est += c->fan_sum[cell + b];
            s -= n;
        } else {
            est += std::min<unsigned long long>(c->fan_sum[cell + b], (unsigned long long)s * ((2ull << b) - 1));
            s = 0;
        }
    }
    return est;
}

uint32_t grid_for(unsigned long long items, int num_cus) {
    unsigned long long blocks = (items + kBlock - 1) / kBlock;
    unsigned long long cap = (unsigned long long)num_cus * 16;
    return (uint32_t)std::max<unsigned long long>(1, std::min(blocks, cap));
}

// Host data to device memory `d` through the pinned staging of the call.
void stage_copy(tfsm_ctx *c, void *d, const void *h, size_t bytes) {
    void *st = c->stage.alloc(bytes);
    std::memcpy(st, h, bytes);
    HIPX(hipMemcpyAsync(d, st, bytes, hipMemcpyHostToDevice, c->stream));
}

template <class T>
T *upload(tfsm_ctx *c, const std::vector<T> &v) {
    T *d = c->work.alloc<T>(std::max<size_t>(v.size(), 1));
    if (!v.empty()) stage_copy(c, d, v.data(), v.size() * sizeof(T));
    return d;
}

// One phase's uploads sent with one copy and its zero fills done in one
// launch.  The device block mirrors a pinned staging block byte for byte, so
// an upload's device address is known when it is added; an upload beyond the
// block's capacity goes on its own.  Fills are split into pieces of at most
// kFillPiece words, one workgroup each.
constexpr unsigned long long kFillPiece = 1u << 14;

struct PhaseBatch {
    tfsm_ctx *c;
    char *host, *dev;
    size_t cap, used = 0;
    std::vector<CopyDesc> zeros;
    PhaseBatch(tfsm_ctx *c_, size_t cap_);
    template <class T>
    T *add(const T *h, size_t count) {
        const size_t bytes = (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~(size_t)255;
        if (used + bytes > cap) {
            T *d = c->work.alloc<T>(std::max<size_t>(count, 1));
            if (count) stage_copy(c, d, h, count * sizeof(T));
            return d;
        }
        T *d = reinterpret_cast<T *>(dev + used);
        if (count) std::memcpy(host + used, h, count * sizeof(T));
        used += bytes;
        return d;
    }
    template <class T>
    T *add(const std::vector<T> &v) {
        return add(v.data(), v.size());
    }
    void zero(void *d, size_t bytes) {       // bytes: a multiple of 4
        uint32_t *w = static_cast<uint32_t *>(d);
        for (unsigned long long o = 0, n = bytes / 4; o < n; o += kFillPiece)
            zeros.push_back({nullptr, w + o, std::min(kFillPiece, n - o)});
    }
    void flush();
};

PhaseBatch::PhaseBatch(tfsm_ctx *c_, size_t cap_) : c(c_), cap((cap_ + 255) & ~(size_t)255) {
    host = static_cast<char *>(c->stage.alloc(std::max<size_t>(cap, 256)));
    dev = c->work.alloc<char>(std::max<size_t>(cap, 256));
}

void PhaseBatch::flush() {
    const CopyDesc *dz = zeros.empty() ? nullptr : add(zeros);
    if (used) HIPX(hipMemcpyAsync(dev, host, used, hipMemcpyHostToDevice, c->stream));
    if (dz) {
        hipLaunchKernelGGL(k_copy_batch, dim3((uint32_t)std::min<size_t>(zeros.size(), 4096)), dim3(kBlock), 0,
                           c->stream, dz, (uint32_t)zeros.size());
        HIPX(hipGetLastError());
    }
    used = cap;                 // later uploads go on their own
    zeros.clear();
}

struct HostWork {
    std::vector<unsigned long long> base{0};
    std::vector<uint32_t> pat, x;
    // `align`: the range takes a multiple of that many items (the excess
    // items fall outside the range's own count, which kernels read)
    void add(uint32_t p, uint32_t xx, uint64_t count, uint64_t align = 1) {
        if (count == 0) return;
        pat.push_back(p);
        x.push_back(xx);
        base.push_back(base.back() + (count + align - 1) / align * align);
    }
    WorkTable to_device(PhaseBatch &b) const {
        WorkTable w;
        w.base = b.add(base);
        w.pat = b.add(pat);
        w.x = b.add(x);
        w.M = (uint32_t)pat.size();
        w.total = base.back();
        return w;
    }
    // device bytes of to_device, with the rounding of the batch
    size_t bytes() const { return 3 * 256 + base.size() * 8 + 2 * pat.size() * 4; }
};

// Pattern checks: sizes, labels, edge endpoints, no duplicate edges, connected.
void check_pattern(const tfsm_ctx *c, const tfsm_pattern &p) {
    if (p.num_vertices < 2 || p.num_vertices > kMaxV)
        throw ArgError{TFSM_ERR_LIMIT, "pattern vertex count out of range [2, 16]"};
    if (p.num_edges > kMaxE) throw ArgError{TFSM_ERR_LIMIT, "pattern has more than 64 edges"};
    if (p.num_edges < p.num_vertices - 1) throw ArgError{TFSM_ERR_ARGUMENT, "pattern is not connected"};
    for (uint32_t i = 0; i < p.num_vertices; ++i)
        if (p.vlabel[i] >= c->L) throw ArgError{TFSM_ERR_ARGUMENT, "pattern vertex label out of range"};
    uint32_t adj[kMaxV] = {0};
    for (uint32_t e = 0; e < p.num_edges; ++e) {
        uint32_t a = p.edge_a[e], b = p.edge_b[e];
        if (a >= p.num_vertices || b >= p.num_vertices || a == b)
            throw ArgError{TFSM_ERR_ARGUMENT, "pattern edge endpoint out of range"};
        if (p.edge_label[e] >= c->EL) throw ArgError{TFSM_ERR_ARGUMENT, "pattern edge label out of range"};
        if (adj[a] >> b & 1) throw ArgError{TFSM_ERR_ARGUMENT, "duplicate pattern edge"};
        adj[a] |= 1u << b;
        adj[b] |= 1u << a;
    }
    uint32_t seen = 1, frontier = 1;
    while (frontier) {
        uint32_t nxt = 0;
        for (uint32_t i = 0; i < p.num_vertices; ++i)
            if (frontier >> i & 1) nxt |= adj[i];
        frontier = nxt & ~seen;
        seen |= nxt;
    }
    if (seen != (1u << p.num_vertices) - 1) throw ArgError{TFSM_ERR_ARGUMENT, "pattern is not connected"};
}

// GQL order and backward neighbours per start vertex (reference
// generateGQLQueryPlan(int, ui*&) and generateBN(...), gmatch.h:692-727,
// 803-828): next = the unvisited vertex adjacent to the visited set with the
// fewest candidates, ties to the higher pattern degree, then the lower id;
// bn lists follow the pattern's ascending adjacency.
void make_plans(PatStatic &P, const uint32_t *slot_of /* [kMaxV*kMaxV] */) {
    const uint32_t nv = P.nv;
    for (uint32_t start = 0; start < nv; ++start) {
        bool visited[kMaxV] = {false}, adjacent[kMaxV] = {false};
        uint8_t order[kMaxV];
        auto visit = [&](uint32_t x) {
            visited[x] = true;
            for (uint32_t j = 0; j < P.nbr_n[x]; ++j) adjacent[P.nbr_v[x][j]] = true;
        };
        order[0] = (uint8_t)start;
        visit(start);
        for (uint32_t i = 1; i < nv; ++i) {
            int next = -1;
            for (uint32_t j = 0; j < nv; ++j) {
                if (visited[j] || !adjacent[j]) continue;
                if (next < 0 || P.dom_size[j] < P.dom_size[next] ||
                    (P.dom_size[j] == P.dom_size[next] && P.deg[j] > P.deg[next]))
                    next = (int)j;
            }
            order[i] = (uint8_t)next;
            visit((uint32_t)next);
        }
        bool vis[kMaxV] = {false};
        vis[order[0]] = true;
        uint32_t sec = 0;
        for (uint32_t d = 0; d < nv; ++d) {
            P.order[start][d] = order[d];
            P.bn_n[start][d] = 0;
            P.sec_base[start][d] = (uint8_t)std::min(sec, 255u);
            uint16_t before = 0;
            uint8_t after = 0;
            for (uint32_t e = 0; e < nv; ++e) {
                if (e == d || P.vlab[order[e]] != P.vlab[order[d]]) continue;
                if (e < d) before |= (uint16_t)(1u << e);
                else after = 1;
            }
            P.same_before[start][d] = before;
            P.same_after[start][d] = after;
            if (d == 0) continue;
            const uint32_t x = order[d];
            for (uint32_t j = 0; j < P.nbr_n[x]; ++j) {
                const uint32_t y = P.nbr_v[x][j];
                if (vis[y]) P.bn_slot[start][d][P.bn_n[start][d]++] = (uint8_t)slot_of[y * kMaxV + x];
            }
            sec += P.bn_n[start][d] - 1;
            vis[x] = true;
        }
    }
}

struct EvalCall {
    const tfsm_pattern *ps;
    const uint64_t *parents;
    uint32_t n;
    uint32_t freq;
    tfsm_eval_options opt;
    tfsm_pattern_result *out;
    bool level_call = false;        // from tfsm_eval_level (the default driver differs)
    bool force_host = false;        // redo of a device-driven call whose filter did not converge
};

// Waits for the stream by polling an event.  A blocking wait adds tens of
// microseconds of wake-up latency, paid two or three times per eval call on
// the critical path; polling costs one host core while the GPU works.
void sync_spin(tfsm_ctx *c) {
    HIPX(hipEventRecord(c->ev_sync, c->stream));
    hipError_t e;
    while ((e = hipEventQuery(c->ev_sync)) == hipErrorNotReady) {
    }
    HIPX(e);
}

// The Fraction-Score table (k_vertex_frac), computed once per context.
void ensure_scores(tfsm_ctx *c) {
    if (c->d_score) return;
    if (c->EL > kFracMaxEL) throw ArgError{TFSM_ERR_LIMIT, "fraction scores support at most 64 edge labels"};
    const unsigned long long cells = (unsigned long long)c->n * c->L;
    double *d = nullptr;
    uint32_t *d_off = nullptr;
    HIPX(hipMalloc(&d, std::max<unsigned long long>(cells, 1) * sizeof(double)));
    try {
        if (cells) {
            HIPX(hipMalloc(&d_off, c->bucket_off.size() * sizeof(uint32_t)));
            HIPX(hipMemcpyAsync(d_off, c->bucket_off.data(), c->bucket_off.size() * sizeof(uint32_t),
                                hipMemcpyHostToDevice, c->stream));
            hipLaunchKernelGGL(k_vertex_frac, dim3(grid_for(cells, c->num_cus)), dim3(kBlock), 0, c->stream, c->g,
                               c->d_bucket, d_off, c->L, d);
            HIPX(hipGetLastError());
            HIPX(hipStreamSynchronize(c->stream));
            HIPX(hipFree(d_off));
        }
    } catch (...) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(d_off);
        (void)hipFree(d);
        throw;
    }
    c->d_score = d;
}

// The double nearest an exact Fraction-Score sum (ties to even).
double frac_value(const FracSum &f) {
    const unsigned __int128 x = ((unsigned __int128)f.hi << 64) | f.lo;
    return std::ldexp((double)x, -96);
}

// Stats, per-pattern results and retained child domains of a finished call,
// from the states, counters and flags read back (both drivers).
// Buffers a device-driven call's sequence filled with every pattern's
// domains (one per pattern, null when the host copies what it keeps):
// finish_call takes those of the patterns it retains, and the rest go back
// to the cache when the call ends, on every path.
struct KeptBuffers {
    tfsm_ctx *c;
    std::vector<std::pair<void *, size_t>> buf;
    explicit KeptBuffers(tfsm_ctx *c_) : c(c_) {}
    ~KeptBuffers() {
        for (auto &b : buf)
            if (b.first) c->cache.put(b.first, b.second);
    }
};

struct Finished {
    std::vector<PatStatic> &hs;
    const PatState *hst;
    const SearchCounters *hctr;
    const uint32_t *hpool;          // pool used, pool overflow, filter passes run
    const std::vector<int> &decided;
    const FracSum *hfrac;
    unsigned long long total_tasks;
    uint32_t verify_launches;
    uint32_t level_mode;
    bool timed;                     // the stream events of every phase were recorded
    KeptBuffers *kept;              // the sequence's copies of every domain, or null
};

int finish_call(tfsm_ctx *c, const EvalCall &call, const Finished &f) {
    const uint32_t n = call.n;
    const tfsm_eval_options &opt = call.opt;
    const bool fraction = opt.mode == TFSM_MODE_FRACTION;
    std::vector<PatStatic> &hs = f.hs;
    const PatState *hst = f.hst;
    const SearchCounters *hctr = f.hctr;
    const uint32_t *hpool = f.hpool;
    const std::vector<int> &decided = f.decided;
    const FracSum *hfrac = f.hfrac;
    const unsigned long long total_tasks = f.total_tasks;
    const uint32_t verify_launches = f.verify_launches;
    hipStream_t s = c->stream;
    std::memset(&c->stats, 0, sizeof(c->stats));
    c->stats.level_mode = f.level_mode;
    c->stats.filter_passes = hpool[2];
    if (f.timed) {
        HIPX(hipEventElapsedTime(&c->stats.filter_ms, c->ev[0], c->ev[1]));
        HIPX(hipEventElapsedTime(&c->stats.table_ms, c->ev[1], c->ev[2]));
    }
    if (total_tasks) {
        if (f.timed) {
            HIPX(hipEventElapsedTime(&c->stats.search_ms, c->ev[4], c->ev[5]));
            HIPX(hipEventElapsedTime(&c->stats.rank_ms, c->ev[2], c->ev[4]));
        }
        c->stats.search_launches = verify_launches;
        c->stats.search_splits = std::min(hctr->split_tail, kSplitCap);
        c->stats.search_watchdog = hctr->watchdog ? 1u : 0u;
        if (verify_launches == 2) {
            c->stats.search_splits += std::min(hctr[1].split_tail, kSplitCap);
            if (hctr[1].watchdog) c->stats.search_watchdog = 1;
        }
    }
    {
        // algorithmic bytes (definitions in tfsm_gpu.h)
        // the pool reservation is the neighbour ids the table build scanned
        unsigned long long tb = 4ull * hpool[0], sb = 0;
        for (uint32_t p = 0; p < n; ++p) {
            if (decided[p] != TFSM_DECIDED_SEARCH) continue;
            const PatStatic &P = hs[p];
            const PatState &S = hst[p];
            for (uint32_t sl = 0; sl < P.nslots; ++sl) {
                const unsigned long long rows = P.dom_size[P.slot_a[sl]];
                tb += 16ull * rows + 4ull * S.aux_total[sl];
            }
            for (uint32_t u = 0; u < P.nv; ++u) {
                if (!P.dom_size[u]) continue;
                double per = 0;
                for (uint32_t sl = 0; sl < P.nslots; ++sl)
                    if (P.slot_a[sl] == u) per += 8.0 + 4.0 * (double)S.aux_total[sl] / (double)P.dom_size[u];
                sb += (unsigned long long)(per * (double)(S.valid_cnt[u] + S.invalid_cnt[u]));
            }
        }
        c->stats.table_bytes = tb;
        c->stats.search_bytes = sb;
    }

    // ---------------- results and retained child domains ----------------
    std::vector<CopyDesc> keep;                 // the retained domains' copies, one launch
    for (uint32_t p = 0; p < n; ++p) {
        const PatStatic &P = hs[p];
        const PatState &S = hst[p];
        tfsm_pattern_result &r = call.out[p];
        std::memset(&r, 0, sizeof(r));
        r.decided_by = decided[p];
        uint32_t lo = 0xffffffffu, hi = 0xffffffffu;
        for (uint32_t u = 0; u < P.nv; ++u) {
            r.domain_size[u] = P.dom_size[u];
            if (decided[p] == TFSM_DECIDED_EXPRESS) {
                r.valid_count[u] = P.dom_size[u];
            } else if (decided[p] == TFSM_DECIDED_SEARCH) {
                r.valid_count[u] = S.valid_cnt[u];
                r.invalid_count[u] = S.invalid_cnt[u];
            }
            lo = std::min(lo, r.valid_count[u]);
            hi = std::min(hi, r.domain_size[u] - r.invalid_count[u]);
            r.verdicts += r.valid_count[u] + r.invalid_count[u];
        }
        r.support_lower = lo;
        r.support_upper = hi;
        r.searches = S.searches;
        if (fraction) {
            // a pattern stopped by the valid-count bound (status 2) is
            // infrequent; every other search- or express-decided one has all
            // its verdicts, so its sums are exact
            double mn = 0;
            for (uint32_t u = 0; u < P.nv; ++u) {
                r.fraction[u] = frac_value(hfrac[(size_t)p * kMaxV + u]);
                mn = u == 0 ? r.fraction[u] : std::min(mn, r.fraction[u]);
            }
            r.fraction_support = mn;
            r.frequent = decided[p] != TFSM_DECIDED_FILTER && !(decided[p] == TFSM_DECIDED_SEARCH && S.status == 2) &&
                         mn >= (double)call.freq;
        } else if (decided[p] == TFSM_DECIDED_SEARCH && opt.mode == TFSM_MODE_CAPPED && S.status != 0) {
            r.frequent = S.status == 1;
        } else {
            r.frequent = decided[p] != TFSM_DECIDED_FILTER && lo >= call.freq;
        }
        if (r.frequent && opt.retain_domains && f.kept) {
            // the sequence wrote the domain and its invalid bits here
            Retained R;
            R.nv = P.nv;
            R.buf = f.kept->buf[p].first;
            R.cls = f.kept->buf[p].second;
            f.kept->buf[p].first = nullptr;
            for (uint32_t u = 0; u < P.nv; ++u) {
                R.vlab[u] = P.vlab[u];
                R.size[u] = P.dom_size[u];
                R.dom[u] = P.keep_dom[u];
                R.invalid[u] = P.keep_inv[u];
            }
            r.domain_handle = c->next_handle++;
            c->retained.emplace(r.domain_handle, R);
        } else if (r.frequent && opt.retain_domains) {
            size_t words = 0;
            for (uint32_t u = 0; u < P.nv; ++u) words += P.dom_size[u] + (P.dom_size[u] + 31) / 32;
            Retained R;
            R.nv = P.nv;
            for (uint32_t u = 0; u < P.nv; ++u) R.vlab[u] = P.vlab[u];
            R.buf = c->cache.get(std::max<size_t>(words, 1) * sizeof(uint32_t), &R.cls);
            uint32_t *b = static_cast<uint32_t *>(R.buf);
            for (uint32_t u = 0; u < P.nv; ++u) {
                const size_t wds = (P.dom_size[u] + 31) / 32;
                R.size[u] = P.dom_size[u];
                R.dom[u] = b; b += P.dom_size[u];
                R.invalid[u] = b; b += wds;
                if (P.dom_size[u]) keep.push_back({P.dom[u], R.dom[u], P.dom_size[u]});
                if (wds) keep.push_back({decided[p] == TFSM_DECIDED_SEARCH ? P.invalid_bits[u] : nullptr, R.invalid[u], wds});
            }
            r.domain_handle = c->next_handle++;
            c->retained.emplace(r.domain_handle, R);
        }
    }
    if (!keep.empty()) {
        const CopyDesc *d_keep = upload(c, keep);
        hipLaunchKernelGGL(k_copy_batch, dim3((uint32_t)std::min<size_t>(keep.size(), 4096)), dim3(kBlock), 0, s,
                           d_keep, (uint32_t)keep.size());
        HIPX(hipGetLastError());
    }
    c->last_static = hs;
    c->last_state.assign(hst, hst + n);
    c->last_pool_used = hpool[0];
    c->last_decided = decided;
    // the retained-domain copies stay queued: the next call waits for the stream before it
    // reuses the staging, and every function that reads device memory from the host syncs first
    return TFSM_OK;
}

// ---------------------------------------------------------------------------
// Device-driven calls (TFSM_LEVEL_DEVICE / TFSM_LEVEL_GRAPH)
// ---------------------------------------------------------------------------
constexpr int kNotTaken = 1000;     // not eligible: the host-driven path runs the call
constexpr int kRedoHost = 1001;     // the filter needed more passes: redo host-driven
constexpr int kRedoDevice = 1002;   // the pool was grown: redo the call

// Blocks for `items` work items of one thread each, rounded up to a power of
// two (so that levels of similar size share a captured sequence), at most cap.
uint32_t grid_class(unsigned long long items, uint32_t cap) {
    const unsigned long long blocks = (items + kBlock - 1) / kBlock;
    uint32_t g = 1;
    while (g < blocks && g < cap) g <<= 1;
    return std::min(g, cap);
}

constexpr uint32_t kGraphScanTiles = ((uint32_t)kPrioBins * kGraphPrioBlocks + kScanTile - 1) / kScanTile;
static_assert(kGraphScanTiles <= 4 * kBlock, "the tile sums fit the one-block scan");

// The launch sequence of a device-driven call.  Every kernel reads the
// context's descriptor, so the sequence depends only on the shape; with
// `events` the phase events are recorded between the kernels.
void launch_level(tfsm_ctx *c, const LevelShape &sh, bool events) {
    hipStream_t s = c->stream;
    const uint32_t slot = c->level_slot;
    if (events) HIPX(hipEventRecord(c->ev[0], s));
    hipLaunchKernelGGL(k_fills_g, dim3(sh.gfill), dim3(kBlock), 0, s, slot);
    hipLaunchKernelGGL(k_ldf_g, dim3(sh.gldf), dim3(kBlock), 0, s, slot);
    for (uint32_t i = 0; i < sh.passes; ++i) hipLaunchKernelGGL(k_prune_g, dim3(sh.gsrc), dim3(kBlock), 0, s, slot, i);
    hipLaunchKernelGGL(k_compact_g, dim3(sh.gpair), dim3(kBlock), 0, s, slot);    // and the plan
    if (events) HIPX(hipEventRecord(c->ev[1], s));
    hipLaunchKernelGGL(k_aux_build_g, dim3(sh.gaux), dim3(kBlock), 0, s, slot);
    if (events) HIPX(hipEventRecord(c->ev[2], s));
    if (sh.learned) {
        hipLaunchKernelGGL(k_rank_scores_g, dim3(sh.gsample), dim3(kBlock), 0, s, slot, 1u);
        hipLaunchKernelGGL(k_rank_scores_g, dim3(sh.grank), dim3(kBlock), 0, s, slot, 0u);
        hipLaunchKernelGGL(k_priority_g, dim3(kGraphPrioBlocks), dim3(kBlock), 0, s, slot);
        hipLaunchKernelGGL(k_scan_g, dim3(kGraphScanTiles), dim3(kBlock), 0, s, slot, 0u);
        hipLaunchKernelGGL(k_scan_g, dim3(1), dim3(kBlock), 0, s, slot, 1u);
        hipLaunchKernelGGL(k_scan_g, dim3(kGraphScanTiles), dim3(kBlock), 0, s, slot, 2u);
        hipLaunchKernelGGL(k_prio_scatter_g, dim3(kGraphPrioBlocks), dim3(kBlock), 0, s, slot);
    }
    if (events) HIPX(hipEventRecord(c->ev[4], s));
    for (uint32_t part = 0; part < (sh.learned ? 2u : 1u); ++part) {
        if (sh.me8)
            hipLaunchKernelGGL((k_verify_g<8, tfsm_lane::kOrderAdaptive, 8>), dim3(sh.vgrid), dim3(kBlock), 0, s, slot,
                               part);
        else
            hipLaunchKernelGGL((k_verify_g<8, tfsm_lane::kOrderAdaptive, 12>), dim3(sh.vgrid), dim3(kBlock), 0, s,
                               slot, part);
    }
    if (events) {
        HIPX(hipEventRecord(c->ev[5], s));
        HIPX(hipEventRecord(c->ev[3], s));
    }
    if (sh.keep) hipLaunchKernelGGL(k_keep_g, dim3(sh.gpair), dim3(kBlock), 0, s, slot);
    HIPX(hipGetLastError());
}

// The context's captured sequence for a shape: captured and instantiated on
// first use, then replayed.
hipGraphExec_t level_graph(tfsm_ctx *c, const LevelShape &sh) {
    for (auto &g : c->graphs)
        if (g.first == sh) return g.second;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    HIPX(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    bool launched = true;
    try {
        launch_level(c, sh, false);
    } catch (const HipError &) {
        launched = false;
    }
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (launched && e == hipSuccess && graph) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (!launched || e != hipSuccess || !exec) {
        (void)hipGetLastError();
        throw HipError{e == hipSuccess ? hipErrorUnknown : e, "capture of the level sequence"};
    }
    if (c->graphs.size() >= 16) {
        (void)hipGraphExecDestroy(c->graphs.front().second);
        c->graphs.erase(c->graphs.begin());
    }
    c->graphs.emplace_back(sh, exec);
    return exec;
}

// One call with no host round trip inside it: the filter, the device-side
// decisions and layout (plan_level), the tables, the learned order and the
// search, in one launch sequence over regions sized by the candidates'
// sources; the host reads the states and counters back once at the end.
int run_device(tfsm_ctx *c, const EvalCall &call, std::vector<PatStatic> &hs, const uint32_t *slot_of,
               const std::vector<char> &distinct, const std::vector<char> &acyclic, uint32_t *bits_region,
               size_t bits_words, PatState *d_st, const std::vector<uint32_t> &ex_off, const std::vector<PushSrc> &ex,
               int attempt) {
    const uint32_t n = call.n;
    const tfsm_eval_options &opt = call.opt;
    const uint32_t lm = opt.level_mode == TFSM_LEVEL_DEFAULT
                            ? (call.level_call ? (uint32_t)TFSM_LEVEL_GRAPH : (uint32_t)TFSM_LEVEL_HOST)
                            : opt.level_mode;
    if (lm != TFSM_LEVEL_DEVICE && lm != TFSM_LEVEL_GRAPH) return kNotTaken;
    const bool learned = opt.task_order == TFSM_ORDER_LEARNED;
    const bool wave_phase = (opt.long_steps == 0 ? kDefaultLongSteps : opt.long_steps) != 0xffffffffu;
    if (c->level_slot == kNoSlot || opt.trace || opt.mode == TFSM_MODE_FRACTION || wave_phase ||
        n > kDeviceMaxPatterns)
        return kNotTaken;
    // regions sized by the sources (every domain is at most its source)
    unsigned long long aux_ub = 0, ver_ub = 0, task_ub = 0, units_ub = 0, M = 0;
    uint32_t max_edges = 0;
    for (uint32_t p = 0; p < n; ++p) {
        PatStatic &P = hs[p];
        if (!P.adaptive || P.nv > 8) return kNotTaken;
        P.express = distinct[p] && acyclic[p];
        P.aux_base = (uint32_t)std::min<unsigned long long>(aux_ub, 0xffffffffull);
        P.verdict_base = (uint32_t)std::min<unsigned long long>(ver_ub, 0xffffffffull);
        P.aux_m0 = (uint32_t)M;
        for (uint32_t sl = 0; sl < P.nslots; ++sl) {
            aux_ub += 2ull * P.src_n[P.slot_a[sl]];
            units_ub += (P.src_n[P.slot_a[sl]] + 63ull) / 64 * 64;
        }
        for (uint32_t u = 0; u < P.nv; ++u) {
            ver_ub += 2ull * ((P.src_n[u] + 31ull) / 32);
            task_ub += P.src_n[u];
        }
        M += P.nslots;
        max_edges = std::max(max_edges, P.nslots / 2);
    }
    const size_t region_bytes = 4 * (aux_ub + ver_ub) + (learned ? 20 * task_ub : 0);
    if (aux_ub >= (1ull << 31) || ver_ub >= (1ull << 31) || task_ub >= (1ull << 31) || units_ub >= (1ull << 31) ||
        region_bytes > kDeviceRegionBudget)
        return kNotTaken;
    hipStream_t s = c->stream;
    // DP-iso passes of the sequence: prune_batch, or a margin over the
    // host-driven batch by default; a level that needs more is redone
    // host-driven (a pass after the fixpoint returns at once)
    const uint32_t passes = opt.prune_batch ? opt.prune_batch : kPruneBatch + 2;
    c->trace_len = 0;

    // regions and work lists
    uint32_t *aux_region = c->work.alloc<uint32_t>(std::max<unsigned long long>(aux_ub, 1));
    uint32_t *verdict_region = c->work.alloc<uint32_t>(std::max<unsigned long long>(ver_ub, 1));
    unsigned long long *task_base = c->work.alloc<unsigned long long>(n + 1);
    unsigned long long *aux_base = c->work.alloc<unsigned long long>(M + 1);
    const uint32_t nchg = n * kMaxV;
    uint32_t *d_chg = c->work.alloc<uint32_t>(3 * (size_t)nchg);
    HostWork src_work, ldf_work;
    std::vector<uint32_t> pair_p, pair_u, aux_pat, aux_x;
    for (uint32_t p = 0; p < n; ++p) {
        for (uint32_t u = 0; u < hs[p].nv; ++u) {
            src_work.add(p, u, hs[p].src_n[u]);
            const uint32_t px = hs[p].push_x[u];
            if (px == kNoPush) ldf_work.add(p, u, hs[p].src_n[u]);
            else ldf_work.add(p, u | kPushItem, hs[p].src_n[px]);
            pair_p.push_back(p);
            pair_u.push_back(u);
        }
        for (uint32_t sl = 0; sl < hs[p].nslots; ++sl) {
            aux_pat.push_back(p);
            aux_x.push_back(sl);
        }
    }
    // retained domains: when their bounds fit, every pattern gets a buffer
    // for its domains at their source sizes, filled in the sequence (the
    // compaction writes the domains, its last kernel their invalid bits), so
    // the call ends with no copies left to launch
    KeptBuffers kept(c);
    if (opt.retain_domains) {
        unsigned long long keep_words = 0;
        for (uint32_t p = 0; p < n; ++p)
            for (uint32_t u = 0; u < hs[p].nv; ++u) keep_words += hs[p].src_n[u] + (hs[p].src_n[u] + 31ull) / 32;
        if (keep_words * 4 <= kDeviceKeepBudget) {
            kept.buf.resize(n);
            for (uint32_t p = 0; p < n; ++p) {
                PatStatic &P = hs[p];
                size_t words = 0;
                for (uint32_t u = 0; u < P.nv; ++u) words += P.src_n[u] + (P.src_n[u] + 31) / 32;
                kept.buf[p].first = c->cache.get(std::max<size_t>(words, 1) * sizeof(uint32_t), &kept.buf[p].second);
                uint32_t *w = static_cast<uint32_t *>(kept.buf[p].first);
                for (uint32_t u = 0; u < P.nv; ++u) {
                    P.keep_dom[u] = w; w += P.src_n[u];
                    P.keep_inv[u] = w; w += (P.src_n[u] + 31) / 32;
                }
            }
        }
    }
    LevelArgs *L = c->h_level.ensure(1);
    std::memset(L, 0, sizeof(*L));
    const size_t out_bytes = kOutStates + (n + 1) * sizeof(PatState);
    const uint32_t group_max = opt.group_lanes == 0 ? kDefaultGroupLanes : opt.group_lanes;
    const size_t gz = std::min<unsigned long long>(task_ub, kGroupCap);
    PhaseBatch b(c, n * sizeof(PatStatic) + src_work.bytes() + ldf_work.bytes() + 4 * ex_off.size() +
                        sizeof(PushSrc) * ex.size() +
                        8 * pair_p.size() + 8 * aux_pat.size() + 6 * 256 +
                        sizeof(CopyDesc) * ((bits_words + ver_ub) / kFillPiece + (n + 1) * sizeof(PatState) / 4 / kFillPiece +
                                            (size_t)n * kMaxV * (2 + kPrioQ) / kFillPiece + 2 * gz / kFillPiece + 24) +
                        out_bytes + 256);
    // the block read back after the call, uploaded with the batch: the phase
    // clock (its minimum slots at the largest value), both search launches'
    // counters, the pool, overflow and pass words, and the pattern states
    // (with the filter's last change flag after them), otherwise zeros
    unsigned char *d_out;
    {
        std::vector<unsigned long long> out0((out_bytes + 7) / 8, 0ull);
        for (uint32_t w = 0; w < kClockWays; ++w) out0[w] = out0[3 * kClockWays + w] = ~0ull;
        d_out = reinterpret_cast<unsigned char *>(b.add(out0));
    }
    unsigned long long *d_clk = reinterpret_cast<unsigned long long *>(d_out);
    SearchCounters *d_ctr = reinterpret_cast<SearchCounters *>(d_out + kOutCtr);
    uint32_t *d_flags = reinterpret_cast<uint32_t *>(d_out + kOutFlags);
    d_st = reinterpret_cast<PatState *>(d_out + kOutStates);
    L->ps = b.add(hs.data(), n);
    L->wsrc = src_work.to_device(b);
    L->wldf = ldf_work.to_device(b);
    L->ex_off = ex.empty() ? nullptr : b.add(ex_off);
    L->ex = ex.empty() ? nullptr : b.add(ex);
    L->pairs_p = b.add(pair_p);
    L->pairs_u = b.add(pair_u);
    L->aux_pat = b.add(aux_pat);
    L->aux_x = b.add(aux_x);
    b.zero(bits_region, bits_words * sizeof(uint32_t));
    b.zero(c->d_pflags, passes * sizeof(uint32_t));
    b.zero(d_chg, 3 * (size_t)nchg * sizeof(uint32_t));
    if (ver_ub) b.zero(verdict_region, ver_ub * sizeof(uint32_t));
    if (group_max > 1 && gz) {
        b.zero(c->d_piece_cnt + kSplitCap, gz * sizeof(uint32_t));
        b.zero(c->d_piece_cnt + kSplitCap + kGroupCap, gz * sizeof(uint32_t));
    }
    float *rank_sums = nullptr;
    if (learned) {
        rank_sums = c->work.alloc<float>((size_t)n * kMaxV * (2 + kPrioQ));
        b.zero(rank_sums, (size_t)n * kMaxV * (2 + kPrioQ) * sizeof(float));
    }
    L->nfills = (uint32_t)b.zeros.size();
    L->fills = b.add(b.zeros);
    b.zeros.clear();
    L->clk = d_clk;
    L->P = n;
    L->g = c->g;
    L->st = d_st;
    L->pflags = c->d_pflags;
    L->chg = d_chg;
    L->nchg = nchg;
    L->passes = passes;
    L->npairs = (uint32_t)pair_p.size();
    L->passes_run = d_flags + 2;
    L->blocks_done = d_flags + 3;
    L->aux_region = aux_region;
    L->verdict_region = verdict_region;
    L->task_base = task_base;
    L->aux_base = aux_base;
    L->aux_M = (uint32_t)M;
    L->pool = c->d_pool;
    L->pool_used = d_flags;
    L->overflow = d_flags + 1;
    L->pool_cap = c->pool_cap;
    uint32_t *hist = nullptr;
    if (learned) {
        L->W = c->d_rank_w;
        L->logits = c->work.alloc<float>(std::max<unsigned long long>(task_ub, 1));
        L->sums = rank_sums;
        L->qsums = rank_sums + (size_t)n * kMaxV;
        L->ssums = L->qsums + (size_t)n * kMaxV * kPrioQ;
        L->keys = c->work.alloc<uint32_t>(std::max<unsigned long long>(task_ub, 1));
        L->pu = c->work.alloc<uint32_t>(std::max<unsigned long long>(task_ub, 1));
        L->order = c->work.alloc<uint32_t>(std::max<unsigned long long>(task_ub, 1));
        L->order_pu = c->work.alloc<uint32_t>(std::max<unsigned long long>(task_ub, 1));
        L->nhist = kPrioBins * kGraphPrioBlocks;
        L->ntiles = kGraphScanTiles;
        hist = c->work.alloc<uint32_t>(L->nhist);
        L->hist = hist;
        L->tiles = c->work.alloc<uint32_t>(kGraphScanTiles);
    }
    // the search launches (the same settings as the host-driven sequence)
    const uint32_t per_cu = opt.grid_waves_per_cu
                                ? std::min((uint32_t)c->verify_blocks8a, std::max(1u, opt.grid_waves_per_cu / (kBlock / 64)))
                                : (uint32_t)c->verify_blocks8a;
    // (verify_blocks8a: the fewer resident blocks of the ME = 8 and ME = 12 instances)
    for (uint32_t part = 0; part < 2; ++part) {
        VerifyArgs &va = L->va[part];
        va.ps = L->ps;
        va.st = d_st;
        va.task_base = task_base;
        va.P = n;
        va.epoch = ++c->epoch;
        va.ctr = d_ctr + part;
        va.pool = c->d_pool;
        va.aux_region = aux_region;
        va.max_edges = max_edges;
        va.overflow = d_flags + 1;
        va.steps_per_round = opt.steps_per_round ? opt.steps_per_round : kDefaultStepsPerRound;
        va.refill_min = opt.refill_min ? opt.refill_min : kDefaultRefillMin;
        va.refill_max = opt.refill_max ? opt.refill_max : kDefaultRefillMax;
        va.split_steps = opt.split_steps == 0 ? kDefaultSplitSteps : opt.split_steps;
        if (opt.split_steps == 0xffffffffu) va.split_steps = 0;
        va.split_fanout = opt.split_fanout == 0 ? kDefaultSplitFanout : std::min(opt.split_fanout, 64u);
        va.records = c->d_records;
        va.split_cap = kSplitCap;
        va.piece_cnt = c->d_piece_cnt;
        va.piece_incomplete = c->d_piece_inc;
        va.group_max = group_max;
        va.hot_group_max = opt.hot_group_lanes ? opt.hot_group_lanes : group_max;
        va.group_base = kSplitCap + part * kGroupCap;
        va.group_cap = kGroupCap;
        va.long_steps = 0xffffffffu;
        if (learned) {
            va.order = L->order;
            va.order_pu = L->order_pu;
            va.abort_stopped = 1;
            va.part = part + 1;
            // after the scan, the first deferred bin of block 0 holds the
            // number of first-part entries; the first key after the hot tier,
            // the hot tier's size
            va.split = hist + (size_t)kDeferredKey * kGraphPrioBlocks;
            va.hot = hist + (size_t)kHotKeys * kGraphPrioBlocks;
        }
    }
    LevelShape sh;
    std::memset(&sh, 0, sizeof(sh));
    sh.learned = learned;
    sh.passes = passes;
    sh.gfill = grid_class((unsigned long long)L->nfills * kBlock, 4096);
    sh.gsrc = grid_class(src_work.base.back(), (uint32_t)c->num_cus * 16);
    sh.gldf = grid_class(ldf_work.base.back(), (uint32_t)c->num_cus * 16);
    sh.gpair = grid_class((unsigned long long)pair_p.size() * kBlock, (uint32_t)c->num_cus * 16);
    // the table build's waves loop over the work (the sizes are known only on
    // the device): two blocks per CU; more resident waves than that made the
    // build slower (its scattered loads contend in the L1)
    sh.gaux = grid_class(units_ub, (uint32_t)c->num_cus * 2);
    // scoring: every block first loads the weights into LDS, so the waves
    // loop over the tasks with eight blocks per CU at most
    sh.gsample = grid_class((task_ub + kSample - 1) / kSample, (uint32_t)c->num_cus * 8);
    sh.grank = grid_class(task_ub, (uint32_t)c->num_cus * 8);
    sh.vgrid = (uint32_t)c->num_cus * per_cu;
    sh.me8 = max_edges <= 8;
    sh.keep = !kept.buf.empty();
    // the batch's uploads, then the descriptor, then the sequence
    b.flush();
    HIPX(hipMemcpyAsync(c->d_level, L, sizeof(LevelArgs), hipMemcpyHostToDevice, s));
    // stream events between the kernels (kernel by kernel) or the phase
    // clock the kernels keep (a graph: events recorded inside a captured
    // sequence carry no times the host can read here)
    const bool graph = lm == TFSM_LEVEL_GRAPH;
    if (graph) HIPX(hipGraphLaunch(level_graph(c, sh), s));
    else launch_level(c, sh, true);
    // one readback: the clock, the counters, the pool and pass words and the
    // states (with the filter's change flag)
    unsigned char *hout = c->h_out.ensure(out_bytes);
    HIPX(hipMemcpyAsync(hout, d_out, out_bytes, hipMemcpyDeviceToHost, s));
    const unsigned long long *hclk = reinterpret_cast<const unsigned long long *>(hout);
    const SearchCounters *hctr = reinterpret_cast<const SearchCounters *>(hout + kOutCtr);
    const uint32_t *hpool = reinterpret_cast<const uint32_t *>(hout + kOutFlags);
    const PatState *hst = reinterpret_cast<const PatState *>(hout + kOutStates);
    sync_spin(c);
    if (*reinterpret_cast<const uint32_t *>(hst + n)) return kRedoHost;
    if (hpool[1]) {
        // the tables did not fit: grow the pool to twice the measured need and redo the call
        if (attempt >= 4) throw ArgError{TFSM_ERR_LIMIT, "table pool overflow"};
        const unsigned long long need = (unsigned long long)hpool[0] * 2;
        if (need > 0xfffffff0ull) throw ArgError{TFSM_ERR_LIMIT, "tables exceed 2^32 entries"};
        HIPX(hipFree(c->d_pool));
        c->d_pool = nullptr;
        c->pool_cap = (uint32_t)std::max<unsigned long long>(need, c->pool_cap);
        HIPX(hipMalloc(&c->d_pool, (size_t)c->pool_cap * sizeof(uint32_t)));
        return kRedoDevice;
    }
    // the device's layout, repeated on the host for the results, the
    // retained domains and the inspection functions
    std::vector<int> decided(n);
    unsigned long long total_tasks = 0;
    for (uint32_t p = 0; p < n; ++p) {
        decided[p] = (int)plan_pattern(hs[p], hst[p].dom_size, aux_region, verdict_region);
        if ((uint32_t)decided[p] != hst[p].decided) throw ArgError{TFSM_ERR_HIP, "device layout differs from the host's"};
        if (decided[p] == TFSM_DECIDED_SEARCH) {
            if (n == 1) make_plans(hs[p], slot_of + (size_t)p * kMaxV * kMaxV);
            total_tasks += hs[p].task_prefix[hs[p].nv];
        }
    }
    const int rc = finish_call(c, call, Finished{hs, hst, hctr, hpool, decided, nullptr, total_tasks,
                                                 learned ? 2u : 1u, lm, !graph, sh.keep ? &kept : nullptr});
    // the stream is idle when the sequence wrote the kept domains or none
    // were retained; otherwise finish_call queued their copies
    c->stream_idle = sh.keep || !opt.retain_domains;
    if (graph) {
        // the phase clock in ms: first block .. compaction .. tables .. search start .. end
        const double per_ms = (double)c->wall_khz;
        // each stamp: the minimum (starts) or maximum (ends) over its words
        unsigned long long st[5];
        for (uint32_t k = 0; k < 5; ++k) {
            const bool start = k == 0 || k == 3;
            st[k] = start ? ~0ull : 0ull;
            for (uint32_t w = 0; w < kClockWays; ++w) {
                const unsigned long long x = hclk[k * kClockWays + w];
                st[k] = start ? std::min(st[k], x) : std::max(st[k], x);
            }
        }
        auto span = [&](int a, int b) { return st[b] > st[a] ? (float)((double)(st[b] - st[a]) / per_ms) : 0.f; };
        c->stats.filter_ms = span(0, 1);
        c->stats.table_ms = span(1, 2);
        if (total_tasks) {
            c->stats.rank_ms = span(2, 3);
            c->stats.search_ms = span(3, 4);
        }
    }
    return rc;
}

int run_eval(tfsm_ctx *c, const EvalCall &call, int attempt = 0) {
    const uint32_t n = call.n;
    const tfsm_eval_options &opt = call.opt;
    if (opt.mode != TFSM_MODE_CAPPED && opt.mode != TFSM_MODE_EXACT && opt.mode != TFSM_MODE_FRACTION)
        throw ArgError{TFSM_ERR_ARGUMENT, "unknown mode"};
    const bool fraction = opt.mode == TFSM_MODE_FRACTION;
    if (opt.task_order != TFSM_ORDER_REFERENCE && opt.task_order != TFSM_ORDER_LEARNED)
        throw ArgError{TFSM_ERR_ARGUMENT, "unknown task order"};
    if (opt.reserved != 0) throw ArgError{TFSM_ERR_ARGUMENT, "reserved option field must be 0"};
    if (opt.group_lanes > 64 || (opt.group_lanes & (opt.group_lanes - 1)))
        throw ArgError{TFSM_ERR_ARGUMENT, "group_lanes must be 0 or a power of two up to 64"};
    if (opt.hot_group_lanes > 64 || (opt.hot_group_lanes & (opt.hot_group_lanes - 1)))
        throw ArgError{TFSM_ERR_ARGUMENT, "hot_group_lanes must be 0 or a power of two up to 64"};
    if (opt.match_order != TFSM_MATCH_ADAPTIVE && opt.match_order != TFSM_MATCH_REFERENCE)
        throw ArgError{TFSM_ERR_ARGUMENT, "unknown matching order"};
    if (opt.steps_per_round > 64) throw ArgError{TFSM_ERR_ARGUMENT, "steps_per_round must be at most 64"};
    {
        const uint32_t lo = opt.refill_min ? opt.refill_min : kDefaultRefillMin;
        const uint32_t hi = opt.refill_max ? opt.refill_max : kDefaultRefillMax;
        if (lo > hi || hi > 4096) throw ArgError{TFSM_ERR_ARGUMENT, "refill_min <= refill_max <= 4096 required"};
    }
    if (opt.prune_batch > 64) throw ArgError{TFSM_ERR_ARGUMENT, "prune_batch must be at most 64"};
    if (opt.level_mode > TFSM_LEVEL_GRAPH) throw ArgError{TFSM_ERR_ARGUMENT, "unknown level mode"};
    const uint32_t prune_batch = opt.prune_batch ? opt.prune_batch : kPruneBatch;
    const bool learned = opt.task_order == TFSM_ORDER_LEARNED;
    if (learned && !c->ranker_installed)
        throw ArgError{TFSM_ERR_ARGUMENT, "learned task order requires an installed ranker"};
    if (call.freq == 0) throw ArgError{TFSM_ERR_ARGUMENT, "freq must be at least 1"};
    for (uint32_t i = 0; i < n; ++i) {
        check_pattern(c, call.ps[i]);
        uint64_t h = call.parents ? call.parents[i] : 0;
        if (h) {
            auto it = c->retained.find(h);
            if (it == c->retained.end()) throw ArgError{TFSM_ERR_HANDLE, "unknown parent domain handle"};
            if (it->second.nv > call.ps[i].num_vertices)
                throw ArgError{TFSM_ERR_ARGUMENT, "parent has more vertices than the pattern"};
        }
    }
    c->work.reset();
    hipStream_t s = c->stream;
    // nothing of an earlier call still reads the staging (a device-driven
    // call that queued nothing after its final sync left the stream idle)
    if (!c->stream_idle) sync_spin(c);
    c->stream_idle = false;
    c->stage.reset();
    if (fraction) ensure_scores(c);
    HIPX(hipEventRecord(c->ev[0], s));

    // ---------------- host description of every pattern ----------------
    std::vector<PatStatic> hs(n);
    std::vector<uint32_t> slot_of((size_t)n * kMaxV * kMaxV, 0xffffffffu);
    const uint32_t push_pct = opt.ldf_push == 0 ? kLdfPushPct : opt.ldf_push == 0xffffffffu ? 0 : opt.ldf_push;
    std::vector<char> distinct(n), acyclic(n);
    size_t bits_words = 0, dom_words = 0;
    for (uint32_t p = 0; p < n; ++p) {
        const tfsm_pattern &tp = call.ps[p];
        PatStatic &P = hs[p];
        std::memset(&P, 0, sizeof(P));
        P.nv = tp.num_vertices;
        P.freq = call.freq;
        P.mode = opt.mode;
        for (uint32_t i = 0; i < P.nv; ++i) P.vlab[i] = tp.vlabel[i];
        uint32_t el_of[kMaxV][kMaxV];
        bool adj[kMaxV][kMaxV] = {{false}};
        for (uint32_t e = 0; e < tp.num_edges; ++e) {
            const uint32_t a = tp.edge_a[e], b = tp.edge_b[e];
            adj[a][b] = adj[b][a] = true;
            el_of[a][b] = el_of[b][a] = tp.edge_label[e];
            uint32_t s0 = P.nslots++, s1 = P.nslots++;
            P.slot_a[s0] = (uint8_t)a; P.slot_b[s0] = (uint8_t)b;
            P.slot_key[s0] = P.vlab[b] * c->EL + tp.edge_label[e];
            P.slot_a[s1] = (uint8_t)b; P.slot_b[s1] = (uint8_t)a;
            P.slot_key[s1] = P.vlab[a] * c->EL + tp.edge_label[e];
            slot_of[(size_t)p * kMaxV * kMaxV + a * kMaxV + b] = s0;
            slot_of[(size_t)p * kMaxV * kMaxV + b * kMaxV + a] = s1;
        }
        for (uint32_t a = 0; a < P.nv; ++a) {
            uint32_t cnt_by_label[kMaxV], lab_of[kMaxV], nl = 0;
            for (uint32_t b = 0; b < P.nv; ++b) {
                if (!adj[a][b]) continue;
                P.nbr_v[a][P.nbr_n[a]] = (uint8_t)b;
                P.nbr_key[a][P.nbr_n[a]] = P.vlab[b] * c->EL + el_of[a][b];
                P.need_mask[a] |= 1ull << (P.nbr_key[a][P.nbr_n[a]] & 63);
                P.nbr_n[a]++;
                uint32_t k = 0;
                while (k < nl && lab_of[k] != P.vlab[b]) ++k;
                if (k == nl) { lab_of[nl] = P.vlab[b]; cnt_by_label[nl] = 0; ++nl; }
                cnt_by_label[k]++;
            }
            P.deg[a] = P.nbr_n[a];
            P.nlf_n[a] = nl;
            for (uint32_t k = 0; k < nl; ++k) { P.nlf_lab[a][k] = lab_of[k]; P.nlf_cnt[a][k] = cnt_by_label[k]; }
        }
        const uint64_t h = call.parents ? call.parents[p] : 0;
        const Retained *par = h ? &c->retained.at(h) : nullptr;
        for (uint32_t u = 0; u < P.nv; ++u) {
            if (par && u < par->nv) {
                P.src[u] = par->dom[u];
                P.src_invalid[u] = par->invalid[u];
                P.src_n[u] = par->size[u];
            } else {
                P.src[u] = c->d_bucket + c->bucket_off[P.vlab[u]];
                P.src_mask[u] = c->d_keymask_bucket + c->bucket_off[P.vlab[u]];
                P.src_invalid[u] = nullptr;
                P.src_n[u] = c->bucket_off[P.vlab[u] + 1] - c->bucket_off[P.vlab[u]];
            }
            bits_words += 2 * (size_t)c->W;
            dom_words += std::max<uint32_t>(P.src_n[u], 1);
        }
        // a label-bucket vertex next to a vertex with a parent domain takes
        // its LDF candidates from that domain's neighbours when a bound on
        // the neighbours walked is at most push_pct percent of its bucket
        // (the neighbour with the smallest bound)
        for (uint32_t u = 0; u < P.nv; ++u) {
            P.push_x[u] = kNoPush;
            if (P.src_mask[u] == nullptr || push_pct == 0 || c->fan_n.empty()) continue;
            unsigned long long best = (unsigned long long)P.src_n[u] * push_pct / 100 + 1;
            for (uint32_t j = 0; j < P.nbr_n[u]; ++j) {
                const uint32_t x = P.nbr_v[u][j];
                if (P.src_mask[x] != nullptr) continue;
                const uint32_t key = P.nbr_key[u][j] % c->EL + P.vlab[u] * c->EL;
                const unsigned long long est = fan_bound(c, P.vlab[x], key, P.src_n[x]);
                if (est >= best) continue;
                best = est;
                P.push_x[u] = (uint8_t)x;
                P.push_key[u] = key;
            }
        }
        bool dl = true;
        for (uint32_t a = 0; a < P.nv; ++a)
            for (uint32_t b = a + 1; b < P.nv; ++b) if (P.vlab[a] == P.vlab[b]) dl = false;
        for (uint32_t a = 0; a < P.nv; ++a)
            for (uint32_t b = 0; b < P.nv; ++b)
                if (a != b && P.vlab[a] == P.vlab[b]) P.same_lab[a] |= (uint16_t)(1u << b);
        // adaptive matching order: at most kCutMax edges, and the non-tree
        // edges (the non-primary rows over all depths) fit the cursor slots
        // of the search kernel instance (8 for <= 8 vertices, else 16) and
        // the 4-bit depth metadata
        {
            const uint32_t slots = P.nv <= 8 ? 8u : 15u;
            P.adaptive = opt.match_order == TFSM_MATCH_ADAPTIVE && tp.num_edges <= tfsm_lane::kCutMax &&
                         tp.num_edges - (P.nv - 1) <= slots;
            for (uint32_t e = 0; e < (uint32_t)kCutEdges; ++e) {
                const uint32_t b = e < tp.num_edges ? (tp.edge_a[e] | tp.edge_b[e] << 4) : 0xffu;
                P.edge_pack[e / 4] |= b << (8 * (e % 4));
            }
        }
        distinct[p] = dl;
        acyclic[p] = tp.num_edges == tp.num_vertices - 1;
    }
    // by default, pushing is kept to calls whose pushed buckets hold many
    // items per GPU thread: a call of a few patterns checks every bucket item
    // on its own thread in about one round, while a pushed item walks a whole
    // adjacency segment (thousands of entries at a hub), so there the walks'
    // latency costs more than the bucket's work (R-MAT scale 20, one pattern
    // per call: 4x the filter time)
    if (opt.ldf_push == 0) {
        unsigned long long pushed_items = 0;
        for (uint32_t p = 0; p < n; ++p)
            for (uint32_t u = 0; u < hs[p].nv; ++u)
                if (hs[p].push_x[u] != kNoPush) pushed_items += hs[p].src_n[u];
        if (pushed_items < (unsigned long long)c->num_cus * kBlock * kPushMinItemsPerThread)
            for (uint32_t p = 0; p < n; ++p)
                for (uint32_t u = 0; u < hs[p].nv; ++u) hs[p].push_x[u] = kNoPush;
    }
    // one region for all bitmaps (zeroed at once) and one for all lists
    uint32_t *bits_region = c->work.alloc<uint32_t>(bits_words);
    uint32_t *dom_region = c->work.alloc<uint32_t>(dom_words);
    {
        size_t bo = 0, dofs = 0;
        for (uint32_t p = 0; p < n; ++p)
            for (uint32_t u = 0; u < hs[p].nv; ++u) {
                hs[p].bits[u] = bits_region + bo; bo += 2 * (size_t)c->W;   // (bits, rank) pairs
                hs[p].dom[u] = dom_region + dofs; dofs += std::max<uint32_t>(hs[p].src_n[u], 1);
            }
    }
    // the pattern states, then one word: the filter's last change flag
    PatState *d_st = c->work.alloc<PatState>(n + 1);

    // push-down sources per (pattern, vertex): the retained child domains
    // the candidates must lie in
    std::vector<uint32_t> ex_off;
    std::vector<PushSrc> ex;
    if (opt.num_pushdown) {
        if (!opt.pushdown) throw ArgError{TFSM_ERR_ARGUMENT, "num_pushdown without a pushdown array"};
        std::vector<std::vector<PushSrc>> per((size_t)n * kMaxV);
        for (uint32_t i = 0; i < opt.num_pushdown; ++i) {
            const tfsm_pushdown &d = opt.pushdown[i];
            if (d.reserved != 0) throw ArgError{TFSM_ERR_ARGUMENT, "reserved pushdown field must be 0"};
            if (d.pattern >= n || d.vertex >= call.ps[d.pattern].num_vertices)
                throw ArgError{TFSM_ERR_ARGUMENT, "pushdown entry names no vertex of the call"};
            auto it = c->retained.find(d.handle);
            if (it == c->retained.end()) throw ArgError{TFSM_ERR_HANDLE, "unknown pushdown domain handle"};
            const Retained &R = it->second;
            if (d.source_vertex >= R.nv) throw ArgError{TFSM_ERR_ARGUMENT, "pushdown source vertex out of range"};
            if (call.ps[d.pattern].vlabel[d.vertex] != R.vlab[d.source_vertex])
                throw ArgError{TFSM_ERR_ARGUMENT, "pushdown source vertex has another label"};
            per[(size_t)d.pattern * kMaxV + d.vertex].push_back(
                PushSrc{R.dom[d.source_vertex], R.invalid[d.source_vertex], R.size[d.source_vertex], 0});
        }
        ex_off.assign((size_t)n * kMaxV + 1, 0);
        for (size_t k = 0; k < per.size(); ++k) {
            ex_off[k + 1] = ex_off[k] + (uint32_t)per[k].size();
            ex.insert(ex.end(), per[k].begin(), per[k].end());
        }
    }

    // ---------------- device-driven calls ----------------
    if (!call.force_host) {
        const int r = run_device(c, call, hs, slot_of.data(), distinct, acyclic, bits_region, bits_words, d_st, ex_off,
                                 ex, attempt);
        if (r == kRedoHost) {
            EvalCall h = call;
            h.force_host = true;
            return run_eval(c, h, 0);
        }
        if (r == kRedoDevice) return run_eval(c, call, attempt + 1);
        if (r != kNotTaken) return r;
    }

    // ---------------- filter: LDF, DP-iso to the fixpoint, compaction ----------------
    HostWork src_work, ldf_work;
    std::vector<uint32_t> pair_p, pair_u;
    for (uint32_t p = 0; p < n; ++p)
        for (uint32_t u = 0; u < hs[p].nv; ++u) {
            src_work.add(p, u, hs[p].src_n[u]);
            const uint32_t px = hs[p].push_x[u];
            if (px == kNoPush) ldf_work.add(p, u, hs[p].src_n[u]);
            else ldf_work.add(p, u | kPushItem, hs[p].src_n[px]);
            pair_p.push_back(p);
            pair_u.push_back(u);
        }
    // this phase's uploads in one copy, its fills in one launch
    PhaseBatch b1(c, n * sizeof(PatStatic) + src_work.bytes() + ldf_work.bytes() + 2 * 256 + 8 * pair_p.size() +
                         4 * ex_off.size() +
                         sizeof(PushSrc) * ex.size() + 512 +
                         sizeof(CopyDesc) * (bits_words / kFillPiece + (n + 1) * sizeof(PatState) / 4 / kFillPiece + 8) +
                         512);
    PatStatic *d_ps = b1.add(hs.data(), n);
    WorkTable wsrc = src_work.to_device(b1);
    const WorkTable wldf = ldf_work.to_device(b1);
    const uint32_t *d_ex_off = ex.empty() ? nullptr : b1.add(ex_off);
    const PushSrc *d_ex = ex.empty() ? nullptr : b1.add(ex);
    const uint32_t *d_pair_p = b1.add(pair_p), *d_pair_u = b1.add(pair_u);
    b1.zero(bits_region, bits_words * sizeof(uint32_t));
    b1.zero(d_st, (n + 1) * sizeof(PatState));
    b1.zero(c->d_flags, 3 * sizeof(uint32_t));        // pool used, pool overflow, filter passes run
    if (wsrc.total) b1.zero(c->d_pflags, kMaxPrunePasses * sizeof(uint32_t));
    // per pass, the pattern vertices that lost candidates (three arrays in rotation)
    const uint32_t nchg = n * kMaxV;
    uint32_t *d_chg = c->work.alloc<uint32_t>(3 * (size_t)nchg);
    b1.zero(d_chg, 3 * (size_t)nchg * sizeof(uint32_t));
    b1.flush();
    const uint32_t gsrc = grid_for(wsrc.total, c->num_cus);
    if (wldf.total)
        hipLaunchKernelGGL(k_ldf, dim3(grid_for(wldf.total, c->num_cus)), dim3(kBlock), 0, s, c->g, d_ps, wldf,
                           d_ex_off, (const PushSrc *)d_ex);
    HIPX(hipGetLastError());
    // DP-iso passes in batches without host round trips: a pass returns at
    // once when the one before it removed nothing (the fixpoint is reached);
    // the last flag of a batch, read with the compacted domains, says whether
    // another batch is needed (compaction is then simply redone)
    PatState *hst = c->h_state.ensure(n + 1);
    uint32_t *d_more = reinterpret_cast<uint32_t *>(d_st + n);
    const uint32_t *h_more = reinterpret_cast<const uint32_t *>(hst + n);
    uint32_t passes = 0;
    for (;;) {
        for (uint32_t i = 0; wsrc.total && i < prune_batch; ++i, ++passes) {
            if (passes >= kMaxPrunePasses) throw ArgError{TFSM_ERR_LIMIT, "filter did not converge"};
            hipLaunchKernelGGL(k_prune, dim3(gsrc), dim3(kBlock), 0, s, c->g, d_ps, wsrc, c->d_pflags + passes,
                               passes ? c->d_pflags + passes - 1 : nullptr,
                               passes ? d_chg + (size_t)((passes + 2) % 3) * nchg : nullptr,
                               d_chg + (size_t)(passes % 3) * nchg, d_chg + (size_t)((passes + 1) % 3) * nchg, nchg,
                               c->d_flags + 2);
        }
        if (pair_p.size() >= kCompactStageMinPairs)
            hipLaunchKernelGGL(k_compact<true>, dim3((uint32_t)pair_p.size()), dim3(kBlock), 0, s, c->g, d_ps, d_st,
                               d_pair_p, d_pair_u, passes ? c->d_pflags + passes - 1 : (const uint32_t *)nullptr, d_more);
        else
            hipLaunchKernelGGL(k_compact<false>, dim3((uint32_t)pair_p.size()), dim3(kBlock), 0, s, c->g, d_ps, d_st,
                               d_pair_p, d_pair_u, passes ? c->d_pflags + passes - 1 : (const uint32_t *)nullptr, d_more);
        HIPX(hipGetLastError());
        HIPX(hipMemcpyAsync(hst, d_st, (n + 1) * sizeof(PatState), hipMemcpyDeviceToHost, s));
        sync_spin(c);
        if (h_more[0] == 0) break;
    }
    HIPX(hipEventRecord(c->ev[1], s));

    // ---------------- decisions and plans ----------------
    const bool wave_phase = (opt.long_steps == 0 ? kDefaultLongSteps : opt.long_steps) != 0xffffffffu;
    std::vector<int> decided(n);
    size_t aux_words = 0, verdict_words = 0;
    for (uint32_t p = 0; p < n; ++p) {
        PatStatic &P = hs[p];
        uint32_t mn = 0xffffffffu;
        for (uint32_t u = 0; u < P.nv; ++u) {
            P.dom_size[u] = hst[p].dom_size[u];
            mn = std::min(mn, P.dom_size[u]);
        }
        if (distinct[p] && acyclic[p]) decided[p] = TFSM_DECIDED_EXPRESS;     // comper.h:1619-1653
        else if (mn < call.freq) decided[p] = TFSM_DECIDED_FILTER;            // isCandidateSetValid(freq)
        else decided[p] = TFSM_DECIDED_SEARCH;
        if (decided[p] != TFSM_DECIDED_SEARCH) continue;
        // the reference order's plans: read by its search and by the
        // wave-cooperative phase, and kept for the inspection of a
        // one-pattern call (tfsm_last_order); the adaptive order needs none
        if (!P.adaptive || n == 1 || wave_phase) make_plans(P, slot_of.data() + (size_t)p * kMaxV * kMaxV);
        P.task_prefix[0] = 0;
        for (uint32_t u = 0; u < P.nv; ++u) {
            P.task_prefix[u + 1] = P.task_prefix[u] + P.dom_size[u];
            verdict_words += 2 * (size_t)((P.dom_size[u] + 31) / 32);
        }
        for (uint32_t sl = 0; sl < P.nslots; ++sl) aux_words += 2 * (size_t)P.dom_size[P.slot_a[sl]];
    }
    uint32_t *aux_region = c->work.alloc<uint32_t>(std::max<size_t>(aux_words, 1));
    uint32_t *verdict_region = c->work.alloc<uint32_t>(std::max<size_t>(verdict_words, 1));
    HostWork aux_work;
    std::vector<unsigned long long> task_base{0};
    {
        size_t ao = 0, vo = 0;
        for (uint32_t p = 0; p < n; ++p) {
            PatStatic &P = hs[p];
            unsigned long long ntask = 0;
            if (decided[p] == TFSM_DECIDED_SEARCH) {
                for (uint32_t sl = 0; sl < P.nslots; ++sl) {
                    if (sl < 2u * kCutEdges) P.cut_ao[sl] = (uint32_t)ao;
                    P.aux_off[sl] = aux_region + ao;
                    ao += 2 * (size_t)P.dom_size[P.slot_a[sl]];
                    aux_work.add(p, sl, P.dom_size[P.slot_a[sl]], 64);   // one slot per wave
                }
                for (uint32_t u = 0; u < P.nv; ++u) {
                    const size_t wds = (P.dom_size[u] + 31) / 32;
                    P.valid_bits[u] = verdict_region + vo; vo += wds;
                    P.invalid_bits[u] = verdict_region + vo; vo += wds;
                }
                ntask = P.task_prefix[P.nv];
            }
            task_base.push_back(task_base.back() + ntask);
        }
    }
    // Fraction-Score sums: one item per 32 candidates of each vertex of a
    // search- or express-decided pattern
    HostWork frac_work;
    if (fraction)
        for (uint32_t p = 0; p < n; ++p)
            if (decided[p] == TFSM_DECIDED_SEARCH || decided[p] == TFSM_DECIDED_EXPRESS)
                for (uint32_t u = 0; u < hs[p].nv; ++u)
                    frac_work.add(p, u | (decided[p] == TFSM_DECIDED_EXPRESS ? 1u << 8 : 0u),
                                  (hs[p].dom_size[u] + 31) / 32);
    const size_t frac_bytes = fraction ? (size_t)n * kMaxV * sizeof(FracSum) : 0;
    // this phase's uploads (the descriptors with their plans, the table
    // work, the task bases) in one copy, the verdict bitmaps' fill in one
    // launch; the later kernels take the new descriptor copy
    PhaseBatch b2(c, n * sizeof(PatStatic) + aux_work.bytes() + 8 * task_base.size() + 3 * 256 +
                         (fraction ? frac_work.bytes() : 0) +
                         sizeof(CopyDesc) * ((verdict_words + frac_bytes / 4) / kFillPiece + 34) + 512);
    d_ps = b2.add(hs.data(), n);
    const WorkTable waux = aux_work.to_device(b2);
    const unsigned long long *d_task_base = b2.add(task_base);
    if (verdict_words) b2.zero(verdict_region, verdict_words * sizeof(uint32_t));
    FracSum *d_frac = nullptr;
    WorkTable wfrac{};
    if (fraction) {
        d_frac = c->work.alloc<FracSum>((size_t)n * kMaxV);
        b2.zero(d_frac, frac_bytes);
        wfrac = frac_work.to_device(b2);
    }
    // the learned order's per-vertex sums (and their probability-level
    // split) and the search counters start at zero
    // sums, then their probability-level split, then the sample pass's sums
    float *rank_sums = nullptr;
    if (learned && task_base.back()) {
        rank_sums = c->work.alloc<float>((size_t)n * kMaxV * (2 + kPrioQ));
        b2.zero(rank_sums, (size_t)n * kMaxV * (2 + kPrioQ) * sizeof(float));
    }
    if (task_base.back()) b2.zero(c->d_ctr, 2 * sizeof(SearchCounters));    // one per search launch
    // the piece counters of lane-grouped entries start at zero, one region per search launch
    const uint32_t group_max = opt.group_lanes == 0 ? kDefaultGroupLanes : opt.group_lanes;
    if (task_base.back() && group_max > 1 && !opt.trace) {
        const size_t gz = std::min<unsigned long long>(task_base.back(), kGroupCap);
        b2.zero(c->d_piece_cnt + kSplitCap, gz * sizeof(uint32_t));
        b2.zero(c->d_piece_cnt + kSplitCap + kGroupCap, gz * sizeof(uint32_t));
    }
    b2.flush();

    // ---------------- auxiliary tables ----------------
    if (aux_work.base.back()) {
        const uint32_t gaux = grid_for(waux.total, c->num_cus);
        // no host round trip: an overflow of the pool is found with the
        // final results (the search kernels do nothing on an overflowed
        // pool), and the call is then redone with a larger pool
        hipLaunchKernelGGL(k_aux_build, dim3(gaux), dim3(kBlock), 0, s, c->g, d_ps, d_st, waux, c->d_pool,
                           c->d_flags, c->pool_cap, c->d_flags + 1);
        HIPX(hipGetLastError());
    }
    HIPX(hipEventRecord(c->ev[2], s));

    // ---------------- existence search ----------------
    const unsigned long long total_tasks = task_base.back();
    uint32_t verify_launches = 0;
    c->trace_len = 0;
    if (opt.trace) {
        if (total_tasks > c->trace_cap) {
            if (c->d_trace) HIPX(hipFree(c->d_trace));
            c->d_trace = nullptr;
            c->trace_cap = std::max<unsigned long long>(total_tasks, 1 << 20);
            HIPX(hipMalloc(&c->d_trace, c->trace_cap * sizeof(uint32_t)));
        }
        c->trace_len = total_tasks;
        if (total_tasks) HIPX(hipMemsetAsync(c->d_trace, 0, total_tasks * sizeof(uint32_t), s));
    }
    if (total_tasks) {
        VerifyArgs va;
        va.ps = d_ps;
        va.st = d_st;
        va.task_base = d_task_base;
        va.P = n;
        va.epoch = 0;
        va.ctr = c->d_ctr;
        va.pool = c->d_pool;
        va.aux_region = aux_region;
        va.overflow = c->d_flags + 1;
        va.trace = opt.trace ? c->d_trace : nullptr;
        va.steps_per_round = opt.steps_per_round ? opt.steps_per_round : kDefaultStepsPerRound;
        va.refill_min = opt.refill_min ? opt.refill_min : kDefaultRefillMin;
        va.refill_max = opt.refill_max ? opt.refill_max : kDefaultRefillMax;
        va.split_steps = opt.split_steps == 0 ? kDefaultSplitSteps : opt.split_steps;
        if (opt.split_steps == 0xffffffffu) va.split_steps = 0;
        va.split_fanout = opt.split_fanout == 0 ? kDefaultSplitFanout : std::min(opt.split_fanout, 64u);
        va.records = c->d_records;
        va.split_cap = kSplitCap;
        va.piece_cnt = c->d_piece_cnt;
        va.piece_incomplete = c->d_piece_inc;
        va.order = nullptr;
        va.order_pu = nullptr;
        va.split = nullptr;
        va.hot = nullptr;
        va.part = 0;
        va.abort_stopped = 0;
        va.group_max = group_max;
        va.hot_group_max = opt.hot_group_lanes ? opt.hot_group_lanes : group_max;
        va.group_base = kSplitCap;
        va.group_cap = kGroupCap;
        if (total_tasks >= 0xffffffffull) throw ArgError{TFSM_ERR_LIMIT, "more than 2^32 tasks in one call"};
        // features (traced calls and the learned order), scores, priority sort
        const uint32_t gt = grid_for(total_tasks, c->num_cus);
        float *feat = nullptr;
        if (opt.trace) {
            if (total_tasks > c->feat_cap) {
                if (c->d_feat) HIPX(hipFree(c->d_feat));
                c->d_feat = nullptr;
                c->feat_cap = std::max<unsigned long long>(total_tasks, 1 << 18);
                HIPX(hipMalloc(&c->d_feat, c->feat_cap * kFeat * sizeof(float)));
            }
            c->feat_len = total_tasks;
            feat = c->d_feat;
        }
        if (opt.trace)
            hipLaunchKernelGGL(k_features, dim3(gt), dim3(kBlock), 0, s, c->g, d_ps, d_st, va.task_base, n,
                               total_tasks, feat);
        if (learned) {
            float *logits = c->work.alloc<float>(total_tasks);
            float *sums = rank_sums;
            float *qsums = sums + (size_t)n * kMaxV;
            float *ssums = qsums + (size_t)n * kMaxV * kPrioQ;
            uint32_t *keys = c->work.alloc<uint32_t>(total_tasks);
            // counting sort: pg blocks own contiguous task chunks (pg <= kPrioBlocks)
            const uint32_t pg = (uint32_t)std::min<unsigned long long>(kPrioBlocks, (total_tasks + kBlock - 1) / kBlock);
            uint32_t *hist = c->work.alloc<uint32_t>((size_t)kPrioBins * pg);
            uint32_t *order = c->work.alloc<uint32_t>(total_tasks);
            uint32_t *pu = c->work.alloc<uint32_t>(total_tasks);          // (pattern << 4) | vertex per task
            uint32_t *order_pu = c->work.alloc<uint32_t>(total_tasks);
            if (n >= (1u << 28)) throw ArgError{TFSM_ERR_LIMIT, "more than 2^28 patterns in one call"};
            // with no margin the cutoff is computed inside k_priority
            uint8_t *cutoff = kDeferMargin > 0.f ? c->work.alloc<uint8_t>((size_t)n * kMaxV) : nullptr;
            if (feat) {
                // traced: the feature rows are kept for training
                hipLaunchKernelGGL(k_mlp_forward, dim3(gt), dim3(kBlock), 0, s, feat, total_tasks, c->d_rank_w,
                                   logits);
                hipLaunchKernelGGL(k_group_sums, dim3(gt), dim3(kBlock), 0, s, d_ps, va.task_base, n, total_tasks,
                                   logits, sums, qsums);
            } else {
                // every kSample-th task first, then the tasks of the patterns it
                // does not expect frequent
                const uint32_t gs = grid_for((total_tasks + kSample - 1) / kSample, c->num_cus);
                hipLaunchKernelGGL(k_rank_scores, dim3(gs), dim3(kBlock), 0, s, c->g, d_ps, va.task_base, n,
                                   c->d_rank_w, logits, ssums, qsums, pu, kSample, (const float *)nullptr);
                hipLaunchKernelGGL(k_rank_scores, dim3(gt), dim3(kBlock), 0, s, c->g, d_ps, va.task_base, n,
                                   c->d_rank_w, logits, sums, qsums, pu, 1u, (const float *)ssums);
            }
            if (cutoff)
                hipLaunchKernelGGL(k_defer_cutoff, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, d_ps, n, sums,
                                   qsums, cutoff);
            hipLaunchKernelGGL(k_priority, dim3(pg), dim3(kBlock), 0, s, d_ps, va.task_base, n,
                               logits, sums, cutoff, keys, pu, hist, feat == nullptr,
                               feat == nullptr ? (const float *)ssums : (const float *)nullptr);
            const uint32_t nhist = kPrioBins * pg, ntiles = (nhist + kScanTile - 1) / kScanTile;
            static_assert((size_t)kPrioBins * kPrioBlocks <= (size_t)kScanTile * 4 * kBlock,
                          "the tile sums fit the one-block scan (4 per thread)");
            uint32_t *tiles = c->work.alloc<uint32_t>(ntiles);
            hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(kBlock), 0, s, hist, nhist, tiles);
            hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, tiles, ntiles);
            hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(kBlock), 0, s, hist, nhist, tiles);
            hipLaunchKernelGGL(k_prio_scatter, dim3(pg), dim3(kBlock), 0, s, keys, pu, total_tasks, hist, order,
                               order_pu);
            HIPX(hipGetLastError());
            va.order = order;
            va.order_pu = order_pu;
            va.abort_stopped = 1;
            // after the scan, the first deferred bin of block 0 holds the
            // number of first-part entries
            va.split = hist + (size_t)kDeferredKey * pg;
            // and the first key after the hot tier, the hot tier's size
            va.hot = hist + (size_t)kHotKeys * pg;
        }
        uint32_t epoch = ++c->epoch;
        HIPX(hipEventRecord(c->ev[4], s));
        uint32_t max_nv = 0;
        bool all_adaptive = true;       // the instance without the reference order's code
        va.max_edges = 0;
        for (uint32_t p = 0; p < n; ++p)
            if (decided[p] == TFSM_DECIDED_SEARCH) {
                max_nv = std::max(max_nv, hs[p].nv);
                all_adaptive = all_adaptive && hs[p].adaptive;
                va.max_edges = std::max(va.max_edges, hs[p].nslots / 2);
            }
        // persistent grid: at most the resident blocks per CU (registers and
        // LDS: verify_lds_bytes(8) = 60 KB per block for the 8-deep variant,
        // verify_lds_bytes(16) = 120 KB for the 16-deep one, of 160 KB per CU)
        const uint32_t resident =
            (uint32_t)(max_nv > 8 ? c->verify_blocks16 : all_adaptive ? c->verify_blocks8a : c->verify_blocks8);
        const uint32_t per_cu = opt.grid_waves_per_cu
                                    ? std::min(resident, std::max(1u, opt.grid_waves_per_cu / (kBlock / 64)))
                                    : resident;
        const uint32_t grid = (uint32_t)c->num_cus * per_cu;
        // root tasks still running after long_steps steps finish in the
        // wave-cooperative search (patterns of up to 8 vertices)
        va.long_cap = 0;
        va.long_list = nullptr;
        va.long_steps = opt.long_steps == 0 ? kDefaultLongSteps : opt.long_steps;
        if (va.long_steps != 0xffffffffu && max_nv <= 8) {
            va.long_cap = (uint32_t)std::min<unsigned long long>(total_tasks, kLongCap);
            va.long_list = c->work.alloc<uint32_t>(va.long_cap);
        }
        va.long_donate_every = opt.long_donate_every == 0 ? kDefaultLongDonateEvery : opt.long_donate_every;
        if (opt.long_donate_every == 0xffffffffu) va.long_donate_every = 0;
        va.long_fanout = opt.long_fanout == 0 ? kDefaultLongFanout : std::min(opt.long_fanout, 64u);
        // the second phase's blocks (4 waves each) all fit at once (20 waves per CU)
        const uint32_t long_blocks_per_cu =
            std::min(5u, std::max(1u, (opt.long_waves_per_cu ? opt.long_waves_per_cu : kDefaultLongWavesPerCu) / 4));
        // the learned order runs in two launches: the tasks the ranker expects
        // to decide the patterns, then the deferred rest (mostly skipped)
        const bool two_parts = va.order && va.split && !va.long_cap;
        auto launch_verify = [&]() {
            if (max_nv > 8)
                hipLaunchKernelGGL((k_verify<16, tfsm_lane::kOrderMixed, 12>), dim3(grid), dim3(kBlock), 0, s, va,
                                   epoch);
            else if (all_adaptive && va.max_edges <= 8)
                hipLaunchKernelGGL((k_verify<8, tfsm_lane::kOrderAdaptive, 8>), dim3(grid), dim3(kBlock), 0, s, va,
                                   epoch);
            else if (all_adaptive)
                hipLaunchKernelGGL((k_verify<8, tfsm_lane::kOrderAdaptive, 12>), dim3(grid), dim3(kBlock), 0, s, va,
                                   epoch);
            else
                hipLaunchKernelGGL((k_verify<8, tfsm_lane::kOrderMixed, 12>), dim3(grid), dim3(kBlock), 0, s, va,
                                   epoch);
            HIPX(hipGetLastError());
        };
        va.part = two_parts ? 1u : 0u;
        launch_verify();
        if (two_parts) {
            // the deferred launch keeps its own counters, zeroed with the first's
            epoch = ++c->epoch;
            va.ctr = c->d_ctr + 1;
            va.part = 2;
            va.group_base = kSplitCap + kGroupCap;
            launch_verify();
        }
        verify_launches = two_parts ? 2u : 1u;
        if (va.long_cap) {
            hipLaunchKernelGGL(k_verify_long<8>, dim3((uint32_t)c->num_cus * long_blocks_per_cu), dim3(kBlock), 0,
                               s, va, epoch);
            HIPX(hipGetLastError());
        }
        HIPX(hipEventRecord(c->ev[5], s));
        if (opt.trace)
            hipLaunchKernelGGL(k_trace_final, dim3(gt), dim3(kBlock), 0, s, d_ps, va.task_base, n, total_tasks,
                               c->d_trace);
    }
    if (fraction && wfrac.total) {
        hipLaunchKernelGGL(k_frac_sum, dim3(grid_for(wfrac.total, c->num_cus)), dim3(kBlock), 0, s, d_ps, wfrac,
                           (const double *)c->d_score, c->L, d_frac);
        HIPX(hipGetLastError());
    }
    HIPX(hipEventRecord(c->ev[3], s));
    HIPX(hipMemcpyAsync(hst, d_st, n * sizeof(PatState), hipMemcpyDeviceToHost, s));
    FracSum *hfrac = nullptr;
    if (fraction) {
        hfrac = c->h_frac.ensure((size_t)n * kMaxV);
        HIPX(hipMemcpyAsync(hfrac, d_frac, frac_bytes, hipMemcpyDeviceToHost, s));
    }
    // the two launches' counters and the pool words follow each other: one copy
    SearchCounters *hctr = c->h_ctr.ensure(3);
    HIPX(hipMemcpyAsync(hctr, c->d_ctr, 2 * sizeof(SearchCounters) + 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    const uint32_t *hpool = reinterpret_cast<const uint32_t *>(hctr + 2);   // pool used, pool overflow, passes
    sync_spin(c);
    if (hpool[1]) {
        // the tables did not fit: grow the pool to twice the measured need and redo the call
        if (attempt >= 4) throw ArgError{TFSM_ERR_LIMIT, "table pool overflow"};
        const unsigned long long need = (unsigned long long)hpool[0] * 2;
        if (need > 0xfffffff0ull) throw ArgError{TFSM_ERR_LIMIT, "tables exceed 2^32 entries"};
        HIPX(hipFree(c->d_pool));
        c->d_pool = nullptr;
        c->pool_cap = (uint32_t)std::max<unsigned long long>(need, c->pool_cap);
        HIPX(hipMalloc(&c->d_pool, (size_t)c->pool_cap * sizeof(uint32_t)));
        return run_eval(c, call, attempt + 1);
    }
    return finish_call(c, call, Finished{hs, hst, hctr, hpool, decided, hfrac, total_tasks, verify_launches,
                                         TFSM_LEVEL_HOST, true, nullptr});
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

const char *tfsm_status_string(int status) {
    switch (status) {
        case TFSM_OK: return "ok";
        case TFSM_ERR_ARGUMENT: return "invalid argument";
        case TFSM_ERR_HIP: return "HIP runtime error";
        case TFSM_ERR_NO_DEVICE: return "no usable GPU";
        case TFSM_ERR_LIMIT: return "input beyond a supported limit";
        case TFSM_ERR_HANDLE: return "unknown domain handle";
        default: return "unknown status";
    }
}

const char *tfsm_last_error(tfsm_ctx *ctx) { return ctx ? ctx->err.c_str() : ""; }

int tfsm_ctx_create(const tfsm_csr *g, tfsm_ctx **out) {
    if (!g || !out) return TFSM_ERR_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return TFSM_ERR_NO_DEVICE;
    if (g->num_vertices == 0 || g->num_vlabels == 0 || g->num_elabels == 0) return TFSM_ERR_ARGUMENT;
    if (!g->offsets || (g->num_adj && (!g->nbrs || !g->elabels)) || !g->vlabels) return TFSM_ERR_ARGUMENT;
    if (g->num_adj >= 0xffffffffull) return TFSM_ERR_LIMIT;
    const uint64_t K = (uint64_t)g->num_vlabels * g->num_elabels;
    if (K > 4096) return TFSM_ERR_LIMIT;
    auto *c = new tfsm_ctx;
    int rc = guarded(c, [&]() -> int {
        HIPX(hipGetDevice(&c->device));
        hipDeviceProp_t prop;
        HIPX(hipGetDeviceProperties(&prop, c->device));
        c->num_cus = prop.multiProcessorCount;
        // the persistent search grid is sized to what stays resident
        // (registers and LDS of each instance)
        HIPX(hipOccupancyMaxActiveBlocksPerMultiprocessor(&c->verify_blocks8, k_verify<8, tfsm_lane::kOrderMixed, 12>,
                                                          kBlock, 0));
        int b8 = 1;
        HIPX(hipOccupancyMaxActiveBlocksPerMultiprocessor(&c->verify_blocks8a,
                                                          k_verify<8, tfsm_lane::kOrderAdaptive, 12>, kBlock, 0));
        HIPX(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b8, k_verify<8, tfsm_lane::kOrderAdaptive, 8>, kBlock, 0));
        c->verify_blocks8a = std::min(c->verify_blocks8a, b8);
        HIPX(hipOccupancyMaxActiveBlocksPerMultiprocessor(&c->verify_blocks16, k_verify<16, tfsm_lane::kOrderMixed, 12>,
                                                          kBlock, 0));
        c->verify_blocks8 = std::max(1, c->verify_blocks8);
        c->verify_blocks8a = std::max(1, c->verify_blocks8a);
        c->verify_blocks16 = std::max(1, c->verify_blocks16);
        HIPX(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (auto &e : c->ev) HIPX(hipEventCreate(&e));
        HIPX(hipEventCreateWithFlags(&c->ev_sync, hipEventDisableTiming));
        c->n = g->num_vertices;
        c->L = g->num_vlabels;
        c->EL = g->num_elabels;
        c->K = (uint32_t)K;
        c->W = (c->n + 31) / 32;
        c->nadj = g->num_adj;
        // validate and partition the adjacency by (label(w), edge label), keeping id order
        if (g->offsets[0] != 0 || g->offsets[c->n] != g->num_adj) throw ArgError{TFSM_ERR_ARGUMENT, "bad offsets"};
        for (uint32_t v = 0; v < c->n; ++v) {
            if (g->offsets[v + 1] < g->offsets[v]) throw ArgError{TFSM_ERR_ARGUMENT, "offsets not monotone"};
            if (g->vlabels[v] >= c->L) throw ArgError{TFSM_ERR_ARGUMENT, "vertex label out of range"};
        }
        std::vector<uint32_t> seg((size_t)c->n * (c->K + 1));
        std::vector<uint32_t> nbr(g->num_adj);
        std::vector<uint32_t> cnt(c->K + 1);
        std::vector<unsigned long long> keymask(c->n, 0);
        if ((size_t)c->L * c->K <= kFanMaxCells) {
            c->fan_n.assign((size_t)c->L * c->K * kFanBins, 0);
            c->fan_sum.assign((size_t)c->L * c->K * kFanBins, 0);
        }
        for (uint32_t v = 0; v < c->n; ++v) {
            const uint32_t b = g->offsets[v], e = g->offsets[v + 1];
            std::fill(cnt.begin(), cnt.end(), 0);
            for (uint32_t k = b; k < e; ++k) {
                const uint32_t w = g->nbrs[k];
                if (w >= c->n || g->elabels[k] >= c->EL) throw ArgError{TFSM_ERR_ARGUMENT, "adjacency entry out of range"};
                if (k > b && g->nbrs[k - 1] > w) throw ArgError{TFSM_ERR_ARGUMENT, "adjacency not ascending"};
                cnt[g->vlabels[w] * c->EL + g->elabels[k] + 1]++;
            }
            uint32_t *sv = &seg[(size_t)v * (c->K + 1)];
            uint32_t acc = b;
            for (uint32_t k = 0; k < c->K; ++k) {
                sv[k] = acc;
                acc += cnt[k + 1];
                if (!cnt[k + 1]) continue;
                keymask[v] |= 1ull << (k & 63);
                if (!c->fan_n.empty()) {
                    const size_t cell = ((size_t)g->vlabels[v] * c->K + k) * kFanBins + (31 - __builtin_clz(cnt[k + 1]));
                    c->fan_n[cell]++;
                    c->fan_sum[cell] += cnt[k + 1];
                }
            }
            sv[c->K] = acc;
            std::vector<uint32_t> fill(sv, sv + c->K);
            for (uint32_t k = b; k < e; ++k) {
                const uint32_t w = g->nbrs[k];
                nbr[fill[g->vlabels[w] * c->EL + g->elabels[k]]++] = w;
            }
        }
        std::vector<uint32_t> bucket(c->n);
        c->bucket_off.assign(c->L + 1, 0);
        for (uint32_t v = 0; v < c->n; ++v) c->bucket_off[g->vlabels[v] + 1]++;
        for (uint32_t l = 0; l < c->L; ++l) c->bucket_off[l + 1] += c->bucket_off[l];
        {
            std::vector<uint32_t> f(c->bucket_off.begin(), c->bucket_off.end() - 1);
            for (uint32_t v = 0; v < c->n; ++v) bucket[f[g->vlabels[v]]++] = v;
        }
        HIPX(hipMalloc(&c->d_seg, seg.size() * sizeof(uint32_t)));
        HIPX(hipMalloc(&c->d_nbr, std::max<size_t>(nbr.size(), 1) * sizeof(uint32_t)));
        HIPX(hipMalloc(&c->d_bucket, bucket.size() * sizeof(uint32_t)));
        HIPX(hipMalloc(&c->d_keymask, std::max<size_t>(keymask.size(), 1) * sizeof(unsigned long long)));
        HIPX(hipMalloc(&c->d_keymask_bucket, std::max<size_t>(keymask.size(), 1) * sizeof(unsigned long long)));
        if (!keymask.empty()) {
            HIPX(hipMemcpy(c->d_keymask, keymask.data(), keymask.size() * sizeof(unsigned long long),
                           hipMemcpyHostToDevice));
            std::vector<unsigned long long> kb(c->n);
            for (uint32_t i = 0; i < c->n; ++i) kb[i] = keymask[bucket[i]];
            HIPX(hipMemcpy(c->d_keymask_bucket, kb.data(), kb.size() * sizeof(unsigned long long),
                           hipMemcpyHostToDevice));
        }
        HIPX(hipMemcpy(c->d_seg, seg.data(), seg.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        if (!nbr.empty()) HIPX(hipMemcpy(c->d_nbr, nbr.data(), nbr.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIPX(hipMemcpy(c->d_bucket, bucket.data(), bucket.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        c->g.n = c->n;
        c->g.W = c->W;
        c->g.K = c->K;
        c->g.EL = c->EL;
        c->g.seg = c->d_seg;
        c->g.nbr = c->d_nbr;
        c->g.keymask = c->d_keymask;
        c->pool_cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(8 * g->num_adj, 1u << 22), 0xfffffff0ull);
        HIPX(hipMalloc(&c->d_pool, (size_t)c->pool_cap * sizeof(uint32_t)));
        HIPX(hipMalloc(&c->d_pflags, kMaxPrunePasses * sizeof(uint32_t)));
        // the counters of the two search launches, then the pool words
        HIPX(hipMalloc(&c->d_ctr, 3 * sizeof(SearchCounters)));
        c->d_flags = reinterpret_cast<uint32_t *>(c->d_ctr + 2);
        HIPX(hipMalloc(&c->d_records, (size_t)kSplitCap * sizeof(SplitRecord)));
        // piece counters: kSplitCap for records, then two regions of
        // kGroupCap for lane-grouped root tasks
        const size_t npieces = (size_t)kSplitCap + 2 * (size_t)kGroupCap;
        HIPX(hipMalloc(&c->d_piece_cnt, npieces * sizeof(uint32_t)));
        HIPX(hipMalloc(&c->d_piece_inc, npieces * sizeof(uint32_t)));
        HIPX(hipMemset(c->d_records, 0, (size_t)kSplitCap * sizeof(SplitRecord)));
        HIPX(hipMemset(c->d_piece_cnt, 0, npieces * sizeof(uint32_t)));
        HIPX(hipMemset(c->d_piece_inc, 0, npieces * sizeof(uint32_t)));
        HIPX(hipMalloc(&c->d_rank_w, kParams * sizeof(uint16_t)));
        // the no-allocation gradient's scratch: this device's copy of the array
        HIPX(hipGetSymbolAddress(reinterpret_cast<void **>(&c->d_static_grad), HIP_SYMBOL(scratch)));
        // the descriptor of device-driven level calls: one slot of this device's c_level
        HIPX(hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, c->device));
        if (c->wall_khz <= 0) c->wall_khz = 100000;
        c->level_slot = acquire_slot(c->device);
        if (c->level_slot != kNoSlot) {
            LevelArgs *base = nullptr;
            HIPX(hipGetSymbolAddress(reinterpret_cast<void **>(&base), HIP_SYMBOL(c_level)));
            c->d_level = base + c->level_slot;
        }
        return TFSM_OK;
    });
    if (rc != TFSM_OK) {
        tfsm_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return TFSM_OK;
}

int tfsm_ctx_set_pool_capacity(tfsm_ctx *c, uint64_t entries) {
    if (!c || entries == 0) return TFSM_ERR_ARGUMENT;
    if (entries > 0xfffffff0ull) return TFSM_ERR_LIMIT;
    return guarded(c, [&]() -> int {
        HIPX(hipStreamSynchronize(c->stream));
        HIPX(hipFree(c->d_pool));
        c->d_pool = nullptr;
        c->pool_cap = (uint32_t)entries;
        HIPX(hipMalloc(&c->d_pool, (size_t)c->pool_cap * sizeof(uint32_t)));
        return TFSM_OK;
    });
}

int tfsm_ctx_destroy(tfsm_ctx *c) {
    if (!c) return TFSM_ERR_ARGUMENT;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &kv : c->retained) c->cache.put(kv.second.buf, kv.second.cls);
    c->retained.clear();
    (void)hipFree(c->d_seg);
    (void)hipFree(c->d_nbr);
    (void)hipFree(c->d_bucket);
    (void)hipFree(c->d_keymask);
    (void)hipFree(c->d_keymask_bucket);
    (void)hipFree(c->d_score);
    (void)hipFree(c->d_pool);
    (void)hipFree(c->d_ctr);
    (void)hipFree(c->d_records);
    (void)hipFree(c->d_piece_cnt);
    (void)hipFree(c->d_piece_inc);
    (void)hipFree(c->d_pflags);
    (void)hipFree(c->d_rank_w);
    (void)hipFree(c->d_feat);
    (void)hipFree(c->d_trace);
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
    c->graphs.clear();
    release_slot(c->device, c->level_slot);
    for (auto &e : c->ev) if (e) (void)hipEventDestroy(e);
    if (c->ev_sync) (void)hipEventDestroy(c->ev_sync);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TFSM_OK;
}

int tfsm_eval_pattern(tfsm_ctx *ctx, const tfsm_pattern *p, uint64_t parent_handle, uint32_t freq,
                      const tfsm_eval_options *opt, tfsm_pattern_result *out) {
    if (!ctx || !p || !opt || !out) return TFSM_ERR_ARGUMENT;
    return guarded(ctx, [&]() { return run_eval(ctx, EvalCall{p, &parent_handle, 1, freq, *opt, out}); }, true);
}

// A level whose candidate bitmaps (two words per 32 data vertices per
// pattern vertex) pass kLevelBitmapBudget bytes is evaluated in consecutive
// slices of patterns, each one fused launch sequence: on a 1M-vertex graph a
// level of 30K three-vertex candidates would otherwise need 25 GB of them.
constexpr size_t kLevelBitmapBudget = size_t(1) << 30;

int tfsm_eval_level(tfsm_ctx *ctx, const tfsm_pattern *ps, const uint64_t *parent_handles, uint32_t n, uint32_t freq,
                    const tfsm_eval_options *opt, tfsm_pattern_result *outs) {
    if (!ctx || !opt || (n && (!ps || !outs))) return TFSM_ERR_ARGUMENT;
    if (n == 0) return TFSM_OK;
    return guarded(ctx, [&]() -> int {
        for (uint32_t i = 0; i < n; ++i) check_pattern(ctx, ps[i]);      // no slice runs on bad input
        const size_t per_vertex = 2 * (size_t)ctx->W * sizeof(uint32_t);
        tfsm_call_stats sum{};
        uint32_t slices = 0;
        for (uint32_t i = 0; i < n;) {
            size_t bytes = (size_t)ps[i].num_vertices * per_vertex;
            uint32_t j = i + 1;
            while (j < n && bytes + (size_t)ps[j].num_vertices * per_vertex <= kLevelBitmapBudget)
                bytes += (size_t)ps[j++].num_vertices * per_vertex;
            if (i == 0 && j < n && opt->trace)
                throw ArgError{TFSM_ERR_LIMIT, "a traced level must fit one slice (split the level)"};
            // the slice's push-down entries, renumbered to the slice's patterns
            tfsm_eval_options so = *opt;
            std::vector<tfsm_pushdown> sp;
            if (i != 0 || j != n) {
                if (opt->num_pushdown && !opt->pushdown)
                    throw ArgError{TFSM_ERR_ARGUMENT, "num_pushdown without a pushdown array"};
                for (uint32_t k = 0; k < opt->num_pushdown; ++k)
                    if (opt->pushdown[k].pattern >= i && opt->pushdown[k].pattern < j) {
                        sp.push_back(opt->pushdown[k]);
                        sp.back().pattern -= i;
                    }
                so.num_pushdown = (uint32_t)sp.size();
                so.pushdown = sp.data();
            }
            EvalCall ec{ps + i, parent_handles ? parent_handles + i : nullptr, j - i, freq, so, outs + i};
            ec.level_call = true;
            const int rc = run_eval(ctx, ec);
            if (rc != TFSM_OK) return rc;
            const tfsm_call_stats &s = ctx->stats;
            sum.filter_ms += s.filter_ms;
            sum.table_ms += s.table_ms;
            sum.search_ms += s.search_ms;
            sum.rank_ms += s.rank_ms;
            sum.search_launches += s.search_launches;
            sum.table_bytes += s.table_bytes;
            sum.search_bytes += s.search_bytes;
            sum.search_splits += s.search_splits;
            sum.search_watchdog |= s.search_watchdog;
            sum.filter_passes += s.filter_passes;
            sum.level_mode = s.level_mode;
            ++slices;
            i = j;
        }
        if (slices > 1) ctx->stats = sum;
        return TFSM_OK;
    }, true);
}

int64_t tfsm_fraction_scores(tfsm_ctx *ctx, double *out, uint64_t cap) {
    if (!ctx) return TFSM_ERR_ARGUMENT;
    int64_t cells = 0;
    const int rc = guarded(ctx, [&]() -> int {
        HIPX(hipStreamSynchronize(ctx->stream));
        ensure_scores(ctx);
        cells = (int64_t)ctx->n * ctx->L;
        const uint64_t k = std::min<uint64_t>(cap, (uint64_t)cells);
        if (out && k) {
            HIPX(hipMemcpyAsync(out, ctx->d_score, k * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            HIPX(hipStreamSynchronize(ctx->stream));
        }
        return TFSM_OK;
    });
    return rc != TFSM_OK ? rc : cells;
}

int tfsm_domain_read(tfsm_ctx *ctx, uint64_t handle, uint32_t u, uint32_t *out, uint32_t cap, uint32_t *size) {
    if (!ctx) return TFSM_ERR_ARGUMENT;
    auto it = ctx->retained.find(handle);
    if (it == ctx->retained.end()) return fail(ctx, TFSM_ERR_HANDLE, "unknown domain handle");
    const Retained &R = it->second;
    if (u >= R.nv) return fail(ctx, TFSM_ERR_ARGUMENT, "pattern vertex out of range");
    return guarded(ctx, [&]() -> int {
               // an eval call returns with its retained-domain copies still queued
               HIPX(hipStreamSynchronize(ctx->stream));
        const uint32_t n = R.size[u];
        std::vector<uint32_t> dom(n), inv((n + 31) / 32);
        if (n) {
            HIPX(hipMemcpy(dom.data(), R.dom[u], n * sizeof(uint32_t), hipMemcpyDeviceToHost));
            HIPX(hipMemcpy(inv.data(), R.invalid[u], inv.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        }
        uint32_t k = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (inv[i >> 5] >> (i & 31) & 1) continue;
            if (out && k < cap) out[k] = dom[i];
            ++k;
        }
        if (size) *size = k;
        return TFSM_OK;
    });
}

int tfsm_domain_release(tfsm_ctx *ctx, uint64_t handle) {
    if (!ctx) return TFSM_ERR_ARGUMENT;
    auto it = ctx->retained.find(handle);
    if (it == ctx->retained.end()) return fail(ctx, TFSM_ERR_HANDLE, "unknown domain handle");
    ctx->cache.put(it->second.buf, it->second.cls);
    ctx->retained.erase(it);
    return TFSM_OK;
}

int64_t tfsm_last_domain(tfsm_ctx *ctx, uint32_t u, uint32_t *out, uint32_t cap) {
    if (!ctx || ctx->last_static.size() != 1 || u >= ctx->last_static[0].nv) return TFSM_ERR_ARGUMENT;
    const PatStatic &P = ctx->last_static[0];
    const uint32_t n = P.dom_size[u];
    int rc = guarded(ctx, [&]() -> int {
                 // an eval call returns with its retained-domain copies still queued
                 HIPX(hipStreamSynchronize(ctx->stream));
        if (out && n) HIPX(hipMemcpy(out, P.dom[u], std::min(n, cap) * sizeof(uint32_t), hipMemcpyDeviceToHost));
        return TFSM_OK;
    });
    return rc == TFSM_OK ? (int64_t)n : rc;
}

int64_t tfsm_last_states(tfsm_ctx *ctx, uint32_t u, uint8_t *out, uint32_t cap) {
    if (!ctx || ctx->last_static.size() != 1 || u >= ctx->last_static[0].nv) return TFSM_ERR_ARGUMENT;
    const PatStatic &P = ctx->last_static[0];
    const int decided = ctx->last_decided[0];
    const uint32_t n = P.dom_size[u];
    int rc = guarded(ctx, [&]() -> int {
                 // an eval call returns with its retained-domain copies still queued
                 HIPX(hipStreamSynchronize(ctx->stream));
        std::vector<uint32_t> vb((n + 31) / 32, 0), ib((n + 31) / 32, 0);
        if (decided == TFSM_DECIDED_SEARCH && n) {
            HIPX(hipMemcpy(vb.data(), P.valid_bits[u], vb.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
            HIPX(hipMemcpy(ib.data(), P.invalid_bits[u], ib.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        }
        for (uint32_t i = 0; i < n && i < cap && out; ++i) {
            if (decided == TFSM_DECIDED_EXPRESS) out[i] = 1;
            else out[i] = (vb[i >> 5] >> (i & 31) & 1) ? 1 : ((ib[i >> 5] >> (i & 31) & 1) ? 2 : 0);
        }
        return TFSM_OK;
    });
    return rc == TFSM_OK ? (int64_t)n : rc;
}

int64_t tfsm_last_aux(tfsm_ctx *ctx, uint32_t a, uint32_t b, uint32_t *off, uint32_t off_cap, uint32_t *idx,
                      uint32_t idx_cap) {
    if (!ctx || ctx->last_static.size() != 1) return TFSM_ERR_ARGUMENT;
    const PatStatic &P = ctx->last_static[0];
    if (a >= P.nv || b >= P.nv || ctx->last_decided[0] != TFSM_DECIDED_SEARCH) return -1;
    int slot = -1;
    for (uint32_t s = 0; s < P.nslots; ++s)
        if (P.slot_a[s] == a && P.slot_b[s] == b) slot = (int)s;
    if (slot < 0) return -1;
    const uint32_t rows = P.dom_size[a];
    const uint32_t total = ctx->last_state[0].aux_total[slot];
    int rc = guarded(ctx, [&]() -> int {
                 // an eval call returns with its retained-domain copies still queued
                 HIPX(hipStreamSynchronize(ctx->stream));
        // the device keeps each row as a (start, end) pair of pool positions;
        // callers get the table as a plain CSR
        std::vector<uint32_t> se(2 * (size_t)rows), pool(idx ? ctx->last_pool_used : 0u);
        if (rows) HIPX(hipMemcpy(se.data(), P.aux_off[slot], se.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (idx && total && !pool.empty())
            HIPX(hipMemcpy(pool.data(), ctx->d_pool, pool.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        uint32_t acc = 0;
        for (uint32_t i = 0; i <= rows; ++i) {
            if (off && i < off_cap) off[i] = acc;
            if (i == rows) break;
            const uint32_t lo = se[2 * i], hi = std::max(lo, se[2 * i + 1]);
            if (!idx) {
                acc += hi - lo;
                continue;
            }
            for (uint32_t q = lo; q < hi; ++q, ++acc)
                if (acc < idx_cap) idx[acc] = q < pool.size() ? pool[q] : 0xffffffffu;
        }
        return TFSM_OK;
    });
    return rc == TFSM_OK ? (int64_t)total : rc;
}

int tfsm_last_order(tfsm_ctx *ctx, uint32_t u, uint32_t *order) {
    if (!ctx || !order || ctx->last_static.size() != 1 || u >= ctx->last_static[0].nv) return TFSM_ERR_ARGUMENT;
    const PatStatic &P = ctx->last_static[0];
    for (uint32_t d = 0; d < P.nv; ++d) order[d] = P.order[u][d];
    return TFSM_OK;
}

int64_t tfsm_read_trace(tfsm_ctx *ctx, uint32_t *out, uint64_t cap) {
    if (!ctx) return TFSM_ERR_ARGUMENT;
    const uint64_t n = ctx->trace_len;
    int rc = guarded(ctx, [&]() -> int {
                 // an eval call returns with its retained-domain copies still queued
                 HIPX(hipStreamSynchronize(ctx->stream));
        if (out && n) HIPX(hipMemcpy(out, ctx->d_trace, std::min(n, cap) * sizeof(uint32_t), hipMemcpyDeviceToHost));
        return TFSM_OK;
    });
    return rc == TFSM_OK ? (int64_t)n : rc;
}

int tfsm_last_stats(tfsm_ctx *ctx, tfsm_call_stats *out) {
    if (!ctx || !out) return TFSM_ERR_ARGUMENT;
    *out = ctx->stats;
    return TFSM_OK;
}

// ---------------------------------------------------------------------------
// Ranker
// ---------------------------------------------------------------------------
static uint16_t host_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t rounding = ((u >> 16) & 1u) + 0x7fffu;
    return (uint16_t)((u + rounding) >> 16);
}

static float host_f32(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

static_assert(sizeof(tfsm_ranker_weights) == kParams * sizeof(float), "ranker layout");

int tfsm_ranker_set(tfsm_ctx *ctx, const tfsm_ranker_weights *w) {
    if (!ctx || !w) return TFSM_ERR_ARGUMENT;
    return guarded(ctx, [&]() -> int {
        const float *f = reinterpret_cast<const float *>(w);
        std::vector<uint16_t> h(kParams);
        for (int i = 0; i < kParams; ++i) {
            if (!std::isfinite(f[i])) throw ArgError{TFSM_ERR_ARGUMENT, "ranker weight is not finite"};
            h[i] = host_bf16(f[i]);
        }
        // queued work of an earlier call may still read the installed weights
        HIPX(hipMemcpyAsync(ctx->d_rank_w, h.data(), kParams * sizeof(uint16_t), hipMemcpyHostToDevice, ctx->stream));
        HIPX(hipStreamSynchronize(ctx->stream));
        ctx->ranker_installed = true;
        return TFSM_OK;
    });
}

int tfsm_ranker_get(tfsm_ctx *ctx, tfsm_ranker_weights *w) {
    if (!ctx || !w) return TFSM_ERR_ARGUMENT;
    if (!ctx->ranker_installed) return fail(ctx, TFSM_ERR_ARGUMENT, "no ranker installed");
    return guarded(ctx, [&]() -> int {
        std::vector<uint16_t> h(kParams);
        HIPX(hipMemcpyAsync(h.data(), ctx->d_rank_w, kParams * sizeof(uint16_t), hipMemcpyDeviceToHost, ctx->stream));
        HIPX(hipStreamSynchronize(ctx->stream));
        float *f = reinterpret_cast<float *>(w);
        for (int i = 0; i < kParams; ++i) f[i] = host_f32(h[i]);
        return TFSM_OK;
    });
}

namespace {

// Device copies of n feature rows and labels in the scratch arena.  An eval
// call returns with work still queued on the context's stream that reads the
// arena (the retained-domain copies and their descriptor list), so the
// stream is drained before the arena is reused, and the copies go on that
// stream.
void stage_rows(tfsm_ctx *c, const float *x, const float *y, uint64_t n, float **dx, float **dy) {
    sync_spin(c);
    c->work.reset();
    *dx = c->work.alloc<float>(n * kFeat);
    HIPX(hipMemcpyAsync(*dx, x, n * kFeat * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (y) {
        *dy = c->work.alloc<float>(n);
        HIPX(hipMemcpyAsync(*dy, y, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    // the host buffers are caller-owned and pageable: the copies must finish
    // before the call returns
    HIPX(hipStreamSynchronize(c->stream));
}

}  // namespace

int tfsm_ranker_forward(tfsm_ctx *ctx, const float *x, uint64_t n, float *logits) {
    if (!ctx || (n && (!x || !logits))) return TFSM_ERR_ARGUMENT;
    if (!ctx->ranker_installed) return fail(ctx, TFSM_ERR_ARGUMENT, "no ranker installed");
    if (!n) return TFSM_OK;
    return guarded(ctx, [&]() -> int {
        float *dx, *dy = nullptr;
        stage_rows(ctx, x, nullptr, n, &dx, &dy);
        float *dl = ctx->work.alloc<float>(n);
        hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for(n, ctx->num_cus)), dim3(kBlock), 0, ctx->stream, dx, n,
                           ctx->d_rank_w, dl);
        HIPX(hipGetLastError());
        HIPX(hipMemcpyAsync(logits, dl, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        HIPX(hipStreamSynchronize(ctx->stream));
        return TFSM_OK;
    });
}

int tfsm_ranker_gradient_noalloc(tfsm_ctx *ctx, const float *d_x, const float *d_y, uint64_t n, float *d_grad,
                                 float *d_loss) {
    if (!ctx || !d_x || !d_y || !d_grad || !n) return TFSM_ERR_ARGUMENT;
    if (!ctx->ranker_installed) return fail(ctx, TFSM_ERR_ARGUMENT, "no ranker installed");
    if (n > TFSM_NOALLOC_MAX_ROWS)
        return fail(ctx, TFSM_ERR_LIMIT, "tfsm_ranker_gradient_noalloc takes at most TFSM_NOALLOC_MAX_ROWS rows");
    if (!ctx->d_static_grad) return fail(ctx, TFSM_ERR_HIP, "no scratch address for this device");
    return guarded(ctx, [&]() -> int {
        // the sums start at zero (a fill kernel: no memset node if a caller captures the stream)
        float *sums = ctx->d_static_grad;
        hipLaunchKernelGGL(k_fill_f32, dim3((kStaticGradWords + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream,
                           sums, kStaticGradWords, 0.f);
        hipLaunchKernelGGL(k_mlp_backward, dim3(grid_for(n, ctx->num_cus)), dim3(kBlock), 0, ctx->stream, d_x, d_y,
                           (const uint32_t *)nullptr, n, ctx->d_rank_w, sums, sums + kParams);
        hipLaunchKernelGGL(k_scale_grad, dim3((kParams + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, sums,
                           1.f / (float)n, d_grad, d_loss);
        HIPX(hipGetLastError());
        HIPX(hipStreamSynchronize(ctx->stream));
        return TFSM_OK;
    });
}

int tfsm_ranker_gradient(tfsm_ctx *ctx, const float *x, const float *y, uint64_t n, tfsm_ranker_weights *grad,
                         float *loss) {
    if (!ctx || !grad || !n || !x || !y) return TFSM_ERR_ARGUMENT;
    if (!ctx->ranker_installed) return fail(ctx, TFSM_ERR_ARGUMENT, "no ranker installed");
    return guarded(ctx, [&]() -> int {
        float *dx, *dy;
        stage_rows(ctx, x, y, n, &dx, &dy);
        float *dg = ctx->work.alloc<float>(kParams + 1);
        HIPX(hipMemsetAsync(dg, 0, (kParams + 1) * sizeof(float), ctx->stream));
        hipLaunchKernelGGL(k_mlp_backward, dim3(grid_for(n, ctx->num_cus)), dim3(kBlock), 0, ctx->stream, dx, dy,
                           (const uint32_t *)nullptr, n, ctx->d_rank_w, dg, dg + kParams);
        HIPX(hipGetLastError());
        std::vector<float> h(kParams + 1);
        HIPX(hipMemcpyAsync(h.data(), dg, (kParams + 1) * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        HIPX(hipStreamSynchronize(ctx->stream));
        float *g = reinterpret_cast<float *>(grad);
        for (int i = 0; i < kParams; ++i) g[i] = h[i] / (float)n;
        if (loss) *loss = h[kParams] / (float)n;
        return TFSM_OK;
    });
}

int tfsm_ranker_train(tfsm_ctx *ctx, const float *x, const float *y, uint64_t n, uint32_t epochs, uint32_t batch,
                      float lr, uint64_t seed, float *loss) {
    if (!ctx || !x || !y || !n || !epochs || !batch || !(lr > 0.f)) return TFSM_ERR_ARGUMENT;
    if (!ctx->ranker_installed) return fail(ctx, TFSM_ERR_ARGUMENT, "install initial weights first");
    if (n >= 0xffffffffull) return fail(ctx, TFSM_ERR_LIMIT, "more than 2^32 training rows");
    return guarded(ctx, [&]() -> int {
        float *dx, *dy;
        stage_rows(ctx, x, y, n, &dx, &dy);
        float *master = ctx->work.alloc<float>(kParams);
        float *m1 = ctx->work.alloc<float>(kParams);
        float *m2 = ctx->work.alloc<float>(kParams);
        float *dg = ctx->work.alloc<float>(kParams + 1);
        uint32_t *didx = ctx->work.alloc<uint32_t>(n);
        std::vector<uint16_t> hw(kParams);
        std::vector<float> hm(kParams);
        // every copy and fill on the context's stream (a non-blocking stream
        // does not wait for the null stream)
        HIPX(hipMemcpyAsync(hw.data(), ctx->d_rank_w, kParams * sizeof(uint16_t), hipMemcpyDeviceToHost, ctx->stream));
        HIPX(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < kParams; ++i) hm[i] = host_f32(hw[i]);
        HIPX(hipMemcpyAsync(master, hm.data(), kParams * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        HIPX(hipMemsetAsync(m1, 0, kParams * sizeof(float), ctx->stream));
        HIPX(hipMemsetAsync(m2, 0, kParams * sizeof(float), ctx->stream));
        std::vector<uint32_t> perm(n);
        for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
        uint64_t state = seed ? seed : 0x9E3779B97F4A7C15ull;
        auto next = [&]() {   // splitmix64
            uint64_t z = (state += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        uint32_t step = 0;
        float last = 0.f;
        for (uint32_t e = 0; e < epochs; ++e) {
            for (uint64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[next() % (i + 1)]);
            HIPX(hipMemcpyAsync(didx, perm.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
            float *eloss = ctx->work.alloc<float>(1);
            HIPX(hipMemsetAsync(eloss, 0, sizeof(float), ctx->stream));
            for (uint64_t b = 0; b < n; b += batch) {
                const uint64_t nb = std::min<uint64_t>(batch, n - b);
                HIPX(hipMemsetAsync(dg, 0, kParams * sizeof(float), ctx->stream));
                hipLaunchKernelGGL(k_mlp_backward, dim3(grid_for(nb, ctx->num_cus)), dim3(kBlock), 0, ctx->stream,
                                   dx, dy, didx + b, nb, ctx->d_rank_w, dg, eloss);
                ++step;
                hipLaunchKernelGGL(k_adam, dim3((kParams + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, master,
                                   m1, m2, dg, 1.f / (float)nb, lr, 0.9f, 0.999f, 1e-8f, step, ctx->d_rank_w);
                HIPX(hipGetLastError());
            }
            HIPX(hipMemcpyAsync(&last, eloss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
            HIPX(hipStreamSynchronize(ctx->stream));
            last /= (float)n;
        }
        if (loss) *loss = last;
        return TFSM_OK;
    });
}

int64_t tfsm_read_features(tfsm_ctx *ctx, float *out, uint64_t cap_rows) {
    if (!ctx) return TFSM_ERR_ARGUMENT;
    const uint64_t n = ctx->feat_len;
    int rc = guarded(ctx, [&]() -> int {
                 // an eval call returns with its retained-domain copies still queued
                 HIPX(hipStreamSynchronize(ctx->stream));
        if (out && n)
            HIPX(hipMemcpy(out, ctx->d_feat, std::min(n, cap_rows) * kFeat * sizeof(float), hipMemcpyDeviceToHost));
        return TFSM_OK;
    });
    return rc == TFSM_OK ? (int64_t)n : rc;
}

}  // extern "C"
