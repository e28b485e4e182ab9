"""ctypes bindings for the MI355X MNI evaluation library and its host side.

`libtfsm_gpu.so` is the C ABI of include/tfsm_gpu.h (HIP kernels for gfx950);
`libtfsm_host.so` holds the loader, the gSpan growth and the miner.  Both are
built in-tree into t-fsm_amd/_build/ by `make -C t-fsm_amd`.  Loading fails
loudly when a library is missing: there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

MAX_V = 16
MAX_E = 64
MODE_CAPPED = 0
MODE_EXACT = 1
MODE_FRACTION = 2
ORDER_REFERENCE = 0
ORDER_LEARNED = 1
MATCH_ADAPTIVE = 0
MATCH_REFERENCE = 1
LEVEL_DEFAULT = 0
LEVEL_HOST = 1
LEVEL_DEVICE = 2
LEVEL_GRAPH = 3
DECIDED_FILTER, DECIDED_EXPRESS, DECIDED_SEARCH = 0, 1, 2

STATUS = {0: "ok", -1: "invalid argument", -2: "HIP runtime error", -3: "no usable GPU",
          -4: "input beyond a supported limit", -5: "unknown domain handle"}


class TfsmError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{STATUS.get(code, code)} ({code}) {msg}".strip())
        self.code = code


class Csr(ctypes.Structure):
    _fields_ = [("num_vertices", ctypes.c_uint32), ("num_adj", ctypes.c_uint64),
                ("offsets", ctypes.POINTER(ctypes.c_uint32)), ("nbrs", ctypes.POINTER(ctypes.c_uint32)),
                ("elabels", ctypes.POINTER(ctypes.c_uint32)), ("vlabels", ctypes.POINTER(ctypes.c_uint32)),
                ("num_vlabels", ctypes.c_uint32), ("num_elabels", ctypes.c_uint32)]


class Pattern(ctypes.Structure):
    _fields_ = [("num_vertices", ctypes.c_uint32), ("num_edges", ctypes.c_uint32),
                ("vlabel", ctypes.c_uint32 * MAX_V), ("edge_a", ctypes.c_uint32 * MAX_E),
                ("edge_b", ctypes.c_uint32 * MAX_E), ("edge_label", ctypes.c_uint32 * MAX_E)]


class EvalOptions(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_uint32), ("retain_domains", ctypes.c_uint32), ("task_order", ctypes.c_uint32),
                ("trace", ctypes.c_uint32), ("grid_waves_per_cu", ctypes.c_uint32), ("split_steps", ctypes.c_uint32),
                ("split_fanout", ctypes.c_uint32), ("long_steps", ctypes.c_uint32),
                ("long_waves_per_cu", ctypes.c_uint32), ("long_donate_every", ctypes.c_uint32),
                ("long_fanout", ctypes.c_uint32), ("match_order", ctypes.c_uint32),
                ("group_lanes", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("num_pushdown", ctypes.c_uint32), ("pushdown", ctypes.c_void_p),
                ("steps_per_round", ctypes.c_uint32), ("refill_min", ctypes.c_uint32),
                ("refill_max", ctypes.c_uint32), ("prune_batch", ctypes.c_uint32), ("level_mode", ctypes.c_uint32),
                ("hot_group_lanes", ctypes.c_uint32), ("ldf_push", ctypes.c_uint32)]


class PatternResult(ctypes.Structure):
    _fields_ = [("frequent", ctypes.c_int32), ("decided_by", ctypes.c_int32),
                ("support_lower", ctypes.c_uint32), ("support_upper", ctypes.c_uint32),
                ("domain_size", ctypes.c_uint32 * MAX_V), ("valid_count", ctypes.c_uint32 * MAX_V),
                ("invalid_count", ctypes.c_uint32 * MAX_V), ("searches", ctypes.c_uint64),
                ("verdicts", ctypes.c_uint64), ("domain_handle", ctypes.c_uint64),
                ("fraction", ctypes.c_double * MAX_V), ("fraction_support", ctypes.c_double)]


class CallStats(ctypes.Structure):
    _fields_ = [("filter_ms", ctypes.c_float), ("table_ms", ctypes.c_float), ("search_ms", ctypes.c_float),
                ("search_launches", ctypes.c_uint32), ("table_bytes", ctypes.c_uint64),
                ("search_bytes", ctypes.c_uint64), ("search_splits", ctypes.c_uint32),
                ("search_watchdog", ctypes.c_uint32), ("rank_ms", ctypes.c_float), ("level_mode", ctypes.c_uint32),
                ("filter_passes", ctypes.c_uint32)]


RANKER_FEATURES = 24
RANKER_HIDDEN = 64
RANKER_PARAMS = RANKER_HIDDEN * RANKER_FEATURES + 2 * RANKER_HIDDEN + 1


class RankerWeights(ctypes.Structure):
    _fields_ = [("w1", ctypes.c_float * (RANKER_HIDDEN * RANKER_FEATURES)), ("b1", ctypes.c_float * RANKER_HIDDEN),
                ("w2", ctypes.c_float * RANKER_HIDDEN), ("b2", ctypes.c_float)]

    def to_numpy(self):
        import numpy as np
        return np.frombuffer(bytes(self), dtype=np.float32).copy()

    @staticmethod
    def from_numpy(flat):
        import numpy as np
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        assert flat.size == RANKER_PARAMS
        w = RankerWeights()
        ctypes.memmove(ctypes.addressof(w), flat.ctypes.data, flat.nbytes)
        return w


# every symbol include/tfsm_gpu.h declares
GPU_SYMBOLS = ["tfsm_ctx_create", "tfsm_ctx_destroy", "tfsm_ctx_set_pool_capacity", "tfsm_eval_pattern", "tfsm_eval_level",
               "tfsm_domain_read", "tfsm_domain_release", "tfsm_last_domain", "tfsm_last_states",
               "tfsm_last_aux", "tfsm_last_order", "tfsm_read_trace", "tfsm_last_stats",
               "tfsm_status_string", "tfsm_last_error", "tfsm_ranker_set", "tfsm_ranker_get",
               "tfsm_ranker_forward", "tfsm_ranker_gradient", "tfsm_ranker_train", "tfsm_read_features",
               "tfsm_ranker_gradient_noalloc", "tfsm_fraction_scores"]

_gpu = None
_host = None


def _load(name):
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with `make -C t-fsm_amd` (or __graft_entry__.build())")
    return ctypes.CDLL(path)


def gpu_lib():
    """The C ABI library; its argument types are declared once here."""
    global _gpu
    if _gpu is None:
        lib = _load("libtfsm_gpu.so")
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        lib.tfsm_ctx_create.argtypes = [P(Csr), P(vp)]
        lib.tfsm_ctx_destroy.argtypes = [vp]
        lib.tfsm_ctx_set_pool_capacity.argtypes = [vp, ctypes.c_uint64]
        lib.tfsm_eval_pattern.argtypes = [vp, P(Pattern), ctypes.c_uint64, ctypes.c_uint32, P(EvalOptions),
                                          P(PatternResult)]
        lib.tfsm_eval_level.argtypes = [vp, P(Pattern), P(ctypes.c_uint64), ctypes.c_uint32, ctypes.c_uint32,
                                        P(EvalOptions), P(PatternResult)]
        lib.tfsm_domain_read.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint32, P(ctypes.c_uint32), ctypes.c_uint32,
                                         P(ctypes.c_uint32)]
        lib.tfsm_domain_release.argtypes = [vp, ctypes.c_uint64]
        lib.tfsm_last_domain.argtypes = [vp, ctypes.c_uint32, P(ctypes.c_uint32), ctypes.c_uint32]
        lib.tfsm_last_domain.restype = ctypes.c_int64
        lib.tfsm_last_states.argtypes = [vp, ctypes.c_uint32, P(ctypes.c_uint8), ctypes.c_uint32]
        lib.tfsm_last_states.restype = ctypes.c_int64
        lib.tfsm_last_aux.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, P(ctypes.c_uint32), ctypes.c_uint32,
                                      P(ctypes.c_uint32), ctypes.c_uint32]
        lib.tfsm_last_aux.restype = ctypes.c_int64
        lib.tfsm_last_order.argtypes = [vp, ctypes.c_uint32, P(ctypes.c_uint32)]
        lib.tfsm_read_trace.argtypes = [vp, P(ctypes.c_uint32), ctypes.c_uint64]
        lib.tfsm_read_trace.restype = ctypes.c_int64
        lib.tfsm_last_stats.argtypes = [vp, P(CallStats)]
        lib.tfsm_status_string.restype = ctypes.c_char_p
        lib.tfsm_last_error.argtypes = [vp]
        lib.tfsm_last_error.restype = ctypes.c_char_p
        fp = P(ctypes.c_float)
        lib.tfsm_ranker_set.argtypes = [vp, P(RankerWeights)]
        lib.tfsm_ranker_get.argtypes = [vp, P(RankerWeights)]
        lib.tfsm_ranker_forward.argtypes = [vp, fp, ctypes.c_uint64, fp]
        lib.tfsm_ranker_gradient.argtypes = [vp, fp, fp, ctypes.c_uint64, P(RankerWeights), fp]
        lib.tfsm_ranker_train.argtypes = [vp, fp, fp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_float, ctypes.c_uint64, fp]
        lib.tfsm_read_features.argtypes = [vp, fp, ctypes.c_uint64]
        # device pointers (integers) for the no-allocation gradient
        lib.tfsm_ranker_gradient_noalloc.argtypes = [vp, vp, vp, ctypes.c_uint64, vp, vp]
        lib.tfsm_read_features.restype = ctypes.c_int64
        lib.tfsm_fraction_scores.argtypes = [vp, P(ctypes.c_double), ctypes.c_uint64]
        lib.tfsm_fraction_scores.restype = ctypes.c_int64
        _gpu = lib
    return _gpu


def host_lib():
    global _host
    if _host is None:
        gpu_lib()
        lib = _load("libtfsm_host.so")
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        lib.tfsmh_load.argtypes = [ctypes.c_char_p, ctypes.c_int]
        lib.tfsmh_load.restype = vp
        lib.tfsmh_load_cached.argtypes = [ctypes.c_char_p, ctypes.c_int, P(ctypes.c_int)]
        lib.tfsmh_load_cached.restype = vp
        lib.tfsmh_free.argtypes = [vp]
        lib.tfsmh_info.argtypes = [vp, P(ctypes.c_uint64)]
        lib.tfsmh_csr.argtypes = [vp, P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32),
                                  P(ctypes.c_uint32), P(ctypes.c_int32), P(ctypes.c_int32)]
        lib.tfsmh_edge_stats.argtypes = [vp, P(ctypes.c_int32)]
        lib.tfsmh_ctx_create.argtypes = [vp, P(vp)]
        lib.tfsmh_mine.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, P(ctypes.c_double),
                                   P(EvalOptions)]
        lib.tfsmh_mine.restype = ctypes.c_int64
        lib.tfsmh_is_min_code.argtypes = [P(ctypes.c_int32), ctypes.c_int]
        lib.tfsmh_extend.argtypes = [vp, ctypes.c_int, P(ctypes.c_int32), ctypes.c_int, P(ctypes.c_int32),
                                     ctypes.c_int]
        lib.tfsmh_last_error.restype = ctypes.c_char_p
        lib.tfsmh_load_ranker.argtypes = [vp, ctypes.c_char_p]
        lib.tfsmh_load_tuning.argtypes = [ctypes.c_char_p, P(EvalOptions)]
        _host = lib
    return _host


def _check(ctx, rc):
    if rc != 0:
        msg = gpu_lib().tfsm_last_error(ctx) if ctx else b""
        raise TfsmError(rc, (msg or b"").decode())
    return rc


class Graph:
    """Pruned data graph loaded and relabelled by the product's host loader."""

    def __init__(self, path, freq, cache=False):
        """cache: load through the binary cache beside the file
        (`<path>.f<freq>.csr`, written when absent or stale); `cache_hit`
        then says whether it was used."""
        self.lib = host_lib()
        self.cache_hit = False
        if cache:
            hit = ctypes.c_int(0)
            self.h = self.lib.tfsmh_load_cached(path.encode(), int(freq), ctypes.byref(hit))
            self.cache_hit = bool(hit.value)
        else:
            self.h = self.lib.tfsmh_load(path.encode(), int(freq))
        if not self.h:
            raise TfsmError(-1, self.lib.tfsmh_last_error().decode())
        info = (ctypes.c_uint64 * 5)()
        self.lib.tfsmh_info(self.h, info)
        self.n, self.num_adj, self.num_vlabels, self.num_elabels, self.num_edge_stats = [int(x) for x in info]
        self.freq = int(freq)

    def csr(self):
        import numpy as np
        off = np.zeros(self.n + 1, np.uint32)
        nbr = np.zeros(max(self.num_adj, 1), np.uint32)
        el = np.zeros(max(self.num_adj, 1), np.uint32)
        vl = np.zeros(max(self.n, 1), np.uint32)
        vv = np.zeros(max(self.num_vlabels, 1), np.int32)
        ev = np.zeros(max(self.num_elabels, 1), np.int32)
        p32 = lambda a, t=ctypes.c_uint32: a.ctypes.data_as(ctypes.POINTER(t))
        self.lib.tfsmh_csr(self.h, p32(off), p32(nbr), p32(el), p32(vl), p32(vv, ctypes.c_int32),
                           p32(ev, ctypes.c_int32))
        return dict(offsets=off, nbrs=nbr[:self.num_adj], elabels=el[:self.num_adj], vlabels=vl[:self.n],
                    vlabel_values=vv[:self.num_vlabels], elabel_values=ev[:self.num_elabels])

    def edge_stats(self):
        import numpy as np
        out = np.zeros(max(5 * self.num_edge_stats, 1), np.int32)
        self.lib.tfsmh_edge_stats(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return out[:5 * self.num_edge_stats].reshape(-1, 5)

    def extend(self, codes, max_nodes=-1):
        """Canonical children of a DFS code [(from, to, fl, el, tl), ...]."""
        flat = (ctypes.c_int32 * (5 * len(codes)))(*[x for c in codes for x in c])
        cap = 1 << 20
        out = (ctypes.c_int32 * cap)()
        n = self.lib.tfsmh_extend(self.h, int(max_nodes), flat, len(codes), out, cap)
        if n < 0:
            raise TfsmError(-4, "extension buffer too small")
        kids, k, m = [], 0, 5 * (len(codes) + 1)
        for _ in range(n):
            kids.append([tuple(out[k + 5 * i:k + 5 * i + 5]) for i in range(len(codes) + 1)])
            k += m
        return kids

    def size2_codes(self):
        """DFS codes of the frequent single-edge patterns."""
        return [[(0, 1, int(r[0]), int(r[1]), int(r[2]))] for r in self.edge_stats()]

    def close(self):
        if self.h:
            self.lib.tfsmh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """A device context holding the graph in HBM."""

    def __init__(self, graph):
        self.lib = gpu_lib()
        self.graph = graph
        h = ctypes.c_void_p()
        rc = host_lib().tfsmh_ctx_create(graph.h, ctypes.byref(h))
        if rc != 0:
            raise TfsmError(rc, "tfsm_ctx_create")
        self.h = h

    def eval_pattern(self, pat, freq, mode=MODE_CAPPED, parent=0, retain=False, trace=False, learned=False,
                     **tuning):
        opt = EvalOptions(mode=mode, retain_domains=int(retain),
                          task_order=ORDER_LEARNED if learned else ORDER_REFERENCE, trace=int(trace), **tuning)
        res = PatternResult()
        _check(self.h, self.lib.tfsm_eval_pattern(self.h, ctypes.byref(pat), ctypes.c_uint64(parent), int(freq),
                                                  ctypes.byref(opt), ctypes.byref(res)))
        return res

    def eval_level(self, pats, freq, mode=MODE_CAPPED, parents=None, retain=False, trace=False, learned=False,
                   **tuning):
        """tuning: grid_waves_per_cu, split_steps, split_fanout, long_steps,
        long_waves_per_cu, long_donate_every, long_fanout (0: tuned default),
        match_order (MATCH_ADAPTIVE or MATCH_REFERENCE)."""
        n = len(pats)
        arr = (Pattern * n)(*pats)
        par = (ctypes.c_uint64 * n)(*(parents or [0] * n))
        res = (PatternResult * n)()
        opt = EvalOptions(mode=mode, retain_domains=int(retain),
                          task_order=ORDER_LEARNED if learned else ORDER_REFERENCE, trace=int(trace), **tuning)
        _check(self.h, self.lib.tfsm_eval_level(self.h, arr, par, n, int(freq), ctypes.byref(opt), res))
        return list(res)

    def set_pool_capacity(self, entries):
        _check(self.h, self.lib.tfsm_ctx_set_pool_capacity(self.h, ctypes.c_uint64(int(entries))))

    def last_domain(self, u):
        import numpy as np
        n = self.lib.tfsm_last_domain(self.h, u, None, 0)
        _check(self.h, 0 if n >= 0 else int(n))
        out = np.zeros(max(n, 1), np.uint32)
        self.lib.tfsm_last_domain(self.h, u, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n)
        return out[:n]

    def last_states(self, u):
        import numpy as np
        n = self.lib.tfsm_last_states(self.h, u, None, 0)
        _check(self.h, 0 if n >= 0 else int(n))
        out = np.zeros(max(n, 1), np.uint8)
        self.lib.tfsm_last_states(self.h, u, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), n)
        return out[:n]

    def last_aux(self, a, b):
        import numpy as np
        total = self.lib.tfsm_last_aux(self.h, a, b, None, 0, None, 0)
        if total < 0:
            return None
        rows = len(self.last_domain(a))
        off = np.zeros(rows + 1, np.uint32)
        idx = np.zeros(max(total, 1), np.uint32)
        self.lib.tfsm_last_aux(self.h, a, b, off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), rows + 1,
                               idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), total)
        return off, idx[:total]

    def last_order(self, u, nv):
        out = (ctypes.c_uint32 * MAX_V)()
        _check(self.h, self.lib.tfsm_last_order(self.h, u, out))
        return list(out)[:nv]

    def last_stats(self):
        st = CallStats()
        _check(self.h, self.lib.tfsm_last_stats(self.h, ctypes.byref(st)))
        return st

    def read_trace(self):
        import numpy as np
        n = self.lib.tfsm_read_trace(self.h, None, 0)
        out = np.zeros(max(n, 1), np.uint32)
        self.lib.tfsm_read_trace(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n)
        return out[:n]

    def read_features(self):
        import numpy as np
        n = self.lib.tfsm_read_features(self.h, None, 0)
        out = np.zeros((max(n, 1), RANKER_FEATURES), np.float32)
        self.lib.tfsm_read_features(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n)
        return out[:n]

    def ranker_load(self, path):
        _check(self.h, host_lib().tfsmh_load_ranker(self.h, path.encode()))

    def ranker_set(self, flat):
        w = RankerWeights.from_numpy(flat)
        _check(self.h, self.lib.tfsm_ranker_set(self.h, ctypes.byref(w)))

    def ranker_get(self):
        w = RankerWeights()
        _check(self.h, self.lib.tfsm_ranker_get(self.h, ctypes.byref(w)))
        return w.to_numpy()

    def ranker_forward(self, X):
        import numpy as np
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(len(X), np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        _check(self.h, self.lib.tfsm_ranker_forward(self.h, X.ctypes.data_as(fp), len(X), out.ctypes.data_as(fp)))
        return out

    def ranker_gradient(self, X, y):
        import numpy as np
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = RankerWeights()
        loss = ctypes.c_float()
        fp = ctypes.POINTER(ctypes.c_float)
        _check(self.h, self.lib.tfsm_ranker_gradient(self.h, X.ctypes.data_as(fp), y.ctypes.data_as(fp), len(X),
                                                     ctypes.byref(g), ctypes.byref(loss)))
        return g.to_numpy(), loss.value

    def ranker_train(self, X, y, epochs=4, batch=4096, lr=3e-3, seed=1):
        import numpy as np
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = ctypes.c_float()
        fp = ctypes.POINTER(ctypes.c_float)
        _check(self.h, self.lib.tfsm_ranker_train(self.h, X.ctypes.data_as(fp), y.ctypes.data_as(fp), len(X), epochs,
                                                  batch, lr, seed, ctypes.byref(loss)))
        return loss.value

    def fraction_scores(self):
        """The Fraction-Score table, (num_vertices, num_vlabels) float64."""
        import numpy as np
        n = self.lib.tfsm_fraction_scores(self.h, None, 0)
        _check(self.h, 0 if n >= 0 else int(n))
        out = np.zeros(max(n, 1), np.float64)
        self.lib.tfsm_fraction_scores(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
        return out[:n].reshape(self.graph.n, -1) if self.graph.n else out[:0].reshape(0, 0)

    def domain_read(self, handle, u):
        import numpy as np
        size = ctypes.c_uint32()
        _check(self.h, self.lib.tfsm_domain_read(self.h, handle, u, None, 0, ctypes.byref(size)))
        out = np.zeros(max(size.value, 1), np.uint32)
        _check(self.h, self.lib.tfsm_domain_read(self.h, handle, u, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                                 size.value, ctypes.byref(size)))
        return out[:size.value]

    def domain_release(self, handle):
        _check(self.h, self.lib.tfsm_domain_release(self.h, handle))

    def mine(self, freq, max_nodes=-1, threads=1, batched=False, mode=MODE_CAPPED, learned=False,
             results_path=None, keys_path=None, pushdown=False, **tuning):
        """tuning: search-kernel shape for every evaluation (see eval_level);
        pushdown: push-down pruning from cached frequent sub-patterns."""
        stats = (ctypes.c_double * 20)()
        opt = EvalOptions(**tuning)
        n = host_lib().tfsmh_mine(self.graph.h, self.h, int(freq), int(max_nodes), int(threads), int(batched),
                                  int(mode), int(learned), int(pushdown), results_path.encode() if results_path else None,
                                  keys_path.encode() if keys_path else None, stats, ctypes.byref(opt))
        if n < 0:
            raise TfsmError(int(n), host_lib().tfsmh_last_error().decode())
        names = ["mining_s", "evaluated", "frequent", "verdicts", "searches", "eval_calls", "filter_ms", "table_ms",
                 "search_ms", "search_launches", "bytes_table", "bytes_search", "levels", "max_size", "domain_sum",
                 "pushdown_sources", "filter_passes", "device_calls", "graph_calls"]
        return int(n), dict(zip(names, list(stats)[:len(names)]))

    def close(self):
        if self.h:
            self.lib.tfsm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_tuning(path):
    """Search-kernel options of a tuning file (keyword arguments for mine and
    eval_level), read by the same loader as `run -tune`."""
    opt = EvalOptions()
    if host_lib().tfsmh_load_tuning(path.encode(), ctypes.byref(opt)) != 0:
        raise TfsmError(-1, host_lib().tfsmh_last_error().decode())
    names = ["grid_waves_per_cu", "split_steps", "split_fanout", "long_steps", "long_waves_per_cu",
             "long_donate_every", "long_fanout", "match_order", "group_lanes", "steps_per_round", "refill_min",
             "refill_max", "prune_batch", "level_mode", "hot_group_lanes", "ldf_push"]
    return {k: getattr(opt, k) for k in names if getattr(opt, k)}


def make_pattern(vlabels, edges):
    """Pattern from dense labels and (a, b, dense edge label) triples."""
    p = Pattern()
    p.num_vertices = len(vlabels)
    p.num_edges = len(edges)
    for i, l in enumerate(vlabels):
        p.vlabel[i] = l
    for i, (a, b, l) in enumerate(edges):
        p.edge_a[i], p.edge_b[i], p.edge_label[i] = a, b, l
    return p


def parse_key(key):
    """DFS-code key 'from,to,fl,el,tl;...' -> (vertex labels, [(a, b, el)]) in original label values."""
    codes = [tuple(int(x) for x in c.split(",")) for c in key.strip(";").split(";") if c]
    vl = {}
    edges = []
    for f, t, fl, el, tl in codes:
        vl[f] = fl
        vl[t] = tl
        edges.append((f, t, el))
    return [vl[i] for i in range(len(vl))], edges
