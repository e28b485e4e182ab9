"""ctypes bindings of the CPU oracle (oracle.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
REF_PARALLEL = os.path.join(HERE, "_ref", "run_parallel")
REF_SERIAL = os.path.join(HERE, "_ref", "run_serial")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise OSError(f"{LIB} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        u32p, i32p = P(ctypes.c_uint32), P(ctypes.c_int32)
        L.orc_load.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.orc_load.restype = vp
        L.orc_graph_free.argtypes = [vp]
        L.orc_num_vertices.argtypes = [vp]
        L.orc_num_vertices.restype = ctypes.c_uint32
        L.orc_num_adj.argtypes = [vp]
        L.orc_num_adj.restype = ctypes.c_uint64
        L.orc_csr.argtypes = [vp, u32p, u32p, i32p, i32p]
        L.orc_edge_stats.argtypes = [vp, i32p, ctypes.c_int]
        L.orc_eval_pattern.argtypes = [vp, ctypes.c_int, i32p, ctypes.c_int, i32p, ctypes.c_int, u32p, u32p,
                                       ctypes.c_int, ctypes.c_int]
        L.orc_eval_pattern.restype = vp
        L.orc_eval_free.argtypes = [vp]
        for f in ["orc_eval_frequent", "orc_eval_decided_by"]:
            getattr(L, f).argtypes = [vp]
        for f in ["orc_eval_support_lower", "orc_eval_support_upper"]:
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = ctypes.c_uint32
        for f in ["orc_eval_searches", "orc_eval_verdicts"]:
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = ctypes.c_uint64
        L.orc_eval_domain.argtypes = [vp, ctypes.c_int, u32p, ctypes.c_uint32]
        L.orc_eval_domain.restype = ctypes.c_uint32
        L.orc_eval_states.argtypes = [vp, ctypes.c_int, P(ctypes.c_uint8), ctypes.c_uint32]
        L.orc_eval_states.restype = ctypes.c_uint32
        L.orc_eval_child_domain.argtypes = [vp, ctypes.c_int, u32p, ctypes.c_uint32]
        L.orc_eval_child_domain.restype = ctypes.c_uint32
        L.orc_eval_aux.argtypes = [vp, ctypes.c_int, ctypes.c_int, u32p, u32p]
        L.orc_eval_aux.restype = ctypes.c_int64
        L.orc_eval_order.argtypes = [vp, ctypes.c_int, u32p]
        L.orc_eval_bn.argtypes = [vp, ctypes.c_int, ctypes.c_int, u32p]
        L.orc_eval_bn.restype = ctypes.c_int
        L.orc_bruteforce_mni.argtypes = [vp, ctypes.c_int, i32p, ctypes.c_int, i32p, u32p]
        L.orc_bruteforce_mni.restype = ctypes.c_uint32
        L.orc_mine.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                               P(ctypes.c_double), P(ctypes.c_uint64), P(ctypes.c_uint64)]
        L.orc_mine.restype = ctypes.c_int64
        dp = P(ctypes.c_double)
        L.orc_fraction_scores.argtypes = [vp, dp, ctypes.c_uint64]
        L.orc_fraction_scores.restype = ctypes.c_int64
        L.orc_edge_fraction.argtypes = [vp, dp, ctypes.c_int]
        L.orc_exact_sum.argtypes = [dp, ctypes.c_uint64]
        L.orc_exact_sum.restype = ctypes.c_double
        L.orc_eval_fraction.argtypes = [vp, dp]
        L.orc_eval_fraction.restype = ctypes.c_double
        _lib = L
    return _lib


def _u32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _i32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def exact_sum(values):
    """The double nearest the exact sum (the oracle's Shewchuk partials)."""
    a = np.ascontiguousarray(values, np.float64)
    return lib().orc_exact_sum(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(a))


class OracleGraph:
    def __init__(self, path, freq):
        self.h = lib().orc_load(path.encode(), int(freq))
        if not self.h:
            raise OSError(f"oracle could not load {path}")
        self.freq = int(freq)
        self.n = lib().orc_num_vertices(self.h)
        self.num_adj = lib().orc_num_adj(self.h)

    def csr(self):
        off = np.zeros(self.n + 1, np.uint32)
        nbr = np.zeros(max(self.num_adj, 1), np.uint32)
        el = np.zeros(max(self.num_adj, 1), np.int32)
        vl = np.zeros(max(self.n, 1), np.int32)
        lib().orc_csr(self.h, _u32(off), _u32(nbr), _i32(el), _i32(vl))
        return off, nbr[:self.num_adj], el[:self.num_adj], vl[:self.n]

    def edge_stats(self):
        k = lib().orc_edge_stats(self.h, None, 0)
        out = np.zeros(max(5 * k, 1), np.int32)
        lib().orc_edge_stats(self.h, _i32(out), k)
        return out[:5 * k].reshape(-1, 5)

    def fraction_scores(self):
        """(n, L) float64: score(v, l), labels in ascending order."""
        k = lib().orc_fraction_scores(self.h, None, 0)
        out = np.zeros(max(k, 1), np.float64)
        lib().orc_fraction_scores(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), k)
        return out[:k].reshape(self.n, -1) if self.n else out[:0].reshape(0, 0)

    def edge_fraction(self):
        """(k, 2) float64: the size-2 sums over candA and candB, in edge_stats order."""
        k = lib().orc_edge_fraction(self.h, None, 0)
        out = np.zeros(max(2 * k, 1), np.float64)
        lib().orc_edge_fraction(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), k)
        return out[:2 * k].reshape(-1, 2)

    def eval(self, vlabels, edges, freq, mode=0, parent_domains=None):
        """vlabels/edges in original label values; parent_domains: list of sorted id arrays."""
        vl = np.asarray(vlabels, np.int32)
        ed = np.asarray([x for e in edges for x in e], np.int32)
        if parent_domains:
            off = np.zeros(len(parent_domains) + 1, np.uint32)
            off[1:] = np.cumsum([len(d) for d in parent_domains])
            dom = np.concatenate([np.asarray(d, np.uint32) for d in parent_domains]) if off[-1] else np.zeros(1, np.uint32)
            h = lib().orc_eval_pattern(self.h, len(vl), _i32(vl), len(edges), _i32(ed), len(parent_domains), _u32(off),
                                       _u32(dom), int(freq), int(mode))
        else:
            h = lib().orc_eval_pattern(self.h, len(vl), _i32(vl), len(edges), _i32(ed), 0, None, None, int(freq),
                                       int(mode))
        return OracleEval(h, len(vl))

    def bruteforce(self, vlabels, edges):
        vl = np.asarray(vlabels, np.int32)
        ed = np.asarray([x for e in edges for x in e], np.int32)
        sizes = np.zeros(len(vl), np.uint32)
        mni = lib().orc_bruteforce_mni(self.h, len(vl), _i32(vl), len(edges), _i32(ed), _u32(sizes))
        return int(mni), sizes

    def mine(self, freq, max_nodes, threads, mode, out_path):
        sec = ctypes.c_double()
        ev = ctypes.c_uint64()
        vd = ctypes.c_uint64()
        n = lib().orc_mine(self.h, int(freq), int(max_nodes), int(threads), int(mode), out_path.encode(),
                           ctypes.byref(sec), ctypes.byref(ev), ctypes.byref(vd))
        return int(n), dict(mining_s=sec.value, evaluated=ev.value, verdicts=vd.value)

    def close(self):
        if self.h:
            lib().orc_graph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OracleEval:
    def __init__(self, h, nv):
        self.h = h
        self.nv = nv

    @property
    def frequent(self):
        return bool(lib().orc_eval_frequent(self.h))

    @property
    def decided_by(self):
        return lib().orc_eval_decided_by(self.h)

    @property
    def lower(self):
        return int(lib().orc_eval_support_lower(self.h))

    @property
    def upper(self):
        return int(lib().orc_eval_support_upper(self.h))

    @property
    def searches(self):
        return int(lib().orc_eval_searches(self.h))

    @property
    def verdicts(self):
        return int(lib().orc_eval_verdicts(self.h))

    def fraction(self):
        """Mode 2: (per-vertex summed scores, their minimum)."""
        out = np.zeros(self.nv, np.float64)
        mn = lib().orc_eval_fraction(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out, mn

    def domain(self, u):
        n = lib().orc_eval_domain(self.h, u, None, 0)
        out = np.zeros(max(n, 1), np.uint32)
        lib().orc_eval_domain(self.h, u, _u32(out), n)
        return out[:n]

    def states(self, u):
        n = lib().orc_eval_states(self.h, u, None, 0)
        out = np.zeros(max(n, 1), np.uint8)
        lib().orc_eval_states(self.h, u, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), n)
        return out[:n]

    def child_domain(self, u):
        n = lib().orc_eval_child_domain(self.h, u, None, 0)
        out = np.zeros(max(n, 1), np.uint32)
        lib().orc_eval_child_domain(self.h, u, _u32(out), n)
        return out[:n]

    def aux(self, a, b):
        total = lib().orc_eval_aux(self.h, a, b, None, None)
        if total < 0:
            return None
        rows = len(self.domain(a))
        off = np.zeros(rows + 1, np.uint32)
        idx = np.zeros(max(total, 1), np.uint32)
        lib().orc_eval_aux(self.h, a, b, _u32(off), _u32(idx))
        return off, idx[:total]

    def order(self, u):
        out = np.zeros(self.nv, np.uint32)
        lib().orc_eval_order(self.h, u, _u32(out))
        return list(out)

    def bn(self, u, d):
        out = np.zeros(self.nv, np.uint32)
        k = lib().orc_eval_bn(self.h, u, d, _u32(out))
        return None if k < 0 else [int(x) for x in out[:k]]

    def close(self):
        if self.h:
            lib().orc_eval_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
