"""Push-down pruning on the GPU (the reference's sub-pattern cache,
parallel/decompose.h:19-94 and comper.h:1582-1612).

* A pattern evaluated with push-down sources (the retained child domains of an
  isomorphic copy of itself, vertices renumbered) gets exactly the domains,
  per-candidate verdicts and supports the oracle computes when the same child
  domains are its candidate sources: the kernel keeps a candidate only when it
  lies in every source's child domain, with the vertex mapping applied.
* Whole mining runs with push-down pruning find the reference's pattern sets;
  in exact mode (every candidate decided, so a frequent pattern's child
  domains are its exact valid sets and do not depend on scheduling) the
  filtered domains sum to no more than without it.  In capped mode a parent's
  proven-invalid candidates depend on when its vertices reached freq, so
  domain sums of two runs are not comparable there.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
import tfsm
from conftest import golden_keys, workload_path
from test_gpu_parity import variants

pytestmark = pytest.mark.gpu


class Push(ctypes.Structure):
    _fields_ = [("pattern", ctypes.c_uint32), ("vertex", ctypes.c_uint32), ("handle", ctypes.c_uint64),
                ("source_vertex", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


def test_pushdown_sources_parity(workdir):
    path = workload_path("gnm2k", 1, workdir)
    freq = 40
    g = tfsm.Graph(path, freq)
    ctx = tfsm.Context(g)
    og = orc.OracleGraph(path, freq)
    csr = g.csr()
    vv = [int(x) for x in csr["vlabel_values"]]
    ev = [int(x) for x in csr["elabel_values"]]
    rng = np.random.default_rng(3)
    lib = tfsm.gpu_lib()
    checked = shrunk = 0
    for vl, es in variants(golden_keys("gnm2k_f40_m4"), 40):
        n = len(vl)
        perm = rng.permutation(n)                   # vertex x of the pattern is perm[x] of the copy
        cvl = [0] * n
        for x in range(n):
            cvl[perm[x]] = vl[x]
        ces = [(int(perm[a]), int(perm[b]), l) for a, b, l in es]
        copy = tfsm.make_pattern([vv.index(l) for l in cvl], [(a, b, ev.index(l)) for a, b, l in ces])
        cres = ctx.eval_pattern(copy, freq, mode=tfsm.MODE_EXACT, retain=True)
        if not cres.domain_handle:
            continue                                # infrequent copy: nothing retained
        child = [ctx.domain_read(cres.domain_handle, int(perm[x])) for x in range(n)]
        items = (Push * n)(*[Push(0, x, cres.domain_handle, int(perm[x]), 0) for x in range(n)])
        pat = tfsm.make_pattern([vv.index(l) for l in vl], [(a, b, ev.index(l)) for a, b, l in es])
        opt = tfsm.EvalOptions(mode=tfsm.MODE_EXACT, num_pushdown=n,
                               pushdown=ctypes.cast(items, ctypes.c_void_p))
        res = tfsm.PatternResult()
        rc = lib.tfsm_eval_pattern(ctx.h, ctypes.byref(pat), ctypes.c_uint64(0), freq, ctypes.byref(opt),
                                   ctypes.byref(res))
        assert rc == 0, lib.tfsm_last_error(ctx.h)
        ref = og.eval(vl, es, freq, mode=1, parent_domains=child)
        plain = og.eval(vl, es, freq, mode=1)
        assert (res.support_lower, res.support_upper) == (ref.lower, ref.upper) == (plain.lower, plain.upper)
        for u in range(n):
            assert np.array_equal(ctx.last_domain(u), ref.domain(u)), f"{vl} {es} domain {u}"
            if ref.decided_by == 2:
                assert np.array_equal(ctx.last_states(u), ref.states(u)), f"{vl} {es} verdicts {u}"
            shrunk += len(ref.domain(u)) < len(plain.domain(u))
        ctx.domain_release(cres.domain_handle)
        checked += 1
    assert checked > 10
    # a bad source is an argument error
    bad = (Push * 1)(Push(0, 0, 987654321, 0, 0))
    opt = tfsm.EvalOptions(mode=tfsm.MODE_EXACT, num_pushdown=1, pushdown=ctypes.cast(bad, ctypes.c_void_p))
    res = tfsm.PatternResult()
    assert lib.tfsm_eval_pattern(ctx.h, ctypes.byref(pat), ctypes.c_uint64(0), freq, ctypes.byref(opt),
                                 ctypes.byref(res)) == -5
    ctx.close()
    g.close()


def read_keys(path, min_vertices):
    keys = set(l.split("\t")[0] for l in open(path) if l.strip())
    return set(k for k in keys if k.count(";") >= min_vertices - 1)


@pytest.mark.parametrize("name,seed,freq,maxn,golden,minv", [
    ("gnm2k", 1, 40, 4, "gnm2k_f40_m4", 2),
    ("powerlaw100k", 2, 9000, 5, "powerlaw100k_f9000_m5", 2),
])
@pytest.mark.parametrize("batched", [False, True], ids=["plain", "batched"])
def test_pushdown_mining(name, seed, freq, maxn, golden, minv, batched, workdir, tmp_path):
    g = tfsm.Graph(workload_path(name, seed, workdir), freq)
    ctx = tfsm.Context(g)
    stats = {}
    for mode in (tfsm.MODE_CAPPED, tfsm.MODE_EXACT):
        for pd in (False, True):
            kp = str(tmp_path / f"keys_{mode}_{int(pd)}.txt")
            n, st = ctx.mine(freq, maxn, threads=4, batched=batched, keys_path=kp, pushdown=pd, mode=mode)
            assert read_keys(kp, minv) == golden_keys(golden), f"mode={mode} pushdown={pd}"
            if mode == tfsm.MODE_EXACT:
                stats[pd] = st
    print("domain sum", stats[False]["domain_sum"], "->", stats[True]["domain_sum"],
          "searches", stats[False]["searches"], "->", stats[True]["searches"],
          "push-down entries", stats[True]["pushdown_sources"])
    assert stats[True]["pushdown_sources"] > 0
    assert stats[True]["domain_sum"] <= stats[False]["domain_sum"]
    ctx.close()
    g.close()
