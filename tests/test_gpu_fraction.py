"""Fraction-Score support on the GPU (TFSM_MODE_FRACTION) against the CPU
oracle and the reference's fraction-score build.

* the device score table equals the oracle's bit for bit (and, on R-MAT with
  four edge labels, the reference loader's digest);
* per pattern: the verdict equals the oracle's mode-2 verdict, and each
  per-vertex sum equals math.fsum of the terms over the oracle's exact valid
  set (both sides are the correctly rounded exact sum, so equality is exact);
* whole runs reproduce the reference build's pattern sets (plain and batched).
"""
import hashlib
import math
import os

import numpy as np
import pytest

import oracle as orc
import tfsm
from conftest import golden_keys, golden_manifest, workload_path
from test_fraction import multi_label_graph, read_fixture, score_lines

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("workload,seed,freq", [("readme", 0, 1), ("gnm2k", 1, 40), ("powerlaw100k", 2, 6000),
                                                ("multilabel", 0, 1)])
def test_scores_equal_oracle(workload, seed, freq, workdir):
    path = multi_label_graph(workdir) if workload == "multilabel" else workload_path(workload, seed, workdir)
    print(f"scores {workload} seed {seed} freq {freq}", flush=True)
    g = tfsm.Graph(path, freq)
    ctx = tfsm.Context(g)
    mine = ctx.fraction_scores()
    theirs = orc.OracleGraph(path, freq).fraction_scores()
    assert mine.shape == theirs.shape
    assert np.array_equal(mine.view(np.uint64), theirs.view(np.uint64))
    ctx.close()


def test_scores_rmat_reference_digest(workdir):
    """R-MAT scale 20 (16 vertex labels, 4 edge labels): every score equals
    the reference loader's (the fixture's digest of all 4.4M entries)."""
    print("scores rmat20 seed 4 freq 5300", flush=True)
    g = tfsm.Graph(workload_path("rmat20", 4, workdir), 5300)
    ctx = tfsm.Context(g)
    sc = ctx.fraction_scores()
    labels = [int(x) for x in g.csr()["vlabel_values"]]
    n, (count, digest), head, _ = read_fixture("rmat20_f5300")
    assert sc.shape[0] == n
    lines = score_lines(sc, labels)
    assert lines[:len(head)] == head
    assert len(lines) == count
    assert hashlib.sha256("".join(l + "\n" for l in lines).encode()).hexdigest() == digest
    ctx.close()


def term_sums(og, vl, es, labels_sorted, scores):
    """Per u: math.fsum over the oracle's exact valid set of min over u's
    pattern neighbours of score(v, label)."""
    ex = og.eval(vl, es, 1, mode=1)
    out = []
    for u in range(len(vl)):
        nb = [b if a == u else a for a, b, _ in es if u in (a, b)]
        cols = [labels_sorted.index(vl[x]) for x in nb]
        dom, st = ex.domain(u), ex.states(u)
        valid = dom[st == 1] if ex.decided_by == 2 else dom
        out.append(math.fsum(float(scores[v, cols].min()) for v in valid))
    return out, ex


@pytest.mark.parametrize("name,limit", [("gnm2k_frac_f40_m4", 400), ("gnm2k_frac_f60_m4", 300)])
def test_pattern_verdicts_and_sums(name, limit, workdir):
    m = golden_manifest()[name]
    path = workload_path(m["workload"], m["seed"], workdir)
    freq = m["freq"]
    print(f"patterns {name} limit {limit}", flush=True)
    g = tfsm.Graph(path, freq)
    ctx = tfsm.Context(g)
    og = orc.OracleGraph(path, freq)
    csr = g.csr()
    vv, ev = [int(x) for x in csr["vlabel_values"]], [int(x) for x in csr["elabel_values"]]
    scores = og.fraction_scores()
    # the frequent patterns and, for each, a relabelled (mostly infrequent) variant
    cands = []
    for k in sorted(golden_keys(name))[:limit // 2]:
        vl, es = tfsm.parse_key(k)
        cands.append((vl, es))
        if vl[-1] != vl[0]:
            cands.append((vl[:-1] + [vl[0]], es))
    pats = [tfsm.make_pattern([vv.index(l) for l in vl], [(a, b, ev.index(l)) for a, b, l in es]) for vl, es in cands]
    res = ctx.eval_level(pats, freq, mode=tfsm.MODE_FRACTION)
    nfreq = nexp = 0
    for (vl, es), r in zip(cands, res):
        where = f"{vl} {es}"
        ref = og.eval(vl, es, freq, mode=2)
        assert bool(r.frequent) == ref.frequent, where
        nfreq += int(ref.frequent)
        if r.decided_by == tfsm.DECIDED_FILTER:
            continue
        if r.decided_by == tfsm.DECIDED_SEARCH and r.support_upper < freq:
            continue                                # stopped by the valid-count bound: sums partial
        sums, ex = term_sums(og, vl, es, vv, scores)
        assert r.decided_by == ex.decided_by, where
        nexp += int(r.decided_by == tfsm.DECIDED_EXPRESS)
        for u in range(len(vl)):
            assert r.fraction[u] == sums[u], f"{where}: u {u} {r.fraction[u]!r} vs {sums[u]!r}"
        assert r.fraction_support == min(sums), where
    assert nfreq >= limit // 4
    ctx.close()


FRAC_CASES = ["readme_frac_f1_m3", "gnm2k_frac_f40_m4", "gnm2k_frac_f60_m4", "powerlaw100k_frac_f3000_m5"]


@pytest.mark.parametrize("name", FRAC_CASES)
@pytest.mark.parametrize("batched", [False, True])
def test_fraction_pattern_sets(name, batched, workdir):
    man = golden_manifest()
    if name not in man:
        pytest.skip(f"{name}: no reference fixture")
    m = man[name]
    path = workload_path(m["workload"], m["seed"], workdir)
    print(f"mine {name} batched {batched}", flush=True)
    g = tfsm.Graph(path, m["freq"])
    ctx = tfsm.Context(g)
    kp = os.path.join(str(workdir), f"{name}_{int(batched)}_frac.keys")
    n, _ = ctx.mine(m["freq"], m["max_nodes"], threads=4, batched=batched, mode=tfsm.MODE_FRACTION, keys_path=kp)
    keys = {}
    for line in open(kp):
        k, lo, hi, fr = line.rstrip("\n").split("\t")
        keys[k] = (int(lo), float.fromhex(fr))
    assert n == m["reported_total"]
    assert {k for k in keys if k.count(";") >= 2} == golden_keys(name)
    for k, (lo, fr) in keys.items():
        assert fr >= m["freq"], k
    ctx.close()
