"""Fraction-Score support (the reference's fraction-score build) in the CPU
oracle, pinned to the reference itself and to what the definition fixes:

* the per-vertex scores equal, bit for bit, the ones the reference's own
  compute_fraction_score writes (fixtures from tests/golden/ref_fraction_dump.cpp);
* the size-2 filter keeps exactly the reference's entries, with sums equal to
  the reference's up to its addition-order rounding;
* sums are the correctly rounded exact sums (Python's math.fsum);
* closed forms on unions of stars, and brute-force images on tiny graphs;
* mining reproduces the reference build's pattern sets.
"""
import hashlib
import itertools
import math
import os
import random

import numpy as np
import pytest

import oracle as orc
from conftest import GOLDEN, golden_keys, golden_manifest, workload_path

FRAC_FIXTURES = {"readme_f1": ("readme", 0, 1), "gnm2k_f40": ("gnm2k", 1, 40),
                 "powerlaw100k_f6000": ("powerlaw100k", 2, 6000)}


def c_hex(x):
    """C99 printf %a of a positive normal double (glibc form: trailing zeros dropped)."""
    h = float(x).hex()                       # 0x1.8000000000000p-1
    mant, exp = h[2:].split("p")
    if "." in mant:
        mant = mant.rstrip("0").rstrip(".")
    return f"0x{mant}p{exp}"


def read_fixture(name):
    n, h, s, e = None, None, [], []
    for line in open(os.path.join(GOLDEN, f"frac_{name}.txt")):
        t = line.split()
        if not t or t[0] == "#":
            continue
        if t[0] == "N":
            n = int(t[1])
        elif t[0] == "H":
            h = (int(t[1]), t[2])
        elif t[0] == "S":
            s.append(line.rstrip("\n"))
        elif t[0] == "E":
            e.append((int(t[1]), int(t[2]), int(t[3]), int(t[4]), int(t[5]), float.fromhex(t[6]),
                      float.fromhex(t[7])))
    return n, h, s, e


def score_lines(scores, labels):
    """The dump's S lines from a dense (n, L) table: one per nonzero entry
    (a present score is at least 2^-32, an absent one 0)."""
    lines = []
    for v in range(scores.shape[0]):
        for j in np.nonzero(scores[v])[0]:
            lines.append(f"S {v} {labels[j]} {c_hex(scores[v, j])}")
    return lines


def oracle_labels(og):
    return sorted(set(int(x) for x in og.csr()[3]))


@pytest.mark.parametrize("name", list(FRAC_FIXTURES))
def test_scores_match_reference(name, workdir):
    workload, seed, freq = FRAC_FIXTURES[name]
    og = orc.OracleGraph(workload_path(workload, seed, workdir), freq)
    n, (count, digest), s_head, _ = read_fixture(name)
    assert og.n == n
    lines = score_lines(og.fraction_scores(), oracle_labels(og))
    assert lines[:len(s_head)] == s_head
    assert len(lines) == count
    assert hashlib.sha256("".join(l + "\n" for l in lines).encode()).hexdigest() == digest


@pytest.mark.parametrize("name", list(FRAC_FIXTURES))
def test_size2_filter_matches_reference(name, workdir):
    workload, seed, freq = FRAC_FIXTURES[name]
    og = orc.OracleGraph(workload_path(workload, seed, workdir), freq)
    _, _, _, ref = read_fixture(name)
    stats, sums = og.edge_stats(), og.edge_fraction()
    kept = {}
    for (la, el, lb, na, nb), (sa, sb) in zip(stats, sums):
        if sa >= freq and sb >= freq:
            kept[(int(la), int(el), int(lb))] = (int(na), int(nb), sa, sb)
    assert sorted(kept) == sorted((r[0], r[1], r[2]) for r in ref)
    for la, el, lb, na, nb, sa, sb in ref:
        k = kept[(la, el, lb)]
        assert k[:2] == (na, nb)
        # the reference adds in hash-set order: equal up to that order's rounding
        assert abs(k[2] - sa) <= max(na, 1) * 2.0 ** -52 * sa
        assert abs(k[3] - sb) <= max(nb, 1) * 2.0 ** -52 * sb


def test_exact_sum_is_fsum(workdir):
    rng = np.random.default_rng(7)
    for _ in range(300):
        x = 1.0 / rng.integers(1, 40, size=int(rng.integers(1, 200))).astype(np.float64)
        x = np.minimum(1.0, x * rng.integers(1, 4, size=len(x)))
        assert orc.exact_sum(x) == math.fsum(x)
    og = orc.OracleGraph(workload_path("gnm2k", 1, workdir), 40)
    sc = og.fraction_scores()
    for (la, el, lb, na, nb), (sa, sb) in zip(og.edge_stats()[:6], og.edge_fraction()[:6]):
        assert sa > 0 and sb > 0
    col = sc[:, 0][sc[:, 0] > 0]
    assert orc.exact_sum(col) == math.fsum(col)


def write_lg(path, labels, edges):
    with open(path, "w") as f:
        f.write("t # 0\n")
        for v, l in enumerate(labels):
            f.write(f"v {v} {l}\n")
        for a, b, l in edges:
            f.write(f"e {a} {b} {l}\n")


def multi_label_graph(workdir):
    """300 vertices, 3 vertex labels, 3 edge labels, parallel edges between
    some pairs under several labels and a few repeated lines."""
    path = os.path.join(str(workdir), "frac_multilabel.lg")
    if not os.path.exists(path):
        rnd = random.Random(5)
        n = 300
        labels = [rnd.randrange(3) for _ in range(n)]
        edges = []
        for _ in range(1400):
            a, b = rnd.sample(range(n), 2)
            edges.append((a, b, rnd.randrange(3)))
        for a, b, l in edges[:150]:
            edges.append((a, b, (l + 1) % 3))       # the same pair under another label
        edges += edges[:20]                         # repeated lines
        write_lg(path, labels, edges)
    return path


REF_DUMP = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "ref_fraction_dump")


@pytest.mark.skipif(not os.path.exists(REF_DUMP), reason="reference harness not built (make -f oracle/ref.mk)")
def test_scores_match_reference_multilabel(workdir):
    """Several edge labels, parallel and repeated edges: the oracle's scores
    and size-2 sums against the reference loader run on the same file."""
    path = multi_label_graph(workdir)
    dump = os.path.join(str(workdir), "frac_multilabel.dump")
    import subprocess
    subprocess.run([REF_DUMP, path, "1", dump], check=True, stdout=subprocess.DEVNULL, cwd=str(workdir))
    ref_s = [l.rstrip("\n") for l in open(dump) if l.startswith("S ")]
    ref_e = [l.split() for l in open(dump) if l.startswith("E ")]
    og = orc.OracleGraph(path, 1)
    assert score_lines(og.fraction_scores(), oracle_labels(og)) == ref_s
    kept = {}
    for (la, el, lb, na, nb), (sa, sb) in zip(og.edge_stats(), og.edge_fraction()):
        if sa >= 1 and sb >= 1:
            kept[(int(la), int(el), int(lb))] = (sa, sb)
    assert sorted(kept) == sorted((int(t[1]), int(t[2]), int(t[3])) for t in ref_e)
    for t in ref_e:
        sa, sb = kept[(int(t[1]), int(t[2]), int(t[3]))]
        assert abs(sa - float.fromhex(t[6])) <= int(t[4]) * 2.0 ** -52 * sa
        assert abs(sb - float.fromhex(t[7])) <= int(t[5]) * 2.0 ** -52 * sb


@pytest.mark.parametrize("k", [4, 3])
def test_star_closed_form(k, workdir):
    """m disjoint stars, centre label 0 with k leaves of label 1: a leaf's
    score for label 0 is 1/k, a centre's for label 1 is min(1, k) = 1, so the
    edge 0-1 and the path 1-0-1 both have Fraction-Score support m (exact
    for k a power of two), while their MNI is m and m*k-ish."""
    m = 7
    labels, edges = [], []
    for s in range(m):
        c = len(labels)
        labels.append(0)
        for _ in range(k):
            edges.append((c, len(labels), 0))
            labels.append(1)
    path = os.path.join(str(workdir), f"stars_{k}.lg")
    write_lg(path, labels, edges)
    og = orc.OracleGraph(path, 1)
    sc = og.fraction_scores()
    lab = np.array(og.csr()[3])
    assert np.all(sc[lab == 0, 1] == 1.0) and np.all(sc[lab == 0, 0] == 0.0)
    assert np.all(sc[lab == 1, 0] == 1.0 / k) and np.all(sc[lab == 1, 1] == 0.0)
    (sa, sb), = og.edge_fraction()
    assert sa == m
    assert sb == math.fsum([1.0 / k] * (m * k))
    if k == 4:
        assert sb == m
    e = og.eval([1, 0, 1], [(0, 1, 0), (1, 2, 0)], m, mode=2)
    per, mn = e.fraction()
    assert e.frequent == (mn >= m)
    assert per[1] == m
    assert abs(per[0] - m) < 1e-12 and abs(per[2] - m) < 1e-12
    if k == 4:
        assert e.frequent and mn == m
    assert not og.eval([1, 0, 1], [(0, 1, 0), (1, 2, 0)], m + 1, mode=2).frequent


def brute_images(labels, adj, pv, pe):
    """Images of each pattern vertex over all injective label- and
    edge-label-preserving maps."""
    n = len(labels)
    imgs = [set() for _ in pv]
    for m in itertools.permutations(range(n), len(pv)):
        if any(labels[m[i]] != pv[i] for i in range(len(pv))):
            continue
        if all(adj.get((m[a], m[b])) is not None and l in adj[(m[a], m[b])] for a, b, l in pe):
            for i, v in enumerate(m):
                imgs[i].add(v)
    return imgs


def test_verdicts_match_bruteforce(workdir):
    """Mode 2 verdict == (min over u of fsum over u's images of the term) >= freq,
    for every freq, on tiny random graphs."""
    rnd = random.Random(11)
    patterns = [([0, 1, 0], [(0, 1, 0), (1, 2, 0)]), ([0, 0, 1], [(0, 1, 0), (1, 2, 0)]),
                ([0, 1, 1], [(0, 1, 0), (1, 2, 0), (0, 2, 0)]), ([1, 0], [(0, 1, 0)]),
                ([0, 1, 0, 1], [(0, 1, 0), (1, 2, 0), (2, 3, 0)])]
    checked = 0
    for trial in range(12):
        n = 9
        labels = [rnd.randrange(2) for _ in range(n)]
        edges = set()
        while len(edges) < 14:
            a, b = rnd.sample(range(n), 2)
            edges.add((min(a, b), max(a, b)))
        edges = [(a, b, 0) for a, b in sorted(edges)]
        path = os.path.join(str(workdir), f"tiny_frac_{trial}.lg")
        write_lg(path, labels, edges)
        og = orc.OracleGraph(path, 1)
        assert og.n == n                 # every label is frequent at freq 1: ids unchanged
        sc = og.fraction_scores()
        labs = oracle_labels(og)
        adj = {}
        for a, b, l in edges:
            adj.setdefault((a, b), set()).add(l)
            adj.setdefault((b, a), set()).add(l)
        for pv, pe in patterns:
            imgs = brute_images(labels, adj, pv, pe)
            nbr = [[b if a == u else a for a, b, _ in pe if u in (a, b)] for u in range(len(pv))]
            sums = [math.fsum(min(sc[v, labs.index(pv[x])] for x in nbr[u]) for v in imgs[u])
                    for u in range(len(pv))]
            for freq in range(1, 6):
                e = og.eval(pv, pe, freq, mode=2)
                assert e.frequent == (min(sums) >= freq), (trial, pv, pe, freq, sums)
                checked += 1
            full = og.eval(pv, pe, 1, mode=1)
            assert [len(x) for x in imgs] == [int((full.states(u) == 1).sum()) for u in range(len(pv))]
    assert checked == 12 * len(patterns) * 5


FRAC_GOLDEN = ["readme_frac_f1_m3", "gnm2k_frac_f40_m4", "gnm2k_frac_f60_m4", "powerlaw100k_frac_f3000_m5"]


@pytest.mark.parametrize("name", FRAC_GOLDEN)
def test_oracle_fraction_pattern_sets(name, workdir):
    m = golden_manifest()[name]
    og = orc.OracleGraph(workload_path(m["workload"], m["seed"], workdir), m["freq"])
    out = os.path.join(str(workdir), f"oracle_{name}.keys")
    n, _ = og.mine(m["freq"], m["max_nodes"], 8, 2, out)
    keys, sup = set(), {}
    for line in open(out):
        k, lo, hi, fr = line.rstrip("\n").split("\t")
        keys.add(k)
        sup[k] = (int(lo), float.fromhex(fr))
    assert n == m["reported_total"]
    big = {k for k in keys if k.count(";") >= 2}
    assert big == golden_keys(name)
    for k, (lo, fr) in sup.items():
        assert fr >= m["freq"] and lo >= fr, k         # a term is at most 1: support <= MNI count
