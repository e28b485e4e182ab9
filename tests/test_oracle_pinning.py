"""Pins the oracle's evaluation against what the mathematics fixes.

* exact MNI == brute-force enumeration of every injective, label- and
  edge-label-preserving map, on small seeded random graphs and patterns;
* capped verdicts agree with the brute-force MNI;
* filtered domains are sound (they contain every brute-force image) and
  arc-consistent (every candidate has a neighbour in every neighbouring
  domain over the right edge label);
* auxiliary tables equal a direct recomputation from their definition;
* closed forms: a single-label clique K_n gives MNI n for every connected
  pattern of at most n vertices; a cycle C_n gives MNI n for paths; a star
  with k leaves in a star graph has centre image 1;
* the size-2 statistics equal min(|A|, |B|) over endpoint sets computed here.
A dropped term, a wrong index or a transposed table fails one of these.
"""
import itertools
import os
import random

import numpy as np
import pytest

import oracle as orc


def write_lg(path, labels, edges):
    with open(path, "w") as f:
        f.write(f"t {len(labels)} {len(edges)}\n")
        deg = [0] * len(labels)
        for a, b, _ in edges:
            deg[a] += 1
            deg[b] += 1
        for i, l in enumerate(labels):
            f.write(f"v {i} {l} {deg[i]}\n")
        for a, b, l in edges:
            f.write(f"e {a} {b} {l}\n")


def random_graph(rng, n, m, nl, nel):
    labels = [rng.randrange(nl) for _ in range(n)]
    es = set()
    while len(es) < m:
        a, b = rng.randrange(n), rng.randrange(n)
        if a != b:
            es.add((min(a, b), max(a, b)))
    return labels, [(a, b, 1 + rng.randrange(nel)) for a, b in sorted(es)]


def random_pattern(rng, k, nl, nel, extra):
    vl = [rng.randrange(nl) for _ in range(k)]
    es = [(rng.randrange(i), i, 1 + rng.randrange(nel)) for i in range(1, k)]
    have = {(a, b) for a, b, _ in es}
    for _ in range(extra):
        a, b = sorted(rng.sample(range(k), 2))
        if (a, b) not in have:
            have.add((a, b))
            es.append((a, b, 1 + rng.randrange(nel)))
    return vl, es


def pruned(og):
    off, nbr, el, vl = og.csr()
    adj = {}
    for v in range(og.n):
        for k in range(off[v], off[v + 1]):
            adj[(v, int(nbr[k]))] = int(el[k])
    return off, nbr, el, vl, adj


@pytest.mark.parametrize("seed", range(12))
def test_exact_mni_equals_bruteforce(seed, tmp_path):
    rng = random.Random(seed)
    labels, edges = random_graph(rng, 11, 24, 2, 1 + seed % 2)
    path = str(tmp_path / "g.lg")
    write_lg(path, labels, edges)
    og = orc.OracleGraph(path, 1)
    off, nbr, el, vl, adj = pruned(og)
    for t in range(6):
        k = 2 + (t % 3)
        pv, pe = random_pattern(rng, k, 2, 1 + seed % 2, t % 2)
        mni, images = og.bruteforce(pv, pe)
        for freq in (1, 2, max(1, mni), mni + 1):
            ev = og.eval(pv, pe, freq, mode=1)
            if ev.decided_by != 0:
                assert ev.lower == ev.upper == mni, (pv, pe, freq)
            else:
                assert mni < freq
            assert ev.frequent == (mni >= freq)
            cap = og.eval(pv, pe, freq, mode=0)
            assert cap.frequent == (mni >= freq), (pv, pe, freq)
            # soundness: every brute-force image survives the filter
            for u in range(k):
                dom = set(int(x) for x in ev.domain(u))
                if ev.decided_by != 0 or dom:
                    assert len(dom) >= images[u]
            # arc consistency of the filtered domains
            doms = [set(int(x) for x in ev.domain(u)) for u in range(k)]
            if all(doms):
                for a, b, l in pe:
                    for (x, y) in ((a, b), (b, a)):
                        for v in doms[x]:
                            assert any(adj.get((v, w)) == l for w in doms[y]), (pv, pe, x, v)


@pytest.mark.parametrize("seed", range(6))
def test_aux_table_definition(seed, tmp_path):
    rng = random.Random(100 + seed)
    labels, edges = random_graph(rng, 14, 40, 2, 2)
    path = str(tmp_path / "g.lg")
    write_lg(path, labels, edges)
    og = orc.OracleGraph(path, 1)
    _, _, _, _, adj = pruned(og)
    for t in range(5):
        pv, pe = random_pattern(rng, 3 + t % 2, 2, 2, 1)
        ev = og.eval(pv, pe, 1, mode=1)
        if ev.decided_by != 2:
            continue
        doms = [list(ev.domain(u)) for u in range(len(pv))]
        elab = {(a, b): l for a, b, l in pe}
        elab.update({(b, a): l for a, b, l in pe})
        for a in range(len(pv)):
            for b in range(len(pv)):
                t_ = ev.aux(a, b)
                if (a, b) not in elab:
                    assert t_ is None
                    continue
                off, idx = t_
                for j, v in enumerate(doms[a]):
                    row = [i for i, w in enumerate(doms[b]) if adj.get((int(v), int(w))) == elab[(a, b)]]
                    assert list(idx[off[j]:off[j + 1]]) == row, (a, b, j)


def test_closed_forms(tmp_path):
    # single-label clique K_6: every connected pattern of <= 6 vertices has MNI 6
    n = 6
    path = str(tmp_path / "k6.lg")
    write_lg(path, [0] * n, [(a, b, 1) for a, b in itertools.combinations(range(n), 2)])
    og = orc.OracleGraph(path, 1)
    for k in range(2, n + 1):
        path_p = [(i, i + 1, 1) for i in range(k - 1)]
        clique_p = [(a, b, 1) for a, b in itertools.combinations(range(k), 2)]
        for pe in (path_p, clique_p):
            ev = og.eval([0] * k, pe, 1, mode=1)
            assert ev.lower == ev.upper == n
    # cycle C_8 with one label: paths of 2..8 vertices have MNI 8, the 8-cycle too
    path = str(tmp_path / "c8.lg")
    write_lg(path, [0] * 8, [(i, (i + 1) % 8, 1) for i in range(8)])
    og = orc.OracleGraph(path, 1)
    for k in range(2, 9):
        ev = og.eval([0] * k, [(i, i + 1, 1) for i in range(k - 1)], 1, mode=1)
        assert ev.lower == 8
    ev = og.eval([0] * 8, [(i, (i + 1) % 8, 1) for i in range(8)], 1, mode=1)
    assert ev.lower == 8
    ev = og.eval([0] * 3, [(0, 1, 1), (1, 2, 1), (0, 2, 1)], 1, mode=1)   # no triangle in C_8
    assert ev.upper == 0 and not ev.frequent
    # star with 5 leaves (centre label 1, leaves label 0); pattern star with 3 leaves
    path = str(tmp_path / "star.lg")
    write_lg(path, [1] + [0] * 5, [(0, i, 1) for i in range(1, 6)])
    og = orc.OracleGraph(path, 1)
    ev = og.eval([1, 0, 0, 0], [(0, 1, 1), (0, 2, 1), (0, 3, 1)], 1, mode=1)
    assert ev.states(0).tolist() == [1] and ev.lower == 1
    assert len(ev.domain(1)) == 5 and (ev.states(1) == 1).all()
    ev = og.eval([1] + [0] * 6, [(0, i, 1) for i in range(1, 7)], 1, mode=1)  # needs 6 leaves
    assert not ev.frequent and ev.upper == 0


@pytest.mark.parametrize("seed", range(4))
def test_size2_statistics(seed, tmp_path):
    rng = random.Random(200 + seed)
    labels, edges = random_graph(rng, 30, 80, 3, 2)
    path = str(tmp_path / "g.lg")
    write_lg(path, labels, edges)
    for freq in (1, 3, 6):
        og = orc.OracleGraph(path, freq)
        keep = {l for l in set(labels) if labels.count(l) >= freq}
        ekeep = {l for l in {e[2] for e in edges} if sum(1 for e in edges if e[2] == l) >= freq}
        # renumbering of kept vertices in id order; dropped vertices map to 0
        newid, k = {}, 0
        for v, l in enumerate(labels):
            if l in keep:
                newid[v] = k
                k += 1
        sets = {}
        for a, b, l in edges:
            la, lb = labels[a], labels[b]
            ia, ib = newid.get(a, 0), newid.get(b, 0)
            if la < lb:
                s = sets.setdefault((la, l, lb), (set(), set()))
                s[0].add(ia); s[1].add(ib)
            elif la == lb:
                s = sets.setdefault((la, l, lb), (set(), set()))
                s[0].update((ia, ib)); s[1].update((ia, ib))
            else:
                s = sets.setdefault((lb, l, la), (set(), set()))
                s[0].add(ib); s[1].add(ia)
        want = sorted((k_[0], k_[1], k_[2], len(A), len(B)) for k_, (A, B) in sets.items()
                      if len(A) >= freq and len(B) >= freq)
        got = sorted(tuple(int(x) for x in r) for r in og.edge_stats())
        assert got == want
        assert og.n == sum(1 for l in labels if l in keep)
        assert og.num_adj == 2 * sum(1 for a, b, l in edges if labels[a] in keep and labels[b] in keep and l in ekeep)


# ---------------------------------------------------------------------------
# The reference's own matching engine (tests/golden/gmatch_*.txt, written by
# tests/golden/make_gmatch_fixtures.py from parallel/gmatch.h compiled here):
# filtered domains (DPisoFilter, gmatch.h:417-472), auxiliary tables
# (buildTable, gmatch.h:995-1130), GQL orders and backward-neighbour lists
# (gmatch.h:692-727, 803-828) must equal the oracle's.
# ---------------------------------------------------------------------------
def read_gmatch_fixture(path):
    pats, cur = [], None
    for line in open(path):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "P":
            cur = dict(key=t[1], ok=None, dom={}, tab={}, order={}, bn={})
            pats.append(cur)
        elif t[0] == "F":
            cur["ok"] = t[1] == "1"
        elif t[0] == "D":
            u, n = int(t[1]), int(t[2])
            cur["dom"][u] = [int(x) for x in t[3:3 + n]]
        elif t[0] == "T":
            a, b, rows, ents = (int(x) for x in t[1:5])
            vals = [int(x) for x in t[5:]]
            cur["tab"][(a, b)] = (vals[:rows + 1], vals[rows + 1:rows + 1 + ents])
        elif t[0] == "O":
            cur["order"][int(t[1])] = [int(x) for x in t[2:]]
        elif t[0] == "B":
            cur["bn"][(int(t[1]), int(t[2]))] = [int(x) for x in t[3:]]
    return pats


def parse_key(key):
    vl, es = {}, []
    for code in key.strip(";").split(";"):
        a, b, la, el, lb = (int(x) for x in code.split(","))
        vl[a], vl[b] = la, lb
        es.append((a, b, el))
    return [vl[i] for i in range(len(vl))], es


@pytest.mark.parametrize("name,workload,seed,freq", [("readme_f1", "readme", 0, 1), ("gnm2k_f40", "gnm2k", 1, 40)])
def test_oracle_matches_reference_matcher(name, workload, seed, freq, workdir):
    from conftest import GOLDEN, workload_path
    pats = read_gmatch_fixture(os.path.join(GOLDEN, f"gmatch_{name}.txt"))
    assert len(pats) >= 15
    og = orc.OracleGraph(workload_path(workload, seed, workdir), freq)
    searched = tabled = 0
    for p in pats:
        vl, es = parse_key(p["key"])
        # support 1, as the fixture's DPisoFilter call: every pattern with
        # non-empty domains goes on to tables and plans
        r = og.eval(vl, es, 1, mode=1)
        if not p["ok"]:
            assert r.upper == 0, p["key"]          # some domain is empty
            continue
        for u, dom in p["dom"].items():
            assert list(r.domain(u)) == dom, f"{p['key']} domain {u}"
        if r.decided_by != 2:
            continue                               # express path: no tables or plans
        searched += 1
        for (a, b), (off, idx) in p["tab"].items():
            mine = r.aux(a, b)
            assert mine is not None, f"{p['key']} table {a}->{b} missing"
            assert list(mine[0]) == off and list(mine[1]) == idx, f"{p['key']} table {a}->{b}"
            tabled += 1
        for u, order in p["order"].items():
            assert [int(x) for x in r.order(u)] == order, f"{p['key']} order {u}"
        for (u, d), bn in p["bn"].items():
            assert r.bn(u, d) == bn, f"{p['key']} bn {u} {d}"
    assert searched >= 5 and tabled >= 10, (searched, tabled)
