"""Host side of the product on the CPU: loader, generators and the
minimum-DFS-code test behind pattern growth."""
import itertools
import random

import numpy as np
import pytest

import oracle as orc
import tfsm
from conftest import workload_path


@pytest.mark.parametrize("name,seed,freq", [("readme", 0, 1), ("readme", 0, 2), ("gnm2k", 1, 40),
                                            ("gnm2k", 1, 150), ("powerlaw100k", 2, 9000)])
def test_loader_matches_oracle(name, seed, freq, workdir):
    path = workload_path(name, seed, workdir)
    g = tfsm.Graph(path, freq)
    og = orc.OracleGraph(path, freq)
    c = g.csr()
    off, nbr, el, vl = og.csr()
    assert g.n == og.n and g.num_adj == og.num_adj
    assert np.array_equal(c["offsets"], off)
    assert np.array_equal(c["nbrs"], nbr)
    assert np.array_equal(c["vlabel_values"][c["vlabels"]], vl)
    assert np.array_equal(c["elabel_values"][c["elabels"]], el)
    assert np.array_equal(g.edge_stats(), og.edge_stats())


def test_generators_are_seeded_and_well_formed(workdir):
    a = workload_path("gnm2k", 1, workdir)
    with open(a) as f:
        lines = f.read().split("\n")
    assert lines[0] == "t 2000 8000"
    edges = [tuple(map(int, l.split()[1:3])) for l in lines if l.startswith("e ")]
    assert len(edges) == 8000 and len(set(edges)) == 8000 and all(u < v for u, v in edges)
    import subprocess, os
    b = os.path.join(str(workdir), "gnm2k_again.lg")
    subprocess.run([os.path.join(tfsm.BUILD, "tfsm_gen"), "gnm2k", b, "1"], check=True, stdout=subprocess.DEVNULL)
    assert open(a).read() == open(b).read()
    p = workload_path("powerlaw100k", 2, workdir)
    deg = np.zeros(100000, np.int64)
    lab = np.zeros(100000, np.int64)
    ne = 0
    with open(p) as f:
        for line in f:
            if line[0] == "v":
                _, i, l, d = line.split()
                lab[int(i)] = int(l)
                deg[int(i)] = int(d)
            elif line[0] == "e":
                ne += 1
    assert ne == 1100000 and deg.max() <= 1500 and abs(deg.mean() - 22.0) < 1e-9
    counts = np.bincount(lab, minlength=29)
    zipf = 1.0 / np.arange(1, 30)
    zipf /= zipf.sum()
    assert np.abs(counts / 100000 - zipf).max() < 0.01


# ---- minimum DFS code: brute force over every DFS traversal (gSpan order) ----

def dfs_codes(vl, adj):
    """Every DFS code of a connected labelled graph: forward edge to a newly
    discovered vertex, then its backward edges to earlier vertices (ascending),
    for every root and every order of visiting neighbours."""
    n = len(vl)
    out = []

    def rec(order, pos, stack, code):
        if len(order) == n:
            out.append(list(code))
            return
        # extend from the deepest stack vertex that has an unvisited neighbour
        for depth in range(len(stack) - 1, -1, -1):
            x = stack[depth]
            nbrs = [y for y in adj[x] if y not in pos]
            if nbrs:
                for y in nbrs:
                    new_pos = dict(pos)
                    new_pos[y] = len(order)
                    c = code + [(pos[x], new_pos[y], vl[x], adj[x][y], vl[y])]
                    for z in sorted((z for z in adj[y] if z in pos and z != x), key=lambda z: pos[z]):
                        c.append((new_pos[y], pos[z], vl[y], adj[y][z], vl[z]))
                    rec(order + [y], new_pos, stack[:depth + 1] + [y], c)
                return

    for r in range(n):
        rec([r], {r: 0}, [r], [])
    return out


def edge_key(e):
    f, t, fl, el, tl = e
    return (0, t, el) if f > t else (1, -f, fl, el, tl)


def code_key(code):
    return [edge_key(e) for e in code]


def random_connected(rng, k, nl, nel, extra):
    vl = [rng.randrange(nl) for _ in range(k)]
    adj = {i: {} for i in range(k)}
    for i in range(1, k):
        j = rng.randrange(i)
        l = 1 + rng.randrange(nel)
        adj[i][j] = adj[j][i] = l
    for _ in range(extra):
        a, b = rng.sample(range(k), 2)
        if b not in adj[a]:
            l = 1 + rng.randrange(nel)
            adj[a][b] = adj[b][a] = l
    return vl, adj


def is_min(code):
    flat = (tfsm.ctypes.c_int32 * (5 * len(code)))(*[x for e in code for x in e])
    r = tfsm.host_lib().tfsmh_is_min_code(flat, len(code))
    assert r >= 0
    return bool(r)


@pytest.mark.parametrize("seed", range(25))
def test_min_dfs_code_bruteforce(seed):
    rng = random.Random(seed)
    k = 3 + seed % 3
    vl, adj = random_connected(rng, k, 2 + seed % 2, 1 + (seed // 10) % 2, seed % 4)
    codes = dfs_codes(vl, adj)
    best = min(codes, key=code_key)
    assert is_min(best)
    uniq = {tuple(c) for c in codes}
    for c in uniq:
        assert is_min(list(c)) == (list(c) == best), c
