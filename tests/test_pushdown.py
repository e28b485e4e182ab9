"""Push-down pruning on the host: canonical forms and the decomposition.

* canonical_form: every vertex renumbering of a small labelled graph gets the
  same key, and the returned orders map the two graphs onto each other
  (labels and labelled edges agree position by position); graphs that are not
  isomorphic (checked by brute force over all bijections) get different keys;
* decompose (the reference's parallel/decompose.h:19-94): per pattern edge,
  the connected part without that edge that holds the last vertex, when it
  has more than one vertex, recomputed here from the definition.
"""
import ctypes
import itertools
import random

import numpy as np
import pytest

import tfsm


def canon(n, lab, el):
    lib = tfsm.host_lib()
    key = ctypes.create_string_buffer(4096)
    order = np.zeros(n, np.int32)
    L = np.asarray(lab, np.int32)
    E = np.asarray(el, np.int32).reshape(-1)
    k = lib.tfsmh_canonical(n, L.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                            E.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), key, 4096,
                            order.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert k > 0
    return key.value.decode(), list(order)


def random_graph(rng, n, nl, nel, p):
    lab = [rng.randrange(nl) for _ in range(n)]
    el = [[-1] * n for _ in range(n)]
    for a in range(n):
        for b in range(a + 1, n):
            if rng.random() < p:
                el[a][b] = el[b][a] = rng.randrange(nel)
    return lab, el


def permuted(lab, el, perm):
    n = len(lab)
    lab2 = [0] * n
    el2 = [[-1] * n for _ in range(n)]
    for a in range(n):
        lab2[perm[a]] = lab[a]
        for b in range(n):
            el2[perm[a]][perm[b]] = el[a][b]
    return lab2, el2


def isomorphic(l1, e1, l2, e2):
    n = len(l1)
    for pm in itertools.permutations(range(n)):
        if all(l1[a] == l2[pm[a]] for a in range(n)) and all(e1[a][b] == e2[pm[a]][pm[b]]
                                                            for a in range(n) for b in range(n)):
            return True
    return False


def test_canonical_form_is_invariant_and_maps():
    rng = random.Random(5)
    for _ in range(300):
        n = rng.randrange(2, 7)
        lab, el = random_graph(rng, n, rng.randrange(1, 3), rng.randrange(1, 3), 0.5)
        k1, o1 = canon(n, lab, el)
        perm = list(range(n))
        rng.shuffle(perm)
        lab2, el2 = permuted(lab, el, perm)
        k2, o2 = canon(n, lab2, el2)
        assert k1 == k2
        # position k: vertex o1[k] of the first graph, o2[k] of the second
        for a in range(n):
            assert lab[o1[a]] == lab2[o2[a]]
            for b in range(n):
                assert el[o1[a]][o1[b]] == el2[o2[a]][o2[b]]


def test_canonical_form_separates():
    rng = random.Random(6)
    seen = 0
    for _ in range(400):
        n = rng.randrange(2, 6)
        l1, e1 = random_graph(rng, n, 2, 2, 0.5)
        l2, e2 = random_graph(rng, n, 2, 2, 0.5)
        same = isomorphic(l1, e1, l2, e2)
        assert (canon(n, l1, e1)[0] == canon(n, l2, e2)[0]) == same
        seen += not same
    assert seen > 100


def decompose(code):
    lib = tfsm.host_lib()
    c = np.asarray([x for e in code for x in e], np.int32)
    out = np.zeros(4096, np.int32)
    k = lib.tfsmh_decompose(c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(code),
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 4096)
    assert k >= 0
    res, i = [], 0
    for _ in range(k):
        m = int(out[i])
        res.append(sorted(int(x) for x in out[i + 1:i + 1 + m]))
        i += 1 + m
    return res


def test_decompose_follows_the_definition(workdir):
    from conftest import golden_keys
    checked = 0
    for key in sorted(golden_keys("gnm2k_f40_m4")):
        vl, es = tfsm.parse_key(key)
        code = [(a, b, vl[a], el, vl[b]) for a, b, el in es]
        n = len(vl)
        want = []
        for e in range(len(es)):
            adj = {v: set() for v in range(n)}
            for f, (a, b, _) in enumerate(es):
                if f != e:
                    adj[a].add(b)
                    adj[b].add(a)
            comp, stack = {n - 1}, [n - 1]
            while stack:
                x = stack.pop()
                for y in adj[x] - comp:
                    comp.add(y)
                    stack.append(y)
            if len(comp) > 1:
                want.append(sorted(comp))
        assert decompose(code) == want, key
        checked += 1
    assert checked > 50
