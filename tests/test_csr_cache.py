"""Binary cache of the pruned data graph: a cached load gives the same graph as
a fresh parse; a stale, truncated or other-threshold cache is not used."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import tfsm

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(REPO, "t-fsm_amd", "_build", "tfsm_gen")


def same_graph(a, b):
    ca, cb = a.csr(), b.csr()
    for k in ca:
        assert np.array_equal(ca[k], cb[k]), k
    assert np.array_equal(a.edge_stats(), b.edge_stats())
    assert (a.n, a.num_adj, a.num_vlabels, a.num_elabels) == (b.n, b.num_adj, b.num_vlabels, b.num_elabels)


@pytest.fixture
def lg(tmp_path):
    path = str(tmp_path / "gnm2k.lg")
    subprocess.run([GEN, "gnm2k", path, "3"], check=True, stdout=subprocess.DEVNULL)
    return path


def test_cache_round_trip(lg):
    fresh = tfsm.Graph(lg, 150)
    first = tfsm.Graph(lg, 150, cache=True)
    assert not first.cache_hit and os.path.exists(lg + ".f150.csr")
    second = tfsm.Graph(lg, 150, cache=True)
    assert second.cache_hit
    same_graph(fresh, first)
    same_graph(fresh, second)


def test_cache_is_per_threshold(lg):
    tfsm.Graph(lg, 150, cache=True)
    other = tfsm.Graph(lg, 260, cache=True)
    assert not other.cache_hit and os.path.exists(lg + ".f260.csr")
    same_graph(tfsm.Graph(lg, 260), other)


def test_stale_cache_is_rebuilt(lg):
    tfsm.Graph(lg, 150, cache=True)
    # a different graph under the same name: the cache's source hash no longer matches
    subprocess.run([GEN, "gnm2k", lg, "4"], check=True, stdout=subprocess.DEVNULL)
    g = tfsm.Graph(lg, 150, cache=True)
    assert not g.cache_hit
    same_graph(tfsm.Graph(lg, 150), g)
    assert tfsm.Graph(lg, 150, cache=True).cache_hit


def test_truncated_cache_is_ignored(lg):
    tfsm.Graph(lg, 150, cache=True)
    cache = lg + ".f150.csr"
    size = os.path.getsize(cache)
    with open(cache, "r+b") as f:
        f.truncate(size - 4)
    g = tfsm.Graph(lg, 150, cache=True)
    assert not g.cache_hit
    same_graph(tfsm.Graph(lg, 150), g)
    assert os.path.getsize(cache) == size


def test_cli_uses_the_cache(lg, tmp_path):
    """The CLI's -csrCache flag writes the cache on the first run and reads it
    on the second; both runs find the same patterns (no GPU needed to load:
    the run stops at context creation here when no GPU is present)."""
    run = os.path.join(REPO, "t-fsm_amd", "_build", "run")
    for _ in range(2):
        subprocess.run([run, "-file", lg, "-freq", "150", "-maxNodes", "3", "-csrCache", "1",
                        "-out", str(tmp_path / "out.txt")], stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    assert os.path.exists(lg + ".f150.csr")
    same_graph(tfsm.Graph(lg, 150), tfsm.Graph(lg, 150, cache=True))
