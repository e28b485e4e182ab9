"""Whole mining runs on the GPU against the reference's golden pattern sets and
the oracle's exact supports.

Pattern sets must equal the reference binaries' sets exactly (plain and
batched paths, capped and exact support modes).  In exact mode every frequent
pattern's MNI support must equal the oracle's exact-mode support.
"""
import os

import pytest

import oracle as orc
import tfsm
from conftest import golden_keys, golden_manifest, workload_path

pytestmark = pytest.mark.gpu

CASES = ["readme_f1_m3", "readme_f2_m4", "gnm2k_f150_m4", "gnm2k_f80_m4", "gnm2k_f40_m4",
         "powerlaw100k_f9000_m5"]


def read_keys(path):
    out = {}
    for line in open(path):
        k, lo, hi = line.rstrip("\n").split("\t")
        out[k] = (int(lo), int(hi))
    return out


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("batched", [False, True])
def test_pattern_sets(name, batched, workdir):
    m = golden_manifest()[name]
    path = workload_path(m["workload"], m["seed"], workdir)
    g = tfsm.Graph(path, m["freq"])
    ctx = tfsm.Context(g)
    keys_path = os.path.join(str(workdir), f"{name}_{int(batched)}.keys")
    n, stats = ctx.mine(m["freq"], m["max_nodes"], threads=4, batched=batched, mode=tfsm.MODE_CAPPED,
                        keys_path=keys_path)
    mine = set(read_keys(keys_path))
    assert n == m["reported_total"]
    assert mine == golden_keys(name)
    ctx.close()


@pytest.mark.parametrize("name", ["gnm2k_f80_m4", "powerlaw100k_f9000_m5"])
def test_exact_supports(name, workdir):
    m = golden_manifest()[name]
    path = workload_path(m["workload"], m["seed"], workdir)
    g = tfsm.Graph(path, m["freq"])
    ctx = tfsm.Context(g)
    kp = os.path.join(str(workdir), f"{name}_exact.keys")
    ctx.mine(m["freq"], m["max_nodes"], threads=4, batched=True, mode=tfsm.MODE_EXACT, keys_path=kp)
    og = orc.OracleGraph(path, m["freq"])
    op = os.path.join(str(workdir), f"{name}_oracle_exact.keys")
    og.mine(m["freq"], m["max_nodes"], 8, 1, op)
    mine, theirs = read_keys(kp), read_keys(op)
    assert set(mine) == set(theirs) == golden_keys(name)
    for k in mine:
        assert mine[k][0] == mine[k][1], k          # exact mode: lower == upper == MNI
        assert mine[k] == theirs[k], k
    ctx.close()


@pytest.mark.parametrize("tuning", [dict(long_steps=64, long_donate_every=1, long_fanout=2)], ids=["second-phase"])
def test_exact_supports_split(tuning, workdir):
    """Whole batched runs with long searches moved to the wave-cooperative
    phase after 64 steps and its pieces offered at every round: every level's
    patterns share one launch, and the exact supports still equal the
    oracle's."""
    name = "powerlaw100k_f9000_m5"
    m = golden_manifest()[name]
    path = workload_path(m["workload"], m["seed"], workdir)
    g = tfsm.Graph(path, m["freq"])
    ctx = tfsm.Context(g)
    kp = os.path.join(str(workdir), f"{name}_exact_split.keys")
    ctx.mine(m["freq"], m["max_nodes"], threads=4, batched=True, mode=tfsm.MODE_EXACT, keys_path=kp, **tuning)
    og = orc.OracleGraph(path, m["freq"])
    op = os.path.join(str(workdir), f"{name}_oracle_exact_s.keys")
    og.mine(m["freq"], m["max_nodes"], 8, 1, op)
    mine, theirs = read_keys(kp), read_keys(op)
    assert set(mine) == set(theirs) == golden_keys(name)
    for k in mine:
        assert mine[k] == theirs[k], k
    ctx.close()


@pytest.mark.parametrize("name", ["rmat20_f20000_m5", "rmat20_f9000_m5"])
def test_rmat20_pattern_sets(name, workdir):
    """R-MAT scale 20 (1M vertices, 16.8M edges, 16 vertex and 4 edge labels):
    plain and batched runs on one loaded graph against the reference's sets.
    At freq 9000 seven single edges are frequent and every size-3 candidate
    is evaluated on the GPU and rejected; at freq 20000 nothing is frequent."""
    m = golden_manifest()[name]
    path = workload_path(m["workload"], m["seed"], workdir)
    print(f"{name}: loading", flush=True)
    g = tfsm.Graph(path, m["freq"])
    ctx = tfsm.Context(g)
    for batched in (False, True):
        print(f"{name}: mining batched={batched}", flush=True)
        kp = os.path.join(str(workdir), f"{name}_{int(batched)}.keys")
        n, stats = ctx.mine(m["freq"], m["max_nodes"], threads=4, batched=batched, mode=tfsm.MODE_CAPPED,
                            keys_path=kp)
        assert n == m["reported_total"]
        assert set(read_keys(kp)) == golden_keys(name)
        print(f"{name}: batched={batched} evaluated={stats['evaluated']} mining_s={stats['mining_s']:.3f}",
              flush=True)
    ctx.close()
    g.close()


@pytest.mark.parametrize("learned", [False, True], ids=["batched", "learned"])
def test_powerlaw_f6000_m6_pattern_sets(learned, workdir):
    """The batched configuration: powerlaw100k at freq 6000, maxNodes 6 (631
    candidate patterns over five levels, 334 frequent).  Batched runs, with
    and without the learned order, find the reference's pattern set.  The
    golden file holds the patterns of three or more vertices; the run also
    reports the four frequent single edges, as the reference's total does."""
    name = "powerlaw100k_f6000_m6"
    m = golden_manifest()[name]
    path = workload_path(m["workload"], m["seed"], workdir)
    g = tfsm.Graph(path, m["freq"])
    ctx = tfsm.Context(g)
    if learned:
        ctx.ranker_load(os.path.join(os.path.dirname(tfsm.__file__), "rankers", "powerlaw.f32"))
    kp = os.path.join(str(workdir), f"{name}_{int(learned)}.keys")
    n, stats = ctx.mine(m["freq"], m["max_nodes"], threads=4, batched=True, learned=learned, mode=tfsm.MODE_CAPPED,
                        keys_path=kp)
    assert n == m["reported_total"]
    keys = set(read_keys(kp))
    assert set(k for k in keys if k.count(";") > 1) == golden_keys(name)
    assert len(keys) - len(golden_keys(name)) == 4
    ctx.close()


def test_rmat20_f5300_through_csr_cache(workdir):
    """R-MAT scale 20 seed 4 at freq 5300 (45,089 candidates over two levels,
    658 frequent) on the HBM-resident graph: the first load parses the file
    and writes the binary CSR cache, the second reads it; batched, then
    batched with the R-MAT ranker (trained on seeds 11 and 12), each find
    the reference's pattern set (parallel build; the golden file holds the
    patterns of three or more vertices)."""
    name = "rmat20_f5300_m5_parallel"
    m = golden_manifest()[name]
    path = workload_path(m["workload"], m["seed"], workdir)
    for learned in (False, True):
        print(f"{name}: loading (cache), learned={learned}", flush=True)
        g = tfsm.Graph(path, m["freq"], cache=True)
        if learned:
            assert g.cache_hit
        ctx = tfsm.Context(g)
        if learned:
            ctx.ranker_load(os.path.join(os.path.dirname(tfsm.__file__), "rankers", "rmat.f32"))
        kp = os.path.join(str(workdir), f"{name}_{int(learned)}.keys")
        n, stats = ctx.mine(m["freq"], m["max_nodes"], threads=4, batched=True, learned=learned,
                            mode=tfsm.MODE_CAPPED, keys_path=kp)
        print(f"{name}: learned={learned} evaluated={stats['evaluated']} mining_s={stats['mining_s']:.3f}",
              flush=True)
        assert n == m["reported_total"]
        keys = set(read_keys(kp))
        assert set(k for k in keys if k.count(";") > 1) == golden_keys(name)
        ctx.close()
        g.close()
    assert os.path.exists(path + f".f{m['freq']}.csr")


def test_rmat20_seed5_learned_equals_oracle(workdir):
    """R-MAT scale 20 seed 5 at freq 5300 (99,339 candidates, 1,365
    frequent), the graph the R-MAT ranker is measured on and was not trained
    on: batched with the learned order finds every frequent pattern this
    build's CPU oracle found on the same graph."""
    name = "rmat20s5_f5300_m5_oracle"
    m = golden_manifest()[name]
    path = workload_path(m["workload"], m["seed"], workdir)
    print(f"{name}: loading", flush=True)
    g = tfsm.Graph(path, m["freq"], cache=True)
    ctx = tfsm.Context(g)
    ctx.ranker_load(os.path.join(os.path.dirname(tfsm.__file__), "rankers", "rmat.f32"))
    kp = os.path.join(str(workdir), f"{name}.keys")
    n, stats = ctx.mine(m["freq"], m["max_nodes"], threads=4, batched=True, learned=True,
                        mode=tfsm.MODE_CAPPED, keys_path=kp)
    print(f"{name}: evaluated={stats['evaluated']} mining_s={stats['mining_s']:.3f}", flush=True)
    assert n == m["reported_total"]
    assert set(read_keys(kp)) == golden_keys(name)
    ctx.close()
    g.close()
