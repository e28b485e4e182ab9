"""The learned task ranker on the GPU, through the C ABI.

Forward logits and gradients of the BF16-weight / FP32-accumulation
perceptron against the FP64 reference (oracle/mlp_ref.py) at the tolerances
written below; training on a separable problem; the feature kernel against
host-computed values; and the learned task order's results: pattern sets equal
to the reference's golden sets and exact supports equal to the oracle's, with
any weights.  The gain test runs the shipped ranker (trained on other seeds)
on the 100K power-law graph (seed 2, freq 9000, 5 vertices).
"""
import os

import numpy as np
import pytest

import mlp_ref
import oracle as orc
import tfsm
from conftest import PKG, golden_keys, golden_manifest, workload_path

pytestmark = pytest.mark.gpu

RANKER = os.path.join(PKG, "rankers", "powerlaw.f32")


def random_weights(seed, scale=0.3):
    return np.random.default_rng(seed).normal(0, scale, tfsm.RANKER_PARAMS).astype(np.float32)


def feature_like(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.3, 0.6, size=(n, tfsm.RANKER_FEATURES)).astype(np.float32)
    X[:, 23] = 1.0
    return X


@pytest.fixture(scope="module")
def ctx(workdir):
    g = tfsm.Graph(workload_path("gnm2k", 1, workdir), 80)
    c = tfsm.Context(g)
    yield c
    c.close()
    g.close()


def test_set_get_rounds_to_bf16(ctx):
    w = random_weights(3)
    ctx.ranker_set(w)
    assert np.array_equal(ctx.ranker_get().view(np.uint32), mlp_ref.bf16_round(w).view(np.uint32))
    bad = w.copy()
    bad[5] = np.nan
    with pytest.raises(tfsm.TfsmError):
        ctx.ranker_set(bad)


def test_forward_matches_reference(ctx):
    w = random_weights(4)
    ctx.ranker_set(w)
    X = feature_like(5000, 5)
    z = ctx.ranker_forward(X)
    ref = mlp_ref.forward(mlp_ref.bf16_round(w), X)
    # FP32 accumulation of 24 + 64 products: |error| <= 1e-4 (1 + |z|)
    assert np.all(np.abs(z - ref) <= 1e-4 * (1 + np.abs(ref)))


def test_gradient_matches_reference(ctx):
    w = random_weights(6)
    ctx.ranker_set(w)
    X = feature_like(3000, 7)
    y = (np.random.default_rng(8).random(3000) < 0.3).astype(np.float32)
    g, loss = ctx.ranker_gradient(X, y)
    ref_loss, ref_g = mlp_ref.loss_and_grad(mlp_ref.bf16_round(w), X, y)
    assert loss == pytest.approx(ref_loss, rel=1e-4)
    # FP32 sums of 3000 rows in atomic order: within 1e-4 of the gradient's scale
    tol = 1e-4 * np.abs(ref_g).max()
    assert np.all(np.abs(g - ref_g) <= tol + 1e-4 * np.abs(ref_g))


def test_ranker_call_after_retaining_level(workdir):
    """A level call returns with its retained-domain copies still queued on the
    stream and reading the scratch arena; a ranker call made at once reuses
    that arena.  The retained domains must still equal the oracle's child
    domains (C(u) minus proven-invalid candidates)."""
    g = tfsm.Graph(workload_path("gnm2k", 1, workdir), 40)
    c = tfsm.Context(g)
    og = orc.OracleGraph(workload_path("gnm2k", 1, workdir), 40)
    csr = g.csr()
    vv = [int(x) for x in csr["vlabel_values"]]
    ev = [int(x) for x in csr["elabel_values"]]
    cases = []
    for key in sorted(golden_keys("gnm2k_f40_m4")):
        vl, es = tfsm.parse_key(key)
        if len(vl) >= 3:
            cases.append((vl, es))
        if len(cases) >= 40:
            break
    pats = [tfsm.make_pattern([vv.index(l) for l in vl], [(a, b, ev.index(l)) for a, b, l in es])
            for vl, es in cases]
    c.ranker_set(random_weights(21))
    X = feature_like(200000, 22)
    for _ in range(2):
        res = c.eval_level(pats, 40, mode=tfsm.MODE_EXACT, retain=True)
        c.ranker_forward(X)                     # reuses the arena at once
        checked = 0
        for (vl, es), r in zip(cases, res):
            if not r.domain_handle:
                continue
            ref = og.eval(vl, es, 40, mode=1)
            for u in range(len(vl)):
                assert np.array_equal(c.domain_read(r.domain_handle, u), ref.child_domain(u)), f"{vl} {es} {u}"
            c.domain_release(r.domain_handle)
            checked += 1
        assert checked > 10
    c.close()
    g.close()


def test_training_reduces_loss(ctx):
    rng = np.random.default_rng(9)
    X = feature_like(20000, 10)
    y = ((X[:, :4] @ np.array([1.0, -1.0, 0.5, 2.0])) > 0.6).astype(np.float32)
    w0 = random_weights(11, 0.1)
    w0[-1] = 0.0
    ctx.ranker_set(w0)
    _, loss0 = ctx.ranker_gradient(X, y)
    ctx.ranker_train(X, y, epochs=4, batch=512, lr=3e-3, seed=12)
    _, loss1 = ctx.ranker_gradient(X, y)
    acc = ((ctx.ranker_forward(X) > 0) == (y > 0.5)).mean()
    assert loss1 < 0.5 * loss0 and acc > 0.93, (loss0, loss1, acc)
    # the trained weights stay BF16-exact
    w = ctx.ranker_get()
    assert np.array_equal(w.view(np.uint32), mlp_ref.bf16_round(w).view(np.uint32))
    del rng


def test_features_match_host(workdir):
    path = workload_path("gnm2k", 1, workdir)
    g = tfsm.Graph(path, 80)
    c = tfsm.Context(g)
    csr = g.csr()
    deg = np.diff(csr["offsets"].astype(np.int64))
    vv = [int(x) for x in csr["vlabel_values"]]
    ev = [int(x) for x in csr["elabel_values"]]
    searched = 0
    for key in sorted(golden_keys("gnm2k_f80_m4")):
        vl, es = tfsm.parse_key(key)
        if len(vl) < 3:
            continue
        pat = tfsm.make_pattern([vv.index(l) for l in vl], [(a, b, ev.index(l)) for a, b, l in es])
        res = c.eval_pattern(pat, 80, mode=tfsm.MODE_EXACT, trace=True)
        if res.decided_by != 2:
            continue
        X = c.read_features()
        doms = [c.last_domain(u) for u in range(len(vl))]
        assert len(X) == sum(len(d) for d in doms)
        rows = 0
        for u, d in enumerate(doms):
            x = X[rows:rows + len(d)]
            rows += len(d)
            lf = np.log2(81.0)
            np.testing.assert_allclose(x[:, 0], np.log2(1 + deg[d]) * 0.1, atol=1e-5)
            np.testing.assert_allclose(x[:, 1], np.log2(1 + len(d)) / 16, atol=1e-5)
            np.testing.assert_allclose(x[:, 2], lf / 16, atol=1e-5)
            np.testing.assert_allclose(x[:, 3], np.log2(1 + len(d)) - lf, atol=1e-4)
            np.testing.assert_allclose(x[:, 4], len(vl) / 8, atol=1e-6)
            np.testing.assert_allclose(x[:, 9], np.arange(len(d)) / len(d), atol=1e-6)
            np.testing.assert_allclose(x[:, 21], u / len(vl), atol=1e-6)
            assert np.all(x[:, 23] == 1.0)
        searched += 1
        if searched >= 6:
            break
    assert searched
    c.close()
    g.close()


def test_learned_order_needs_a_ranker(workdir):
    g = tfsm.Graph(workload_path("gnm2k", 1, workdir), 80)
    c = tfsm.Context(g)
    with pytest.raises(tfsm.TfsmError):
        c.mine(80, 4, batched=True, learned=True)
    c.close()
    g.close()


def read_keys(path):
    out = {}
    for line in open(path):
        k, lo, hi = line.rstrip("\n").split("\t")
        out[k] = (int(lo), int(hi))
    return out


@pytest.mark.parametrize("name", ["readme_f2_m4", "gnm2k_f80_m4", "gnm2k_f40_m4", "powerlaw100k_f9000_m5"])
@pytest.mark.parametrize("weights", ["random", "shipped"])
def test_learned_pattern_sets(name, weights, workdir):
    m = golden_manifest()[name]
    g = tfsm.Graph(workload_path(m["workload"], m["seed"], workdir), m["freq"])
    c = tfsm.Context(g)
    if weights == "random":
        c.ranker_set(random_weights(13))
    else:
        c.ranker_load(RANKER)
    kp = os.path.join(str(workdir), f"{name}_learned_{weights}.keys")
    n, _ = c.mine(m["freq"], m["max_nodes"], threads=4, batched=True, learned=True, keys_path=kp)
    assert n == m["reported_total"]
    assert set(read_keys(kp)) == golden_keys(name)
    c.close()
    g.close()


def test_learned_plain_path(workdir):
    m = golden_manifest()["gnm2k_f80_m4"]
    g = tfsm.Graph(workload_path(m["workload"], m["seed"], workdir), m["freq"])
    c = tfsm.Context(g)
    c.ranker_load(RANKER)
    kp = os.path.join(str(workdir), "gnm2k_f80_learned_plain.keys")
    c.mine(m["freq"], m["max_nodes"], threads=4, batched=False, learned=True, keys_path=kp)
    assert set(read_keys(kp)) == golden_keys("gnm2k_f80_m4")
    c.close()
    g.close()


@pytest.mark.parametrize("name", ["gnm2k_f80_m4", "powerlaw100k_f9000_m5"])
def test_learned_supports(name, workdir):
    """Exact mode: supports equal the oracle's.  Capped mode: every frequent
    pattern's bounds bracket the exact MNI, and the lower bound reaches freq."""
    m = golden_manifest()[name]
    path = workload_path(m["workload"], m["seed"], workdir)
    g = tfsm.Graph(path, m["freq"])
    c = tfsm.Context(g)
    c.ranker_load(RANKER)
    ke = os.path.join(str(workdir), f"{name}_learned_exact.keys")
    kc = os.path.join(str(workdir), f"{name}_learned_capped.keys")
    c.mine(m["freq"], m["max_nodes"], threads=4, batched=True, mode=tfsm.MODE_EXACT, learned=True, keys_path=ke)
    c.mine(m["freq"], m["max_nodes"], threads=4, batched=True, mode=tfsm.MODE_CAPPED, learned=True, keys_path=kc)
    og = orc.OracleGraph(path, m["freq"])
    op = os.path.join(str(workdir), f"{name}_oracle_exact_l.keys")
    og.mine(m["freq"], m["max_nodes"], 8, 1, op)
    exact, capped, theirs = read_keys(ke), read_keys(kc), read_keys(op)
    assert set(exact) == set(capped) == set(theirs)
    for k, (lo, hi) in exact.items():
        assert (lo, hi) == theirs[k], k
    for k, (lo, hi) in capped.items():
        mni = theirs[k][0]
        assert m["freq"] <= lo <= mni <= hi, (k, lo, mni, hi)
    c.close()
    g.close()


def test_learned_gain(workdir):
    """The shipped ranker on a graph it was not trained on: the same pattern
    set with fewer existence searches than the reference task order (best of
    three runs each), because the tasks it expects to decide each pattern run
    first and the deferred rest is skipped once the patterns are decided.
    Search-kernel time is reported beside it (the bench compares the paths)."""
    m = golden_manifest()["powerlaw100k_f9000_m5"]
    g = tfsm.Graph(workload_path(m["workload"], m["seed"], workdir), m["freq"])
    c = tfsm.Context(g)
    c.ranker_load(RANKER)
    out = {}
    for learned in (False, True):
        best = None
        for r in range(3):
            kp = os.path.join(str(workdir), f"gain_{int(learned)}.keys")
            _, st = c.mine(m["freq"], m["max_nodes"], threads=4, batched=True, learned=learned, keys_path=kp)
            assert set(read_keys(kp)) == golden_keys("powerlaw100k_f9000_m5")
            if best is None or st["searches"] < best["searches"]:
                best = st
        out[learned] = best
    print({k: (v["search_ms"], v["searches"]) for k, v in out.items()})
    assert out[True]["searches"] < 0.95 * out[False]["searches"]
    c.close()
    g.close()
