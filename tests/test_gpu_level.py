"""Device-driven level calls against the host-driven sequence and the oracle.

A device-driven call runs the filter, the decisions and layout, the tables,
the learned order and the search as one launch sequence with no host round
trip inside it, either launched kernel by kernel (LEVEL_DEVICE) or as one
captured hipGraph replayed for every level of the same shape (LEVEL_GRAPH).
Results must be bit-identical to the host-driven sequence (LEVEL_HOST): per
pattern decisions, domain sizes and, in exact mode, every valid and invalid
count; one-pattern calls are checked against the oracle's domains, tables
and per-candidate verdicts.
"""
import os

import pytest

import tfsm
from conftest import golden_keys, golden_manifest, workload_path
from test_gpu_parity import check_exact, env_for, variants

pytestmark = pytest.mark.gpu

DEVICE_MODES = [tfsm.LEVEL_DEVICE, tfsm.LEVEL_GRAPH]
MODE_IDS = ["device", "graph"]
LEVELS = [("gnm2k", 1, 40, "gnm2k_f40_m4", 120), ("powerlaw100k", 2, 9000, "powerlaw100k_f9000_m5", 40)]


def level_patterns(env, golden, limit):
    cases = [c for c in variants(golden_keys(golden), limit) if env.dense(*c) is not None]
    return cases, [env.dense(vl, es) for vl, es in cases]


def same_results(a, b, exact, where):
    for i, (x, y) in enumerate(zip(a, b)):
        w = f"{where}: pattern {i}"
        assert x.decided_by == y.decided_by and bool(x.frequent) == bool(y.frequent), w
        assert list(x.domain_size) == list(y.domain_size), w
        if exact:
            assert (x.support_lower, x.support_upper) == (y.support_lower, y.support_upper), w
            assert list(x.valid_count) == list(y.valid_count), w
            assert list(x.invalid_count) == list(y.invalid_count), w


@pytest.mark.parametrize("level_mode", DEVICE_MODES, ids=MODE_IDS)
@pytest.mark.parametrize("name,seed,freq,golden,limit", LEVELS)
def test_level_equals_host_driven(name, seed, freq, golden, limit, level_mode, workdir):
    env = env_for(name, seed, freq, workdir)
    cases, pats = level_patterns(env, golden, limit)
    print(f"level of {len(pats)} patterns, mode {level_mode}", flush=True)
    for mode in (tfsm.MODE_EXACT, tfsm.MODE_CAPPED):
        host = env.ctx.eval_level(pats, env.freq, mode=mode, level_mode=tfsm.LEVEL_HOST)
        assert env.ctx.last_stats().level_mode == tfsm.LEVEL_HOST
        dev = env.ctx.eval_level(pats, env.freq, mode=mode, level_mode=level_mode)
        st = env.ctx.last_stats()
        # the device-driven sequence ran (or the call was redone host-driven
        # because its filter needed more passes than the sequence holds)
        assert st.level_mode in (level_mode, tfsm.LEVEL_HOST)
        same_results(dev, host, mode == tfsm.MODE_EXACT, f"{name} mode {mode}")
        if mode == tfsm.MODE_EXACT:
            for (vl, es), r in zip(cases, dev):
                ref = env.og.eval(vl, es, env.freq, mode=1)
                assert (r.support_lower, r.support_upper) == (ref.lower, ref.upper), f"{vl} {es}"


@pytest.mark.parametrize("level_mode", DEVICE_MODES, ids=MODE_IDS)
def test_one_pattern_calls_against_oracle(level_mode, workdir):
    """One-pattern device-driven calls: the host repeats the device's layout,
    so domains, tables, orders and per-candidate verdicts read through the
    inspection functions equal the oracle's."""
    env = env_for("gnm2k", 1, 40, workdir)
    ran = 0
    for vl, es in variants(golden_keys("gnm2k_f40_m4"), 30):
        if check_exact(env, vl, es, level_mode=level_mode) is not None:
            ran += env.ctx.last_stats().level_mode == level_mode
    assert ran > 10


def test_graph_replays(workdir):
    """The captured sequence replayed for levels of different sizes, in turn
    and repeatedly, with host-driven calls and synchronisations between the
    replays: every replay equals the host-driven result."""
    env = env_for("gnm2k", 1, 40, workdir)
    _, pats = level_patterns(env, "gnm2k_f40_m4", 120)
    levels = [pats[:7], pats[7:60], pats, pats[60:61]]
    want = [env.ctx.eval_level(lv, env.freq, mode=tfsm.MODE_EXACT, level_mode=tfsm.LEVEL_HOST) for lv in levels]
    graphs = 0
    for rnd in range(3):
        for lv, w in zip(levels, want):
            got = env.ctx.eval_level(lv, env.freq, mode=tfsm.MODE_EXACT, level_mode=tfsm.LEVEL_GRAPH)
            graphs += env.ctx.last_stats().level_mode == tfsm.LEVEL_GRAPH
            same_results(got, w, True, f"replay {rnd}, level of {len(lv)}")
            env.ctx.eval_pattern(lv[0], env.freq, mode=tfsm.MODE_EXACT)
    assert graphs >= 9


@pytest.mark.parametrize("level_mode", DEVICE_MODES, ids=MODE_IDS)
def test_learned_level_equals_host_driven(level_mode, workdir):
    env = env_for("powerlaw100k", 2, 9000, workdir)
    env.ctx.ranker_load(os.path.join(os.path.dirname(tfsm.__file__), "rankers", "powerlaw.f32"))
    _, pats = level_patterns(env, "powerlaw100k_f9000_m5", 40)
    host = env.ctx.eval_level(pats, env.freq, mode=tfsm.MODE_EXACT, learned=True, level_mode=tfsm.LEVEL_HOST)
    dev = env.ctx.eval_level(pats, env.freq, mode=tfsm.MODE_EXACT, learned=True, level_mode=level_mode)
    same_results(dev, host, True, "learned exact")
    host = env.ctx.eval_level(pats, env.freq, mode=tfsm.MODE_CAPPED, learned=True, level_mode=tfsm.LEVEL_HOST)
    dev = env.ctx.eval_level(pats, env.freq, mode=tfsm.MODE_CAPPED, learned=True, level_mode=level_mode)
    same_results(dev, host, False, "learned capped")


def test_unconverged_filter_is_redone(workdir):
    """With one DP-iso pass in the sequence (prune_batch 1), a level whose
    first pass removes a candidate has not shown the fixpoint: it is redone
    host-driven, with the same results."""
    env = env_for("powerlaw100k", 2, 9000, workdir)
    _, pats = level_patterns(env, "powerlaw100k_f9000_m5", 40)
    host = env.ctx.eval_level(pats, env.freq, mode=tfsm.MODE_EXACT, level_mode=tfsm.LEVEL_HOST)
    redone = 0
    for chunk in (pats, pats[: len(pats) // 2], pats[len(pats) // 2:]):
        got = env.ctx.eval_level(chunk, env.freq, mode=tfsm.MODE_EXACT, level_mode=tfsm.LEVEL_GRAPH, prune_batch=1)
        redone += env.ctx.last_stats().level_mode == tfsm.LEVEL_HOST
        want = host if chunk is pats else env.ctx.eval_level(chunk, env.freq, mode=tfsm.MODE_EXACT,
                                                              level_mode=tfsm.LEVEL_HOST)
        same_results(got, want, True, f"prune_batch 1, level of {len(chunk)}")
    assert redone >= 1


def test_redone_level_retains_domains(workdir):
    """A device-driven call that keeps domains sets buffers aside before it
    runs; when its filter does not reach the fixpoint (prune_batch 1) the
    call is redone host-driven, the buffers go back, and the retained
    domains are the host-driven call's, read back equal."""
    env = env_for("powerlaw100k", 2, 9000, workdir)
    _, pats = level_patterns(env, "powerlaw100k_f9000_m5", 40)
    for rnd in range(2):
        got = env.ctx.eval_level(pats, env.freq, mode=tfsm.MODE_EXACT, retain=True, level_mode=tfsm.LEVEL_GRAPH,
                                 prune_batch=1)
        assert env.ctx.last_stats().level_mode == tfsm.LEVEL_HOST
        want = env.ctx.eval_level(pats, env.freq, mode=tfsm.MODE_EXACT, retain=True, level_mode=tfsm.LEVEL_HOST)
        same_results(got, want, True, f"redone level with retained domains, round {rnd}")
        kept = 0
        for g, w, p in zip(got, want, pats):
            assert bool(g.domain_handle) == bool(w.domain_handle)
            if g.domain_handle:
                for u in range(p.num_vertices):
                    assert list(env.ctx.domain_read(g.domain_handle, u)) == list(env.ctx.domain_read(w.domain_handle, u))
                kept += 1
        for r in got + want:
            if r.domain_handle:
                env.ctx.domain_release(r.domain_handle)
        assert kept > 0


@pytest.mark.parametrize("level_mode", [tfsm.LEVEL_HOST, tfsm.LEVEL_DEVICE, tfsm.LEVEL_GRAPH],
                         ids=["host", "device", "graph"])
@pytest.mark.parametrize("name", ["gnm2k_f40_m4", "powerlaw100k_f9000_m5"])
def test_mining_level_modes(name, level_mode, workdir):
    m = golden_manifest()[name]
    path = workload_path(m["workload"], m["seed"], workdir)
    g = tfsm.Graph(path, m["freq"])
    ctx = tfsm.Context(g)
    kp = os.path.join(str(workdir), f"{name}_lm{level_mode}.keys")
    for learned in (False, True):
        if learned:
            ctx.ranker_load(os.path.join(os.path.dirname(tfsm.__file__), "rankers", "powerlaw.f32"))
        n, st = ctx.mine(m["freq"], m["max_nodes"], threads=4, batched=True, learned=learned, keys_path=kp,
                         level_mode=level_mode)
        keys = set(l.split("\t")[0] for l in open(kp) if l.strip())
        assert n == m["reported_total"] and keys == golden_keys(name), (learned, level_mode)
        if level_mode == tfsm.LEVEL_HOST:
            assert st["device_calls"] == 0
        else:
            assert st["device_calls"] > 0
        assert (st["graph_calls"] > 0) == (level_mode == tfsm.LEVEL_GRAPH)
    ctx.close()


@pytest.mark.parametrize("hot", [8, 64])
def test_hot_tier_groups(hot, workdir):
    """The learned order's hot tier with up to `hot` lanes per task (other
    tasks up to 4): exact supports and verdict counts equal the host-driven
    sequence's with the default groups."""
    env = env_for("powerlaw100k", 2, 9000, workdir)
    env.ctx.ranker_load(os.path.join(os.path.dirname(tfsm.__file__), "rankers", "powerlaw.f32"))
    _, pats = level_patterns(env, "powerlaw100k_f9000_m5", 40)
    want = env.ctx.eval_level(pats, env.freq, mode=tfsm.MODE_EXACT, learned=True, level_mode=tfsm.LEVEL_HOST)
    for lm in (tfsm.LEVEL_HOST, tfsm.LEVEL_GRAPH):
        got = env.ctx.eval_level(pats, env.freq, mode=tfsm.MODE_EXACT, learned=True, level_mode=lm, group_lanes=4,
                                 hot_group_lanes=hot)
        same_results(got, want, True, f"hot tier groups of {hot}, level mode {lm}")


@pytest.mark.parametrize("level_mode", [tfsm.LEVEL_HOST, tfsm.LEVEL_GRAPH], ids=["host", "graph"])
@pytest.mark.parametrize("name", ["gnm2k_f40_m4", "powerlaw100k_f9000_m5"])
def test_ldf_push_mining(name, level_mode, workdir):
    """Every new pattern vertex next to a parent domain takes its LDF
    candidates from that domain's adjacency (ldf_push 1000000): the mined
    pattern set equals the reference's, in both level drivers."""
    m = golden_manifest()[name]
    path = workload_path(m["workload"], m["seed"], workdir)
    g = tfsm.Graph(path, m["freq"])
    ctx = tfsm.Context(g)
    kp = os.path.join(str(workdir), f"{name}_push{level_mode}.keys")
    print(f"{name} push always, level mode {level_mode}", flush=True)
    n, st = ctx.mine(m["freq"], m["max_nodes"], threads=4, batched=True, keys_path=kp, level_mode=level_mode,
                     ldf_push=1000000)
    keys = set(l.split("\t")[0] for l in open(kp) if l.strip())
    assert n == m["reported_total"] and keys == golden_keys(name)
    ctx.close()


@pytest.mark.parametrize("level_mode", DEVICE_MODES, ids=MODE_IDS)
@pytest.mark.parametrize("mode", [tfsm.MODE_EXACT, tfsm.MODE_CAPPED], ids=["exact", "capped"])
def test_level_retained_domains(mode, level_mode, workdir):
    """Retained domains of a device-driven level (copied by the sequence's
    last kernel into buffers set aside before the call) read back, in exact
    mode, equal to the host-driven call's and to the oracle's child domains
    (the domain minus the candidates proven invalid).  A capped search stops
    once the pattern is decided, so its invalid set depends on the schedule:
    the read-back lies between the exact child domain and the whole domain.
    Two calls in a row, so the second takes buffers the first released."""
    env = env_for("gnm2k", 1, 40, workdir)
    cases, pats = level_patterns(env, "gnm2k_f40_m4", 120)
    for rnd in range(2):
        host = env.ctx.eval_level(pats, env.freq, mode=mode, retain=True, level_mode=tfsm.LEVEL_HOST)
        # enough DP-iso passes in the sequence that the call is not redone host-driven
        dev = env.ctx.eval_level(pats, env.freq, mode=mode, retain=True, level_mode=level_mode, prune_batch=32)
        assert env.ctx.last_stats().level_mode == level_mode
        kept = 0
        for (vl, es), h, d in zip(cases, host, dev):
            assert bool(h.frequent) == bool(d.frequent)
            if not d.frequent:
                assert d.domain_handle == 0
                continue
            ref = env.og.eval(vl, es, env.freq, mode=1)
            for u in range(len(vl)):
                got = list(env.ctx.domain_read(d.domain_handle, u))
                w = f"{vl} {es} vertex {u}"
                if mode == tfsm.MODE_EXACT:
                    assert got == list(env.ctx.domain_read(h.domain_handle, u)), w
                    assert got == list(ref.child_domain(u)), w
                else:
                    assert got == sorted(got) and set(ref.child_domain(u)) <= set(got) <= set(ref.domain(u)), w
            kept += 1
        for r in host + dev:
            if r.domain_handle:
                env.ctx.domain_release(r.domain_handle)
        assert kept > 10, rnd
