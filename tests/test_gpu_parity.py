"""Per-pattern parity of the HIP path against the CPU oracle, through the C ABI.

Every verdict, domain and table is integer data, so the bar is bit-exact
equality: filtered domains C(u), auxiliary tables (the reference's
edge_matrix), matching orders, per-candidate verdicts and the exact MNI
support.  Patterns: every frequent pattern of the reference's golden sets
(three or more vertices) plus cycle-closing and relabelled variants of them,
so frequent, infrequent, filter-rejected and express-path cases all occur.
"""
import numpy as np
import pytest

import oracle as orc
import tfsm
from conftest import golden_keys, workload_path

pytestmark = pytest.mark.gpu

CASES = [
    ("readme", 0, 1, "readme_f1_m3"),
    ("readme", 0, 2, "readme_f2_m4"),
    ("gnm2k", 1, 80, "gnm2k_f80_m4"),
    ("gnm2k", 1, 40, "gnm2k_f40_m4"),
    ("powerlaw100k", 2, 9000, "powerlaw100k_f9000_m5"),
]


class Env:
    def __init__(self, path, freq):
        self.freq = freq
        self.g = tfsm.Graph(path, freq)
        self.ctx = tfsm.Context(self.g)
        self.og = orc.OracleGraph(path, freq)
        csr = self.g.csr()
        self.vv = [int(x) for x in csr["vlabel_values"]]
        self.ev = [int(x) for x in csr["elabel_values"]]

    def dense(self, vlabels, edges):
        if any(l not in self.vv for l in vlabels) or any(e[2] not in self.ev for e in edges):
            return None
        return tfsm.make_pattern([self.vv.index(l) for l in vlabels],
                                 [(a, b, self.ev.index(l)) for a, b, l in edges])


_envs = {}


def env_for(name, seed, freq, workdir):
    key = (name, seed, freq)
    if key not in _envs:
        _envs[key] = Env(workload_path(name, seed, workdir), freq)
    return _envs[key]


def variants(keys, limit):
    """Golden patterns (>= 3 vertices) and derived candidates, deterministic."""
    out = []
    for k in sorted(keys):
        vl, es = tfsm.parse_key(k)
        if len(vl) < 3:
            continue
        out.append((vl, es))
        # close a cycle between the first and the last vertex when not adjacent
        adj = {(a, b) for a, b, _ in es} | {(b, a) for a, b, _ in es}
        if (0, len(vl) - 1) not in adj:
            out.append((vl, es + [(0, len(vl) - 1, es[0][2])]))
        # relabel the last vertex to the first vertex's label
        if vl[-1] != vl[0]:
            out.append((vl[:-1] + [vl[0]], es))
        if len(out) >= limit:
            break
    return out[:limit]


def check_exact(env, vl, es, **opts):
    pat = env.dense(vl, es)
    if pat is None:
        return None
    res = env.ctx.eval_pattern(pat, env.freq, mode=tfsm.MODE_EXACT, **opts)
    ref = env.og.eval(vl, es, env.freq, mode=1)
    where = f"pattern {vl} {es}"
    assert res.decided_by == ref.decided_by, where
    for u in range(len(vl)):
        assert np.array_equal(env.ctx.last_domain(u), ref.domain(u)), f"{where}: domain {u}"
    if ref.decided_by == 0:
        assert res.support_upper == ref.upper and not res.frequent, where
        return res
    assert (res.support_lower, res.support_upper) == (ref.lower, ref.upper), where
    assert bool(res.frequent) == ref.frequent, where
    if ref.decided_by == 2:
        for a in range(len(vl)):
            for b in range(len(vl)):
                mine, theirs = env.ctx.last_aux(a, b), ref.aux(a, b)
                assert (mine is None) == (theirs is None), f"{where}: table {a}->{b} presence"
                if mine is not None:
                    assert np.array_equal(mine[0], theirs[0]) and np.array_equal(mine[1], theirs[1]), \
                        f"{where}: table {a}->{b}"
            assert env.ctx.last_order(a, len(vl)) == ref.order(a), f"{where}: order {a}"
        for u in range(len(vl)):
            assert np.array_equal(env.ctx.last_states(u), ref.states(u)), f"{where}: verdicts {u}"
    return res


MATCH = [tfsm.MATCH_ADAPTIVE, tfsm.MATCH_REFERENCE]
MATCH_IDS = ["adaptive", "reference"]


@pytest.mark.parametrize("match", MATCH, ids=MATCH_IDS)
@pytest.mark.parametrize("name,seed,freq,golden", CASES)
def test_exact_parity(name, seed, freq, golden, match, workdir):
    env = env_for(name, seed, freq, workdir)
    cases = variants(golden_keys(golden), 60)
    assert cases
    kinds = set()
    for vl, es in cases:
        r = check_exact(env, vl, es, match_order=match)
        if r is not None:
            kinds.add(r.decided_by)
    assert kinds  # at least one pattern was evaluated


@pytest.mark.parametrize("name,seed,freq,golden", CASES)
def test_capped_verdicts(name, seed, freq, golden, workdir):
    env = env_for(name, seed, freq, workdir)
    for vl, es in variants(golden_keys(golden), 60):
        pat = env.dense(vl, es)
        if pat is None:
            continue
        res = env.ctx.eval_pattern(pat, freq, mode=tfsm.MODE_CAPPED)
        ref = env.og.eval(vl, es, freq, mode=0)
        assert bool(res.frequent) == ref.frequent, f"{vl} {es}"
        if res.frequent:
            assert res.support_lower >= freq
        else:
            assert res.support_upper < freq


# ldf_push: 0 is the library's bound (push when at most 100% of the bucket),
# 1000000 pushes from every neighbouring parent domain, 0xffffffff never does
@pytest.mark.parametrize("push", [0, 1000000, 0xffffffff], ids=["default", "always", "never"])
def test_parent_domains(push, workdir):
    """A child's candidates start from the parent's domains minus its
    proven-invalid ones; a new vertex's candidates come from its label bucket
    or from the adjacency of its neighbour's parent domain, with the same
    filtered domains either way."""
    env = env_for("gnm2k", 1, 40, workdir)
    keys = sorted(k for k in golden_keys("gnm2k_f40_m4") if k.count(";") == 2)
    checked = 0
    for k in keys[:25]:
        vl, es = tfsm.parse_key(k)
        parent = env.dense(vl, es)
        pres = env.ctx.eval_pattern(parent, env.freq, mode=tfsm.MODE_EXACT, retain=True)
        ref_parent = env.og.eval(vl, es, env.freq, mode=1)
        assert pres.frequent and ref_parent.frequent
        pdoms = [ref_parent.child_domain(u) for u in range(len(vl))]
        for u in range(len(vl)):
            assert np.array_equal(env.ctx.domain_read(pres.domain_handle, u), pdoms[u])
        # child: forward edge from the last vertex to a new vertex with label vl[0]
        cvl, ces = vl + [vl[0]], es + [(len(vl) - 1, len(vl), es[0][2])]
        child = env.dense(cvl, ces)
        cres = env.ctx.eval_pattern(child, env.freq, mode=tfsm.MODE_EXACT, parent=pres.domain_handle, ldf_push=push)
        cref = env.og.eval(cvl, ces, env.freq, mode=1, parent_domains=pdoms)
        for u in range(len(cvl)):
            assert np.array_equal(env.ctx.last_domain(u), cref.domain(u)), f"{cvl} {ces} domain {u}"
        assert (cres.support_lower, cres.support_upper) == (cref.lower, cref.upper)
        env.ctx.domain_release(pres.domain_handle)
        checked += 1
    assert checked > 5


def test_level_equals_single(workdir):
    """The batched level path returns the same verdicts and supports as per-pattern calls."""
    env = env_for("gnm2k", 1, 40, workdir)
    cases = [c for c in variants(golden_keys("gnm2k_f40_m4"), 120) if env.dense(*c) is not None]
    pats = [env.dense(vl, es) for vl, es in cases]
    for mode in (tfsm.MODE_EXACT, tfsm.MODE_CAPPED):
        batch = env.ctx.eval_level(pats, env.freq, mode=mode)
        for (vl, es), p, b in zip(cases, pats, batch):
            s = env.ctx.eval_pattern(p, env.freq, mode=mode)
            assert b.frequent == s.frequent and b.decided_by == s.decided_by, f"{vl} {es}"
            assert list(b.domain_size) == list(s.domain_size)
            if mode == tfsm.MODE_EXACT:
                assert (b.support_lower, b.support_upper) == (s.support_lower, s.support_upper)
                assert list(b.valid_count) == list(s.valid_count)


NEVER = 0xffffffff


@pytest.mark.parametrize("match", MATCH, ids=MATCH_IDS)
@pytest.mark.parametrize("split_steps,fanout", [
    (64, 1),           # record donation every 64 steps
    (64, 16),          # each donation cut into 16 records
    (NEVER, 1),        # no splitting at all
    (0, 0),            # tuned defaults
])
def test_split_search_parity(split_steps, fanout, match, workdir):
    """Sub-tree donation through the record queue at a tiny step budget, cut
    into one or many records, and no splitting at all, must all give the
    oracle's exact per-candidate verdicts, in either matching order (a record
    of the adaptive order carries the pattern vertex of each depth)."""
    env = env_for("powerlaw100k", 2, 9000, workdir)
    opts = dict(split_steps=split_steps, split_fanout=fanout, match_order=match)
    for vl, es in variants(golden_keys("powerlaw100k_f9000_m5"), 15):
        pat = env.dense(vl, es)
        if pat is None:
            continue
        res = env.ctx.eval_pattern(pat, env.freq, mode=tfsm.MODE_EXACT, **opts)
        ref = env.og.eval(vl, es, env.freq, mode=1)
        assert (res.support_lower, res.support_upper) == (ref.lower, ref.upper), f"{vl} {es}"
        if ref.decided_by == 2:
            for u in range(len(vl)):
                assert np.array_equal(env.ctx.last_states(u), ref.states(u)), f"{vl} {es} verdicts {u}"
        cap = env.ctx.eval_pattern(pat, env.freq, mode=tfsm.MODE_CAPPED, **opts)
        assert bool(cap.frequent) == ref.frequent


@pytest.mark.parametrize("match", MATCH, ids=MATCH_IDS)
@pytest.mark.parametrize("group,split_steps,fanout", [
    (2, NEVER, 1),
    (8, NEVER, 1),
    (64, NEVER, 1),
    (8, 16, 2),        # grouped lanes that also donate records
    (1, 0, 0),         # one lane per task
])
def test_lane_group_parity(group, split_steps, fanout, match, workdir):
    """Root tasks searched by lane groups (up to `group` lanes interleaving a
    task's depth-1 candidates, the group's lanes pieces of one counter) give
    the oracle's exact per-candidate verdicts and the capped verdicts, one
    pattern per call and a whole level per call.  These calls have fewer
    tasks than resident lanes, so the groups are used."""
    env = env_for("gnm2k", 1, 40, workdir)
    opts = dict(group_lanes=group, split_steps=split_steps, split_fanout=fanout, match_order=match)
    cases = [c for c in variants(golden_keys("gnm2k_f40_m4"), 60) if env.dense(*c) is not None]
    for vl, es in cases:
        pat = env.dense(vl, es)
        res = env.ctx.eval_pattern(pat, env.freq, mode=tfsm.MODE_EXACT, **opts)
        ref = env.og.eval(vl, es, env.freq, mode=1)
        assert (res.support_lower, res.support_upper) == (ref.lower, ref.upper), f"{vl} {es}"
        if ref.decided_by == 2:
            for u in range(len(vl)):
                assert np.array_equal(env.ctx.last_states(u), ref.states(u)), f"{vl} {es} verdicts {u}"
        cap = env.ctx.eval_pattern(pat, env.freq, mode=tfsm.MODE_CAPPED, **opts)
        assert bool(cap.frequent) == ref.frequent, f"{vl} {es}"
    pats = [env.dense(vl, es) for vl, es in cases]
    lev = env.ctx.eval_level(pats, env.freq, mode=tfsm.MODE_EXACT, **opts)
    for (vl, es), r in zip(cases, lev):
        ref = env.og.eval(vl, es, env.freq, mode=1)
        assert (r.support_lower, r.support_upper) == (ref.lower, ref.upper), f"level {vl} {es}"


def test_lane_group_option_checked(workdir):
    env = env_for("gnm2k", 1, 40, workdir)
    vl, es = sorted(variants(golden_keys("gnm2k_f40_m4"), 5))[0]
    for bad in (3, 128):
        with pytest.raises(tfsm.TfsmError) as e:
            env.ctx.eval_pattern(env.dense(vl, es), env.freq, group_lanes=bad)
        assert e.value.code == -1


def test_adaptive_order_less_work(workdir):
    """The adaptive matching order gives the reference order's exact verdicts
    with fewer DFS steps in total (traced per task) on the power-law graph's
    four- and five-vertex patterns."""
    env = env_for("powerlaw100k", 2, 9000, workdir)
    steps = {tfsm.MATCH_ADAPTIVE: 0, tfsm.MATCH_REFERENCE: 0}
    for vl, es in variants(golden_keys("powerlaw100k_f9000_m5"), 40):
        pat = env.dense(vl, es)
        if pat is None or len(vl) < 4:
            continue
        states = {}
        for match in steps:
            res = env.ctx.eval_pattern(pat, env.freq, mode=tfsm.MODE_EXACT, trace=True, match_order=match)
            if res.decided_by != tfsm.DECIDED_SEARCH:
                break
            tr = env.ctx.read_trace()
            steps[match] += int((tr & ((1 << 30) - 1)).sum())
            states[match] = [env.ctx.last_states(u) for u in range(len(vl))]
        if len(states) == 2:
            for a, b in zip(states[tfsm.MATCH_ADAPTIVE], states[tfsm.MATCH_REFERENCE]):
                assert np.array_equal(a, b), f"{vl} {es}"
    print("traced steps", steps)
    assert 0 < steps[tfsm.MATCH_ADAPTIVE] < steps[tfsm.MATCH_REFERENCE]


@pytest.mark.parametrize("long_opts", [
    dict(long_steps=64),                                           # second phase, default offers
    dict(long_steps=64, long_donate_every=1, long_fanout=2),       # offers at every round
    dict(long_steps=256, long_donate_every=NEVER, long_waves_per_cu=4),
    dict(long_steps=NEVER),                                        # first phase only
], ids=["long64", "long64-offer1", "long256-nooffer", "never"])
@pytest.mark.parametrize("name,seed,freq,golden,limit", [
    ("gnm2k", 1, 40, "gnm2k_f40_m4", 60),
    ("powerlaw100k", 2, 9000, "powerlaw100k_f9000_m5", 40),
])
def test_long_search_parity(long_opts, name, seed, freq, golden, limit, workdir):
    """Root searches handed to the wave-cooperative second phase after 64 or
    256 steps (its pieces offered to idle waves at the default interval, at
    every round, or never), or never: the oracle's exact per-candidate
    verdicts and the capped verdicts either way."""
    env = env_for(name, seed, freq, workdir)
    for vl, es in variants(golden_keys(golden), limit):
        pat = env.dense(vl, es)
        if pat is None:
            continue
        res = env.ctx.eval_pattern(pat, env.freq, mode=tfsm.MODE_EXACT, **long_opts)
        ref = env.og.eval(vl, es, env.freq, mode=1)
        assert (res.support_lower, res.support_upper) == (ref.lower, ref.upper), f"{vl} {es}"
        if ref.decided_by == 2:
            for u in range(len(vl)):
                assert np.array_equal(env.ctx.last_states(u), ref.states(u)), f"{vl} {es} verdicts {u}"
        cap = env.ctx.eval_pattern(pat, env.freq, mode=tfsm.MODE_CAPPED, **long_opts)
        assert bool(cap.frequent) == ref.frequent


def test_split_level_parity(workdir):
    """A whole level in one launch with aggressive splitting: exact supports
    and verdict counts equal the per-pattern evaluations without splitting.
    (An earlier in-wave work-sharing scheme, since removed, faulted on this
    case; the record queue it runs now is checked on the CPU first by
    tests/test_lane_dfs.py::test_level_record_queue.)"""
    env = env_for("powerlaw100k", 2, 9000, workdir)
    pats = []
    for vl, es in variants(golden_keys("powerlaw100k_f9000_m5"), 40):
        pat = env.dense(vl, es)
        if pat is not None and len(vl) >= 4:
            pats.append(pat)
    assert pats
    lev = env.ctx.eval_level(pats, env.freq, mode=tfsm.MODE_EXACT, split_steps=64, split_fanout=4)
    for pat, r in zip(pats, lev):
        s = env.ctx.eval_pattern(pat, env.freq, mode=tfsm.MODE_EXACT, split_steps=NEVER)
        assert (r.support_lower, r.support_upper) == (s.support_lower, s.support_upper)
        assert list(r.valid_count) == list(s.valid_count) and list(r.invalid_count) == list(s.invalid_count)


def test_tables_outgrow_the_pool(workdir):
    """With a table pool far too small, a call's tables overflow it; the call is
    redone with the pool enlarged, and the results equal the oracle's (a
    single pattern and a whole level)."""
    env = Env(workload_path("gnm2k", 1, workdir), 40)
    env.ctx.set_pool_capacity(256)
    cases = variants(golden_keys("gnm2k_f40_m4"), 12)
    for vl, es in cases:
        check_exact(env, vl, es)
    env.ctx.set_pool_capacity(256)
    pats = [p for p in (env.dense(vl, es) for vl, es in cases) if p is not None]
    lev = env.ctx.eval_level(pats, env.freq, mode=tfsm.MODE_EXACT)
    for (vl, es), r in zip([c for c in cases if env.dense(*c) is not None], lev):
        ref = env.og.eval(vl, es, env.freq, mode=1)
        assert (r.support_lower, r.support_upper) == (ref.lower, ref.upper), f"{vl} {es}"
    env.ctx.close()


def test_large_pattern_and_errors(workdir):
    """Nine-vertex paths use the 16-deep search kernel; malformed input is an error code."""
    env = env_for("gnm2k", 1, 40, workdir)
    vl = [env.vv[i % len(env.vv)] for i in range(9)]
    es = [(i, i + 1, env.ev[0]) for i in range(8)]
    res = env.ctx.eval_pattern(env.dense(vl, es), 1, mode=tfsm.MODE_EXACT)
    ref = env.og.eval(vl, es, 1, mode=1)
    assert (res.support_lower, res.support_upper) == (ref.lower, ref.upper)
    bad = tfsm.make_pattern([0, 0, 0], [(0, 1, 0)])          # disconnected
    with pytest.raises(tfsm.TfsmError) as e:
        env.ctx.eval_pattern(bad, 10)
    assert e.value.code == -1
    with pytest.raises(tfsm.TfsmError) as e:
        env.ctx.eval_pattern(tfsm.make_pattern([0, 0], [(0, 1, 0)]), 10, parent=987654321)
    assert e.value.code == -5
