"""Train the task ranker on the search's own traces.

Runs exact-mode, level-by-level mining on seeded training graphs with per-task
tracing on: every (pattern, u, candidate) task of every level contributes its
24 features (computed on the GPU by the same kernel the learned order uses)
and its final verdict (1 valid, 0 invalid).  The perceptron is then trained on
the GPU by the library's own forward/backward/Adam kernels and written as a
weight file (1665 little-endian float32 values, BF16-exact) for `run -learned`
and the learned eval order.

    python t-fsm_amd/train_ranker.py --out t-fsm_amd/rankers/powerlaw.f32 \\
        --workload powerlaw100k --seeds 11,12 --freqs 9000,7000 --max-nodes 5

Levels of tens of thousands of patterns over a million-vertex graph (R-MAT
scale 20) have billions of tasks; `--sample-patterns N` traces a seeded
sample of at most N patterns of each level in exact mode and grows the
next level from the whole level evaluated in capped mode:

    python t-fsm_amd/train_ranker.py --out t-fsm_amd/rankers/rmat.f32 \\
        --workload rmat20 --seeds 11,12 --freqs 5300 --max-nodes 5 --sample-patterns 150
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import tfsm  # noqa: E402


def dense_pattern(codes, vv, ev):
    vl = {}
    for f, t, fl, el, tl in codes:
        vl[f], vl[t] = fl, tl
    if any(l not in vv for l in vl.values()) or any(c[3] not in ev for c in codes):
        return None
    return tfsm.make_pattern([vv.index(vl[i]) for i in range(len(vl))],
                             [(f, t, ev.index(el)) for f, t, fl, el, tl in codes])


def level_traces(graph, ctx, freq, max_nodes, mode=tfsm.MODE_EXACT, sample=0, seed=0):
    """Mines level by level (batched) and yields (features, final states, steps) per level.
    sample > 0: trace at most that many patterns of each level (a seeded
    choice) and grow from the whole level evaluated in capped mode."""
    rng = np.random.default_rng(seed)
    csr = graph.csr()
    vv = [int(x) for x in csr["vlabel_values"]]
    ev = [int(x) for x in csr["elabel_values"]]
    level = [(k, 0) for s in graph.size2_codes() for k in graph.extend(s, max_nodes)]
    handles = []
    while level:
        pats, keep = [], []
        for k, h in level:
            p = dense_pattern(k, vv, ev)
            if p is not None:
                pats.append(p)
                keep.append((k, h))
        if not pats:
            break
        parents = [h for _, h in keep]
        if sample and len(pats) > sample:
            pick = np.sort(rng.choice(len(pats), sample, replace=False))
            ctx.eval_level([pats[i] for i in pick], freq, mode=mode, parents=[parents[i] for i in pick], trace=True)
            tr = ctx.read_trace()
            X = ctx.read_features()
            yield X, tr >> 30, tr & ((1 << 30) - 1)
            res = ctx.eval_level(pats, freq, mode=tfsm.MODE_CAPPED, parents=parents, retain=True)
        else:
            res = ctx.eval_level(pats, freq, mode=mode, parents=parents, retain=True, trace=True)
            tr = ctx.read_trace()
            X = ctx.read_features()
            yield X, tr >> 30, tr & ((1 << 30) - 1)
        for h in handles:
            ctx.domain_release(h)
        handles = [r.domain_handle for r in res if r.frequent and r.domain_handle]
        nxt = []
        for (k, _), r in zip(keep, res):
            if r.frequent:
                nxt += [(kk, r.domain_handle) for kk in graph.extend(k, max_nodes)]
        level = nxt
    for h in handles:
        ctx.domain_release(h)


def collect(workload, seeds, freqs, max_nodes, tmp, per_level=400_000, sample=0):
    """Decided tasks' (features, valid) rows; at most per_level rows sampled per level."""
    rng = np.random.default_rng(0)
    Xs, ys = [], []
    for seed in seeds:
        lg = os.path.join(tmp, f"{workload}_{seed}.lg")
        subprocess.run([os.path.join(tfsm.BUILD, "tfsm_gen"), workload, lg, str(seed)], check=True,
                       stdout=subprocess.DEVNULL)
        for freq in freqs:
            g = tfsm.Graph(lg, freq)
            ctx = tfsm.Context(g)
            for X, state, _ in level_traces(g, ctx, freq, max_nodes, sample=sample, seed=seed):
                keep = np.flatnonzero(state > 0)
                if len(keep) > per_level:
                    keep = np.sort(rng.choice(keep, per_level, replace=False))
                Xs.append(X[keep])
                ys.append((state[keep] == 1).astype(np.float32))
            ctx.close()
            g.close()
    return np.concatenate(Xs), np.concatenate(ys)


def init_weights(rng, pos_rate):
    F, H = tfsm.RANKER_FEATURES, tfsm.RANKER_HIDDEN
    w1 = rng.normal(0.0, np.sqrt(2.0 / F), size=(H, F))
    b1 = np.zeros(H)
    w2 = rng.normal(0.0, np.sqrt(1.0 / H), size=H)
    b2 = np.log(pos_rate / max(1e-6, 1.0 - pos_rate))
    return np.concatenate([w1.ravel(), b1, w2, [b2]]).astype(np.float32)


def auc(scores, labels):
    order = np.argsort(scores)
    ranks = np.empty(len(scores))
    ranks[order] = np.arange(1, len(scores) + 1)
    pos = labels > 0.5
    npos, nneg = pos.sum(), (~pos).sum()
    if npos == 0 or nneg == 0:
        return float("nan")
    return float((ranks[pos].sum() - npos * (npos + 1) / 2) / (npos * nneg))


def train(X, y, seed=1, epochs=6, batch=4096, lr=2e-3, max_rows=2_000_000, ctx=None):
    rng = np.random.default_rng(seed)
    if len(X) > max_rows:
        pick = rng.choice(len(X), max_rows, replace=False)
        X, y = X[pick], y[pick]
    perm = rng.permutation(len(X))
    cut = int(0.9 * len(X))
    tr, va = perm[:cut], perm[cut:]
    ctx.ranker_set(init_weights(rng, float(y[tr].mean())))
    loss = ctx.ranker_train(X[tr], y[tr], epochs=epochs, batch=batch, lr=lr, seed=seed)
    z = ctx.ranker_forward(X[va])
    acc = float(((z > 0) == (y[va] > 0.5)).mean())
    return ctx.ranker_get(), dict(rows=int(len(X)), train_loss=float(loss), valid_accuracy=acc,
                                  valid_auc=auc(z, y[va]), positive_rate=float(y.mean()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--workload", default="powerlaw100k")
    ap.add_argument("--seeds", default="11,12")
    ap.add_argument("--freqs", default="9000,7000")
    ap.add_argument("--max-nodes", type=int, default=5)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--sample-patterns", type=int, default=0,
                    help="trace at most this many patterns per level (0: every pattern)")
    args = ap.parse_args()
    seeds = [int(s) for s in args.seeds.split(",")]
    freqs = [int(f) for f in args.freqs.split(",")]
    with tempfile.TemporaryDirectory() as tmp:
        X, y = collect(args.workload, seeds, freqs, args.max_nodes, tmp, sample=args.sample_patterns)
        lg = os.path.join(tmp, f"{args.workload}_{seeds[0]}.lg")
        g = tfsm.Graph(lg, freqs[0])
        ctx = tfsm.Context(g)
        w, info = train(X, y, epochs=args.epochs, ctx=ctx)
        ctx.close()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    w.astype("<f4").tofile(args.out)
    info.update(workload=args.workload, seeds=seeds, freqs=freqs, max_nodes=args.max_nodes, epochs=args.epochs,
                sample_patterns=args.sample_patterns)
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(info, f, indent=1)
    print(json.dumps(info))


if __name__ == "__main__":
    main()
