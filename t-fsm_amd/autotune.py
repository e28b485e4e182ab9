"""Autotuner for the search kernel's shape on one workload.

Times whole mining runs (graph resident in HBM) for each candidate shape
(resident waves per CU, step budget of the record splitting) and keeps the
fastest.  A shape is accepted only if its run returns exactly the pattern set
of the first (default) shape: the search's verdicts never depend on its shape,
and the check makes sure.  Writes a tuning file that `run -tune FILE`,
`Context.mine(..., **tfsm.load_tuning(FILE))` and bench.py apply.

    python t-fsm_amd/autotune.py --workload powerlaw100k --seed 2 --freq 9000 \\
        --max-nodes 5 --ranker t-fsm_amd/rankers/powerlaw.f32 \\
        --out t-fsm_amd/tuning/powerlaw100k_f9000_m5.json
"""
import argparse
import itertools
import json
import os
import statistics
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import tfsm  # noqa: E402

NEVER = 0xffffffff
# Shapes the tuner may choose from: workgroups of 4 waves, 1 or 2 resident per
# CU (the library clamps to what stays resident), and record splitting after
# 64 to 1024 steps with one record per donation, or none.
SPACE = {
    "grid_waves_per_cu": [4, 8],
    "split_steps": [64, 128, 256, 1024, NEVER],
}


def keys_of(path):
    with open(path) as f:
        return sorted(l.split("\t")[0] for l in f if l.strip())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="powerlaw100k")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--lg", help="an .lg file instead of a generated workload")
    ap.add_argument("--freq", type=int, required=True)
    ap.add_argument("--max-nodes", type=int, default=5)
    ap.add_argument("--ranker", help="time the learned order (the ranker file to install)")
    ap.add_argument("--repeats", type=int, default=5)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()

    tmp = tempfile.mkdtemp(prefix="tfsm_tune_")
    lg = args.lg
    if not lg:
        lg = os.path.join(tmp, f"{args.workload}_{args.seed}.lg")
        subprocess.run([os.path.join(tfsm.BUILD, "tfsm_gen"), args.workload, lg, str(args.seed)], check=True,
                       stdout=subprocess.DEVNULL)
    g = tfsm.Graph(lg, args.freq)
    ctx = tfsm.Context(g)
    learned = bool(args.ranker)
    if learned:
        ctx.ranker_load(args.ranker)

    shapes = [dict(zip(SPACE, v)) for v in itertools.product(*SPACE.values())]
    reference_keys, results = None, []
    for shape in shapes:
        kp = os.path.join(tmp, "keys.txt")
        times = []
        ctx.mine(args.freq, args.max_nodes, threads=16, batched=True, learned=learned, **shape)   # warm-up
        for r in range(args.repeats):
            _, st = ctx.mine(args.freq, args.max_nodes, threads=16, batched=True, learned=learned,
                             keys_path=kp if r == 0 else None, **shape)
            times.append(st["mining_s"])
        keys = keys_of(kp)
        if reference_keys is None:
            reference_keys = keys
        ok = keys == reference_keys
        results.append(dict(options=shape, median_ms=round(1e3 * statistics.median(times), 3),
                            min_ms=round(1e3 * min(times), 3), same_patterns=ok))
        print(json.dumps(results[-1]), flush=True)
    good = [r for r in results if r["same_patterns"]]
    best = min(good, key=lambda r: r["median_ms"])
    out = dict(workload=dict(name=args.lg or args.workload, seed=args.seed, freq=args.freq,
                             max_nodes=args.max_nodes, learned=learned),
               options=best["options"], measured=results)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print("chosen", json.dumps(best))
    ctx.close()
    g.close()


if __name__ == "__main__":
    main()
