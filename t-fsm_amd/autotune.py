"""Autotuner for the search kernel's shape on one workload.

Times whole mining runs (graph resident in HBM) for candidate shapes and keeps
the fastest.  The knobs are the level driver, the resident waves per CU, the step budget
of the record splitting, the lanes of a lane group (and of the hot tier's), the DFS steps between
bookkeeping rounds, the sizes of a wave's root-task refills and the DP-iso
passes launched between convergence checks.  The space is searched one knob at a
time from the library's defaults (coordinate descent, two sweeps), keeping
each knob's best value before the next.  A shape is accepted only if its run
returns exactly the pattern set of the first (default) shape: the search's
verdicts never depend on its shape, and the check makes sure.  Writes a tuning
file that `run -tune FILE`, `Context.mine(..., **tfsm.load_tuning(FILE))` and
bench.py apply.

    python t-fsm_amd/autotune.py --workload powerlaw100k --seed 2 --freq 9000 \\
        --max-nodes 5 --ranker t-fsm_amd/rankers/powerlaw.f32 \\
        --out t-fsm_amd/tuning/powerlaw100k_f9000_m5.json
"""
import argparse
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
# Values the tuner may choose from, the library's default first: workgroups of
# 4 waves, 1 or 2 resident per CU (the library clamps to what stays
# resident); record splitting after 64 to 1024 steps, or none; lane groups of
# up to 1-16 lanes, and up to 8-64 for the learned order's hot tier (0: as the
# others); 2-16 steps per bookkeeping round; refills of 16-128 (at
# least) and 128-1024 (at most) lanes' worth of tasks; 2-9 filter passes per
# convergence check (a device-driven call runs that many passes); the
# level driver: device-driven as a replayed hipGraph (0, the default),
# host-driven (1) or device-driven kernel by kernel (2); and the bound, in
# percent of a label bucket, under which a new pattern vertex's LDF
# candidates come from a neighbouring parent domain's adjacency (0: 100).
SPACE = {
    "level_mode": [0, 1, 2],
    "grid_waves_per_cu": [4, 8],
    "split_steps": [128, 64, 256, 1024, NEVER],
    "group_lanes": [8, 1, 2, 4, 16],
    "hot_group_lanes": [0, 8, 16, 32, 64],
    "steps_per_round": [8, 2, 4, 16],
    "refill_min": [128, 16, 32, 64],
    "refill_max": [256, 128, 512, 1024],
    "prune_batch": [6, 2, 4, 9],
    "ldf_push": [0, NEVER, 25, 400],
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
    ap.add_argument("--fix", action="append", default=[], metavar="KNOB=VALUE",
                    help="hold a knob at one value (e.g. level_mode=0: every level as one replayed hipGraph)")
    args = ap.parse_args()
    for f in args.fix:
        knob, value = f.split("=")
        if knob not in SPACE:
            ap.error(f"unknown knob {knob}")
        SPACE[knob] = [int(value)]

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

    reference_keys, results, seen = None, [], {}

    def measure(shape):
        key = tuple(sorted(shape.items()))
        if key in seen:
            return seen[key]
        kp = os.path.join(tmp, "keys.txt")
        times = []
        ctx.mine(args.freq, args.max_nodes, threads=16, batched=True, learned=learned, **shape)   # warm-up
        for r in range(args.repeats):
            _, st = ctx.mine(args.freq, args.max_nodes, threads=16, batched=True, learned=learned,
                             keys_path=kp if r == 0 else None, **shape)
            times.append(st["mining_s"])
        keys = keys_of(kp)
        nonlocal reference_keys
        if reference_keys is None:
            reference_keys = keys
        res = dict(options=dict(shape), median_ms=round(1e3 * statistics.median(times), 3),
                   min_ms=round(1e3 * min(times), 3), same_patterns=keys == reference_keys)
        results.append(res)
        seen[key] = res
        print(json.dumps(res), flush=True)
        return res

    best = measure({k: v[0] for k, v in SPACE.items()})
    for sweep in range(2):
        for knob, values in SPACE.items():
            for v in values:
                shape = dict(best["options"], **{knob: v})
                if shape["refill_min"] > shape["refill_max"]:
                    continue
                r = measure(shape)
                if r["same_patterns"] and r["median_ms"] < best["median_ms"]:
                    best = r
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
