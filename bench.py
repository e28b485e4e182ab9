"""Benchmark: MNI frequent-subgraph mining of the 100K-vertex power-law graph
(Chung-Lu, 1.1M edges, Zipf(1) labels over 29 values, freq 9000, maxNodes 5)
on one MI355X.

One step = one whole mining run (size-2 level from the edge statistics, then
every candidate pattern of every extension level evaluated on the GPU) with
the graph already resident in HBM.  The batched path with the learned task
order (one fused launch sequence per level, tasks ordered by the shipped
ranker, t-fsm_amd/rankers/powerlaw.f32, trained on seeds 11 and 12 only) is
the reported value; the plain path (one launch sequence per pattern,
reference task order) and the batched path without the ranker are timed
beside it.  The CPU baseline is the reference's own parallel build (compiled
from its sources into oracle/_ref), timed once on this host's cores.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]

`--workload powerlaw100k_f6000_m6` times the batched configuration instead
(the same graph at freq 6000, maxNodes 6: 631 candidate patterns, 334
frequent); its reference run takes minutes, so it has no CPU baseline here.
"""
import argparse
import json
import os
import statistics
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "t-fsm_amd"))

WORKLOADS = {
    # golden: the reference's pattern set; min_vertices: the smallest patterns it lists
    "powerlaw100k_f9000_m5": dict(name="powerlaw100k", seed=2, freq=9000, max_nodes=5, vertices=100000,
                                  edges=1100000, label="powerlaw100k_chunglu_freq9000_maxnodes5", min_vertices=2,
                                  tuning="powerlaw100k_f9000_m5.json"),
    "powerlaw100k_f6000_m6": dict(name="powerlaw100k", seed=2, freq=6000, max_nodes=6, vertices=100000,
                                  edges=1100000, label="powerlaw100k_chunglu_freq6000_maxnodes6", min_vertices=3,
                                  tuning="powerlaw100k_f6000_m6.json"),
}
WORKLOAD = WORKLOADS["powerlaw100k_f9000_m5"]
HBM_PEAK_GBPS = 8000.0
RANKER = os.path.join(REPO, "t-fsm_amd", "rankers", "powerlaw.f32")


def host_info():
    """Host cores: what the OS shows, what this process may use, and the CPU
    share this GPU is allotted (OMP_NUM_THREADS on a GPU box: 16 host cores
    per GPU; worker pools there must stay within it)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or affinity
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": affinity, "share": share, "model": model,
            "threads": max(1, min(affinity, share))}


def host_threads():
    return host_info()["threads"]


def profiled_traffic(workload):
    """HBM bytes per search launch from the committed counter summary of this
    workload (two rocprofv3 --pmc runs, FETCH_SIZE x 2.0 and WRITE_SIZE x 1.0,
    the scales a 256 MiB copy gives), or None when there is none."""
    path = os.path.join(REPO, "profiles", f"k_verify_{workload}.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch = "):
                return int(line.split("=")[1]), os.path.relpath(path, REPO)
    return None, None


def golden_keys(workload):
    path = os.path.join(REPO, "tests", "golden", workload + ".keys")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return set(l.strip() for l in f if l.strip())


def read_keys(path, min_vertices=2):
    # a key lists its edges, each ended by ';': a pattern of v >= 3 vertices has >= 2 edges
    with open(path) as f:
        keys = set(l.split("\t")[0] for l in f if l.strip())
    return keys if min_vertices <= 2 else set(k for k in keys if k.count(";") >= min_vertices - 1)


def cpu_baseline(lg, threads, runs=3):
    """The reference's parallel build on the host cores (`-thread` = the
    threads this job may use), `runs` full runs of about 1.5-2 s each; its
    [TIME] Mining Time (loading excluded) per run and their median."""
    ref = os.path.join(REPO, "oracle", "_ref", "run_parallel")
    if not os.path.exists(ref):
        return None
    mining, walls, frequent = [], [], None
    for _ in range(runs):
        with tempfile.TemporaryDirectory() as tmp:
            t0 = time.time()
            out = subprocess.run([ref, "-file", lg, "-freq", str(WORKLOAD["freq"]), "-thread", str(threads),
                                  "-maxNodes", str(WORKLOAD["max_nodes"])], cwd=tmp, capture_output=True, text=True,
                                 timeout=600).stdout
            walls.append(time.time() - t0)
        for line in out.split("\n"):
            if "Mining Time" in line:
                mining.append(float(line.split(":")[1].split()[0]))
            if "# Frequent Patterns" in line:
                frequent = int(line.split(":")[1])
    if len(mining) != runs:
        return None
    return {"mining_s": statistics.median(mining), "runs_s": [round(x, 3) for x in mining],
            "wall_s": statistics.median(walls), "frequent": frequent}


def oracle_baseline(lg, threads, runs=3):
    """This build's own CPU oracle (oracle/, a multithreaded restatement of the
    reference's algorithm) mining the same job on the same cores, capped
    mode; median of `runs` runs (about 1-3 s each)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    try:
        import oracle as orc
    except OSError:
        return None
    og = orc.OracleGraph(lg, WORKLOAD["freq"])
    secs, n = [], 0
    with tempfile.TemporaryDirectory() as tmp:
        for _ in range(runs):
            n, st = og.mine(WORKLOAD["freq"], WORKLOAD["max_nodes"], threads, 0, os.path.join(tmp, "keys.txt"))
            secs.append(st["mining_s"])
    og.close()
    return {"mining_s": statistics.median(secs), "runs_s": [round(x, 3) for x in secs], "frequent": n}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--paths", default="plain,batched,learned",
                    help="paths to time (comma list of plain, batched, learned; the value needs learned)")
    ap.add_argument("--workload", default="powerlaw100k_f9000_m5", choices=sorted(WORKLOADS))
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first copy this many MiB device to device (a kernel of known traffic for counter scaling)")
    args = ap.parse_args()
    global WORKLOAD
    WORKLOAD = WORKLOADS[args.workload]
    if args.workload != "powerlaw100k_f9000_m5":
        args.no_cpu_baseline = True
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards across patterns "
              "of a level and across (u, v) tasks, but only the one-GPU share is built and timed)", file=sys.stderr)
        sys.exit(2)

    import torch
    import tfsm

    tmp = tempfile.mkdtemp(prefix="tfsm_bench_")
    lg = os.path.join(tmp, "powerlaw100k.lg")
    subprocess.run([os.path.join(REPO, "t-fsm_amd", "_build", "tfsm_gen"), WORKLOAD["name"], lg,
                    str(WORKLOAD["seed"])], check=True, stdout=subprocess.DEVNULL)
    graph = tfsm.Graph(lg, WORKLOAD["freq"])
    ctx = tfsm.Context(graph)             # graph now resident in HBM
    ctx.ranker_load(RANKER)
    # the search-kernel shape chosen by t-fsm_amd/autotune.py for this workload
    # (library defaults when the file is absent); the same for every path
    tuning_file = os.path.join(REPO, "t-fsm_amd", "tuning", WORKLOAD["tuning"]) if WORKLOAD["tuning"] else None
    tuning = tfsm.load_tuning(tuning_file) if tuning_file and os.path.exists(tuning_file) else {}
    gold = golden_keys(args.workload)
    threads = host_threads()

    def run(batched, learned, steps, warmup, tag, **extra):
        keys_path = os.path.join(tmp, f"keys_{tag}.txt")
        opts = dict(tuning, **extra)
        for _ in range(warmup):
            ctx.mine(WORKLOAD["freq"], WORKLOAD["max_nodes"], threads=threads, batched=batched, learned=learned,
                     **opts)
        torch.cuda.synchronize()
        agg = dict(evaluated=0, frequent=0, verdicts=0, searches=0, search_ms=0.0, search_launches=0,
                   bytes_search=0, filter_ms=0.0, table_ms=0.0, bytes_table=0, graph_calls=0, eval_calls=0)
        t0 = time.perf_counter()
        for i in range(steps):
            n, st = ctx.mine(WORKLOAD["freq"], WORKLOAD["max_nodes"], threads=threads, batched=batched,
                             learned=learned, keys_path=keys_path if i == 0 else None, **opts)
            for k in agg:
                agg[k] += st[k]
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        keys = read_keys(keys_path, WORKLOAD["min_vertices"])
        parity = gold is None or keys == gold
        return elapsed, agg, parity, n

    if args.copy_mib:
        src = torch.empty(args.copy_mib << 18, dtype=torch.int32, device="cuda")
        src.fill_(1)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
    paths = set(args.paths.split(","))
    none = (float("nan"), None, True, 0)
    plain_s, plain_agg, plain_ok, _ = run(False, False, args.steps, args.warmup, "plain") if "plain" in paths else none
    batch_s, batch_agg, batch_ok, _ = run(True, False, args.steps, args.warmup, "batched") if "batched" in paths else none
    learn_s, agg, learn_ok, nfreq = run(True, True, args.steps, args.warmup, "learned")
    # The roofline pass: the learned path again with the same device-driven
    # launch sequence launched kernel by kernel, so that HIP events on the
    # library's stream bracket the search launches (inside a replayed hipGraph
    # the library times its phases with the device's wall clock instead)
    rk = max(5, args.steps // 4)
    _, ragg, roof_ok, _ = run(True, True, rk, 1, "roofline", level_mode=tfsm.LEVEL_DEVICE)

    K = args.steps
    evaluated = agg["evaluated"] / K
    value = agg["evaluated"] / learn_s
    plain_value = plain_agg["evaluated"] / plain_s if plain_agg else None
    batch_value = batch_agg["evaluated"] / batch_s if batch_agg else None

    def side(a, sec, v):
        if not a:
            return None
        return {"value": round(v, 3), "ms_per_step": round(1e3 * sec / K, 3),
                "verifications_per_s": round(a["verdicts"] / sec, 1), "searches_per_step": a["searches"] / K,
                "search_ms_per_step": round(a["search_ms"] / K, 3)}
    ms_search = ragg["search_ms"] / max(1, ragg["search_launches"])
    bytes_per_launch = ragg["bytes_search"] / max(1, ragg["search_launches"])
    achieved = bytes_per_launch / (ms_search * 1e-3) / 1e9 if ms_search > 0 else 0.0

    traffic, traffic_src = profiled_traffic(args.workload)

    cpu = None
    if not args.no_cpu_baseline:
        host = host_info()
        ref = cpu_baseline(lg, threads)
        if ref and ref.get("mining_s"):
            orc = oracle_baseline(lg, threads)
            cpu = {"value": round(evaluated / ref["mining_s"], 3), "unit": "patterns/s", "cores": threads,
                   "kind": "reference",
                   "sample": (f"the whole job, 3 full runs of the reference parallel build (oracle/_ref/run_parallel,"
                              f" -thread {threads}) on the same graph; value = the {int(evaluated)} candidate "
                              f"patterns this job evaluates / the median [TIME] Mining Time {ref['mining_s']:.3f} s "
                              f"(it includes the reference's 0.1 s idle-poll floor); it reported "
                              f"{ref.get('frequent')} frequent patterns. -thread is this GPU's host-core share "
                              f"({host['share']}) of the {host['nproc']} cores the host shows"),
                   "mining_s": ref["mining_s"], "runs_s": ref["runs_s"],
                   "host": {"nproc": host["nproc"], "affinity": host["affinity"], "share": host["share"],
                            "lscpu_model": host["model"]},
                   "oracle": None if orc is None else {
                       "value": round(evaluated / orc["mining_s"], 3), "mining_s": round(orc["mining_s"], 4),
                       "runs_s": orc["runs_s"], "threads": threads, "frequent": orc["frequent"],
                       "what": "this build's CPU oracle (oracle/oracle.cpp), capped mining of the same job"}}

    line = {
        "metric": "patterns mined/sec + MNI verifications/sec, batched+learned vs plain GPU",
        "value": round(value, 3),
        "unit": "patterns/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * learn_s / K, 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded Chung-Lu generator, t-fsm_amd/host/gen.h)",
        "config": {"workload": WORKLOAD["label"], "vertices": WORKLOAD["vertices"],
                   "edges": WORKLOAD["edges"], "freq": WORKLOAD["freq"], "max_nodes": WORKLOAD["max_nodes"],
                   "path": "batched+learned", "ranker": "t-fsm_amd/rankers/powerlaw.f32 (trained on seeds 11, 12)",
                   "tuning": tuning, "host_threads": threads},
        "verifications_per_s": round(agg["verdicts"] / learn_s, 1),
        "frequent_patterns_per_s": round(nfreq * K / learn_s, 3),
        "patterns_evaluated_per_step": evaluated,
        "frequent_patterns": nfreq,
        "searches_per_step": agg["searches"] / K,
        "pattern_set_matches_reference": bool(learn_ok and batch_ok and plain_ok and roof_ok),
        # every level of the value path ran device-driven as one replayed hipGraph launch
        "level_graph_calls_per_step": agg["graph_calls"] / K,
        "eval_calls_per_step": agg["eval_calls"] / K,
        "plain": side(plain_agg, plain_s, plain_value),
        "batched": side(batch_agg, batch_s, batch_value),
        # device wall clock: first search block's start to the last search wave's end, per level
        "search_ms_per_step": round(agg["search_ms"] / K, 3),
        # the auxiliary-table phase of the same path (k_aux_build after the device-side plan, on the
        # device's wall clock), and the table build's algorithmic bytes (neighbour ids scanned, segment
        # bounds read, row pairs and entries written)
        "table_phase": {"ms_per_step": round(agg["table_ms"] / K, 3),
                        "bytes_per_step": int(agg["bytes_table"] / K),
                        "GB_per_s": round(agg["bytes_table"] / (agg["table_ms"] * 1e-3) / 1e9, 2)
                        if agg["table_ms"] > 0 else None},
        "speedup_vs_plain": round(value / plain_value, 3) if plain_value else None,
        "speedup_batched_vs_plain": round(batch_value / plain_value, 3) if plain_value and batch_value else None,
        "roofline": {"bound": "hbm", "kernel": "k_verify (existence search)", "achieved": round(achieved, 3),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 6),
                     "traffic": traffic,
                     "traffic_source": (f"{traffic_src}: rocprofv3 --pmc FETCH_SIZE (x 2.0) and WRITE_SIZE (x 1.0), "
                                        "mean per k_verify dispatch of the same path") if traffic else None,
                     "launches": ragg["search_launches"] // rk,
                     "avg_launch_ms": round(ms_search, 4), "algorithmic_bytes_per_launch": int(bytes_per_launch),
                     "timing": (f"HIP events on the library's stream around the search launches of {rk} learned "
                                "runs with the same device-driven sequence launched kernel by kernel")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
