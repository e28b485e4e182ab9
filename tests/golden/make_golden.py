"""Writes the golden pattern-set fixtures from the reference T-FSM binaries.

Runs oracle/_ref/run_serial, run_parallel and run_fraction (built from the
reference sources by `make -f oracle/ref.mk`) on the seeded workloads written
by t-fsm_amd/_build/tfsm_gen, and stores each frequent-pattern set as sorted
DFS-code keys "from,to,from_label,edge_label,to_label;..." (one pattern per
line) in tests/golden/<name>.keys, with tests/golden/manifest.json recording
the command, the reference's own reported count and a sha256 of the keys.

    python tests/golden/make_golden.py [name ...]
    python tests/golden/make_golden.py --oracle-check name [threads]

The second form cross-checks a fixture with this build's CPU oracle
(oracle/oracle.cpp, capped mining with the given threads): it mines the same
configuration, compares the oracle's pattern set (restricted to what the
fixture holds) and total with the fixture's, and records the outcome under
"oracle_check" in the fixture's manifest entry.  Fixtures from the parallel
build, which can drop patterns on its timeout path (parallel/comper.h:1450),
are pinned this way where the serial build takes too long.

The serial build writes grami_results.txt (serial/grami.h:622-635) with every
frequent pattern.  The parallel build prints patterns of three or more
vertices between `~~~~~` / `#####` markers (parallel/comper.h:1027-1031,
1095-1101) and only counts the size-2 patterns, so its fixtures hold the
patterns of three or more vertices plus the reported total.  The
Fraction-Score build (oracle/_ref/run_fraction, from fraction-score/) prints
as the parallel build does.  Every reference run's console output is kept in
tests/golden/_raw/<name>.stdout (not committed) before it is parsed: with
several threads the parallel build can interleave its pattern lines, and the
parse then stops with the line it could not read.
"""
import fcntl
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
GEN = os.path.join(REPO, "t-fsm_amd", "_build", "tfsm_gen")
REF = os.path.join(REPO, "oracle", "_ref")

# name: (workload, seed, freq, maxNodes, reference build, threads)
CONFIGS = {
    "readme_f1_m3": ("readme", 0, 1, 3, "serial", 1),
    "readme_f2_m4": ("readme", 0, 2, 4, "serial", 1),
    "gnm2k_f150_m4": ("gnm2k", 1, 150, 4, "serial", 1),
    "gnm2k_f80_m4": ("gnm2k", 1, 80, 4, "serial", 1),
    "gnm2k_f40_m4": ("gnm2k", 1, 40, 4, "serial", 1),
    "powerlaw100k_f9000_m5": ("powerlaw100k", 2, 9000, 5, "serial", 1),
    "powerlaw100k_f6000_m6": ("powerlaw100k", 2, 6000, 6, "parallel", 8),
    # the same configuration from the serial build: its complete list checks
    # the parallel build's (which can lose patterns on its timeout path)
    "powerlaw100k_f6000_m6_serial": ("powerlaw100k", 2, 6000, 6, "serial", 1),
    # R-MAT scale 20 (16 vertex labels, 4 edge labels): at freq 20000 the
    # reference finds nothing (no single-edge type reaches 9200 distinct
    # endpoints per side), so the fixture uses freq 9000
    "rmat20_f9000_m5": ("rmat20", 4, 9000, 5, "serial", 1),
    "rmat20_f20000_m5": ("rmat20", 4, 20000, 5, "serial", 1),
    # freq 5300: with uniform labels the single-edge supports lie in a narrow
    # band (8.2K-9.2K) and size-3 supports fall off sharply between 5200
    # (1203 frequent) and 5400 (1); at 5300, 114 three-vertex patterns are
    # frequent and their 11.8K children are evaluated, none frequent.  Seed 5
    # is the graph the learned order is tested on (not trained on).
    "rmat20_f5300_m5": ("rmat20", 4, 5300, 5, "serial", 1),
    "rmat20s5_f5300_m5": ("rmat20", 5, 5300, 5, "serial", 1),
    # the same two from the parallel build (patterns of three or more
    # vertices), which finishes in minutes where the serial one takes hours
    "rmat20_f5300_m5_parallel": ("rmat20", 4, 5300, 5, "parallel", 2),
    "rmat20s5_f5300_m5_parallel": ("rmat20", 5, 5300, 5, "parallel", 2),
    # seed 5 mined by this build's CPU oracle (8 threads, every frequent
    # pattern): the reference's parallel run printed its two threads' pattern
    # lines into each other, and its serial run takes longer than a day
    "rmat20s5_f5300_m5_oracle": ("rmat20", 5, 5300, 5, "oracle", 8),
    # Fraction-Score support (the reference's fraction-score build, which
    # prints like the parallel build; one thread keeps its output lines whole)
    "readme_frac_f1_m3": ("readme", 0, 1, 3, "fraction", 1),
    "gnm2k_frac_f40_m4": ("gnm2k", 1, 40, 4, "fraction", 1),
    "gnm2k_frac_f60_m4": ("gnm2k", 1, 60, 4, "fraction", 1),
    "powerlaw100k_frac_f3000_m5": ("powerlaw100k", 2, 3000, 5, "fraction", 1),
}


def key_of(vl, es):
    return "".join(f"{a},{b},{vl[a]},{el},{vl[b]};" for a, b, el in es)


def parse_serial(path):
    keys, vl, es = [], {}, []
    for line in open(path):
        t = line.split()
        if not t:
            continue
        if t[0].startswith("====="):
            if es:
                keys.append(key_of(vl, es))
            vl, es = {}, []
        elif t[0] == "v":
            vl[int(t[1])] = int(t[2])
        elif t[0] == "e":
            es.append((int(t[1]), int(t[2]), int(t[3])))
    if es:
        keys.append(key_of(vl, es))
    return keys


def parse_parallel(text):
    keys, cur = [], None
    for line in text.split("\n"):
        if line in ("~~~~~", "#####"):
            if cur is None:
                cur = []
            else:
                vl, es = {}, []
                for l in cur:
                    t = l.split()
                    if len(t) == 3 and t[0] == "v":
                        vl[int(t[1])] = int(t[2])
                    elif len(t) == 4 and t[0] == "e":
                        es.append((int(t[1]), int(t[2]), int(t[3])))
                    else:
                        raise ValueError(f"interleaved output line {l!r}")
                keys.append(key_of(vl, es))
                cur = None
        elif cur is not None:
            cur.append(line)
    return keys


def graph_file(workload, seed, tmp):
    if workload == "readme":
        return os.path.join(HERE, "readme_sample.lg")
    path = os.path.join(tmp, f"{workload}_{seed}.lg")
    if not os.path.exists(path):
        subprocess.run([GEN, workload, path, str(seed)], check=True, stdout=subprocess.DEVNULL)
    return path


def update_manifest(name, entry):
    manifest_path = os.path.join(HERE, "manifest.json")
    # several invocations may run at once: merge into the file under a lock
    with open(manifest_path + ".lock", "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        disk = json.load(open(manifest_path)) if os.path.exists(manifest_path) else {}
        disk[name] = entry
        with open(manifest_path, "w") as f:
            json.dump(disk, f, indent=1, sort_keys=True)


def oracle_check(name, threads):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    manifest = json.load(open(os.path.join(HERE, "manifest.json")))
    entry = manifest[name]
    want = set(l.strip() for l in open(os.path.join(HERE, name + ".keys")) if l.strip())
    with tempfile.TemporaryDirectory() as tmp:
        lg = graph_file(entry["workload"], entry["seed"], tmp)
        t0 = time.time()
        og = orc.OracleGraph(lg, entry["freq"])
        out = os.path.join(tmp, "oracle.keys")
        og.mine(entry["freq"], entry["max_nodes"], threads, 0, out)
        got = set(l.rstrip("\n").split("\t")[0] for l in open(out) if l.strip())
        seconds = round(time.time() - t0, 2)
    held = got if entry["keys_hold"] == "all frequent patterns" else set(k for k in got if k.count(";") > 1)
    entry["oracle_check"] = dict(threads=threads, seconds=seconds, total=len(got), keys=len(held),
                                 equal=held == want and len(got) == entry["reported_total"])
    print(name, entry["oracle_check"])
    update_manifest(name, entry)
    return 0 if entry["oracle_check"]["equal"] else 1


def keep_raw(name, text):
    raw = os.path.join(HERE, "_raw")
    os.makedirs(raw, exist_ok=True)
    with open(os.path.join(raw, name + ".stdout"), "w") as f:
        f.write(text)


def main(names):
    manifest_path = os.path.join(HERE, "manifest.json")
    manifest = json.load(open(manifest_path)) if os.path.exists(manifest_path) else {}
    with tempfile.TemporaryDirectory() as tmp:
        for name in names or CONFIGS:
            workload, seed, freq, maxn, build, threads = CONFIGS[name]
            lg = graph_file(workload, seed, tmp)
            t0 = time.time()
            if build == "oracle":
                sys.path.insert(0, os.path.join(REPO, "oracle"))
                import oracle as orc
                og = orc.OracleGraph(lg, freq)
                og.mine(freq, maxn, threads, 0, os.path.join(tmp, "oracle.keys"))
                keys = [l.rstrip("\n").split("\t")[0] for l in open(os.path.join(tmp, "oracle.keys")) if l.strip()]
                reported = len(set(keys))
            elif build == "serial":
                cmd = [os.path.join(REF, "run_serial"), "-file", lg, "-freq", str(freq), "-maxNodes", str(maxn)]
                out = subprocess.run(cmd, cwd=tmp, check=True, capture_output=True, text=True).stdout
                keep_raw(name, out)
                keep_raw(name + ".grami_results", open(os.path.join(tmp, "grami_results.txt")).read())
                keys = parse_serial(os.path.join(tmp, "grami_results.txt"))
                if workload in ("readme", "gnm2k"):
                    # the reference's own output file, for the CLI output test
                    shutil.copy(os.path.join(tmp, "grami_results.txt"),
                                os.path.join(HERE, name + ".grami_results.txt"))
                reported = int([l for l in out.split("\n") if "# Frequent Patterns" in l][0].split(":")[1])
            else:
                exe = "run_fraction" if build == "fraction" else "run_parallel"
                cmd = [os.path.join(REF, exe), "-file", lg, "-freq", str(freq), "-thread", str(threads),
                       "-maxNodes", str(maxn)]
                out = subprocess.run(cmd, cwd=tmp, check=True, capture_output=True, text=True).stdout
                keep_raw(name, out)
                keys = parse_parallel(out)
                reported = int([l for l in out.split("\n") if "[INFO] # Frequent Patterns" in l][0].split(":")[1])
            keys = sorted(set(keys))
            body = "".join(k + "\n" for k in keys)
            with open(os.path.join(HERE, name + ".keys"), "w") as f:
                f.write(body)
            manifest[name] = dict(workload=workload, seed=seed, freq=freq, max_nodes=maxn, reference_build=build,
                                  threads=threads, reported_total=reported, keys=len(keys),
                                  keys_hold="all frequent patterns" if build in ("serial", "oracle") else
                                  "frequent patterns with three or more vertices",
                                  sha256=hashlib.sha256(body.encode()).hexdigest(),
                                  seconds=round(time.time() - t0, 2))
            print(name, manifest[name])
            update_manifest(name, manifest[name])


if __name__ == "__main__":
    if sys.argv[1:2] == ["--oracle-check"]:
        sys.exit(oracle_check(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 8))
    main(sys.argv[1:])
