"""Fixtures from the reference's own matching engine (lyuheng/T-FSM
parallel/gmatch.h): filtered domains, auxiliary tables, GQL orders and
backward-neighbour lists of a set of patterns, as the reference computes them
at pattern activation (parallel/comper.h:1655-1685; no push-down pruning).

Compiles tests/golden/ref_gmatch_dump.cpp against the reference sources where
they lie (/root/reference/parallel; g++, the reference's -std=c++14 -fopenmp,
-march=x86-64-v3 for its AVX2 intersection code) into
oracle/_ref/ref_gmatch_dump, runs it on seeded graphs written by the
product's generator, and writes tests/golden/gmatch_<name>.txt.  The CPU
oracle is pinned to these files by tests/test_oracle_pinning.py.

    python tests/golden/make_gmatch_fixtures.py
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/parallel"
OUT = os.path.join(REPO, "oracle", "_ref")
GEN = os.path.join(REPO, "t-fsm_amd", "_build", "tfsm_gen")
for _p in ("tests", "oracle", "t-fsm_amd"):
    sys.path.insert(0, os.path.join(REPO, _p))

# name: (workload, seed, freq the graph is pruned with, golden key file, patterns)
FIXTURES = {
    "readme_f1": ("readme", 0, 1, "readme_f1_m3", 20),
    "gnm2k_f40": ("gnm2k", 1, 40, "gnm2k_f40_m4", 24),
}


def build():
    os.makedirs(OUT, exist_ok=True)
    exe = os.path.join(OUT, "ref_gmatch_dump")
    subprocess.run(["g++", "-std=c++14", "-O2", "-march=x86-64-v3", "-fopenmp", "-w", "-I", REF, "-o", exe,
                    os.path.join(HERE, "ref_gmatch_dump.cpp"), os.path.join(REF, "intersection",
                                                                          "computesetintersection.cpp"),
                    "-lpthread"], check=True)
    return exe


def key_of(vl, es):
    return "".join(f"{a},{b},{vl[a]},{el},{vl[b]};" for a, b, el in es)


def patterns(golden, limit):
    """Golden patterns of three or more vertices plus a cycle-closing and a
    relabelled variant of each (the GPU parity tests' variants)."""
    from test_gpu_parity import variants
    keys = set(l.strip() for l in open(os.path.join(HERE, golden + ".keys")) if l.strip())
    return [key_of(vl, es) for vl, es in variants(keys, limit)]


def main():
    exe = build()
    with tempfile.TemporaryDirectory() as tmp:
        for name, (workload, seed, freq, golden, limit) in FIXTURES.items():
            if workload == "readme":
                lg = os.path.join(HERE, "readme_sample.lg")
            else:
                lg = os.path.join(tmp, f"{workload}_{seed}.lg")
                subprocess.run([GEN, workload, lg, str(seed)], check=True, stdout=subprocess.DEVNULL)
            spec = os.path.join(tmp, name + ".spec")
            with open(spec, "w") as f:
                f.write("".join(k + "\n" for k in patterns(golden, limit)))
            out = os.path.join(HERE, f"gmatch_{name}.txt")
            with open(out + ".tmp", "w") as f:
                f.write(f"# reference gmatch.h dump: {workload} seed {seed}, graph pruned at freq {freq}\n")
            subprocess.run([exe, lg, str(freq), spec, out + ".body"], check=True)
            with open(out + ".tmp", "a") as f, open(out + ".body") as b:
                f.write(b.read())
            os.remove(out + ".body")
            os.replace(out + ".tmp", out)
            print(name, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
