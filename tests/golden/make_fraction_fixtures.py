"""Fixtures from the reference's own Fraction-Score loader (lyuheng/T-FSM
fraction-score/graph.h): the per-vertex scores of compute_fraction_score
(graph.h:895-922) and the size-2 entries that survive its Fraction-Score
filter (graph.h:789-826) with the sums it compares.

Runs oracle/_ref/ref_fraction_dump (tests/golden/ref_fraction_dump.cpp built
against the reference sources by `make -f oracle/ref.mk`) on seeded graphs
written by the product's generator, and writes tests/golden/frac_<name>.txt:
    N <vertices>
    H <entries> <sha256 of the dump's S lines>     (every score, bit for bit)
    S ...                                          the first 64 S lines
    E ...                                          every surviving size-2 entry
The CPU oracle is pinned to these files by tests/test_fraction.py.

    python tests/golden/make_fraction_fixtures.py [name ...]
"""
import hashlib
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(REPO, "oracle", "_ref", "ref_fraction_dump")
GEN = os.path.join(REPO, "t-fsm_amd", "_build", "tfsm_gen")

# name: (workload, seed, freq)
FIXTURES = {
    "readme_f1": ("readme", 0, 1),
    "gnm2k_f40": ("gnm2k", 1, 40),
    "powerlaw100k_f6000": ("powerlaw100k", 2, 6000),
    # 16 vertex labels and 4 edge labels: scores merge four edge-label ranges
    "rmat20_f5300": ("rmat20", 4, 5300),
}
KEEP_S = 64


def main(names):
    if not os.path.exists(EXE):
        sys.exit(f"{EXE} is missing: run `make -f oracle/ref.mk` first")
    with tempfile.TemporaryDirectory() as tmp:
        for name in names or FIXTURES:
            workload, seed, freq = FIXTURES[name]
            if workload == "readme":
                lg = os.path.join(HERE, "readme_sample.lg")
            else:
                lg = os.path.join(tmp, f"{workload}_{seed}.lg")
                if not os.path.exists(lg):
                    subprocess.run([GEN, workload, lg, str(seed)], check=True, stdout=subprocess.DEVNULL)
            dump = os.path.join(tmp, name + ".txt")
            subprocess.run([EXE, lg, str(freq), dump], cwd=tmp, check=True, stdout=subprocess.DEVNULL)
            lines = open(dump).read().split("\n")
            n_line = [l for l in lines if l.startswith("N ")]
            s_lines = [l for l in lines if l.startswith("S ")]
            e_lines = [l for l in lines if l.startswith("E ")]
            digest = hashlib.sha256("".join(l + "\n" for l in s_lines).encode()).hexdigest()
            with open(os.path.join(HERE, f"frac_{name}.txt"), "w") as f:
                f.write(f"# {workload} seed {seed} freq {freq}: tests/golden/make_fraction_fixtures.py\n")
                f.write(n_line[0] + "\n")
                f.write(f"H {len(s_lines)} {digest}\n")
                for l in s_lines[:KEEP_S] + e_lines:
                    f.write(l + "\n")
            print(name, n_line[0], len(s_lines), "scores,", len(e_lines), "size-2 entries")


if __name__ == "__main__":
    main(sys.argv[1:])
