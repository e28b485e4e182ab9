"""The `run` command line on the GPU against the reference's own output.

`t-fsm_amd/_build/run` takes the reference's flags and writes
grami_results.txt in the serial build's format (serial/grami.h:622-635).  Its
file, as a set of pattern blocks, must equal the file the reference's serial
build wrote for the same graph and flags (tests/golden/*.grami_results.txt,
made by tests/golden/make_golden.py), for the plain path, the batched path,
and the batched path loading the pruned graph through the binary CSR cache
(written on the first run, read on the second).
"""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, PKG, golden_manifest, workload_path

pytestmark = pytest.mark.gpu

RUN = os.path.join(PKG, "_build", "run")


def blocks(path):
    """The pattern blocks of a grami_results.txt, each a tuple of its lines, sorted."""
    out, cur = [], None
    for line in open(path):
        line = line.strip()
        if line.startswith("====="):
            if cur:
                out.append(tuple(cur))
            cur = []
        elif line:
            cur.append(line)
    if cur:
        out.append(tuple(cur))
    return sorted(out)


@pytest.mark.parametrize("name,workload,seed", [("readme_f2_m4", "readme", 0), ("gnm2k_f150_m4", "gnm2k", 1)])
@pytest.mark.parametrize("flags", [[], ["-gpu", "batched"], ["-gpu", "batched", "-csrCache", "1"]],
                         ids=["plain", "batched", "batched-cache"])
def test_run_matches_reference_results(name, workload, seed, flags, workdir, tmp_path):
    m = golden_manifest()[name]
    lg = str(tmp_path / "g.lg")
    shutil.copy(workload_path(workload, seed, workdir), lg)
    want = blocks(os.path.join(GOLDEN, name + ".grami_results.txt"))
    runs = 2 if "-csrCache" in flags else 1
    for i in range(runs):
        out = str(tmp_path / f"grami_results_{i}.txt")
        r = subprocess.run([RUN, "-file", lg, "-freq", str(m["freq"]), "-thread", "4", "-maxNodes",
                            str(m["max_nodes"]), "-out", out] + flags, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        assert f"[INFO] # Frequent Patterns: {m['reported_total']}" in r.stdout, r.stdout
        assert blocks(out) == want, f"run {i}: results differ from the reference's"
    if runs == 2:
        assert os.path.exists(lg + f".f{m['freq']}.csr")
