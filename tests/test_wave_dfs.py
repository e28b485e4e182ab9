"""CPU check of the wave-cooperative existence search (t-fsm_amd/csrc/wave_dfs.h).

The device kernel and this check compile the same search source; here it runs
with an emulated 64-lane wave over the oracle's filtered domains, auxiliary
tables and matching orders, with every access bounds-checked (and the harness
built with the host address and undefined-behaviour sanitizers).  Every
verdict must equal the oracle's exact-mode verdict, and every embedding found
must be injective and use only table edges; with donation on (every round or
every few rounds, one or several pieces per donation), the verdict combined
over a task's pieces must be the same.
"""
import os
import subprocess

import pytest

import tfsm
from conftest import GOLDEN, ORACLE, REPO, golden_keys, workload_path

HARNESS_SRC = os.path.join(REPO, "tests", "wave_dfs_check.cpp")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    lib = os.path.join(ORACLE, "_build", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", ORACLE], check=True, stdout=subprocess.DEVNULL)
    out = str(tmp_path_factory.mktemp("wave") / "wave_dfs_check")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-o", out, HARNESS_SRC, lib, f"-Wl,-rpath,{os.path.dirname(lib)}"], check=True)
    return out


def variants(keys, limit):
    """Golden patterns (>= 3 vertices) and cycle-closing / relabelled variants."""
    out = []
    for k in sorted(keys):
        vl, es = tfsm.parse_key(k)
        if len(vl) < 3:
            continue
        out.append((vl, es))
        adj = {(a, b) for a, b, _ in es} | {(b, a) for a, b, _ in es}
        if (0, len(vl) - 1) not in adj:
            out.append((vl, es + [(0, len(vl) - 1, es[0][2])]))
        if vl[-1] != vl[0]:
            out.append((vl[:-1] + [vl[0]], es))
        if len(out) >= limit:
            break
    return out[:limit]


DONATION = [(0, 1), (1, 1), (3, 4)]     # (donate every N rounds (0: never), pieces per donation)


def run(harness, path, freq, samples, pats, tmp, donate=(0, 1)):
    spec = os.path.join(str(tmp), "spec.txt")
    with open(spec, "w") as f:
        f.write(f"{path} {freq} {samples}\n")
        for vl, es in pats:
            f.write(" ".join(str(x) for x in [len(vl)] + list(vl) + [len(es)] + [y for e in es for y in e]) + "\n")
    r = subprocess.run([harness, spec, str(donate[0]), str(donate[1])], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    f = r.stdout.split()
    patterns, checked, donations = int(f[1]), int(f[3]), int(f[7])
    assert patterns > 0 and checked > 0, r.stdout
    assert donate[0] > 0 or donations == 0, r.stdout
    return patterns, checked, donations


@pytest.mark.parametrize("donate", DONATION)
@pytest.mark.parametrize("golden,freq", [("readme_f1_m3", 1), ("readme_f2_m4", 2)])
def test_readme(harness, golden, freq, donate, tmp_path):
    path = os.path.join(GOLDEN, "readme_sample.lg")
    run(harness, path, freq, 0, variants(golden_keys(golden), 40), tmp_path, donate)


@pytest.mark.parametrize("donate", DONATION)
@pytest.mark.parametrize("golden,freq", [("gnm2k_f80_m4", 80), ("gnm2k_f40_m4", 40)])
def test_gnm(harness, golden, freq, donate, workdir, tmp_path):
    path = workload_path("gnm2k", 1, workdir)
    _, _, donations = run(harness, path, freq, 0, variants(golden_keys(golden), 200), tmp_path, donate)
    assert (donations > 0) == (donate[0] > 0)


def test_powerlaw_sampled(harness, workdir, tmp_path):
    """Four- and five-vertex patterns with long searches, 10 sampled
    candidates per vertex (the oracle's exact evaluation bounds the count)."""
    path = workload_path("powerlaw100k", 2, workdir)
    pats = [p for p in variants(golden_keys("powerlaw100k_f9000_m5"), 60) if len(p[0]) >= 4][:14]
    _, _, donations = run(harness, path, 9000, 10, pats, tmp_path, (8, 2))
    assert donations > 0


@pytest.mark.parametrize("seed,instances,donate", [(1, 30, (0, 1)), (2, 30, (0, 1)), (3, 8, (3, 4)), (4, 8, (1, 2))])
def test_random_long_rows(harness, seed, instances, donate):
    """Synthetic tables with rows far longer than the 64-entry chunks and
    windows, against a plain recursive search written in the harness: the same
    verdicts, and without pieces the same candidates entered in the same order
    (so a window or cursor that skips or repeats a table entry shows up)."""
    r = subprocess.run([harness, "--random", str(seed), str(instances), str(donate[0]), str(donate[1])],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    f = r.stdout.split()
    checked, found = int(f[2]), int(f[8])
    assert checked > 1000 and 0 < found < checked, r.stdout
