"""CPU check of the lane-per-task existence search (t-fsm_amd/csrc/lane_dfs.h).

The search kernel and this check compile the same search source; here it
runs on the CPU over the oracle's filtered domains, auxiliary tables and
matching orders, with every access bounds-checked (and the harness built with
the host address and undefined-behaviour sanitizers).  In the reference
matching order and in the adaptive one, every verdict must equal the oracle's
exact-mode verdict and every embedding found must be injective and use only
table edges; with pieces donated (every step or every few steps, one or
several pieces per donation), the verdict combined over a task's pieces must
be the same.
"""
import os
import subprocess

import pytest

from conftest import GOLDEN, ORACLE, REPO, golden_keys, workload_path
from test_wave_dfs import variants

HARNESS_SRC = os.path.join(REPO, "tests", "lane_dfs_check.cpp")


def build_harness(tmp_path_factory, me):
    lib = os.path.join(ORACLE, "_build", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", ORACLE], check=True, stdout=subprocess.DEVNULL)
    out = str(tmp_path_factory.mktemp("lane") / f"lane_dfs_check_{me}")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-Wno-unknown-pragmas", "-fsanitize=address,undefined,bounds",
                    "-fno-sanitize-recover=all", f"-DLANE_ME={me}", "-o", out, HARNESS_SRC, lib,
                    f"-Wl,-rpath,{os.path.dirname(lib)}"], check=True)
    return out


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    return build_harness(tmp_path_factory, 12)


@pytest.fixture(scope="module")
def harness8(tmp_path_factory):
    """The search instance for launches whose patterns have at most 8 edges
    (per-edge arrays of 8; patterns with more run in the reference order)."""
    return build_harness(tmp_path_factory, 8)


DONATION = [(0, 1), (1, 1), (2, 3)]     # (donate every N steps (0: never), pieces per donation)


def run(harness, path, freq, samples, pats, tmp, donate=(0, 1), group=1):
    spec = os.path.join(str(tmp), "spec.txt")
    with open(spec, "w") as f:
        f.write(f"{path} {freq} {samples}\n")
        for vl, es in pats:
            f.write(" ".join(str(x) for x in [len(vl)] + list(vl) + [len(es)] + [y for e in es for y in e]) + "\n")
    r = subprocess.run([harness, spec, str(donate[0]), str(donate[1]), str(group)], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    f = r.stdout.split()
    out = dict(zip(f[0::2], (int(x) for x in f[1::2])))
    assert out["patterns"] > 0 and out["checked"] > 0, r.stdout
    assert donate[0] > 0 or out["pieces"] == 0, r.stdout
    return out


@pytest.mark.parametrize("donate", DONATION)
@pytest.mark.parametrize("golden,freq", [("readme_f1_m3", 1), ("readme_f2_m4", 2)])
def test_readme(harness, golden, freq, donate, tmp_path):
    path = os.path.join(GOLDEN, "readme_sample.lg")
    run(harness, path, freq, 0, variants(golden_keys(golden), 40), tmp_path, donate)


@pytest.mark.parametrize("donate", DONATION)
@pytest.mark.parametrize("golden,freq", [("gnm2k_f80_m4", 80), ("gnm2k_f40_m4", 40)])
def test_gnm(harness, golden, freq, donate, workdir, tmp_path):
    path = workload_path("gnm2k", 1, workdir)
    out = run(harness, path, freq, 0, variants(golden_keys(golden), 200), tmp_path, donate)
    assert (out["pieces"] > 0) == (donate[0] > 0), out


@pytest.mark.parametrize("me", [12, 8])
@pytest.mark.parametrize("donate", [(0, 1), (16, 2)])
def test_powerlaw_sampled(harness, harness8, me, workdir, tmp_path, donate):
    """Four- and five-vertex patterns of the power-law graph, 40 sampled
    candidates per vertex; the adaptive order takes far fewer steps."""
    path = workload_path("powerlaw100k", 2, workdir)
    pats = [p for p in variants(golden_keys("powerlaw100k_f9000_m5"), 60) if len(p[0]) >= 4][:14]
    out = run(harness if me == 12 else harness8, path, 9000, 40, pats, tmp_path, donate)
    assert out["steps_adaptive"] < out["steps_reference"], out
    assert (out["pieces"] > 0) == (donate[0] > 0), out


@pytest.mark.parametrize("seed,instances,donate,me", [(1, 30, (0, 1), 12), (2, 30, (0, 1), 12), (3, 6, (16, 4), 12),
                                                      (4, 4, (4, 2), 12), (6, 30, (0, 1), 8), (7, 6, (16, 4), 8)])
def test_random(harness, harness8, seed, instances, donate, me):
    """Synthetic tables with empty, short and long rows, against a plain
    recursive search written in the harness: the same verdicts in both orders,
    and in the reference order without pieces the same candidates entered in
    the same order (so a cursor that skips or repeats a table entry shows
    up)."""
    r = subprocess.run([harness if me == 12 else harness8, "--random", str(seed), str(instances), str(donate[0]),
                        str(donate[1])], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    f = r.stdout.split()
    out = dict(zip(f[1::2], (int(x) for x in f[2::2])))
    assert out["checked"] > 1000 and 0 < out["found"] < out["checked"], r.stdout
    assert (out["pieces"] > 0) == (donate[0] > 0), r.stdout


@pytest.mark.parametrize("group,donate", [(2, (0, 1)), (4, (0, 1)), (8, (3, 2)), (64, (0, 1))])
def test_lane_groups(harness, group, donate, workdir, tmp_path):
    """Root tasks run as `group` interleaved lanes (lane p of the group tests
    depth-1 candidates p, p + group, ...), each free to donate pieces: the
    combined verdict equals the oracle's exact verdict, on G(2000, 8000) and
    on sampled power-law candidates."""
    out = run(harness, workload_path("gnm2k", 1, workdir), 40, 0, variants(golden_keys("gnm2k_f40_m4"), 120),
              tmp_path, donate, group)
    assert out["mismatches"] == 0 and (out["pieces"] > 0) == (donate[0] > 0), out
    pats = [p for p in variants(golden_keys("powerlaw100k_f9000_m5"), 60) if len(p[0]) >= 4][:8]
    out = run(harness, workload_path("powerlaw100k", 2, workdir), 9000, 20, pats, tmp_path, donate, group)
    assert out["mismatches"] == 0, out


@pytest.mark.parametrize("group", [2, 4, 16])
def test_random_lane_groups(harness, group):
    """Interleaved lane groups on synthetic tables against the plain recursive search."""
    r = subprocess.run([harness, "--random", "5", "20", "0", "1", str(group)], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    f = r.stdout.split()
    out = dict(zip(f[1::2], (int(x) for x in f[2::2])))
    assert out["checked"] > 1000 and out["mismatches"] == 0 and 0 < out["found"] < out["checked"], r.stdout


@pytest.mark.parametrize("reference_order,me", [(0, 12), (1, 12), (0, 8)], ids=["adaptive", "reference", "adaptive-me8"])
def test_level_record_queue(harness, harness8, reference_order, me, workdir, tmp_path):
    """Several power-law patterns in one launch, as the GPU parity test of a
    whole level with split_steps=64 and four records per donation runs them:
    256 virtual lanes take root tasks of every pattern, then records from one
    FIFO (donated every 64 steps once the root tasks are out and no record
    waits), each record starting a piece of its own task's pattern from the
    prefix it carries; per-task piece counters give the verdicts.  With
    every access bounds-checked (and the sanitizers on), the verdicts equal
    the oracle's exact ones and records are used."""
    path = workload_path("powerlaw100k", 2, workdir)
    pats = [p for p in variants(golden_keys("powerlaw100k_f9000_m5"), 40) if len(p[0]) >= 4][7:14]
    spec = os.path.join(str(tmp_path), "level.spec")
    with open(spec, "w") as f:
        f.write(f"{path} 9000 0\n")
        for vl, es in pats:
            f.write(" ".join(str(x) for x in [len(vl)] + list(vl) + [len(es)] + [y for e in es for y in e]) + "\n")
    r = subprocess.run([harness if me == 12 else harness8, "--level", spec, "64", "4", "256", "1",
                        str(reference_order)], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout + r.stderr
    f = r.stdout.split()
    out = dict(zip(f[0::2], (int(x) for x in f[1::2])))
    assert out["level_patterns"] >= 5 and out["mismatches"] == 0 and out["records"] > 0, r.stdout
