"""The CPU oracle (with the host pattern growth) reproduces the reference
binaries' frequent-pattern sets exactly, in both support modes."""
import os

import pytest

import oracle as orc
from conftest import golden_keys, golden_manifest, workload_path

CASES = ["readme_f1_m3", "readme_f2_m4", "gnm2k_f150_m4", "gnm2k_f80_m4", "gnm2k_f40_m4",
         "powerlaw100k_f9000_m5"]


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("mode", [0, 1])
def test_oracle_pattern_sets(name, mode, workdir):
    m = golden_manifest()[name]
    path = workload_path(m["workload"], m["seed"], workdir)
    og = orc.OracleGraph(path, m["freq"])
    out = os.path.join(str(workdir), f"oracle_{name}_{mode}.keys")
    n, _ = og.mine(m["freq"], m["max_nodes"], 8, mode, out)
    keys = {}
    for line in open(out):
        k, lo, hi = line.rstrip("\n").split("\t")
        keys[k] = (int(lo), int(hi))
    assert n == m["reported_total"]
    assert set(keys) == golden_keys(name)
    for k, (lo, hi) in keys.items():
        assert lo >= m["freq"], k
        if mode == 1:
            assert lo == hi, k              # exact mode: the support is the MNI itself


def test_parallel_fixture_counts():
    """The parallel-build fixture lists the patterns of three or more vertices;
    with the size-2 count it adds up to the reference's reported total."""
    m = golden_manifest()["powerlaw100k_f6000_m6"]
    keys = golden_keys("powerlaw100k_f6000_m6")
    assert len(keys) == m["keys"] == 330
    assert all(k.count(";") >= 2 for k in keys)
    assert m["reported_total"] == 334


def test_parallel_fixture_matches_serial():
    """The serial build mined the same configuration (all frequent patterns,
    written by serial/grami.h:622-635): its patterns of three or more
    vertices are exactly the parallel build's fixture, and its total is the
    parallel build's reported total, so that fixture lost nothing on the
    parallel build's timeout path."""
    serial = golden_keys("powerlaw100k_f6000_m6_serial")
    parallel = golden_keys("powerlaw100k_f6000_m6")
    m = golden_manifest()
    assert len(serial) == m["powerlaw100k_f6000_m6_serial"]["reported_total"] == 334
    assert {k for k in serial if k.count(";") >= 2} == parallel
    assert len(serial) == m["powerlaw100k_f6000_m6"]["reported_total"]
