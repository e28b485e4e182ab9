"""The miner's worker pool under ThreadSanitizer: background growth jobs and
waited-for jobs interleaved as the level loop uses them, every item run
exactly once, and no data race reported (CPU only)."""
import os
import subprocess

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(REPO, "tests", "worker_pool_check.cpp")
HOST = os.path.join(REPO, "t-fsm_amd", "host")


def test_worker_pool_tsan(tmp_path):
    exe = str(tmp_path / "worker_pool_check")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=thread", "-pthread", "-I", HOST, "-o", exe, SRC],
                   check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.strip() == "ok"
