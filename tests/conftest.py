import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "t-fsm_amd")
ORACLE = os.path.join(REPO, "oracle")
GOLDEN = os.path.join(REPO, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def workdir(tmp_path_factory):
    return tmp_path_factory.mktemp("tfsm")


_graphs = {}


def workload_path(name, seed, workdir):
    """Seeded `.lg` file written by the product's generator (cached per session)."""
    if name == "readme":
        return os.path.join(GOLDEN, "readme_sample.lg")
    key = (name, seed)
    if key not in _graphs:
        path = os.path.join(str(workdir), f"{name}_{seed}.lg")
        subprocess.run([os.path.join(PKG, "_build", "tfsm_gen"), name, path, str(seed)], check=True,
                       stdout=subprocess.DEVNULL)
        _graphs[key] = path
    return _graphs[key]


def golden_keys(name):
    with open(os.path.join(GOLDEN, name + ".keys")) as f:
        return set(l.strip() for l in f if l.strip())


def golden_manifest():
    import json
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)
