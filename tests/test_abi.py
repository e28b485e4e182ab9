"""The C-ABI library loads on a host without a GPU and exports every entry
point include/tfsm_gpu.h declares (no compute is called)."""
import ctypes
import os
import re
import subprocess

import pytest

import tfsm
from conftest import REPO


def declared_functions():
    text = open(os.path.join(REPO, "include", "tfsm_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tfsm_[a-z_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 12
    lib = ctypes.CDLL(os.path.join(tfsm.BUILD, "libtfsm_gpu.so"))
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(tfsm.GPU_SYMBOLS) == names


def test_status_strings_and_null_arguments():
    lib = tfsm.gpu_lib()
    assert lib.tfsm_status_string(0) == b"ok"
    assert lib.tfsm_status_string(-5) == b"unknown domain handle"
    # argument checks run before any device call
    assert lib.tfsm_eval_pattern(None, None, 0, 1, None, None) == -1
    assert lib.tfsm_ctx_destroy(None) == -1


def test_host_library_exports():
    lib = tfsm.host_lib()
    for name in ["tfsmh_load", "tfsmh_mine", "tfsmh_extend", "tfsmh_is_min_code", "tfsmh_ctx_create"]:
        assert hasattr(lib, name)


REFERENCE_PARALLEL = "/root/reference/parallel"


@pytest.mark.skipif(not os.path.isdir(REFERENCE_PARALLEL), reason="the reference sources are not on this host")
def test_reference_binding_compiles_and_links(tmp_path):
    """t-fsm_amd/integration/gpu_eval.h, the binding a maintainer adds to the
    reference's parallel/ tree, compiles against the reference's own headers
    (graph.h's Graph, vertex_t, edge_t, Pattern and dfs_code_t) and links
    against libtfsm_gpu.so.  Nothing is run (no GPU here)."""
    tu = tmp_path / "bind.cpp"
    tu.write_text(
        '#include "graph.h"\n'
        '#include "gpu_eval.h"\n'
        "int main() {\n"
        "    GpuGraph G; Graph g; Pattern p; bool f; uint64_t h;\n"
        "    std::vector<Pattern *> ps{&p}; std::vector<uint64_t> par{0}, kids; std::vector<char> fr;\n"
        "    if (gpu_graph_init(G, g) != TFSM_OK) return 1;\n"
        "    gpu_eval(G, p, 0, 1, &f, &h);\n"
        "    gpu_eval_level(G, ps, par, 1, fr, kids);\n"
        "    gpu_graph_destroy(G);\n"
        "    return 0;\n"
        "}\n")
    out = tmp_path / "bind"
    r = subprocess.run(["g++", "-std=c++14", "-fopenmp", "-Wall", "-Werror=return-type",
                        "-I", REFERENCE_PARALLEL, "-I", os.path.join(REPO, "include"),
                        "-I", os.path.join(REPO, "t-fsm_amd", "integration"), str(tu), "-o", str(out),
                        "-L", tfsm.BUILD, "-ltfsm_gpu", "-Wl,-rpath," + tfsm.BUILD],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_struct_layouts_match_the_header(tmp_path):
    """Every struct the Python binding passes across the boundary has the
    header's size and field offsets (a C program compiled against
    include/tfsm_gpu.h prints them)."""
    structs = {"tfsm_csr": tfsm.Csr, "tfsm_pattern": tfsm.Pattern, "tfsm_eval_options": tfsm.EvalOptions,
               "tfsm_pattern_result": tfsm.PatternResult, "tfsm_call_stats": tfsm.CallStats,
               "tfsm_ranker_weights": tfsm.RankerWeights}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "tfsm_gpu.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'    printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines.append("    return 0;\n}\n")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = str(tmp_path / "layout")
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), "-o", exe, str(src)], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)
