"""The C-ABI library loads on a host without a GPU and exports every entry
point include/tfsm_gpu.h declares (no compute is called)."""
import ctypes
import os
import re

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
