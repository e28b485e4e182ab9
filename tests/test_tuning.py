"""Tuning files (t-fsm_amd/autotune.py's output) through the loader that
`run -tune` uses: known options land in the options struct, anything else is
an error."""
import json

import pytest

import tfsm


def write(tmp_path, obj, name="t.json"):
    p = tmp_path / name
    p.write_text(obj if isinstance(obj, str) else json.dumps(obj))
    return str(p)


def test_options_loaded(tmp_path):
    p = write(tmp_path, {"workload": {"freq": 9000}, "options": {"grid_waves_per_cu": 8, "split_steps": 4294967295},
                         "measured": [{"options": {"grid_waves_per_cu": 12}}]})
    assert tfsm.load_tuning(p) == {"grid_waves_per_cu": 8, "split_steps": 4294967295}


def test_shipped_tuning_files_load():
    import glob
    import os
    files = glob.glob(os.path.join(tfsm.HERE, "tuning", "*.json"))
    for f in files:
        opts = tfsm.load_tuning(f)
        assert set(opts) <= {"grid_waves_per_cu", "split_steps", "split_fanout", "long_steps", "group_lanes",
                             "steps_per_round", "refill_min", "refill_max", "prune_batch"}
        assert opts.get("refill_min", 128) <= opts.get("refill_max", 256)


@pytest.mark.parametrize("bad", [
    {"options": {"grid_waves": 8}},            # unknown name
    {"options": {"split_steps": -1}},          # not an unsigned integer
    {"workload": {}},                          # no options
    "not json {",
])
def test_bad_files(tmp_path, bad):
    with pytest.raises(tfsm.TfsmError):
        tfsm.load_tuning(write(tmp_path, bad))


def test_missing_file(tmp_path):
    with pytest.raises(tfsm.TfsmError):
        tfsm.load_tuning(str(tmp_path / "none.json"))
