"""The C-ABI library loads and exports every entry point include/chkn.h declares (no GPU
work is done here)."""
import ctypes
import os
import re

from chkn import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "chkn.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void|long long|char\s*\*|const char\*)\s*\**\s*(chkn_\w+)\s*\(",
                                 text, flags=re.M)))


def test_every_declared_symbol_is_exported():
    # Both gas-model flavours: ideal air and the reacting A->B gas.
    names = declared_functions()
    assert len(names) >= 15
    for path in (lib.LIB_PATH, lib.LIB_AB_PATH):
        so = ctypes.CDLL(path)
        for n in names:
            assert hasattr(so, n), (path, n)
    assert set(names) <= set(lib.EXPORTS), set(names) - set(lib.EXPORTS)


def test_load_job_errors_without_gpu(tmp_path):
    so = lib.load()
    h = ctypes.c_void_p()
    rc = so.chkn_load_job(str(tmp_path / "missing").encode(), 0, 1, ctypes.byref(h))
    assert rc == -2
    assert b"Job directory is not present" in so.chkn_last_error(None)


def test_gas_model_mismatch_is_a_config_error(tmp_path):
    from chkn import cases, prep
    cases.sod_box(4, 4, 4, max_step=1)
    prep.config.gas_model.R = 287.0  # R = 287 rather than 287.1 (config.cu:372-386)
    prep.config.gas_model.Cv = 287.0 / 0.4
    prep.write_initial_files(True, str(tmp_path))
    so = lib.load()
    h = ctypes.c_void_p()
    rc = so.chkn_load_job(str(tmp_path).encode(), 0, 1, ctypes.byref(h))
    assert rc == -3
    assert so.chkn_last_error(None) == b"Incorrect gas model data"
    # ... and the ideal-air job in the reacting-gas build.
    cases.sod_box(4, 4, 4, max_step=1)
    prep.write_initial_files(True, str(tmp_path))
    so = lib.load("ab")
    rc = so.chkn_load_job(str(tmp_path).encode(), 0, 1, ctypes.byref(h))
    assert rc == -3
    assert so.chkn_last_error(None) == b"Incorrect gas model data"


def test_tuner_candidates_are_valid_launch_options():
    from chkn import tune
    c = tune.candidates()
    assert len(c) == len(tune.TILES) * len(tune.TILE_K) * len(tune.ORDERS) * len(tune.PERSISTENT)
    for o in c:
        assert set(o) == {"tile_i", "tile_j", "tile_k", "tile_order", "persistent"}
        assert o["tile_i"] * o["tile_j"] <= 1024
