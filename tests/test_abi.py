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
    # Every flavour: ideal air and the reacting A->B gas, each in FP64 and FP32.
    names = declared_functions()
    assert len(names) >= 15
    for flavour, path in lib.LIB_PATHS.items():
        so = ctypes.CDLL(path)
        for n in names:
            assert hasattr(so, n), (path, n)
        assert lib.load(flavour).chkn_real_size() == (4 if flavour.endswith("_f32") else 8)
    assert set(names) <= set(lib.EXPORTS), set(names) - set(lib.EXPORTS)


def test_fp32_build_reads_text_jobs_only(tmp_path):
    # block.cu:794-796: the FP32 build refuses binary grid files, with the reference's text.
    from chkn import cases, prep
    cases.sod_box(4, 4, 4, max_step=1)
    prep.write_initial_files(True, str(tmp_path))
    so = lib.load("air_f32")
    h = ctypes.c_void_p()
    rc = so.chkn_load_job(str(tmp_path).encode(), 0, 1, ctypes.byref(h))
    assert rc == -2
    assert so.chkn_last_error(None) == b"Do not ask to FP32 code to read grid file with binary data."


def test_fp32_reacting_build_rejects_the_prep_gas_model(tmp_path):
    # The FP32 A->B build's Cv = R/(g - 1.0) from float g and R is 1434.9996..., outside the
    # 1e-4 check against the 1435 the prep writes (config.cu:369-386): the reference's FP32
    # build refuses the job the same way (oracle/_ref/chkn_ref_ab_f32, where it is built).
    import subprocess
    from chkn import cases, prep
    cases.ab_reacting(max_step=1)
    prep.write_initial_files(False, str(tmp_path))
    so = lib.load("ab_f32")
    h = ctypes.c_void_p()
    rc = so.chkn_load_job(str(tmp_path).encode(), 0, 0, ctypes.byref(h))
    assert rc == -3
    assert so.chkn_last_error(None) == b"Incorrect gas model data"
    ref = os.path.join(ROOT, "oracle", "_ref", "chkn_ref_ab_f32")
    if os.path.exists(ref):
        r = subprocess.run([ref, "run", str(tmp_path), "0", "1", "-", "textin"], capture_output=True, text=True)
        assert r.returncode == 3 and "Incorrect gas model data" in r.stderr


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
    base = len(tune.TILES) * len(tune.TILE_K) * len(tune.ORDERS) * len(tune.PERSISTENT)
    assert len(c) == base + len(tune.VERTEX_FRAME_TILES)
    for o in c[:base]:
        assert set(o) == {"tile_i", "tile_j", "tile_k", "tile_order", "persistent"}
        assert o["tile_i"] * o["tile_j"] <= 1024
    for o in c[base:]:
        assert o["vertex_frames"] and (o["tile_i"], o["tile_j"]) in tune.VERTEX_FRAME_TILES
    assert len(tune.candidates(x_order=1)) == base  # the vertex-frame form is second-order only


def test_block_size_limit_at_its_boundary():
    # FP64: the vertex count (nic+1)(njc+1)(nkc+1), which bounds every per-direction face
    # count, must stay below 2^29 (32-bit byte offsets of 8-byte numbers); viscous jobs also
    # need 4 x (all faces) below 2^29.
    so = lib.load()
    assert so.chkn_check_block_dims(1023, 1023, 510, 0) == 0   # 2^20 * 511 vertices
    assert so.chkn_check_block_dims(1023, 1023, 511, 0) == -1  # exactly 2^29 vertices
    assert so.chkn_last_error(None) == b"block too large for 32-bit array offsets"
    assert so.chkn_check_block_dims(2047, 511, 511, 0) == -1   # 2^11 * 2^9 * 2^9
    assert so.chkn_check_block_dims(2046, 511, 511, 0) == 0
    nft = lambda a, b, c: (a + 1) * b * c + a * (b + 1) * c + a * b * (c + 1)
    assert 4 * nft(256, 256, 600) < 1 << 29 <= 4 * nft(256, 256, 700)
    assert so.chkn_check_block_dims(256, 256, 600, 1) == 0
    assert so.chkn_check_block_dims(256, 256, 700, 0) == 0
    assert so.chkn_check_block_dims(256, 256, 700, 1) == -1
    assert so.chkn_check_block_dims(1, 4, 4, 0) == -1          # fewer than 2 cells


def test_blocks_beyond_32_bit_offsets_are_rejected():
    # The kernels address each per-block array with 32-bit byte offsets: a block with 2^29 or
    # more faces in one direction (FP64) is refused at set-up, before any device call.
    from chkn import cases
    job = cases.sod_box(4, 4, 4, max_step=1)
    keep = []
    d = lib._desc(job, keep)
    dims = keep[0]["blk_dims"]
    dims[:3] = [1024, 1024, 512]
    so = lib.load()
    h = ctypes.c_void_p()
    rc = so.chkn_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == -1
    assert so.chkn_last_error(None) == b"block too large for 32-bit array offsets"


def test_libraries_are_built_from_the_sources_in_the_tree():
    # Each flavour carries the digest of the sources it was compiled from; a library left over
    # from other sources (a stale build beside newer sources) fails here.
    want = lib.tree_digest()
    for flavour in lib.LIB_PATHS:
        assert lib.source_digest(flavour) == want, flavour
