"""The C oracle against golden vectors written by the reference's own sources.

The vectors come from tests/golden/make_golden.py, which drives the reference's
face/cell/block/bcs code (compiled from the reference tree into oracle/_ref).  Every
comparison is bitwise: the oracle restates the reference's arithmetic operation by
operation.  The FP32 fixtures come from the reference's FLOAT_NUMBERS build and check the
oracle's FP32 build (float arithmetic, with the few double-evaluated expressions the
reference's FP32 build has).
"""
import os

import numpy as np
import pytest

import oracle_bind as ob
from chkn import cases, prep
from chkn.job import BL_PROFILE_FILE, JobData, write_bl_profile

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def test_flux_vectors_bitwise():
    d = np.load(os.path.join(GOLD, "flux_vectors.npz"))
    rec, F = d["rec"], d["F"]
    assert len(rec) >= 1000
    got = np.array([ob.face_flux(r[:36], r[36:45], int(r[45]), int(r[46])) for r in rec])
    assert np.array_equal(bits(got), bits(F)), np.max(np.abs(got - F))


def test_geometry_vectors_bitwise():
    d = np.load(os.path.join(GOLD, "geom_vectors.npz"))
    for p, out in zip(d["p"], d["out"]):
        hexo = ob.hex_props(p)
        assert np.array_equal(bits(hexo), bits(out[:7]))
        quad = ob.quad_props(p[:12])
        assert np.array_equal(bits(quad), bits(out[7:20]))


def _cfl_ramp():
    cases.ramp(20, 6, 8, max_step=40)
    prep.config.dt_init = 2.0e-5
    return JobData.from_prep()


def _first_order():
    cases.mixed_bc_array(max_step=15)
    prep.config.x_order = 1
    return JobData.from_prep()


def _two_cell():
    cases.ramp(2, 3, 2, max_step=20)
    return JobData.from_prep()


STEP_JOBS = {
    "sod": (lambda: cases.sod_box(16, 8, 8, max_step=30), 30),
    "mixed": (lambda: cases.mixed_bc_array(max_step=25), 25),
    "ramp_cfl": (_cfl_ramp, 40),
    "xorder1": (_first_order, 15),
    "tiny_n2": (_two_cell, 20),
    "viscous": (lambda: cases.viscous_plate(max_step=30), 30),
    "viscous_x1": (lambda: cases.viscous_plate(max_step=20, x_order=1, seed=8), 20),
    "bl_plate": (lambda: cases.bl_plate(max_step=30), 30),
    "bl_plate_x1": (lambda: cases.bl_plate(max_step=20, x_order=1, seed=6), 20),
    "sbp": (lambda: cases.sbp_channel(max_step=25), 25),
    "sbp_viscous": (lambda: _viscous_sbp(), 20),
    "ab_reacting": (lambda: cases.ab_reacting(max_step=30), 30),
    "ab_inert": (lambda: cases.ab_reacting(max_step=20, seed=5, reacting=False), 20),
}


def _viscous_sbp():
    cases.sbp_channel(max_step=20, seed=5)
    prep.config.viscous = True
    return JobData.from_prep()


def build_step_job(name, job_dir):
    """Build the named job, write its files (and boundary-layer profile) into job_dir and
    read it back as the golden run did.  Returns (JobData, steps, profile or None)."""
    builder, steps = STEP_JOBS[name]
    built = builder()
    prof = built[1] if isinstance(built, tuple) else None
    prep.write_initial_files(True, job_dir)
    if prof is not None:
        write_bl_profile(os.path.join(job_dir, BL_PROFILE_FILE), prof)
    return JobData.from_job_dir(job_dir, bl_profile=prof), steps, prof


def job_digest(job):
    import hashlib
    f = job.flat()
    h = hashlib.sha256()
    for key in sorted(f):
        h.update(key.encode())
        h.update(np.ascontiguousarray(f[key]).tobytes())
    for key in ("x_order", "cfl_count", "max_step", "max_time", "dt_init", "viscous"):
        h.update(repr(getattr(job, key)).encode())
    return h.hexdigest()


def load_step_fixture(name):
    return np.load(os.path.join(GOLD, "step_%s.npz" % name))


# Step jobs with fixtures from the reference's FP32 build (step_f32_<name>.npz).
F32_STEP_JOBS = ["sod", "mixed", "ramp_cfl", "xorder1", "viscous", "bl_plate", "sbp", "ab_reacting"]


def _golden_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLD, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def build_step_job_f32(name, job_dir):
    """The named job in gzipped-text files (the FP32 build reads no binary files), read back
    as the FP32 golden run read it; A->B jobs carry the FP32 build's Cv (make_golden.py
    f32_gas_model).  Returns (JobData, steps, profile or None)."""
    from chkn.job import gas_flavour
    builder, steps = STEP_JOBS[name]
    built = builder()
    prof = built[1] if isinstance(built, tuple) else None
    job = built[0] if isinstance(built, tuple) else built
    prep.write_initial_files(False, job_dir)
    if gas_flavour(job.gas_model) == "ab":
        _golden_module().f32_gas_model(job_dir)
    if prof is not None:
        write_bl_profile(os.path.join(job_dir, BL_PROFILE_FILE), prof)
    return JobData.from_job_dir(job_dir, binary=False, bl_profile=prof), steps, prof


def bits32(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.int32)


@pytest.mark.parametrize("name", F32_STEP_JOBS)
def test_f32_march_matches_reference_bitwise(name, tmp_path):
    jd, steps, _ = build_step_job_f32(name, str(tmp_path))
    gold = np.load(os.path.join(GOLD, "step_f32_%s.npz" % name))
    assert str(gold["digest"]) == job_digest(jd), "job generator changed; regenerate golden"
    o = ob.Oracle(jd, fp32=True)
    rc, bad, dts = o.march(steps)
    assert rc == 0
    got = np.asarray(dts[:steps])
    assert np.array_equal(got, got.astype(np.float32))  # the FP32 oracle's dt is a float
    assert np.array_equal(bits32(got), bits32(gold["dt"]))
    for b in range(len(jd.blocks)):
        io = o.iovars(b)
        assert np.array_equal(io, io.astype(np.float32))
        assert np.array_equal(bits32(io), bits32(gold["io%d" % b])), "block %d" % b


def test_f32_fixtures_differ_from_fp64():
    # The FP32 trajectories are genuinely float: they differ from the FP64 ones in the low
    # bits but agree to float precision (a check that the two fixture sets are not swapped).
    for name in ("sod", "viscous", "ab_reacting"):
        g64, g32 = load_step_fixture(name), np.load(os.path.join(GOLD, "step_f32_%s.npz" % name))
        a, b = np.asarray(g64["io0"]), np.asarray(g32["io0"], dtype=np.float64)
        assert not np.array_equal(a.astype(np.float32), g32["io0"])
        rho64, rho32 = a[6], b[6]
        assert np.max(np.abs(rho32 - rho64) / np.abs(rho64)) < 1e-4


@pytest.mark.parametrize("name", sorted(STEP_JOBS))
def test_march_matches_reference_bitwise(name, tmp_path):
    # The golden run read the job back from binary files; do the same.
    jd, steps, _ = build_step_job(name, str(tmp_path))
    gold = load_step_fixture(name)
    assert str(gold["digest"]) == job_digest(jd), "job generator changed; regenerate golden"
    o = ob.Oracle(jd)
    rc, bad, dts = o.march(steps)
    assert rc == 0
    assert np.array_equal(bits(dts[:steps]), bits(gold["dt"]))
    for b in range(len(jd.blocks)):
        assert np.array_equal(bits(o.iovars(b)), bits(gold["io%d" % b])), "block %d" % b
