"""The HIP march against the reference build's own output, on the GPU.

Each tests/golden/step_<name>.npz holds what the reference's CPU driver, compiled from its
sources, computed for a small job: the dt of every step and the final IO variables.  The job
directory is written again and loaded through chkn_load_job (the reference's file formats;
the boundary-layer profile is read from the working directory, as the reference reads it),
marched for the same steps, and compared bit for bit: the kernels keep the reference's
operation order without contraction, so the dt sequence and every IO variable are equal.
The FP32 flavour is compared the same way with tests/golden/step_f32_<name>.npz, written by
the reference's FP32 build from gzipped-text job files.
"""
import os

import numpy as np
import pytest

from chkn.lib import Sim
from test_oracle_golden import F32_STEP_JOBS, STEP_JOBS, build_step_job, build_step_job_f32, job_digest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", sorted(STEP_JOBS))
def test_hip_march_matches_reference_build(name, tmp_path, monkeypatch):
    jd, steps, _ = build_step_job(name, str(tmp_path))
    monkeypatch.chdir(tmp_path)
    gold = np.load(os.path.join(GOLD, "step_%s.npz" % name))
    assert str(gold["digest"]) == job_digest(jd)
    sim = Sim(job_dir=str(tmp_path))
    rep = sim.step(steps)
    assert rep.step == steps
    t = 0.0
    for dt in gold["dt"]:
        t += float(dt)
    assert rep.t == t
    worst = 0.0
    same = True
    for b in range(len(jd.blocks)):
        a, r = sim.iovars(b), gold["io%d" % b]
        worst = max(worst, float(np.max(np.abs(a - r) / np.maximum(np.abs(r), 1e-300))))
        same = same and np.array_equal(a.view(np.uint64), np.ascontiguousarray(r).view(np.uint64))
    sim.close()
    print("%s: max rel diff %g" % (name, worst))
    assert worst <= 1e-9
    assert same


@pytest.mark.parametrize("name", F32_STEP_JOBS)
def test_hip_fp32_march_matches_reference_fp32_build(name, tmp_path, monkeypatch):
    # The FP32 flavour (libchkn_f32.so / libchkn_ab_f32.so) against the reference's FP32
    # build (FLOAT_NUMBERS) on the same gzipped-text job: every dt and IO variable as float.
    jd, steps, _ = build_step_job_f32(name, str(tmp_path))
    monkeypatch.chdir(tmp_path)
    gold = np.load(os.path.join(GOLD, "step_f32_%s.npz" % name))
    assert str(gold["digest"]) == job_digest(jd)
    sim = Sim(job_dir=str(tmp_path), binary=False, fp32=True)
    rep = sim.step(steps)
    assert rep.step == steps
    t = np.float32(0.0)
    for dt in gold["dt"]:
        t = np.float32(t + dt)
    assert rep.t == float(t)
    for b in range(len(jd.blocks)):
        a = sim.iovars(b)
        assert np.array_equal(a, a.astype(np.float32))
        assert np.array_equal(a.astype(np.float32).view(np.int32), np.asarray(gold["io%d" % b]).view(np.int32)), b
    sim.close()
