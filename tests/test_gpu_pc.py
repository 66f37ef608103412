"""The two-stage predictor-corrector step (time_scheme 1) of the HIP path against the oracle's.

The reference steps TVD-RK3 only; the predictor-corrector follows Lorikeet
(src/puffin/cell.d:294-303) and the oracle's form of it is pinned in test_oracle_pc.py.  Here
the HIP stages (fused tile kernel, face-list form, per-block and persistent launches, vertex
frames, ensembles) are compared with the oracle bit for bit, as the RK3 step is.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_bind as ob
from chkn import cases, prep
from chkn.job import JobData
from chkn.lib import BadCellError, Sim

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def pc(job):
    job.time_scheme = 1
    return job


def match(job, steps, launch=None):
    sim = Sim(job)
    if launch:
        sim.set_launch(**launch)
    rep = sim.step(steps)
    orc = ob.Oracle(job)
    rc, _, dts = orc.march(steps)
    assert rc == 0 and rep.step == orc.step == steps
    assert rep.t == orc.t and rep.dt == dts[steps - 1]
    for b in range(len(job.blocks)):
        assert np.array_equal(sim.iovars(b), orc.iovars(b)), b
        if job.blocks[b].active:
            assert np.array_equal(sim.conserved(b), orc.conserved(b)), b
    sim.close()


def _cfl_ramp():
    cases.ramp(20, 6, 8, max_step=40)
    prep.config.dt_init = 2.0e-5
    return JobData.from_prep()


def _first_order():
    cases.mixed_bc_array(max_step=15)
    prep.config.x_order = 1
    return JobData.from_prep()


JOBS = {
    "sod": (lambda: cases.sod_box(16, 8, 8, max_step=30), 30),
    "mixed_bc": (lambda: cases.mixed_bc_array(max_step=25), 25),
    "ramp_cfl": (_cfl_ramp, 40),
    "x_order_1": (_first_order, 15),
    "viscous_plate": (lambda: cases.viscous_plate(max_step=20), 20),
    "bl_plate": (lambda: cases.bl_plate(max_step=15)[0], 15),
    "sbp": (lambda: cases.sbp_channel(max_step=15), 15),
    "uneven_array": (lambda: cases.ramp_array(seed=3, lo=4, hi=9, max_step=22, i_walls=True), 22),
}


@pytest.mark.parametrize("name", sorted(JOBS))
def test_predictor_corrector_matches_oracle(name):
    builder, steps = JOBS[name]
    match(pc(builder()), steps)


@pytest.mark.parametrize("launch", [
    dict(per_block_launch=True),
    dict(persistent=True, use_graph=True),
    dict(tile_i=64, tile_j=4, tile_k=3),
    dict(tile_i=32, tile_j=8, use_graph=True, vertex_frames=True),
])
def test_predictor_corrector_launch_options(launch):
    match(pc(cases.ramp_array(seed=5, lo=4, hi=12, max_step=25)), 25, launch)


def test_predictor_corrector_reacting_gas():
    match(pc(cases.ab_reacting(max_step=20)), 20)


def test_predictor_corrector_scale_ramp_at_full_size():
    # 256x128x128 Mach 5 ramp, 11 steps through the first CFL update, as the bench runs it.
    match(pc(cases.ramp(256, 128, 128, max_step=11)), 11, dict(tile_i=32, tile_j=8, use_graph=True))


def test_predictor_corrector_ensemble_members_match_single_runs():
    jobs = [pc(j) for j in cases.ramp_ensemble(seed=2, n=3, nic=(8, 14), njk=(6, 9), max_step=12)]
    sim = Sim(jobs=jobs)
    sim.step(12)
    for q, job in enumerate(jobs):
        first, count = sim.instance_blocks(q)
        orc = ob.Oracle(job)
        rc, _, _ = orc.march(12)
        assert rc == 0
        for b in range(count):
            assert np.array_equal(sim.iovars(first + b), orc.iovars(b))
    sim.close()


def test_predictor_corrector_bad_cell_report():
    job = pc(cases.sod_box(8, 4, 4, max_step=5))
    io = job.blocks[0].iovars.copy()
    io[6, 9] = -1.0
    job.blocks[0].iovars = io
    sim = Sim(job)
    with pytest.raises(BadCellError) as ei:
        sim.step(5)
    rc, bad, _ = ob.Oracle(job).march(5)
    assert rc // 100 == ei.value.stage and bad == ei.value.count
    sim.close()


def test_time_scheme_key_in_job_files(tmp_path):
    # config.json carries "time_scheme": "predictor_corrector"; chkn-run marches it and writes
    # the same final flow file as the oracle's state.
    cases.sod_box(12, 4, 4, max_step=16)
    prep.config.time_scheme = "predictor_corrector"
    job = JobData.from_prep()
    assert job.time_scheme == 1
    d = str(tmp_path / "pcjob")
    prep.write_initial_files(True, d)
    run = os.path.join(ROOT, "gdtk-rmhd_amd", "chkn", "chkn-run")
    r = subprocess.run([run, "--job=" + d, "--binary"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    orc = ob.Oracle(job)
    rc, _, _ = orc.march(16)
    assert rc == 0
    last = sorted(os.listdir(os.path.join(d, "flow")))[-1]
    got = np.fromfile(os.path.join(d, "flow", last, "flow-0000-0000-0000.bin")).reshape(13, -1)
    assert np.array_equal(got, orc.iovars(0))
