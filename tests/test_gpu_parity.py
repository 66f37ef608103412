"""HIP step (libchkn, through the C-ABI) against the CPU oracle, on the GPU.

Tolerances (FP64 throughout): after N steps the IO fields agree to 1e-9 relative, and a
single step from an identical state agrees to 1e-12 relative.  The kernels are built
without FMA contraction and keep the reference's operation order, so the results are
expected to be bit-identical; the tests report the observed maximum difference.
"""
import numpy as np
import pytest

import oracle_bind as ob
from chkn import cases, prep
from chkn.job import JobData
from chkn.lib import BadCellError, Sim

pytestmark = pytest.mark.gpu

FIELD_TOL = 1.0e-9
STEP_TOL = 1.0e-12


def rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def compare(job, steps, launch=None, tol=FIELD_TOL):
    sim = Sim(job)
    if launch:
        sim.set_launch(**launch)
    rep = sim.step(steps)
    orc = ob.Oracle(job)
    rc, _, dts = orc.march(steps)
    assert rc == 0
    assert rep.step == orc.step
    assert rep.t == orc.t, (rep.t, orc.t)
    worst = 0.0
    for b in range(len(job.blocks)):
        worst = max(worst, rel(sim.iovars(b), orc.iovars(b)))
        if job.blocks[b].active:
            worst = max(worst, rel(sim.conserved(b), orc.conserved(b)))
    sim.close()
    print("max rel diff %g" % worst)
    assert worst <= tol
    return worst


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
    "two_cell_block": (lambda: cases.ramp(2, 3, 2, max_step=20), 20),
    "uneven_array": (lambda: cases.ramp_array(seed=3, lo=4, hi=9, max_step=22, i_walls=True), 22),
}


@pytest.mark.parametrize("name", sorted(JOBS))
def test_fields_match_oracle(name):
    builder, steps = JOBS[name]
    compare(builder(), steps)


@pytest.mark.parametrize("launch", [
    dict(per_block_launch=True),
    dict(use_graph=True),
    dict(persistent=True),
    dict(tile_i=64, tile_j=4, tile_k=3),
    dict(tile_i=16, tile_j=16, tile_k=1),
    dict(tile_i=64, tile_j=8, use_graph=True, per_block_launch=True),
    dict(tile_i=32, tile_j=8, vertex_frames=True),
    dict(tile_i=16, tile_j=16, tile_k=2, use_graph=True, vertex_frames=True),
])
def test_launch_options_do_not_change_results(launch):
    compare(cases.ramp_array(seed=5, lo=4, hi=12, max_step=25), 25, launch=launch)


def test_single_step_residual():
    # From an identical developed state (the oracle's state after 12 steps), one step of
    # the HIP path agrees with one oracle step to 1e-12 relative in U.
    job = cases.ramp(24, 8, 10, max_step=30)
    orc = ob.Oracle(job)
    orc.march(12)
    job.blocks[0].iovars = orc.iovars(0)
    sim = Sim(job)
    sim.step(1)
    orc2 = ob.Oracle(job)
    orc2.march(1)
    assert rel(sim.conserved(0), orc2.conserved(0)) <= STEP_TOL
    sim.close()


def test_bad_cell_reports_stage_and_count():
    job = cases.sod_box(8, 4, 4, max_step=5)
    io = job.blocks[0].iovars.copy()
    io[6, 5] = -1.0  # negative density in one cell
    io[6, 17] = -2.0
    job.blocks[0].iovars = io
    sim = Sim(job)
    with pytest.raises(BadCellError) as ei:
        sim.step(5)
    orc = ob.Oracle(job)
    rc, bad, _ = orc.march(5)
    assert rc // 100 == ei.value.stage
    assert bad == ei.value.count
    assert "Stage %d bad cell count: %d" % (ei.value.stage, bad) in str(ei.value)
    sim.close()


def test_set_iovars_on_one_block_keeps_the_others_current():
    # Replacing one block's state after steps: the other blocks still report their stepped
    # state (not the creation-time one).  Re-setting them too re-encodes every block from its
    # primitives, as a restart does (simulate.cu:196-203), and the next steps then match an
    # oracle restarted from the same states -- at x_order 1, whose first stage reads the held
    # gas states.
    job = _first_order()
    sim = Sim(job)
    sim.step(6)
    before = [sim.iovars(b) for b in range(len(job.blocks))]
    new0 = before[0].copy()
    new0[4] *= 1.001  # p of block 0
    new0[6] *= 1.001  # rho: same T
    sim.set_iovars(0, new0)
    for b in range(1, len(job.blocks)):
        assert np.array_equal(sim.iovars(b), before[b]), b
    for b in range(1, len(job.blocks)):
        sim.set_iovars(b, before[b])
    sim.step(4)  # steps 6-9: the same dt as a restart's first steps (no CFL update yet)
    for b, jb in enumerate(job.blocks):
        jb.iovars = new0 if b == 0 else before[b]
    orc = ob.Oracle(job)
    rc, _, _ = orc.march(4)
    assert rc == 0
    for b in range(len(job.blocks)):
        assert np.array_equal(sim.iovars(b), orc.iovars(b)), b
    sim.close()


def test_restart_continues_after_chkn_step_calls():
    # Several chkn_step calls (splitting CFL segments) equal one long call.
    job = _cfl_ramp()
    a = Sim(job)
    for n in (3, 7, 1, 9, 20):
        a.step(n)
    b = Sim(job)
    b.step(40)
    assert np.array_equal(a.conserved(0), b.conserved(0))
    a.close()
    b.close()


def test_large_sod_conserves_mass_and_energy():
    # All-wall box: total mass and energy are conserved to round-off (size-independent check).
    job = cases.sod_box(128, 64, 64, max_step=40)
    sim = Sim(job)
    vol = sim.iovars(0)[3]
    U0 = sim.conserved(0)
    sim.step(40)
    U1 = sim.conserved(0)
    for q in (0, 4):
        m0, m1 = float(np.dot(U0[q], vol)), float(np.dot(U1[q], vol))
        assert abs(m1 - m0) <= 1e-11 * abs(m0)
    sim.close()


def test_sod_direction_symmetry():
    # The same tube along i, j and k gives the same profile (sod.py / sod_j.py / sod_k.py).
    profs = []
    for axis in range(3):
        dims = [8, 8, 8]
        dims[axis] = 48
        job = cases.sod_box(*dims, max_step=30, axis=axis)
        sim = Sim(job)
        sim.step(30)
        p = sim.iovars(0)[4].reshape(dims[2], dims[1], dims[0])
        line = [p[4, 4, :], p[4, :, 4], p[:, 4, 4]][axis]
        profs.append(line)
        sim.close()
    assert np.allclose(profs[0], profs[1], rtol=1e-12)
    assert np.allclose(profs[0], profs[2], rtol=1e-12)


def test_scale_ramp_at_full_size_matches_oracle():
    # The headline workload at its real size (one 256x128x128 block, Mach 5 over the 15-degree
    # ramp, 1% seeded density perturbation), 11 steps: the CFL reduction at step 10 and the
    # step after it, on the bench's tuned launch shape.  Bit for bit against the oracle,
    # fields and time-step sequence.
    # Both the fixed 4-plane k-chunks and the default layout (per-XCD slabs of 6, 5, 5 planes).
    job = cases.ramp(256, 128, 128, max_step=11)
    orc = ob.Oracle(job)
    rc, _, dts = orc.march(11)
    assert rc == 0 and orc.step == 11
    assert dts[10] != dts[9]  # the step after the reduction runs on the new dt
    for launch in (dict(tile_i=32, tile_j=8, tile_k=4, use_graph=True), dict(tile_i=32, tile_j=8, use_graph=True),
                   dict(tile_i=32, tile_j=8, use_graph=True, vertex_frames=True)):
        sim = Sim(job)
        sim.set_launch(**launch)
        rep = sim.step(11)
        assert rep.step == 11
        assert rep.t == orc.t and rep.dt == dts[10], (launch, rep.t, orc.t, rep.dt, dts[10])
        assert np.array_equal(sim.iovars(0), orc.iovars(0)), launch
        assert np.array_equal(sim.conserved(0), orc.conserved(0)), launch
        sim.close()


def test_plain_config_at_full_size_matches_oracle():
    # The plain workload at its own size: one 64x32x32 block, 3D Sod tube in a slip-wall box,
    # 200 steps.  dt grows 1.5x every 10 steps from 1e-6 until the CFL limit binds (near step
    # 80); fields, conserved state and the whole dt sequence bit for bit against the oracle.
    job = cases.sod_box(64, 32, 32, max_step=200)
    orc = ob.Oracle(job)
    rc, _, dts = orc.march(200)
    assert rc == 0 and orc.step == 200
    ramp = [1.0e-6 * 1.5 ** (s // 10) for s in range(200)]
    limited = [s for s in range(200) if dts[s] < ramp[s] * (1 - 1e-12)]
    assert limited and 60 <= limited[0] <= 100, limited[:3]  # the CFL limit is reached
    sim = Sim(job)
    sim.set_launch(use_graph=True)
    dt_seen = []
    for _ in range(20):
        rep = sim.step(10)
        dt_seen.append(rep.dt)
    assert rep.step == 200 and rep.t == orc.t
    assert dt_seen == [dts[10 * q + 9] for q in range(20)]
    assert np.array_equal(sim.iovars(0), orc.iovars(0))
    assert np.array_equal(sim.conserved(0), orc.conserved(0))
    sim.close()


def test_batched_array_at_full_size_matches_oracle():
    # The batched workload at its real size: a 5x4x2 array of 40 uneven blocks (6.44M cells,
    # slab sizes seeded in [16, 96]) with its seeded exchange / wall connection graph, one
    # launch per stage; 11 steps (through the first CFL update) bit for bit against the oracle.
    job = cases.ramp_array(seed=7, max_step=11)
    sim = Sim(job)
    sim.set_launch(tile_i=32, tile_j=8, tile_k=4, use_graph=True)
    rep = sim.step(11)
    orc = ob.Oracle(job)
    rc, _, dts = orc.march(11)
    assert rc == 0 and rep.step == orc.step == 11 and rep.t == orc.t and rep.dt == dts[10]
    for b in range(len(job.blocks)):
        assert np.array_equal(sim.iovars(b), orc.iovars(b)), b
        if job.blocks[b].active:
            assert np.array_equal(sim.conserved(b), orc.conserved(b)), b
    sim.close()


def test_ramp_slab_matches_oracle():
    # One GPU's share of the 256x128x128 ramp split 8 ways along i (the 32x128x128 slab with
    # the ramp corner), run standalone: 3 steps bit for bit against the oracle.
    job = cases.ramp_slab(64, 32, max_step=3)
    sim = Sim(job)
    rep = sim.step(3)
    orc = ob.Oracle(job)
    rc, _, _ = orc.march(3)
    assert rc == 0 and rep.step == orc.step == 3 and rep.t == orc.t
    assert np.array_equal(sim.iovars(0), orc.iovars(0))
    assert np.array_equal(sim.conserved(0), orc.conserved(0))
    sim.close()
