"""Viscous fluxes, inflow functions, the SBP-ASF flux and the face-list stage form on the GPU
(through the C-ABI).

The viscous plate (inflow, outflow, periodic exchange, a no-slip adiabatic wall and a
no-slip fixed-T wall, warped grid, varied initial state) against the oracle at x_order 2 and
1, to the step tolerances of test_gpu_parity (1e-9 relative after N steps); launch options
and ensembles must not change the result (bit for bit).  First-order jobs whose file and
inflow gas states differ from the equation of state run in the same face-list form.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from chkn import cases, prep
from chkn.job import JobData
from chkn.lib import ChknError, Sim
from test_gpu_parity import FIELD_TOL, compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("x_order,steps", [(2, 30), (1, 20)])
def test_viscous_plate_matches_oracle(x_order, steps):
    compare(cases.viscous_plate(max_step=steps, x_order=x_order, seed=3 + x_order), steps)


def test_viscous_plate_single_step_residual():
    # One step from the oracle's developed state (no file gas states left) to 1e-12 in U.
    job = cases.viscous_plate(max_step=40)
    orc = ob.Oracle(job)
    orc.march(15)
    for b in range(len(job.blocks)):
        job.blocks[b].iovars = orc.iovars(b)
    sim = Sim(job)
    sim.step(1)
    orc2 = ob.Oracle(job)
    orc2.march(1)
    for b in range(len(job.blocks)):
        a, r = sim.conserved(b), orc2.conserved(b)
        assert float(np.max(np.abs(a - r) / np.maximum(np.abs(r), 1e-300))) <= 1e-12
    sim.close()


def _run(job_or_jobs, steps, launch=None, splits=None):
    sim = Sim(jobs=job_or_jobs) if isinstance(job_or_jobs, list) else Sim(job_or_jobs)
    if launch:
        sim.set_launch(**launch)
    for n in (splits or [steps]):
        sim.step(n)
    out = [sim.conserved(b) for b in range(sim.nblocks)]
    sim.close()
    return out


@pytest.mark.parametrize("launch", [dict(use_graph=True), dict(per_block_launch=True),
                                    dict(persistent=True, tile_i=16, tile_j=4)])
def test_viscous_launch_options_and_split_calls(launch):
    job = cases.viscous_plate(max_step=25)
    ref = _run(job, 25)
    got = _run(job, 25, launch=launch, splits=[1, 6, 18])
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def test_viscous_ensemble_member_equals_single():
    jobs = [cases.viscous_plate(max_step=20, seed=5), cases.viscous_plate(nics=(6, 9), nkcs=(5, 7),
                                                                            max_step=14, seed=9)]
    ens = _run(jobs, 20)
    off = 0
    for job in jobs:
        single = _run(job, 20)
        for b in range(len(job.blocks)):
            assert np.array_equal(single[b], ens[off + b])
        off += len(job.blocks)


def test_first_order_with_file_and_inflow_gas_states():
    # mixed_bc_array: inflow at p = 5e3, T = 260 (rho, e do not give back p), an inactive
    # block read from the file, cells whose file T differs from e/Cv.
    cases.mixed_bc_array(max_step=15)
    prep.config.x_order = 1
    job = JobData.from_prep()
    worst = compare(job, 15)
    assert worst <= FIELD_TOL


def test_twall_per_face_needs_its_values():
    # NaN marks a TWall given per face; without the values the job is rejected.
    job = cases.viscous_plate(max_step=5)
    job.blocks[1].bc_twall = [0.0, 0.0, 0.0, 0.0, math.nan, 0.0]
    with pytest.raises(ChknError, match="twall_faces holds too few values"):
        Sim(job)


def test_viscous_gradient_is_rejected():
    sim = Sim(cases.viscous_plate(max_step=5))
    with pytest.raises(ChknError, match="adjoint of the viscous fluxes"):
        sim.grad_forward(2, 2)
    sim.close()


@pytest.mark.parametrize("x_order,steps", [(2, 30), (1, 20)])
def test_boundary_layer_inflow_function_matches_oracle(x_order, steps):
    # Inflow through the laminar boundary-layer profile (spline in y), a no-slip wall with
    # TWall per face, inflow above, periodic span.
    job, _ = cases.bl_plate(max_step=steps, x_order=x_order, seed=2 + x_order)
    compare(job, steps)


def test_boundary_layer_launch_options():
    job, _ = cases.bl_plate(max_step=20)
    ref = _run(job, 20)
    got = _run(job, 20, launch=dict(use_graph=True), splits=[3, 17])
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def test_inviscid_inflow_function_matches_oracle():
    # The inflow-function ghosts in the fused stage kernel (no viscous fluxes).
    job, _ = cases.bl_plate(max_step=25)
    job.viscous = False
    compare(job, 25)


def test_missing_boundary_layer_file_and_twall_file(tmp_path, monkeypatch):
    cases.bl_plate(max_step=5)
    prep.write_initial_files(True, str(tmp_path))
    monkeypatch.chdir(tmp_path)
    with pytest.raises(ChknError, match="Boundary-layer file is not present"):
        Sim(job_dir=str(tmp_path))
    (tmp_path / "InflowBoundaryLayerBC.data").write_text("1013\n12\n" + "0 300 0\n" * 12)
    (tmp_path / "TWall-0000-0000-0000-kminus.gz").unlink()
    with pytest.raises(ChknError, match="TWall file"):
        Sim(job_dir=str(tmp_path))


def test_sbp_asf_matches_oracle_and_launch_options():
    # The central SBP-ASF flux (face.cu:268-342), inflow/outflow and periodic exchanges.
    job = cases.sbp_channel(max_step=25)
    compare(job, 25)
    ref = _run(job, 25)
    got = _run(job, 25, launch=dict(use_graph=True), splits=[2, 23])
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def test_reacting_gas_matches_oracle():
    # The A->B reacting gas build (libchkn_ab.so): gas constants and the reaction step.
    compare(cases.ab_reacting(max_step=30), 30)


def test_reacting_ensemble_and_launch_options():
    # A reacting and an inert instance in one handle equal each run alone; hipGraph replay
    # and split calls change nothing.
    jobs = [cases.ab_reacting(max_step=24, seed=2), cases.ab_reacting(max_step=24, seed=5, reacting=False)]
    ens = _run(jobs, 24, launch=dict(use_graph=True), splits=[5, 19])
    off = 0
    for job in jobs:
        single = _run(job, 24)
        for b in range(len(job.blocks)):
            assert np.array_equal(single[b], ens[off + b])
        off += len(job.blocks)


def test_viscous_closed_box_conserves_mass_at_size():
    # A 128x64x64 Sod box with no-slip adiabatic walls on every side, viscous: no mass crosses
    # a wall (the convective mass flux of a reflected ghost is zero and the viscous flux has no
    # mass term), so the total mass changes only by the round-off of the cell updates.  The
    # walls make the flow three-dimensional, so that round-off grows to ~1e-12 relative over 30
    # steps at 32x16x16 in the oracle (against ~1e-15 for the one-dimensional inviscid tube);
    # a leak through the walls would be orders of magnitude larger than the 1e-9 allowed.
    cases.sod_box(128, 64, 64, max_step=30)
    for fb in prep.fluidBlocksList:
        for face in list(fb.bcs):
            fb.bcs[face] = prep.WallNoSlipAdiabaticBC()
    prep.config.viscous = True
    job = JobData.from_prep()
    sim = Sim(job)
    vol = sim.iovars(0)[3]
    m0 = float(np.dot(sim.conserved(0)[0], vol))
    sim.step(30)
    U1 = sim.conserved(0)
    assert np.all(np.isfinite(U1))
    m1 = float(np.dot(U1[0], vol))
    assert abs(m1 - m0) <= 1e-9 * abs(m0)
    sim.close()


def test_viscous_ramp_matches_oracle_mid_size():
    # The viscous ramp of the viscous workload at 48x24x24, against the oracle.
    compare(cases.ramp(48, 24, 24, max_step=6, viscous=True), 6)


def test_viscous_ramp_at_full_size_matches_oracle():
    # The viscous workload at its real size (one 256x128x128 block, Mach 5 ramp, no-slip ramp
    # wall at 300 K, Sutherland air), 2 steps bit for bit against the oracle.
    job = cases.ramp(256, 128, 128, max_step=2, viscous=True)
    sim = Sim(job)
    rep = sim.step(2)
    orc = ob.Oracle(job)
    rc, _, _ = orc.march(2)
    assert rc == 0 and rep.step == orc.step == 2 and rep.t == orc.t
    assert np.array_equal(sim.iovars(0), orc.iovars(0))
    assert np.array_equal(sim.conserved(0), orc.conserved(0))
    sim.close()
