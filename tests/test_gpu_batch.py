"""Ensembles (chkn_create_batch): several independent jobs in one handle, every stage of
every instance computed by one launch.  Each instance must reproduce, bit for bit, what a
handle holding that job alone computes, and the first instance is also checked against the
oracle.
"""
import numpy as np
import pytest

import oracle_bind as ob
from chkn import cases, prep
from chkn.job import JobData
from chkn.lib import ChknError, Sim

pytestmark = pytest.mark.gpu


def _ramp_cfl(seed, max_step, cfl):
    cases.ramp(20, 6, 8, max_step=max_step, seed=seed)
    prep.config.dt_init = 2.0e-5
    prep.config.cfl_count = 5
    prep.config.cfl_list = [(0.0, cfl)]
    return JobData.from_prep()


def _jobs():
    """Four jobs with different grids, block arrays, boundary conditions, inflow states,
    max_step and CFL numbers (all share x_order 2 and cfl_count 5)."""
    jobs = [_ramp_cfl(1, 23, 0.5)]
    jobs.append(cases.ramp_array(seed=3, lo=4, hi=9, max_step=17))
    prep.config.cfl_count = 5
    jobs[-1] = JobData.from_prep()
    cases.mixed_bc_array(max_step=30)
    prep.config.cfl_count = 5
    jobs.append(JobData.from_prep())
    jobs.append(_ramp_cfl(4, 30, 0.35))
    return jobs


def _same(a, b):
    return np.array_equal(a, b) or float(np.max(np.abs(a - b))) == 0.0


@pytest.mark.parametrize("launch", [dict(), dict(use_graph=True), dict(per_block_launch=True),
                                    dict(persistent=True, tile_i=16, tile_j=4, tile_k=3)])
def test_batch_equals_separate_handles(launch):
    jobs = _jobs()
    batch = Sim(jobs=jobs)
    batch.set_launch(**launch)
    assert batch.ninstances == len(jobs)
    # Uneven stepping: the segment boundaries inside chkn_step calls must not matter.
    for n in (7, 3, 100):
        batch.step(n)
    for q, job in enumerate(jobs):
        single = Sim(job)
        single.set_launch(**launch)
        single.step(100)
        rb, rs = batch.instance_report(q), single.report()
        assert (rb.step, rb.t, rb.dt) == (rs.step, rs.t, rs.dt), q
        assert rb.step == job.max_step
        first, count = batch.instance_blocks(q)
        assert count == len(job.blocks)
        for b in range(count):
            assert _same(batch.iovars(first + b), single.iovars(b)), (q, b)
            if job.blocks[b].active:
                assert _same(batch.conserved(first + b), single.conserved(b)), (q, b)
        single.close()
    batch.close()


def test_batch_first_instance_matches_oracle():
    jobs = _jobs()
    batch = Sim(jobs=jobs)
    batch.step(40)
    orc = ob.Oracle(jobs[0])
    rc, _, _ = orc.march(40)
    assert rc == 0
    rep = batch.instance_report(0)
    assert rep.step == orc.step and rep.t == orc.t
    for b in range(len(jobs[0].blocks)):
        assert _same(batch.iovars(b), orc.iovars(b))
    batch.close()


def test_batch_rejects_mixed_x_order():
    cases.sod_box(8, 4, 4, max_step=5)
    a = JobData.from_prep()
    prep.config.x_order = 1
    b = JobData.from_prep()
    with pytest.raises(ChknError, match="x_order"):
        Sim(jobs=[a, b])
