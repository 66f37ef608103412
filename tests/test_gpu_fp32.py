"""The FP32 flavours (libchkn_f32.so, libchkn_ab_f32.so) on the GPU.

The reference builds the same sources with float numbers (number.cu:10-27, FLOAT_NUMBERS).
Its own FP32 build is checked against this build in test_gpu_golden.py; here the FP32
oracle (oracle/liborc_f32.so, pinned to the same reference fixtures) checks larger and
multi-block jobs, launch options and ensembles bit for bit, and the FP32 file formats
(4-byte binary output, text input only) are exercised through the library.
"""
import os

import numpy as np
import pytest

import oracle_bind as ob
from chkn import cases, prep
from chkn.job import JobData
from chkn.lib import ChknError, Sim

pytestmark = pytest.mark.gpu


def f32_bits(a):
    a = np.asarray(a)
    assert np.array_equal(a, a.astype(np.float32)), "values are not floats"
    return a.astype(np.float32).view(np.int32)


def compare_f32(job, steps, launch=None, splits=None):
    sim = Sim(job, fp32=True)
    if launch:
        sim.set_launch(**launch)
    for n in (splits or [steps]):
        rep = sim.step(n)
    orc = ob.Oracle(job, fp32=True)
    rc, _, dts = orc.march(steps)
    assert rc == 0
    assert rep.step == orc.step and rep.t == orc.t, (rep.t, orc.t)
    for b in range(len(job.blocks)):
        assert np.array_equal(f32_bits(sim.iovars(b)), f32_bits(orc.iovars(b))), "iovars block %d" % b
        if job.blocks[b].active:
            assert np.array_equal(f32_bits(sim.conserved(b)), f32_bits(orc.conserved(b))), "U block %d" % b
    sim.close()


@pytest.mark.parametrize("name", ["uneven_array", "ramp_mid", "viscous_ramp", "x_order_1"])
def test_fp32_matches_fp32_oracle(name):
    if name == "uneven_array":
        compare_f32(cases.ramp_array(seed=3, lo=4, hi=9, max_step=22, i_walls=True), 22)
    elif name == "ramp_mid":
        compare_f32(cases.ramp(64, 32, 32, max_step=25), 25)
    elif name == "viscous_ramp":
        compare_f32(cases.ramp(32, 16, 16, max_step=12, viscous=True), 12)
    else:
        cases.mixed_bc_array(max_step=15)
        prep.config.x_order = 1
        compare_f32(JobData.from_prep(), 15)


@pytest.mark.parametrize("launch", [dict(use_graph=True), dict(per_block_launch=True),
                                    dict(persistent=True, tile_i=16, tile_j=4)])
def test_fp32_launch_options_and_split_calls(launch):
    compare_f32(cases.ramp_array(seed=5, lo=4, hi=9, max_step=24, i_walls=True), 24, launch=launch,
                splits=[1, 9, 14])


@pytest.mark.parametrize("name", ["uneven_array", "ramp_mid"])
def test_fp32_predictor_corrector_matches_oracle(name):
    # The two-stage predictor-corrector (time_scheme 1) in float numbers.
    if name == "uneven_array":
        job = cases.ramp_array(seed=3, lo=4, hi=9, max_step=22, i_walls=True)
    else:
        job = cases.ramp(64, 32, 32, max_step=25)
    job.time_scheme = 1
    compare_f32(job, job.max_step)


def test_fp32_reacting_matches_oracle():
    # The FP32 build's own Cv = R/(g - 1.0) in float g and R; the prep's 1435 is outside its
    # 1e-4 gas-model check (test_abi.py), as it is for the reference's FP32 build.
    job = cases.ab_reacting(max_step=30)
    g, r = float(np.float32(job.gas_model[0])), float(np.float32(job.gas_model[1]))
    job.gas_model = list(job.gas_model[:5]) + [float(np.float32(r / (g - 1.0)))]
    compare_f32(job, 30)


def test_fp32_ensemble_member_equals_single():
    jobs = [cases.ramp(24, 8, 8, max_step=20), cases.sod_box(16, 8, 8, max_step=14)]
    ens = Sim(jobs=jobs, fp32=True)
    ens.step(20)
    off = 0
    for job in jobs:
        single = Sim(job, fp32=True)
        single.step(20)
        for b in range(len(job.blocks)):
            assert np.array_equal(single.conserved(b), ens.conserved(off + b))
        single.close()
        off += len(job.blocks)
    ens.close()


def test_fp32_binary_output_is_four_byte(tmp_path):
    # Text job in, binary snapshot out: the flow file holds 13 x ncells 4-byte numbers
    # (block.cu:938-955 writes sizeof(number)), equal to the handle's IO variables.
    cases.sod_box(16, 8, 8, max_step=10)
    d = str(tmp_path)
    prep.write_initial_files(False, d)
    sim = Sim(job_dir=d, binary=False, fp32=True)
    sim.march(binary=True)
    sim.finalize(binary=True)
    io = sim.iovars(0)
    sim.close()
    path = os.path.join(d, "flow", "t0001", "flow-0000-0000-0000.bin")
    assert os.path.getsize(path) == 13 * 16 * 8 * 8 * 4
    got = np.fromfile(path, dtype=np.float32).reshape(13, -1)
    assert np.array_equal(got.view(np.int32), f32_bits(io).reshape(13, -1))


def test_fp32_gradients_are_fp64_only():
    sim = Sim(cases.ramp(16, 8, 8, max_step=4), fp32=True)
    with pytest.raises(ChknError, match="FP64 flavour only"):
        sim.grad_forward(2, 2)
    sim.close()
