"""The forward-facing-step known answer on the GPU (examples/chicken/test_chicken.py:53-71):
the shock at t = 1e-3 s within a tenth of a cell of 0.41369441539674984, through the library
(bit for bit against the oracle as well) and through the job-file workflow the reference's
test uses (prep writes gzipped-text files, chkn-run marches them, the t0001 snapshot is
read back)."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_bind as ob
from chkn import cases, prep
from chkn.lib import Sim
from chkn.post import PostData
from test_oracle_ffs import FFS_SHOCK_X, ffs_line

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_ffs_shock_position_library():
    job = cases.forward_step(max_time=1.0e-3)
    sim = Sim(job)
    rep = sim.step(100000)
    assert rep.t >= 1.0e-3
    iov = sim.iovars(0)
    xs, dx = cases.shock_position(*ffs_line(iov))
    assert abs(xs - FFS_SHOCK_X) < 0.1 * dx, (xs, dx)
    orc = ob.Oracle(job)
    rc, _, _ = orc.march(100000)
    assert rc == 0 and orc.step == rep.step and orc.t == rep.t
    for b in range(len(job.blocks)):
        assert np.array_equal(sim.iovars(b), orc.iovars(b)), b
    sim.close()


def test_ffs_shock_position_job_files(tmp_path):
    cases.forward_step(max_time=1.0e-3)  # one snapshot, at the end (dt_plot = max_time)
    d = str(tmp_path / "ffs")
    prep.write_initial_files(False, d)
    run = os.path.join(ROOT, "gdtk-rmhd_amd", "chkn", "chkn-run")
    r = subprocess.run([run, "--job=" + d], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "Done in" in r.stdout, r.stderr
    # The progress line every print_count (20) steps, in the reference's format
    # (simulate.cu:311-326).
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("Step=")]
    assert lines and lines[0].startswith("Step=20 ")
    assert all(re.fullmatch(r"Step=\d+ t=\d\.\d{3}e[-+]\d\d dt=\d\.\d{3}e[-+]\d\d cfl=0\.500 "
                            r"WC=\d+\.\d{3}s WCtFT=-?\d+\.\d{3}s WCtMS=\d+\.\d{3}s", ln) for ln in lines), lines[:2]
    assert [int(ln.split()[0][5:]) for ln in lines] == list(range(20, 20 * len(lines) + 1, 20))
    post = PostData()
    post.read_config(d)
    post.read_flow_blocks(d, 1, False)
    fd = post.flows["0,0,0"]
    x = fd["posx"].reshape(2, 20, 60)[0, 0, :]
    p = fd["p"].reshape(2, 20, 60)[0, 0, :]
    xs, dx = cases.shock_position(x, p)
    assert abs(xs - FFS_SHOCK_X) < 0.1 * dx, (xs, dx)
