"""The reference's integration known answer on the oracle: Mach 3 flow over the forward-facing
step (examples/chicken/forward-facing-step), t = 1e-3 s.  Along the first row of cells of the
block upstream of the step, the interpolated position of the normal-shock pressure
1047025 Pa is 0.41369441539674984 to a tenth of the cell size
(examples/chicken/test_chicken.py:53-71).  This pins the oracle's whole trajectory (CFL dt
sequence, inflow, exchange across 21 active blocks, inactive blocks and walls) against a
number the reference itself holds."""
import numpy as np

import oracle_bind as ob
from chkn import cases

FFS_SHOCK_X = 0.41369441539674984  # examples/chicken/test_chicken.py:68


def ffs_line(iov, nic=60, njc=20, nkc=2):
    """x and p along i of row j = 0, layer k = 0 of block (0, 0, 0)."""
    return iov[0].reshape(nkc, njc, nic)[0, 0, :], iov[4].reshape(nkc, njc, nic)[0, 0, :]


def test_forward_facing_step_shock_position():
    job = cases.forward_step(max_time=1.0e-3)
    assert job.active_cells == 21 * 60 * 20 * 2
    o = ob.Oracle(job)
    rc, _, _ = o.march(100000)
    assert rc == 0 and o.t >= 1.0e-3 and o.t - o.dt < 1.0e-3
    x, p = ffs_line(o.iovars(0))
    xs, dx = cases.shock_position(x, p)
    assert xs is not None
    assert abs(xs - FFS_SHOCK_X) < 0.1 * dx, (xs, dx)
