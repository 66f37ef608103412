"""The oracle's two-stage predictor-corrector step (time_scheme 1).

The reference (chicken) always steps TVD-RK3; the predictor-corrector is the scheme of its
sibling solver Lorikeet (src/puffin/cell.d:294-303: U1 = U0 + dt L0, then
U1 = U0 + 0.5 dt (L0 + L1)).  No reference build holds it, so the oracle's form is pinned by
what the mathematics fixes: second-order convergence in dt on a smooth problem (first order
would be a lost corrector, third order the RK3 step), against the oracle's own TVD-RK3 run at
a much smaller dt on the same grid; a one-step gap to RK3 of third order; and discrete
conservation to round-off where the fluxes telescope (a periodic tube).
"""
import numpy as np
import pytest

import oracle_bind as ob
from chkn import cases, prep
from chkn.job import JobData


def periodic_wave(n, dt, steps, scheme, x_order=1):
    """Smooth acoustic wave riding a 100 m/s stream through a periodic (self-exchanging)
    tube, fixed dt (no CFL update within the run), subsonic everywhere so the flux's Mach
    and upwind switches never change branch."""
    prep.reset()
    c = prep.config
    c.max_step, c.max_time, c.dt_init = steps, 1.0, dt
    c.cfl_count, c.x_order = 1 << 20, x_order
    prep.add_cfl_value(0.0, 0.5)
    g = cases.box_grid(n, 4, 4, L=(1.0, 0.2, 0.2))
    x = prep.cell_centres(g).x
    s = np.sin(2.0 * np.pi * x)
    p = 1.0e5 * (1.0 + 0.014 * s)
    rho = 1.0e5 / (287.1 * 300.0) * (1.0 + 0.01 * s)
    st = cases._state_arrays(p, p / (287.1 * rho), 100.0 + 3.0 * s)
    prep.FluidBlock(grid=g, initialState=st, bcs={"iminus": prep.ExchangeBC(), "iplus": prep.ExchangeBC()})
    job = JobData.from_prep()
    job.time_scheme = scheme
    return job


def final_state(job, steps):
    o = ob.Oracle(job)
    rc, _, dts = o.march(steps)
    assert rc == 0 and o.step == steps and np.all(dts[:steps] == job.dt_init)
    return o.conserved(0)


@pytest.mark.parametrize("scheme,order", [(1, 2), (0, 3)])
def test_temporal_convergence_order(scheme, order):
    n, dt0, steps = 24, 2.0e-5, 40
    ref = final_state(periodic_wave(n, dt0 / 32, steps * 32, 0), steps * 32)
    err = []
    for m in (1, 2, 4):
        u = final_state(periodic_wave(n, dt0 / m, steps * m, scheme), steps * m)
        err.append(np.max(np.abs(u[0] - ref[0])))
    ratios = [err[0] / err[1], err[1] / err[2]]
    expect = 2.0 ** order
    assert all(abs(r / expect - 1.0) < 0.1 for r in ratios), (err, ratios)


def test_predictor_corrector_conserves_in_a_periodic_tube():
    # Periodic in i (the block exchanges with itself: both copies of the seam face see the
    # same stencil, so their fluxes telescope exactly), slip walls in j and k with no
    # transverse velocity: mass, x-momentum and energy stay fixed to round-off over 300 steps
    # of developing flow, at x_order 2.
    job = periodic_wave(24, 4.0e-5, 300, 1, x_order=2)
    o = ob.Oracle(job)
    u0, vol = o.conserved(0), o.iovars(0)[3]
    rc, _, _ = o.march(300)
    assert rc == 0
    u = o.conserved(0)
    for q in (0, 1, 4):
        tot0, tot = np.sum(u0[q] * vol), np.sum(u[q] * vol)
        assert abs(tot - tot0) <= 1e-14 * abs(tot0), q
    assert np.max(np.abs(u[0] - u0[0])) > 1e-3 * np.max(u0[0])  # the flow did move


def test_predictor_corrector_differs_from_rk3_by_a_third_order_term():
    # One step from the same state: the two schemes share the predictor and differ by
    # O(dt^3) (both are exact to second order): halving dt divides the gap by about 8.
    gaps = []
    for dt in (2.0e-5, 1.0e-5):
        a = final_state(periodic_wave(24, dt, 1, 1), 1)
        b = final_state(periodic_wave(24, dt, 1, 0), 1)
        gaps.append(np.max(np.abs(a[0] - b[0])))
    assert gaps[1] > 0 and 6.0 < gaps[0] / gaps[1] < 10.0, gaps
