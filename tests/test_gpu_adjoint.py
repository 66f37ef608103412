"""Adjoint (reverse-mode) gradients of the unrolled step, on the GPU.

* flux level: the device vector-Jacobian product of the face flux (AUSMDV at x_order 1 and 2,
  and SBP-ASF) against the complex-step derivative of the oracle's flux:
  <J^T Fbar, v> = <Fbar, J v> to 1e-12 relative;
* unroll level: d(wall-pressure loss)/d(inflow Mach, p, T) against central finite
  differences (relative step 1e-7, time-step sequence frozen) to 1e-5 relative, and the
  derivative along a seeded random smooth direction of U0 to 1e-3 at step 1e-7 and 1e-5 at
  step 1e-9 (see _fd_rows_u0 for why the two);
* the gradient is deterministic and depends neither on the checkpoint interval nor on how
  many segments the forward pass keeps taped (bit for bit), and an ensemble member's gradient
  equals that of the job run alone (bit for bit).
"""
import numpy as np
import pytest

import oracle_bind as ob
from chkn import cases, prep
from chkn.job import JobData
from chkn.lib import Sim, flux_vjp_points

torch = pytest.importorskip("torch")
from chkn import adjoint  # noqa: E402

pytestmark = pytest.mark.gpu

FD_REL = 1.0e-7
FD_TOL = 1.0e-5


def _points(rng, n, speed):
    prims = np.empty((n, 4, 6))
    frames = np.empty((n, 9))
    for t in range(n):
        base_rho, base_e = rng.uniform(0.1, 1.0), rng.uniform(1.5e5, 3.0e5)
        for c in range(4):
            prims[t, c] = [base_rho * (1 + 0.2 * rng.uniform(-1, 1)), base_e * (1 + 0.2 * rng.uniform(-1, 1)),
                           0.0, speed * rng.uniform(-1, 1), speed * rng.uniform(-1, 1), speed * rng.uniform(-1, 1)]
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        t1 = rng.normal(size=3)
        t1 -= t1.dot(nrm) * nrm
        t1 /= np.linalg.norm(t1)
        frames[t] = np.concatenate([nrm, t1, np.cross(nrm, t1)])
    return prims, frames


@pytest.mark.parametrize("x_order", [0, 1, 2])  # 0: the SBP-ASF flux
def test_flux_vjp_matches_complex_step(x_order):
    rng = np.random.default_rng(10 + x_order)
    worst = 0.0
    for speed in (80.0, 400.0, 1500.0):
        prims, frames = _points(rng, 64, speed)
        fbar = rng.normal(size=(64, 6))
        b = flux_vjp_points(prims, frames, fbar, x_order=x_order)
        for t in range(64):
            v = rng.normal(size=(4, 6)) * (np.abs(prims[t]) + 1.0)
            v[:, 2] = 0.0  # YB = 0: the limiter's kink at zero differences
            _, jv = ob.face_flux_cs(prims[t], frames[t], v, x_order=x_order)
            lhs = float(np.sum(b[t] * v))
            rhs = float(np.dot(fbar[t], jv))
            scale = float(np.sum(np.abs(b[t] * v))) + 1e-300
            worst = max(worst, abs(lhs - rhs) / scale)
    assert worst < 1e-12, worst


def _vary_state(job, seed, scale):
    """Give the initial velocity and internal energy (hence pressure) smooth variations.  On
    a uniform field the scheme sits on its kinks -- |ab| = 0 in every limiter, |pL - pR| = 0
    in AUSMDV's switch -- where central differences of step 1e-7 average two one-sided
    slopes instead of measuring the derivative."""
    for b in job.blocks:
        io = b.iovars
        x, y, z = io[0], io[1], io[2]
        u = np.full_like(x, scale)
        io[7] = io[7] * (1.0 + 0.02 * np.sin(4.0 * x + 9.0 * y + 5.0 * z + 0.7 * seed))
        io[10] = io[10] + 0.02 * u * np.sin(7.0 * x + 3.0 * z + 0.3 * seed)
        io[11] = io[11] + 0.01 * u * np.sin(5.0 * x + 11.0 * y)
        io[12] = io[12] + 0.01 * u * np.cos(3.0 * x + 13.0 * z)
    return job


def _small_ramp(max_step=60, seed=2):
    cases.ramp(12, 4, 6, max_step=max_step, seed=seed)
    prep.config.dt_init = 2.0e-6
    return _vary_state(JobData.from_prep(), seed, 1486.8)


def _smooth_direction(prob, seed):
    """A seeded random direction for U0 that varies smoothly in space (random Fourier modes
    per conserved quantity), scaled by |U0| cell by cell."""
    rng = np.random.default_rng(seed)
    parts = []
    for job in prob.jobs:
        for b in job.blocks:
            x, y, z = (np.asarray(b.iovars[r]) for r in range(3))
            for q in range(6):
                f = np.zeros_like(x)
                for _ in range(4):
                    k = rng.integers(1, 6, size=3)
                    f += rng.normal() * np.sin(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + rng.uniform(0, 6.3))
                parts.append(f)
    v = torch.as_tensor(np.concatenate(parts), device="cuda")
    return v * prob.u0.abs()


class Problem:
    """A job with its loss as a function of (U0, inflow (Mach, p, T)) via the adjoint."""

    def __init__(self, jobs, nsteps, C, mpt):
        self.jobs = jobs
        self.sim = Sim(jobs=jobs) if len(jobs) > 1 else Sim(jobs[0])
        self.u0 = adjoint.initial_state(self.sim)
        self.loss_fn = adjoint.WallPressure(self.sim, jobs)
        self.nsteps, self.C = nsteps, C
        self.mpt = [np.asarray(m, dtype=np.float64) for m in mpt]
        self.dts = None
        self.gas = adjoint.Gas.of(jobs[0])
        # Each inflow flow state takes the (Mach, p, T) of its instance.
        self.state_inst = [q for q, j in enumerate(jobs) for _ in j.flow_states]

    def inflow(self, mpts):
        return torch.stack([adjoint.inflow_from_mpt(mpts[q], gas=self.gas) for q in self.state_inst])

    def loss(self, u0, mpts, C=None):
        uN = adjoint.unroll(self.sim, u0, self.inflow(mpts), self.nsteps, C or self.C, self.dts)
        if self.dts is None:
            self.dts = uN.grad_fn.dts
        return self.loss_fn(uN), uN

    def gradient(self, C=None):
        u0 = self.u0.clone().requires_grad_(True)
        mpts = [torch.tensor(m, dtype=torch.float64, requires_grad=True) for m in self.mpt]
        L, _ = self.loss(u0, mpts, C)
        L.backward()
        return float(L.detach()), u0.grad.clone(), [m.grad.clone() for m in mpts]

    def value(self, u0=None, mpts=None):
        with torch.no_grad():
            mp = [torch.tensor(m, dtype=torch.float64) for m in (mpts or self.mpt)]
            L, _ = self.loss(self.u0 if u0 is None else u0, mp)
        return float(L)


def _fd_rows_inflow(prob, grad_mpt, inst=0):
    rows = []
    for k in range(3):
        h = FD_REL * abs(prob.mpt[inst][k])
        mp = [m.copy() for m in prob.mpt]
        mm = [m.copy() for m in prob.mpt]
        mp[inst][k] += h
        mm[inst][k] -= h
        fd = (prob.value(mpts=mp) - prob.value(mpts=mm)) / (2 * h)
        rows.append(("dL/d" + "MpT"[k], fd, float(grad_mpt[inst][k]), FD_TOL))
    return rows


def _fd_rows_u0(prob, grad_u0, component=None):
    """Central differences along a smooth direction of U0.  The limited scheme has kinks
    (|ab| in each van Albada limiter, |pL - pR| and the upwind switch in AUSMDV) at which
    its derivative jumps; a perturbation of every cell by 1e-7 relative crosses a few of
    them, which costs central differences up to about 1e-4 of accuracy, while the step 1e-9
    does not (and roundoff still allows 1e-5 there).  Both are checked.  With `component`
    the direction moves that conserved quantity only (5: rho YB)."""
    v = _smooth_direction(prob, 7)
    tag = "U0"
    if component is not None:
        keep = []
        for job in prob.jobs:
            for b in job.blocks:
                n = b.nic * b.njc * b.nkc
                keep.append(np.repeat(np.arange(6) == component, n))
        v = v * torch.as_tensor(np.concatenate(keep), device="cuda")
        tag = "U0[%d]" % component
    g = float(torch.dot(grad_u0, v))
    rows = []
    for h, tol in ((FD_REL, 1e-3), (1e-9, FD_TOL)):
        fd = (prob.value(u0=prob.u0 + h * v) - prob.value(u0=prob.u0 - h * v)) / (2 * h)
        rows.append(("dL/d%s.v (step %.0e)" % (tag, h), fd, g, tol))
    return rows


def _assert_rows(rows):
    for name, fd, g, tol in rows:
        print("%-22s fd %.12e adjoint %.12e rel %.2e" % (name, fd, g, abs(fd - g) / max(abs(g), abs(fd))))
    for name, fd, g, tol in rows:
        assert abs(fd - g) <= tol * max(abs(g), abs(fd)), (name, fd, g)


# The U0 check uses a short unroll (4 steps), where the kinks are few.


def test_unroll_gradient_matches_finite_differences():
    job = _small_ramp()
    prob = Problem([job], nsteps=40, C=7, mpt=[[5.0, 2.0e3, 220.0]])
    L, gu0, gm = prob.gradient()
    assert L > 0 and np.all(np.isfinite(gm[0].numpy()))
    assert abs(prob.value() - L) <= 1e-15 * abs(L)  # replay with the recorded dts
    rows = _fd_rows_inflow(prob, gm)
    short = Problem([job], nsteps=4, C=3, mpt=[[5.0, 2.0e3, 220.0]])
    _, gu0s, _ = short.gradient()
    _assert_rows(rows + _fd_rows_u0(short, gu0s))


def test_gradient_through_exchange_walls_outflow_and_inactive_block():
    cases.mixed_bc_array(max_step=40)
    job = _vary_state(JobData.from_prep(), 3, 808.0)
    prob = Problem([job], nsteps=12, C=5, mpt=[[2.5, 5.0e3, 260.0]])
    _, _, gm = prob.gradient()
    rows = _fd_rows_inflow(prob, gm)
    short = Problem([job], nsteps=4, C=3, mpt=[[2.5, 5.0e3, 260.0]])
    _, gu0s, _ = short.gradient()
    _assert_rows(rows + _fd_rows_u0(short, gu0s))


def test_gradient_is_deterministic_and_independent_of_checkpoint_interval():
    job = _small_ramp()
    prob = Problem([job], nsteps=25, C=4, mpt=[[5.0, 2.0e3, 220.0]])
    _, g1, m1 = prob.gradient()
    _, g2, m2 = prob.gradient()
    _, g3, m3 = prob.gradient(C=25)
    assert torch.equal(g1, g2) and torch.equal(m1[0], m2[0])
    assert torch.equal(g1, g3) and torch.equal(m1[0], m3[0])


def test_resident_tapes_give_the_recomputed_gradient():
    # The forward pass tapes the last checkpoint segments that fit the tape budget and the
    # backward pass recomputes the others.  All seven segments resident, none, and three give
    # the same gradient bit for bit: the taped states are the forward's own and the recomputed
    # ones repeat its arithmetic.
    job = _small_ramp()
    prob = Problem([job], nsteps=25, C=4, mpt=[[5.0, 2.0e3, 220.0]])
    seg = 6 * (12 * 4 * 6) * (3 * 4 + 1) * 8  # bytes of one segment's tape
    out = []
    for budget, nres in ((None, 7), (0, 0), (3 * seg + seg // 2, 3), (None, 7)):
        prob.sim.set_tape_budget(budget)
        L, g, m = prob.gradient()
        assert prob.sim.grad_resident_segments == nres, budget
        out.append((L, g, m[0]))
    for L, g, m in out[1:]:
        assert L == out[0][0] and torch.equal(g, out[0][1]) and torch.equal(m, out[0][2])


def test_ensemble_member_gradient_equals_single_job():
    jobs = [_small_ramp(seed=2), _small_ramp(seed=5)]
    mpt = [[5.0, 2.0e3, 220.0], [5.0, 2.0e3, 220.0]]
    ens = Problem(jobs, nsteps=20, C=6, mpt=mpt)
    _, gu_e, gm_e = ens.gradient()
    n0 = 6 * 12 * 4 * 6
    for q in range(2):
        one = Problem([jobs[q]], nsteps=20, C=6, mpt=[mpt[q]])
        _, gu, gm = one.gradient()
        assert torch.equal(gm[0], gm_e[q])
        assert torch.equal(gu, gu_e[q * n0:(q + 1) * n0])


def _ramp_x1():
    cases.ramp(12, 4, 6, max_step=60, seed=4)
    prep.config.dt_init = 2.0e-6
    prep.config.x_order = 1
    return _vary_state(JobData.from_prep(), 4, 1486.8)


def test_fused_gather_matches_the_separate_gather():
    # The fused form of a reversed stage (the k patches finish their own cells) against the
    # face pass of every direction followed by the separate gather launch: the same gradient
    # bit for bit.  The ramp (inflow, outflow, walls) at both reconstruction orders and with
    # the predictor-corrector; the SBP-ASF channel (exchange in i, periodic in j); the reacting
    # A->B pair of blocks; a two-job ensemble.  The mixed-BC array exchanges through k sides
    # and runs the separate form whatever is asked.
    ramp_mpt = [5.0, 2.0e3, 220.0]
    ab = cases.ab_reacting(max_step=40, k_walls=True)
    ab_mpt = [964.302 / (adjoint.Gas.of(ab).gamma * adjoint.Gas.of(ab).R * 300.0) ** 0.5, 86.1e3, 300.0]
    runs = [([_small_ramp()], [ramp_mpt]), ([_ramp_x1()], [ramp_mpt]), ([_pc(_small_ramp())], [ramp_mpt]),
            ([cases.sbp_channel(max_step=40, k_walls=True)], [[0.4, 5.0e4, 300.0]]), ([ab], [ab_mpt]),
            ([_small_ramp(seed=2), _small_ramp(seed=5)], [ramp_mpt, ramp_mpt])]
    for jobs, mpts in runs:
        prob = Problem(jobs, nsteps=10, C=4, mpt=mpts)
        assert prob.sim.gather_form == 0
        L, g, m = prob.gradient()
        prob.sim.set_gather_form(True)
        assert prob.sim.gather_form == 1
        L2, g2, m2 = prob.gradient()
        print("max |fused - separate| / max |g|: %.3e" % float((g - g2).abs().max() / g.abs().max()))
        assert L == L2 and torch.equal(g, g2)
        assert all(torch.equal(a, b) for a, b in zip(m, m2))
        assert bool(torch.any(g != 0)) and all(bool(torch.any(a != 0)) for a in m)
        prob.sim.close()
    cases.mixed_bc_array(max_step=40)
    sim = Sim(JobData.from_prep())
    assert sim.gather_form == 1
    sim.close()


@pytest.mark.parametrize("dims", [(5, 3, 2), (13, 16, 13), (14, 17, 27), (30, 5, 14)])
def test_fused_gather_on_small_and_ragged_blocks(dims):
    # Patch edges in the fused form: k lines of 2 cells (both ends in one cell pair), lines
    # of exactly one patch's 13 owned cells, lines whose last patch owns one cell, i and j
    # lines shorter or longer than a 16-face patch.  Fused against separate, bit for bit.
    cases.ramp(*dims, max_step=40, seed=4)
    prep.config.dt_init = 2.0e-6
    job = _vary_state(JobData.from_prep(), 4, 1486.8)
    prob = Problem([job], nsteps=6, C=4, mpt=[[5.0, 2.0e3, 220.0]])
    assert prob.sim.gather_form == 0
    L, g, m = prob.gradient()
    prob.sim.set_gather_form(True)
    L2, g2, m2 = prob.gradient()
    assert L == L2 and torch.equal(g, g2) and torch.equal(m[0], m2[0])
    assert bool(torch.any(g != 0)) and bool(torch.all(m[0] != 0))
    prob.sim.close()


def _inflow_z(mpt):
    """(Mach, p, T) -> the inflow state of a stream along +z."""
    w = adjoint.inflow_from_mpt(mpt)
    return torch.stack([w[0], w[1], w[2], w[4], w[5], w[3]])


def test_fused_gather_with_inflow_through_a_k_side():
    # In the fused form the k patches finish their cells, including the accumulator of each
    # inflow line of a k side (the inflow state's gradient).  A Mach 2 stream entering through
    # kminus: against the separate gather bit for bit, and dL/d(Mach, p, T) of a linear loss on
    # the final energy and z momentum against central differences.
    job = cases.k_inflow_box(max_step=40)
    sim = Sim(job)
    assert sim.gather_form == 0
    u0 = adjoint.initial_state(sim)
    n = 10 * 6 * 14
    wc = torch.as_tensor(np.random.default_rng(3).uniform(0.5, 1.5, size=n), device="cuda")
    mpt0 = [2.0, 5.0e4, 300.0]
    dts = [None]

    def run(mp, u=None):
        uN = adjoint.unroll(sim, u0 if u is None else u, _inflow_z(mp)[None], 12, 5, dts[0])
        if dts[0] is None:
            dts[0] = uN.grad_fn.dts
        U = uN.view(6, n)
        return (U[4] * wc).sum() + 1.0e3 * (U[3] * wc).sum()

    def grad(separate):
        sim.set_gather_form(separate)
        mp = torch.tensor(mpt0, dtype=torch.float64, requires_grad=True)
        u = u0.clone().requires_grad_(True)
        L = run(mp, u)
        L.backward()
        return float(L.detach()), u.grad.clone(), mp.grad.clone()

    L1, gu1, gm1 = grad(False)
    L2, gu2, gm2 = grad(True)
    assert sim.gather_form == 1
    assert L1 == L2 and torch.equal(gu1, gu2) and torch.equal(gm1, gm2)
    assert bool(torch.all(gm1 != 0))
    sim.set_gather_form(False)
    # Two relative steps, as for U0: perturbations of 1e-7 cross a few limiter kinks here
    # (observed 1.3e-5 for T at 1e-7, 2.6e-6 at 1e-8; steps of 1e-5 and 1e-6 err by 1e-4 to
    # 1e-3 and 1e-9 by roundoff).
    rows = []
    for k in range(3):
        for rel, tol in ((FD_REL, 1e-4), (1e-8, FD_TOL)):
            h = rel * mpt0[k]
            mp_p, mp_m = list(mpt0), list(mpt0)
            mp_p[k] += h
            mp_m[k] -= h
            with torch.no_grad():
                fd = (float(run(torch.tensor(mp_p, dtype=torch.float64))) -
                      float(run(torch.tensor(mp_m, dtype=torch.float64)))) / (2 * h)
            rows.append(("dL/d%s (step %.0e)" % ("MpT"[k], rel), fd, float(gm1[k]), tol))
    _assert_rows(rows)
    sim.close()


def test_first_order_gradient_matches_finite_differences():
    cases.mixed_bc_array(max_step=40)
    prep.config.x_order = 1
    job = _vary_state(JobData.from_prep(), 5, 808.0)
    prob = Problem([job], nsteps=10, C=4, mpt=[[2.5, 5.0e3, 260.0]])
    _, _, gm = prob.gradient()
    rows = _fd_rows_inflow(prob, gm)
    short = Problem([job], nsteps=4, C=3, mpt=[[2.5, 5.0e3, 260.0]])
    _, gu0s, _ = short.gradient()
    _assert_rows(rows + _fd_rows_u0(short, gu0s))


def test_sbp_asf_gradient_matches_finite_differences():
    # The SBP-ASF flux's adjoint in the unroll: the smooth channel on two warped blocks
    # (inflow, outflow, periodic exchange in j, slip walls in k carrying the loss).
    job = cases.sbp_channel(max_step=40, k_walls=True)
    prob = Problem([job], nsteps=20, C=6, mpt=[[0.4, 5.0e4, 300.0]])
    L, _, gm = prob.gradient()
    assert L > 0 and abs(prob.value() - L) <= 1e-15 * abs(L)
    rows = _fd_rows_inflow(prob, gm)
    short = Problem([job], nsteps=4, C=3, mpt=[[0.4, 5.0e4, 300.0]])
    _, gu0s, _ = short.gradient()
    _assert_rows(rows + _fd_rows_u0(short, gu0s))


@pytest.mark.parametrize("time_scheme", [0, 1])
def test_reacting_gradient_matches_finite_differences(time_scheme):
    # The A->B gas (libchkn_ab) with the reaction after every step: the hot slug above the
    # ignition temperature reacts, the Mach 3 inflow at 300 K does not.  The reaction's adjoint
    # (chem_vjp) sits between the reversed steps; a recomputed segment repeats the reaction.
    # Both time schemes: TVD-RK3 and the predictor-corrector.
    job = cases.ab_reacting(max_step=40, k_walls=True)
    job.time_scheme = time_scheme
    gas = adjoint.Gas.of(job)
    mpt = [[964.302 / (gas.gamma * gas.R * 300.0) ** 0.5, 86.1e3, 300.0]]
    prob = Problem([job], nsteps=12, C=5, mpt=mpt)
    L, g1, gm = prob.gradient()
    assert L > 0 and abs(prob.value() - L) <= 1e-15 * abs(L)
    rows = _fd_rows_inflow(prob, gm)
    short = Problem([job], nsteps=4, C=3, mpt=mpt)
    _, gu0s, _ = short.gradient()
    # rho YB is non-zero only in the burning slug: along it the gradient runs through the
    # reaction's factor f = exp(-alpha dt) once per step (about 2 % over the four steps).
    _assert_rows(rows + _fd_rows_u0(short, gu0s) + _fd_rows_u0(short, gu0s, component=5))
    # Every segment recomputed: the same gradient bit for bit.
    prob.sim.set_tape_budget(0)
    L2, g2, gm2 = prob.gradient()
    assert prob.sim.grad_resident_segments == 0
    assert L2 == L and torch.equal(g1, g2) and torch.equal(gm[0], gm2[0])


def _pc(job):
    job.time_scheme = 1  # the two-stage predictor-corrector
    return job


def test_predictor_corrector_gradient_matches_finite_differences():
    # The adjoint of U1 = U0 + dt L0, U = U0 + dt/2 (L0 + L1): the ramp and the mixed-BC array
    # (exchange, walls, outflow, an inactive block), against central differences of the
    # predictor-corrector forward (which equals its oracle bit for bit).
    job = _pc(_small_ramp())
    prob = Problem([job], nsteps=40, C=7, mpt=[[5.0, 2.0e3, 220.0]])
    L, _, gm = prob.gradient()
    assert abs(prob.value() - L) <= 1e-15 * abs(L)
    rows = _fd_rows_inflow(prob, gm)
    short = Problem([job], nsteps=4, C=3, mpt=[[5.0, 2.0e3, 220.0]])
    _, gu0s, _ = short.gradient()
    rows += _fd_rows_u0(short, gu0s)
    cases.mixed_bc_array(max_step=40)
    mixed = _pc(_vary_state(JobData.from_prep(), 3, 808.0))
    pm = Problem([mixed], nsteps=12, C=5, mpt=[[2.5, 5.0e3, 260.0]])
    _, _, gmm = pm.gradient()
    _assert_rows(rows + _fd_rows_inflow(pm, gmm))


def test_predictor_corrector_gradient_is_independent_of_taping():
    # Two stage states per step on a tape (2C + 1 slots): every segment resident, none, and
    # three, and two checkpoint intervals give the same gradient bit for bit.
    job = _pc(_small_ramp())
    prob = Problem([job], nsteps=25, C=4, mpt=[[5.0, 2.0e3, 220.0]])
    seg = 6 * (12 * 4 * 6) * (2 * 4 + 1) * 8  # bytes of one segment's tape
    out = []
    for budget, nres in ((None, 7), (0, 0), (3 * seg + seg // 2, 3)):
        prob.sim.set_tape_budget(budget)
        L, g, m = prob.gradient()
        assert prob.sim.grad_resident_segments == nres, budget
        out.append((L, g, m[0]))
    prob.sim.set_tape_budget(0)
    L, g, m = prob.gradient(C=9)
    out.append((L, g, m[0]))
    for L, g, m in out[1:]:
        assert L == out[0][0] and torch.equal(g, out[0][1]) and torch.equal(m, out[0][2])
    # Not the RK3 gradient: the scheme is the predictor-corrector's.
    rk = Problem([_small_ramp()], nsteps=25, C=4, mpt=[[5.0, 2.0e3, 220.0]])
    _, grk, _ = rk.gradient()
    assert not torch.equal(grk, out[0][1])


def test_work_on_the_callers_stream_gives_identical_results():
    # chkn_set_stream: a torch side stream, and the null stream with hipGraph replay requested
    # (it runs without capture there), give the same fields as the handle's own stream.
    job = cases.ramp(24, 8, 10, max_step=20)
    ref = Sim(job)
    ref.step(20)
    side = Sim(job)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        side.set_stream(torch.cuda.current_stream())
        side.step(7)
        side.step(13)
        side.set_stream(None)
    null = Sim(job)
    null.set_launch(use_graph=True)
    null.set_stream(0)
    null.step(20)
    null.set_stream(None)
    for sim in (side, null):
        assert np.array_equal(sim.conserved(0), ref.conserved(0))
        sim.close()
    ref.close()


def test_gradient_at_full_size_matches_finite_differences():
    # The differentiation workload at its real size: one 128x64x64 block, Mach 5 ramp, loss
    # after a 500-step unroll checkpointed every 50 steps.  dL/d(Mach, p, T) against central
    # differences (relative step 1e-7, recorded time steps).  Over 500 steps the perturbed runs
    # cross more limiter and upwind kinks than the short unrolls above, so the bound is 1e-4.
    job = cases.ramp(128, 64, 64, max_step=1 << 30)
    p, T, _, vx = (float(v) for v in job.flow_states[0][:4])
    prob = Problem([job], nsteps=500, C=50, mpt=[[vx / (1.4 * 287.1 * T) ** 0.5, p, T]])
    L, gu0, gm = prob.gradient()
    assert L > 0 and np.all(np.isfinite(gm[0].numpy())) and bool(torch.isfinite(gu0).all())
    rows = [(n, fd, g, 1e-4) for n, fd, g, _ in _fd_rows_inflow(prob, gm)]
    _assert_rows(rows)


def test_ensemble_gradient_at_full_size_matches_finite_differences():
    # The batched differentiation workload at its real size: 64 independent ramp jobs (Mach
    # 3-8, grids 32-96 x 32-64 x 32-64, 9.0M cells), per-job time steps, 300-step unroll
    # checkpointed every 50 steps.  Two members' dL/d(Mach, p, T) against central differences
    # of the whole ensemble's loss (the members are independent, so the other members' terms
    # cancel).  Over 300 steps of Mach 3-8 flow the perturbed runs cross limiter and upwind
    # kinks, and central differences converge to the adjoint value as the step shrinks (for
    # job 37's dL/dT, 3.9e-3, 6.1e-4, 1.4e-4 at relative steps 1e-6, 1e-7, 1e-8), until
    # round-off of the whole-ensemble loss (3.9e5) takes over near 1e-8.  So: every component
    # within 1e-3 at step 1e-7, and the closer of steps 1e-7 and 1e-8 within 2e-4.
    jobs = cases.ramp_ensemble(seed=11, n=64, max_step=1 << 30)
    mpts = []
    for j in jobs:
        p, T, _, vx = (float(v) for v in j.flow_states[0][:4])
        mpts.append([vx / (1.4 * 287.1 * T) ** 0.5, p, T])
    prob = Problem(jobs, nsteps=300, C=50, mpt=mpts)
    L, gu0, gm = prob.gradient()
    assert L > 0 and all(np.all(np.isfinite(g.numpy())) for g in gm)
    for q in (0, 37):
        for k in range(3):
            g = float(gm[q][k])
            err = []
            for rel in (1e-7, 1e-8):
                h = rel * abs(prob.mpt[q][k])
                mp = [m.copy() for m in prob.mpt]
                mm = [m.copy() for m in prob.mpt]
                mp[q][k] += h
                mm[q][k] -= h
                fd = (prob.value(mpts=mp) - prob.value(mpts=mm)) / (2 * h)
                err.append(abs(fd - g) / abs(g))
                print("job %d dL/d%s step %.0e fd %.10e adjoint %.10e rel %.1e" % (q, "MpT"[k], rel, fd, g, err[-1]))
            assert err[0] <= 1e-3 and min(err) <= 2e-4, (q, k, err)


def test_backward_of_a_replaced_unroll_is_refused():
    # The handle keeps the checkpoints and dts of its latest unroll only.  Chaining two unrolls
    # on one Sim and differentiating through both would reverse the first with the second's
    # tape: the backward must say so instead of returning a wrong gradient.
    job = _small_ramp(max_step=1000)
    sim = Sim(job)
    u0 = adjoint.initial_state(sim).requires_grad_(True)
    inflow = torch.as_tensor(np.asarray([sim.get_inflow_state(0)]), dtype=torch.float64,
                             device="cuda").requires_grad_(True)
    g0 = sim.grad_generation
    u1 = adjoint.unroll(sim, u0, inflow, 3, 2)
    assert sim.grad_generation == g0 + 1
    u2 = adjoint.unroll(sim, u1, inflow, 3, 2)
    loss = adjoint.WallPressure(sim, [job])(u2)
    with pytest.raises(RuntimeError, match="replaced by a later unroll"):
        loss.backward()
    # One unroll per Sim is fine.
    u1 = adjoint.unroll(sim, u0, inflow, 3, 2)
    adjoint.WallPressure(sim, [job])(u1).backward()
    assert torch.isfinite(u0.grad).all()
    sim.close()
