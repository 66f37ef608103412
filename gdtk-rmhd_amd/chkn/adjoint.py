"""Differentiable unrolls of the chicken step: a torch.autograd.Function over libchkn's
gradient entry points (chkn_grad_forward / chkn_grad_backward), the wall-pressure loss and
the map from inflow (Mach, p, T) to the inflow state the solver uses.

    sim = Sim(job)
    u0 = initial_state(sim)                       # flat float64 device tensor
    mpt = torch.tensor([5.0, 2e3, 220.0], dtype=torch.float64, requires_grad=True)
    uN = unroll(sim, u0, inflow_from_mpt(mpt)[None], nsteps=500, checkpoint_every=50)
    loss = WallPressure(sim, [job])(uN)
    loss.backward()                               # gradients w.r.t. u0 (if it requires grad) and mpt

The time-step sequence of the forward pass is recorded and held fixed in the gradient;
pass it back through `dts` to replay the same sequence (finite-difference checks do this).
"""
import numpy as np
import torch

WALL_CODES = (0, 1, 2)


class Gas:
    """The gas model's constants as the library applies them (gas.cu:20-48): ratio of specific
    heats, gas constant, heat of reaction q and Cv = R / (gamma - 1)."""

    def __init__(self, gamma, R, q=0.0):
        self.gamma, self.R, self.q = float(gamma), float(R), float(q)
        self.cv = self.R / (self.gamma - 1.0)

    @classmethod
    def of(cls, job):
        """The gas model of a job (JobData.gas_model: gamma, R, q, alpha, Ti, Cv)."""
        gm = job.gas_model
        return cls(gm[0], gm[1], gm[2])


AIR = Gas(1.4, 287.1)


def inflow_from_mpt(mpt, YB=0.0, gas=AIR):
    """(Mach, p, T) -> (rho, e, YB, vx, vy, vz) with the operation order of update_from_pT
    (gas.cu:119-126) and velx = Mach * a: the state chkn.cases.freestream describes."""
    M, p, T = mpt[0], mpt[1], mpt[2]
    e = T * gas.cv - YB * gas.q
    rho = p / (T * gas.R)
    a = torch.sqrt(gas.gamma * gas.R * T)
    z = torch.zeros_like(T)
    return torch.stack([rho, e, z + YB, M * a, z, z])


def initial_state(sim):
    """The handle's current conserved state as a flat float64 device tensor."""
    u = torch.empty(sim.state_size, dtype=torch.float64, device="cuda")
    sim.read_conserved(u)
    return u


class _Unroll(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, inflow, sim, nsteps, checkpoint_every, dts):
        src = u0.detach().contiguous()
        sim.set_stream(torch.cuda.current_stream())  # ordered with torch's work on the device
        sim.write_conserved(src)
        w = inflow.detach().to("cpu", torch.float64).numpy()
        for s in range(w.shape[0]):
            sim.set_inflow_state(s, w[s])
        ctx.dts = sim.grad_forward(nsteps, checkpoint_every, dts)
        ctx.generation = sim.grad_generation
        uN = torch.empty_like(u0)
        sim.read_conserved(uN)
        sim.set_stream(None)
        ctx.sim = sim
        ctx.inflow_device = inflow.device
        return uN

    @staticmethod
    def backward(ctx, g):
        # The handle holds the checkpoints and dts of its latest unroll only: reversing an
        # earlier one with them would give a wrong gradient without any error.
        if ctx.sim.grad_generation != ctx.generation:
            raise RuntimeError("this unroll's tape was replaced by a later unroll on the same Sim: "
                               "call backward before the next unroll, or unroll on separate Sims")
        lam = g.contiguous()
        gu0 = torch.empty_like(lam)
        ctx.sim.set_stream(torch.cuda.current_stream())
        gin = ctx.sim.grad_backward(lam, gu0)
        ctx.sim.set_stream(None)
        return gu0, torch.as_tensor(gin, dtype=torch.float64, device=ctx.inflow_device), None, None, None, None


def unroll(sim, u0, inflow, nsteps, checkpoint_every=50, dts=None):
    """State after nsteps steps from u0 with the inflow states `inflow` ((nstates, 6):
    rho, e, YB, vx, vy, vz), differentiable in u0 and inflow.  The dts used, shape
    (nsteps, ninstances), are on the result's grad_fn.dts when it requires grad."""
    out = _Unroll.apply(u0, inflow, sim, int(nsteps), int(checkpoint_every), dts)
    return out


class WallPressure:
    """L = sum over the kminus faces of every block whose kminus side is a wall of
    p(cell at k = 0) * face area: the pressure force on the ramp wall."""

    def __init__(self, sim, jobs):
        self.gas = Gas.of(jobs[0])  # the instances of a handle share the library's gas model
        self.terms = []  # (offset of the block in the flat state, ncells, nic*njc, areas)
        off = 0
        b = 0
        for job in jobs:
            for blk in job.blocks:
                n = blk.nic * blk.njc * blk.nkc
                if blk.active and blk.bc[4] in WALL_CODES:
                    areas = torch.as_tensor(sim.boundary_face_areas(b, 4), dtype=torch.float64,
                                            device="cuda")
                    self.terms.append((off, n, blk.nic * blk.njc, areas))
                off += 6 * n
                b += 1

    def per_block(self, u):
        out = []
        for off, n, nw, areas in self.terms:
            U = u[off:off + 6 * n].view(6, n)[:, :nw]
            rho = U[0]
            vx, vy, vz = U[1] / rho, U[2] / rho, U[3] / rho
            e = U[4] / rho - 0.5 * (vx * vx + vy * vy + vz * vz)
            T = (e + (U[5] / rho) * self.gas.q) / self.gas.cv
            p = rho * self.gas.R * T
            out.append(torch.sum(p * areas))
        return out

    def __call__(self, u):
        return torch.stack(self.per_block(u)).sum()
