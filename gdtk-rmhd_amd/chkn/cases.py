"""Synthetic chicken workloads (all FP64, ideal air gamma=1.4, R=287.1).

* ``sod_box``        -- 3D Sod-type shock tube in a slip-walled box.
* ``ramp``           -- Mach 5 flow over a 15-degree compression ramp with a seeded 1%
                        density perturbation (sum of 8 Fourier modes).
* ``ramp_array``     -- the ramp split into an array of tensor-product uneven blocks with a
                        seeded exchange/wall connection graph.
* ``mixed_bc_array`` -- a small block array exercising every built boundary condition.
* ``forward_step``   -- Mach 3 flow over a forward-facing step (the reference's integration
                        example), 25 active blocks plus the inactive blocks inside the step.

Each function describes the job through chkn.prep and returns ``JobData``.
"""
import math

import numpy as np

from . import prep
from .geom import StructuredGrid, TFIVolume, Vector3, cell_centres
from .job import JobData

GAMMA, R_GAS = 1.4, 287.1
CV = R_GAS / (GAMMA - 1.0)


def _state_arrays(p, T, velx, vely=0.0, velz=0.0, rho=None):
    """Per-cell flow arrays with the prep's gas-state formulas (chkn_prep.py:328-334)."""
    p = np.asarray(p, dtype=float)
    T = np.asarray(T, dtype=float)
    shape = np.broadcast(p, T, np.asarray(velx)).shape
    out = {"p": np.broadcast_to(p, shape), "T": np.broadcast_to(T, shape)}
    out["e"] = CV * out["T"] + 0.0 * 0.0
    out["rho"] = out["p"] / (R_GAS * out["T"]) if rho is None else np.broadcast_to(rho, shape)
    out["a"] = np.sqrt(GAMMA * R_GAS * out["T"])
    out["YB"] = np.zeros(shape)
    out["velx"] = np.broadcast_to(np.asarray(velx, dtype=float), shape)
    out["vely"] = np.broadcast_to(np.asarray(vely, dtype=float), shape)
    out["velz"] = np.broadcast_to(np.asarray(velz, dtype=float), shape)
    return {k: np.ascontiguousarray(v, dtype=float) for k, v in out.items()}


def _common(max_step, cfl=0.5, dt_init=1.0e-6, max_time=1.0, title=""):
    prep.reset()
    c = prep.config
    c.title = title
    c.max_step = max_step
    c.max_time = max_time
    c.dt_init = dt_init
    c.cfl_count = 10
    c.x_order = 2
    c.flux_calc = "ausmdv"
    prep.add_cfl_value(0.0, cfl)
    prep.add_dt_plot(0.0, max_time, max_time)


def box_grid(nic, njc, nkc, L=(1.0, 0.5, 0.5), origin=(0.0, 0.0, 0.0)):
    x0, y0, z0 = origin
    Lx, Ly, Lz = L
    vol = TFIVolume(p000=Vector3(x0, y0, z0), p100=Vector3(x0 + Lx, y0, z0),
                    p110=Vector3(x0 + Lx, y0 + Ly, z0), p010=Vector3(x0, y0 + Ly, z0),
                    p001=Vector3(x0, y0, z0 + Lz), p101=Vector3(x0 + Lx, y0, z0 + Lz),
                    p111=Vector3(x0 + Lx, y0 + Ly, z0 + Lz), p011=Vector3(x0, y0 + Ly, z0 + Lz))
    return StructuredGrid(pvolume=vol, niv=nic + 1, njv=njc + 1, nkv=nkc + 1)


def sod_box(nic=64, njc=32, nkc=32, max_step=200, axis=0, title="3D Sod shock tube"):
    """Sod-type tube along `axis`: x<0.5 p=1e5 Pa T=348 K, else p=1e4 Pa T=278 K, at rest,
    slip walls on all six faces."""
    _common(max_step, title=title)
    dims = [nic, njc, nkc]
    L = [0.5, 0.5, 0.5]
    L[axis] = 1.0
    grid = box_grid(*dims, L=tuple(L))
    cc = cell_centres(grid)
    coord = (cc.x, cc.y, cc.z)[axis]
    high = coord < 0.5
    st = _state_arrays(np.where(high, 1.0e5, 1.0e4), np.where(high, 348.0, 278.0), 0.0)
    prep.FluidBlock(i=0, j=0, k=0, grid=grid, initialState=st)
    return JobData.from_prep()


def ramp_vertices(NI, NJ, NK, length=1.0, corner=0.25, angle_deg=15.0, top=0.5, span=0.25):
    """Vertices of the ramp domain: uniform in x and y; in z uniform between the wall
    z_w(x) = max(0, (x - corner) tan(angle)) and the top."""
    x = np.linspace(0.0, length, NI + 1)
    y = np.linspace(0.0, span, NJ + 1)
    t = np.linspace(0.0, 1.0, NK + 1)
    zw = np.maximum(0.0, (x - corner) * math.tan(math.radians(angle_deg)))
    X = np.broadcast_to(x[:, None, None], (NI + 1, NJ + 1, NK + 1))
    Y = np.broadcast_to(y[None, :, None], (NI + 1, NJ + 1, NK + 1))
    Z = zw[:, None, None] + t[None, None, :] * (top - zw)[:, None, None]
    Z = np.broadcast_to(Z, (NI + 1, NJ + 1, NK + 1))
    return StructuredGrid(x=X.copy(), y=Y.copy(), z=Z.copy())


def freestream(mach=5.0, p=2.0e3, T=220.0):
    a = math.sqrt(GAMMA * R_GAS * T)
    return prep.FlowState(p=p, T=T, velx=mach * a)


def perturbed_state(grid, fs, seed=1, amplitude=0.01, nmodes=8):
    """Freestream with rho = rho_inf (1 + amplitude * S / sum|a_m|), p held fixed."""
    rng = np.random.default_rng(seed)
    kvec = rng.integers(1, 5, size=(nmodes, 3))
    amp = rng.uniform(0.5, 1.0, size=nmodes)
    phase = rng.uniform(0.0, 2.0 * math.pi, size=nmodes)
    cc = cell_centres(grid)
    S = np.zeros_like(cc.x)
    for m in range(nmodes):
        S += amp[m] * np.sin(2.0 * math.pi * (kvec[m, 0] * cc.x + kvec[m, 1] * cc.y + kvec[m, 2] * cc.z)
                             + phase[m])
    rho_inf = fs.gas.p / (R_GAS * fs.gas.T)
    rho = rho_inf * (1.0 + amplitude * S / amp.sum())
    T = fs.gas.p / (R_GAS * rho)
    return _state_arrays(fs.gas.p, T, fs.vel.x, rho=rho)


def ramp(nic=256, njc=128, nkc=128, max_step=500, seed=1, mach=5.0, p=2.0e3, T=220.0,
         title="Mach 5 flow over a 15-degree ramp", viscous=False, TWall=300.0):
    """Ramp flow on one block.  viscous: viscous fluxes on, and the ramp surface a no-slip
    wall held at TWall (the side walls stay slip walls)."""
    _common(max_step, title=title + " (seed %d)" % seed)
    grid = ramp_vertices(nic, njc, nkc)
    fs = freestream(mach, p, T)
    st = perturbed_state(grid, fs, seed=seed)
    wall = prep.WallNoSlipFixedTBC(TWall) if viscous else prep.WallWithSlipBC()
    prep.config.viscous = bool(viscous)
    bcs = {"iminus": prep.InflowBC(fs), "iplus": prep.OutflowBC(), "kminus": wall,
           "kplus": prep.OutflowBC(), "jminus": prep.WallWithSlipBC(), "jplus": prep.WallWithSlipBC()}
    prep.FluidBlock(i=0, j=0, k=0, grid=grid, initialState=st, bcs=bcs)
    return JobData.from_prep()


def ramp_slab(i0=64, ni=32, nic=256, njc=128, nkc=128, max_step=500, seed=1, mach=5.0, p=2.0e3,
              T=220.0, title="Mach 5 ramp, one i-slab"):
    """The i-slab [i0, i0+ni) of the ramp block of ramp(nic, njc, nkc): what one GPU would own
    if the block were split along i over nic/ni GPUs (the default slab holds the ramp corner).
    Standalone, its i-sides take the freestream (inflow) and an outflow where the neighbouring
    slabs' exchange boundaries would be."""
    _common(max_step, title=title + " %d-%d of %d (seed %d)" % (i0, i0 + ni, nic, seed))
    grid = ramp_vertices(nic, njc, nkc).subgrid(i0, 0, 0, ni + 1, njc + 1, nkc + 1)
    fs = freestream(mach, p, T)
    st = perturbed_state(grid, fs, seed=seed)
    bcs = {"iminus": prep.InflowBC(fs), "iplus": prep.OutflowBC(), "kminus": prep.WallWithSlipBC(),
           "kplus": prep.OutflowBC(), "jminus": prep.WallWithSlipBC(), "jplus": prep.WallWithSlipBC()}
    prep.FluidBlock(i=0, j=0, k=0, grid=grid, initialState=st, bcs=bcs)
    return JobData.from_prep()


def uneven_sizes(seed, nib=5, njb=4, nkb=2, lo=16, hi=96):
    rng = np.random.default_rng(seed)
    return (rng.integers(lo, hi + 1, size=nib), rng.integers(lo, hi + 1, size=njb),
            rng.integers(lo, hi + 1, size=nkb))


def ramp_array(seed=7, nib=5, njb=4, nkb=2, lo=16, hi=96, max_step=300, p_exchange=0.8,
               sizes=None, i_walls=False, title="Mach 5 ramp on uneven blocks"):
    """Ramp flow on an nib x njb x nkb array of blocks whose slab sizes are seeded in
    [lo, hi]; each internal j- and k-interface is an exchange pair with probability
    p_exchange, else a pair of slip walls (plates along the stream).  Internal i-interfaces
    are exchange pairs unless i_walls: a wall across the Mach 5 stream leaves the block
    behind it expanding into vacuum, which ends the run with a bad cell within a few hundred
    steps."""
    _common(max_step, title=title + " (seed %d)" % seed)
    nics, njcs, nkcs = sizes if sizes is not None else uneven_sizes(seed, nib, njb, nkb, lo, hi)
    nib, njb, nkb = len(nics), len(njcs), len(nkcs)
    rng = np.random.default_rng(seed + 1000)
    NI, NJ, NK = int(sum(nics)), int(sum(njcs)), int(sum(nkcs))
    grid = ramp_vertices(NI, NJ, NK)
    fs = freestream()
    st = perturbed_state(grid, fs, seed=seed)
    # Interface kinds: True = exchange.  Index [dir][i][j][k] for the plus face of block ijk.
    link = {d: rng.random((nib, njb, nkb)) < p_exchange for d in range(3)}
    if not i_walls:
        link[0][:] = True
    io = np.concatenate([[0], np.cumsum(nics)])
    jo = np.concatenate([[0], np.cumsum(njcs)])
    ko = np.concatenate([[0], np.cumsum(nkcs)])
    for k in range(nkb):
        for j in range(njb):
            for i in range(nib):
                bcs = {}
                bcs["iminus"] = prep.InflowBC(fs) if i == 0 else (
                    prep.ExchangeBC() if link[0][i - 1, j, k] else prep.WallWithSlipBC())
                bcs["iplus"] = prep.OutflowBC() if i == nib - 1 else (
                    prep.ExchangeBC() if link[0][i, j, k] else prep.WallWithSlipBC())
                bcs["jminus"] = prep.WallWithSlipBC() if j == 0 else (
                    prep.ExchangeBC() if link[1][i, j - 1, k] else prep.WallWithSlipBC())
                bcs["jplus"] = prep.WallWithSlipBC() if j == njb - 1 else (
                    prep.ExchangeBC() if link[1][i, j, k] else prep.WallWithSlipBC())
                bcs["kminus"] = prep.WallWithSlipBC() if k == 0 else (
                    prep.ExchangeBC() if link[2][i, j, k - 1] else prep.WallWithSlipBC())
                bcs["kplus"] = prep.OutflowBC() if k == nkb - 1 else (
                    prep.ExchangeBC() if link[2][i, j, k] else prep.WallWithSlipBC())
                sub = grid.subgrid(i0=io[i], j0=jo[j], k0=ko[k], niv=nics[i] + 1, njv=njcs[j] + 1,
                                   nkv=nkcs[k] + 1)
                init = {key: v[io[i]:io[i + 1], jo[j]:jo[j + 1], ko[k]:ko[k + 1]] if np.ndim(v) == 3 else v
                        for key, v in st.items()}
                prep.FluidBlock(i=i, j=j, k=k, grid=sub, initialState=init, bcs=bcs)
    return JobData.from_prep()


def mixed_bc_array(max_step=20, seed=3, title="mixed boundary conditions"):
    """A 3x2x2 array of small warped blocks using inflow, outflow, slip walls, exchange
    (including a periodic wrap in j) and an inactive block."""
    _common(max_step, title=title, cfl=0.4)
    nics, njcs, nkcs = [6, 5, 4], [4, 6], [5, 3]
    NI, NJ, NK = sum(nics), sum(njcs), sum(nkcs)
    rng = np.random.default_rng(seed)
    g = ramp_vertices(NI, NJ, NK, angle_deg=10.0)
    # Warp the interior vertices a little so that faces are not planar.
    jit = 0.15 / max(NI, NJ, NK)
    for a in (g.vertices.x, g.vertices.y, g.vertices.z):
        a[1:-1, 1:-1, 1:-1] += jit * rng.uniform(-1, 1, size=a[1:-1, 1:-1, 1:-1].shape)
    fs = freestream(mach=2.5, p=5.0e3, T=260.0)
    st = perturbed_state(g, fs, seed=seed, amplitude=0.05)
    st["vely"] = st["velx"] * 0.1
    st["velz"] = -st["velx"] * 0.05
    io, jo, ko = np.cumsum([0] + nics), np.cumsum([0] + njcs), np.cumsum([0] + nkcs)
    for k in range(2):
        for j in range(2):
            for i in range(3):
                bcs = {"iminus": prep.InflowBC(fs) if i == 0 else prep.ExchangeBC(),
                       "iplus": prep.OutflowBC() if i == 2 else prep.ExchangeBC(),
                       "jminus": prep.ExchangeBC(), "jplus": prep.ExchangeBC(),
                       "kminus": prep.WallWithSlipBC() if k == 0 else prep.ExchangeBC(),
                       "kplus": prep.OutflowBC() if k == 1 else prep.ExchangeBC()}
                active = not (i == 2 and j == 1 and k == 1)
                if not active:
                    bcs = {}
                # Blocks next to the inactive one see it through exchange (its state is frozen).
                sub = g.subgrid(i0=io[i], j0=jo[j], k0=ko[k], niv=nics[i] + 1, njv=njcs[j] + 1,
                                nkv=nkcs[k] + 1)
                init = {key: v[io[i]:io[i + 1], jo[j]:jo[j + 1], ko[k]:ko[k + 1]] for key, v in st.items()}
                prep.FluidBlock(i=i, j=j, k=k, grid=sub, initialState=init, bcs=bcs, active=active)
    return JobData.from_prep()


def ramp_ensemble(seed=11, n=64, max_step=300, nic=(32, 96), njk=(32, 64)):
    """n independent single-block ramp jobs with seeded inflow Mach in [3, 8], p in
    [1e3, 1e4] Pa, T in [200, 300] K and grids of nic in nic-range by njc, nkc in njk-range."""
    rng = np.random.default_rng(seed)
    jobs = []
    for q in range(n):
        mach = float(rng.uniform(3.0, 8.0))
        p = float(rng.uniform(1.0e3, 1.0e4))
        T = float(rng.uniform(200.0, 300.0))
        ni = int(rng.integers(nic[0], nic[1] + 1))
        nj = int(rng.integers(njk[0], njk[1] + 1))
        nk = int(rng.integers(njk[0], njk[1] + 1))
        jobs.append(ramp(ni, nj, nk, max_step=max_step, seed=seed * 1000 + q, mach=mach, p=p, T=T,
                         title="ensemble member %d: Mach %.3f, p %.1f Pa, T %.2f K" % (q, mach, p, T)))
    return jobs


def viscous_plate(nics=(10, 8), njc=4, nkcs=(8, 6), max_step=30, seed=5, x_order=2, TWall=350.0,
                  title="viscous flow over a flat plate"):
    """Mach 2 air over a plate on a 2x1x2 array of warped blocks: inflow and outflow in i,
    a periodic exchange in j, the kminus plate a no-slip adiabatic wall (first block) and a
    no-slip wall at TWall (second block), outflow on top; exchange on the internal
    interfaces.  The initial state varies smoothly in density, energy and velocity."""
    _common(max_step, title=title + " (seed %d)" % seed, cfl=0.4, dt_init=1.0e-8)
    prep.config.viscous = True
    prep.config.x_order = x_order
    NI, NK = sum(nics), sum(nkcs)
    g = box_grid(NI, njc, NK, L=(0.02, 0.004, 0.01))
    rng = np.random.default_rng(seed)
    jit = 0.1 / max(NI, njc, NK)
    v = g.vertices
    for a, Lx in ((v.x, 0.02), (v.y, 0.004), (v.z, 0.01)):
        a[1:-1, 1:-1, 1:-1] += Lx * jit * rng.uniform(-1, 1, size=a[1:-1, 1:-1, 1:-1].shape)
    fs = freestream(mach=2.0, p=1.0e4, T=300.0)
    cc = cell_centres(g)
    X, Y, Z = cc.x / 0.02, cc.y / 0.004, cc.z / 0.01
    rho0 = fs.gas.p / (R_GAS * fs.gas.T)
    rho = rho0 * (1.0 + 0.03 * np.sin(2 * np.pi * (X + 0.5 * Z) + 0.3 * seed))
    T = fs.gas.T * (1.0 + 0.02 * np.cos(2 * np.pi * (Y + Z)))
    st = _state_arrays(rho * R_GAS * T, T, fs.vel.x * (1.0 + 0.05 * np.sin(2 * np.pi * Z)), rho=rho)
    st["vely"] = 0.02 * fs.vel.x * np.sin(2 * np.pi * (X + Y))
    st["velz"] = 0.01 * fs.vel.x * np.cos(2 * np.pi * (X - Z))
    io = np.cumsum([0] + list(nics))
    ko = np.cumsum([0] + list(nkcs))
    for k in range(2):
        for i in range(2):
            bcs = {"iminus": prep.InflowBC(fs) if i == 0 else prep.ExchangeBC(),
                   "iplus": prep.OutflowBC() if i == 1 else prep.ExchangeBC(),
                   "jminus": prep.ExchangeBC(), "jplus": prep.ExchangeBC(),
                   "kminus": prep.ExchangeBC() if k == 1 else (
                       prep.WallNoSlipAdiabaticBC() if i == 0 else prep.WallNoSlipFixedTBC(TWall)),
                   "kplus": prep.OutflowBC() if k == 1 else prep.ExchangeBC()}
            sub = g.subgrid(i0=io[i], j0=0, k0=ko[k], niv=nics[i] + 1, njv=njc + 1, nkv=nkcs[k] + 1)
            init = {key: val[io[i]:io[i + 1], :, ko[k]:ko[k + 1]] if np.ndim(val) == 3 else val
                    for key, val in st.items()}
            prep.FluidBlock(i=i, j=0, k=k, grid=sub, initialState=init, bcs=bcs)
    return JobData.from_prep()


def bl_profile_numbers(p_e=1.013e3, T_e=300.0, u_e=1390.0, T_peak=480.0, delta=0.004):
    """A laminar boundary-layer profile in the layout of InflowBoundaryLayerBC.data: p_e,
    then (y, T, velx) at 12 samples from the wall (y = 0, velx = 0, T = T_e) to the edge
    (y = delta, free-stream values), samples clustered toward the wall."""
    out = [p_e]
    for i in range(12):
        s = i / 11.0
        y = delta * s ** 1.4
        eta = y / delta
        T = T_e + (T_peak - T_e) * math.sin(math.pi * eta) * (1.0 - eta) * 1.6
        u = u_e * math.tanh(3.2 * eta) / math.tanh(3.2)
        out += [y, T, u]
    return out


def bl_plate(nics=(10, 8), njc=3, nkc=10, max_step=30, seed=4, x_order=2, twall_fun=True):
    """Supersonic flow over a plate entering through a laminar boundary-layer profile
    (the layout of examples/chicken/boundary-layer/short.py on two blocks): i along x,
    j across a thin periodic span (-z), k from the plate (y = 0) upward, clustered to the
    wall.  iminus: inflow_function laminar_boundary_layer; kminus: no-slip wall at a TWall
    varying along x (given per face, as the TWall files hold it) or 300 K; kplus: inflow;
    iplus: outflow; j: periodic exchange.  Initial state: the free stream with smooth
    variations.  Returns (JobData, profile numbers); the job reads the profile given."""
    _common(max_step, title="laminar boundary-layer inflow (seed %d)" % seed, cfl=0.4, dt_init=1.0e-8)
    prep.config.viscous = True
    prep.config.x_order = x_order
    L0, L, H, W = 0.5, 0.53, 0.012, 0.002
    NI = sum(nics)
    xs = L0 + (L - L0) * np.linspace(0.0, 1.0, NI + 1)
    ys = H * np.linspace(0.0, 1.0, nkc + 1) ** 1.5
    zs = W * (1.0 - np.linspace(0.0, 1.0, njc + 1))
    X, Z, Y = np.meshgrid(xs, zs, ys, indexing="ij")
    rng = np.random.default_rng(seed)
    X = X + 0.02 * (L - L0) / NI * rng.uniform(-1, 1, X.shape) * (ys[None, None, :] > 0)
    g = StructuredGrid(x=X, y=Y, z=Z)
    fs = prep.FlowState(p=1.013e3, T=300.0, velx=1390.0)
    cc = cell_centres(g)
    T = 300.0 * (1.0 + 0.03 * np.sin(2 * np.pi * (cc.x - L0) / (L - L0)) * np.exp(-cc.y / H))
    st = _state_arrays(1.013e3, T, 1390.0 * (1.0 - 0.3 * np.exp(-cc.y / (0.2 * H))))
    st["vely"] = 10.0 * np.sin(2 * np.pi * cc.y / H)
    twall = (lambda x, y, z: 300.0 + 150.0 * (x - L0) / (L - L0)) if twall_fun else 300.0
    io = np.cumsum([0] + list(nics))
    for i in range(2):
        bcs = {"iminus": prep.InflowFunctionBC("laminar_boundary_layer") if i == 0 else prep.ExchangeBC(),
               "iplus": prep.OutflowBC() if i == 1 else prep.ExchangeBC(),
               "jminus": prep.ExchangeBC(), "jplus": prep.ExchangeBC(),
               "kminus": prep.WallNoSlipFixedTBC(twall), "kplus": prep.InflowBC(fs)}
        sub = g.subgrid(i0=io[i], j0=0, k0=0, niv=nics[i] + 1, njv=njc + 1, nkv=nkc + 1)
        init = {key: val[io[i]:io[i + 1]] if np.ndim(val) == 3 else val for key, val in st.items()}
        prep.FluidBlock(i=i, j=0, k=0, grid=sub, initialState=init, bcs=bcs)
    prof = bl_profile_numbers()
    return JobData.from_prep(bl_profile=prof), prof


def sbp_channel(nics=(10, 8), njc=6, nkc=5, max_step=25, seed=3, flux_calc="sbp_asf", k_walls=False):
    """Smooth subsonic flow through a periodic channel on two warped blocks, for the
    central SBP-ASF flux (face.cu:268-342): inflow at iminus, outflow at iplus, periodic
    exchange in j and k (slip walls in k with k_walls), a smooth density, pressure and
    velocity wave."""
    _common(max_step, title="SBP-ASF channel (seed %d)" % seed, cfl=0.4, dt_init=2.0e-7)
    prep.config.flux_calc = flux_calc
    NI = sum(nics)
    g = box_grid(NI, njc, nkc, L=(0.2, 0.05, 0.04))
    rng = np.random.default_rng(seed)
    v = g.vertices
    for a, Lx in ((v.x, 0.2), (v.y, 0.05), (v.z, 0.04)):
        a[1:-1, 1:-1, 1:-1] += Lx * 0.05 / NI * rng.uniform(-1, 1, size=a[1:-1, 1:-1, 1:-1].shape)
    fs = freestream(mach=0.4, p=5.0e4, T=300.0)
    cc = cell_centres(g)
    X, Y, Z = cc.x / 0.2, cc.y / 0.05, cc.z / 0.04
    wave = np.sin(2 * np.pi * (X + 0.3 * Y)) * np.cos(2 * np.pi * Z)
    p = fs.gas.p * (1.0 + 0.02 * wave)
    T = fs.gas.T * (1.0 + 0.01 * wave)
    st = _state_arrays(p, T, fs.vel.x * (1.0 + 0.05 * wave))
    st["vely"] = 3.0 * np.sin(2 * np.pi * (X + Z))
    st["velz"] = 2.0 * np.cos(2 * np.pi * (X - Y))
    io = np.cumsum([0] + list(nics))
    for i in range(2):
        bcs = {"iminus": prep.InflowBC(fs) if i == 0 else prep.ExchangeBC(),
               "iplus": prep.OutflowBC() if i == 1 else prep.ExchangeBC(),
               "jminus": prep.ExchangeBC(), "jplus": prep.ExchangeBC(),
               "kminus": prep.WallWithSlipBC() if k_walls else prep.ExchangeBC(),
               "kplus": prep.WallWithSlipBC() if k_walls else prep.ExchangeBC()}
        sub = g.subgrid(i0=io[i], j0=0, k0=0, niv=nics[i] + 1, njv=njc + 1, nkv=nkc + 1)
        init = {key: val[io[i]:io[i + 1]] if np.ndim(val) == 3 else val for key, val in st.items()}
        prep.FluidBlock(i=i, j=0, k=0, grid=sub, initialState=init, bcs=bcs)
    return JobData.from_prep()


def k_inflow_box(nic=10, njc=6, nkc=14, max_step=40, seed=5):
    """A Mach 2 stream (p = 50 kPa, T = 300 K) entering a warped box through its kminus side
    and leaving through kplus; slip walls on the i sides, periodic in j.  The initial state is
    the stream with a smooth seeded variation of pressure, temperature and velocity."""
    _common(max_step, title="k inflow box (seed %d)" % seed, dt_init=5.0e-7)
    g = box_grid(nic, njc, nkc, L=(0.1, 0.06, 0.28))
    rng = np.random.default_rng(seed)
    v = g.vertices
    for a, L in ((v.x, 0.1), (v.y, 0.06), (v.z, 0.28)):
        a[1:-1, 1:-1, 1:-1] += L * 0.04 / nkc * rng.uniform(-1, 1, size=a[1:-1, 1:-1, 1:-1].shape)
    speed = 2.0 * math.sqrt(GAMMA * R_GAS * 300.0)
    inflow = prep.FlowState(p=5.0e4, T=300.0, velz=speed)
    cc = cell_centres(g)
    X, Y, Z = cc.x / 0.1, cc.y / 0.06, cc.z / 0.28
    wave = np.sin(2 * np.pi * (Z + 0.3 * X)) * np.cos(2 * np.pi * Y)
    st = _state_arrays(5.0e4 * (1.0 + 0.02 * wave), 300.0 * (1.0 + 0.01 * wave),
                       4.0 * np.sin(2 * np.pi * (X + Z)), 3.0 * np.cos(2 * np.pi * (Y - Z)),
                       speed * (1.0 + 0.03 * wave))
    bcs = {"iminus": prep.WallWithSlipBC(), "iplus": prep.WallWithSlipBC(),
           "jminus": prep.ExchangeBC(), "jplus": prep.ExchangeBC(),
           "kminus": prep.InflowBC(inflow), "kplus": prep.OutflowBC()}
    prep.FluidBlock(i=0, j=0, k=0, grid=g, initialState=st, bcs=bcs)
    return JobData.from_prep()


def ab_reacting(nics=(16, 12), njc=6, nkc=5, max_step=30, seed=2, reacting=True, k_walls=False):
    """The reacting A->B gas (gas.cu:25-48, examples/chicken/oblique-detonation-wave): a
    Mach 3 stream (p = 86.1 kPa, T = 300 K) into gas at 28.7 kPa holding a hot, partly
    burnt slug above the ignition temperature, on two blocks: inflow at iminus, outflow at
    iplus, slip walls in j, periodic in k (slip walls in k with k_walls)."""
    _common(max_step, title="A->B reacting gas (seed %d)" % seed, cfl=0.5, dt_init=5.0e-7)
    gm = prep.GasModel(gamma=6.0 / 5.0, R=287.0, q=300000.0, alpha=1000.0, Ti=362.58)
    prep.config.gas_model = gm
    prep.config.reacting = bool(reacting)
    NI = sum(nics)
    g = box_grid(NI, njc, nkc, L=(0.32, 0.08, 0.08))
    rng = np.random.default_rng(seed)
    v = g.vertices
    for a, Lx in ((v.x, 0.32), (v.y, 0.08), (v.z, 0.08)):
        a[1:-1, 1:-1, 1:-1] += Lx * 0.04 / NI * rng.uniform(-1, 1, size=a[1:-1, 1:-1, 1:-1].shape)
    inflow = prep.FlowState(p=86.1e3, T=300.0, velx=964.302)
    cc = cell_centres(g)
    bump = np.exp(-((cc.x - 0.16) / 0.05) ** 2) * (1.0 + 0.2 * np.sin(2 * np.pi * cc.z / 0.08))
    p = np.full_like(bump, 28.7e3)
    T = 300.0 + 320.0 * bump
    YB = 0.3 * bump
    st = {"p": p, "T": T, "YB": YB, "e": T * gm.Cv - YB * gm.q, "rho": p / (T * gm.R),
          "a": np.sqrt(gm.gamma * gm.R * T), "velx": 40.0 * bump, "vely": np.zeros_like(T),
          "velz": 5.0 * np.sin(2 * np.pi * cc.x / 0.32)}
    io = np.cumsum([0] + list(nics))
    for i in range(2):
        bcs = {"iminus": prep.InflowBC(inflow) if i == 0 else prep.ExchangeBC(),
               "iplus": prep.OutflowBC() if i == 1 else prep.ExchangeBC(),
               "jminus": prep.WallWithSlipBC(), "jplus": prep.WallWithSlipBC(),
               "kminus": prep.WallWithSlipBC() if k_walls else prep.ExchangeBC(),
               "kplus": prep.WallWithSlipBC() if k_walls else prep.ExchangeBC()}
        sub = g.subgrid(i0=io[i], j0=0, k0=0, niv=nics[i] + 1, njv=njc + 1, nkv=nkc + 1)
        init = {key: np.ascontiguousarray(val[io[i]:io[i + 1]]) for key, val in st.items()}
        prep.FluidBlock(i=i, j=0, k=0, grid=sub, initialState=init, bcs=bcs)
    return JobData.from_prep()


def forward_step(max_time=1.0e-3, dx=0.01, max_step=6000, dt_init=1.0e-6, cfl=0.5):
    """Mach 3 air (p = 101.325 kPa, T = 300 K) entering a 3 x 1 x 0.1 m channel from x = 0,
    with a step of height 0.2 m filling x >= 0.6 (the forward-facing-step example of the
    reference, examples/chicken/forward-facing-step).  Cells are dx cubes in x and y and two
    layers in z.  The array is 5 x 5 blocks of 0.6 x 0.2 m: column 0 (upstream of the step) and
    rows 1-4 above the step are active, the four blocks inside the step (row 0, columns 1-4)
    are inactive at rest.  Inflow on the upstream face, outflow at x = 3, slip walls
    elsewhere (the step's face and top, the channel floor, roof and side planes)."""
    _common(max_step, cfl=cfl, dt_init=dt_init, max_time=max_time,
            title="Forward-facing step with supersonic flow.")
    at_rest = prep.FlowState(p=101.325e3, T=300.0)
    inflow = prep.FlowState(p=101.325e3, T=300.0, velx=3.0 * at_rest.gas.a)
    ni, nj = int(round(0.6 / dx)), int(round(0.2 / dx))
    grid = box_grid(5 * ni, 5 * nj, 2, L=(3.0, 1.0, 0.1))
    inflow_bc = prep.InflowBC(inflow)
    for i in range(5):
        for j in range(5):
            active = i == 0 or j > 0
            bcs = {}
            if active:
                side = {"iminus": (i == 0, inflow_bc), "iplus": (i == 4, prep.OutflowBC()),
                        "jminus": (j == 0 or (j == 1 and i > 0), prep.WallWithSlipBC()),
                        "jplus": (j == 4, prep.WallWithSlipBC())}
                for name, (outer, bc) in side.items():
                    bcs[name] = bc if outer else prep.ExchangeBC()
                if i == 0 and j == 0:
                    bcs["iplus"] = prep.WallWithSlipBC()  # the step's upstream face
            sub = grid.subgrid(i0=i * ni, j0=j * nj, k0=0, niv=ni + 1, njv=nj + 1, nkv=3)
            prep.FluidBlock(i=i, j=j, k=0, grid=sub, initialState=inflow if active else at_rest,
                            bcs=bcs, active=active, label="ffs-%d-%d" % (i, j))
    return JobData.from_prep()


def shock_position(x, p, p_shock=1047025.0, rise=0.01):
    """Position of a normal shock on a line of cells (x, p ordered downstream-to-upstream or
    upstream-to-downstream): the first cell whose pressure exceeds the line's first value by
    `rise`, then linear interpolation between it and the cell before to the pressure p_shock
    (examples/chicken/test_chicken.py:53-66 states the measure).  Returns (x_shock, cell
    spacing there), or (None, None) when no cell rises."""
    x, p = np.asarray(x, float), np.asarray(p, float)
    up = np.nonzero(p > p[0] * (1.0 + rise))[0]
    if len(up) == 0 or up[0] == 0:
        return None, None
    s = up[0]
    return (x[s] - x[s - 1]) / (p[s] - p[s - 1]) * (p_shock - p[s - 1]) + x[s - 1], x[s] - x[s - 1]
