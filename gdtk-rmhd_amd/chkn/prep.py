"""Job preparation for the chicken step: the user-facing API of chkn_prep.py.

A job script (Python) is executed with the names of this module in scope: ``config``,
``FlowState``, ``FluidBlock``, ``makeFBArray``, the boundary-condition classes,
``add_cfl_value``, ``add_dt_plot`` and the geometry helpers ``Vector3``, ``TFIVolume``,
``StructuredGrid``.  ``write_initial_files`` then writes config.json, the grid and flow
files and times.data in the formats the solver reads (chkn_prep.py:856-932).

Names, defaults and the JSON text follow chkn_prep.py (GasModel :67-88, GlobalConfig
:93-246, GasState/FlowState :311-374, boundary conditions :383-599, FluidBlock :602-780,
makeFBArray :784-851).  Two deliberate differences: cell centres and volumes written to
the flow file are in the flow file's own cell order (i fastest) -- the solver recomputes
both from the grid (block.cu:640-647) -- and an initial state may also be a dict of
per-cell arrays, so large blocks need no per-cell Python call.

Use ``reset()`` to start a new job in the same process.
"""
import collections
import gzip
import json
import math
import os
from copy import copy

import numpy as np

from .geom import StructuredGrid, TFIVolume, Vector3, cell_centres, hex_volumes

IOVAR_NAMES = ["posx", "posy", "posz", "vol", "p", "T", "rho", "e", "YB", "a",
               "velx", "vely", "velz"]

FaceNames = "iminus iplus jminus jplus kminus kplus"
Face = collections.namedtuple("_", FaceNames)(*range(6))
FaceList = FaceNames.split()


class GasModel:
    """Ideal gas; must match the solver's compiled model (gas.cu:86-93)."""

    __slots__ = "gamma", "R", "Cv", "q", "alpha", "Ti"

    def __init__(self, gamma=1.4, R=287.1, q=0.0, alpha=0.0, Ti=0.0):
        self.gamma = gamma
        self.R = R
        self.Cv = R / (gamma - 1.0)
        self.q = q
        self.alpha = alpha
        self.Ti = Ti

    def to_json(self):
        return '{"gamma": %g, "R": %g, "Cv": %g, "q": %g, "alpha": %g, "Ti": %g}' % (
            self.gamma, self.R, self.Cv, self.q, self.alpha, self.Ti)


class GlobalConfig:
    """Simulation control parameters; the script edits the module-level ``config``."""

    count = 0
    __slots__ = ("job_name", "title", "gas_model", "nib", "njb", "nkb", "blk_ids", "nics", "njcs",
                 "nkcs", "dt_init", "cfl_list", "cfl_count", "print_count", "dt_plot_list",
                 "max_time", "max_step", "x_order", "t_order", "flux_calc", "viscous", "reacting",
                 "source_terms", "iovar_names", "threads_per_gpu_block")

    def __init__(self):
        if GlobalConfig.count >= 1:
            raise Exception("Already have a GlobalConfig object defined.")
        self.job_name = ""
        self.title = "Chicken Simulation."
        self.gas_model = GasModel()
        self.nib = self.njb = self.nkb = None
        self.blk_ids = None
        self.nics = self.njcs = self.nkcs = None
        self.dt_init = 1.0e-6
        self.cfl_list = []
        self.cfl_count = 10
        self.print_count = 20
        self.dt_plot_list = []
        self.max_time = 1.0e-3
        self.max_step = 100
        self.x_order = 2
        self.t_order = 2
        self.flux_calc = "ausmdv"
        self.reacting = False
        self.viscous = False
        self.source_terms = "none"
        self.threads_per_gpu_block = 128
        self.iovar_names = list(IOVAR_NAMES)
        GlobalConfig.count += 1

    def write(self, fp):
        """Global part of config.json (chkn_prep.py:151-196)."""
        fp.write('  "title": "%s",\n' % self.title)
        fp.write('  "gas_model": %s,\n' % self.gas_model.to_json())
        fp.write('  "max_time": %e,\n' % self.max_time)
        fp.write('  "max_step": %d,\n' % self.max_step)
        fp.write('  "dt_init": %e,\n' % self.dt_init)
        if len(self.cfl_list) == 0:
            add_cfl_value(0.0, 0.5)
        n_cfl = len(self.cfl_list)
        fp.write('  "n_cfl": %d,\n' % n_cfl)
        fp.write('  "cfl_times": %s,\n' % json.dumps([c[0] for c in self.cfl_list]))
        vals = [c[1] for c in self.cfl_list]
        assert min(vals) > 0.0, "Require positive nonzero cfl values."
        fp.write('  "cfl_values": %s,\n' % json.dumps(vals))
        fp.write('  "cfl_count": %d,\n' % self.cfl_count)
        fp.write('  "print_count": %d,\n' % self.print_count)
        fp.write('  "x_order": %d,\n' % self.x_order)
        fp.write('  "t_order": %d,\n' % self.t_order)
        fp.write('  "flux_calc": "%s",\n' % self.flux_calc)
        fp.write('  "reacting": %s,\n' % json.dumps(self.reacting))
        fp.write('  "viscous": %s,\n' % json.dumps(self.viscous))
        fp.write('  "source_terms": "%s",\n' % self.source_terms)
        fp.write('  "threads_per_gpu_block": %d,\n' % self.threads_per_gpu_block)
        if len(self.dt_plot_list) == 0:
            add_dt_plot(0.0, self.max_time, self.max_time)
        n = len(self.dt_plot_list)
        fp.write('  "n_dt_plot": %d,\n' % n)
        fp.write('  "t_change": %s,\n' % json.dumps([d[0] for d in self.dt_plot_list]))
        fp.write('  "dt_plot": %s,\n' % json.dumps([d[1] for d in self.dt_plot_list]))
        fp.write('  "dt_hist": %s,\n' % json.dumps([d[2] for d in self.dt_plot_list]))
        fp.write('  "iovar_names": %s,\n' % json.dumps(self.iovar_names))

    def check_array_of_fluid_blocks(self):
        """Place blocks in the block array; slab sizes must agree (chkn_prep.py:198-246)."""
        imax = jmax = kmax = 0
        for b in fluidBlocksList:
            imax, jmax, kmax = max(b.i, imax), max(b.j, jmax), max(b.k, kmax)
        self.nib, self.njb, self.nkb = imax + 1, jmax + 1, kmax + 1
        self.blk_ids = np.full((self.nib, self.njb, self.nkb), -1, dtype=int)
        self.nics = np.zeros(self.nib, dtype=int)
        self.njcs = np.zeros(self.njb, dtype=int)
        self.nkcs = np.zeros(self.nkb, dtype=int)
        for b in fluidBlocksList:
            i, j, k = b.i, b.j, b.k
            if self.blk_ids[i, j, k] >= 0:
                raise RuntimeError("Already a block at i=%d, j=%d, k=%d" % (i, j, k))
            self.blk_ids[i, j, k] = b.indx
            nic, njc, nkc = b.grid.niv - 1, b.grid.njv - 1, b.grid.nkv - 1
            if self.nics[i] == 0: self.nics[i] = nic
            if self.njcs[j] == 0: self.njcs[j] = njc
            if self.nkcs[k] == 0: self.nkcs[k] = nkc
            if nic != self.nics[i]:
                raise RuntimeError("FluidBlock at i=%d j=%d k=%d has inconsistent nic=%d" % (i, j, k, nic))
            if njc != self.njcs[j]:
                raise RuntimeError("FluidBlock at i=%d j=%d k=%d has inconsistent njc=%d" % (i, j, k, njc))
            if nkc != self.nkcs[k]:
                raise RuntimeError("FluidBlock at i=%d j=%d k=%d has inconsistent nkc=%d" % (i, j, k, nkc))


def add_cfl_value(t_change, cfl_value):
    """Append a (t, cfl) pair; times must increase (chkn_prep.py:254-273)."""
    if len(config.cfl_list) > 0:
        if t_change > config.cfl_list[-1][0]:
            config.cfl_list.append((t_change, cfl_value))
        else:
            print("Warning: did not add t_change,cfl_value tuple (", t_change, cfl_value, ").")
    else:
        config.cfl_list.append((t_change, cfl_value))
    return len(config.cfl_list)


def add_dt_plot(t_change, dt_plot, dt_his=1.0e-6):
    """Append a (t, dt_plot, dt_his) triple; times must increase (chkn_prep.py:276-300)."""
    if len(config.dt_plot_list) > 0:
        if t_change > config.dt_plot_list[-1][0]:
            config.dt_plot_list.append((t_change, dt_plot, dt_his))
        else:
            print("Warning: did not add dt_plot tuple (", t_change, dt_plot, dt_his, ").")
    else:
        config.dt_plot_list.append((t_change, dt_plot, dt_his))
    return len(config.dt_plot_list)


class GasState:
    """p, T, YB with rho, e, a from the gas model (chkn_prep.py:311-342)."""

    __slots__ = "indx", "p", "T", "rho", "e", "YB", "a"

    def __init__(self, p, T, YB=0.0):
        self.p = p
        self.T = T
        self.YB = YB
        self.update_from_pT(config.gas_model)

    def update_from_pT(self, gmodel):
        self.e = gmodel.Cv * self.T + self.YB * gmodel.q
        self.rho = self.p / (gmodel.R * self.T)
        self.a = np.sqrt(gmodel.gamma * gmodel.R * self.T)

    def __repr__(self):
        return "GasState(p={}, T={}, YB={})".format(self.p, self.T, self.YB)

    def to_json(self):
        return '{"p": %g, "T": %g, "rho": %g, "e": %g, "YB": %g, "a": %g}' % (
            self.p, self.T, self.rho, self.e, self.YB, self.a)


class FlowState:
    """Gas state plus velocity (chkn_prep.py:345-374)."""

    __slots__ = "indx", "label", "gas", "vel"

    def __init__(self, gas=None, p=100.0e3, T=300.0, YB=0.0, vel=None, velx=0.0, vely=0.0,
                 velz=0.0, label=""):
        self.gas = copy(gas) if isinstance(gas, GasState) else GasState(p=p, T=T, YB=YB)
        self.vel = copy(vel) if isinstance(vel, Vector3) else Vector3(velx, vely, velz)
        self.label = label

    def __repr__(self):
        return "FlowState(gas={}, vel={})".format(self.gas, self.vel)

    def to_json(self):
        return '{"gas": %s, "vel": [%g, %g, %g]}' % (self.gas.to_json(), self.vel.x, self.vel.y,
                                                    self.vel.z)


class BoundaryCondition:
    __slots__ = ["tag"]

    def to_json(self):
        return '{"tag": "%s"}' % self.tag


class ExchangeBC(BoundaryCondition):
    def __init__(self):
        self.tag = "exchange"

    def __repr__(self):
        return "ExchangeBC()"


class WallWithSlipBC(BoundaryCondition):
    def __init__(self):
        self.tag = "wall_with_slip"

    def __repr__(self):
        return "WallWithSlipBC()"


class WallNoSlipAdiabaticBC(BoundaryCondition):
    def __init__(self):
        self.tag = "wall_no_slip_adiabatic"

    def __repr__(self):
        return "WallNoSlipAdiabaticBC()"


class WallNoSlipFixedTBC(BoundaryCondition):
    """Fixed-temperature no-slip wall; TWall a number or a function of (x, y, z)."""

    __slots__ = ["tag", "form", "value", "fun", "values"]

    def __init__(self, TWall):
        self.tag = "wall_no_slip_fixed_T"
        self.fun = None
        self.value = 0.0
        if callable(TWall):
            self.form = "fun"
            self.fun = TWall
        elif isinstance(TWall, (int, float)):
            self.form = "value"
            self.value = float(TWall)
        else:
            raise RuntimeError("TWall should be a single number or a function of (x,y,z)")

    def __repr__(self):
        return "WallNoSlipFixedTBC(form=%s)" % self.form

    def to_json(self):
        return '{"tag": "%s", "form": "%s", "value": %g}' % (self.tag, self.form, self.value)

    def generate_values(self, grid, face):
        """TWall at the midpoint of every face of the block side, in the order the solver
        reads them (chkn_prep.py:479-545, block.cu:566-611): i-sides k-major, j-sides
        k-major, k-sides j-major.  Midpoints are 0.25*(p0+p1+p2+p3) of the four face
        vertices, with the vertex order of the reference's j- and k-sides; its i-side slices
        are malformed (they cannot broadcast), so the i-sides take the same pattern."""
        v = grid.vertices
        nic, njc, nkc = grid.niv - 1, grid.njv - 1, grid.nkv - 1
        def mid(p0, p1, p2, p3):
            return tuple(0.25 * (p0[q] + p1[q] + p2[q] + p3[q]) for q in range(3))
        xyz = (v.x, v.y, v.z)
        values = []
        if face in ("iminus", "iplus"):
            i = 0 if face == "iminus" else nic
            m = mid([a[i, :-1, :-1] for a in xyz], [a[i, 1:, :-1] for a in xyz],
                    [a[i, 1:, 1:] for a in xyz], [a[i, :-1, 1:] for a in xyz])
            for k in range(nkc):
                for j in range(njc):
                    values.append(self.fun(m[0][j, k], m[1][j, k], m[2][j, k]))
        elif face in ("jminus", "jplus"):
            j = 0 if face == "jminus" else njc
            m = mid([a[1:, j, :-1] for a in xyz], [a[:-1, j, :-1] for a in xyz],
                    [a[:-1, j, 1:] for a in xyz], [a[1:, j, 1:] for a in xyz])
            for k in range(nkc):
                for i in range(nic):
                    values.append(self.fun(m[0][i, k], m[1][i, k], m[2][i, k]))
        elif face in ("kminus", "kplus"):
            k = 0 if face == "kminus" else nkc
            m = mid([a[:-1, :-1, k] for a in xyz], [a[1:, :-1, k] for a in xyz],
                    [a[:-1, 1:, k] for a in xyz], [a[1:, 1:, k] for a in xyz])
            for j in range(njc):
                for i in range(nic):
                    values.append(self.fun(m[0][i, j], m[1][i, j], m[2][i, j]))
        else:
            raise RuntimeError("Unknown face: " + str(face))
        return [float(val) for val in values]


def write_twall_file(fileName, values):
    """TWall-*.gz: one value per line, '%g' as chkn_prep.py:469-472 writes them."""
    with gzip.open(fileName, "wt") as f:
        for val in values:
            f.write("%g\n" % val)


class InflowBC(BoundaryCondition):
    __slots__ = ["tag", "fs", "fsi"]

    def __init__(self, fs):
        self.tag = "inflow"
        if type(fs) is FlowState:
            self.fs = fs
        else:
            raise RuntimeError("Inflow boundary condition expects a FlowState object.")

    def __repr__(self):
        return "InflowBC(fs={})".format(self.fs)

    def to_json(self):
        return '{"tag": "%s", "flow_state_index": %d}' % (self.tag, self.fsi)


class InflowFunctionBC(BoundaryCondition):
    __slots__ = ["tag", "fun_name"]

    def __init__(self, fun_name):
        self.tag = "inflow_function"
        self.fun_name = fun_name

    def __repr__(self):
        return "InflowFunctionBC(fun_name={})".format(self.fun_name)

    def to_json(self):
        return '{"tag": "%s", "fun_name": "%s"}' % (self.tag, self.fun_name)


class OutflowBC(BoundaryCondition):
    def __init__(self):
        self.tag = "outflow"

    def __repr__(self):
        return "OutflowBC()"


def _resolve_bcs(bcs):
    """Default walls, then the user's entries.  An integer key Face.iplus also sets jplus,
    as chkn_prep.py:671-677 and :815-821 do."""
    out = {f: WallWithSlipBC() for f in FaceList}
    for key in bcs.keys():
        if key in [Face.iminus, "iminus"]: out["iminus"] = bcs[key]
        if key in [Face.iplus, "iplus"]: out["iplus"] = bcs[key]
        if key in [Face.jminus, "jminus"]: out["jminus"] = bcs[key]
        if key in [Face.iplus, "jplus"]: out["jplus"] = bcs[key]
        if key in [Face.kminus, "kminus"]: out["kminus"] = bcs[key]
        if key in [Face.kplus, "kplus"]: out["kplus"] = bcs[key]
    return out


class FluidBlock:
    """One structured block: grid, initial flow and boundary conditions."""

    __slots__ = ("indx", "active", "i", "j", "k", "grid", "initialState", "nic", "njc", "nkc",
                 "bcs", "label", "cellc", "cellv", "flow", "twall_values")

    def __init__(self, i=0, j=0, k=0, grid=None, initialState=None, bcs={}, active=True, label=""):
        if not isinstance(grid, StructuredGrid):
            raise RuntimeError("Need to supply a StructuredGrid object.")
        self.grid = grid
        self.cellc = cell_centres(grid)
        self.cellv = hex_volumes(grid)
        self.nic, self.njc, self.nkc = grid.niv - 1, grid.njv - 1, grid.nkv - 1
        n = self.nic * self.njc * self.nkc
        if isinstance(initialState, FlowState):
            s = initialState
            self.flow = {"p": np.full(n, s.gas.p), "T": np.full(n, s.gas.T),
                         "rho": np.full(n, s.gas.rho), "e": np.full(n, s.gas.e),
                         "YB": np.full(n, s.gas.YB), "a": np.full(n, s.gas.a),
                         "velx": np.full(n, s.vel.x), "vely": np.full(n, s.vel.y),
                         "velz": np.full(n, s.vel.z)}
        elif isinstance(initialState, dict):
            self.flow = {}
            for key in ("p", "T", "rho", "e", "YB", "a", "velx", "vely", "velz"):
                a = np.asarray(initialState[key], dtype=float)
                if a.ndim == 3:
                    a = a.transpose().ravel()
                self.flow[key] = np.ascontiguousarray(np.broadcast_to(a, (n,)), dtype=float)
        elif callable(initialState):
            self.flow = {key: np.zeros(n) for key in ("p", "T", "rho", "e", "YB", "a", "velx",
                                                        "vely", "velz")}
            idx = 0
            for kk in range(self.nkc):
                for jj in range(self.njc):
                    for ii in range(self.nic):
                        st = initialState(self.cellc.x[ii, jj, kk], self.cellc.y[ii, jj, kk],
                                          self.cellc.z[ii, jj, kk])
                        for key, val in (("p", st.gas.p), ("T", st.gas.T), ("rho", st.gas.rho),
                                         ("e", st.gas.e), ("YB", st.gas.YB), ("a", st.gas.a),
                                         ("velx", st.vel.x), ("vely", st.vel.y), ("velz", st.vel.z)):
                            self.flow[key][idx] = val
                        idx += 1
        else:
            raise RuntimeError("Need to supply a FlowState object or a function that produces"
                               " a FlowState object as a function of (x,y,z) location.")
        self.bcs = _resolve_bcs(bcs)
        # Per-face TWall of function-form walls, kept per block (chkn_prep.py:688-692 keeps
        # them on the boundary-condition object, which blocks may share).
        self.twall_values = {face: bc.generate_values(grid, face) for face, bc in self.bcs.items()
                             if type(bc) is WallNoSlipFixedTBC and bc.form == "fun"}
        for bc in bcs.values():
            if type(bc) is InflowBC:
                bc.fsi = len(flowStatesList)
                flowStatesList.append(bc.fs)
        self.i, self.j, self.k = i, j, k
        self.active = active
        self.label = label
        self.indx = len(fluidBlocksList)
        fluidBlocksList.append(self)

    def to_json(self):
        s = '{"i": %d, "j": %d, "k": %d,' % (self.i, self.j, self.k)
        s += ' "bcs": {' + ",".join('"%s": %s' % (f, self.bcs[f].to_json()) for f in FaceList) + "},"
        s += ' "active": %s,' % json.dumps(self.active)
        s += ' "label": "%s"}' % self.label
        return s

    def iovars(self):
        """(13, ncells) array in flow-file order."""
        n = self.nic * self.njc * self.nkc
        out = np.empty((13, n))
        out[0] = self.cellc.x.transpose().ravel()
        out[1] = self.cellc.y.transpose().ravel()
        out[2] = self.cellc.z.transpose().ravel()
        out[3] = self.cellv.transpose().ravel()
        for q, key in enumerate(("p", "T", "rho", "e", "YB", "a", "velx", "vely", "velz")):
            out[4 + q] = self.flow[key]
        return out

    def write_flow_data_to_file(self, fileName, binaryData):
        write_flow_file(fileName, self.iovars(), binaryData)


def write_flow_file(fileName, iovars, binaryData):
    """Flow file: the 13 IO variables one after another (chkn_prep.py:749-780)."""
    iovars = np.ascontiguousarray(iovars, dtype=float)
    if binaryData:
        iovars.tofile(fileName)
    else:
        import gzip
        with gzip.open(fileName, "wt") as f:
            np.savetxt(f, iovars.ravel())


def makeFBArray(i0=0, j0=0, k0=0, ni=1, nj=1, nk=1, grid=None, initialState=None, bcs={},
                active=True, label=""):
    """Split one grid into ni x nj x nk blocks joined by exchange boundaries
    (chkn_prep.py:784-851)."""
    if not isinstance(grid, StructuredGrid):
        raise RuntimeError("Need to supply a StructuredGrid object to subdivide.")
    if not (isinstance(initialState, (FlowState, dict)) or callable(initialState)):
        raise RuntimeError("Need to supply a FlowState object or a function for initialState.")
    _bcs = _resolve_bcs(bcs)
    blks = []
    nic = (grid.niv - 1) // ni
    njc = (grid.njv - 1) // nj
    nkc = (grid.nkv - 1) // nk
    for k in range(nk):
        k0v = k * nkc
        nkv = nkc + 1 if k < nk - 1 else grid.nkv - k0v
        for j in range(nj):
            j0v = j * njc
            njv = njc + 1 if j < nj - 1 else grid.njv - j0v
            for i in range(ni):
                i0v = i * nic
                niv = nic + 1 if i < ni - 1 else grid.niv - i0v
                newBCs = {
                    "iminus": _bcs["iminus"] if i == 0 else ExchangeBC(),
                    "iplus": _bcs["iplus"] if i == ni - 1 else ExchangeBC(),
                    "jminus": _bcs["jminus"] if j == 0 else ExchangeBC(),
                    "jplus": _bcs["jplus"] if j == nj - 1 else ExchangeBC(),
                    "kminus": _bcs["kminus"] if k == 0 else ExchangeBC(),
                    "kplus": _bcs["kplus"] if k == nk - 1 else ExchangeBC(),
                }
                sub = grid.subgrid(i0=i0v, j0=j0v, k0=k0v, niv=niv, njv=njv, nkv=nkv)
                init = initialState
                if isinstance(initialState, dict):
                    init = {key: np.asarray(v)[i0v:i0v + niv - 1, j0v:j0v + njv - 1, k0v:k0v + nkv - 1]
                            if np.ndim(v) == 3 else v for key, v in initialState.items()}
                blks.append(FluidBlock(i=i0 + i, j=j0 + j, k=k0 + k, grid=sub, initialState=init,
                                       bcs=newBCs, active=active,
                                       label=label + "-%d-%d-%d" % (i, j, k)))
    return blks


def config_json_text():
    """The config.json text (chkn_prep.py:868-895)."""
    import io
    config.check_array_of_fluid_blocks()
    fp = io.StringIO()
    fp.write("{\n")
    config.write(fp)
    n_fs = len(flowStatesList)
    fp.write('  "n_flow_states": %d,\n' % n_fs)
    fp.write('  "flow_states": [\n')
    for i, fs in enumerate(flowStatesList):
        fp.write("    %s" % fs.to_json())
        fp.write(",\n" if i < n_fs - 1 else "\n")
    fp.write("  ],\n")
    n_fb = len(fluidBlocksList)
    fp.write('  "n_fluid_blocks": %d,\n' % n_fb)
    fp.write('  "nib": %d,\n' % config.nib)
    fp.write('  "njb": %d,\n' % config.njb)
    fp.write('  "nkb": %d,\n' % config.nkb)
    fp.write('  "blk_ids": %s,\n' % json.dumps(config.blk_ids.tolist()))
    fp.write('  "nics": %s,\n' % json.dumps(config.nics.tolist()))
    fp.write('  "njcs": %s,\n' % json.dumps(config.njcs.tolist()))
    fp.write('  "nkcs": %s,\n' % json.dumps(config.nkcs.tolist()))
    fp.write('  "fluid_blocks": [\n')
    for i, fb in enumerate(fluidBlocksList):
        fp.write("    %s" % fb.to_json())
        fp.write(",\n" if i < n_fb - 1 else "\n")
    fp.write("  ]\n")
    fp.write("}\n")
    return fp.getvalue()


def write_initial_files(binaryData, job_dir=None):
    """Write config.json, grids, flow/t0000 and times.data into the job directory."""
    job = job_dir or config.job_name
    os.makedirs(job, exist_ok=True)
    with open(os.path.join(job, "config.json"), "w") as fp:
        fp.write(config_json_text())
    os.makedirs(os.path.join(job, "grid"), exist_ok=True)
    for fb in fluidBlocksList:
        name = os.path.join(job, "grid", "grid-%04d-%04d-%04d" % (fb.i, fb.j, fb.k))
        if binaryData:
            fb.grid.write_to_binary_file(name + ".bin")
        else:
            fb.grid.write_to_gzip_file(name + ".gz")
    flow_dir = os.path.join(job, "flow", "t0000")
    os.makedirs(flow_dir, exist_ok=True)
    for fb in fluidBlocksList:
        name = os.path.join(flow_dir, "flow-%04d-%04d-%04d" % (fb.i, fb.j, fb.k))
        fb.write_flow_data_to_file(name + (".bin" if binaryData else ".gz"), binaryData)
        for face, values in fb.twall_values.items():  # chkn_prep.py:919-925
            write_twall_file(os.path.join(job, "TWall-%04d-%04d-%04d-%s.gz" % (fb.i, fb.j, fb.k, face)),
                             values)
    with open(os.path.join(job, "times.data"), "w") as fp:
        fp.write("# tindx t\n")
        fp.write("0 0.0\n")


def reset():
    """Start a fresh job description (module-level state, as in chkn_prep.py)."""
    global config, fluidBlocksList, flowStatesList
    GlobalConfig.count = 0
    config = GlobalConfig()
    fluidBlocksList = []
    flowStatesList = []


config = None
fluidBlocksList = []
flowStatesList = []
reset()


def script_namespace():
    """Names visible to a job script."""
    import sys
    mod = sys.modules[__name__]
    ns = {k: getattr(mod, k) for k in dir(mod) if not k.startswith("_")}
    ns["math"] = math
    ns["np"] = np
    return ns


def run_script(path):
    """Execute a job script with this module's names in scope (chkn_prep.py:958)."""
    import sys
    mod = sys.modules[__name__]
    with open(path, "rb") as f:
        code = compile(f.read(), path, "exec")
    exec(code, mod.__dict__)


def main(argv=None):
    """chkn-prep --job=<name> [--binary]"""
    import getopt
    import sys
    import time
    argv = sys.argv[1:] if argv is None else argv
    opts, _ = getopt.getopt(argv, "hf:b", ["help", "job=", "binary"])
    uo = dict(opts)
    if not opts or "--help" in uo or "-h" in uo:
        print("Usage: chkn-prep [--help | -h] [--job=<jobName> | -f <jobName>] [--binary | -b]")
        return 0
    job = uo.get("--job", uo.get("-f", ""))
    if not job:
        raise Exception("Job name is not specified.")
    config.job_name, _ = os.path.splitext(job)
    run_script(config.job_name + ".py")
    config.check_array_of_fluid_blocks()
    write_initial_files(("--binary" in uo) or ("-b" in uo))
    print("Done in {:.3f} seconds.".format(time.process_time()))
    return 0
