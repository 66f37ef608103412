"""Job preparation for the chicken step (the chkn-prep front-end).

A job script is ordinary Python run with this module's names in scope (``config``,
``FlowState``, ``FluidBlock``, ``makeFBArray``, the boundary-condition classes,
``add_cfl_value``, ``add_dt_plot``, ``Face`` and the geometry helpers ``Vector3``,
``TFIVolume``, ``StructuredGrid``); ``write_initial_files`` then writes the job directory
the solver reads: config.json, grid/, flow/t0000/, times.data and any TWall-*.gz files.

What is fixed by the solver's file contract and the job-script API of the reference
(src/chicken/chkn_prep.py) is kept: the user-facing names, keyword arguments and defaults,
the config.json text byte for byte (key order and number formats read by config.cu:347-525),
the gas-state formulas whose results land in the flow files, and one quirk of the reference's
boundary-condition lookup (an integer key ``Face.iplus`` also sets ``jplus``,
chkn_prep.py:671-677).  Everything else is this module's own design:

* config.json is produced from a table of keys and formatters (``_GLOBAL_KEYS``);
* the block array is checked by collecting slab sizes per axis (``_BlockArray``);
* boundary conditions declare their JSON fields in a class table (``_fields``);
* initial states (uniform, per-cell arrays, or a function of position) are all turned into
  per-cell arrays by one helper, and ``makeFBArray`` splits a grid through per-axis ranges.

Cell centres and volumes are written to the flow file in its own cell order (i fastest); the
solver recomputes both from the grid (block.cu:640-647).  ``reset()`` starts a new job.
"""
import collections
import gzip
import io
import itertools
import json
import math
import os
import sys
from copy import copy

import numpy as np

from .geom import StructuredGrid, TFIVolume, Vector3, cell_centres, hex_volumes

IOVAR_NAMES = ["posx", "posy", "posz", "vol", "p", "T", "rho", "e", "YB", "a",
               "velx", "vely", "velz"]
# The nine flow variables of a cell after its centre and volume (cell.cu:66-84).
FLOW_KEYS = ("p", "T", "rho", "e", "YB", "a", "velx", "vely", "velz")

FaceNames = "iminus iplus jminus jplus kminus kplus"
FaceList = FaceNames.split()
Face = collections.namedtuple("Face", FaceList)(*range(6))


# --------------------------------------------------------------------------- gas and flow
class GasModel:
    """Gas constants written to config.json; the solver checks them against the model it
    was built with (config.cu:369-386).  Defaults: ideal air."""

    __slots__ = "gamma", "R", "Cv", "q", "alpha", "Ti"

    def __init__(self, gamma=1.4, R=287.1, q=0.0, alpha=0.0, Ti=0.0):
        self.gamma, self.R, self.q, self.alpha, self.Ti = gamma, R, q, alpha, Ti
        self.Cv = R / (gamma - 1.0)

    def to_json(self):
        return "{%s}" % ", ".join('"%s": %g' % (k, getattr(self, k))
                                  for k in ("gamma", "R", "Cv", "q", "alpha", "Ti"))


class GasState:
    """(p, T, YB) with rho, e and a from the current gas model."""

    __slots__ = "indx", "p", "T", "rho", "e", "YB", "a"

    def __init__(self, p, T, YB=0.0):
        self.p, self.T, self.YB = p, T, YB
        self.update_from_pT(config.gas_model)

    def update_from_pT(self, gmodel):
        # Ideal-gas relations; the same operations as the solver's update_from_pT
        # (gas.cu:119-126), so inflow states agree with what it derives from p and T.
        self.e = gmodel.Cv * self.T + self.YB * gmodel.q
        self.rho = self.p / (gmodel.R * self.T)
        self.a = np.sqrt(gmodel.gamma * gmodel.R * self.T)

    def __repr__(self):
        return "GasState(p=%r, T=%r, YB=%r)" % (self.p, self.T, self.YB)

    def to_json(self):
        return "{%s}" % ", ".join('"%s": %g' % (k, getattr(self, k))
                                  for k in ("p", "T", "rho", "e", "YB", "a"))


class FlowState:
    """A gas state and a velocity; either given whole or built from p, T, YB and components."""

    __slots__ = "indx", "label", "gas", "vel"

    def __init__(self, gas=None, p=100.0e3, T=300.0, YB=0.0, vel=None, velx=0.0, vely=0.0,
                 velz=0.0, label=""):
        self.gas = copy(gas) if isinstance(gas, GasState) else GasState(p=p, T=T, YB=YB)
        self.vel = copy(vel) if isinstance(vel, Vector3) else Vector3(velx, vely, velz)
        self.label = label

    def __repr__(self):
        return "FlowState(gas=%r, vel=%r)" % (self.gas, self.vel)

    def to_json(self):
        v = self.vel
        return '{"gas": %s, "vel": [%g, %g, %g]}' % (self.gas.to_json(), v.x, v.y, v.z)

    def cell_values(self):
        """The nine flow variables of a cell in this state, in FLOW_KEYS order."""
        g, v = self.gas, self.vel
        return (g.p, g.T, g.rho, g.e, g.YB, g.a, v.x, v.y, v.z)


# --------------------------------------------------------------------------- schedules
def _append_schedule(entries, entry, what):
    """Schedules are piecewise constant in time: an entry is kept only if its start time
    follows the last one's."""
    if entries and not entry[0] > entries[-1][0]:
        print("Warning: %s %r ignored: its start time does not follow %g." % (what, entry, entries[-1][0]))
    else:
        entries.append(entry)
    return len(entries)


def add_cfl_value(t_change, cfl_value):
    """From time t_change on, use this CFL number."""
    return _append_schedule(config.cfl_list, (t_change, cfl_value), "cfl entry")


def add_dt_plot(t_change, dt_plot, dt_his=1.0e-6):
    """From time t_change on, write snapshots every dt_plot (dt_his: history interval)."""
    return _append_schedule(config.dt_plot_list, (t_change, dt_plot, dt_his), "dt_plot entry")


# --------------------------------------------------------------------------- global config
def _js(value):
    return json.dumps(value)


def _quoted(value):
    return '"%s"' % value


def _cfl_times(c):
    return _js([e[0] for e in c.cfl_list])


def _cfl_values(c):
    vals = [e[1] for e in c.cfl_list]
    if min(vals) <= 0.0:
        raise ValueError("CFL values must be positive, got %r" % (vals,))
    return _js(vals)


# config.json keys of the global part, in the order the file holds them, with a function
# giving each value's text from the config.
_GLOBAL_KEYS = (
    ("title", lambda c: _quoted(c.title)),
    ("gas_model", lambda c: c.gas_model.to_json()),
    ("max_time", lambda c: "%e" % c.max_time),
    ("max_step", lambda c: "%d" % c.max_step),
    ("dt_init", lambda c: "%e" % c.dt_init),
    ("n_cfl", lambda c: "%d" % len(c.cfl_list)),
    ("cfl_times", _cfl_times),
    ("cfl_values", _cfl_values),
    ("cfl_count", lambda c: "%d" % c.cfl_count),
    ("print_count", lambda c: "%d" % c.print_count),
    ("x_order", lambda c: "%d" % c.x_order),
    ("t_order", lambda c: "%d" % c.t_order),
    ("time_scheme", lambda c: _quoted(c.time_scheme)),
    ("flux_calc", lambda c: _quoted(c.flux_calc)),
    ("reacting", lambda c: _js(c.reacting)),
    ("viscous", lambda c: _js(c.viscous)),
    ("source_terms", lambda c: _quoted(c.source_terms)),
    ("threads_per_gpu_block", lambda c: "%d" % c.threads_per_gpu_block),
    ("n_dt_plot", lambda c: "%d" % len(c.dt_plot_list)),
    ("t_change", lambda c: _js([e[0] for e in c.dt_plot_list])),
    ("dt_plot", lambda c: _js([e[1] for e in c.dt_plot_list])),
    ("dt_hist", lambda c: _js([e[2] for e in c.dt_plot_list])),
    ("iovar_names", lambda c: _js(c.iovar_names)),
)

# Keys the reference's config.json does not have, written only when they differ from their
# default: "time_scheme" selects the two-stage predictor-corrector ("predictor_corrector")
# instead of the reference's TVD-RK3 ("rk3"); chicken itself reads t_order and ignores it.
_OPTIONAL_KEYS = {"time_scheme": "rk3"}
TIME_SCHEMES = ("rk3", "predictor_corrector")

# Attribute defaults of a new config (callables give a fresh mutable value each time).
_CONFIG_DEFAULTS = (
    ("job_name", ""), ("title", "Chicken Simulation."), ("gas_model", GasModel),
    ("nib", None), ("njb", None), ("nkb", None), ("blk_ids", None),
    ("nics", None), ("njcs", None), ("nkcs", None),
    ("dt_init", 1.0e-6), ("cfl_list", list), ("cfl_count", 10), ("print_count", 20),
    ("dt_plot_list", list), ("max_time", 1.0e-3), ("max_step", 100),
    ("x_order", 2), ("t_order", 2), ("time_scheme", "rk3"), ("flux_calc", "ausmdv"), ("reacting", False),
    ("viscous", False), ("source_terms", "none"), ("threads_per_gpu_block", 128),
    ("iovar_names", lambda: list(IOVAR_NAMES)),
)


class GlobalConfig:
    """Simulation control parameters.  There is one per job (the module-level ``config``);
    a job script sets its attributes."""

    count = 0
    __slots__ = tuple(name for name, _ in _CONFIG_DEFAULTS)

    def __init__(self):
        if GlobalConfig.count:
            raise RuntimeError("a job has one GlobalConfig: use the module's `config` (reset() "
                               "starts a new job)")
        for name, default in _CONFIG_DEFAULTS:
            setattr(self, name, default() if callable(default) else default)
        GlobalConfig.count += 1

    def write(self, fp):
        """The global part of config.json.  Empty schedules get the defaults: CFL 0.5 from
        t = 0, and one snapshot interval of max_time."""
        if not self.cfl_list:
            add_cfl_value(0.0, 0.5)
        if not self.dt_plot_list:
            add_dt_plot(0.0, self.max_time, self.max_time)
        if self.time_scheme not in TIME_SCHEMES:
            raise ValueError("time_scheme must be one of %s, got %r" % (TIME_SCHEMES, self.time_scheme))
        for key, text in _GLOBAL_KEYS:
            if key in _OPTIONAL_KEYS and getattr(self, key) == _OPTIONAL_KEYS[key]:
                continue
            fp.write('  "%s": %s,\n' % (key, text(self)))

    def check_array_of_fluid_blocks(self):
        """Place the blocks in their (i, j, k) array and record the slab sizes."""
        arr = _BlockArray(fluidBlocksList)
        self.nib, self.njb, self.nkb = arr.shape
        self.blk_ids = arr.ids
        self.nics, self.njcs, self.nkcs = arr.sizes


class _BlockArray:
    """Blocks by array position.  Every block of array column i has the same nic, of row j
    the same njc and of layer k the same nkc (the solver's exchange copies whole faces)."""

    def __init__(self, blocks):
        if not blocks:
            raise RuntimeError("no FluidBlock has been defined")
        self.shape = tuple(1 + max(getattr(b, a) for b in blocks) for a in "ijk")
        self.ids = np.full(self.shape, -1, dtype=int)
        self.sizes = [np.zeros(n, dtype=int) for n in self.shape]
        for b in blocks:
            pos = (b.i, b.j, b.k)
            if self.ids[pos] >= 0:
                raise RuntimeError("blocks %d and %d both sit at array position %s"
                                   % (self.ids[pos], b.indx, pos))
            self.ids[pos] = b.indx
            for axis, (slab, cells) in enumerate(zip(pos, (b.nic, b.njc, b.nkc))):
                have = self.sizes[axis][slab]
                if have and have != cells:
                    raise RuntimeError("block %d at %s has %d cells along %s, its slab has %d"
                                       % (b.indx, pos, cells, "ijk"[axis], have))
                self.sizes[axis][slab] = cells


# --------------------------------------------------------------------------- boundaries
class BoundaryCondition:
    """A block side's boundary condition.  ``tag`` is the name config.json uses
    (face.cu:41-59); ``_fields`` lists (attribute, json key, format) of its extra entries."""

    __slots__ = ()
    tag = ""
    _fields = ()

    def to_json(self):
        parts = ['"tag": "%s"' % self.tag]
        parts += ['"%s": %s' % (key, fmt % getattr(self, attr)) for attr, key, fmt in self._fields]
        return "{%s}" % ", ".join(parts)

    def __repr__(self):
        args = ", ".join("%s=%r" % (attr, getattr(self, attr, None)) for attr, _, _ in self._fields)
        return "%s(%s)" % (type(self).__name__, args)


class ExchangeBC(BoundaryCondition):
    """Copy from the neighbouring block in the array (periodic at the array's edges)."""
    tag = "exchange"


class WallWithSlipBC(BoundaryCondition):
    """Inviscid wall: the normal velocity is reflected."""
    tag = "wall_with_slip"


class WallNoSlipAdiabaticBC(BoundaryCondition):
    """Viscous wall without heat flux."""
    tag = "wall_no_slip_adiabatic"


class WallNoSlipFixedTBC(BoundaryCondition):
    """Viscous wall at temperature TWall: a number, or a function of (x, y, z) evaluated at
    each face centre and written to TWall-*.gz files."""

    __slots__ = ("form", "value", "fun")
    tag = "wall_no_slip_fixed_T"
    _fields = (("form", "form", '"%s"'), ("value", "value", "%g"))

    def __init__(self, TWall):
        if callable(TWall):
            self.form, self.value, self.fun = "fun", 0.0, TWall
        elif isinstance(TWall, (int, float)):
            self.form, self.value, self.fun = "value", float(TWall), None
        else:
            raise TypeError("TWall must be a number or a function of (x, y, z), got %r" % (TWall,))

    def generate_values(self, grid, face):
        """TWall at every face of a block side, in the order the solver reads them
        (block.cu:566-611): i-sides over (j, k), j-sides over (i, k), k-sides over (i, j),
        the first index fastest.  A face's point is the mean of its four vertices."""
        xyz = [np.asarray(a) for a in (grid.vertices.x, grid.vertices.y, grid.vertices.z)]
        p0, p1, p2, p3 = _side_quads(xyz, face)
        pts = [0.25 * (p0[q] + p1[q] + p2[q] + p3[q]) for q in range(3)]
        return [float(self.fun(x, y, z)) for x, y, z in zip(*(p.T.ravel() for p in pts))]


def _side_quads(xyz, face):
    """Vertex arrays (p0, p1, p2, p3), each three coordinate arrays over the side's faces
    indexed by the two in-plane indices; the order of the four vertices is that of the
    midpoint sums of the TWall files."""
    side = FaceList.index(face) if isinstance(face, str) else face
    d, plus = divmod(side, 2)
    n = xyz[0].shape[d] - 1 if plus else 0
    cut = [slice(None)] * 3
    cut[d] = n
    plane = [a[tuple(cut)] for a in xyz]  # the side's vertex plane, two remaining axes
    lo, hi = slice(None, -1), slice(1, None)
    if d == 0:      # (j, k): (j,k) (j+1,k) (j+1,k+1) (j,k+1)
        corners = ((lo, lo), (hi, lo), (hi, hi), (lo, hi))
    elif d == 1:    # (i, k): (i+1,k) (i,k) (i,k+1) (i+1,k+1)
        corners = ((hi, lo), (lo, lo), (lo, hi), (hi, hi))
    else:           # (i, j): (i,j) (i+1,j) (i,j+1) (i+1,j+1)
        corners = ((lo, lo), (hi, lo), (lo, hi), (hi, hi))
    return [[c[a, b] for c in plane] for a, b in corners]


def write_twall_file(fileName, values):
    """A TWall-*.gz file: one value per line, 6 significant digits."""
    with gzip.open(fileName, "wt") as f:
        f.writelines("%g\n" % v for v in values)


class InflowBC(BoundaryCondition):
    """Ghost cells hold the flow state fs (listed in config.json's flow_states)."""

    __slots__ = ("fs", "fsi")
    tag = "inflow"
    _fields = (("fsi", "flow_state_index", "%d"),)

    def __init__(self, fs):
        if type(fs) is not FlowState:
            raise TypeError("InflowBC needs a FlowState, got %r" % (fs,))
        self.fs = fs


class InflowFunctionBC(BoundaryCondition):
    """Ghost cells from a named function of position (face.cu:51-59)."""

    __slots__ = ("fun_name",)
    tag = "inflow_function"
    _fields = (("fun_name", "fun_name", '"%s"'),)

    def __init__(self, fun_name):
        self.fun_name = fun_name


class OutflowBC(BoundaryCondition):
    """Ghost cells copy the adjacent interior cell."""
    tag = "outflow"


def _resolve_bcs(bcs):
    """The six sides' conditions: slip walls unless given, by name or by Face index.  The
    integer key Face.iplus (1) also sets jplus, as the reference's lookup does, so job
    scripts give the same files under both front-ends."""
    out = {name: WallWithSlipBC() for name in FaceList}
    aliases = {name: {name, Face[q]} for q, name in enumerate(FaceList)}
    aliases["jplus"] = {"jplus", Face.iplus}
    for key, bc in bcs.items():
        for name in FaceList:
            if key in aliases[name]:
                out[name] = bc
    return out


# --------------------------------------------------------------------------- blocks
def _initial_flow(initialState, centres, n):
    """Per-cell arrays of the nine flow variables (cell order, i fastest) from a FlowState,
    a dict of arrays (per cell, either flat in cell order or [i, j, k]) or a function of the
    cell centre that returns a FlowState."""
    if isinstance(initialState, FlowState):
        return {key: np.full(n, v) for key, v in zip(FLOW_KEYS, initialState.cell_values())}
    if isinstance(initialState, dict):
        flow = {}
        for key in FLOW_KEYS:
            a = np.asarray(initialState[key], dtype=float)
            a = a.transpose().ravel() if a.ndim == 3 else a
            flow[key] = np.ascontiguousarray(np.broadcast_to(a, (n,)), dtype=float)
        return flow
    if callable(initialState):
        pts = zip(*(c.transpose().ravel() for c in (centres.x, centres.y, centres.z)))
        rows = np.array([initialState(x, y, z).cell_values() for x, y, z in pts], dtype=float)
        return {key: np.ascontiguousarray(rows[:, q]) for q, key in enumerate(FLOW_KEYS)}
    raise TypeError("initialState must be a FlowState, a dict of per-cell arrays or a function "
                    "of (x, y, z) giving a FlowState")


class FluidBlock:
    """One structured block of the job: grid, initial flow and the six boundary conditions,
    at position (i, j, k) of the block array."""

    __slots__ = ("indx", "active", "i", "j", "k", "grid", "initialState", "nic", "njc", "nkc",
                 "bcs", "label", "cellc", "cellv", "flow", "twall_values")

    def __init__(self, i=0, j=0, k=0, grid=None, initialState=None, bcs={}, active=True, label=""):
        if not isinstance(grid, StructuredGrid):
            raise TypeError("FluidBlock needs a StructuredGrid, got %r" % (grid,))
        self.grid = grid
        self.nic, self.njc, self.nkc = grid.niv - 1, grid.njv - 1, grid.nkv - 1
        self.cellc = cell_centres(grid)
        self.cellv = hex_volumes(grid)
        self.flow = _initial_flow(initialState, self.cellc, self.nic * self.njc * self.nkc)
        self.bcs = _resolve_bcs(bcs)
        # TWall per face of function-form walls, kept with the block (a wall object may be
        # shared by several blocks).
        self.twall_values = {face: bc.generate_values(grid, face) for face, bc in self.bcs.items()
                             if isinstance(bc, WallNoSlipFixedTBC) and bc.form == "fun"}
        for bc in bcs.values():
            if isinstance(bc, InflowBC):
                bc.fsi = len(flowStatesList)
                flowStatesList.append(bc.fs)
        self.i, self.j, self.k = i, j, k
        self.active, self.label = active, label
        self.indx = len(fluidBlocksList)
        fluidBlocksList.append(self)

    def to_json(self):
        sides = ",".join('"%s": %s' % (f, self.bcs[f].to_json()) for f in FaceList)
        return ('{"i": %d, "j": %d, "k": %d, "bcs": {%s}, "active": %s, "label": "%s"}'
                % (self.i, self.j, self.k, sides, _js(self.active), self.label))

    def iovars(self):
        """The 13 IO variables, (13, ncells), in flow-file order."""
        geo = [self.cellc.x, self.cellc.y, self.cellc.z, self.cellv]
        return np.vstack([g.transpose().ravel() for g in geo] + [self.flow[k] for k in FLOW_KEYS])

    def write_flow_data_to_file(self, fileName, binaryData):
        write_flow_file(fileName, self.iovars(), binaryData)


def write_flow_file(fileName, iovars, binaryData):
    """A flow file: the 13 IO variables one after another, float64 (.bin) or one value per
    line (.gz)."""
    iovars = np.ascontiguousarray(iovars, dtype=float)
    if binaryData:
        iovars.tofile(fileName)
    else:
        with gzip.open(fileName, "wt") as f:
            np.savetxt(f, iovars.ravel())


def _split_axis(nv, parts):
    """Vertex ranges (start, count) of `parts` consecutive blocks along an axis of nv
    vertices: equal cell counts, the last block taking the remainder."""
    step = (nv - 1) // parts
    if step < 1:
        raise ValueError("cannot split %d cells into %d blocks" % (nv - 1, parts))
    starts = [q * step for q in range(parts)]
    return [(s, (step if q < parts - 1 else nv - 1 - s) + 1) for q, s in enumerate(starts)]


def makeFBArray(i0=0, j0=0, k0=0, ni=1, nj=1, nk=1, grid=None, initialState=None, bcs={},
                active=True, label=""):
    """Split one grid into ni x nj x nk blocks placed at array positions (i0+i, j0+j, k0+k);
    inner sides are exchange boundaries, outer sides take `bcs`.  Returns the blocks."""
    if not isinstance(grid, StructuredGrid):
        raise TypeError("makeFBArray needs a StructuredGrid to split, got %r" % (grid,))
    if not (isinstance(initialState, (FlowState, dict)) or callable(initialState)):
        raise TypeError("initialState must be a FlowState, a dict of per-cell arrays or a "
                        "function of (x, y, z) giving a FlowState")
    outer = _resolve_bcs(bcs)
    ranges = [_split_axis(nv, parts) for nv, parts in ((grid.niv, ni), (grid.njv, nj), (grid.nkv, nk))]
    blocks = []
    for k, j, i in itertools.product(range(nk), range(nj), range(ni)):
        pos, counts = (i, j, k), (ni, nj, nk)
        sides = {}
        for q, name in enumerate(FaceList):
            d, plus = divmod(q, 2)
            at_edge = pos[d] == (counts[d] - 1 if plus else 0)
            sides[name] = outer[name] if at_edge else ExchangeBC()
        (si, ni_v), (sj, nj_v), (sk, nk_v) = ranges[0][i], ranges[1][j], ranges[2][k]
        sub = grid.subgrid(i0=si, j0=sj, k0=sk, niv=ni_v, njv=nj_v, nkv=nk_v)
        init = initialState
        if isinstance(initialState, dict):
            cells = (slice(si, si + ni_v - 1), slice(sj, sj + nj_v - 1), slice(sk, sk + nk_v - 1))
            init = {key: (np.asarray(v)[cells] if np.ndim(v) == 3 else v) for key, v in initialState.items()}
        blocks.append(FluidBlock(i=i0 + i, j=j0 + j, k=k0 + k, grid=sub, initialState=init,
                                 bcs=sides, active=active, label="%s-%d-%d-%d" % (label, i, j, k)))
    return blocks


# --------------------------------------------------------------------------- job files
def _json_list(items):
    """Entries of a config.json list, one per line."""
    return ",\n".join("    %s" % it for it in items) + ("\n" if items else "")


def config_json_text():
    """The text of config.json for the job described so far."""
    config.check_array_of_fluid_blocks()
    fp = io.StringIO()
    fp.write("{\n")
    config.write(fp)
    fp.write('  "n_flow_states": %d,\n' % len(flowStatesList))
    fp.write('  "flow_states": [\n%s  ],\n' % _json_list([fs.to_json() for fs in flowStatesList]))
    fp.write('  "n_fluid_blocks": %d,\n' % len(fluidBlocksList))
    for key, val in (("nib", config.nib), ("njb", config.njb), ("nkb", config.nkb)):
        fp.write('  "%s": %d,\n' % (key, val))
    for key, arr in (("blk_ids", config.blk_ids), ("nics", config.nics), ("njcs", config.njcs),
                     ("nkcs", config.nkcs)):
        fp.write('  "%s": %s,\n' % (key, _js(arr.tolist())))
    fp.write('  "fluid_blocks": [\n%s  ]\n' % _json_list([fb.to_json() for fb in fluidBlocksList]))
    fp.write("}\n")
    return fp.getvalue()


def write_initial_files(binaryData, job_dir=None):
    """Write config.json, the grids, flow/t0000, TWall files and times.data into job_dir
    (default: the job name)."""
    job = job_dir or config.job_name
    for sub in ("grid", os.path.join("flow", "t0000")):
        os.makedirs(os.path.join(job, sub), exist_ok=True)
    with open(os.path.join(job, "config.json"), "w") as fp:
        fp.write(config_json_text())
    ext = ".bin" if binaryData else ".gz"
    for fb in fluidBlocksList:
        tag = "%04d-%04d-%04d" % (fb.i, fb.j, fb.k)
        grid_file = os.path.join(job, "grid", "grid-" + tag + ext)
        if binaryData:
            fb.grid.write_to_binary_file(grid_file)
        else:
            fb.grid.write_to_gzip_file(grid_file)
        fb.write_flow_data_to_file(os.path.join(job, "flow", "t0000", "flow-" + tag + ext), binaryData)
        for face, values in fb.twall_values.items():
            write_twall_file(os.path.join(job, "TWall-%s-%s.gz" % (tag, face)), values)
    with open(os.path.join(job, "times.data"), "w") as fp:
        fp.write("# tindx t\n0 0.0\n")


def reset():
    """Start a new job description (the module holds one job at a time, as job scripts
    expect)."""
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
    """The names a job script sees."""
    ns = {k: v for k, v in globals().items() if not k.startswith("_")}
    ns.update(math=math, np=np)
    return ns


def run_script(path):
    """Run a job script with this module's names in scope."""
    with open(path, "rb") as f:
        code = compile(f.read(), path, "exec")
    exec(code, globals())


USAGE = "Usage: chkn-prep [--help | -h] [--job=<jobName> | -f <jobName>] [--binary | -b]"


def main(argv=None):
    """chkn-prep --job=<name> [--binary]: run <name>.py and write the job directory <name>."""
    import getopt
    import time
    opts, _ = getopt.getopt(sys.argv[1:] if argv is None else argv, "hf:b", ["help", "job=", "binary"])
    given = dict(opts)
    if not opts or given.keys() & {"--help", "-h"}:
        print(USAGE)
        return 0
    job = given.get("--job") or given.get("-f")
    if not job:
        raise SystemExit("chkn-prep: no job given\n" + USAGE)
    config.job_name = os.path.splitext(job)[0]
    run_script(config.job_name + ".py")
    write_initial_files(bool(given.keys() & {"--binary", "-b"}))
    print("chkn-prep: wrote job %s in %.3f s" % (config.job_name, time.process_time()))
    return 0


if __name__ == "__main__":
    sys.exit(main())
