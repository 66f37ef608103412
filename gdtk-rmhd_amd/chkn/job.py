"""Array form of a chicken job, as the solver reads it from config.json and the block files.

``JobData`` holds exactly what read_config_file (config.cu:347-525) and the grid/flow
readers (block.cu:785-930) extract from a job directory, so the same object can start the
HIP step in memory (``chkn_create``) or be compared with a job loaded from disk.
"""
import json
import os

import numpy as np

from .geom import StructuredGrid

BC_CODES = {"wall_with_slip": 0, "wall_no_slip_adiabatic": 1, "wall_no_slip": 1,
            "wall_no_slip_fixed_T": 2, "exchange": 3, "inflow": 4, "outflow": 5,
            "inflow_function": 6}
FACE_NAMES = ("iminus", "iplus", "jminus", "jplus", "kminus", "kplus")


def bc_code(tag):
    """face.cu:38-49: unknown tags mean wall_with_slip."""
    return BC_CODES.get(tag, 0)


class JobBlock:
    __slots__ = ("i", "j", "k", "nic", "njc", "nkc", "active", "bc", "bc_fs", "bc_twall", "vertices",
                 "iovars", "label")

    @property
    def ncells(self):
        return self.nic * self.njc * self.nkc


class JobData:
    def __init__(self):
        self.title = ""
        self.nib = self.njb = self.nkb = 0
        self.blk_ids = None
        self.blocks = []
        self.flow_states = np.zeros((0, 6))
        self.x_order = 2
        self.t_order = 2
        self.flux_calc = 0
        self.cfl_count = 10
        self.max_step = 100
        self.print_count = 20
        self.max_time = 1.0e-3
        self.dt_init = 1.0e-6
        self.cfl_times = [0.0]
        self.cfl_values = [0.5]
        self.dt_plot_times = [0.0]
        self.dt_plot_values = [1.0e-3]
        self.t0 = 0.0
        self.start_tindx = 0
        self.reacting = False
        self.viscous = False

    # ------------------------------------------------------------------ construction
    def _apply_config(self, cfg):
        self.title = cfg["title"]
        self.nib, self.njb, self.nkb = cfg["nib"], cfg["njb"], cfg["nkb"]
        self.blk_ids = np.array(cfg["blk_ids"], dtype=np.int32).reshape(self.nib, self.njb, self.nkb)
        fs = []
        for s in cfg["flow_states"]:
            fs.append([s["gas"]["p"], s["gas"]["T"], s["gas"]["YB"]] + list(s["vel"]))
        self.flow_states = np.array(fs, dtype=float).reshape(-1, 6)
        self.x_order = cfg["x_order"]
        self.t_order = cfg["t_order"]
        self.flux_calc = 1 if cfg["flux_calc"] in ("sbp_asf", "sbp-asf") else 0
        self.cfl_count = cfg["cfl_count"]
        self.max_step = cfg["max_step"]
        self.print_count = cfg["print_count"]
        self.max_time = cfg["max_time"]
        self.dt_init = cfg["dt_init"]
        self.cfl_times = list(cfg["cfl_times"])
        self.cfl_values = list(cfg["cfl_values"])
        self.dt_plot_times = list(cfg["t_change"])
        self.dt_plot_values = list(cfg["dt_plot"])
        self.reacting = bool(cfg["reacting"])
        self.viscous = bool(cfg["viscous"])
        blocks = []
        for b in cfg["fluid_blocks"]:
            jb = JobBlock()
            jb.i, jb.j, jb.k = b["i"], b["j"], b["k"]
            jb.nic, jb.njc, jb.nkc = cfg["nics"][jb.i], cfg["njcs"][jb.j], cfg["nkcs"][jb.k]
            jb.active = bool(b["active"])
            jb.bc = [bc_code(b["bcs"][f]["tag"]) for f in FACE_NAMES]
            jb.bc_fs = [b["bcs"][f].get("flow_state_index", 0) if jb.bc[q] == 4 else 0
                        for q, f in enumerate(FACE_NAMES)]
            # config.cu:463-467; NaN marks a TWall given as a function (not built).
            jb.bc_twall = [(float("nan") if b["bcs"][f].get("form") == "fun" else float(b["bcs"][f].get("value", 0.0)))
                           if jb.bc[q] == 2 else 0.0 for q, f in enumerate(FACE_NAMES)]
            jb.label = b.get("label", "")
            blocks.append(jb)
        self.blocks = blocks

    @classmethod
    def from_prep(cls):
        """From the job described through chkn.prep (no files written)."""
        from . import prep
        cfg = json.loads(prep.config_json_text())
        jd = cls()
        jd._apply_config(cfg)
        for jb, fb in zip(jd.blocks, prep.fluidBlocksList):
            jb.vertices = fb.grid.xyz_flat()
            jb.iovars = fb.iovars()
        return jd

    @classmethod
    def from_job_dir(cls, job, tindx=0, binary=True):
        with open(os.path.join(job, "config.json")) as f:
            cfg = json.load(f)
        jd = cls()
        jd._apply_config(cfg)
        ext = ".bin" if binary else ".gz"
        for jb in jd.blocks:
            g = os.path.join(job, "grid", "grid-%04d-%04d-%04d%s" % (jb.i, jb.j, jb.k, ext))
            grid = StructuredGrid.read_from_binary_file(g) if binary else StructuredGrid.read_from_gzip_file(g)
            jb.vertices = grid.xyz_flat()
            jb.iovars = read_flow_file(os.path.join(job, "flow", "t%04d" % tindx,
                                                    "flow-%04d-%04d-%04d%s" % (jb.i, jb.j, jb.k, ext)),
                                       jb.ncells, binary)
        jd.t0, jd.start_tindx = read_times(job, tindx)
        return jd

    # ------------------------------------------------------------------ flattening
    def flat(self):
        """Concatenated int32/float64 arrays in the order of chkn_desc."""
        ints = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        dbl = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        return {
            "blk_ids": ints(self.blk_ids.ravel()),
            "blk_ijk": ints([[b.i, b.j, b.k] for b in self.blocks]).ravel(),
            "blk_dims": ints([[b.nic, b.njc, b.nkc] for b in self.blocks]).ravel(),
            "active": ints([1 if b.active else 0 for b in self.blocks]),
            "bc_codes": ints([b.bc for b in self.blocks]).ravel(),
            "bc_fs": ints([b.bc_fs for b in self.blocks]).ravel(),
            "flow_states": dbl(self.flow_states.ravel() if len(self.flow_states) else np.zeros(6)),
            "vertices": dbl(np.concatenate([np.asarray(b.vertices).ravel() for b in self.blocks])),
            "iovars": dbl(np.concatenate([np.asarray(b.iovars).ravel() for b in self.blocks])),
            "cfl_times": dbl(self.cfl_times),
            "cfl_values": dbl(self.cfl_values),
            "dt_plot_times": dbl(self.dt_plot_times),
            "dt_plot_values": dbl(self.dt_plot_values),
            "bc_twall": dbl([getattr(b, "bc_twall", [0.0] * 6) for b in self.blocks]).ravel(),
        }

    @property
    def active_cells(self):
        return sum(b.ncells for b in self.blocks if b.active)


def read_flow_file(path, ncells, binary=True):
    """(13, ncells) IO variables (block.cu:882-930)."""
    if binary:
        v = np.fromfile(path, dtype=np.float64, count=13 * ncells)
    else:
        import gzip
        with gzip.open(path, "rt") as f:
            v = np.array(f.read().split(), dtype=float)[:13 * ncells]
    return v.reshape(13, ncells)


def read_times(job, tindx):
    """(t, tindx) for time index tindx from times.data (simulate.cu:134-149)."""
    t = 0.0
    path = os.path.join(job, "times.data")
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                if not line.strip() or "#" in line:
                    continue
                parts = line.split()
                if int(parts[0]) == tindx:
                    t = float(parts[1])
    return t, tindx
