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
# config.json "time_scheme" values (a key of this build; the reference always steps TVD-RK3).
TIME_SCHEMES = ("rk3", "predictor_corrector")
# face.cu:51-68: inflow functions; unknown names mean none.
BC_FUNCTIONS = {"none": 0, "supersonic_vortex": 1, "laminar_boundary_layer": 2,
                "manufactured_solution": 3}
BL_SAMPLES = 12  # config.cu:31
# gas.cu:20-48: the two gas models the solver is built for (gamma, R, q, alpha, Ti).
IDEAL_AIR = (1.4, 287.1, 0.0, 0.0, 0.0)
AB_REACTING_GAS = (6.0 / 5.0, 287.0, 300000.0, 1000.0, 362.58)


def gas_flavour(gas_model):
    """"ab" when the gas model (gamma, R, q, alpha, Ti, ...) is the reacting A->B gas, else
    "air" (the library is built for one of them, as the reference's executables are)."""
    if gas_model is not None and all(abs(a - b) < 1.0e-4 for a, b in zip(AB_REACTING_GAS, gas_model[:5])):
        return "ab"
    return "air"
BL_PROFILE_FILE = "InflowBoundaryLayerBC.data"


def bc_code(tag):
    """face.cu:38-49: unknown tags mean wall_with_slip."""
    return BC_CODES.get(tag, 0)


def read_bl_profile(path):
    """Laminar boundary-layer profile (config.cu:37-60): p_e, the sample count (must be 12),
    then y, T, velx per sample.  Returns the 37 numbers p_e, y0, T0, velx0, y1, ..."""
    if not os.path.isfile(path):
        raise RuntimeError("Boundary-layer file is not present in current directory.")
    with open(path) as f:
        tok = f.read().split()
    if int(tok[1]) != BL_SAMPLES:
        raise RuntimeError("Boundary-layer file has unexpected number of sample points.")
    return [float(tok[0])] + [float(v) for v in tok[2:2 + 3 * BL_SAMPLES]]


def write_bl_profile(path, numbers):
    """Write a profile in the file layout read_bl_profile reads (17 significant digits)."""
    with open(path, "w") as f:
        f.write("%.17g\n%d\n" % (numbers[0], BL_SAMPLES))
        for i in range(BL_SAMPLES):
            f.write("%.17g %.17g %.17g\n" % tuple(numbers[1 + 3 * i:4 + 3 * i]))


def boundary_faces(blk, side):
    """Number of faces on a block side."""
    d = side // 2
    return (blk.njc * blk.nkc, blk.nic * blk.nkc, blk.nic * blk.njc)[d]


class JobBlock:
    __slots__ = ("i", "j", "k", "nic", "njc", "nkc", "active", "bc", "bc_fs", "bc_twall", "vertices",
                 "iovars", "label", "bc_fun", "twall_faces")

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
        self.time_scheme = 0  # TIME_SCHEMES index: 0 TVD-RK3 (the reference), 1 predictor-corrector
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
        self.bl_profile = None  # read_bl_profile() numbers when a side uses the BL function
        self.gas_model = list(IDEAL_AIR) + [IDEAL_AIR[1] / (IDEAL_AIR[0] - 1.0)]  # + Cv

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
        self.time_scheme = TIME_SCHEMES.index(cfg.get("time_scheme", "rk3"))
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
        gm = cfg["gas_model"]
        self.gas_model = [float(gm[k]) for k in ("gamma", "R", "q", "alpha", "Ti", "Cv")]
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
            jb.bc_fun = [BC_FUNCTIONS.get(b["bcs"][f].get("fun_name", "none"), 0) if jb.bc[q] == 6 else 0
                         for q, f in enumerate(FACE_NAMES)]
            jb.twall_faces = {}  # side -> TWall per boundary face, for function-form walls
            jb.label = b.get("label", "")
            blocks.append(jb)
        self.blocks = blocks

    def _read_side_data(self, twall_reader, bl_profile=None):
        """Per-face TWall of function-form walls and, when a side uses the laminar
        boundary-layer inflow, the profile given or else the one in the current directory
        (config.cu:476)."""
        for jb in self.blocks:
            for q, f in enumerate(FACE_NAMES):
                if jb.bc[q] == 2 and np.isnan(jb.bc_twall[q]):
                    jb.twall_faces[q] = np.asarray(twall_reader(jb, f), dtype=float)[:boundary_faces(jb, q)]
        if any(jb.bc[q] == 6 and jb.bc_fun[q] == 2 for jb in self.blocks for q in range(6)):
            self.bl_profile = (list(bl_profile) if bl_profile is not None
                               else read_bl_profile(os.path.join(".", BL_PROFILE_FILE)))

    @classmethod
    def from_prep(cls, bl_profile=None):
        """From the job described through chkn.prep (no files written).  bl_profile: the
        boundary-layer profile numbers (else ./InflowBoundaryLayerBC.data when used)."""
        from . import prep
        cfg = json.loads(prep.config_json_text())
        jd = cls()
        jd._apply_config(cfg)
        for jb, fb in zip(jd.blocks, prep.fluidBlocksList):
            jb.vertices = fb.grid.xyz_flat()
            jb.iovars = fb.iovars()
        # The values as the TWall files hold them ('%g').
        blocks = {(fb.i, fb.j, fb.k): fb for fb in prep.fluidBlocksList}
        jd._read_side_data(lambda jb, f: [float("%g" % v) for v in blocks[(jb.i, jb.j, jb.k)].twall_values[f]],
                           bl_profile)
        return jd

    @classmethod
    def from_job_dir(cls, job, tindx=0, binary=True, bl_profile=None):
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

        def twall(jb, f):  # block.cu:566-611
            import gzip
            path = os.path.join(job, "TWall-%04d-%04d-%04d-%s.gz" % (jb.i, jb.j, jb.k, f))
            with gzip.open(path, "rt") as fh:
                return [float(v) for v in fh.read().split()]
        jd._read_side_data(twall, bl_profile)
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
            "bc_fun": ints([getattr(b, "bc_fun", [0] * 6) for b in self.blocks]).ravel(),
            "bl_profile": dbl(self.bl_profile if self.bl_profile is not None else []),
            "gas_model": dbl(self.gas_model),
            "twall_faces": dbl(np.concatenate([np.zeros(0)] + [
                b.twall_faces[q] for b in self.blocks for q in sorted(getattr(b, "twall_faces", {}))])),
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
