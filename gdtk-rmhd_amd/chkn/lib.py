"""ctypes binding of libchkn.so (include/chkn.h), the HIP/gfx950 chicken step.

There is no fallback: if the shared library is missing this module raises at import of
``load()``, so nothing silently runs on the CPU.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CHKN_LIB selects another build of the library (for A/B measurements of kernel variants).
LIB_PATH = os.environ.get("CHKN_LIB") or os.path.join(_HERE, "libchkn.so")
# The reacting A->B gas flavour (the reference's AB_REACTING_GAS build).
LIB_AB_PATH = os.path.join(_HERE, "libchkn_ab.so")
# FP32 flavours (the reference's FLOAT_NUMBERS builds): every state, geometry and time value
# is a float; the C-ABI still takes and returns doubles (exactly widened on output).
LIB_F32_PATH = os.path.join(_HERE, "libchkn_f32.so")
LIB_AB_F32_PATH = os.path.join(_HERE, "libchkn_ab_f32.so")
LIB_PATHS = {"air": LIB_PATH, "ab": LIB_AB_PATH, "air_f32": LIB_F32_PATH, "ab_f32": LIB_AB_F32_PATH}

_I = ctypes.POINTER(ctypes.c_int)
_D = ctypes.POINTER(ctypes.c_double)


class ChknDesc(ctypes.Structure):
    _fields_ = [("nblocks", ctypes.c_int), ("nib", ctypes.c_int), ("njb", ctypes.c_int),
                ("nkb", ctypes.c_int), ("blk_ids", _I), ("blk_ijk", _I), ("blk_dims", _I),
                ("active", _I), ("bc_codes", _I), ("bc_fs", _I), ("n_flow_states", ctypes.c_int),
                ("flow_states", _D), ("vertices", _D), ("iovars", _D), ("x_order", ctypes.c_int),
                ("flux_calc", ctypes.c_int), ("t_order", ctypes.c_int), ("cfl_count", ctypes.c_int),
                ("max_step", ctypes.c_int), ("print_count", ctypes.c_int),
                ("max_time", ctypes.c_double), ("dt_init", ctypes.c_double),
                ("n_cfl", ctypes.c_int), ("cfl_times", _D), ("cfl_values", _D),
                ("n_dt_plot", ctypes.c_int), ("dt_plot_times", _D), ("dt_plot_values", _D),
                ("t0", ctypes.c_double), ("start_tindx", ctypes.c_int), ("reacting", ctypes.c_int),
                ("viscous", ctypes.c_int), ("bc_twall", _D), ("bc_fun", _I), ("bl_profile", _D),
                ("n_twall_faces", ctypes.c_int), ("twall_faces", _D), ("gas_model", _D),
                ("time_scheme", ctypes.c_int)]


class ChknReport(ctypes.Structure):
    _fields_ = [("step", ctypes.c_int), ("t", ctypes.c_double), ("dt", ctypes.c_double),
                ("bad_stage", ctypes.c_int), ("bad_count", ctypes.c_int),
                ("steps_done", ctypes.c_int)]


class ChknLaunchOpts(ctypes.Structure):
    _fields_ = [("tile_i", ctypes.c_int), ("tile_j", ctypes.c_int), ("tile_k", ctypes.c_int),
                ("per_block_launch", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("persistent", ctypes.c_int), ("tile_order", ctypes.c_int),
                ("vertex_frames", ctypes.c_int)]


EXPORTS = {
    "chkn_create": (ctypes.c_int, [ctypes.POINTER(ChknDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "chkn_create_batch": (ctypes.c_int, [ctypes.POINTER(ChknDesc), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "chkn_num_instances": (ctypes.c_int, [ctypes.c_void_p]),
    "chkn_instance_blocks": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _I, _I]),
    "chkn_get_instance_report": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                                ctypes.POINTER(ChknReport)]),
    "chkn_state_size": (ctypes.c_longlong, [ctypes.c_void_p]),
    "chkn_set_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "chkn_reset_stream": (ctypes.c_int, [ctypes.c_void_p]),
    "chkn_num_flow_states": (ctypes.c_int, [ctypes.c_void_p]),
    "chkn_read_conserved_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "chkn_write_conserved_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "chkn_set_inflow_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    "chkn_get_inflow_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    "chkn_boundary_face_areas": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D]),
    "chkn_grad_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D, _D]),
    "chkn_grad_backward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, _D]),
    "chkn_flux_vjp_points": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _D, _D]),
    "chkn_flux_vjp_points_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _D, _D]),
    "chkn_load_job": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_void_p)]),
    "chkn_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ChknReport)]),
    "chkn_march": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ChknReport)]),
    "chkn_finalize": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "chkn_write_flow": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "chkn_get_iovars": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    "chkn_get_conserved": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    "chkn_set_iovars": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    "chkn_set_launch_opts": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ChknLaunchOpts)]),
    "chkn_get_report": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ChknReport)]),
    "chkn_num_blocks": (ctypes.c_int, [ctypes.c_void_p]),
    "chkn_block_dims": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _I]),
    "chkn_active_cells": (ctypes.c_longlong, [ctypes.c_void_p]),
    "chkn_last_timing": (ctypes.c_int, [ctypes.c_void_p, _D, ctypes.POINTER(ctypes.c_longlong)]),
    "chkn_real_size": (ctypes.c_int, []),
    "chkn_set_verbosity": (ctypes.c_int, [ctypes.c_int]),
    "chkn_grad_generation": (ctypes.c_longlong, [ctypes.c_void_p]),
    "chkn_grad_set_tape_budget": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong]),
    "chkn_grad_set_gather_form": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "chkn_grad_gather_form": (ctypes.c_int, [ctypes.c_void_p]),
    "chkn_grad_resident_segments": (ctypes.c_int, [ctypes.c_void_p]),
    "chkn_check_block_dims": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "chkn_device_bytes": (ctypes.c_longlong, [ctypes.c_void_p]),
    "chkn_source_digest": (ctypes.c_char_p, []),
    "chkn_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "chkn_destroy": (None, [ctypes.c_void_p]),
}



_libs = {}


def load(flavour="air"):
    """Load libchkn.so (flavour "air"), libchkn_ab.so ("ab", the reacting A->B gas) or their
    FP32 builds ("air_f32", "ab_f32") and declare the C-ABI; raises if it has not been built."""
    if flavour not in LIB_PATHS:
        raise ValueError("unknown library flavour %r" % (flavour,))
    if flavour not in _libs:
        path = LIB_PATHS[flavour]
        if not os.path.exists(path):
            raise RuntimeError("%s is not built (expected at %s); run __graft_entry__.build() "
                               "or make -C gdtk-rmhd_amd" % (os.path.basename(path), path))
        lib = ctypes.CDLL(path)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _libs[flavour] = lib
    return _libs[flavour]


def tree_digest():
    """The digest chkn_source_digest() of a library built from the sources in this tree:
    SHA-256 over gdtk-rmhd_amd/csrc/*.h, *.hip and *.cpp in name order, then include/chkn.h
    (the Makefile's SRC_DIGEST), first 16 hex digits."""
    import glob
    import hashlib
    pkg = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    files = []
    for pat in ("*.h", "*.hip", "*.cpp"):
        files += glob.glob(os.path.join(pkg, "csrc", pat))
    files = sorted(files, key=lambda f: "csrc/" + os.path.basename(f))
    h = hashlib.sha256()
    for f in files + [os.path.join(os.path.dirname(pkg), "include", "chkn.h")]:
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def source_digest(flavour="air"):
    """chkn_source_digest() of a built library flavour."""
    return load(flavour).chkn_source_digest().decode()


class ChknError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


class BadCellError(ChknError):
    """A stage produced bad cells: the reference's "Stage k bad cell count: n"."""

    def __init__(self, stage, count, msg):
        super().__init__(count, msg)
        self.stage = stage
        self.count = count


def _ptr(a, typ):
    return a.ctypes.data_as(typ)


def _check_dev(t, n):
    import torch
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() == n):
        raise ValueError("expected a contiguous float64 device tensor of %d elements" % n)


def flux_vjp_points(prims, frames, fbar, x_order=2, noalloc=False):
    """Face-flux vector-Jacobian products on the device (test hook of the adjoint kernels).
    prims (n, 4, 6), frames (n, 9), fbar (n, 6) -> (n, 4, 6).  noalloc: the entry point that
    stages through the library's fixed device array (at most 4096 points) instead of
    allocating."""
    lib = load()
    n = len(prims)
    inp = np.ascontiguousarray(np.concatenate([np.reshape(prims, (n, 24)), np.reshape(frames, (n, 9)),
                                               np.reshape(fbar, (n, 6))], axis=1), dtype=np.float64)
    out = np.zeros((n, 24))
    fn = lib.chkn_flux_vjp_points_noalloc if noalloc else lib.chkn_flux_vjp_points
    rc = fn(n, x_order, _ptr(inp, _D), _ptr(out, _D))
    if rc != 0:
        raise ChknError(rc, lib.chkn_last_error(None).decode())
    return out.reshape(n, 4, 6)


def _desc(job, keep):
    """Fill a ChknDesc from a JobData; the flat arrays are appended to keep."""
    f = job.flat()
    keep.append(f)
    d = ChknDesc()
    d.nblocks = len(job.blocks)
    d.nib, d.njb, d.nkb = job.nib, job.njb, job.nkb
    for key in ("blk_ids", "blk_ijk", "blk_dims", "active", "bc_codes", "bc_fs"):
        setattr(d, key, _ptr(f[key], _I))
    d.n_flow_states = len(job.flow_states)
    for key in ("flow_states", "vertices", "iovars", "cfl_times", "cfl_values",
                "dt_plot_times", "dt_plot_values", "bc_twall"):
        setattr(d, key, _ptr(f[key], _D))
    d.x_order, d.flux_calc, d.t_order = job.x_order, job.flux_calc, job.t_order
    d.cfl_count, d.max_step, d.print_count = job.cfl_count, job.max_step, job.print_count
    d.max_time, d.dt_init = job.max_time, job.dt_init
    d.n_cfl = len(job.cfl_times)
    d.n_dt_plot = len(job.dt_plot_times)
    d.t0, d.start_tindx = job.t0, job.start_tindx
    d.reacting, d.viscous = int(job.reacting), int(job.viscous)
    d.bc_fun = _ptr(f["bc_fun"], _I)
    d.bl_profile = _ptr(f["bl_profile"], _D) if len(f["bl_profile"]) else None
    d.n_twall_faces = len(f["twall_faces"])
    d.twall_faces = _ptr(f["twall_faces"], _D) if len(f["twall_faces"]) else None
    d.gas_model = _ptr(f["gas_model"], _D)
    d.time_scheme = int(job.time_scheme)
    return d


class Sim:
    """One simulation handle (chkn_sim): a single job (``job`` or ``job_dir``) or an ensemble
    of independent jobs stepped together (``jobs``)."""

    def __init__(self, job=None, job_dir=None, tindx=0, binary=True, jobs=None, fp32=False):
        """fp32: run the FP32 build (as the reference's chkn-run-fp32); its job files must be
        gzipped text (binary=False), as the reference's FP32 build reads no binary files."""
        from .job import gas_flavour
        if job_dir is not None:
            import json
            with open(os.path.join(job_dir, "config.json")) as fp:
                gm = json.load(fp).get("gas_model", {})
            flavour = gas_flavour([gm.get(k, 0.0) for k in ("gamma", "R", "q", "alpha", "Ti")])
        else:
            flavour = gas_flavour((jobs or [job])[0].gas_model)
        lib = load(flavour + ("_f32" if fp32 else ""))
        h = ctypes.c_void_p()
        self._keep = []
        if job_dir is not None:
            rc = lib.chkn_load_job(os.fsencode(job_dir), tindx, 1 if binary else 0, ctypes.byref(h))
        elif jobs is not None:
            arr = (ChknDesc * len(jobs))(*[_desc(j, self._keep) for j in jobs])
            self._keep.append(arr)
            rc = lib.chkn_create_batch(arr, len(jobs), ctypes.byref(h))
        else:
            d = _desc(job, self._keep)
            rc = lib.chkn_create(ctypes.byref(d), ctypes.byref(h))
        if rc != 0:
            raise ChknError(rc, lib.chkn_last_error(None).decode())
        self._lib = lib
        self.h = h
        self.nblocks = lib.chkn_num_blocks(h)
        self.ninstances = lib.chkn_num_instances(h)

    def _check(self, rc, rep=None):
        if rc > 0 and rep is not None and rep.bad_stage:
            raise BadCellError(rep.bad_stage, rep.bad_count, self._lib.chkn_last_error(self.h).decode())
        if rc != 0:
            raise ChknError(rc, self._lib.chkn_last_error(self.h).decode())

    def step(self, nsteps=1):
        rep = ChknReport()
        self._check(self._lib.chkn_step(self.h, int(nsteps), ctypes.byref(rep)), rep)
        return rep

    def march(self, binary=True):
        rep = ChknReport()
        self._check(self._lib.chkn_march(self.h, 1 if binary else 0, ctypes.byref(rep)), rep)
        return rep

    def finalize(self, binary=True):
        self._check(self._lib.chkn_finalize(self.h, 1 if binary else 0))

    def write_flow(self, tindx, binary=True):
        self._check(self._lib.chkn_write_flow(self.h, tindx, 1 if binary else 0))

    def dims(self, blk):
        d = np.zeros(3, dtype=np.int32)
        self._check(self._lib.chkn_block_dims(self.h, blk, _ptr(d, _I)))
        return tuple(int(v) for v in d)

    def ncells(self, blk):
        a, b, c = self.dims(blk)
        return a * b * c

    def iovars(self, blk):
        out = np.empty((13, self.ncells(blk)))
        self._check(self._lib.chkn_get_iovars(self.h, blk, _ptr(out, _D)))
        return out

    def conserved(self, blk):
        out = np.empty((6, self.ncells(blk)))
        self._check(self._lib.chkn_get_conserved(self.h, blk, _ptr(out, _D)))
        return out

    def set_iovars(self, blk, io):
        io = np.ascontiguousarray(io, dtype=np.float64)
        self._check(self._lib.chkn_set_iovars(self.h, blk, _ptr(io, _D)))

    def set_launch(self, tile_i=0, tile_j=0, tile_k=0, per_block_launch=False, use_graph=False,
                   persistent=False, tile_order=0, vertex_frames=False):
        o = ChknLaunchOpts(tile_i, tile_j, tile_k, int(per_block_launch), int(use_graph),
                           int(persistent), int(tile_order), int(vertex_frames))
        self._check(self._lib.chkn_set_launch_opts(self.h, ctypes.byref(o)))

    def report(self):
        rep = ChknReport()
        self._check(self._lib.chkn_get_report(self.h, ctypes.byref(rep)))
        return rep

    def instance_blocks(self, inst):
        """(first, count) global block numbers of instance inst."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.chkn_instance_blocks(self.h, inst, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def instance_report(self, inst):
        rep = ChknReport()
        self._check(self._lib.chkn_get_instance_report(self.h, inst, ctypes.byref(rep)))
        return rep

    # ---------------------------------------------------------------- gradients
    @property
    def state_size(self):
        """Doubles in a flat conserved-state array (every block's 6*ncells, block order)."""
        return int(self._lib.chkn_state_size(self.h))

    @property
    def num_flow_states(self):
        return int(self._lib.chkn_num_flow_states(self.h))

    def read_conserved(self, out):
        """Copy the conserved state into `out`, a contiguous float64 device tensor."""
        _check_dev(out, self.state_size)
        self._check(self._lib.chkn_read_conserved_dev(self.h, ctypes.c_void_p(out.data_ptr())))

    def write_conserved(self, src):
        """Set the conserved state (and W = decode(U)) from a float64 device tensor."""
        _check_dev(src, self.state_size)
        self._check(self._lib.chkn_write_conserved_dev(self.h, ctypes.c_void_p(src.data_ptr())))

    def set_inflow_state(self, state, w6):
        w = np.ascontiguousarray(w6, dtype=np.float64).reshape(6)
        self._check(self._lib.chkn_set_inflow_state(self.h, int(state), _ptr(w, _D)))

    def get_inflow_state(self, state):
        w = np.zeros(6)
        self._check(self._lib.chkn_get_inflow_state(self.h, int(state), _ptr(w, _D)))
        return w

    def boundary_face_areas(self, blk, side):
        a, b, c = self.dims(blk)
        d = side // 2
        n = (b * c) if d == 0 else ((a * c) if d == 1 else (a * b))
        out = np.zeros(n)
        self._check(self._lib.chkn_boundary_face_areas(self.h, int(blk), int(side), _ptr(out, _D)))
        return out

    def grad_forward(self, nsteps, checkpoint_every, dts=None):
        """Advance nsteps steps recording checkpoints; returns the dts used, (nsteps, ninst)."""
        out = np.zeros((nsteps, self.ninstances))
        dptr = None
        if dts is not None:
            dts = np.ascontiguousarray(dts, dtype=np.float64).reshape(nsteps, self.ninstances)
            dptr = _ptr(dts, _D)
        self._check(self._lib.chkn_grad_forward(self.h, int(nsteps), int(checkpoint_every), dptr,
                                                _ptr(out, _D)))
        return out

    def set_tape_budget(self, nbytes):
        """Device bytes grad_forward may fill with resident tapes of its last checkpoint
        segments (reversed without recomputation): None or a negative value for half of the
        free device memory (the default), 0 to recompute every segment."""
        self._check(self._lib.chkn_grad_set_tape_budget(self.h, -1 if nbytes is None else int(nbytes)))

    def set_gather_form(self, separate):
        """separate=False (the default): the k-direction face pass of each reversed stage
        finishes the cells itself where the blocks allow it; True: a separate per-cell gather
        launch.  The gradient is the same bit for bit."""
        self._check(self._lib.chkn_grad_set_gather_form(self.h, 1 if separate else 0))

    @property
    def gather_form(self):
        """0: the reversed stages run in the fused form; 1: with a separate gather launch."""
        rc = self._lib.chkn_grad_gather_form(self.h)
        self._check(min(rc, 0))
        return rc

    @property
    def grad_resident_segments(self):
        """Checkpoint segments of the last grad_forward held on resident tapes."""
        return int(self._lib.chkn_grad_resident_segments(self.h))

    @property
    def grad_generation(self):
        """Number of grad_forward calls so far; grad_backward reverses the latest one."""
        return int(self._lib.chkn_grad_generation(self.h))

    def grad_backward(self, lam, grad_u0):
        """Reverse the last grad_forward: lam, grad_u0 are flat float64 device tensors.
        Returns dLoss/d(rho, e, YB, vx, vy, vz) of each inflow state, (nstates, 6)."""
        _check_dev(lam, self.state_size)
        _check_dev(grad_u0, self.state_size)
        gin = np.zeros((max(self.num_flow_states, 1), 6))
        self._check(self._lib.chkn_grad_backward(self.h, ctypes.c_void_p(lam.data_ptr()),
                                                 ctypes.c_void_p(grad_u0.data_ptr()),
                                                 _ptr(gin, _D)))
        return gin[:self.num_flow_states]

    def set_stream(self, stream):
        """Issue the handle's work on `stream` (a torch.cuda.Stream, or a raw hipStream_t as
        int; 0 is the null stream), or on its own stream again with None."""
        if stream is None:
            self._check(self._lib.chkn_reset_stream(self.h))
            return
        raw = int(getattr(stream, "cuda_stream", stream))
        self._check(self._lib.chkn_set_stream(self.h, ctypes.c_void_p(raw or None)))

    def last_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        self._check(self._lib.chkn_last_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    @property
    def active_cells(self):
        return int(self._lib.chkn_active_cells(self.h))

    def close(self):
        if getattr(self, "h", None):
            self._lib.chkn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
