"""ctypes binding of oracle/liborc.so, the CPU restatement used as the checker.

Only tests/, the smoke check and the bench's cpu_baseline leg use this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liborc.so")
LIB_AB = os.path.join(ROOT, "oracle", "liborc_ab.so")  # the reacting A->B gas build
# FP32 builds (the reference's FLOAT_NUMBERS): float arithmetic behind the same double API.
LIBS = {"air": LIB, "ab": LIB_AB, "air_f32": os.path.join(ROOT, "oracle", "liborc_f32.so"),
        "ab_f32": os.path.join(ROOT, "oracle", "liborc_ab_f32.so")}
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "chkn_ref")

_I = ctypes.POINTER(ctypes.c_int)
_D = ctypes.POINTER(ctypes.c_double)


class OrcParams(ctypes.Structure):
    _fields_ = [("x_order", ctypes.c_int), ("flux_calc", ctypes.c_int), ("cfl_count", ctypes.c_int),
                ("max_step", ctypes.c_int), ("max_time", ctypes.c_double),
                ("dt_init", ctypes.c_double), ("n_cfl", ctypes.c_int), ("cfl_times", _D),
                ("cfl_values", _D), ("t0", ctypes.c_double), ("viscous", ctypes.c_int),
                ("bc_twall", _D), ("bc_fun", _I), ("bl_profile", _D), ("n_twall_faces", ctypes.c_int),
                ("twall_faces", _D), ("reacting", ctypes.c_int), ("time_scheme", ctypes.c_int)]




_libs = {}


def load(flavour="air"):
    if flavour not in _libs:
        path = LIBS[flavour]
        if not os.path.exists(path):
            raise RuntimeError("%s is not built; run make -C oracle" % path)
        lib = ctypes.CDLL(path)
        lib.orc_create.restype = ctypes.c_void_p
        lib.orc_create.argtypes = [ctypes.c_int] * 4 + [_I] * 6 + [ctypes.c_int, _D, _D, _D,
                                                                  ctypes.POINTER(OrcParams)]
        lib.orc_destroy.argtypes = [ctypes.c_void_p]
        lib.orc_last_error.restype = ctypes.c_char_p
        lib.orc_march.restype = ctypes.c_int
        lib.orc_march.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _I]
        lib.orc_step_count.argtypes = [ctypes.c_void_p]
        lib.orc_time.restype = ctypes.c_double
        lib.orc_time.argtypes = [ctypes.c_void_p]
        lib.orc_dt.restype = ctypes.c_double
        lib.orc_dt.argtypes = [ctypes.c_void_p]
        for name in ("orc_get_iovars", "orc_get_conserved"):
            getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        lib.orc_get_dudt.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D]
        lib.orc_face_flux.argtypes = [_D, _D, ctypes.c_int, ctypes.c_int, _D]
        if hasattr(lib, "orc_face_flux_cs"):  # FP64 builds only
            lib.orc_face_flux_cs.argtypes = [_D, _D, ctypes.c_int, _D, ctypes.c_double, _D, _D]
        lib.orc_hex_props.argtypes = [_D, _D]
        lib.orc_quad_props.argtypes = [_D, _D]
        lib.orc_decode.argtypes = [_D, _D]
        lib.orc_decode.restype = ctypes.c_int
        lib.orc_encode.argtypes = [_D, _D]
        _libs[flavour] = lib
    return _libs[flavour]


def _p(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    """CPU restatement of the reference march on a JobData (fp32: the FP32 build)."""

    def __init__(self, job, fp32=False):
        from chkn.job import gas_flavour
        lib = load(gas_flavour(job.gas_model) + ("_f32" if fp32 else ""))
        f = job.flat()
        self._keep = f
        prm = OrcParams(job.x_order, job.flux_calc, job.cfl_count, job.max_step, job.max_time,
                        job.dt_init, len(job.cfl_times), _p(f["cfl_times"], _D),
                        _p(f["cfl_values"], _D), job.t0, int(job.viscous), _p(f["bc_twall"], _D),
                        _p(f["bc_fun"], _I), _p(f["bl_profile"], _D) if len(f["bl_profile"]) else None,
                        len(f["twall_faces"]), _p(f["twall_faces"], _D) if len(f["twall_faces"]) else None,
                        int(job.reacting), int(job.time_scheme))
        self._prm = prm
        h = lib.orc_create(len(job.blocks), job.nib, job.njb, job.nkb, _p(f["blk_ids"], _I),
                           _p(f["blk_ijk"], _I), _p(f["blk_dims"], _I), _p(f["active"], _I),
                           _p(f["bc_codes"], _I), _p(f["bc_fs"], _I), len(job.flow_states),
                           _p(f["flow_states"], _D), _p(f["vertices"], _D), _p(f["iovars"], _D),
                           ctypes.byref(prm))
        if not h:
            raise RuntimeError(lib.orc_last_error().decode())
        self.lib, self.h, self.job = lib, h, job

    def march(self, nsteps):
        dts = np.zeros(max(nsteps, 1))
        bad = ctypes.c_int(0)
        rc = self.lib.orc_march(self.h, nsteps, _p(dts, _D), ctypes.byref(bad))
        return rc, bad.value, dts

    @property
    def step(self):
        return self.lib.orc_step_count(self.h)

    @property
    def t(self):
        return self.lib.orc_time(self.h)

    @property
    def dt(self):
        return self.lib.orc_dt(self.h)

    def iovars(self, blk):
        n = self.job.blocks[blk].ncells
        out = np.empty((13, n))
        self.lib.orc_get_iovars(self.h, blk, _p(out, _D))
        return out

    def conserved(self, blk):
        n = self.job.blocks[blk].ncells
        out = np.empty((6, n))
        self.lib.orc_get_conserved(self.h, blk, _p(out, _D))
        return out

    def dudt(self, blk, level=0):
        n = self.job.blocks[blk].ncells
        out = np.empty((6, n))
        self.lib.orc_get_dudt(self.h, blk, level, _p(out, _D))
        return out

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def face_flux(states, frame, x_order=2, flux_calc=0):
    lib = load()
    s = np.ascontiguousarray(states, dtype=float).reshape(36)
    fr = np.ascontiguousarray(frame, dtype=float).reshape(9)
    F = np.zeros(6)
    lib.orc_face_flux(_p(s, _D), _p(fr, _D), x_order, flux_calc, _p(F, _D))
    return F


def face_flux_cs(prims, frame, direction, x_order=2, h=1.0e-30):
    """(Re F, J.direction) of the face flux at prims (4 x 6: rho, e, YB, vx, vy, vz) by the
    complex step."""
    lib = load()
    s = np.ascontiguousarray(prims, dtype=float).reshape(24)
    fr = np.ascontiguousarray(frame, dtype=float).reshape(9)
    d = np.ascontiguousarray(direction, dtype=float).reshape(24)
    Fre, jv = np.zeros(6), np.zeros(6)
    lib.orc_face_flux_cs(_p(s, _D), _p(fr, _D), x_order, _p(d, _D), h, _p(Fre, _D), _p(jv, _D))
    return Fre, jv


def hex_props(p8):
    lib = load()
    p = np.ascontiguousarray(p8, dtype=float).reshape(24)
    out = np.zeros(7)
    lib.orc_hex_props(_p(p, _D), _p(out, _D))
    return out


def quad_props(p4):
    lib = load()
    p = np.ascontiguousarray(p4, dtype=float).reshape(12)
    out = np.zeros(13)
    lib.orc_quad_props(_p(p, _D), _p(out, _D))
    return out


def decode(U):
    lib = load()
    u = np.ascontiguousarray(U, dtype=float).reshape(6)
    w = np.zeros(9)
    bad = lib.orc_decode(_p(u, _D), _p(w, _D))
    return bad, w


def encode(W9):
    lib = load()
    w = np.ascontiguousarray(W9, dtype=float).reshape(9)
    u = np.zeros(6)
    lib.orc_encode(_p(w, _D), _p(u, _D))
    return u
