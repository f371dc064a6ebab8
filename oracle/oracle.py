"""ctypes wrapper of the CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()``
and bench.py's cpu_baseline leg, always to check or time the reference's
algorithm -- never by the product package.  See aeroflux_oracle.c for the function-by-function
citations of the Go reference.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

F_P, F_U, F_V, F_W, F_DIV, F_TMP0, F_TMP1, F_TMP2 = range(8)


class _Cfg(C.Structure):
    _fields_ = [("NX", C.c_int64), ("NY", C.c_int64), ("NZ", C.c_int64), ("Ghost", C.c_int64),
                ("H", C.c_double), ("Viscosity", C.c_double), ("Density", C.c_double),
                ("DT", C.c_double), ("Workers", C.c_int64), ("GaussSeidelIters", C.c_int64)]


_libs: dict[str, C.CDLL] = {}


def load(path: str = LIB) -> C.CDLL:
    if path not in _libs:
        if not os.path.exists(path):
            from . import build as _b
            _b.build(out=path) if path != LIB else _b.build()
        lib = C.CDLL(path)
        lib.orc_normalize.argtypes = [C.POINTER(_Cfg), C.c_char_p, C.c_size_t]
        lib.orc_mac_bytes.restype = C.c_int64
        lib.orc_mac_bytes.argtypes = [C.c_int64] * 4
        lib.orc_new.restype = C.c_void_p
        lib.orc_new.argtypes = [C.POINTER(_Cfg)]
        lib.orc_free.argtypes = [C.c_void_p]
        lib.orc_field_desc.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64)]
        lib.orc_field_ptr.restype = C.POINTER(C.c_double)
        lib.orc_field_ptr.argtypes = [C.c_void_p, C.c_int]
        lib.orc_arena_bytes.restype = C.c_int64
        lib.orc_arena_bytes.argtypes = [C.c_void_p]
        for name in ("orc_step", "orc_project"):
            getattr(lib, name).restype = C.c_int
            getattr(lib, name).argtypes = [C.c_void_p]
        lib.orc_sweep_axis.restype = C.c_int
        lib.orc_sweep_axis.argtypes = [C.c_void_p, C.c_int, C.c_double]
        for name in ("orc_apply_velocity_bc", "orc_commit_velocity", "orc_compute_divergence",
                     "orc_apply_pressure_bc", "orc_subtract_gradient", "orc_solve_pressure"):
            getattr(lib, name).restype = None
            getattr(lib, name).argtypes = [C.c_void_p]
        lib.orc_gs_sweep.restype = None
        lib.orc_gs_sweep.argtypes = [C.c_void_p, C.c_int]
        for name in ("orc_max_divergence", "orc_kinetic_energy", "orc_max_velocity", "orc_cfl"):
            getattr(lib, name).restype = C.c_double
            getattr(lib, name).argtypes = [C.c_void_p]
        _libs[path] = lib
    return _libs[path]


def normalize(cfg: dict, path: str = LIB) -> dict:
    """Config.Normalize restated; raises ValueError with the reference message."""
    c = _Cfg(**{k: cfg.get(k, 0) for k, _ in _Cfg._fields_})
    buf = C.create_string_buffer(512)
    rc = load(path).orc_normalize(C.byref(c), buf, 512)
    if rc:
        raise ValueError(buf.value.decode())
    return {k: getattr(c, k) for k, _ in _Cfg._fields_}


class OracleEngine:
    """CPU engine with the reference's arena layout; fields are zero-copy
    numpy views shaped (NZ, NY, NX) of the current descriptors."""

    def __init__(self, cfg: dict, path: str = LIB):
        self.lib = load(path)
        self.cfg = normalize(cfg, path)
        self.h = self.lib.orc_new(C.byref(_Cfg(**self.cfg)))
        if not self.h:
            raise MemoryError("oracle arena allocation failed")

    def close(self):
        if self.h:
            self.lib.orc_free(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def desc(self, fid: int) -> dict:
        out = (C.c_int64 * 7)()
        self.lib.orc_field_desc(self.h, fid, out)
        return dict(zip(("Off", "NX", "NY", "NZ", "SY", "SZ", "Count"), list(out)))

    def field(self, fid: int) -> np.ndarray:
        d = self.desc(fid)
        ptr = self.lib.orc_field_ptr(self.h, fid)
        return np.ctypeslib.as_array(ptr, shape=(d["NZ"], d["NY"], d["NX"]))

    def arena(self) -> np.ndarray:
        n = self.lib.orc_arena_bytes(self.h)
        ptr = C.cast(self.lib.orc_field_ptr(self.h, F_P), C.POINTER(C.c_uint8))
        return np.ctypeslib.as_array(ptr, shape=(n,))

    def step(self) -> int:
        return self.lib.orc_step(self.h)

    def project(self) -> int:
        return self.lib.orc_project(self.h)

    def sweep_axis(self, ax: int, dt: float) -> int:
        return self.lib.orc_sweep_axis(self.h, ax, dt)

    def apply_velocity_bc(self): self.lib.orc_apply_velocity_bc(self.h)
    def commit_velocity(self): self.lib.orc_commit_velocity(self.h)
    def compute_divergence(self): self.lib.orc_compute_divergence(self.h)
    def apply_pressure_bc(self): self.lib.orc_apply_pressure_bc(self.h)
    def subtract_gradient(self): self.lib.orc_subtract_gradient(self.h)
    def solve_pressure(self): self.lib.orc_solve_pressure(self.h)
    def gs_sweep(self, color: int): self.lib.orc_gs_sweep(self.h, color)
    def max_divergence(self) -> float: return self.lib.orc_max_divergence(self.h)
    def kinetic_energy(self) -> float: return self.lib.orc_kinetic_energy(self.h)
    def max_velocity(self) -> float: return self.lib.orc_max_velocity(self.h)
    def cfl(self) -> float: return self.lib.orc_cfl(self.h)
