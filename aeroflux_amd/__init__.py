"""Python mirror of the AeroFlux Go API over the MI355X C ABI.

The names, argument meanings and error behaviour follow the reference's
exported API (AeroFlex/aeroflux.go:14-439, grid.go:8-193, project.go:210-258,
stencil.go:319-338): ``Config`` + ``Config.Normalize()``, ``New(cfg)``,
``Engine.Step()``, ``MaxDivergence()``, ``KineticEnergy()``,
``MaxVelocity()``, ``CFL()``, ``Close()``, ``EnableSparseSolver()``,
``Engine.Mesh`` with its ``Field`` descriptors and ``Data``.  Go ``error``
returns become :class:`AeroFluxError` exceptions carrying the same
``"aeroflux: ..."`` messages.

Every call goes through ``libaeroflux.so`` (hand-written HIP for gfx950).
There is no CPU fallback: without the built library or a GPU the calls
raise.

Differences forced by device memory, by design:

* ``Mesh.Data(field)`` returns a host *copy* of the padded field shaped
  ``(NZ, NY, NX)`` (Go returns a live view aliasing the arena,
  grid.go:147-149); write a modified array back with ``Mesh.SetData``.
* ``Field`` values name a role (P, U, ...); ``Data`` always reads the
  current descriptor of that role, as ``m.Data(&m.U)`` does in Go.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, fields as dc_fields

import numpy as np

__all__ = [
    "AeroFluxError", "Config", "Field", "MAC", "Engine", "New", "MACBytes",
    "FIELD_P", "FIELD_U", "FIELD_V", "FIELD_W", "FIELD_DIV", "FIELD_TMP0",
    "FIELD_TMP1", "FIELD_TMP2", "AXIS_X", "AXIS_Y", "AXIS_Z", "COLOR_RED",
    "COLOR_BLACK", "PRECISION_FP64", "PRECISION_MIXED", "lib_path", "load_library",
    "Batch", "BatchConfig", "SlotBytes", "Serve",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libaeroflux.so")

AF_OK, AF_ERR_CONFIG, AF_ERR_FAULT, AF_ERR_OOM, AF_ERR_HIP, AF_ERR_ARG, AF_ERR_CLOSED, \
    AF_ERR_UNSUPPORTED = range(8)

FIELD_P, FIELD_U, FIELD_V, FIELD_W, FIELD_DIV, FIELD_TMP0, FIELD_TMP1, FIELD_TMP2 = range(8)
FIELD_NAMES = ("P", "U", "V", "W", "Div", "Tmp0", "Tmp1", "Tmp2")
AXIS_X, AXIS_Y, AXIS_Z = range(3)
COLOR_RED, COLOR_BLACK = range(2)
PRECISION_FP64, PRECISION_MIXED = range(2)
STREAM_READ, STREAM_COPY, STREAM_2R1W, STREAM_4R3W = range(4)
SERVE_FCFS, SERVE_LEAST_WORK_PREEMPT, SERVE_LEAST_WORK_STEPS, SERVE_LEAST_WORK = 0, 1, 2, 3
DIRECT_AUTO, DIRECT_FFT, DIRECT_DENSE = 0, 1, 2


class AeroFluxError(RuntimeError):
    """A Go ``error`` from the reference API, raised.  ``code`` is the C ABI
    return code (AF_ERR_*)."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class _CConfig(C.Structure):
    _fields_ = [
        ("NX", C.c_int64), ("NY", C.c_int64), ("NZ", C.c_int64), ("Ghost", C.c_int64),
        ("H", C.c_double), ("Viscosity", C.c_double), ("Density", C.c_double),
        ("DT", C.c_double), ("Workers", C.c_int64), ("GaussSeidelIters", C.c_int64),
        ("UseSparseSolver", C.c_int32), ("Lock", C.c_int32),
        ("Precision", C.c_int32), ("MixedMaxCycles", C.c_int32), ("MixedSweeps", C.c_int32),
        ("MixedFixSweeps", C.c_int32), ("MixedTarget", C.c_double),
        ("TileK", C.c_int32), ("TileRows", C.c_int32), ("TileZ", C.c_int32),
        ("SweepShape", C.c_int32), ("SweepSlab", C.c_int32), ("PassSlab", C.c_int32),
    ]


class _CField(C.Structure):
    _fields_ = [("Off", C.c_int64), ("NX", C.c_int64), ("NY", C.c_int64), ("NZ", C.c_int64),
                ("SY", C.c_int64), ("SZ", C.c_int64), ("Count", C.c_int64)]


class _CBatchConfig(C.Structure):
    _fields_ = [
        ("Precision", C.c_int32), ("MaxSlots", C.c_int32), ("MemoryBudget", C.c_int64),
        ("GaussSeidelIters", C.c_int64), ("MixedMaxCycles", C.c_int32),
        ("MixedSweeps", C.c_int32), ("MixedFixSweeps", C.c_int32), ("MixedChecked", C.c_int32),
        ("SweepShape", C.c_int32), ("Workgroups", C.c_int32), ("SweepStaged", C.c_int32),
        ("ServePolicy", C.c_int32), ("PreemptRatio", C.c_double), ("SnapshotBudget", C.c_int64),
        ("StepCellBudget", C.c_int64), ("SolveGroupBytes", C.c_int64),
        ("StepMaxWait", C.c_int32),
    ]


class _CBatchInfo(C.Structure):
    _fields_ = [
        ("Slots", C.c_int32), ("SweepShape", C.c_int32), ("SweepWorkgroups", C.c_int32),
        ("PassWorkgroups", C.c_int32), ("OtherWorkgroups", C.c_int32), ("Tuned", C.c_int32),
        ("Cells", C.c_int64), ("BytesInUse", C.c_int64), ("BytesLargestFree", C.c_int64),
        ("MemoryBudget", C.c_int64), ("SweepStaged", C.c_int32), ("StagedSlots", C.c_int32),
        ("SolveGroups", C.c_int32),
    ]


class SnapshotHeader(C.Structure):
    """af_snapshot_header: the 128 bytes at the start of a snapshot."""
    _fields_ = [
        ("Magic", C.c_char * 8), ("Version", C.c_uint32), ("HeaderBytes", C.c_uint32),
        ("NX", C.c_int64), ("NY", C.c_int64), ("NZ", C.c_int64), ("Ghost", C.c_int64),
        ("H", C.c_double), ("DT", C.c_double), ("Viscosity", C.c_double),
        ("Density", C.c_double), ("PayloadBytes", C.c_int64), ("Checksum", C.c_uint64),
        ("Reserved", C.c_uint8 * 32),
    ]


def _raw_buffer(buf) -> tuple[int, int]:
    """(address, bytes) of a snapshot buffer: a numpy array, bytes or
    bytearray in host memory, or an (address, bytes) pair for memory from
    the HIP runtime the library links (device or pinned host)."""
    if isinstance(buf, tuple) and len(buf) == 2:
        return int(buf[0]), int(buf[1])
    if isinstance(buf, np.ndarray):
        if not buf.flags.c_contiguous:
            raise AeroFluxError(AF_ERR_ARG, "aeroflux: snapshot array must be contiguous")
        return int(buf.ctypes.data), int(buf.nbytes)
    if isinstance(buf, (bytes, bytearray, memoryview)):
        arr = np.frombuffer(buf, dtype=np.uint8)
        return int(arr.ctypes.data), int(arr.nbytes)
    raise AeroFluxError(AF_ERR_ARG, f"aeroflux: unsupported snapshot buffer {type(buf)!r}")


class _CServeJob(C.Structure):
    _fields_ = [
        ("NX", C.c_int64), ("NY", C.c_int64), ("NZ", C.c_int64), ("Steps", C.c_int64),
        ("Viscosity", C.c_double), ("DT", C.c_double), ("H", C.c_double),
        ("Seed", C.c_uint64), ("Arrival", C.c_double), ("MixedTarget", C.c_double),
    ]


class _CServeResult(C.Structure):
    _fields_ = [
        ("Admitted", C.c_double), ("Done", C.c_double), ("MaxDivergence", C.c_double),
        ("KineticEnergy", C.c_double), ("Status", C.c_int32), ("Slot", C.c_int32),
        ("FirstStep", C.c_int64), ("DoneStep", C.c_int64),
    ]


class _CServeStats(C.Structure):
    _fields_ = [
        ("Wall", C.c_double), ("Steps", C.c_int64), ("CellUpdates", C.c_int64),
        ("MaxResident", C.c_int32), ("Preemptions", C.c_int32), ("PeakBytes", C.c_int64),
    ]


# af_serve_progress: (const af_serve_stats *, int64_t jobs_done, void *user).
_ServeProgress = C.CFUNCTYPE(None, C.POINTER(_CServeStats), C.c_int64, C.c_void_p)


_lib = None

# (name, restype, argtypes) of every exported symbol in include/aeroflux.h.
_SIGNATURES = [
    ("af_config_normalize", C.c_int, [C.POINTER(_CConfig)]),
    ("af_mac_bytes", C.c_int64, [C.c_int64] * 4),
    ("af_new", C.c_int, [C.POINTER(_CConfig), C.POINTER(C.c_void_p)]),
    ("af_close", C.c_int, [C.c_void_p]),
    ("af_free", None, [C.c_void_p]),
    ("af_get_config", C.c_int, [C.c_void_p, C.POINTER(_CConfig)]),
    ("af_step", C.c_int, [C.c_void_p]),
    ("af_step_n", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    ("af_max_divergence", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("af_kinetic_energy", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("af_max_velocity", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("af_cfl", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("af_max_divergence_noalloc", C.c_int, [C.c_uint64] * 3 + [C.c_int64] * 4
     + [C.c_double, C.c_uint64, C.c_uint64]),
    ("af_enable_sparse_solver", C.c_int, [C.c_void_p]),
    ("af_enable_direct_solver", C.c_int, [C.c_void_p]),
    ("af_enable_direct_solver_method", C.c_int, [C.c_void_p, C.c_int32]),
    ("af_direct_solver_enabled", C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    ("af_field_desc", C.c_int, [C.c_void_p, C.c_int, C.POINTER(_CField)]),
    ("af_field_read", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64]),
    ("af_field_write", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64]),
    ("af_field_device_ptr", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint64)]),
    ("af_arena_bytes", C.c_int64, [C.c_void_p]),
    ("af_snapshot_bytes", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("af_snapshot", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("af_restore", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("af_apply_velocity_bc", C.c_int, [C.c_void_p]),
    ("af_sweep_axis", C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    ("af_commit_velocity", C.c_int, [C.c_void_p]),
    ("af_compute_divergence", C.c_int, [C.c_void_p]),
    ("af_gs_sweep", C.c_int, [C.c_void_p, C.c_int]),
    ("af_apply_pressure_bc", C.c_int, [C.c_void_p]),
    ("af_solve_pressure", C.c_int, [C.c_void_p]),
    ("af_subtract_gradient", C.c_int, [C.c_void_p]),
    ("af_project", C.c_int, [C.c_void_p]),
    ("af_time_kernel", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    ("af_measure_stream", C.c_int, [C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double)]),
    ("af_mixed_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    ("af_solver_tile", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32)]),
    ("af_sweep_shape", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32)]),
    ("af_sweep_slab", C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    ("af_pass_slab", C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    ("af_batch_new", C.c_int, [C.POINTER(_CBatchConfig), C.POINTER(C.c_void_p)]),
    ("af_batch_close", C.c_int, [C.c_void_p]),
    ("af_batch_free", None, [C.c_void_p]),
    ("af_batch_get_info", C.c_int, [C.c_void_p, C.POINTER(_CBatchInfo)]),
    ("af_batch_slot_bytes", C.c_int64, [C.POINTER(_CConfig), C.c_int32]),
    ("af_batch_add", C.c_int, [C.c_void_p, C.POINTER(_CConfig), C.POINTER(C.c_int32)]),
    ("af_batch_remove", C.c_int, [C.c_void_p, C.c_int32]),
    ("af_batch_step_n", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int32), C.c_int32,
                                  C.POINTER(C.c_int32)]),
    ("af_batch_field_desc", C.c_int, [C.c_void_p, C.c_int32, C.c_int, C.POINTER(_CField)]),
    ("af_batch_field_read", C.c_int, [C.c_void_p, C.c_int32, C.c_int, C.c_void_p, C.c_int64]),
    ("af_batch_field_write", C.c_int, [C.c_void_p, C.c_int32, C.c_int, C.c_void_p, C.c_int64]),
    ("af_batch_apply_velocity_bc", C.c_int, [C.c_void_p, C.c_int32]),
    ("af_batch_max_divergence", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    ("af_batch_kinetic_energy", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    ("af_batch_mixed_stats", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                                       C.POINTER(C.c_int32)]),
    ("af_batch_snapshot_bytes", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]),
    ("af_batch_snapshot", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]),
    ("af_batch_restore", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]),
    ("af_batch_tune", C.c_int, [C.c_void_p]),
    ("af_batch_time_stage", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    ("af_init_band_limited", C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.c_double,
                                        C.c_double]),
    ("af_batch_init_band_limited", C.c_int, [C.c_void_p, C.c_int32, C.c_uint64, C.c_int32,
                                              C.c_double, C.c_double]),
    ("af_serve", C.c_int, [C.c_void_p, C.POINTER(_CServeJob), C.c_int64, C.c_int32, C.c_double,
                            C.c_double, C.c_int32, C.POINTER(_CServeResult),
                            C.POINTER(_CServeStats), _ServeProgress, C.c_double, C.c_void_p]),
    ("af_last_error", C.c_char_p, []),
]


def lib_path() -> str:
    return _LIB_PATH


def load_library():
    """Load ``libaeroflux.so`` (built by ``python -m aeroflux_amd.build``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise ImportError(
                f"aeroflux: {_LIB_PATH} is not built; run `python -m aeroflux_amd.build`")
        lib = C.CDLL(_LIB_PATH)
        for name, res, args in _SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _check(rc: int) -> None:
    if rc != AF_OK:
        msg = load_library().af_last_error().decode()
        raise AeroFluxError(rc, msg or f"aeroflux: error code {rc}")


@dataclass
class Config:
    """Config, aeroflux.go:14-41 (plus the MI355X-build extensions)."""
    NX: int = 0
    NY: int = 0
    NZ: int = 0
    Ghost: int = 0
    H: float = 0.0
    Viscosity: float = 0.0
    Density: float = 0.0
    DT: float = 0.0
    Workers: int = 0
    GaussSeidelIters: int = 0
    UseSparseSolver: bool = False
    Lock: bool = False
    Precision: int = PRECISION_FP64
    MixedMaxCycles: int = 0
    MixedSweeps: int = 0
    MixedFixSweeps: int = 0
    MixedTarget: float = 0.0
    TileK: int = 0
    TileRows: int = 0
    TileZ: int = 0
    SweepShape: int = 0
    SweepSlab: int = 0
    PassSlab: int = 0

    def _to_c(self) -> _CConfig:
        c = _CConfig()
        for f in dc_fields(self):
            v = getattr(self, f.name)
            setattr(c, f.name, int(v) if isinstance(v, bool) else v)
        return c

    def _from_c(self, c: _CConfig) -> None:
        for f in dc_fields(self):
            v = getattr(c, f.name)
            if f.type in ("bool", bool):
                v = bool(v)
            setattr(self, f.name, v)

    def Normalize(self) -> None:
        """Config.Normalize, aeroflux.go:47-113: fills defaults in place,
        raises AeroFluxError on a geometry the solver cannot represent."""
        c = self._to_c()
        rc = load_library().af_config_normalize(C.byref(c))
        self._from_c(c)
        _check(rc)


@dataclass(frozen=True)
class Field:
    """Field, grid.go:12-21.  ``id`` names the role (FIELD_*)."""
    id: int
    Off: int
    NX: int
    NY: int
    NZ: int
    SY: int
    SZ: int
    Count: int

    def Index(self, x: int, y: int, z: int) -> int:  # grid.go:24-26
        return x + y * self.SY + z * self.SZ

    def ByteOffset(self, x: int, y: int, z: int) -> int:  # grid.go:29-31
        return self.Off + self.Index(x, y, z) * 8

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.NZ, self.NY, self.NX)


def MACBytes(nx: int, ny: int, nz: int, ghost: int) -> int:
    """MACBytes, grid.go:138-144."""
    return int(load_library().af_mac_bytes(nx, ny, nz, ghost))


class MAC:
    """The engine's staggered grid (MAC, grid.go:52-73) over HBM."""

    def __init__(self, engine: "Engine"):
        self._e = engine
        cfg = engine.Cfg
        self.NX, self.NY, self.NZ, self.Ghost, self.H = cfg.NX, cfg.NY, cfg.NZ, cfg.Ghost, cfg.H

    def field(self, fid: int) -> Field:
        f = _CField()
        _check(load_library().af_field_desc(self._e._handle(), fid, C.byref(f)))
        return Field(fid, f.Off, f.NX, f.NY, f.NZ, f.SY, f.SZ, f.Count)

    P = property(lambda s: s.field(FIELD_P))
    U = property(lambda s: s.field(FIELD_U))
    V = property(lambda s: s.field(FIELD_V))
    W = property(lambda s: s.field(FIELD_W))
    Div = property(lambda s: s.field(FIELD_DIV))
    Tmp0 = property(lambda s: s.field(FIELD_TMP0))
    Tmp1 = property(lambda s: s.field(FIELD_TMP1))
    Tmp2 = property(lambda s: s.field(FIELD_TMP2))

    @property
    def Bytes(self) -> int:
        return int(load_library().af_arena_bytes(self._e._handle()))

    def Data(self, f: Field) -> np.ndarray:
        """Host copy of the current padded field of f's role, (NZ, NY, NX)."""
        cur = self.field(f.id)
        out = np.empty(cur.shape, dtype=np.float64)
        _check(load_library().af_field_read(self._e._handle(), cur.id,
                                            out.ctypes.data_as(C.c_void_p), cur.Count))
        return out

    def SetData(self, f: Field, data: np.ndarray) -> None:
        """Write a whole padded field (any shape with Count float64s)."""
        cur = self.field(f.id)
        arr = np.ascontiguousarray(data, dtype=np.float64).reshape(-1)
        if arr.size != cur.Count:
            raise AeroFluxError(AF_ERR_ARG, f"aeroflux: field has {cur.Count} elements, "
                                            f"got {arr.size}")
        _check(load_library().af_field_write(self._e._handle(), cur.id,
                                             arr.ctypes.data_as(C.c_void_p), cur.Count))

    def DevicePtr(self, f: Field) -> int:
        out = C.c_uint64()
        _check(load_library().af_field_device_ptr(self._e._handle(), f.id, C.byref(out)))
        return int(out.value)


class Engine:
    """Engine, aeroflux.go:133-172: owns the HBM arena and the step graph."""

    def __init__(self, handle: int, cfg: Config):
        self._h = handle
        self.Cfg = cfg
        self.Mesh = MAC(self)

    def _handle(self) -> int:
        if self._h is None:
            raise AeroFluxError(AF_ERR_CLOSED, "aeroflux: engine is closed")
        return self._h

    # -- public API ---------------------------------------------------------
    def Step(self) -> None:
        """Step, aeroflux.go:348-409."""
        _check(load_library().af_step(self._handle()))

    def StepN(self, n: int, first_fault: bool = False) -> None:
        """n Steps with one synchronisation (no reference counterpart).  A
        contained device fault raises AeroFluxError after the n steps.  With
        ``first_fault`` the flag is read after every step instead (one
        synchronisation per step, as n calls of Step), the call stops at the
        first faulting step and the error's ``step`` names it."""
        lib = load_library()
        if not first_fault:
            _check(lib.af_step_n(self._handle(), n, None))
            return
        first = C.c_int64(-1)
        try:
            _check(lib.af_step_n(self._handle(), n, C.byref(first)))
        except AeroFluxError as err:
            err.step = int(first.value)
            raise

    def _scalar(self, fn) -> float:
        out = C.c_double()
        _check(fn(self._handle(), C.byref(out)))
        return float(out.value)

    def MaxDivergence(self) -> float:  # project.go:210-232
        return self._scalar(load_library().af_max_divergence)

    def KineticEnergy(self) -> float:  # project.go:238-258
        return self._scalar(load_library().af_kinetic_energy)

    def MaxVelocity(self) -> float:  # stencil.go:319-331
        return self._scalar(load_library().af_max_velocity)

    def CFL(self) -> float:  # stencil.go:336-338
        return self._scalar(load_library().af_cfl)

    def EnableSparseSolver(self) -> None:  # sparse_fallback.go:17-19
        _check(load_library().af_enable_sparse_solver(self._handle()))

    def snapshot_bytes(self) -> int:
        n = C.c_int64()
        _check(load_library().af_snapshot_bytes(self._handle(), C.byref(n)))
        return int(n.value)

    def Snapshot(self, out=None):
        """Snapshot of the engine's whole state (af_snapshot): a header and
        the eight padded fields.  Returns a new host uint8 array, or writes
        into `out` (a host numpy array, or an (address, bytes) pair of
        device or pinned memory, at least snapshot_bytes() long) and
        returns it."""
        if out is None:
            out = np.empty(self.snapshot_bytes(), dtype=np.uint8)
        ptr, n = _raw_buffer(out)
        _check(load_library().af_snapshot(self._handle(), C.c_void_p(ptr), n))
        return out

    def Restore(self, snap) -> None:
        """Restores a snapshot taken from an engine or batch slot of the same
        grid (af_restore)."""
        ptr, n = _raw_buffer(snap)
        keep = snap  # noqa: F841 -- the buffer must outlive the call
        _check(load_library().af_restore(self._handle(), C.c_void_p(ptr), n))

    def EnableDirectSolver(self) -> None:
        """Direct pressure solve on the GPU (the pinned Neumann system of the
        reference's Darwin EnableSparseSolver, sparse_darwin.go:19-172),
        used by every later projection.  No counterpart in the Linux
        reference build."""
        _check(load_library().af_enable_direct_solver(self._handle()))

    def EnableDirectSolverMethod(self, method: int) -> None:
        """EnableDirectSolver with the transform chosen: DIRECT_AUTO,
        DIRECT_FFT (power-of-two extents 8..512) or DIRECT_DENSE."""
        _check(load_library().af_enable_direct_solver_method(self._handle(), method))

    def direct_solver_method(self) -> int:
        """0 (red-black solve), DIRECT_FFT or DIRECT_DENSE."""
        on = C.c_int32()
        _check(load_library().af_direct_solver_enabled(self._handle(), C.byref(on)))
        return int(on.value)

    def direct_solver_enabled(self) -> bool:
        return self.direct_solver_method() != 0

    def Close(self) -> None:
        """Close, aeroflux.go:419-439; idempotent."""
        if self._h is not None:
            lib = load_library()
            rc = lib.af_close(self._h)
            lib.af_free(self._h)
            self._h = None
            _check(rc)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.Close()

    def __del__(self):
        try:
            if self._h is not None:
                load_library().af_free(self._h)
                self._h = None
        except Exception:
            pass

    # -- unexported methods the reference's in-package tests call ------------
    def apply_velocity_bc(self) -> None:  # stencil.go:246-315
        _check(load_library().af_apply_velocity_bc(self._handle()))

    def sweep_axis(self, axis: int, dt: float) -> None:  # stencil.go:21-25
        _check(load_library().af_sweep_axis(self._handle(), axis, dt))

    def commit_velocity(self) -> None:  # stencil.go:230-235
        _check(load_library().af_commit_velocity(self._handle()))

    def compute_divergence(self) -> None:  # project.go:42-62
        _check(load_library().af_compute_divergence(self._handle()))

    def gs_sweep(self, color: int) -> None:  # project.go:103-130
        _check(load_library().af_gs_sweep(self._handle(), color))

    def apply_pressure_bc(self) -> None:  # project.go:139-163
        _check(load_library().af_apply_pressure_bc(self._handle()))

    def solve_pressure(self) -> None:  # project.go:73-100
        _check(load_library().af_solve_pressure(self._handle()))

    def subtract_gradient(self) -> None:  # project.go:172-205
        _check(load_library().af_subtract_gradient(self._handle()))

    def project(self) -> None:  # project.go:12-34
        _check(load_library().af_project(self._handle()))

    # -- measurement ----------------------------------------------------------
    def time_kernel(self, kernel: int, reps: int) -> float:
        out = C.c_double()
        _check(load_library().af_time_kernel(self._handle(), kernel, reps, C.byref(out)))
        return float(out.value)

    def solver_tile(self) -> tuple[int, int, int]:
        k, r, z = C.c_int32(), C.c_int32(), C.c_int32()
        _check(load_library().af_solver_tile(self._handle(), C.byref(k), C.byref(r), C.byref(z)))
        return int(k.value), int(r.value), int(z.value)

    def sweep_shape(self) -> tuple[int, int, int]:
        s, r, z = C.c_int32(), C.c_int32(), C.c_int32()
        _check(load_library().af_sweep_shape(self._handle(), C.byref(s), C.byref(r), C.byref(z)))
        return int(s.value), int(r.value), int(z.value)

    def sweep_slab(self) -> int:
        r = C.c_int32()
        _check(load_library().af_sweep_slab(self._handle(), C.byref(r)))
        return int(r.value)

    def pass_slab(self) -> int:
        r = C.c_int32()
        _check(load_library().af_pass_slab(self._handle(), C.byref(r)))
        return int(r.value)

    def mixed_stats(self) -> tuple[float, int]:
        d = C.c_double()
        n = C.c_int32()
        _check(load_library().af_mixed_stats(self._handle(), C.byref(d), C.byref(n)))
        return float(d.value), int(n.value)

    def init_band_limited(self, seed: int, kmax: int = 8, slope: float = -5.0 / 3.0,
                          rms: float = 1.0) -> None:
        """Seeded band-limited velocity built on the GPU, velocity BC applied
        (the field of aeroflux_amd.ic.band_limited, up to summation order)."""
        _check(load_library().af_init_band_limited(self._handle(), seed, kmax, slope, rms))


def New(cfg: Config) -> Engine:
    """New, aeroflux.go:176-213.  ``cfg`` is normalised in a copy."""
    lib = load_library()
    c = cfg._to_c()
    h = C.c_void_p()
    _check(lib.af_new(C.byref(c), C.byref(h)))
    out = _CConfig()
    lib.af_get_config(h, C.byref(out))
    ncfg = Config()
    ncfg._from_c(out)
    return Engine(h.value, ncfg)


# ------------------------------------------------------------------ batching


@dataclass
class BatchConfig:
    """Batch of engines stepped together (no reference counterpart; each slot
    steps as Step, aeroflux.go:348-409, does).  Zero fields take defaults:
    1024 slots, a 1 GiB pool, the Config.Normalize solve schedule, tuned
    sweep shape and workgroup counts, the pressure solve run in groups of
    slots of about 224 MiB of solve buffers each (``SolveGroupBytes``; -1:
    one group)."""
    Precision: int = PRECISION_FP64
    MaxSlots: int = 0
    MemoryBudget: int = 0
    GaussSeidelIters: int = 0
    MixedMaxCycles: int = 0
    MixedSweeps: int = 0
    MixedFixSweeps: int = 0
    MixedChecked: int = 0
    SweepShape: int = 0
    Workgroups: int = 0
    SweepStaged: int = 0
    ServePolicy: int = SERVE_FCFS
    PreemptRatio: float = 0.0
    SnapshotBudget: int = 0
    StepCellBudget: int = 0
    SolveGroupBytes: int = 0
    StepMaxWait: int = 0

    def _to_c(self) -> _CBatchConfig:
        c = _CBatchConfig()
        for f in dc_fields(self):
            v = getattr(self, f.name)
            setattr(c, f.name, float(v) if f.name == "PreemptRatio" else int(v))
        return c




def MeasureStream(kind: int, n: int, reps: int = 20) -> float:
    """Attainable streaming bandwidth in GB/s of one access pattern over
    arrays of n doubles (af_measure_stream)."""
    out = C.c_double()
    _check(load_library().af_measure_stream(kind, n, reps, C.byref(out)))
    return float(out.value)


def SlotBytes(cfg: Config, precision: int) -> int:
    """HBM bytes one batch slot of ``cfg`` takes (0 if cfg is invalid)."""
    c = cfg._to_c()
    return int(load_library().af_batch_slot_bytes(C.byref(c), precision))


class Batch:
    """Many engines of uneven sizes, resident in one HBM pool and stepped
    together: one persistent launch per stage over every slot's tiles, the
    step replayed as one hipGraph.  Each slot's results are bit-identical to
    the same engine stepped alone."""

    def __init__(self, cfg: BatchConfig | None = None):
        lib = load_library()
        c = (cfg or BatchConfig())._to_c()
        h = C.c_void_p()
        _check(lib.af_batch_new(C.byref(c), C.byref(h)))
        self._h = h.value
        self._cfgs: dict[int, Config] = {}

    def _handle(self) -> int:
        if self._h is None:
            raise AeroFluxError(AF_ERR_CLOSED, "aeroflux: batch is closed")
        return self._h

    def Add(self, cfg: Config) -> int:
        """Adds an engine (New for one slot); returns its slot id."""
        lib = load_library()
        c = cfg._to_c()
        _check(lib.af_config_normalize(C.byref(c)))
        slot = C.c_int32()
        _check(lib.af_batch_add(self._handle(), C.byref(c), C.byref(slot)))
        n = Config()
        n._from_c(c)
        self._cfgs[int(slot.value)] = n
        return int(slot.value)

    def Remove(self, slot: int) -> None:
        _check(load_library().af_batch_remove(self._handle(), slot))
        self._cfgs.pop(slot, None)

    def Cfg(self, slot: int) -> Config:
        return self._cfgs[slot]

    def StepN(self, n: int) -> list[int]:
        """n batched steps; raises AeroFluxError (AF_ERR_FAULT) naming the
        faulted slots, as Step does for one engine."""
        cap = max(1, len(self._cfgs))
        out = (C.c_int32 * cap)()
        nf = C.c_int32()
        rc = load_library().af_batch_step_n(self._handle(), n, out, cap, C.byref(nf))
        if rc == AF_ERR_FAULT:
            err = AeroFluxError(rc, load_library().af_last_error().decode())
            err.slots = [int(out[i]) for i in range(min(nf.value, cap))]
            raise err
        _check(rc)
        return []

    def Step(self) -> None:
        self.StepN(1)

    def field(self, slot: int, fid: int) -> Field:
        f = _CField()
        _check(load_library().af_batch_field_desc(self._handle(), slot, fid, C.byref(f)))
        return Field(fid, f.Off, f.NX, f.NY, f.NZ, f.SY, f.SZ, f.Count)

    def Data(self, slot: int, fid: int) -> np.ndarray:
        f = self.field(slot, fid)
        out = np.empty(f.shape, dtype=np.float64)
        _check(load_library().af_batch_field_read(self._handle(), slot, fid,
                                                  out.ctypes.data_as(C.c_void_p), f.Count))
        return out

    def SetData(self, slot: int, fid: int, data: np.ndarray) -> None:
        f = self.field(slot, fid)
        arr = np.ascontiguousarray(data, dtype=np.float64).reshape(-1)
        if arr.size != f.Count:
            raise AeroFluxError(AF_ERR_ARG, f"aeroflux: field has {f.Count} elements, "
                                            f"got {arr.size}")
        _check(load_library().af_batch_field_write(self._handle(), slot, fid,
                                                   arr.ctypes.data_as(C.c_void_p), f.Count))

    def apply_velocity_bc(self, slot: int) -> None:  # stencil.go:246-315
        _check(load_library().af_batch_apply_velocity_bc(self._handle(), slot))

    def _scalar(self, fn, slot: int) -> float:
        out = C.c_double()
        _check(fn(self._handle(), slot, C.byref(out)))
        return float(out.value)

    def MaxDivergence(self, slot: int) -> float:  # project.go:210-232
        return self._scalar(load_library().af_batch_max_divergence, slot)

    def KineticEnergy(self, slot: int) -> float:  # project.go:238-258
        return self._scalar(load_library().af_batch_kinetic_energy, slot)

    def mixed_stats(self, slot: int) -> tuple[float, int]:
        d = C.c_double()
        n = C.c_int32()
        _check(load_library().af_batch_mixed_stats(self._handle(), slot, C.byref(d),
                                                   C.byref(n)))
        return float(d.value), int(n.value)

    def init_band_limited(self, slot: int, seed: int, kmax: int = 8,
                          slope: float = -5.0 / 3.0, rms: float = 1.0) -> None:
        _check(load_library().af_batch_init_band_limited(self._handle(), slot, seed, kmax, slope,
                                                         rms))

    def snapshot_bytes(self, slot: int) -> int:
        n = C.c_int64()
        _check(load_library().af_batch_snapshot_bytes(self._handle(), slot, C.byref(n)))
        return int(n.value)

    def Snapshot(self, slot: int, out=None):
        """Snapshot of one slot, in the engine's format (af_batch_snapshot)."""
        if out is None:
            out = np.empty(self.snapshot_bytes(slot), dtype=np.uint8)
        ptr, n = _raw_buffer(out)
        _check(load_library().af_batch_snapshot(self._handle(), slot, C.c_void_p(ptr), n))
        return out

    def Restore(self, slot: int, snap) -> None:
        ptr, n = _raw_buffer(snap)
        keep = snap  # noqa: F841 -- the buffer must outlive the call
        _check(load_library().af_batch_restore(self._handle(), slot, C.c_void_p(ptr), n))

    def Tune(self) -> None:
        _check(load_library().af_batch_tune(self._handle()))

    def time_stage(self, stage: int, reps: int) -> float:
        out = C.c_double()
        _check(load_library().af_batch_time_stage(self._handle(), stage, reps, C.byref(out)))
        return float(out.value)

    def info(self) -> dict:
        i = _CBatchInfo()
        _check(load_library().af_batch_get_info(self._handle(), C.byref(i)))
        return {name: getattr(i, name) for name, _ in _CBatchInfo._fields_}

    def Close(self) -> None:
        if self._h is not None:
            lib = load_library()
            rc = lib.af_batch_close(self._h)
            lib.af_batch_free(self._h)
            self._h = None
            self._cfgs.clear()
            _check(rc)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.Close()

    def __del__(self):
        try:
            if self._h is not None:
                load_library().af_batch_free(self._h)
                self._h = None
        except Exception:
            pass


def Serve(batch: Batch, jobs: list[dict], kmax: int = 8, slope: float = -5.0 / 3.0,
          rms: float = 1.0, realtime: bool = True, progress=None,
          progress_interval: float = 30.0) -> tuple[list[dict], dict]:
    """Runs step-N jobs through an empty batch with continuous batching
    (af_serve).  Each job is a dict with NX, NY, NZ, Steps, Viscosity, DT,
    H, Seed, Arrival (seconds; sorted) and, for a checked mixed batch,
    MixedTarget.  Returns per-job results (Admitted,
    Done, MaxDivergence, KineticEnergy, Status, Slot, FirstStep, DoneStep)
    and run statistics.
    progress(stats_so_far, jobs_done), if given, is called between batched
    steps at most once per progress_interval seconds."""
    n = len(jobs)
    arr = (_CServeJob * max(n, 1))()
    for i, j in enumerate(jobs):
        for name, _ in _CServeJob._fields_:
            setattr(arr[i], name, j.get(name, 0))
    res = (_CServeResult * max(n, 1))()
    st = _CServeStats()
    if progress is None:
        cb = C.cast(None, _ServeProgress)
    else:
        def report(p, done, _user):
            progress({name: getattr(p.contents, name) for name, _ in _CServeStats._fields_},
                     int(done))
        cb = _ServeProgress(report)
    _check(load_library().af_serve(batch._handle(), arr, n, kmax, slope, rms, int(realtime),
                                   res, C.byref(st), cb, float(progress_interval), None))
    out = [{name: getattr(res[i], name) for name, _ in _CServeResult._fields_} for i in range(n)]
    return out, {name: getattr(st, name) for name, _ in _CServeStats._fields_}
