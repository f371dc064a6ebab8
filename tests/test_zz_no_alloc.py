"""af_max_divergence_noalloc against af_max_divergence, through the C ABI.

The allocation-free form takes the device addresses of an engine's U, V and
W and writes its result to device memory; it must give the same bits as the
engine's own MaxDivergence (itself bit-identical to the oracle's,
project.go:210-232), on two calls in a row, the second on a smaller grid,
with a NaN face (skipped, as the reference skips it) in the second.
"""
import ctypes

import numpy as np
import pytest

import aeroflux_amd as afm
from aeroflux_amd import ic

pytestmark = pytest.mark.gpu


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def _noalloc(lib, hip, e, out_dev):
    m = e.Mesh
    rc = lib.af_max_divergence_noalloc(m.DevicePtr(m.U), m.DevicePtr(m.V), m.DevicePtr(m.W),
                                       m.NX, m.NY, m.NZ, m.Ghost, m.H, out_dev, 0)
    if rc != 0:
        raise afm.AeroFluxError(rc, lib.af_last_error().decode())
    assert hip.hipDeviceSynchronize() == 0
    got = ctypes.c_double()
    assert hip.hipMemcpy(ctypes.byref(got), ctypes.c_void_p(out_dev), 8, 2) == 0  # D2H
    return got.value


def _engine(af, n3, seed):
    nx, ny, nz = n3
    h = 1.0 / max(n3)
    e = af.New(af.Config(NX=nx, NY=ny, NZ=nz, H=h, DT=0.002, Viscosity=1e-4))
    f = ic.taylor_green(nx, ny, nz, 1, h, seed=seed, noise=0.05)
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        e.Mesh.SetData(e.Mesh.field(fid), f[name])
    e.apply_velocity_bc()
    e.StepN(2)
    return e


def test_noalloc_max_divergence_equals_engine(af):
    lib = af.load_library()
    hip = _hip()
    out = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(out), 8) == 0
    try:
        # first call: 96 x 40 x 24 after two steps
        e = _engine(af, (96, 40, 24), seed=4)
        want = e.MaxDivergence()
        got = _noalloc(lib, hip, e, out.value)
        assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64)
        assert got > 0
        e.Close()
        # second call, smaller grid, one NaN face: skipped by both
        e = _engine(af, (20, 12, 9), seed=5)
        u = e.Mesh.Data(e.Mesh.U)
        u[5, 6, 7] = np.nan
        e.Mesh.SetData(e.Mesh.U, u)
        want = e.MaxDivergence()
        got = _noalloc(lib, hip, e, out.value)
        assert np.isfinite(want)
        assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64)
        e.Close()
    finally:
        hip.hipFree(out)


def test_noalloc_rejects_grids_beyond_its_scratch(af):
    lib = af.load_library()
    # ceil(1100/64) * ceil(2000/4) = 18 * 500 partials > 8192: refused before
    # any device work (the addresses are never touched).
    rc = lib.af_max_divergence_noalloc(8, 8, 8, 1100, 2000, 1, 1, 1e-3, 8, 0)
    assert rc == afm.AF_ERR_ARG
    assert "reduction partials" in lib.af_last_error().decode()
    assert lib.af_max_divergence_noalloc(0, 8, 8, 4, 4, 4, 1, 0.25, 8, 0) == afm.AF_ERR_ARG
