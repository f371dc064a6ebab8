"""Ports of the reference's engine tests that the other files do not already
cover, each against the same configuration and property.

* TestFieldsAreArenaResident (engine_test.go:92-103): every field lies inside
  the one arena allocation.
* TestStepIsReproducibleAcrossRuns (engine_test.go:174-192): two runs give
  the same kinetic energy (here: the same bits in every field too).
* TestNoSlipWallsHoldNormalVelocity (engine_test.go:247-268): the normal
  velocity on the walls is exactly zero after steps.
* TestEngineCloseIdempotent (engine_test.go:290-298).
* TestStepZeroAllocs (engine_test.go:271-288), TestDirectSolveZeroAllocs
  (sparse_engine_test.go:87-...): the Go tests count heap allocations per
  Step; the device analogue is that a step allocates no device memory
  (free device memory is unchanged across steps after the first).
* TestPoissonSolveAgainstKnownSolution (poisson_test.go:56-100) checks the
  Darwin factorisation against an exact answer: choose x, form b = A x,
  solve, recover x.  Here the same is done for the GPU direct solve with the
  pinned 3-D system the reference's sparse solver assembles
  (sparse_darwin.go:41-95).
"""
import ctypes

import numpy as np
import pytest

import aeroflux_amd as afm
from aeroflux_amd import ic

from test_gpu_direct import pinned_matrix

pytestmark = pytest.mark.gpu


def vortex_engine(af, **cfg):
    e = af.New(af.Config(**cfg))
    f = ic.seed_vortex(cfg["NX"], cfg["NY"], cfg["NZ"], cfg.get("Ghost", 1), 1.0)
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        e.Mesh.SetData(e.Mesh.field(fid), f[name])
    return e


def test_fields_are_arena_resident(af):
    e = af.New(af.Config(NX=8, NY=8, NZ=8))
    nbytes = e.Mesh.Bytes
    assert nbytes == afm.MACBytes(8, 8, 8, 1)  # MACBytes, grid.go:138-144
    base = e.Mesh.DevicePtr(e.Mesh.field(afm.FIELD_P)) - e.Mesh.field(afm.FIELD_P).Off
    for fid in range(8):
        f = e.Mesh.field(fid)
        p = e.Mesh.DevicePtr(f)
        assert base <= p and p + f.Count * 8 <= base + nbytes, afm.FIELD_NAMES[fid]
        assert 0 <= f.Off and f.Off + f.Count * 8 <= nbytes
    e.Close()


def test_step_is_reproducible_across_runs(af):
    def run():
        e = vortex_engine(af, NX=10, NY=10, NZ=10, H=0.1, DT=0.004, Workers=4,
                          GaussSeidelIters=20)
        for _ in range(8):
            e.Step()
        out = (e.KineticEnergy(), [e.Mesh.Data(e.Mesh.field(k)).tobytes() for k in range(8)])
        e.Close()
        return out
    a, b = run(), run()
    assert a[0] == b[0]
    assert a[1] == b[1]


@pytest.mark.parametrize("direct", [False, True])
def test_no_slip_walls_hold_normal_velocity(af, direct):
    e = vortex_engine(af, NX=8, NY=8, NZ=8, H=0.125, DT=0.005)
    if direct:
        e.EnableDirectSolver()
    for _ in range(3):
        e.Step()
    g = 1
    u = e.Mesh.Data(e.Mesh.field(afm.FIELD_U))
    v = e.Mesh.Data(e.Mesh.field(afm.FIELD_V))
    w = e.Mesh.Data(e.Mesh.field(afm.FIELD_W))
    # arrays are (z, y, x); the wall faces are x = g and x = g + NX for u, etc.
    assert not u[g:g + 8, g:g + 8, g].any() and not u[g:g + 8, g:g + 8, g + 8].any()
    assert not v[g:g + 8, g, g:g + 8].any() and not v[g:g + 8, g + 8, g:g + 8].any()
    assert not w[g, g:g + 8, g:g + 8].any() and not w[g + 8, g:g + 8, g:g + 8].any()
    e.Close()


def test_engine_close_idempotent(af):
    e = af.New(af.Config(NX=8, NY=8, NZ=8))
    e.Close()
    e.Close()


def free_device_bytes():
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("direct", [False, True])
def test_step_allocates_no_device_memory(af, direct):
    e = vortex_engine(af, NX=8, NY=8, NZ=8, H=0.125, DT=0.005, Workers=2, GaussSeidelIters=4)
    if direct:
        e.EnableDirectSolver()
    e.Step()  # the first step captures the step graph
    before = free_device_bytes()
    for _ in range(20):
        e.Step()
    after = free_device_bytes()
    e.Close()
    assert after == before


@pytest.mark.parametrize("shp", [(16, 16, 16, 1), (32, 8, 16, 2), (12, 10, 8, 1)])
def test_direct_solve_recovers_a_known_solution(af, shp):
    """Choose p, form b = A p with the pinned matrix, hand the GPU -b as Div
    (the system is A p = -Div), and recover p.  Extents that are powers of
    two take the FFT passes, the others the dense transforms."""
    nx, ny, nz, g = shp
    h = 1.0 / max(nx, ny, nz)
    x, y, z = np.meshgrid((np.arange(nx) + 0.5) / nx, (np.arange(ny) + 0.5) / ny,
                          (np.arange(nz) + 0.5) / nz, indexing="ij")
    p_true = (np.cos(np.pi * x) * np.cos(2 * np.pi * y) + 0.3 * np.sin(3 * z) + 2.0).ravel(order="F")
    b = pinned_matrix(nx, ny, nz, h) @ p_true
    e = af.New(af.Config(NX=nx, NY=ny, NZ=nz, Ghost=g, H=h, DT=0.01, Viscosity=1e-6))
    e.EnableDirectSolver()
    div = np.zeros(e.Mesh.field(afm.FIELD_DIV).shape)
    div[g:g + nz, g:g + ny, g:g + nx] = -b.reshape(nz, ny, nx)
    e.Mesh.SetData(e.Mesh.field(afm.FIELD_DIV), div)
    e.solve_pressure()
    p = e.Mesh.Data(e.Mesh.field(afm.FIELD_P))[g:g + nz, g:g + ny, g:g + nx].ravel()
    e.Close()
    assert np.abs(p - p_true).max() <= 1e-9 * np.abs(p_true).max()
