"""Direct pressure solve (af_enable_direct_solver).

The reference's Darwin build assembles the interior 7-point Neumann
Laplacian with the first cell pinned and factors it (EnableSparseSolver,
sparse_darwin.go:19-125); solvePressureDirect (:127-172) solves it for -Div
and applies the pressure BC.  Here the same matrix is assembled in Python the
way the reference assembles it and solved with scipy's sparse LU; the GPU's
spectral solve must match it.  Properties: a projection with the direct
solve leaves the velocity divergence-free, and the step graph (replayed)
equals the same stages run one by one.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import aeroflux_amd as afm
from aeroflux_amd import ic

pytestmark = pytest.mark.gpu


def pinned_matrix(nx, ny, nz, h):
    """sparse_darwin.go:41-95: neighbour count on the diagonal (+1 for cell
    0), -1 for each neighbour, all divided by h*h; CellIndex = i + j*NX +
    k*NX*NY (grid.go:183-185)."""
    h2 = h * h
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    col = i + j * nx + k * nx * ny
    diag = ((i > 0).astype(float) + (i < nx - 1) + (j > 0) + (j < ny - 1) + (k > 0) + (k < nz - 1)
            + (col == 0))
    rows, cols, vals = [col], [col], [diag / h2]
    for di, dj, dk, ok in ((1, 0, 0, i < nx - 1), (0, 1, 0, j < ny - 1), (0, 0, 1, k < nz - 1)):
        a = col[ok]
        b = (i[ok] + di) + (j[ok] + dj) * nx + (k[ok] + dk) * nx * ny
        off = np.full(a.size, -1.0 / h2)
        rows += [a, b]
        cols += [b, a]
        vals += [off, off]
    n = nx * ny * nz
    return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


# Grids with a non-power-of-two extent take the dense transforms; the rest
# the FFT passes (every tile shape: extents 8 .. 512, ghost 1 and 2).
GRIDS = [(10, 12, 8, 1), (16, 16, 16, 1), (7, 5, 9, 2), (24, 20, 6, 1),
         (32, 8, 16, 2), (8, 64, 32, 1), (128, 16, 8, 1), (8, 8, 256, 1), (512, 8, 8, 1)]


def is_pow2_grid(shp):
    return all(n in (8, 16, 32, 64, 128, 256, 512) for n in shp[:3])


@pytest.mark.parametrize("shp", GRIDS)
def test_direct_solve_matches_pinned_system(af, shp):
    nx, ny, nz, g = shp
    h = 1.0 / max(nx, ny, nz)
    # The solve does not depend on DT or the viscosity; these keep every grid
    # inside Config.Normalize's diffusion stability limit.
    e = af.New(af.Config(NX=nx, NY=ny, NZ=nz, Ghost=g, H=h, DT=0.01, Viscosity=1e-6))
    e.EnableDirectSolver()
    assert e.direct_solver_enabled()
    assert e.direct_solver_method() == (afm.DIRECT_FFT if is_pow2_grid(shp) else afm.DIRECT_DENSE)
    rng = np.random.default_rng(nx * 100 + ny)
    div = np.zeros(e.Mesh.field(afm.FIELD_DIV).shape)
    interior = rng.standard_normal((nz, ny, nx)) * 50.0
    div[g:g + nz, g:g + ny, g:g + nx] = interior
    e.Mesh.SetData(e.Mesh.field(afm.FIELD_DIV), div)
    e.solve_pressure()
    p = e.Mesh.Data(e.Mesh.field(afm.FIELD_P))
    e.Close()
    ref = spl.spsolve(pinned_matrix(nx, ny, nz, h), -interior.ravel())
    got = p[g:g + nz, g:g + ny, g:g + nx].ravel()
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-10 * scale, (np.abs(got - ref).max(), scale)
    # applyPressureBC (project.go:139-163): first ghost layer mirrors the
    # interior cell with clamped coordinates.
    pin = p[g:g + nz, g:g + ny, g:g + nx]
    assert np.array_equal(p[g:g + nz, g:g + ny, g - 1], pin[:, :, 0])
    assert np.array_equal(p[g:g + nz, g:g + ny, g + nx], pin[:, :, -1])
    assert np.array_equal(p[g - 1, g:g + ny, g:g + nx], pin[0])
    assert np.array_equal(p[g:g + nz, g + ny, g:g + nx], pin[:, -1, :])


@pytest.mark.parametrize("method", [afm.DIRECT_FFT, afm.DIRECT_DENSE])
def test_direct_projection_is_divergence_free(af, method):
    n = 32
    cfg = dict(NX=n, NY=n, NZ=n, H=1.0 / n, DT=0.002, Viscosity=1e-4)
    f = ic.taylor_green(n, n, n, 1, 1.0 / n, seed=4, amp=2.0, noise=0.3)
    e = af.New(af.Config(**cfg))
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        e.Mesh.SetData(e.Mesh.field(fid), f[name])
    e.apply_velocity_bc()
    before = e.MaxDivergence()
    e.EnableDirectSolverMethod(method)
    assert e.direct_solver_method() == method
    e.project()
    after = e.MaxDivergence()
    e.Close()
    assert before > 1.0
    assert after <= 1e-9 * before, (before, after)


def stage_step(e, dt):
    """Step (aeroflux.go:348-409) through the stage calls, no graph."""
    for ax in (0, 1, 2):
        e.apply_velocity_bc()
        e.sweep_axis(ax, 0.5 * dt)
        e.commit_velocity()
    e.project()
    for ax in (2, 1, 0):
        e.apply_velocity_bc()
        e.sweep_axis(ax, 0.5 * dt)
        e.commit_velocity()
    e.project()


@pytest.mark.parametrize("shp", [(24, 20, 16), (32, 16, 64)])
def test_direct_step_graph_equals_stages(af, shp):
    """The direct solve inside the captured step graph, replayed step after
    step, gives the bits of the same stages enqueued one by one (dense
    transforms with their GEMMs in the graph, and the FFT passes)."""
    n, ny, nz = shp
    cfg = dict(NX=n, NY=ny, NZ=nz, H=1.0 / n, DT=0.004, Viscosity=1e-4)
    f = ic.taylor_green(n, ny, nz, 1, 1.0 / n, seed=6, amp=1.5, noise=0.1)
    out = []
    for graph in (True, False):
        e = af.New(af.Config(**cfg))
        for fid, name in ((1, "U"), (2, "V"), (3, "W")):
            e.Mesh.SetData(e.Mesh.field(fid), f[name])
        e.EnableDirectSolver()
        states = []
        for _ in range(3):
            if graph:
                e.Step()
            else:
                stage_step(e, cfg["DT"])
            states.append([e.Mesh.Data(e.Mesh.field(k)).tobytes() for k in range(8)])
        e.Close()
        out.append(states)
    for s, (a, b) in enumerate(zip(*out)):
        assert a == b, f"step {s}"


def test_sparse_solver_keeps_reference_linux_error(af):
    e = af.New(af.Config(NX=8, NY=8, NZ=8))
    with pytest.raises(af.AeroFluxError) as ei:
        e.EnableSparseSolver()
    assert "requires darwin" in str(ei.value)
    assert not e.direct_solver_enabled()
    e.Close()


def seeded(af, cfg):
    e = af.New(af.Config(**cfg))
    f = ic.seed_vortex(cfg["NX"], cfg["NY"], cfg["NZ"], 1, 1.0)
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        e.Mesh.SetData(e.Mesh.field(fid), f[name])
    return e


def test_direct_solver_matches_iterative(af):
    """TestDirectSolverMatchesIterative (sparse_engine_test.go:16-56): one
    projection of the seeded vortex at 12^3; the direct solve must remove
    divergence at least as well as 400 red-black sweeps (the README quotes
    8.7e-6 against 2.1e-12, README.md:422-427)."""
    base = dict(NX=12, NY=12, NZ=12, H=1.0 / 12, DT=0.01)
    it = seeded(af, dict(base, GaussSeidelIters=400))
    it.project()
    iter_div = it.MaxDivergence()
    it.Close()
    d = seeded(af, base)
    d.EnableDirectSolver()
    d.project()
    direct_div = d.MaxDivergence()
    d.Close()
    assert np.isfinite(direct_div)
    assert not (direct_div > iter_div * 2 and direct_div > 1e-8), (direct_div, iter_div)
    assert direct_div < 1e-11, direct_div


def test_direct_solver_steps_stably(af):
    """TestDirectSolverStepsStably (sparse_engine_test.go:60-84): 20 steps
    at 10^3 with the direct solve stay finite; the README reports
    max|div| = 2.5e-14 after them (README.md:686)."""
    e = seeded(af, dict(NX=10, NY=10, NZ=10, H=0.1, DT=0.005))
    e.EnableDirectSolver()
    for _ in range(20):
        e.Step()
    for fid in (afm.FIELD_U, afm.FIELD_V, afm.FIELD_W, afm.FIELD_P):
        assert np.isfinite(e.Mesh.Data(e.Mesh.field(fid))).all()
    div = e.MaxDivergence()
    e.Close()
    assert div < 1e-12, div


def random_div_solve(af, shp, method, seed):
    nx, ny, nz, g = shp
    h = 1.0 / max(nx, ny, nz)
    e = af.New(af.Config(NX=nx, NY=ny, NZ=nz, Ghost=g, H=h, DT=0.01, Viscosity=1e-6))
    e.EnableDirectSolverMethod(method)
    rng = np.random.default_rng(seed)
    div = np.zeros(e.Mesh.field(afm.FIELD_DIV).shape)
    div[g:g + nz, g:g + ny, g:g + nx] = rng.standard_normal((nz, ny, nx)) * 20.0
    e.Mesh.SetData(e.Mesh.field(afm.FIELD_DIV), div)
    e.solve_pressure()
    p = e.Mesh.Data(e.Mesh.field(afm.FIELD_P))
    e.Close()
    return p


@pytest.mark.parametrize("shp", [(64, 32, 16, 1), (16, 128, 8, 2), (256, 16, 32, 1)])
def test_fft_and_dense_transforms_agree(af, shp):
    """The FFT passes and the dense DCT GEMMs solve the same system; on the
    same right-hand side they agree to rounding, ghost layers included."""
    a = random_div_solve(af, shp, afm.DIRECT_FFT, 17)
    b = random_div_solve(af, shp, afm.DIRECT_DENSE, 17)
    scale = np.abs(b).max()
    assert np.abs(a - b).max() <= 1e-12 * scale, (np.abs(a - b).max(), scale)


def test_direct_method_switch_and_unsupported_grid(af):
    e = af.New(af.Config(NX=32, NY=16, NZ=8))
    e.EnableDirectSolverMethod(afm.DIRECT_DENSE)
    assert e.direct_solver_method() == afm.DIRECT_DENSE
    e.EnableDirectSolverMethod(afm.DIRECT_FFT)
    assert e.direct_solver_method() == afm.DIRECT_FFT
    e.EnableDirectSolver()                      # auto keeps the FFT passes here
    assert e.direct_solver_method() == afm.DIRECT_FFT
    e.Close()
    e = af.New(af.Config(NX=24, NY=16, NZ=8))
    with pytest.raises(af.AeroFluxError) as ei:
        e.EnableDirectSolverMethod(afm.DIRECT_FFT)
    assert ei.value.code == afm.AF_ERR_UNSUPPORTED
    assert e.direct_solver_method() == 0
    with pytest.raises(af.AeroFluxError):
        e.EnableDirectSolverMethod(3)
    e.EnableDirectSolver()
    assert e.direct_solver_method() == afm.DIRECT_DENSE
    e.Close()


@pytest.mark.parametrize("shp", [(32, 16, 64, 1), (16, 8, 32, 2)])
def test_fused_divergence_projection_equals_stages(af, shp):
    """With the FFT passes, project computes the divergence inside the first
    transform pass.  It must give the bits of the unfused stages
    (applyVelocityBC, computeDivergence, solvePressure, subtractGradient,
    applyVelocityBC; project.go:12-34), Div included."""
    nx, ny, nz, g = shp
    cfg = dict(NX=nx, NY=ny, NZ=nz, Ghost=g, H=1.0 / nx, DT=0.004, Viscosity=1e-4)
    f = ic.taylor_green(nx, ny, nz, g, 1.0 / nx, seed=8, amp=2.0, noise=0.2)
    out = []
    for fused in (True, False):
        e = af.New(af.Config(**cfg))
        for fid, name in ((1, "U"), (2, "V"), (3, "W")):
            e.Mesh.SetData(e.Mesh.field(fid), f[name])
        e.EnableDirectSolverMethod(afm.DIRECT_FFT)
        if fused:
            e.project()
        else:
            e.apply_velocity_bc()
            e.compute_divergence()
            e.solve_pressure()
            e.subtract_gradient()
            e.apply_velocity_bc()
        out.append([e.Mesh.Data(e.Mesh.field(k)).tobytes() for k in range(8)])
        e.Close()
    for k in range(8):
        assert out[0][k] == out[1][k], afm.FIELD_NAMES[k]
