"""Parity of the HIP path (through the C ABI) with the CPU oracle.

FP64 path: bit-for-bit on every stage and on whole steps (the bar for the
reference's unfused IEEE evaluation), signed zeros and ghost layers
included.  KineticEnergy sums in a fixed tree instead of the reference's
serial order: relative tolerance 1e-12.  Mixed precision: error bounds
stated per field in the test.
"""
import math

import numpy as np
import pytest

import aeroflux_amd as afm
from aeroflux_amd import ic

pytestmark = pytest.mark.gpu

NAMES = ("P", "U", "V", "W", "Div", "Tmp0", "Tmp1", "Tmp2")


def make_pair(af, oracle_lib, fields=None, **kw):
    kw.setdefault("Workers", 4)
    g = af.New(af.Config(**kw))
    o = oracle_lib.OracleEngine({k: v for k, v in kw.items() if not k.startswith("Tile")})
    if fields is not None:
        for fid, name in ((1, "U"), (2, "V"), (3, "W")):
            g.Mesh.SetData(g.Mesh.field(fid), fields[name])
            o.field(fid)[...] = fields[name]
    return g, o


def assert_bitwise(g, o, ids=range(8), what=""):
    for fid in ids:
        a = g.Mesh.Data(g.Mesh.field(fid))
        b = o.field(fid)
        assert a.shape == b.shape
        if a.tobytes() != b.tobytes():
            bad = np.argwhere(a.view(np.uint64) != b.view(np.uint64))
            z, y, x = bad[0]
            raise AssertionError(
                f"{what} field {NAMES[fid]}: {len(bad)} mismatches, first at (x={x},y={y},z={z}) "
                f"gpu={a[z, y, x]!r} oracle={b[z, y, x]!r}")


def random_fields(shape_cfg, seed, scale=1.0, ghosts=True):
    nx, ny, nz, g = shape_cfg
    rng = np.random.default_rng(seed)
    out = {}
    for name, shp in ic.shapes(nx, ny, nz, g).items():
        a = rng.standard_normal(shp) * scale
        a[rng.random(shp) < 0.02] = 0.0
        a[rng.random(shp) < 0.02] = -0.0
        if not ghosts:
            a2 = np.zeros(shp)
            xr, yr, zr = ic.face_ranges(nx, ny, nz, g)[name]
            a2[zr.start:zr.stop, yr.start:yr.stop, xr.start:xr.stop] = \
                a[zr.start:zr.stop, yr.start:yr.stop, xr.start:xr.stop]
            a = a2
        out[name] = a
    return out


SHAPES = [(7, 5, 6, 1), (5, 4, 3, 2), (16, 9, 11, 1), (66, 5, 4, 1)]


@pytest.mark.parametrize("shp", SHAPES)
def test_velocity_bc_bitwise(af, oracle_lib, shp):
    nx, ny, nz, g = shp
    f = random_fields(shp, 1)
    ge, oe = make_pair(af, oracle_lib, f, NX=nx, NY=ny, NZ=nz, Ghost=g)
    ge.apply_velocity_bc()
    oe.apply_velocity_bc()
    assert_bitwise(ge, oe, (1, 2, 3), "velocity bc")
    ge.Close()


@pytest.mark.parametrize("shp", SHAPES)
@pytest.mark.parametrize("ax", [0, 1, 2])
@pytest.mark.parametrize("nu", [0.0, 3e-3])
def test_sweep_bitwise(af, oracle_lib, shp, ax, nu):
    nx, ny, nz, g = shp
    h = 1.0 / max(nx, ny, nz)
    # velocities of several cells per half step: exercises the clamps
    f = random_fields(shp, 2 + ax, scale=2.0)
    dt = 0.5 * h
    ge, oe = make_pair(af, oracle_lib, f, NX=nx, NY=ny, NZ=nz, Ghost=g, H=h, DT=dt,
                       Viscosity=nu if nu == 0 else 0.1 * h * h / dt)
    ge.sweep_axis(ax, dt * 0.5)
    assert oe.sweep_axis(ax, dt * 0.5) == 0
    assert_bitwise(ge, oe, (5, 6, 7), f"sweep ax={ax}")
    ge.Close()


# The one-pass solve layouts TileK picks between by timing: colour-split and
# arena.  Tests that check the solve pin each one.
LAYOUTS = {"split": -2, "arena": -1}


@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("shp", SHAPES)
def test_projection_stages_bitwise(af, oracle_lib, shp, layout):
    nx, ny, nz, g = shp
    f = random_fields(shp, 5)
    kw = dict(NX=nx, NY=ny, NZ=nz, Ghost=g, H=0.37, DT=0.021, Density=1.3, GaussSeidelIters=7,
              TileK=LAYOUTS[layout])
    ge, oe = make_pair(af, oracle_lib, f, **kw)
    rng = np.random.default_rng(11)
    p = rng.standard_normal(oe.field(0).shape)
    ge.Mesh.SetData(ge.Mesh.P, p)
    oe.field(0)[...] = p
    ge.compute_divergence(); oe.compute_divergence()
    assert_bitwise(ge, oe, (4,), "divergence")
    for col in (0, 1, 0):  # non-fused colour passes read the ghosts as stored
        ge.gs_sweep(col); oe.gs_sweep(col)
        assert_bitwise(ge, oe, (0,), f"gs colour {col}")
    ge.apply_pressure_bc(); oe.apply_pressure_bc()
    assert_bitwise(ge, oe, (0,), "pressure bc")
    ge.subtract_gradient(); oe.subtract_gradient()
    assert_bitwise(ge, oe, (1, 2, 3), "gradient")
    # production solve: zeroing and per-pass BCs fused, final fill materialised
    ge.Mesh.SetData(ge.Mesh.P, p)
    oe.field(0)[...] = p
    ge.solve_pressure(); oe.solve_pressure()
    assert_bitwise(ge, oe, (0,), "solve")
    ge.project(); oe.project()
    assert_bitwise(ge, oe, range(8), "project")
    ge.Close()


STEP_CASES = [
    # name, config, initial condition, steps
    ("vortex12_inviscid", dict(NX=12, NY=12, NZ=12, H=1 / 12, DT=0.005, Viscosity=0.0,
                               GaussSeidelIters=30), "vortex", 5),
    ("tg_16x12x10", dict(NX=16, NY=12, NZ=10, H=1 / 16, DT=0.004, Viscosity=1e-4,
                         GaussSeidelIters=24), "tg", 6),
    ("ghost2", dict(NX=9, NY=7, NZ=8, Ghost=2, H=1 / 9, DT=0.01, Viscosity=1e-4,
                    GaussSeidelIters=10), "tg", 4),
    ("fast_flow_clamps", dict(NX=10, NY=10, NZ=10, H=0.1, DT=0.05, Viscosity=1e-4,
                              GaussSeidelIters=12), "tg_strong", 4),
]


def initial(kind, cfg):
    nx, ny, nz, g, h = cfg["NX"], cfg["NY"], cfg["NZ"], cfg.get("Ghost", 1), cfg["H"]
    if kind == "vortex":
        return ic.seed_vortex(nx, ny, nz, g, 1.0)
    if kind == "tg_strong":
        return ic.taylor_green(nx, ny, nz, g, h, seed=9, amp=8.0, noise=0.1)
    return ic.taylor_green(nx, ny, nz, g, h, seed=4)


@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("case", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_step_bitwise(af, oracle_lib, case, layout):
    name, cfg, kind, steps = case
    ge, oe = make_pair(af, oracle_lib, initial(kind, cfg), **cfg, TileK=LAYOUTS[layout])
    ge.apply_velocity_bc(); oe.apply_velocity_bc()
    for s in range(steps):
        ge.Step()
        assert oe.step() == 0
        assert_bitwise(ge, oe, range(8), f"{name} step {s}")
    assert ge.MaxDivergence() == oe.max_divergence()
    assert ge.MaxVelocity() == oe.max_velocity()
    assert ge.CFL() == oe.cfl()
    ke_g, ke_o = ge.KineticEnergy(), oe.kinetic_energy()
    assert abs(ke_g - ke_o) <= 1e-12 * abs(ke_o)
    ge.Close()


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_config0_64cubed_100_steps_bitwise(af, oracle_lib, layout):
    """The reference CPU case: 64^3, H=1/64, DT=0.004, nu=1e-5, seeded
    Taylor-Green + 1e-3 noise, default 24 sweeps, 100 Step() calls."""
    cfg = dict(NX=64, NY=64, NZ=64, H=1 / 64, DT=0.004, Viscosity=1e-5, Workers=16,
               TileK=LAYOUTS[layout])
    f = ic.taylor_green(64, 64, 64, 1, 1 / 64, seed=1)
    ge, oe = make_pair(af, oracle_lib, f, **cfg)
    ge.apply_velocity_bc(); oe.apply_velocity_bc()
    for s in range(100):
        ge.Step()
        assert oe.step() == 0
        if s % 25 == 24:
            assert_bitwise(ge, oe, range(8), f"64^3 step {s}")
    assert ge.MaxDivergence() == oe.max_divergence()
    ge.Close()


def test_step_n_equals_repeated_step_and_replay_is_repeatable(af):
    cfg = dict(NX=24, NY=20, NZ=16, H=1 / 24, DT=0.004, Viscosity=1e-5)
    f = ic.taylor_green(24, 20, 16, 1, 1 / 24, seed=5)
    outs = []
    for mode in ("step", "stepn", "step"):
        e = af.New(af.Config(**cfg))
        for fid, name in ((1, "U"), (2, "V"), (3, "W")):
            e.Mesh.SetData(e.Mesh.field(fid), f[name])
        if mode == "step":
            for _ in range(7):
                e.Step()
        else:
            e.StepN(7)
        outs.append([e.Mesh.Data(e.Mesh.field(i)).tobytes() for i in range(8)])
        e.Close()
    assert outs[0] == outs[1] == outs[2]


def test_descriptor_swap_state_is_honoured(af, oracle_lib):
    """After an odd number of commitVelocity swaps the step graph must run on
    the swapped descriptors (a second graph is captured)."""
    cfg = dict(NX=8, NY=6, NZ=7, H=1 / 8, DT=0.01, Viscosity=1e-4, GaussSeidelIters=6)
    f = ic.taylor_green(8, 6, 7, 1, 1 / 8, seed=6)
    ge, oe = make_pair(af, oracle_lib, f, **cfg)
    ge.sweep_axis(1, 0.005); assert oe.sweep_axis(1, 0.005) == 0
    ge.commit_velocity(); oe.commit_velocity()
    assert ge.Mesh.U.Off == oe.desc(1)["Off"]
    for s in range(3):
        ge.Step(); assert oe.step() == 0
    assert_bitwise(ge, oe, range(8), "swapped")
    ge.Close()


def test_nan_departure_is_a_contained_fault(af):
    """The reference's guard-page/bounds-panic path: a NaN departure
    coordinate makes Step return an error; the engine survives and can be
    reseeded and stepped again.  The counterpart of
    TestGuardPageFaultIsContained (arena_test.go:51-78), whose deliberate
    write past the arena has no safe device form: here the fault is raised
    from the flag, and no access ever leaves the arena."""
    cfg = dict(NX=8, NY=8, NZ=8, H=1 / 8, DT=0.01)
    e = af.New(af.Config(**cfg))
    f = ic.taylor_green(8, 8, 8, 1, 1 / 8, seed=2)
    f["U"][4, 4, 4] = math.nan
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        e.Mesh.SetData(e.Mesh.field(fid), f[name])
    with pytest.raises(afm.AeroFluxError, match="^aeroflux: device fault contained") as ei:
        e.Step()
    assert ei.value.code == afm.AF_ERR_FAULT
    clean = ic.taylor_green(8, 8, 8, 1, 1 / 8, seed=2)
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        e.Mesh.SetData(e.Mesh.field(fid), clean[name])
    e.Step()
    assert math.isfinite(e.MaxDivergence())
    e.Close()
    e.Close()  # idempotent


def test_step_n_fault_reporting(af):
    """StepN(n) reads the fault flag once, after n steps; StepN(n,
    first_fault=True) reads it after every step and names the first faulting
    step.  Both leave the engine usable."""
    cfg = dict(NX=8, NY=8, NZ=8, H=1 / 8, DT=0.01)
    e = af.New(af.Config(**cfg))
    bad = ic.taylor_green(8, 8, 8, 1, 1 / 8, seed=2)
    bad["V"][3, 5, 2] = math.nan
    for first_fault in (False, True):
        for fid, name in ((1, "U"), (2, "V"), (3, "W")):
            e.Mesh.SetData(e.Mesh.field(fid), bad[name])
        with pytest.raises(afm.AeroFluxError, match="^aeroflux: device fault contained") as ei:
            e.StepN(3, first_fault=first_fault)
        assert ei.value.code == afm.AF_ERR_FAULT
        if first_fault:
            assert ei.value.step == 0
    clean = ic.taylor_green(8, 8, 8, 1, 1 / 8, seed=2)
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        e.Mesh.SetData(e.Mesh.field(fid), clean[name])
    e.StepN(3)
    e.StepN(2, first_fault=True)
    assert math.isfinite(e.MaxDivergence())
    e.Close()


def test_close_then_use_errors(af):
    e = af.New(af.Config(NX=4, NY=4, NZ=4))
    e.Close()
    with pytest.raises(afm.AeroFluxError):
        e.Step()


def test_sparse_solver_errors(af):
    e = af.New(af.Config(NX=4, NY=4, NZ=4))
    with pytest.raises(afm.AeroFluxError, match="sparse direct solver requires darwin"):
        e.EnableSparseSolver()
    e.Close()


def test_sparse_solver_flag_steps_like_red_black(af, oracle_lib):
    """UseSparseSolver is accepted by New and ignored, as in the reference's
    non-Darwin build (project.go:19, sparse_fallback.go:20-24): the engine
    steps with the red-black solve, bit for bit as the oracle does."""
    cfg = dict(NX=12, NY=10, NZ=9, H=1 / 12, DT=0.004, Viscosity=1e-4)
    f = ic.taylor_green(12, 10, 9, 1, 1 / 12, seed=21)
    ge, oe = make_pair(af, oracle_lib, f, **cfg, UseSparseSolver=True)
    assert ge.Cfg.UseSparseSolver
    ge.apply_velocity_bc(); oe.apply_velocity_bc()
    for s in range(3):
        ge.Step()
        assert oe.step() == 0
    assert_bitwise(ge, oe, range(8), "UseSparseSolver")
    assert not ge.direct_solver_enabled()
    ge.Close()
    oe.close()


def test_diagnostics_skip_nan_and_scan_ghosts(af, oracle_lib):
    """MaxDivergence skips NaN cells (project.go:226-229, `a > mx` is false
    for NaN) and MaxVelocity scans the whole padded arrays including ghosts
    (stencil.go:319-331): a NaN face and a maximum planted in a ghost corner
    give the oracle's values on the GPU."""
    cfg = dict(NX=20, NY=14, NZ=11, H=1 / 20, DT=0.004, Viscosity=1e-4)
    f = ic.taylor_green(20, 14, 11, 1, 1 / 20, seed=22, noise=0.05)
    f["U"][4, 5, 6] = np.nan
    f["V"][0, 0, 0] = -9.5
    ge, oe = make_pair(af, oracle_lib, f, **cfg)
    d = ge.MaxDivergence()
    assert np.isfinite(d) and d == oe.max_divergence()
    assert ge.MaxVelocity() == oe.max_velocity() == 9.5
    assert ge.CFL() == oe.cfl()
    ge.Close()
    oe.close()


def test_reference_example_on_gpu(af):
    """example_test.go:13-61 on the GPU path."""
    e = af.New(af.Config(NX=16, NY=16, NZ=16, H=1 / 16, DT=0.004, Viscosity=1e-5, Density=1.0,
                         GaussSeidelIters=400))
    f = ic.shear_layer(16, 16, 16, 1)
    e.Mesh.SetData(e.Mesh.U, f["U"])
    for _ in range(10):
        e.Step()
    assert not math.isnan(e.KineticEnergy())
    assert e.MaxDivergence() < 1e-3
    e.Close()


# ------------------------------------------------------------ mixed precision

def run_engine(af, cfg, f, steps):
    e = af.New(af.Config(**cfg))
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        e.Mesh.SetData(e.Mesh.field(fid), f[name])
    e.apply_velocity_bc()
    e.StepN(steps)
    out = {n: e.Mesh.Data(e.Mesh.field(i)) for i, n in enumerate(NAMES) if i < 4}
    d = e.MaxDivergence()
    stats = e.mixed_stats() if cfg.get("Precision") == afm.PRECISION_MIXED else None
    e.Close()
    return out, d, stats


@pytest.mark.parametrize("target_mode", ["one_cycle", "fp64_target"])
def test_mixed_precision_matches_fp64(af, target_mode):
    """FP32 smoothing + FP64 defect correction vs the FP64 path (itself
    bit-exact to the oracle).

    one_cycle (same 24 sweep pairs, FP32): differences are FP32 rounding
    only -- velocities within 1e-6 of max|u|, pressure within 1e-5 of
    max|p|, final max|div| within 1% of the FP64 path's.

    fp64_target (cycles until max|div| <= the FP64 path's final value):
    extra correction sweeps carry the iterate closer to the converged
    solution than the FP64 path's fixed 24 sweeps, so the difference is the
    FP64 path's own unconverged remainder -- velocities within 1e-4 of
    max|u|, pressure within 1e-3 of max|p| -- and the final max|div| must
    not exceed the FP64 path's."""
    n = 32
    base = dict(NX=n, NY=n, NZ=n, H=1 / n, DT=0.004, Viscosity=1e-5)
    f = ic.taylor_green(n, n, n, 1, 1 / n, seed=8)
    ref, d64, _ = run_engine(af, base, f, 5)
    mcfg = dict(base, Precision=afm.PRECISION_MIXED)
    tol_u, tol_p = 1e-6, 1e-5
    if target_mode == "fp64_target":
        mcfg["MixedTarget"] = d64
        tol_u, tol_p = 1e-4, 1e-3
    mix, dmx, stats = run_engine(af, mcfg, f, 5)
    vmax = max(np.abs(ref[k]).max() for k in "UVW")
    for k in "UVW":
        assert np.abs(mix[k] - ref[k]).max() <= tol_u * vmax, k
    assert np.abs(mix["P"] - ref["P"]).max() <= tol_p * np.abs(ref["P"]).max()
    if target_mode == "fp64_target":
        assert dmx <= d64 * (1 + 1e-9)
        assert stats[1] >= 1
    else:
        assert abs(dmx - d64) <= 0.01 * d64


# ------------------------------------------------------ full size (256^3)

def test_config1_256cubed_12_steps_bitwise_and_properties(af, oracle_lib):
    """256^3, H=1/256, DT=0.001, nu=1e-5, Taylor-Green + noise, 24 sweeps:
    12 steps bit for bit against the oracle (P, U, V, W after every step, the
    whole arena after the first and the last; the diagnostics at the end),
    then size-independent properties over further steps: finite, exact zero
    normal velocity on walls, divergence reduced."""
    n = 256
    cfg = dict(NX=n, NY=n, NZ=n, H=1 / n, DT=0.001, Viscosity=1e-5, Workers=16)
    f = ic.taylor_green(n, n, n, 1, 1 / n, seed=1)
    ge, oe = make_pair(af, oracle_lib, f, **cfg)
    ge.apply_velocity_bc(); oe.apply_velocity_bc()
    d0 = ge.MaxDivergence()
    steps = 12
    for s in range(steps):
        ge.Step()
        assert oe.step() == 0
        ids = range(8) if s in (0, steps - 1) else range(4)
        assert_bitwise(ge, oe, ids, f"256^3 step {s}")
    assert ge.MaxDivergence() == oe.max_divergence()
    assert ge.MaxVelocity() == oe.max_velocity()
    assert ge.CFL() == oe.cfl()
    assert abs(ge.KineticEnergy() - oe.kinetic_energy()) <= 1e-12 * oe.kinetic_energy()
    oe.close()
    # A projection from the step's end state leaves max|div| well below
    # the unprojected initial field's (24 sweep pairs do not converge at
    # 256^3, so it is not near zero, and it drifts over many steps).
    assert ge.MaxDivergence() < d0
    ge.StepN(8)
    u = ge.Mesh.Data(ge.Mesh.U)
    assert np.all(u[1:n + 1, 1:n + 1, 1] == 0) and np.all(u[1:n + 1, 1:n + 1, n + 1] == 0)
    for fid in range(4):
        assert np.all(np.isfinite(ge.Mesh.Data(ge.Mesh.field(fid))))
    ge.Close()


# ------------------------------------- red-black schedules and tile shapes
# (0,0,0): colour-split layout, one pass per launch (the default);
# (-1,0,0): arena layout, one pass per launch; (k,rows,z): k passes per
# launch, temporally blocked.

TILES = [(0, 0, 0), (-1, 0, 0), (2, 1, 1), (2, 2, 3), (4, 1, 8), (4, 2, 16), (6, 1, 5),
         (6, 2, 32)]
TILE_SHAPES = [(37, 29, 23, 1), (130, 9, 7, 1), (10, 70, 9, 2)]


@pytest.mark.parametrize("tile", TILES, ids=[f"k{t[0]}r{t[1]}z{t[2]}" for t in TILES])
@pytest.mark.parametrize("shp", TILE_SHAPES, ids=["37x29x23", "130x9x7", "10x70x9g2"])
def test_blocked_solve_bitwise_across_tiles(af, oracle_lib, tile, shp):
    """Every tile shape (passes per launch, rows per thread, z-chunk) gives
    the reference's solve bit for bit -- the GPU form of the reference's
    worker-count determinism test (engine_test.go:134-170)."""
    nx, ny, nz, g = shp
    k, rows, cz = tile
    # k = 0 in the table is the colour-split layout (TileK = -2 pins it; 0
    # would let the timing choose).
    kw = dict(NX=nx, NY=ny, NZ=nz, Ghost=g, H=0.41, DT=0.017, GaussSeidelIters=5,
              TileK=k if k != 0 else -2, TileRows=rows, TileZ=cz)
    ge = af.New(af.Config(**kw))
    assert ge.solver_tile()[0] == k
    oe = oracle_lib.OracleEngine({k2: v for k2, v in kw.items() if not k2.startswith("Tile")})
    rng = np.random.default_rng(13)
    d = rng.standard_normal(oe.field(4).shape)
    p = rng.standard_normal(oe.field(0).shape)
    for fid, arr in ((4, d), (0, p)):
        ge.Mesh.SetData(ge.Mesh.field(fid), arr)
        oe.field(fid)[...] = arr
    ge.solve_pressure(); oe.solve_pressure()
    assert_bitwise(ge, oe, (0, 4), f"blocked solve {tile}")
    ge.Close()


@pytest.mark.parametrize("tile", TILES[1:], ids=[f"k{t[0]}r{t[1]}z{t[2]}" for t in TILES[1:]])
def test_step_bitwise_across_tiles(af, oracle_lib, tile):
    k, rows, cz = tile
    cfg = dict(NX=40, NY=36, NZ=20, H=1 / 40, DT=0.004, Viscosity=1e-4, GaussSeidelIters=9)
    f = ic.taylor_green(40, 36, 20, 1, 1 / 40, seed=12)
    ge, oe = make_pair(af, oracle_lib, f, **cfg)
    ge.Close()
    ge = af.New(af.Config(**cfg, TileK=k, TileRows=rows, TileZ=cz))
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        ge.Mesh.SetData(ge.Mesh.field(fid), f[name])
    for s in range(3):
        ge.Step(); assert oe.step() == 0
    assert_bitwise(ge, oe, range(8), f"step tile {tile}")
    ge.Close()


def test_schedule_is_reported_and_blocked_tiles_are_tuned(af):
    e = af.New(af.Config(NX=48, NY=40, NZ=32))
    assert e.solver_tile()[0] in (0, -1)  # timed: colour-split or arena layout
    e.Close()
    for k, want in ((-2, 0), (-1, -1)):
        e = af.New(af.Config(NX=48, NY=40, NZ=32, TileK=k))
        assert e.solver_tile()[0] == want
        e.Close()
    with pytest.raises(af.AeroFluxError):
        af.New(af.Config(NX=48, NY=40, NZ=32, TileK=-3))
    e = af.New(af.Config(NX=48, NY=40, NZ=32, TileK=4))
    k, rows, cz = e.solver_tile()
    assert k == 4 and rows in (1, 2) and cz in (8, 16, 32, 64)
    e.Close()


def test_mixed_precision_is_reproducible(af):
    """Two runs of the mixed-precision path give the same bits (fixed-order
    reductions, no atomics)."""
    n = 24
    f = ic.taylor_green(n, n, n, 1, 1 / n, seed=3)
    cfg = dict(NX=n, NY=n, NZ=n, H=1 / n, DT=0.004, Viscosity=1e-5,
               Precision=afm.PRECISION_MIXED, MixedTarget=1e-3)
    outs = []
    for _ in range(2):
        out, d, _ = run_engine(af, cfg, f, 3)
        outs.append([out[k].tobytes() for k in "PUVW"])
    assert outs[0] == outs[1]


@pytest.mark.parametrize("shape", range(1, 15))
def test_step_bitwise_across_sweep_shapes(af, oracle_lib, shape):
    """Every advection workgroup shape gives the reference's bits."""
    cfg = dict(NX=20, NY=18, NZ=14, H=1 / 20, DT=0.01, Viscosity=1e-4, GaussSeidelIters=6)
    f = ic.taylor_green(20, 18, 14, 1, 1 / 20, seed=21, amp=3.0, noise=0.05)
    ge, oe = make_pair(af, oracle_lib, f, **cfg)
    ge.Close()
    ge = af.New(af.Config(**cfg, SweepShape=shape))
    assert ge.sweep_shape()[0] == shape - 1
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        ge.Mesh.SetData(ge.Mesh.field(fid), f[name])
    for s in range(2):
        ge.Step(); assert oe.step() == 0
    assert_bitwise(ge, oe, range(8), f"sweep shape {shape}")
    ge.Close()


@pytest.mark.parametrize("slab", [-1, 1, 2, 3, 7])
def test_step_bitwise_across_pass_orders(af, oracle_lib, slab):
    """The colour passes' workgroup order (plane by plane, or y-slabs of 1-7
    workgroup rows walked through z, uneven last slab) does not change a bit:
    FP64 against the oracle, mixed against the plane order."""
    cfg = dict(NX=40, NY=30, NZ=18, H=1 / 40, DT=0.01, Viscosity=1e-4, GaussSeidelIters=6)
    f = ic.taylor_green(40, 30, 18, 1, 1 / 40, seed=23, amp=3.0, noise=0.05)
    ge, oe = make_pair(af, oracle_lib, f, **cfg)
    ge.Close()
    ge = af.New(af.Config(**cfg, PassSlab=slab, TileK=-2))
    assert ge.pass_slab() == slab and ge.solver_tile()[0] == 0
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        ge.Mesh.SetData(ge.Mesh.field(fid), f[name])
    for s in range(2):
        ge.Step(); assert oe.step() == 0
    assert_bitwise(ge, oe, range(8), f"pass slab {slab}")
    ge.Close()
    mixed = []
    for sl in (-1, slab):
        me = af.New(af.Config(**cfg, PassSlab=sl, Precision=afm.PRECISION_MIXED))
        for fid, name in ((1, "U"), (2, "V"), (3, "W")):
            me.Mesh.SetData(me.Mesh.field(fid), f[name])
        me.StepN(2)
        mixed.append([me.Mesh.Data(me.Mesh.field(k)).tobytes() for k in range(8)])
        me.Close()
    assert mixed[0] == mixed[1], f"mixed pass slab {slab}"


LDS_SHAPES = range(9, 15)  # the LDS-staged sweep variants (SweepShape 9..14)


@pytest.mark.parametrize("shape", LDS_SHAPES)
@pytest.mark.parametrize("shp", [(70, 9, 11, 1), (5, 4, 3, 2), (64, 12, 37, 1)])
@pytest.mark.parametrize("ax", [0, 1, 2])
@pytest.mark.parametrize("reach", ["staged", "beyond"])
def test_lds_sweep_bitwise(af, oracle_lib, shape, shp, ax, reach):
    """The LDS-staged sweep serves taps from its staged box when the
    departure moves at most two cells and from global memory beyond; both
    give the reference's bits, as do the wall column and partial tiles
    (x extents 70 and 64 leave a partial or no partial 64-wide tile)."""
    nx, ny, nz, g = shp
    h = 1.0 / max(nx, ny, nz)
    dt = 0.5 * h
    # "staged": |u| * dt/2 / h <= 0.75 cells; "beyond": up to ~6 cells, clamps.
    f = random_fields(shp, 40 + ax, scale=0.5 if reach == "staged" else 4.0)
    kw = dict(NX=nx, NY=ny, NZ=nz, Ghost=g, H=h, DT=dt, Viscosity=0.1 * h * h / dt)
    ge = af.New(af.Config(**kw, SweepShape=shape))
    assert ge.sweep_shape()[0] == shape - 1
    oe = oracle_lib.OracleEngine(dict(kw, Workers=4))
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        ge.Mesh.SetData(ge.Mesh.field(fid), f[name])
        oe.field(fid)[...] = f[name]
    ge.sweep_axis(ax, dt * 0.5)
    assert oe.sweep_axis(ax, dt * 0.5) == 0
    assert_bitwise(ge, oe, (5, 6, 7), f"lds sweep shape {shape} ax={ax}")
    ge.Close()
