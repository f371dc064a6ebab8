"""Batched engines: each slot's results are the same alone and in the mix.

A batch steps uneven grids together (persistent launches over a tile
work-list of every slot, one hipGraph per step).  The bar is bit-identity
with the same engine stepped alone through af_step -- itself bit-exact to
the oracle in FP64 -- for every field of the arena, in FP64 and in both
mixed-precision modes, across membership changes, tuned and fixed launch
shapes, and with a faulting neighbour.
"""
import numpy as np
import pytest

import aeroflux_amd as afm
from aeroflux_amd import ic

pytestmark = pytest.mark.gpu

NAMES = ("P", "U", "V", "W", "Div", "Tmp0", "Tmp1", "Tmp2")

# Uneven grids, one with Ghost = 2, one with an odd x extent.
GRIDS = [(16, 12, 10, 1), (24, 16, 20, 1), (32, 32, 32, 1), (12, 20, 8, 2), (33, 9, 14, 1)]


def cfg_for(shp, i, **extra):
    nx, ny, nz, g = shp
    nu = [1e-5, 1e-4, 3e-6, 2e-5, 5e-5][i % 5]
    h = 1.0 / max(nx, ny, nz)
    dt = min(0.004, 0.5 * h * h / (6 * nu))
    return dict(NX=nx, NY=ny, NZ=nz, Ghost=g, H=h, Viscosity=nu, DT=dt, **extra)


def fields_for(shp, i):
    nx, ny, nz, g = shp
    return ic.taylor_green(nx, ny, nz, g, 1.0 / max(nx, ny, nz), seed=100 + i)


def solo(af, cfg, f, steps):
    e = af.New(af.Config(**cfg))
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        e.Mesh.SetData(e.Mesh.field(fid), f[name])
    e.apply_velocity_bc()
    if steps:
        e.StepN(steps)
    out = [e.Mesh.Data(e.Mesh.field(k)) for k in range(8)]
    extra = (e.MaxDivergence(), e.KineticEnergy())
    stats = e.mixed_stats() if cfg.get("Precision") == afm.PRECISION_MIXED else None
    e.Close()
    return out, extra, stats


def load(b, slot, f):
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        b.SetData(slot, fid, f[name])
    b.apply_velocity_bc(slot)


def assert_same(a, b, what):
    for k in range(8):
        if a[k].tobytes() != b[k].tobytes():
            bad = np.argwhere(a[k].view(np.uint64) != b[k].view(np.uint64))
            z, y, x = bad[0]
            raise AssertionError(f"{what}: field {NAMES[k]} differs at {len(bad)} cells, first "
                                 f"(x={x},y={y},z={z}): {a[k][z, y, x]!r} vs {b[k][z, y, x]!r}")


def slot_state(b, slot):
    return [b.Data(slot, k) for k in range(8)]


PRECISIONS = {
    "fp64": (afm.PRECISION_FP64, {}, {}),
    "mixed_one_cycle": (afm.PRECISION_MIXED, {}, {}),
    "mixed_checked": (afm.PRECISION_MIXED, {"MixedChecked": 1}, {"MixedTarget": 2e-6}),
}


def split_bytes(shp, p):
    nx, ny, nz, _ = shp
    return 4 * nz * ny * ((nx + 1) // 2) * (4 if p == afm.PRECISION_MIXED else 8)


# Solve groups (SolveGroupBytes): the default (every slot in one group at
# these sizes), about a third of the solve buffers per group, one slot per
# group.
GROUPS = {"one_group": 0, "thirds": -3, "each_slot": 1}


@pytest.mark.parametrize("groups", list(GROUPS))
@pytest.mark.parametrize("prec", list(PRECISIONS))
def test_batch_equals_alone(af, prec, groups):
    p, bextra, eextra = PRECISIONS[prec]
    steps = 4
    cfgs = [cfg_for(s, i, Precision=p, **eextra) for i, s in enumerate(GRIDS)]
    fs = [fields_for(s, i) for i, s in enumerate(GRIDS)]
    gb = GROUPS[groups]
    total = sum(split_bytes(s, p) for s in GRIDS)
    if gb < 0:
        gb = (total + 2) // 3
    b = af.Batch(af.BatchConfig(Precision=p, MemoryBudget=256 << 20, SolveGroupBytes=gb,
                                **bextra))
    slots = [b.Add(af.Config(**c)) for c in cfgs]
    for s, f in zip(slots, fs):
        load(b, s, f)
    b.StepN(steps)
    info = b.info()
    assert info["Slots"] == len(GRIDS) and info["Tuned"] == 1
    want = {"one_group": 1, "each_slot": len(GRIDS)}.get(groups)
    if want is None:
        assert 2 <= info["SolveGroups"] <= 3, info
    else:
        assert info["SolveGroups"] == want, info
    for i, (s, c, f) in enumerate(zip(slots, cfgs, fs)):
        ref, (d, ke), stats = solo(af, c, f, steps)
        assert_same(slot_state(b, s), ref, f"{prec} slot {s} grid {GRIDS[i]}")
        assert b.MaxDivergence(s) == d
        assert b.KineticEnergy(s) == ke
        if prec == "mixed_checked":
            assert b.mixed_stats(s) == stats
    b.Close()


@pytest.mark.parametrize("group_bytes", [0, 100000])
def test_membership_changes_between_steps(af, group_bytes):
    """Continuous batching: engines join and leave between steps; each one's
    state equals the same engine stepped alone for the steps it was in.
    With 100 kB solve groups the group count changes with the membership
    (1, 2, 3 ... groups), and each count's step graph is kept and reused."""
    b = af.Batch(af.BatchConfig(MemoryBudget=256 << 20, SolveGroupBytes=group_bytes))
    cfgs = [cfg_for(s, i) for i, s in enumerate(GRIDS)]
    fs = [fields_for(s, i) for i, s in enumerate(GRIDS)]
    live = {}
    # Steps 0-1: engines 0, 1.  Step 2: 2 joins.  Step 3: 0 leaves and 3
    # takes its slot and pool block.  Step 4: 4 joins.
    plan = {0: [0, 1], 2: [2], 3: [3], 4: [4]}
    leave = {3: [0]}
    done = {}
    for step in range(6):
        for i in leave.get(step, []):
            slot, joined = live.pop(i)
            done[i] = (slot_state(b, slot), step - joined)
            b.Remove(slot)
        for i in plan.get(step, []):
            slot = b.Add(af.Config(**cfgs[i]))
            load(b, slot, fs[i])
            live[i] = (slot, step)
        b.StepN(1)
    for i, (slot, joined) in live.items():
        done[i] = (slot_state(b, slot), 6 - joined)
    for i, (state, n) in done.items():
        ref, _, _ = solo(af, cfgs[i], fs[i], n)
        assert_same(state, ref, f"engine {i} after {n} steps in a changing batch")
    b.Close()


def test_tuned_and_fixed_launch_shapes_agree(af):
    cfgs = [cfg_for(s, i) for i, s in enumerate(GRIDS[:3])]
    fs = [fields_for(s, i) for i, s in enumerate(GRIDS[:3])]
    results = []
    for shape, wg in ((0, 0), (1, 8), (8, 1024), (5, 64)):
        b = af.Batch(af.BatchConfig(MemoryBudget=128 << 20, SweepShape=shape, Workgroups=wg))
        slots = [b.Add(af.Config(**c)) for c in cfgs]
        for s, f in zip(slots, fs):
            load(b, s, f)
        b.StepN(2)
        results.append([slot_state(b, s) for s in slots])
        b.Close()
    for r in results[1:]:
        for a, c in zip(results[0], r):
            assert_same(a, c, "launch shape")


# Grids whose x extent fills 64-wide tiles to at least 3/4 take the
# LDS-staged sweep when the batch enables it; (33, 9, 14) and (70, 8, 6)
# stay on the gather beside them.  (64, 7, 33, 2) has Ghost = 2.
STAGED_GRIDS = [(64, 12, 10, 1), (96, 9, 14, 1), (128, 6, 5, 1), (64, 7, 33, 2), (33, 9, 14, 1),
                (70, 8, 6, 2), (48, 10, 9, 1)]


def staged_fields(i, shp, amp):
    nx, ny, nz, g = shp
    return ic.taylor_green(nx, ny, nz, g, 1.0 / max(nx, ny, nz), seed=200 + i, amp=amp,
                           noise=0.3 * amp)


@pytest.mark.parametrize("variant", range(9))
@pytest.mark.parametrize("amp", [1.0, 24.0])
def test_staged_sweep_batch_equals_alone(af, variant, amp):
    """Slots on the LDS-staged sweep -- departures within the staged reach at
    amplitude 1, up to ~6 cells (clamped at the walls) at amplitude 24 -- and
    the slots beside them on the gather give the bits of the same engines
    stepped alone.  variant 0 lets the tuner choose."""
    steps = 2
    cfgs = [cfg_for(s, i) for i, s in enumerate(STAGED_GRIDS)]
    fs = [staged_fields(i, s, amp) for i, s in enumerate(STAGED_GRIDS)]
    b = af.Batch(af.BatchConfig(MemoryBudget=256 << 20, SweepStaged=variant))
    slots = [b.Add(af.Config(**c)) for c in cfgs]
    for s, f in zip(slots, fs):
        load(b, s, f)
    b.StepN(steps)
    info = b.info()
    if variant:
        assert info["SweepStaged"] == variant - 1 and info["StagedSlots"] == 5
    else:
        assert -1 <= info["SweepStaged"] < 8
    for i, (s, c, f) in enumerate(zip(slots, cfgs, fs)):
        ref, (d, _), _ = solo(af, c, f, steps)
        assert_same(slot_state(b, s), ref, f"staged variant {variant} grid {STAGED_GRIDS[i]}")
        assert b.MaxDivergence(s) == d
    b.Close()


def test_staged_sweep_membership_growth(af):
    """The staged sweep runs one workgroup per tile, so its grid is sized to
    the work-list: adding slots past it rebuilds the step graph.  Results
    stay those of the engines alone."""
    b = af.Batch(af.BatchConfig(MemoryBudget=512 << 20, SweepStaged=3))
    grids = [(64, 8, 8, 1), (128, 24, 20, 1), (96, 40, 33, 1)]
    cfgs = [cfg_for(s, i) for i, s in enumerate(grids)]
    fs = [staged_fields(i, s, 1.0) for i, s in enumerate(grids)]
    live = []
    for i in range(len(grids)):
        s = b.Add(af.Config(**cfgs[i]))
        load(b, s, fs[i])
        live.append(s)
        b.StepN(1)
    for i, s in enumerate(live):
        ref, _, _ = solo(af, cfgs[i], fs[i], len(grids) - i)
        assert_same(slot_state(b, s), ref, f"engine {i} of a growing staged batch")
    b.Close()


def test_fault_in_one_slot_is_contained(af):
    """A NaN departure coordinate in one slot raises that slot's flag only;
    the other slots' results are unaffected."""
    cfgs = [cfg_for(s, i) for i, s in enumerate(GRIDS[:3])]
    fs = [fields_for(s, i) for i, s in enumerate(GRIDS[:3])]
    bad = {k: v.copy() for k, v in fs[1].items()}
    bad["U"][5, 5, 5] = np.nan
    b = af.Batch(af.BatchConfig(MemoryBudget=128 << 20, SweepStaged=-1))
    slots = [b.Add(af.Config(**c)) for c in cfgs]
    for s, f in zip(slots, (fs[0], bad, fs[2])):
        load(b, s, f)
    with pytest.raises(af.AeroFluxError) as ei:
        b.StepN(2)
    assert ei.value.code == afm.AF_ERR_FAULT
    assert ei.value.slots == [slots[1]]
    for i in (0, 2):
        ref, _, _ = solo(af, cfgs[i], fs[i], 2)
        assert_same(slot_state(b, slots[i]), ref, f"healthy slot {i} beside a fault")
    b.Close()


def test_fault_in_a_staged_slot_is_contained(af):
    """The same on the LDS-staged sweep: the faulting slot (64 wide, staged)
    is named, the gather and staged slots beside it are unaffected."""
    grids = [STAGED_GRIDS[0], STAGED_GRIDS[1], STAGED_GRIDS[4]]
    cfgs = [cfg_for(s, i) for i, s in enumerate(grids)]
    fs = [staged_fields(i, s, 1.0) for i, s in enumerate(grids)]
    bad = {k: v.copy() for k, v in fs[0].items()}
    bad["V"][4, 6, 30] = np.nan
    b = af.Batch(af.BatchConfig(MemoryBudget=128 << 20, SweepStaged=1))
    slots = [b.Add(af.Config(**c)) for c in cfgs]
    for s, f in zip(slots, (bad, fs[1], fs[2])):
        load(b, s, f)
    with pytest.raises(af.AeroFluxError) as ei:
        b.StepN(2)
    assert ei.value.code == afm.AF_ERR_FAULT
    assert ei.value.slots == [slots[0]]
    for i in (1, 2):
        ref, _, _ = solo(af, cfgs[i], fs[i], 2)
        assert_same(slot_state(b, slots[i]), ref, f"healthy slot {i} beside a staged fault")
    b.Close()


def test_pool_admission_and_reuse(af):
    c = cfg_for((32, 32, 32, 1), 0)
    need = af.SlotBytes(af.Config(**c), afm.PRECISION_FP64)
    assert need > af.MACBytes(32, 32, 32, 1)
    b = af.Batch(af.BatchConfig(MemoryBudget=2 * need + need // 2, MaxSlots=8))
    s0 = b.Add(af.Config(**c))
    s1 = b.Add(af.Config(**c))
    b.SetData(s0, afm.FIELD_U, np.full(b.field(s0, afm.FIELD_U).shape, 3.0))
    with pytest.raises(af.AeroFluxError) as ei:
        b.Add(af.Config(**c))
    assert ei.value.code == afm.AF_ERR_OOM
    b.Remove(s0)
    s2 = b.Add(af.Config(**c))
    assert s2 == s0
    # The reused block starts zeroed (MAC.Zero), like a fresh engine.
    assert all(not b.Data(s2, k).any() for k in range(8))
    info = b.info()
    assert info["Slots"] == 2 and info["BytesInUse"] == 2 * need
    b.Remove(s1)
    b.Remove(s2)
    assert b.info()["BytesInUse"] == 0
    b.Close()


def test_batch_rejects_mismatched_schedule(af):
    b = af.Batch(af.BatchConfig(MemoryBudget=64 << 20))
    with pytest.raises(af.AeroFluxError) as ei:
        b.Add(af.Config(**cfg_for((16, 16, 16, 1), 0, GaussSeidelIters=10)))
    assert ei.value.code == afm.AF_ERR_CONFIG
    with pytest.raises(af.AeroFluxError) as ei:
        b.Add(af.Config(**cfg_for((16, 16, 16, 1), 0, Precision=afm.PRECISION_MIXED)))
    assert ei.value.code == afm.AF_ERR_CONFIG
    with pytest.raises(af.AeroFluxError):
        b.Remove(3)
    # UseSparseSolver is accepted (the red-black solve runs), as by New.
    s = b.Add(af.Config(**cfg_for((16, 16, 16, 1), 0, UseSparseSolver=True)))
    b.Remove(s)
    b.Close()
    with pytest.raises(af.AeroFluxError) as ei:
        b.Add(af.Config(**cfg_for((16, 16, 16, 1), 0)))
    assert ei.value.code == afm.AF_ERR_CLOSED
