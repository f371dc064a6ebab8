"""Arena snapshot and restore (af_snapshot / af_restore and the per-slot
batch calls).  No reference counterpart: the reference has no engine
snapshot.  What pins it:

* resuming from a snapshot gives the bits of never having stopped, for every
  field of the arena, in FP64, mixed precision and with the direct solve;
* a snapshot moves between an engine and a batch slot of the same grid (the
  batch is bit-identical to the engine alone) with the same result;
* the header's checksum equals the documented formula recomputed here in
  numpy over the payload;
* a snapshot for another grid, a foreign or truncated buffer is rejected
  before the arena is touched; a corrupted payload is rejected and the
  fields are zeroed.
"""
import numpy as np
import pytest

import aeroflux_amd as afm
from aeroflux_amd import ic

pytestmark = pytest.mark.gpu

HDR = 128
MODES = {
    "fp64": ({}, False),
    "mixed": ({"Precision": afm.PRECISION_MIXED}, False),
    "mixed_to_target": ({"Precision": afm.PRECISION_MIXED, "MixedTarget": 1e-6,
                         "MixedSweeps": 8, "MixedMaxCycles": 6}, False),
    "direct": ({}, True),
}


def np_checksum(payload: np.ndarray) -> int:
    """Sum mod 2^64 of SplitMix64's output function applied to
    w_i ^ (i * 0x9E3779B97F4A7C15), over the payload's 64-bit words."""
    w = payload.view(np.uint64)
    i = np.arange(w.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = w ^ (i * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
        return int(z.sum(dtype=np.uint64))


def header(snap: np.ndarray) -> afm.SnapshotHeader:
    return afm.SnapshotHeader.from_buffer_copy(bytes(snap[:HDR]))


def cfg_for(nx, ny, nz, g=1, **extra):
    h = 1.0 / max(nx, ny, nz)
    return dict(NX=nx, NY=ny, NZ=nz, Ghost=g, H=h, DT=0.004, Viscosity=1e-4, **extra)


def seeded(af, cfg, seed, direct=False):
    e = af.New(af.Config(**cfg))
    f = ic.taylor_green(cfg["NX"], cfg["NY"], cfg["NZ"], cfg.get("Ghost", 1),
                        cfg["H"], seed=seed, amp=1.5, noise=0.05)
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        e.Mesh.SetData(e.Mesh.field(fid), f[name])
    e.apply_velocity_bc()
    if direct:
        e.EnableDirectSolver()
    return e


def state(e):
    return [e.Mesh.Data(e.Mesh.field(k)).tobytes() for k in range(8)]


def slot_state(b, slot):
    return [b.Data(slot, k).tobytes() for k in range(8)]


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("shp", [(24, 20, 16, 1), (17, 12, 10, 2)])
def test_resume_from_snapshot_is_bitwise(af, mode, shp):
    extra, direct = MODES[mode]
    cfg = cfg_for(*shp, **extra)
    a = seeded(af, cfg, seed=11, direct=direct)
    a.StepN(3)
    snap = a.Snapshot()
    assert snap.size == a.snapshot_bytes()
    a.StepN(4)
    want = state(a)
    a.Close()

    h = header(snap)
    assert snap[:8].tobytes() == b"AFSNAP\x00\x01"
    assert (h.Version, h.HeaderBytes) == (1, HDR)
    assert (h.NX, h.NY, h.NZ, h.Ghost) == shp
    assert (h.H, h.DT) == (cfg["H"], cfg["DT"])
    assert h.PayloadBytes == snap.size - HDR
    assert h.Checksum == np_checksum(snap[HDR:])

    b = seeded(af, cfg, seed=99, direct=direct)   # a different state, overwritten
    b.Restore(snap)
    b.StepN(4)
    got = state(b)
    b.Close()
    for k in range(8):
        assert got[k] == want[k], f"{mode}: field {afm.FIELD_NAMES[k]}"


def test_snapshot_fields_are_the_arena(af):
    cfg = cfg_for(20, 18, 14)
    e = seeded(af, cfg, seed=3)
    e.StepN(2)
    snap = e.Snapshot()
    off = HDR
    for k in range(8):
        d = e.Mesh.Data(e.Mesh.field(k))
        n = d.nbytes
        assert snap[off:off + n].tobytes() == d.tobytes(), afm.FIELD_NAMES[k]
        off += n
    assert off == snap.size
    e.Close()


def hip_runtime():
    """The HIP runtime the library links (already loaded with it)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def test_snapshot_through_device_memory(af):
    import ctypes
    cfg = cfg_for(24, 16, 12)
    a = seeded(af, cfg, seed=5)
    a.StepN(2)
    n = a.snapshot_bytes()
    hip = hip_runtime()
    ptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(ptr), n) == 0
    try:
        a.Snapshot((ptr.value, n))
        host = a.Snapshot()
        back = np.empty(n, dtype=np.uint8)
        assert hip.hipMemcpy(back.ctypes.data, ptr.value, n, 2) == 0   # device to host
        assert back.tobytes() == host.tobytes()
        a.StepN(3)
        want = state(a)
        b = af.New(af.Config(**cfg))
        b.Restore((ptr.value, n))
        b.StepN(3)
        assert state(b) == want
        b.Close()
    finally:
        a.Close()
        hip.hipFree(ptr.value)


def test_restore_rejects_before_touching_the_arena(af):
    cfg = cfg_for(16, 16, 12)
    a = seeded(af, cfg, seed=7)
    snap = a.Snapshot()
    a.Close()

    other = seeded(af, cfg_for(16, 16, 14), seed=8)
    before = state(other)
    with pytest.raises(af.AeroFluxError) as ei:
        other.Restore(snap)
    assert ei.value.code == afm.AF_ERR_ARG and "does not match" in str(ei.value)
    assert state(other) == before
    other.Close()

    e = seeded(af, cfg, seed=9)
    before = state(e)
    bad_magic = snap.copy()
    bad_magic[0] ^= 0xFF
    cases = [(bad_magic, "not an aeroflux snapshot"),
             (snap[:HDR + 100].copy(), "truncated"),
             (snap[:64].copy(), "shorter than its header")]
    for buf, msg in cases:
        with pytest.raises(af.AeroFluxError) as ei:
            e.Restore(buf)
        assert msg in str(ei.value)
        assert state(e) == before
    small = np.empty(e.snapshot_bytes() - 1, dtype=np.uint8)
    with pytest.raises(af.AeroFluxError) as ei:
        e.Snapshot(small)
    assert "needs" in str(ei.value)
    e.Close()


def test_corrupt_payload_is_rejected_and_zeroed(af):
    cfg = cfg_for(16, 12, 10)
    a = seeded(af, cfg, seed=4)
    a.StepN(1)
    snap = a.Snapshot()
    a.Close()
    bad = snap.copy()
    bad[HDR + (bad.size - HDR) // 3] ^= 0x10
    e = seeded(af, cfg, seed=6)
    with pytest.raises(af.AeroFluxError) as ei:
        e.Restore(bad)
    assert "checksum mismatch" in str(ei.value)
    for k in range(8):
        assert not e.Mesh.Data(e.Mesh.field(k)).any()
    e.Restore(snap)   # a good snapshot still restores afterwards
    assert e.Snapshot().tobytes() == snap.tobytes()
    e.Close()


@pytest.mark.parametrize("prec", [afm.PRECISION_FP64, afm.PRECISION_MIXED])
def test_snapshot_moves_between_engine_and_batch(af, prec):
    cfg = cfg_for(24, 16, 20, Precision=prec)
    others = [cfg_for(16, 12, 10, Precision=prec), cfg_for(32, 24, 16, Precision=prec)]
    ref = seeded(af, cfg, seed=21)
    ref.StepN(2)
    snap = ref.Snapshot()
    ref.StepN(3)
    after5 = state(ref)
    ref.StepN(2)
    after7 = state(ref)
    ref.Close()

    b = af.Batch(af.BatchConfig(Precision=prec, MemoryBudget=256 << 20))
    for i, oc in enumerate(others):
        s = b.Add(af.Config(**oc))
        f = ic.taylor_green(oc["NX"], oc["NY"], oc["NZ"], 1, oc["H"], seed=40 + i)
        for fid, name in ((1, "U"), (2, "V"), (3, "W")):
            b.SetData(s, fid, f[name])
        b.apply_velocity_bc(s)
    slot = b.Add(af.Config(**cfg))
    assert b.snapshot_bytes(slot) == snap.size
    b.Restore(slot, snap)                      # engine -> batch slot
    b.StepN(3)
    assert slot_state(b, slot) == after5
    back = b.Snapshot(slot)                    # batch slot -> engine
    b.Close()
    e = af.New(af.Config(**cfg))
    e.Restore(back)
    e.StepN(2)
    assert state(e) == after7
    e.Close()
