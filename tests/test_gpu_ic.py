"""Device band-limited initial condition against the host generator.

The host generator (aeroflux_amd.ic.band_limited) and the GPU generator
draw the same SplitMix64 phases and evaluate the same separable sum; only
the summation order differs, so the fields agree to rounding: max abs
difference <= 1e-11 * rms over the whole padded U/V/W (ghosts after the
velocity BC included).  An engine and a batch slot use the same kernels
and must agree bit for bit.
"""
import numpy as np
import pytest

import aeroflux_amd as afm
from aeroflux_amd import ic

pytestmark = pytest.mark.gpu

CASES = [(16, 12, 10, 1, 3), (24, 16, 20, 2, 11), (33, 9, 14, 1, 5), (64, 64, 64, 1, 2)]


@pytest.mark.parametrize("case", CASES)
def test_device_band_limited_matches_host(af, case):
    nx, ny, nz, g, seed = case
    h = 1.0 / max(nx, ny, nz)
    cfg = af.Config(NX=nx, NY=ny, NZ=nz, Ghost=g, H=h, Viscosity=1e-5, DT=0.004)
    host = ic.band_limited(nx, ny, nz, g, h, seed=seed)
    ref = af.New(cfg)
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        ref.Mesh.SetData(ref.Mesh.field(fid), host[name])
    ref.apply_velocity_bc()
    e = af.New(cfg)
    e.init_band_limited(seed)
    for fid in (1, 2, 3):
        a = e.Mesh.Data(e.Mesh.field(fid))
        b = ref.Mesh.Data(ref.Mesh.field(fid))
        assert np.abs(a - b).max() <= 1e-11, fid
        # Zero exactly where the host field is zero (outside faces and BC box).
        assert np.array_equal(a == 0, b == 0), fid
    # A batch slot builds the same bits.
    b = af.Batch(af.BatchConfig(MemoryBudget=64 << 20))
    s = b.Add(cfg)
    b.init_band_limited(s, seed)
    for fid in range(8):
        assert b.Data(s, fid).tobytes() == e.Mesh.Data(e.Mesh.field(fid)).tobytes(), fid
    b.Close()
    e.Close()
    ref.Close()


def test_band_limited_seeds_differ_and_repeat(af):
    cfg = af.Config(NX=16, NY=16, NZ=16, H=1 / 16, Viscosity=1e-5, DT=0.004)
    out = []
    for seed in (1, 2, 1):
        e = af.New(cfg)
        e.init_band_limited(seed)
        out.append(e.Mesh.Data(e.Mesh.U))
        e.Close()
    assert out[0].tobytes() == out[2].tobytes()
    assert not np.array_equal(out[0], out[1])


def test_band_limited_rejects_bad_arguments(af):
    e = af.New(af.Config(NX=8, NY=8, NZ=8))
    for kw in (dict(kmax=0), dict(rms=0.0), dict(rms=-1.0)):
        with pytest.raises(af.AeroFluxError) as ei:
            e.init_band_limited(1, **kw)
        assert ei.value.code == afm.AF_ERR_ARG
    e.Close()
