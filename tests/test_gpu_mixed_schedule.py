"""The checked mixed-precision projection's three outcomes, each pinned.

A projection with a max|div| target (MixedTarget) runs the first cycle's
FP32 sweep pairs, then checks the FP64 residual of p = e without writing
anything.  Met at once, the projection is the unchecked one-cycle
projection, bit for bit.  Missed, the same FP32 iteration continues by
MixedFixSweeps pairs on the same right-hand side: when that meets the target
the projection is the one-cycle projection with MixedSweeps + MixedFixSweeps
pairs, bit for bit.  Missed again, FP64 defect-correction cycles merge the
solution into P and the gradient reads P (a device flag picks the gradient):
the residual the last check saw must then be the divergence the projection
leaves, max|div| = DT/(rho H^2) max|r| up to rounding, which fails if the
gradient read the wrong pressure.
"""
import numpy as np
import pytest

import aeroflux_amd as afm
from aeroflux_amd import ic

pytestmark = pytest.mark.gpu

GRIDS = [(40, 36, 32, 1), (33, 20, 18, 2)]
S, FIX = 6, 2


def project(af, shp, **extra):
    nx, ny, nz, g = shp
    h = 1.0 / max(nx, ny, nz)
    cfg = dict(NX=nx, NY=ny, NZ=nz, Ghost=g, H=h, DT=0.004, Viscosity=1e-5,
               Precision=afm.PRECISION_MIXED, MixedFixSweeps=FIX, MixedMaxCycles=4)
    cfg.update(extra)
    f = ic.taylor_green(nx, ny, nz, g, h, seed=17, amp=2.0, noise=0.05)
    e = af.New(af.Config(**cfg))
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        e.Mesh.SetData(e.Mesh.field(fid), f[name])
    e.apply_velocity_bc()
    e.project()
    out = [e.Mesh.Data(e.Mesh.field(k)).tobytes() for k in range(4)]
    d = e.MaxDivergence()
    last, cycles = e.mixed_stats()
    e.Close()
    return out, d, last, cycles


@pytest.mark.parametrize("shp", GRIDS)
def test_checked_projection_outcomes(af, shp):
    one_s, d_s, _, _ = project(af, shp, MixedSweeps=S)
    one_s2, d_s2, _, _ = project(af, shp, MixedSweeps=S + FIX)
    assert d_s2 < 0.98 * d_s, (d_s, d_s2)

    # Met at the first check: the one-cycle projection's bits.  The
    # residual-implied figure is the divergence left, up to rounding.
    a, d_a, last_a, cyc_a = project(af, shp, MixedSweeps=S, MixedTarget=1e30)
    assert a == one_s and d_a == d_s and cyc_a == 1
    assert abs(last_a - d_a) <= 1e-9 * d_a, (last_a, d_a)

    # Met after the continued iteration: the S + FIX one-cycle bits.
    tgt = (d_s * d_s2) ** 0.5
    b, d_b, last_b, cyc_b = project(af, shp, MixedSweeps=S, MixedTarget=tgt)
    assert cyc_b == 2 and b == one_s2 and d_b == d_s2 and d_b <= tgt

    # Missed twice: defect-correction cycles, the gradient reads P.
    c, d_c, last_c, cyc_c = project(af, shp, MixedSweeps=S, MixedTarget=0.25 * d_s2)
    assert cyc_c >= 3 and c != one_s2
    assert abs(last_c - d_c) <= 1e-9 * d_c, (last_c, d_c)
    assert d_c < d_s2
    if cyc_c < 4:
        assert d_c <= 0.25 * d_s2 * (1 + 1e-9)
    for k in range(4):
        assert np.all(np.isfinite(np.frombuffer(c[k], dtype=np.float64)))
