"""af_measure_stream: the attainable-bandwidth probe bench.py reports next to
the roofline.  Each pattern runs and gives a positive, finite rate below
what no HBM3E part reaches (the MI355X's published 8 TB/s, with room for
the last-level cache on arrays that fit it)."""
import math

import pytest

import aeroflux_amd as afm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", [afm.STREAM_READ, afm.STREAM_COPY, afm.STREAM_2R1W,
                                  afm.STREAM_4R3W])
def test_stream_measurement_runs(af, kind):
    g = af.MeasureStream(kind, 1 << 22, 5)
    assert math.isfinite(g) and 10.0 < g < 40000.0
