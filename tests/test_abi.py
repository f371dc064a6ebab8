"""CPU checks of the C ABI library: it loads, exports every symbol that
include/aeroflux.h declares, and its host-side logic (Config.Normalize,
MACBytes) matches the reference's tests.  No compute calls (no GPU here)."""
import math
import os
import re

import pytest

import aeroflux_amd as af

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(af.lib_path()):
        from aeroflux_amd import build
        build.build()
    return af.load_library()


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "aeroflux.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(af_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_boundary():
    syms = declared_symbols()
    for s in ("af_new", "af_step", "af_max_divergence", "af_close", "af_config_normalize",
              "af_project", "af_field_read", "af_field_write"):
        assert s in syms


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing


def test_python_mirror_binds_every_symbol():
    bound = {name for name, _, _ in af._SIGNATURES}
    assert bound == set(declared_symbols())


def test_normalize_defaults(lib):
    # engine_test.go:18-35
    c = af.Config(NX=16, NY=16, NZ=16)
    c.Normalize()
    assert c.Ghost == 1 and c.Density == 1.0 and c.H > 0 and c.DT > 0 and c.Workers >= 1
    assert c.GaussSeidelIters == 24 and c.H == 1 / 16 and c.DT == 0.25 / 16


@pytest.mark.parametrize("kw", [
    dict(NX=0, NY=8, NZ=8), dict(NX=-4, NY=8, NZ=8), dict(NX=8, NY=8, NZ=8, Ghost=-1),
    dict(NX=8, NY=8, NZ=8, Viscosity=-1.0), dict(NX=8, NY=8, NZ=8, H=math.nan),
    dict(NX=8, NY=8, NZ=8, DT=math.inf), dict(NX=8, NY=8, NZ=8, Density=-2.0),
    dict(NX=8, NY=8, NZ=8, H=0.1, Viscosity=1.0, DT=1.0),
])
def test_normalize_rejects(lib, kw):
    # engine_test.go:37-59
    with pytest.raises(af.AeroFluxError, match="^aeroflux: ") as ei:
        af.Config(**kw).Normalize()
    assert ei.value.code == af.AF_ERR_CONFIG


def test_normalize_viscous_default_dt(lib):
    c = af.Config(NX=32, NY=32, NZ=32, Viscosity=1e-3)
    c.Normalize()
    assert c.DT == 0.5 * c.H * c.H / (6 * 1e-3)


def test_workers_capped_by_depth(lib):
    c = af.Config(NX=8, NY=8, NZ=2, Workers=64)
    c.Normalize()
    assert c.Workers <= c.NZ


def test_normalize_matches_oracle(lib, oracle_lib):
    for kw in (dict(NX=64, NY=64, NZ=64, H=1 / 64, DT=0.004, Viscosity=1e-5),
               dict(NX=5, NY=7, NZ=3, Viscosity=2e-4), dict(NX=9, NY=2, NZ=4, Ghost=2)):
        c = af.Config(**kw, Workers=4)
        c.Normalize()
        o = oracle_lib.normalize(dict(kw, Workers=4))
        for k in ("NX", "NY", "NZ", "Ghost", "H", "Viscosity", "Density", "DT", "Workers",
                  "GaussSeidelIters"):
            assert getattr(c, k) == o[k], k


def test_mac_bytes_matches_oracle(lib, oracle_lib):
    o = oracle_lib.load()
    for dims in ((64, 32, 16, 1), (256, 256, 256, 1), (5, 7, 3, 2), (1, 1, 1, 1)):
        assert af.MACBytes(*dims) == o.orc_mac_bytes(*dims)
    assert af.MACBytes(0, 4, 4, 1) == 0


def test_new_without_gpu_raises_cleanly(lib):
    """On a host with no HIP device, New must fail loudly (no CPU fallback)."""
    import ctypes
    cnt = ctypes.c_int(0)
    hip = ctypes.CDLL("libamdhip64.so")
    if hip.hipGetDeviceCount(ctypes.byref(cnt)) == 0 and cnt.value > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(af.AeroFluxError) as ei:
        af.New(af.Config(NX=8, NY=8, NZ=8))
    assert ei.value.code == af.AF_ERR_HIP


def test_sparse_solver_flag_is_accepted(lib):
    """New ignores UseSparseSolver (aeroflux.go:176-213; project.go:19 only
    takes the direct solve once EnableSparseSolver succeeded, which it never
    does off Darwin, sparse_fallback.go:17-24): the flag passes
    normalisation and New gets as far as device work, failing on a host
    without a GPU with the same error as without the flag."""
    import ctypes
    c = af.Config(NX=8, NY=8, NZ=8, UseSparseSolver=True)
    c.Normalize()
    assert c.UseSparseSolver
    cnt = ctypes.c_int(0)
    hip = ctypes.CDLL("libamdhip64.so")
    if hip.hipGetDeviceCount(ctypes.byref(cnt)) == 0 and cnt.value > 0:
        pytest.skip("a HIP device is present (the GPU test covers stepping)")
    with pytest.raises(af.AeroFluxError) as ei:
        af.New(af.Config(NX=8, NY=8, NZ=8, UseSparseSolver=True))
    assert ei.value.code == af.AF_ERR_HIP


# Every struct of the C ABI and its ctypes mirror.
ABI_STRUCTS = {
    "af_config": af._CConfig,
    "af_field": af._CField,
    "af_snapshot_header": af.SnapshotHeader,
    "af_batch_config": af._CBatchConfig,
    "af_batch_info": af._CBatchInfo,
    "af_serve_job": af._CServeJob,
    "af_serve_result": af._CServeResult,
    "af_serve_stats": af._CServeStats,
}


@pytest.mark.parametrize("cname", list(ABI_STRUCTS))
def test_struct_layout_matches_c(tmp_path, cname):
    """Each struct as gcc lays it out from include/aeroflux.h equals its
    ctypes mirror: the same size, the same field names in the same order,
    every field at the same offset (af_snapshot_header: 128 bytes)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    mirror = ABI_STRUCTS[cname]
    names = [n for n, _ in mirror._fields_]
    src = tmp_path / "hdr.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "aeroflux.h"\n'
        "int main(void) {\n"
        f'  printf("%zu\\n", sizeof({cname}));\n'
        + "".join(f'  printf("%zu\\n", offsetof({cname}, {n}));\n' for n in names)
        + "  return 0;\n}\n")
    exe = tmp_path / "hdr"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert out[0] == C_sizeof(mirror)
    assert out[1:] == [getattr(mirror, n).offset for n in names]
    if cname == "af_snapshot_header":
        assert out[0] == 128
    # No C field left out of the mirror: the header's member list, in order.
    hdr = open(os.path.join(ROOT, "include", "aeroflux.h")).read()
    body = hdr[hdr.index(f"typedef struct {cname} {{"):]
    body = body[body.index("{") + 1:body.index(f"}} {cname};")]
    import re
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    members = []
    for stmt in body.split(";"):
        decl = [d.strip() for d in re.sub(r"\[[^\]]*\]", "", stmt).split(",")]
        if decl[0]:
            members += [re.findall(r"\w+", d)[-1] for d in decl]
    assert members == names


def C_sizeof(t):
    import ctypes
    return ctypes.sizeof(t)


@pytest.mark.parametrize("kw", [dict(ServePolicy=4), dict(ServePolicy=-1),
                                dict(ServePolicy=1, PreemptRatio=0.5),
                                dict(SnapshotBudget=-1),
                                dict(ServePolicy=2, StepCellBudget=-5),
                                dict(SolveGroupBytes=-2), dict(StepMaxWait=-1)])
def test_batch_serve_policy_is_validated(lib, kw):
    """Rejected before any device call, so this runs without a GPU."""
    with pytest.raises(af.AeroFluxError) as ei:
        af.Batch(af.BatchConfig(**kw))
    assert ei.value.code == af.AF_ERR_CONFIG


def test_stream_measurement_rejects_bad_arguments(lib):
    for args in ((4, 1024, 1), (0, 1023, 1), (0, 1024, 0), (-1, 1024, 1)):
        with pytest.raises(af.AeroFluxError) as ei:
            af.MeasureStream(*args)
        assert ei.value.code == af.AF_ERR_ARG
