"""The C ABI driven from C in the Go binding's call order (tests/abi_harness.c).

The Go package go/aeroflux cannot be built here (no Go toolchain); the
harness issues the same calls in the same order -- normalise, New, field
descriptors, mirror refresh and write-back, Step per call, read-back,
diagnostics, EnableSparseSolver's error, Close twice -- and its output must
equal the oracle's bit for bit (KineticEnergy within 1e-12 relative: fixed
tree against a serial sum).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from aeroflux_amd import ic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_harness(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    lib_dir = os.path.join(ROOT, "aeroflux_amd")
    if not os.path.exists(os.path.join(lib_dir, "libaeroflux.so")):
        pytest.skip("libaeroflux.so not built")
    exe = str(tmp_path / "abi_harness")
    cmd = ["gcc", "-O2", "-Wall", "-Wextra", "-Werror", "-std=c11",
           "-I" + os.path.join(ROOT, "include"), "-o", exe,
           os.path.join(ROOT, "tests", "abi_harness.c"), "-L" + lib_dir, "-laeroflux",
           "-Wl,-rpath," + lib_dir, "-Wl,-rpath-link,/opt/rocm/lib"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr
    return exe


def test_harness_links_and_runs_host_calls(tmp_path):
    """Compiles against include/aeroflux.h, links libaeroflux.so, and runs
    the calls that need no device."""
    exe = build_harness(tmp_path)
    proc = subprocess.run([exe, "-"], capture_output=True, text=True, timeout=60)
    assert proc.returncode == 0, proc.stderr
    assert "host-only checks passed" in proc.stdout


@pytest.mark.gpu
def test_harness_steps_like_the_oracle(tmp_path, oracle_lib):
    nx, ny, nz, h, dt, nu, steps = 20, 14, 11, 1 / 20, 0.004, 1e-4, 4
    exe = build_harness(tmp_path)
    f = ic.taylor_green(nx, ny, nz, 1, h, seed=17, noise=0.05)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        for k in "UVW":
            fh.write(np.ascontiguousarray(f[k], dtype="<f8").tobytes())
    print(f"abi harness: {nx}x{ny}x{nz}, {steps} steps", flush=True)
    proc = subprocess.run([exe, str(inp), str(out), str(nx), str(ny), str(nz), repr(h),
                           repr(dt), repr(nu), str(steps)], capture_output=True, text=True,
                          timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    o = oracle_lib.OracleEngine(dict(NX=nx, NY=ny, NZ=nz, H=h, DT=dt, Viscosity=nu, Workers=4))
    for fid, k in ((1, "U"), (2, "V"), (3, "W")):
        o.field(fid)[...] = f[k]
    o.apply_velocity_bc()
    for _ in range(steps):
        assert o.step() == 0
    raw = np.fromfile(out, dtype="<f8")
    pos = 0
    for fid in range(8):
        want = o.field(fid)
        got = raw[pos:pos + want.size].reshape(want.shape)
        pos += want.size
        assert got.tobytes() == want.tobytes(), fid
    mdiv, ke, mvel, cfl = raw[pos:pos + 4]
    assert pos + 4 == raw.size
    assert mdiv == o.max_divergence() and mvel == o.max_velocity() and cfl == o.cfl()
    assert abs(ke - o.kinetic_energy()) <= 1e-12 * o.kinetic_energy()
    o.close()
