"""Build recipe for the HIP library ``aeroflux_amd/libaeroflux.so`` (gfx950).

Run ``python -m aeroflux_amd.build`` or call :func:`build`.  Every source is
compiled with contraction off (the FP64 path must reproduce the reference's
unfused IEEE operation order bit for bit) and with packed FP32 instructions
disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libaeroflux.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["af_kernels.hip", "af_advect.hip", "af_advect_lds.hip", "af_rbgs_tb.hip", "af_rbgs_split.hip", "af_mixed.hip",
           "af_engine.hip", "af_batch.hip", "af_ic.hip", "af_direct.hip",
           "af_snapshot.hip", "af_measure.hip", "af_dct.hip", "af_noalloc.hip"]

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    # hipcc's host pass warns about the device-only target feature; drop it.
    noise = "packed-fp32-ops"
    for line in (proc.stdout + proc.stderr).splitlines():
        if line.strip() and noise not in line and "warning generated" not in line:
            sys.stderr.write(line + "\n")


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link ``libaeroflux.so``; returns its path."""
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(build_dir, src.replace(".hip", ".o"))
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd))
    # The sources compile independently; run them side by side (at most 8 at once).
    jobs = max(1, min(8, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for fut in [pool.submit(_run, cmd) for cmd in cmds]:
            fut.result()
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lrocblas"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
