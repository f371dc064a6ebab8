"""Build recipe for the CPU oracle ``oracle/liboracle.so``.

TEST INFRASTRUCTURE ONLY (see aeroflux_oracle.c).  The restatement is plain
C11 compiled with ``-O2 -ffp-contract=off``: no fused multiply-add and no
fast-math, so it evaluates the reference's IEEE binary64 operations in source
order.  The Go reference itself cannot be built here (no Go toolchain), so
there is no ``oracle/_ref`` build.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "aeroflux_oracle.c")
LIB = os.path.join(HERE, "liboracle.so")

CFLAGS = ["-O2", "-std=c11", "-fPIC", "-shared", "-pthread", "-ffp-contract=off",
          "-fno-fast-math", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def build(extra: list[str] | None = None, out: str | None = None) -> str:
    """Compile the oracle; ``extra`` flags and ``out`` path are for variants."""
    out = out or LIB
    tmp = out + ".tmp"
    cmd = ["gcc", *CFLAGS, *(extra or []), "-o", tmp, SRC, "-lm"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build())
