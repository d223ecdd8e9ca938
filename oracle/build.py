"""Build the C restatement of the reference hot path (TEST INFRASTRUCTURE).

oracle/csrc/oracle.c -> oracle/liboracle.so (FP64 functions, suffix _f64) and
the FP32 variant (suffix _f32), gcc -O3 -fopenmp. Used only by tests and by
bench.py's CPU-baseline leg.
"""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "csrc" / "oracle.c"
LIB = HERE / "liboracle.so"


def build() -> Path | None:
    if not SRC.exists():
        return None
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-Wall", str(SRC), "-o", str(tmp), "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
