"""Build oracle/liboracle.so from gridding_oracle.c (TEST INFRASTRUCTURE).

gcc only; -ffp-contract=off keeps float32 rounding step by step (no FMA
contraction), which the reference-order restatement relies on.
"""

from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "gridding_oracle.c"
OUT = HERE / "liboracle.so"


def build(verbose: bool = False) -> Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [
        "gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp",
        "-ffp-contract=off", "-fno-fast-math", "-Wall",
        str(SRC), "-o", str(tmp), "-lm",
    ]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{' '.join(cmd)}\n{res.stderr}")
    tmp.replace(OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
