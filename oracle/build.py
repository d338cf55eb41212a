"""Build the oracle (test infrastructure only) with gcc.

Two shared libraries are produced from the same C sources:
  oracle/_build/liboracle64.so  REAL = double (FP64 oracle)
  oracle/_build/liboracle32.so  REAL = float  (FP32 twin)
Both use -ffp-contract=off so every expression rounds as written, and
OpenMP over the pixel loops (the CPU baseline).  The reference itself
(Rust + Vulkan + HLSL/DXC) cannot be built here, so there is no
oracle/_ref build.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build"
SOURCES = ["oracle_kat.c", "oracle_denoise.c"]
FLAGS = ["-O3", "-std=gnu11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
         "-fno-fast-math", "-mavx2", "-mfma", "-Wall", "-Wno-unused-function"]


def extra_sources() -> list[str]:
    return [s for s in sorted(p.name for p in HERE.glob("oracle_*.c")) if s not in SOURCES]


def build(verbose: bool = True) -> None:
    OUT.mkdir(exist_ok=True)
    srcs = [str(HERE / s) for s in SOURCES + extra_sources()]
    for tag, defs in (("64", ["-DORACLE_REAL=double"]), ("32", ["-DORACLE_REAL=float", "-DORACLE_F32"])):
        lib = OUT / f"liboracle{tag}.so"
        tmp = lib.with_suffix(".so.tmp")
        cmd = [os.environ.get("CC", "gcc"), *FLAGS, *defs, "-o", str(tmp), *srcs, "-lm"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"oracle build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, lib)
        if verbose:
            print(f"built {lib}")


if __name__ == "__main__":
    build()
