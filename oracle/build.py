"""ORACLE build (test infrastructure only): compiles the CPU restatements under oracle/ into
oracle/_build/libcaelix_oracle.so with g++. Floating point: -ffp-contract=off, no fast math, so each
float operation is one IEEE-754 binary32 operation in source order.

The reference implementation itself cannot be built here (it needs the Vulkan SDK, DXC and network
fetches of its CMake dependencies), so there is no oracle/_ref build.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

ORACLE = Path(__file__).resolve().parent
REPO = ORACLE.parent
BUILD = ORACLE / "_build"
LIBRARY = BUILD / "libcaelix_oracle.so"


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(str(p) for p in ORACLE.glob("*.cpp"))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fopenmp",
           "-Wall", "-I", str(REPO / "include"), "-o", str(tmp), *sources]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY} from {len(sources)} sources")
    return LIBRARY


if __name__ == "__main__":
    build()
