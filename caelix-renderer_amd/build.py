"""Builds the gfx950 library libcaelix_amd.so in-tree with hipcc.

Every .hip file under csrc/ is compiled for gfx950 only, every .cpp file (host code: the .cxw reader)
with g++, and all are linked into one shared library with a C ABI (include/caelix_amd/caelix_amd.h). Floating point keeps IEEE single operations in source order
(-ffp-contract=off, no fast math), which the bit-exact parity of the light-tree build relies on.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
REPO = PACKAGE.parent
CSRC = PACKAGE / "csrc"
BUILD = PACKAGE / "_build"
LIBRARY = PACKAGE / "libcaelix_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMPILE_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fvisibility=hidden",
    # Packed FP32 instructions stay off (the feature is disabled for every device function).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(REPO / "include"),
    "-I", str(CSRC),
    "-Wall",
    "-Wno-unused-function",
]


def _compile(source: Path) -> Path:
    obj = BUILD / (source.stem + ".o")
    cmd = [HIPCC, *COMPILE_FLAGS, "-c", str(source), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {source.name}:\n{proc.stdout}\n{proc.stderr}")
    return obj


HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-pthread", "-Wall", "-I", str(REPO / "include")]


def _compile_host(source: Path) -> Path:
    obj = BUILD / (source.stem + ".o")
    cmd = ["g++", *HOST_FLAGS, "-c", str(source), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"g++ failed for {source.name}:\n{proc.stdout}\n{proc.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    """Compiles every csrc/*.hip for gfx950 and every csrc/*.cpp (host code) and links libcaelix_amd.so;
    returns its path."""
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    host_sources = sorted(CSRC.glob("*.cpp"))
    workers = min(8, len(sources) + len(host_sources)) or 1
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
        objects = list(pool.map(_compile, sources)) + list(pool.map(_compile_host, host_sources))
    sources = sources + host_sources
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objects), "-lz", "-pthread"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc link failed:\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY} from {len(sources)} sources")
    return LIBRARY


if __name__ == "__main__":
    build()
    sys.exit(0)
