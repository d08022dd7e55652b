"""Builds the C++ tests under tests/cpp/ (host programs over the library's C ABI) into tests/cpp/_build/ with
hipcc; they link libcaelix_amd.so in place (rpath to caelix-renderer_amd/)."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
BUILD = HERE / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(verbose: bool = True) -> list[Path]:
    BUILD.mkdir(exist_ok=True)
    out = []
    for source in sorted(HERE.glob("*.cpp")):
        exe = BUILD / source.stem
        cmd = [HIPCC, "-std=c++17", "-O1", "-Wall", "-Werror", "-I", str(REPO / "include"), str(source),
               "-L", str(REPO / "caelix-renderer_amd"), "-lcaelix_amd",
               "-Wl,-rpath,$ORIGIN/../../../caelix-renderer_amd", "-o", str(exe)]
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            raise RuntimeError(f"hipcc failed for {source.name}:\n{proc.stdout}\n{proc.stderr}")
        out.append(exe)
    if verbose:
        print(f"built {len(out)} C++ test programs in {BUILD}")
    return out


if __name__ == "__main__":
    build()
