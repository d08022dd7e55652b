"""GPU: the light tree adapter driven from C++ through the C ABI (tests/cpp/light_tree_adapter_test.cpp,
built by __graft_entry__.build() into tests/cpp/_build/): begin's decision over every combination, the
build-up from a 4096-byte heap, an edit's recomputed and copied records, draws against point pdfs."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

EXE = Path(__file__).resolve().parent / "cpp" / "_build" / "light_tree_adapter_test"


@pytest.mark.gpu
def test_light_tree_adapter_cpp():
    assert EXE.exists(), "build the C++ tests first (__graft_entry__.build())"
    proc = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    print(proc.stdout)
    print(proc.stderr)
    assert proc.returncode == 0, proc.stderr
    assert "PASSED" in proc.stdout
