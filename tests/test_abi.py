"""CPU: the C-ABI library loads and exports every entry point include/caelix_amd/caelix_amd.h declares;
the layouts match the reference's byte sizes and offsets (light_tree_layout.h, hash_table_layout.h)."""
from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent.parent
HEADER = REPO / "include" / "caelix_amd" / "caelix_amd.h"
LIBRARY = REPO / "caelix-renderer_amd" / "libcaelix_amd.so"


def declared_entry_points() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"CX_API\s+[\w\s\*]+?\b(cx_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_entry_points()
    assert len(names) >= 10
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIBRARY)], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (cx_\w+)", out.stdout))
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_library_loads_and_binds(cx):
    lib = cx.load()
    for name in cx.SIGNATURES:
        assert hasattr(lib, name)
    assert lib.cx_version_string().startswith(b"caelix-mi355x")
    # pure host queries need no GPU
    assert lib.cx_light_bvh_scratch_bytes(4096) >= 88 * 4096


def test_argument_errors_return_invalid_argument_without_gpu(cx):
    lib = cx.load()
    messages = []

    @ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p)
    def log(level, message, user):
        messages.append((level, message.decode()))

    lib.cx_set_log_callback(ctypes.cast(log, ctypes.c_void_p), None)
    try:
        assert lib.cx_light_bvh_build(None, None) == 1
        desc = cx.CxLightBvhBuild(1, 1, 1, 1, 1, 100, 10)  # capacity not a multiple of 64
        assert lib.cx_light_bvh_build(ctypes.byref(desc), None) == 1
        assert lib.cx_light_bvh_scratch_init(None, 64, None) == 1
        assert lib.cx_radix_sort_pairs(None, None, None, None, None, None, 0, None, None) == 1
    finally:
        lib.cx_set_log_callback(None, None)
    assert len(messages) == 4 and all(level == 3 for level, _ in messages)


def test_layout_sizes_match_reference():
    from caelix_amd import layout

    assert layout.LEAF.itemsize == 64 and layout.LEAF.fields["parent"][1] == 60
    assert layout.NODE.itemsize == 224 and layout.NODE.fields["child"][1] == 192
    assert layout.NODE.fields["parent"][1] == 208 and layout.NODE.fields["parent_slot"][1] == 212
    assert layout.HEADER.itemsize == 64 and layout.HEADER.fields["node_count"][1] == 56
    assert layout.STATE.itemsize == 64 and layout.STATE.fields["bounds"][1] == 32
    assert layout.HASH_SLOT.itemsize == 16
    assert layout.node_capacity_for(4096) == 4096 // 8 * 5 + 64
    assert np.dtype(layout.SUMMARY).fields["hi"][1] == 36


def test_header_is_c99_and_c99_check_touches_every_entry_point(tmp_path):
    """tests/abi_c99.c compiles as C99 without a warning (the reference's tests/abi_c99.c check) and names
    every entry point the header declares."""
    source = REPO / "tests" / "abi_c99.c"
    text = source.read_text()
    missing = [n for n in declared_entry_points() if f"(AnyFn){n}," not in text]
    assert not missing, missing
    subprocess.run(["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-c", "-I", str(REPO / "include"),
                    str(source), "-o", str(tmp_path / "abi_c99.o")], check=True, capture_output=True, text=True)
