"""ORACLE — test infrastructure only.

ctypes binding of oracle/_build/libcaelix_oracle.so, the sequential CPU restatement of the reference's
hot-path algorithms. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package, as the checker; the product (caelix-renderer_amd/) never does.

Parity status per function (how each is pinned independently of its own formula) is listed in the
header comment of each C++ source file.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIBRARY_PATH = ORACLE_DIR / "_build" / "libcaelix_oracle.so"

# Layout sizes (include/caelix_amd/layout.h).
LEAF_BYTES = 64
NODE_BYTES = 224
HEADER_BYTES = 64
STATE_BYTES = 64
SCRATCH_SORT_STATE = 88

_lib = None
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY_PATH.exists():
        from . import build as _build
        _build.build(verbose=False)
    lib = ctypes.CDLL(str(LIBRARY_PATH))
    lib.oracle_morton3.restype = _U32
    lib.oracle_morton3.argtypes = [_U32, _U32, _U32]
    lib.oracle_radix_sort_pairs.restype = None
    lib.oracle_radix_sort_pairs.argtypes = [_P, _P, _P, _P, _U32]
    lib.oracle_exclusive_scan.restype = _U32
    lib.oracle_exclusive_scan.argtypes = [_P, _U32]
    lib.oracle_lt_build.restype = None
    lib.oracle_lt_build.argtypes = [_P, _U32, _U32, _P, _P, _P, _P]
    lib.oracle_splitmix64.restype = ctypes.c_uint64
    lib.oracle_splitmix64.argtypes = [ctypes.c_uint64]
    lib.oracle_hash_fingerprint.restype = None
    lib.oracle_hash_fingerprint.argtypes = [_P, _U32, _U32, _P]
    lib.oracle_hash_build_bricks.restype = _U32
    lib.oracle_hash_build_bricks.argtypes = [_P, _U32, _P, _U32]
    lib.oracle_hash_get_brick.restype = _U32
    lib.oracle_hash_get_brick.argtypes = [_P, _U32, _U32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P]
    _lib = lib
    return lib


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def morton3(x: int, y: int, z: int) -> int:
    return int(load().oracle_morton3(x, y, z))


def radix_sort_pairs(keys: np.ndarray, values: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """The reference's 8-pass 4-bit stable LSD sort; returns sorted copies."""
    k0 = np.ascontiguousarray(keys, dtype=np.uint32).copy()
    v0 = np.ascontiguousarray(values, dtype=np.uint32).copy()
    k1 = np.empty_like(k0)
    v1 = np.empty_like(v0)
    load().oracle_radix_sort_pairs(_ptr(k0), _ptr(v0), _ptr(k1), _ptr(v1), len(k0))
    return k0, v0


def light_bvh_build(leaves_u8: np.ndarray, leaf_capacity: int, node_capacity: int, state_u8: np.ndarray,
                    header_u8: np.ndarray | None = None, nodes_u8: np.ndarray | None = None) -> dict:
    """Runs the reference build sequence on copies of the inputs. `leaves_u8` holds leaf_capacity
    LightBvhLeaf records as bytes; `state_u8` the LightTreeState (building, leaves_found, builds).
    Returns raw byte arrays: header, state, nodes, leaves, scratch."""
    leaves = np.zeros(leaf_capacity * LEAF_BYTES, np.uint8)
    leaves[: len(leaves_u8)] = leaves_u8
    state = np.ascontiguousarray(state_u8, dtype=np.uint8).copy()
    header = np.zeros(HEADER_BYTES, np.uint8) if header_u8 is None else header_u8.copy()
    nodes = np.zeros(node_capacity * NODE_BYTES, np.uint8) if nodes_u8 is None else nodes_u8.copy()
    scratch = np.zeros(SCRATCH_SORT_STATE * leaf_capacity, np.uint8)
    load().oracle_lt_build(_ptr(leaves), leaf_capacity, node_capacity, _ptr(header), _ptr(state), _ptr(scratch),
                           _ptr(nodes))
    return {"header": header, "state": state, "nodes": nodes, "leaves": leaves, "scratch": scratch}


def splitmix64(x: int) -> int:
    return int(load().oracle_splitmix64(x))


def hash_fingerprint(words: np.ndarray) -> np.ndarray:
    """words [count, word_count] uint32 -> [count, 2] (key_lo, key_hi)."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    out = np.zeros((w.shape[0], 2), np.uint32)
    load().oracle_hash_fingerprint(_ptr(w), w.shape[1], w.shape[0], _ptr(out))
    return out


def hash_build_bricks(keys_xyz: np.ndarray, slot_count: int) -> tuple[np.ndarray, int]:
    """Sequential ordered insert; returns (slots [slot_count, 4] uint32, longest displacement)."""
    k = np.ascontiguousarray(keys_xyz, dtype=np.int32)
    slots = np.zeros((slot_count, 4), np.uint32)
    longest = load().oracle_hash_build_bricks(_ptr(k), k.shape[0], _ptr(slots), slot_count)
    return slots, int(longest)


def hash_get_brick(slots: np.ndarray, probe_limit: int, x: int, y: int, z: int) -> tuple[int, int]:
    probes = ctypes.c_uint32()
    s = np.ascontiguousarray(slots, dtype=np.uint32)
    payload = load().oracle_hash_get_brick(_ptr(s), s.shape[0], probe_limit, x, y, z, ctypes.byref(probes))
    return int(payload), int(probes.value)
