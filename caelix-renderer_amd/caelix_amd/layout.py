"""numpy views of the GPU layouts in include/caelix_amd/layout.h (byte-identical to the reference's
lib/src/light/light_tree_layout.h and lib/src/rhi/hash_table_layout.h)."""
from __future__ import annotations

import numpy as np

LT_NONE = 0xFFFFFFFF
LT_CHILD_LEAF_BIT = 0x80000000

SUMMARY = np.dtype([("face", "<f4", (6,)), ("lo", "<f4", (3,)), ("hi", "<f4", (3,))])
assert SUMMARY.itemsize == 48

LEAF = np.dtype([("summary", SUMMARY), ("group", "<u4"), ("entry", "<u4"), ("pad", "<u4"), ("parent", "<u4")])
assert LEAF.itemsize == 64

NODE = np.dtype([
    ("child_face", "<f4", (4, 6)),
    ("child_lo", "<f4", (4, 3)),
    ("child_hi", "<f4", (4, 3)),
    ("child", "<u4", (4,)),
    ("parent", "<u4"),
    ("parent_slot", "<u4"),
    ("pad", "<u4", (2,)),
])
assert NODE.itemsize == 224

HEADER = np.dtype([("summary", SUMMARY), ("root", "<u4"), ("leaf_count", "<u4"), ("node_count", "<u4"),
                   ("pad", "<u4")])
assert HEADER.itemsize == 64

STATE = np.dtype([("building", "<u4"), ("leaves_found", "<u4"), ("leaves_dropped", "<u4"), ("builds", "<u4"),
                  ("node_overflow", "<u4"), ("build_error", "<u4"), ("pad0", "<u4", (2,)), ("bounds", "<f4", (6,)), ("pad1", "<u4", (2,))])
assert STATE.itemsize == 64

HASH_SLOT = np.dtype([("key_lo", "<u4"), ("key_hi", "<u4"), ("payload", "<u4"), ("pad", "<u4")])
assert HASH_SLOT.itemsize == 16

# Build scratch array offsets (bytes per leaf of capacity), light_tree_layout.h:403-423.
SCRATCH_KEYS = 0
SCRATCH_VALUES = 4
SCRATCH_SORT_BUFFER_STRIDE = 8
SCRATCH_LEFT = 16
SCRATCH_RIGHT = 20
SCRATCH_PARENT = 24
SCRATCH_SORTED_LEAF_PARENT = 28
SCRATCH_COUNTER = 32
SCRATCH_WIDE = 36
SCRATCH_AGGREGATES = 40
SCRATCH_SORT_STATE = 88


def node_capacity_for(leaf_capacity: int) -> int:
    """lib/src/light/light_tree.cpp:25-28."""
    return leaf_capacity // 8 * 5 + 64


def scratch_u32(scratch: np.ndarray, offset: int, capacity: int) -> np.ndarray:
    """Array `offset` (SCRATCH_*) of a build scratch held as a uint8 numpy array."""
    return scratch[offset * capacity:(offset + 4) * capacity].view("<u4")


def scratch_aggregates(scratch: np.ndarray, capacity: int) -> np.ndarray:
    return scratch[SCRATCH_AGGREGATES * capacity:(SCRATCH_AGGREGATES + 48) * capacity].view(SUMMARY)


# Block value encodings (record_layout.h:27-34): packed and block-ID, opaque bits and id masks.
ENCODING_PACKED = 0
ENCODING_BLOCK_ID = 1
PACKED_OPAQUE_MASK = 0xC000
PACKED_ID_MASK = 0xFF
BLOCK_ID_OPAQUE_MASK = 0x8000
BLOCK_ID_ID_MASK = 0x1FF

# The light tree's tables (light_tree_layout.h:138-389).
CONTROL = np.dtype([("current", "<u4"), ("pass_result", "<u4"), ("needed_units", "<u4"), ("pad", "<u4"),
                    ("light_groups", "<u4"), ("bricks_dropped", "<u4"), ("light_bricks", "<u8"), ("light_faces", "<u8"),
                    ("passes", "<u4"), ("failed_passes", "<u4"), ("pass_flags", "<u4"), ("records_recomputed", "<u4"),
                    ("records_copied", "<u4"), ("trees_changed", "<u4"), ("headers", HEADER, (2,)),
                    ("states", STATE, (2,))])
assert CONTROL.itemsize == 320 and CONTROL.fields["headers"][1] == 64 and CONTROL.fields["states"][1] == 192
ROW = np.dtype([("block_offset", "<u4"), ("block_units", "<u4"), ("generation", "<u4"), ("gpu_flags", "<u4"),
                ("cpu_flags", "<u4"), ("bricks", "<u4"), ("light_faces", "<u4"), ("pad", "<u4")])
assert ROW.itemsize == 32
DESCRIPTOR = np.dtype([("control", "<u8"), ("heap", "<u8", (2,)), ("group_rows", "<u8"), ("group_infos", "<u8"),
                       ("entity_infos", "<u8"), ("entity_motion", "<u8"), ("grids", "<u8"), ("nodes", "<u8", (2,)),
                       ("leaves", "<u8", (2,)), ("light_codes", "<u8"), ("tree_headers", "<u8"), ("group_shift", "<u4"),
                       ("encoding", "<u4"), ("pad", "<u4"), ("group_count", "<u4")])
assert DESCRIPTOR.itemsize == 128 and DESCRIPTOR.fields["group_count"][1] == 124
PASS_NO_CHANGE, PASS_OK, PASS_FAILED = 2, 1, 0
CPU_REBUILD, CPU_RELEASED, CPU_FRESH = 1, 2, 4
GPU_LIGHT_CHANGED = 1
RECORD_STALE_BIT = 1 << 30
