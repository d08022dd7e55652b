"""Python binding of libcaelix_amd.so (the MI355X-native Caelix hot path) over its C ABI.

torch is used only as plumbing: device memory (uint8 tensors) and the current HIP stream. It is
imported before the library is loaded, so that the library binds to the HIP runtime torch already
loaded (both carry the soname libamdhip64.so.7) and device pointers are shared. A missing library is
an error: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import layout

PACKAGE_DIR = Path(__file__).resolve().parent.parent
LIBRARY_PATH = PACKAGE_DIR / "libcaelix_amd.so"

CX_OK = 0

_lib = None


class CxError(RuntimeError):
    pass


def _torch():
    import torch  # noqa: F401  (loads torch's HIP runtime first)
    return torch


class CxLightBvhBuild(ctypes.Structure):
    _fields_ = [
        ("leaves", ctypes.c_void_p),
        ("nodes", ctypes.c_void_p),
        ("header", ctypes.c_void_p),
        ("state", ctypes.c_void_p),
        ("scratch", ctypes.c_void_p),
        ("leaf_capacity", ctypes.c_uint32),
        ("node_capacity", ctypes.c_uint32),
    ]


class CxHashTable(ctypes.Structure):
    _fields_ = [
        ("keys", ctypes.c_void_p),
        ("stack", ctypes.c_void_p),
        ("counters", ctypes.c_void_p),
        ("payloads", ctypes.c_void_p),
        ("slot_count", ctypes.c_uint32),
        ("payload_stride", ctypes.c_uint32),
        ("probe_limit", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
    ]


# Exported C symbols and their signatures (include/caelix_amd/caelix_amd.h).
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_INT = ctypes.c_int
SIGNATURES = {
    "cx_set_log_callback": (None, [_P, _P]),
    "cx_version_string": (ctypes.c_char_p, []),
    "cx_light_bvh_scratch_bytes": (_U64, [_U32]),
    "cx_light_bvh_scratch_init": (_INT, [_P, _U32, _P]),
    "cx_light_bvh_begin": (_INT, [_P, _P, _U32, _U32, _U32, _P]),
    "cx_light_bvh_build": (_INT, [ctypes.POINTER(CxLightBvhBuild), _P]),
    "cx_radix_sort_state_bytes": (_U64, [_U32]),
    "cx_radix_sort_pairs": (_INT, [_P, _P, _P, _P, _P, _P, _U32, _P, _P]),
    "cx_hash_fingerprint": (_INT, [_P, _U32, _U32, _P, _P]),
    "cx_hash_lookup": (_INT, [ctypes.POINTER(CxHashTable), _P, _U32, _U32, _U32, _U32, _P, _P, _P]),
    "cx_hash_maintain": (_INT, [ctypes.POINTER(CxHashTable), _U32, _U32, _U32, _U32, _U32, _U32, _P, _P, _P]),
    "cx_hash_build_scratch_bytes": (_U64, [_U32]),
    "cx_hash_build_bricks": (_INT, [_P, _U32, _P, _U32, _P, _P, _P]),
    "cx_hash_get_bricks": (_INT, [_P, _U32, _P, _P, _U32, _P, _P, _P]),
    "cx_profile_enable": (None, [_INT]),
    "cx_profile_reset": (None, []),
    "cx_profile_read": (_INT, [_U32, _P, _P, _P, _P]),
}


def load(path: Path | str | None = None):
    """Loads libcaelix_amd.so once (after torch) and declares the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    _torch()
    lib_path = Path(path) if path is not None else LIBRARY_PATH
    if not lib_path.exists():
        raise CxError(f"{lib_path} is missing: build it with caelix-renderer_amd/build.py "
                      "(there is no CPU fallback)")
    lib = ctypes.CDLL(str(lib_path))
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


def check(code: int, what: str) -> None:
    if code != CX_OK:
        raise CxError(f"{what} failed with code {code}")


def stream_handle(stream=None) -> int:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def device_bytes(nbytes: int, device="cuda"):
    torch = _torch()
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def to_device(array: np.ndarray, device="cuda"):
    torch = _torch()
    raw = np.ascontiguousarray(array).view(np.uint8).reshape(-1)
    return torch.from_numpy(raw.copy()).to(device)


def to_host(tensor, dtype=np.uint8) -> np.ndarray:
    raw = tensor.detach().cpu().numpy().view(np.uint8).reshape(-1)
    return raw.view(dtype)


class LightBvhBuilder:
    """Device buffers of one light tree and its build (cx_light_bvh_build), sized for a leaf
    capacity: leaves, nodes, header, state, scratch. Mirrors the buffers LightTree owns for one tree
    (lib/src/light/light_tree.h:37-101)."""

    def __init__(self, leaf_capacity: int, node_capacity: int | None = None, device="cuda"):
        if leaf_capacity <= 0 or leaf_capacity % 64 != 0:
            raise ValueError("leaf capacity must be a positive multiple of 64")
        self.lib = load()
        self.leaf_capacity = leaf_capacity
        self.node_capacity = node_capacity if node_capacity is not None else layout.node_capacity_for(leaf_capacity)
        self.leaves = device_bytes(leaf_capacity * 64, device)
        self.nodes = device_bytes(self.node_capacity * 224, device)
        self.header = device_bytes(64, device)
        self.state = device_bytes(64, device)
        self.scratch_bytes = int(self.lib.cx_light_bvh_scratch_bytes(leaf_capacity))
        self.scratch = device_bytes(self.scratch_bytes, device)
        self.header.zero_()
        self.state.zero_()
        self.nodes.zero_()
        check(self.lib.cx_light_bvh_scratch_init(self.scratch.data_ptr(), leaf_capacity, stream_handle()),
              "cx_light_bvh_scratch_init")
        self._desc = CxLightBvhBuild(self.leaves.data_ptr(), self.nodes.data_ptr(), self.header.data_ptr(),
                                     self.state.data_ptr(), self.scratch.data_ptr(), leaf_capacity,
                                     self.node_capacity)

    def upload(self, leaves: np.ndarray, building: int = 1) -> None:
        """Writes the leaves (LEAF records) and sets state.leaves_found / building."""
        n = len(leaves)
        if n > 0:
            self.leaves[: n * 64].copy_(to_device(leaves.view(np.uint8), self.leaves.device))
        state = np.zeros(1, layout.STATE)
        state["building"] = building
        state["leaves_found"] = n
        host = to_host(self.state).view(layout.STATE).copy()
        host["building"] = building
        host["leaves_found"] = n
        self.state.copy_(to_device(host.view(np.uint8), self.state.device))

    def build(self, stream=None) -> None:
        check(self.lib.cx_light_bvh_build(ctypes.byref(self._desc), stream_handle(stream)), "cx_light_bvh_build")

    def read(self) -> dict:
        """Copies every output back: header, state, nodes, leaves, scratch (numpy)."""
        _torch().cuda.synchronize()
        return {
            "header": to_host(self.header).view(layout.HEADER)[0].copy(),
            "state": to_host(self.state).view(layout.STATE)[0].copy(),
            "nodes": to_host(self.nodes).view(layout.NODE).copy(),
            "leaves": to_host(self.leaves).view(layout.LEAF).copy(),
            "scratch": to_host(self.scratch).copy(),
        }


def radix_sort_pairs(keys, values, stream=None):
    """Sorts device uint32 (key, value) tensors in place (stable by key) with the onesweep kernels."""
    torch = _torch()
    lib = load()
    n = keys.numel()
    capacity = max(n, 1)
    keys1 = torch.empty_like(keys)
    values1 = torch.empty_like(values)
    count = torch.tensor([n], dtype=torch.int32, device=keys.device)
    state = device_bytes(int(lib.cx_radix_sort_state_bytes(capacity)), keys.device)
    check(lib.cx_radix_sort_pairs(keys.data_ptr(), values.data_ptr(), keys1.data_ptr(), values1.data_ptr(),
                                  count.data_ptr(), None, capacity, state.data_ptr(), stream_handle(stream)),
          "cx_radix_sort_pairs")


def profile_enable(enabled: bool) -> None:
    load().cx_profile_enable(1 if enabled else 0)


def profile_reset() -> None:
    load().cx_profile_reset()


def profile_read() -> dict:
    lib = load()
    max_stages = 64
    names = (ctypes.c_char_p * max_stages)()
    ms = (ctypes.c_double * max_stages)()
    launches = (ctypes.c_uint32 * max_stages)()
    count = ctypes.c_uint32()
    check(lib.cx_profile_read(max_stages, names, ms, launches, ctypes.byref(count)), "cx_profile_read")
    return {names[i].decode(): (ms[i], launches[i]) for i in range(min(count.value, max_stages))}


class BrickHashTable:
    """A bulk-built brick table (cx_hash_build_bricks) over int32 (x, y, z) brick keys at a load
    factor: slot_count = ceil(count / load)."""

    def __init__(self, keys_xyz, load_factor: float = 0.7, slot_count: int | None = None, stream=None):
        torch = _torch()
        self.lib = load_library = load()
        self.count = int(keys_xyz.shape[0])
        self.slot_count = (int(slot_count) if slot_count is not None
                           else max(int(np.ceil(self.count / load_factor)), self.count + 1))
        device = keys_xyz.device
        self.keys_xyz = keys_xyz.contiguous()
        self.slots = torch.empty(self.slot_count * 4, dtype=torch.int32, device=device)
        self.info = torch.zeros(4, dtype=torch.int32, device=device)
        scratch = device_bytes(int(load_library.cx_hash_build_scratch_bytes(self.count)), device)
        check(load_library.cx_hash_build_bricks(self.keys_xyz.data_ptr(), self.count, self.slots.data_ptr(),
                                                self.slot_count, scratch.data_ptr(), self.info.data_ptr(),
                                                stream_handle(stream)), "cx_hash_build_bricks")
        self._scratch = scratch

    def get(self, keys_xyz, with_probes: bool = False, stream=None):
        torch = _torch()
        n = int(keys_xyz.shape[0])
        payloads = torch.empty(n, dtype=torch.int32, device=keys_xyz.device)
        probes = torch.empty(n, dtype=torch.int32, device=keys_xyz.device) if with_probes else None
        check(self.lib.cx_hash_get_bricks(self.slots.data_ptr(), self.slot_count, self.info.data_ptr(),
                                          keys_xyz.contiguous().data_ptr(), n, payloads.data_ptr(),
                                          probes.data_ptr() if probes is not None else None,
                                          stream_handle(stream)), "cx_hash_get_bricks")
        return payloads, probes


class DynamicHashTable:
    """The reference's table (HashLookupOrInsert, maintenance) over a payload pool of `capacity`
    payloads of `stride` bytes: power-of-two key slots, 4 per payload (hash_table.cpp:47), 8-probe
    window."""

    def __init__(self, capacity: int, key_words: int, stride: int | None = None, device="cuda"):
        torch = _torch()
        self.lib = load()
        self.capacity = capacity
        self.key_words = key_words
        self.stride = stride if stride is not None else 8 + 4 * key_words + 4
        self.stride = (self.stride + 3) // 4 * 4
        slots = 1
        while slots < 4 * capacity:
            slots *= 2
        self.slot_count = slots
        self.keys = torch.zeros(slots * 4, dtype=torch.int32, device=device)
        self.stack = torch.zeros(1 + capacity, dtype=torch.int32, device=device)
        self.counters = torch.zeros(16, dtype=torch.int32, device=device)
        self.payloads = torch.zeros(capacity * self.stride // 4, dtype=torch.int32, device=device)
        self.active = torch.zeros(capacity, dtype=torch.int32, device=device)
        self.dispatch = torch.zeros(4, dtype=torch.int32, device=device)
        self.desc = CxHashTable(self.keys.data_ptr(), self.stack.data_ptr(), self.counters.data_ptr(),
                                self.payloads.data_ptr(), slots, self.stride, 8, 0)
        self.initialised = False

    def maintain(self, frame: int, max_age: int = 128, items_per_group: int = 64):
        check(self.lib.cx_hash_maintain(ctypes.byref(self.desc), self.capacity, self.key_words, frame, max_age,
                                        items_per_group, 0 if self.initialised else 1, self.active.data_ptr(),
                                        self.dispatch.data_ptr(), stream_handle()), "cx_hash_maintain")
        self.initialised = True
        return self.counters_host()

    def lookup(self, words, frame: int, insert: bool):
        """words: int32 tensor [count, key_words]. Returns (results, payloads) tensors."""
        torch = _torch()
        n = int(words.shape[0])
        results = torch.empty(n, dtype=torch.int32, device=words.device)
        payloads = torch.empty(n, dtype=torch.int32, device=words.device)
        check(self.lib.cx_hash_lookup(ctypes.byref(self.desc), words.contiguous().data_ptr(), self.key_words, n,
                                      frame, 1 if insert else 0, results.data_ptr(), payloads.data_ptr(),
                                      stream_handle()), "cx_hash_lookup")
        return results, payloads

    def counters_host(self) -> dict:
        c = to_host(self.counters, np.uint32)
        names = ["live", "active", "inserted", "pending", "failed_slot", "failed_payload", "evicted",
                 "rebuild_dropped"]
        return {name: int(c[i]) for i, name in enumerate(names)}

    def top(self) -> int:
        return int(to_host(self.stack, np.int32)[0])
