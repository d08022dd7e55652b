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


class CxScene(ctypes.Structure):
    _fields_ = [
        ("keys", ctypes.c_void_p), ("info", ctypes.c_void_p), ("occupancy", ctypes.c_void_p),
        ("payload", ctypes.c_void_p), ("slots", ctypes.c_void_p), ("slot_count", ctypes.c_uint32),
        ("bricks", ctypes.c_uint32), ("hash_info", ctypes.c_void_p), ("brick_bits", ctypes.c_void_p),
        ("super_bits", ctypes.c_void_p), ("grid_min", ctypes.c_int32 * 3), ("grid_dim", ctypes.c_uint32),
        ("group_shift", ctypes.c_uint32), ("encoding", ctypes.c_uint32), ("materials", ctypes.c_void_p),
        ("light_codes", ctypes.c_void_p), ("emission_scale", ctypes.c_float), ("pad", ctypes.c_uint32),
        ("brick_emission_scale", ctypes.c_void_p), ("scene_count", ctypes.c_uint32), ("pad2", ctypes.c_uint32),
        ("occupied_box", ctypes.c_void_p), ("materials_ext", ctypes.c_void_p), ("group_of_brick", ctypes.c_void_p),
        ("light_group_rows", ctypes.c_void_p), ("subsurface", ctypes.c_void_p),
    ]


class CxLightPass(ctypes.Structure):
    _fields_ = [("descriptor", ctypes.c_void_p), ("heap_units", ctypes.c_uint64), ("staged_flags", ctypes.c_void_p),
                ("scratch", ctypes.c_void_p), ("group_count", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class CxLightTreeView(ctypes.Structure):
    _fields_ = [("descriptor", ctypes.c_void_p), ("mis", ctypes.c_uint32), ("select", ctypes.c_uint32)]


class CxLightOptions(ctypes.Structure):
    _fields_ = [("enabled", ctypes.c_uint32), ("min_emission", ctypes.c_float), ("heap_bytes", ctypes.c_uint64)]


class CxLightStats(ctypes.Structure):
    _U = ctypes.c_uint32
    _fields_ = [("enabled", _U), ("light_groups", _U), ("light_bricks", ctypes.c_uint64),
                ("light_faces", ctypes.c_uint64), ("bricks_dropped", ctypes.c_uint64),
                ("heap_bytes_used", ctypes.c_uint64), ("heap_bytes_capacity", ctypes.c_uint64), ("passes", _U),
                ("failed_passes", _U), ("records_recomputed", _U), ("records_copied", _U), ("gpu_ms", ctypes.c_float),
                ("pad", _U), ("tree_leaves", _U * 2), ("tree_nodes", _U * 2), ("tree_builds", _U * 2),
                ("tree_leaves_dropped", _U * 2)]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class CxLightSceneTables(ctypes.Structure):
    _fields_ = [("group_infos", ctypes.c_void_p), ("entity_infos", ctypes.c_void_p),
                ("group_capacity", ctypes.c_uint32), ("entity_capacity", ctypes.c_uint32),
                ("brick_capacity", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class CxCamera(ctypes.Structure):
    _fields_ = [
        ("position", ctypes.c_float * 3), ("forward", ctypes.c_float * 3), ("right", ctypes.c_float * 3),
        ("up", ctypes.c_float * 3), ("tan_half_fov_y", ctypes.c_float), ("aspect", ctypes.c_float),
        ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
    ]


class CxTraceParams(ctypes.Structure):
    _fields_ = [("bounces", ctypes.c_uint32), ("frame_seed", ctypes.c_uint32), ("spp", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("sun_dir", ctypes.c_float * 3), ("sun_solid_angle", ctypes.c_float),
                ("sun_radiance", ctypes.c_float * 3), ("sun_cos_radius", ctypes.c_float),
                ("sky_zenith", ctypes.c_float * 3), ("pad0", ctypes.c_float), ("sky_horizon", ctypes.c_float * 3),
                ("pad1", ctypes.c_float), ("sky_ground", ctypes.c_float * 3), ("pad2", ctypes.c_float),
                ("fog_color", ctypes.c_float * 3), ("fog_density", ctypes.c_float),
                ("bounces_transparent", ctypes.c_uint32), ("bounce_nee_probability", ctypes.c_float),
                ("hint", ctypes.c_void_p), ("sky_cubemap", ctypes.c_void_p)]


class CxSkyCubemap(ctypes.Structure):
    _fields_ = [("texels", ctypes.c_void_p), ("size", ctypes.c_uint32), ("mip_levels", ctypes.c_uint32)]


class CxLightHint(ctypes.Structure):
    _fields_ = [("previous", ctypes.c_void_p), ("current", ctypes.c_void_p), ("read", ctypes.c_void_p),
                ("previous_camera", ctypes.POINTER(CxCamera)), ("patch_shift", ctypes.c_uint32),
                ("reset", ctypes.c_uint32)]


assert ctypes.sizeof(CxTraceParams) == 136 and ctypes.sizeof(CxLightHint) == 40 and ctypes.sizeof(CxSkyCubemap) == 16


def cube_texel_direction(face: int, s, t):
    """The direction of face coordinates (s, t) in [-1, 1] (the reference's CubeTexelDirection): +X
    (1, -t, -s), -X (-1, -t, s), +Y (s, 1, t), -Y (s, -1, -t), +Z (s, -t, 1), -Z (-s, -t, -1)."""
    one = np.ones_like(np.asarray(s, np.float64))
    return [np.stack([one, -t, -s], -1), np.stack([-one, -t, s], -1), np.stack([s, one, t], -1),
            np.stack([s, -one, -t], -1), np.stack([s, -t, one], -1), np.stack([-s, -t, -one], -1)][face]


def make_sky_cube(size: int, radiance, mip_levels: int = 1) -> np.ndarray:
    """Cubemap texels in the CxSkyCubemap layout ([texels, 4] float32): mip 0 = radiance(unit direction of
    each texel centre) (a callable on [..., 3] arrays returning [..., 3]), each further mip the 2 x 2
    average of the one above."""
    faces = []
    c = (np.arange(size) + 0.5) / size * 2.0 - 1.0
    s, t = np.meshgrid(c, c, indexing="xy")  # [y, x]
    for f in range(6):
        d = cube_texel_direction(f, s, t)
        d = d / np.linalg.norm(d, axis=-1, keepdims=True)
        faces.append(np.asarray(radiance(d), np.float64))
    mip = np.stack(faces)  # [6, y, x, 3]
    out = []
    for m in range(mip_levels):
        rgba = np.concatenate([mip, np.ones(mip.shape[:-1] + (1,))], axis=-1)
        out.append(rgba.reshape(-1, 4))
        if m + 1 < mip_levels:
            e = mip.shape[1]
            if e > 1:
                mip = mip.reshape(6, e // 2, 2, e // 2, 2, 3).mean(axis=(2, 4))
    return np.concatenate(out).astype(np.float32)

HINT_PATCHES = {1: 0, 2: 1, 4: 2}
DEFAULT_HINT_PATCH = 4  # the reference renderer's kDefaultHintPatch


class LightHint:
    """The light NEE's visibility hint images of one view, as the reference renderer keeps them
    (renderer_light.cpp / docs of the light NEE section 12.1): two images of 16-bit words (sun | lamp << 8,
    k / 255) per patch of `patch` pixels, ping-ponged between frames; the first frame after a creation
    runs with the reset (every primary vertex reads 0x8080); the previous frame's camera reprojects the
    primary surfaces. `read` (optional, per pixel): the word each primary vertex read."""

    def __init__(self, width: int, height: int, patch: int = DEFAULT_HINT_PATCH, keep_read: bool = False):
        if patch not in HINT_PATCHES:
            raise ValueError("hint patch must be 1, 2 or 4")
        torch = _torch()
        self.width, self.height, self.patch = int(width), int(height), int(patch)
        self.shift = HINT_PATCHES[self.patch]
        self.columns = (self.width + self.patch - 1) >> self.shift
        self.rows = (self.height + self.patch - 1) >> self.shift
        self.images = [torch.full((self.rows * self.columns,), 0x8080 - 0x10000, dtype=torch.int16, device="cuda")
                       for _ in range(2)]
        self.read = torch.zeros(self.width * self.height, dtype=torch.int16, device="cuda") if keep_read else None
        self.parity = 0
        self.history = False
        self.camera = None
        self._keep = None

    def frame_struct(self, camera: dict) -> CxLightHint:
        """The hint of the next frame (kept alive until `advance`)."""
        prev = camera_struct(self.camera) if self.camera is not None else None
        h = CxLightHint(self.images[self.parity].data_ptr(), self.images[1 - self.parity].data_ptr(),
                        self.read.data_ptr() if self.read is not None else None,
                        ctypes.pointer(prev) if prev is not None else None, self.shift, 0 if self.history else 1)
        self._keep = (h, prev)
        return h

    def advance(self, camera: dict) -> None:
        self.parity ^= 1
        self.history = True
        self.camera = dict(camera)

    def current(self) -> np.ndarray:
        """[rows, columns, 2] uint8 of the newest image: (sun, lamp)."""
        w = self.images[self.parity].cpu().numpy().view(np.uint16).reshape(self.rows, self.columns)
        return np.stack([w & 0xFF, w >> 8], axis=-1).astype(np.uint8)

    def read_words(self) -> np.ndarray:
        return self.read.cpu().numpy().view(np.uint16).reshape(self.height, self.width)


TRACE_NO_RUSSIAN_ROULETTE = 1
TRACE_SKY = 2
TRACE_NO_SUN_NEE = 4
TRACE_SHADOW_TRANSMISSION = 8
TRACE_DELTA_CHECKERBOARD = 16
TRACE_SUBSURFACE = 32


def sky_params(sun_dir=(0.0, 1.0, 0.0), sun_radiance=(0.0, 0.0, 0.0), sun_radius_deg: float = 0.5,
               zenith=(0.0, 0.0, 0.0), horizon=(0.0, 0.0, 0.0), ground=(0.0, 0.0, 0.0)) -> dict:
    """The sky fields of CxTraceParams: sun_dir normalised, the cone's solid angle 2 pi (1 - cos r) and
    cos r computed in double (as the reference's host does)."""
    d = np.asarray(sun_dir, np.float64)
    d = d / np.linalg.norm(d)
    r = np.radians(sun_radius_deg)
    return {"sun_dir": [float(v) for v in d], "sun_radiance": [float(v) for v in sun_radiance],
            "sun_solid_angle": float(2.0 * np.pi * (1.0 - np.cos(r))), "sun_cos_radius": float(np.cos(r)),
            "sky_zenith": [float(v) for v in zenith], "sky_horizon": [float(v) for v in horizon],
            "sky_ground": [float(v) for v in ground]}


def apply_sky(params: CxTraceParams, sky: dict | None, fog=None) -> None:
    """Sets the CX_TRACE_SKY flag and the sky fields of `params` (sky None: black sky, no sun) and
    the fog of the air, fog = (rgb radiance, density) (None: no fog)."""
    if fog is not None:
        params.fog_color[:] = [float(v) for v in fog[0]]
        params.fog_density = float(fog[1])
    if sky is None:
        return
    params.flags |= TRACE_SKY
    for name in ("sun_dir", "sun_radiance", "sky_zenith", "sky_horizon", "sky_ground"):
        getattr(params, name)[:] = sky[name]
    params.sun_solid_angle = sky["sun_solid_angle"]
    params.sun_cos_radius = sky["sun_cos_radius"]


class CxTraceGradients(ctypes.Structure):
    _fields_ = [("d_color", ctypes.c_void_p), ("grad_albedo", ctypes.c_void_p), ("grad_scale", ctypes.c_void_p)]


class CxTraceOutput(ctypes.Structure):
    _fields_ = [("color", ctypes.c_void_p), ("radiance", ctypes.c_void_p), ("attrib", ctypes.c_void_p),
                ("counters", ctypes.c_void_p), ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_uint64)]


def camera_struct(cam: dict) -> CxCamera:
    c = CxCamera()
    for name in ("position", "forward", "right", "up"):
        getattr(c, name)[:] = [float(v) for v in cam[name]]
    c.tan_half_fov_y = float(cam["tan_half_fov_y"])
    c.aspect = float(cam["aspect"])
    c.width = int(cam["width"])
    c.height = int(cam["height"])
    return c


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
    "cx_light_bvh_build_static": (_INT, [ctypes.POINTER(CxLightBvhBuild), _P]),
    "cx_light_bvh_batch_bytes": (_U64, [_U32]),
    "cx_light_bvh_batch_prepare": (_INT, [ctypes.POINTER(CxLightBvhBuild), _U32, _P, _P]),
    "cx_light_bvh_build_batch": (_INT, [_P, _U32, _U32, _P]),
    "cx_light_bvh_build_keys64": (_INT, [ctypes.POINTER(CxLightBvhBuild), _P]),
    "cx_light_bvh_set_tuning": (_INT, [_P]),
    "cx_light_bvh_get_tuning": (None, [_P]),
    "cx_radix_sort_state_bytes": (_U64, [_U32]),
    "cx_radix_sort_pairs": (_INT, [_P, _P, _P, _P, _P, _P, _U32, _P, _P]),
    "cx_radix_sort_keys64_state_bytes": (_U64, [_U32]),
    "cx_radix_sort_keys64": (_INT, [_P, _P, _P, _P, _U32, _P, _P]),
    "cx_hash_fingerprint": (_INT, [_P, _U32, _U32, _P, _P]),
    "cx_hash_lookup": (_INT, [ctypes.POINTER(CxHashTable), _P, _U32, _U32, _U32, _U32, _P, _P, _P]),
    "cx_hash_maintain": (_INT, [ctypes.POINTER(CxHashTable), _U32, _U32, _U32, _U32, _U32, _U32, _P, _P, _P]),
    "cx_hash_build_scratch_bytes": (_U64, [_U32]),
    "cx_hash_build_bricks": (_INT, [_P, _U32, _P, _U32, _P, _P, _P]),
    "cx_hash_get_bricks": (_INT, [_P, _U32, _P, _P, _U32, _P, _P, _P]),
    "cx_scene_finalize": (_INT, [ctypes.POINTER(CxScene), _P]),
    "cx_scene_finalize_bricks": (_INT, [ctypes.POINTER(CxScene), _P, _U32, _P]),
    "cx_light_pass_scratch_bytes": (_U64, [_U32, _U32]),
    "cx_light_pass": (_INT, [ctypes.POINTER(CxScene), ctypes.POINTER(CxLightPass), _P]),
    "cx_light_gather": (_INT, [ctypes.POINTER(CxLightPass), _U32, _P, _U32, _P, _P]),
    "cx_light_gather_entities": (_INT, [ctypes.POINTER(CxLightPass), _U32, _P, _U32, _P, _P, _P]),
    "cx_light_tree_create": (_INT, [_U32, _U32, ctypes.POINTER(CxLightOptions), _P]),
    "cx_light_tree_destroy": (None, [_P]),
    "cx_light_tree_on_entity_added": (_INT, [_P, _U32, _U32]),
    "cx_light_tree_on_entity_removed": (_INT, [_P, _U32]),
    "cx_light_tree_on_entity_moved": (_INT, [_P, _U32]),
    "cx_light_tree_on_entity_motion": (_INT, [_P, _U32, _U32]),
    "cx_light_tree_on_group_created": (_INT, [_P, _U32]),
    "cx_light_tree_on_group_released": (_INT, [_P, _U32, _U32]),
    "cx_light_tree_on_group_bricks_removed": (_INT, [_P, _U32]),
    "cx_light_tree_on_materials_changed": (_INT, [_P]),
    "cx_light_tree_set_min_emission": (_INT, [_P, ctypes.c_float]),
    "cx_light_tree_prepare": (_INT, [_P, _P, ctypes.POINTER(CxLightSceneTables), _P]),
    "cx_light_tree_stage": (_INT, [_P, _U32, _P]),
    "cx_light_tree_record_transfers": (_INT, [_P, _P]),
    "cx_light_tree_record": (_INT, [_P, ctypes.POINTER(CxScene), _U64, _P]),
    "cx_light_tree_end_flush": (_INT, [_P, _U64]),
    "cx_light_tree_rebuild": (_INT, [_P, _U32, _P]),
    "cx_light_tree_recycle": (_INT, [_P, _U64]),
    "cx_light_tree_get_stats": (None, [_P, ctypes.POINTER(CxLightStats)]),
    "cx_light_tree_pending": (_U32, [_P]),
    "cx_light_tree_descriptor": (_P, [_P]),
    "cx_light_tree_light_codes": (_P, [_P]),
    "cx_light_tree_group_rows": (_P, [_P]),
    "cx_trace_workspace_bytes": (_U64, [_U32, _U32]),
    "cx_trace_frame": (_INT, [ctypes.POINTER(CxScene), ctypes.POINTER(CxLightTreeView), ctypes.POINTER(CxCamera),
                              ctypes.POINTER(CxTraceParams), ctypes.POINTER(CxTraceOutput), _P]),
    "cx_trace_frame_adjoint": (_INT, [ctypes.POINTER(CxScene), ctypes.POINTER(CxLightTreeView), ctypes.POINTER(CxCamera),
                                      ctypes.POINTER(CxTraceParams), ctypes.POINTER(CxTraceGradients),
                                      ctypes.POINTER(CxTraceOutput), _P]),
    "cx_trace_batch_bytes": (_U64, [_U32]),
    "cx_trace_batch_workspace_bytes": (_U64, [_U32, _U32, _U32]),
    "cx_trace_batch_prepare": (_INT, [ctypes.POINTER(CxLightTreeView), ctypes.POINTER(CxCamera), _U32, _P, _P]),
    "cx_trace_batch": (_INT, [ctypes.POINTER(CxScene), _P, _U32, _U32, _U32, ctypes.POINTER(CxTraceParams),
                              ctypes.POINTER(CxTraceOutput), _P]),
    "cx_hash_build_keys": (_INT, [_P, _U32, _U32, _P, _U32, _P, _P, _P]),
    "cx_trace_rays": (_INT, [ctypes.POINTER(CxScene), _P, _U32, _P, _P]),
    "cx_light_sample": (_INT, [ctypes.POINTER(CxLightTreeView), _P, _P, _U32, _P, _P, _P]),
    "cx_light_point_pdf": (_INT, [ctypes.POINTER(CxLightTreeView), _P, _U32, _P, _P]),
    "cx_light_mis_weight": (_INT, [_P, _U32, _P, _P]),
    "cx_sky_cube_sample": (_INT, [_P, _P, _U32, _P, _P]),
    "cx_sky_sh": (_INT, [_P, _P, _P]),
    "cx_light_sample_brick_box": (_INT, [ctypes.POINTER(CxLightTreeView), _P, _P, _U32, _P, _P]),
    "cx_light_brick_pdf": (_INT, [ctypes.POINTER(CxLightTreeView), _P, _U32, _U32, _P, _P]),
    "cx_light_box_support": (_INT, [ctypes.POINTER(CxLightTreeView), _P, _U32, _P, _P]),
    "cx_cxw_open": (_INT, [ctypes.c_char_p, _P, _P]),
    "cx_cxw_close": (None, [_P]),
    "cx_cxw_world": (_INT, [_P, _P]),
    "cx_cxw_entity": (_INT, [_P, _U32, _P]),
    "cx_cxw_sector": (_INT, [_P, _U64, _P]),
    "cx_cxw_decode": (_INT, [_P, _U64, _P, _P, _P, _P, _P, _P]),
    "cx_cxw_decode_all": (_INT, [_P, _U32, _P, _P, _P]),
    "cx_cxw_copy_decoded": (_INT, [_P, _P, _P, _P, _P, _U64]),
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

    def __init__(self, leaf_capacity: int, node_capacity: int | None = None, device="cuda", leaves=None, header=None):
        """`leaves`: an existing uint8 device tensor view of at least leaf_capacity * 64 bytes to build
        over (a scene's slice of a batch's leaves), else the builder allocates its own; `header`: a
        64-byte uint8 view to write the tree header into (the first header of a light tree view)."""
        if leaf_capacity <= 0 or leaf_capacity % 64 != 0:
            raise ValueError("leaf capacity must be a positive multiple of 64")
        self.lib = load()
        self.leaf_capacity = leaf_capacity
        self.node_capacity = node_capacity if node_capacity is not None else layout.node_capacity_for(leaf_capacity)
        if leaves is not None and leaves.numel() < leaf_capacity * 64:
            raise ValueError("the leaves view is smaller than the leaf capacity")
        self.leaves = leaves if leaves is not None else device_bytes(leaf_capacity * 64, device)
        self.nodes = device_bytes(self.node_capacity * 224, device)
        self.header = header if header is not None else device_bytes(64, device)
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

    def build(self, stream=None, keys64: bool = False) -> None:
        """The build; keys64 sorts code << 32 | leaf as one 64-bit key (the same tree)."""
        if keys64:
            check(self.lib.cx_light_bvh_build_keys64(ctypes.byref(self._desc), stream_handle(stream)),
                  "cx_light_bvh_build_keys64")
        else:
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


class LightBvhBatch:
    """Many light trees built by one sequence of launches (cx_light_bvh_build_batch): the builders'
    leaves, states and buffers stay theirs; build() rebuilds every tree whose state is building."""

    def __init__(self, builders, stream=None):
        self.lib = load()
        self.builders = list(builders)
        n = len(self.builders)
        array = (CxLightBvhBuild * n)(*[b._desc for b in self.builders])
        self.descriptors = device_bytes(int(self.lib.cx_light_bvh_batch_bytes(n)), "cuda")
        check(self.lib.cx_light_bvh_batch_prepare(array, n, self.descriptors.data_ptr(), stream_handle(stream)),
              "cx_light_bvh_batch_prepare")
        self.max_capacity = max(b.leaf_capacity for b in self.builders)

    def build(self, stream=None) -> None:
        check(self.lib.cx_light_bvh_build_batch(self.descriptors.data_ptr(), len(self.builders), self.max_capacity,
                                                stream_handle(stream)), "cx_light_bvh_build_batch")


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


def radix_sort_keys64(keys, stream=None):
    """Sorts a device int64 tensor of uint64 key bit patterns (bit 63 clear) in place, stable, with
    the 7-pass onesweep kernels."""
    torch = _torch()
    lib = load()
    n = keys.numel()
    capacity = max(n, 1)
    keys1 = torch.empty_like(keys)
    count = torch.tensor([n], dtype=torch.int32, device=keys.device)
    state = device_bytes(int(lib.cx_radix_sort_keys64_state_bytes(capacity)), keys.device)
    check(lib.cx_radix_sort_keys64(keys.data_ptr(), keys1.data_ptr(), count.data_ptr(), None, capacity,
                                   state.data_ptr(), stream_handle(stream)), "cx_radix_sort_keys64")
    keys.copy_(keys1)


def light_mis_weight(queries: np.ndarray, stream=None) -> np.ndarray:
    """cx_light_mis_weight: queries [n, 9] float32 = (p_area, d, cos_l, n_r, w) -> [n, 2] float32 =
    (w_nee, w_bsdf)."""
    torch = _torch()
    n = len(queries)
    q = torch.from_numpy(np.ascontiguousarray(queries, np.float32)).cuda()
    out = torch.empty(max(n, 1) * 2, dtype=torch.float32, device="cuda")
    check(load().cx_light_mis_weight(q.data_ptr(), n, out.data_ptr(), stream_handle(stream)), "cx_light_mis_weight")
    return out.cpu().numpy().reshape(-1, 2)[:n]


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


def groups_of(keys: np.ndarray, group_shift: int):
    """The render groups of bricks: (group coordinates [G, k], group_of_brick [n] uint32). A group holds the
    (1 << group_shift)^3 bricks of one group coordinate key >> group_shift (per coordinate, arithmetic);
    keys of 4 words (scene, x, y, z) keep the scene word. Groups are numbered in ascending lexicographic
    order of their coordinates (scene-major for a batch)."""
    k = np.ascontiguousarray(keys, np.int64).copy()
    k[:, -3:] >>= int(group_shift)
    coords, inverse = np.unique(k, axis=0, return_inverse=True)
    return coords.astype(np.int32), inverse.reshape(-1).astype(np.uint32)


GROUP_INFO = np.dtype([("coord", "<i4", (3,)), ("entity", "<u4"), ("slot_row_base", "<u4"), ("reverse_base", "<u4"),
                       ("aabb_row_base", "<u4"), ("high_water", "<u4")])
ENTITY_INFO = np.dtype([("grid_min", "<i4", (3,)), ("grid_base", "<u4"), ("grid_dims", "<i4", (3,)), ("life", "<u4"),
                        ("cur_rows", "<f4", (3, 4)), ("prev_rows", "<f4", (3, 4))])
assert GROUP_INFO.itemsize == 32 and ENTITY_INFO.itemsize == 128


def group_infos(coords: np.ndarray, entity_of_group=None) -> np.ndarray:
    g = np.zeros(len(coords), GROUP_INFO)
    g["coord"] = coords[:, -3:]
    g["entity"] = 0 if entity_of_group is None else np.asarray(entity_of_group, np.uint32)
    return g


def entity_infos(count: int, rows=None) -> np.ndarray:
    """`count` entities; rows [count, 3, 4] object-to-world (rotation | translation) or None: identity."""
    e = np.zeros(count, ENTITY_INFO)
    if rows is None:
        rows = np.tile(np.eye(3, 4, dtype=np.float32), (count, 1, 1))
    e["cur_rows"] = rows
    e["prev_rows"] = rows
    return e


def light_codes_for(materials: np.ndarray, min_emission: float, encoding: int) -> np.ndarray:
    """The light code set the tree computes (LightTree::RebuildCodes, light_tree.cpp:185-198): a code whose
    canonical code's emission luminance reaches min_emission. [2048] uint32."""
    codes = np.arange(65536, dtype=np.uint32)
    if encoding == layout.ENCODING_PACKED:
        opaque, ids = codes & np.uint32(layout.PACKED_OPAQUE_MASK), codes & np.uint32(layout.PACKED_ID_MASK)
    else:
        opaque, ids = codes & np.uint32(layout.BLOCK_ID_OPAQUE_MASK), codes & np.uint32(layout.BLOCK_ID_ID_MASK)
    canonical = np.where(opaque != 0, codes, ids)
    e = np.asarray(materials, np.float32).reshape(65536, 8)[canonical, 4:7]
    lum = np.float32(0.2126) * e[:, 0] + np.float32(0.7152) * e[:, 1] + np.float32(0.0722) * e[:, 2]
    bits = (lum >= np.float32(min_emission)).reshape(2048, 32).astype(np.uint32)
    return (bits << np.arange(32, dtype=np.uint32)).sum(axis=1, dtype=np.uint64).astype(np.uint32)


def min_emission_for(data) -> float:
    """The light threshold that makes the tree's code set the scene's light codes (the smallest emission
    luminance among them), checked on every canonical code; ValueError when no threshold does."""
    enc = int(getattr(data, "encoding", 1))
    want = np.asarray(data.light_codes, np.uint32)
    bits = ((want[:, None] >> np.arange(32, dtype=np.uint32)) & 1).reshape(-1).astype(bool)
    e = np.asarray(data.materials, np.float32).reshape(65536, 8)[:, 4:7]
    lum = np.float32(0.2126) * e[:, 0] + np.float32(0.7152) * e[:, 1] + np.float32(0.0722) * e[:, 2]
    if not bits.any():
        return 3.0e38
    threshold = float(lum[bits].min())
    if not threshold > 0.0:
        raise ValueError("a light code with zero emission")
    got = light_codes_for(data.materials, threshold, enc)
    got_bits = ((got[:, None] >> np.arange(32, dtype=np.uint32)) & 1).reshape(-1).astype(bool)
    codes = np.arange(65536, dtype=np.uint32)
    mask = np.uint32(layout.PACKED_OPAQUE_MASK if enc == layout.ENCODING_PACKED else layout.BLOCK_ID_OPAQUE_MASK)
    idm = np.uint32(layout.PACKED_ID_MASK if enc == layout.ENCODING_PACKED else layout.BLOCK_ID_ID_MASK)
    canonical = np.where((codes & mask) != 0, codes, codes & idm) == codes
    if not np.array_equal(got_bits[canonical], bits[canonical]):
        raise ValueError("the scene's light codes are not a luminance threshold of its materials")
    return threshold


_hip = None


def read_device(address: int, nbytes: int) -> np.ndarray:
    """Copies nbytes at a device address (memory the library allocated) to a uint8 array."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    out = np.empty(nbytes, np.uint8)
    if nbytes > 0:
        rc = _hip.hipMemcpy(out.ctypes.data, address, nbytes, 2)  # hipMemcpyDeviceToHost
        if rc != 0:
            raise CxError(f"hipMemcpy from the device failed ({rc})")
    return out


class LightTree:
    """The light tree's host side (cx_light_tree_*, the reference's LightTree): hooks, flush steps and
    stats over the library's adapter. flush() runs prepare, stage, the transfers, the scene's finalize
    (with the tree's codes and rows), record and end_flush; settle() flushes until nothing is pending (a
    failed pass retried after the heap grew, a tree rebuilt after its capacity grew)."""

    def __init__(self, group_shift: int, encoding: int, min_emission: float = 0.0, heap_bytes: int = 0):
        self.lib = load()
        self.handle = ctypes.c_void_p()
        opts = CxLightOptions(1, float(min_emission), int(heap_bytes))
        check(self.lib.cx_light_tree_create(int(group_shift), int(encoding), ctypes.byref(opts),
                                            ctypes.byref(self.handle)), "cx_light_tree_create")
        self.frame = 0

    def __del__(self):
        lib = getattr(self, "lib", None)
        if lib is not None and getattr(self, "handle", None) and self.handle.value:
            lib.cx_light_tree_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def hook(self, name: str, *args) -> None:
        check(getattr(self.lib, "cx_light_tree_" + name)(self.handle, *args), "cx_light_tree_" + name)

    @property
    def descriptor(self) -> int:
        return int(self.lib.cx_light_tree_descriptor(self.handle) or 0)

    @property
    def light_codes(self) -> int:
        return int(self.lib.cx_light_tree_light_codes(self.handle) or 0)

    @property
    def group_rows(self) -> int:
        return int(self.lib.cx_light_tree_group_rows(self.handle) or 0)

    def rebuild(self, tree: int = 0, stream=None) -> None:
        """The gather and build of one tree now (kernel launches only)."""
        check(self.lib.cx_light_tree_rebuild(self.handle, int(tree), stream_handle(stream)), "cx_light_tree_rebuild")

    def pending(self) -> bool:
        return bool(self.lib.cx_light_tree_pending(self.handle))

    def stats(self) -> dict:
        st = CxLightStats()
        self.lib.cx_light_tree_get_stats(self.handle, ctypes.byref(st))
        return st.as_dict()

    def flush(self, scene: CxScene, materials: np.ndarray, tables: CxLightSceneTables, finalize: bool = True,
              stream=None) -> None:
        lib, h = self.lib, stream_handle(stream)
        self.frame += 1
        requeue = ctypes.c_uint32(0)
        mats = np.ascontiguousarray(materials, np.float32)
        check(lib.cx_light_tree_prepare(self.handle, mats.ctypes.data, ctypes.byref(tables), ctypes.byref(requeue)),
              "cx_light_tree_prepare")
        finalize = finalize or requeue.value != 0
        staged = ctypes.c_uint32(0)
        check(lib.cx_light_tree_stage(self.handle, 1 if finalize else 0, ctypes.byref(staged)), "cx_light_tree_stage")
        check(lib.cx_light_tree_record_transfers(self.handle, h), "cx_light_tree_record_transfers")
        scene.light_codes = self.light_codes
        scene.light_group_rows = self.group_rows
        if finalize:
            check(lib.cx_scene_finalize(ctypes.byref(scene), h), "cx_scene_finalize")
        check(lib.cx_light_tree_record(self.handle, ctypes.byref(scene), self.frame, h), "cx_light_tree_record")
        check(lib.cx_light_tree_end_flush(self.handle, self.frame), "cx_light_tree_end_flush")

    def recycle(self) -> None:
        """Waits for the device, then recycles every flushed frame."""
        _torch().cuda.synchronize()
        check(self.lib.cx_light_tree_recycle(self.handle, self.frame), "cx_light_tree_recycle")

    def settle(self, scene: CxScene, materials: np.ndarray, tables: CxLightSceneTables, finalize: bool = True,
               max_flushes: int = 16, stream=None) -> int:
        """flush + recycle until nothing is pending; returns the flushes made."""
        for i in range(max_flushes):
            self.flush(scene, materials, tables, finalize=finalize and i == 0, stream=stream)
            self.recycle()
            if not self.pending():
                return i + 1
        raise CxError("the light tree still has work after %d flushes" % max_flushes)


def light_sample(view: CxLightTreeView, receivers: np.ndarray, seeds: np.ndarray, stream=None):
    torch = _torch()
    lib = load()
    n = len(receivers)
    r = torch.from_numpy(np.ascontiguousarray(receivers, np.float32)).cuda()
    sd = torch.from_numpy(np.ascontiguousarray(seeds, np.uint32).view(np.int32)).cuda()
    out = torch.empty(max(n, 1) * 12, dtype=torch.float32, device="cuda")
    ok = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    check(lib.cx_light_sample(ctypes.byref(view), r.data_ptr(), sd.data_ptr(), n, out.data_ptr(), ok.data_ptr(),
                              stream_handle(stream)), "cx_light_sample")
    return out.cpu().numpy().reshape(-1, 12)[:n], ok.cpu().numpy()[:n]


def light_point_pdf(view: CxLightTreeView, queries: np.ndarray, stream=None) -> np.ndarray:
    torch = _torch()
    lib = load()
    n = len(queries)
    q = torch.from_numpy(np.ascontiguousarray(queries, np.uint32).view(np.int32)).cuda()
    pdf = torch.empty(max(n, 1), dtype=torch.float32, device="cuda")
    check(lib.cx_light_point_pdf(ctypes.byref(view), q.data_ptr(), n, pdf.data_ptr(), stream_handle(stream)),
          "cx_light_point_pdf")
    return pdf.cpu().numpy()[:n]


def light_sample_brick_box(view: CxLightTreeView, boxes: np.ndarray, seeds: np.ndarray, stream=None) -> np.ndarray:
    """LightTreeSampleBrickBox per box (lo, hi, n): [n, 8] uint32 (ok, group, entry, generation, leaf,
    pdf bits, importance bits, 0)."""
    torch = _torch()
    n = len(boxes)
    b = torch.from_numpy(np.ascontiguousarray(boxes, np.float32)).cuda()
    sd = torch.from_numpy(np.ascontiguousarray(seeds, np.uint32).view(np.int32)).cuda()
    out = torch.empty(max(n, 1) * 8, dtype=torch.int32, device="cuda")
    check(load().cx_light_sample_brick_box(ctypes.byref(view), b.data_ptr(), sd.data_ptr(), n, out.data_ptr(),
                                           stream_handle(stream)), "cx_light_sample_brick_box")
    return out.cpu().numpy().view(np.uint32).reshape(-1, 8)[:n]


def light_brick_pdf(view: CxLightTreeView, queries: np.ndarray, box: bool, stream=None) -> np.ndarray:
    torch = _torch()
    n = len(queries)
    q = torch.from_numpy(np.ascontiguousarray(queries, np.uint32).view(np.int32)).cuda()
    pdf = torch.empty(max(n, 1), dtype=torch.float32, device="cuda")
    check(load().cx_light_brick_pdf(ctypes.byref(view), q.data_ptr(), n, 1 if box else 0, pdf.data_ptr(),
                                    stream_handle(stream)), "cx_light_brick_pdf")
    return pdf.cpu().numpy()[:n]


def light_box_support(view: CxLightTreeView, queries: np.ndarray, stream=None) -> np.ndarray:
    torch = _torch()
    n = len(queries)
    q = torch.from_numpy(np.ascontiguousarray(queries, np.uint32).view(np.int32)).cuda()
    out = torch.empty(max(n, 1) * 4, dtype=torch.float32, device="cuda")
    check(load().cx_light_box_support(ctypes.byref(view), q.data_ptr(), n, out.data_ptr(), stream_handle(stream)),
          "cx_light_box_support")
    return out.cpu().numpy().reshape(-1, 4)[:n]


def sky_cube_sample(texels: np.ndarray, size: int, mip_levels: int, queries: np.ndarray) -> np.ndarray:
    """The cubemap read per query (direction, lod): [n, 3] float32."""
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(texels, np.float32)).cuda()
    q = torch.from_numpy(np.ascontiguousarray(queries, np.float32)).cuda()
    n = len(queries)
    out = torch.empty(max(n, 1) * 3, dtype=torch.float32, device="cuda")
    cube = CxSkyCubemap(t.data_ptr(), size, mip_levels)
    check(load().cx_sky_cube_sample(ctypes.byref(cube), q.data_ptr(), n, out.data_ptr(), stream_handle()),
          "cx_sky_cube_sample")
    return out.cpu().numpy().reshape(-1, 3)[:n]


def sky_sh(texels: np.ndarray, size: int, mip_levels: int) -> np.ndarray:
    """The cubemap's irradiance SH: [9, 3] float32."""
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(texels, np.float32)).cuda()
    out = torch.empty(36, dtype=torch.float32, device="cuda")
    cube = CxSkyCubemap(t.data_ptr(), size, mip_levels)
    check(load().cx_sky_sh(ctypes.byref(cube), out.data_ptr(), stream_handle()), "cx_sky_sh")
    return out.cpu().numpy().reshape(9, 4)[:, :3]


def point_queries(queries: np.ndarray, group_of_brick: np.ndarray, group_coords: np.ndarray, group_shift: int):
    """Face queries {x, n, brick record, world voxel, face, block} (12 u32) as LightTreePointPdf queries
    {x, n, group, group-local voxel, face, block} (identity entity frames)."""
    q = np.ascontiguousarray(queries, np.uint32).copy()
    group = np.asarray(group_of_brick, np.uint32)[q[:, 6]]
    origin = group_coords[group][:, -3:].astype(np.int64) * (8 << group_shift)
    q[:, 6] = group
    q[:, 7:10] = (q[:, 7:10].view(np.int32).astype(np.int64) - origin).astype(np.uint32)
    return q


def light_tables(descriptor: int, groups: int, entities: int) -> dict:
    """The tables a light tree descriptor names, read back as uint8 arrays: descriptor, control, both heaps
    (to the end of the rows' blocks), rows, motion words, light codes, nodes and leaves of both trees (to
    their headers' counts)."""
    d = read_device(descriptor, 128).view(layout.DESCRIPTOR)[0]
    control = read_device(int(d["control"]), 320).view(layout.CONTROL)[0]
    rows = read_device(int(d["group_rows"]), 32 * groups).view(layout.ROW)
    units = int(((rows["block_offset"].astype(np.int64) + rows["block_units"]).max()) if groups else 0)
    out = {
        "descriptor": read_device(descriptor, 128),
        "control": read_device(int(d["control"]), 320),
        "rows": rows.view(np.uint8).copy(),
        "motion": read_device(int(d["entity_motion"]), 4 * entities),
        "codes": read_device(int(d["light_codes"]), 8192),
        "heap0": read_device(int(d["heap"][0]), 16 * units),
        "heap1": read_device(int(d["heap"][1]), 16 * units),
    }
    for t in range(2):
        h = control["headers"][t]
        st = control["states"][t]
        nodes = int(h["node_count"]) if int(h["root"]) != layout.LT_NONE else 0
        out[f"nodes{t}"] = read_device(int(d["nodes"][t]), 224 * max(nodes, 1))
        out[f"leaves{t}"] = read_device(int(d["leaves"][t]), 64 * max(int(h["leaf_count"]), 1))
    return out


class TraceScene:
    """A scene resident on the GPU, ready to trace: bricks, bulk hash table, finalized record words,
    resident bitmaps, brick light records, and the light tree over the light bricks (tree 0; tree 1
    empty). Built with the library's kernels from SceneData (caelix_amd.scene)."""

    def __init__(self, data, load_factor: float = 0.7, mis: bool = True, stream=None):
        torch = _torch()
        lib = self.lib = load()
        dev = "cuda"
        self.data = data
        n = self.bricks = int(len(data.keys))
        self.keys = torch.from_numpy(np.ascontiguousarray(data.keys, np.int32)).to(dev)
        self.payload = torch.from_numpy(np.ascontiguousarray(data.payload).view(np.int16)).to(dev)
        self.materials = torch.from_numpy(np.ascontiguousarray(data.materials, np.float32)).to(dev)
        self.light_codes = torch.from_numpy(np.ascontiguousarray(data.light_codes).view(np.int32)).to(dev)
        ext = getattr(data, "materials_ext", None)
        self.materials_ext = None if ext is None else torch.from_numpy(np.ascontiguousarray(ext, np.float32)).to(dev)
        self.hash = BrickHashTable(self.keys, load_factor=load_factor, stream=stream)
        self.info = torch.zeros(n * 2, dtype=torch.int32, device=dev)
        self.occupancy = torch.zeros(n * 8, dtype=torch.int64, device=dev)
        dim = int(data.grid_dim)
        cells = dim ** 3
        self.brick_bits = torch.zeros((cells + 31) // 32, dtype=torch.int32, device=dev)
        self.super_bits = torch.zeros((cells // 512 + 31) // 32 + 1, dtype=torch.int32, device=dev)
        self.occupied_box = torch.zeros(6, dtype=torch.int32, device=dev)
        sc = self.scene = CxScene()
        sc.keys = self.keys.data_ptr()
        sc.info = self.info.data_ptr()
        sc.occupancy = self.occupancy.data_ptr()
        sc.payload = self.payload.data_ptr()
        sc.slots = self.hash.slots.data_ptr()
        sc.slot_count = self.hash.slot_count
        sc.bricks = n
        sc.hash_info = self.hash.info.data_ptr()
        sc.brick_bits = self.brick_bits.data_ptr()
        sc.super_bits = self.super_bits.data_ptr()
        sc.grid_min[:] = list(data.grid_min)
        sc.grid_dim = dim
        sc.group_shift = int(data.group_shift)
        sc.encoding = int(getattr(data, "encoding", 1))  # 0 packed, 1 block-ID
        sc.materials = self.materials.data_ptr()
        sc.light_codes = self.light_codes.data_ptr()
        sc.emission_scale = 1.0
        sc.occupied_box = self.occupied_box.data_ptr()
        sc.materials_ext = None if self.materials_ext is None else self.materials_ext.data_ptr()
        sss = getattr(data, "subsurface", None)
        self.subsurface = None if sss is None else torch.from_numpy(
            np.ascontiguousarray(sss).view(np.uint8).copy()).to(dev)
        sc.subsurface = None if self.subsurface is None else self.subsurface.data_ptr()
        # the light tree: one static entity at the identity frame, its groups, the tree's flushes
        coords, gob = groups_of(data.keys, int(data.group_shift))
        self.group_coords = coords
        self.group_of_brick_host = gob
        self.group_of_brick = torch.from_numpy(gob.view(np.int32)).to(dev)
        self.group_infos = torch.from_numpy(group_infos(coords).view(np.uint8)).to(dev)
        self.entity_infos = torch.from_numpy(entity_infos(1).view(np.uint8)).to(dev)
        sc.group_of_brick = self.group_of_brick.data_ptr()
        self.lights = LightTree(int(data.group_shift), sc.encoding, min_emission_for(data))
        self.lights.hook("on_entity_added", 0, 0)
        for g in range(len(coords)):
            self.lights.hook("on_group_created", g)
        self.tables = CxLightSceneTables(self.group_infos.data_ptr(), self.entity_infos.data_ptr(), len(coords), 1, n)
        self.flushes = self.lights.settle(sc, data.materials, self.tables, stream=stream)
        st = self.lights.stats()
        self.light_bricks = int(st["light_bricks"])
        self.light_faces = int(st["light_faces"])
        v = self.view = CxLightTreeView()
        v.descriptor = self.lights.descriptor
        v.mis = 1 if mis else 0
        torch.cuda.synchronize()

    def set_brick_scale(self, scale):
        """Per-brick emission scales (a float32 device tensor of `bricks` values) or None (1)."""
        self.brick_scale = scale
        self.scene.brick_emission_scale = scale.data_ptr() if scale is not None else None

    def _output(self, camera, color, radiance, attrib, counters):
        w, h = int(camera["width"]), int(camera["height"])
        key = (w, h)
        if getattr(self, "_workspace_key", None) != key:
            self.workspace = device_bytes(int(self.lib.cx_trace_workspace_bytes(max(w, 1), max(h, 1))), "cuda")
            self._workspace_key = key
        return CxTraceOutput(color.data_ptr() if color is not None else None,
                             radiance.data_ptr() if radiance is not None else None,
                             attrib.data_ptr() if attrib is not None else None, counters.data_ptr(),
                             self.workspace.data_ptr(), self.workspace.numel())

    def trace_frame_adjoint(self, camera: dict, bounces: int, seed: int, spp: int, d_color, grad_albedo, grad_scale,
                            flags: int = 0, counters=None, stream=None, fog=None):
        """Adds the path-replay gradients for d_color = dL/dI (rgba per pixel, float32 device tensor)
        to grad_albedo ([256, 4] float32) and grad_scale ([bricks] float32)."""
        torch = _torch()
        counters = counters if counters is not None else torch.zeros(8, dtype=torch.int64, device="cuda")
        cam = camera_struct(camera)
        params = CxTraceParams(bounces, seed, spp, flags)
        apply_sky(params, None, fog)
        grads = CxTraceGradients(d_color.data_ptr(), grad_albedo.data_ptr(), grad_scale.data_ptr())
        out = self._output(camera, None, None, None, counters)
        check(self.lib.cx_trace_frame_adjoint(ctypes.byref(self.scene), ctypes.byref(self.view), ctypes.byref(cam),
                                              ctypes.byref(params), ctypes.byref(grads), ctypes.byref(out),
                                              stream_handle(stream)), "cx_trace_frame_adjoint")
        return counters

    def trace_frame(self, camera: dict, bounces: int, seed: int, spp: int = 1, color=None, radiance=None,
                    attrib=None, counters=None, stream=None, flags: int = 0, sky: dict | None = None, fog=None,
                    bounces_transparent: int = 0, bounce_nee_probability: float = 1.0, hint: LightHint | None = None,
                    sky_cube: tuple | None = None):
        """sky_cube = (texels [n, 4] float32 numpy or a device tensor, size, mip_levels): the host cubemap
        in place of the procedural sky (with `sky`)."""
        torch = _torch()
        w, h = int(camera["width"]), int(camera["height"])
        color = color if color is not None else torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
        radiance = radiance if radiance is not None else torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
        attrib = attrib if attrib is not None else torch.empty(w * h, dtype=torch.int32, device="cuda")
        counters = counters if counters is not None else torch.zeros(8, dtype=torch.int64, device="cuda")
        cam = camera_struct(camera)
        params = CxTraceParams(bounces, seed, spp, flags)
        apply_sky(params, sky, fog)
        params.bounces_transparent = int(bounces_transparent)
        params.bounce_nee_probability = float(bounce_nee_probability)
        cube_keep = None
        if sky_cube is not None:
            texels, size, mips = sky_cube
            if isinstance(texels, np.ndarray):
                texels = torch.from_numpy(np.ascontiguousarray(texels, np.float32)).cuda()
            cube = CxSkyCubemap(texels.data_ptr(), int(size), int(mips))
            cube_keep = (texels, cube)
            params.sky_cubemap = ctypes.cast(ctypes.pointer(cube), ctypes.c_void_p)
        hint_struct = None
        if hint is not None:
            hint_struct = hint.frame_struct(camera)
            params.hint = ctypes.cast(ctypes.pointer(hint_struct), ctypes.c_void_p)
        out = self._output(camera, color, radiance, attrib, counters)
        check(self.lib.cx_trace_frame(ctypes.byref(self.scene), ctypes.byref(self.view), ctypes.byref(cam),
                                      ctypes.byref(params), ctypes.byref(out), stream_handle(stream)), "cx_trace_frame")
        if hint is not None:
            hint.advance(camera)
        if cube_keep is not None:
            torch.cuda.current_stream().synchronize()  # the texels may be a temporary upload
        return color, radiance, attrib, counters

    def trace_rays(self, rays: np.ndarray) -> np.ndarray:
        torch = _torch()
        r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
        n = len(rays)
        hits = torch.empty(max(n, 1) * 9, dtype=torch.int32, device="cuda")
        check(self.lib.cx_trace_rays(ctypes.byref(self.scene), r.data_ptr(), n, hits.data_ptr(), stream_handle()),
              "cx_trace_rays")
        return hits.cpu().numpy().view(np.uint32).reshape(-1, 9)[:n]

    def light_sample(self, receivers: np.ndarray, seeds: np.ndarray):
        """LightTreeSample per receiver (x, n): ([n, 12] float32 = position, normal, pdf, block, face, group,
        entry, voxel in group (the last five as u32 bits), ok [n])."""
        return light_sample(self.view, receivers, seeds)

    def light_point_pdf(self, queries: np.ndarray) -> np.ndarray:
        """LightTreePointPdf per query (x, n, group, voxel in group, face, block)."""
        return light_point_pdf(self.view, queries)

    def light_face_pdf(self, queries: np.ndarray) -> np.ndarray:
        """The face pdf of queries that name a brick record and a world voxel (words 6..9): converted to
        the group and group-local voxel of LightTreePointPdf."""
        return self.light_point_pdf(point_queries(queries, self.group_of_brick_host, self.group_coords,
                                                  int(self.data.group_shift)))

    def light_tables(self) -> dict:
        """The light tree's tables read back through its descriptor (numpy), for the oracle."""
        return light_tables(self.lights.descriptor, len(self.group_coords), 1)

    def tree_read(self, tree: int = 0) -> dict:
        """Tree `tree` as read back: header (HEADER), nodes (NODE) and leaves (LEAF) to their counts."""
        _torch().cuda.synchronize()
        t = self.light_tables()
        control = t["control"].view(layout.CONTROL)[0]
        return {"header": control["headers"][tree].copy(), "nodes": t["nodes%d" % tree].view(layout.NODE),
                "leaves": t["leaves%d" % tree].view(layout.LEAF)}

    def host_arrays(self) -> dict:
        """Every device array of the scene and its light tree, copied to numpy (for the oracle)."""
        _torch().cuda.synchronize()
        return {
            "keys": self.keys.cpu().numpy(),
            "info": self.info.cpu().numpy().view(np.uint32),
            "occupancy": self.occupancy.cpu().numpy().view(np.uint64),
            "payload": self.payload.cpu().numpy().view(np.uint16),
            "slots": self.hash.slots.cpu().numpy().view(np.uint32),
            "hash_info": self.hash.info.cpu().numpy().view(np.uint32),
            "brick_bits": self.brick_bits.cpu().numpy().view(np.uint32),
            "super_bits": self.super_bits.cpu().numpy().view(np.uint32),
            "materials": self.materials.cpu().numpy(),
            "light_codes": self.light_codes.cpu().numpy().view(np.uint32),
            "group_of_brick": self.group_of_brick_host.copy(),
            "group_infos": self.group_infos.cpu().numpy(),
            "entity_infos": self.entity_infos.cpu().numpy(),
            **({} if self.subsurface is None else {"subsurface": self.subsurface.cpu().numpy()}),
            **{"lt_" + k: v for k, v in self.light_tables().items()},
        }
