"""The reference tests' own scenes, rebuilt from their seeds and literals (no reference file is read):

Mt19937Stream        raw std::mt19937 outputs (numpy's legacy RandomState draws full-range uint32
                     values straight from the same generator: init_genrand seeding, tempering)
random_blobs         RandomBlobs (tests/test_scenes.cpp:27-66): blobs of up to 30 face-adjacent keys
random_brick         RandomBrick (test_scenes.cpp:27-37): 30 % air, else 0x8000 | 15 random bits
random_opaque_world  RandomOpaqueWorld (test_scenes.cpp:68-76): keys first, then a brick per key
solid_cube_world     SolidCubeWorld (test_scenes.cpp:78-98): 3^3 bricks around (16, 0, 32) of 0xC000 |
                     14 random bits
packed_materials     the built-in table of the packed encoding (lib/src/scene/material_table.cpp:90-110:
                     0x8000 | rgb5 opaque, 0x4000 | rgb4 | strength emissive with emission rgb4 *
                     2^(3 strength), transparent ids from the glass palette: 0 vacuum, 1 clear glass)
                     and the light codes of a threshold (light_tree.cpp:186-199: luminance >=
                     min_emission, 20 by default)
red_brick            tests/trace_test.cpp:436-445: one brick of packed red 0x8000 | 31 << 10 at (0, 0, 0)
ceiling_room         trace_test.cpp:1118-1143: a grey floor under an emissive ceiling
lamp_box             trace_test.cpp:2929-2961: a closed grey box with a ceiling opening, half glazed,
                     a pillar and five single-voxel lamps of strength 3
look_plus_z, look_down, look_along   the tests' cameras (trace_test.cpp:257-272, 1146-1160, 1794-1820)
"""
from __future__ import annotations

import numpy as np

from caelix_amd.scene import SceneData

PACKED, BLOCK_ID = 0, 1
DIRS = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]


class Mt19937Stream:
    """Successive raw outputs of std::mt19937(seed)."""

    def __init__(self, seed: int):
        self._rs = np.random.RandomState(seed)
        self._buf = np.zeros(0, np.uint32)
        self._pos = 0

    def draw(self, n: int) -> np.ndarray:
        if self._pos + n > len(self._buf):
            rest = self._buf[self._pos:]
            fresh = self._rs.randint(0, 2**32, size=max(n, 1 << 16), dtype=np.uint32)
            self._buf = np.concatenate([rest, fresh])
            self._pos = 0
        out = self._buf[self._pos:self._pos + n]
        self._pos += n
        return out

    def __call__(self) -> int:
        return int(self.draw(1)[0])


def random_blobs(rng: Mt19937Stream, count: int) -> list[tuple[int, int, int]]:
    seen: set = set()
    keys: list = []
    while len(keys) < count:
        first = (rng() % 80 - 40, rng() % 40 - 20, rng() % 80 - 40)
        if first in seen:
            continue
        seen.add(first)
        blob = [first]
        keys.append(first)
        k = 0
        while k < 29 and len(keys) < count:
            base = blob[rng() % len(blob)]
            d = DIRS[rng() % 6]
            nxt = (base[0] + d[0], base[1] + d[1], base[2] + d[2])
            if nxt not in seen:
                seen.add(nxt)
                blob.append(nxt)
                keys.append(nxt)
            k += 1
    return keys


def random_brick(rng: Mt19937Stream) -> np.ndarray:
    r = rng.draw(512).astype(np.uint32)
    return np.where(r % 100 < 30, 0, 0x8000 | ((r >> 8) & 0x7FFF)).astype(np.uint16)


def random_opaque_world(seed: int, count: int) -> dict:
    rng = Mt19937Stream(seed)
    world = {}
    for key in random_blobs(rng, count):
        world[key] = random_brick(rng)
    return world


def solid_cube_world(seed: int) -> dict:
    rng = Mt19937Stream(seed)
    world = {}
    for z in range(31, 34):
        for y in range(-1, 2):
            for x in range(15, 18):
                world[(x, y, z)] = (0xC000 | (rng.draw(512) & 0x3FFF)).astype(np.uint16)
    return world


def fill_box(world: dict, lo, hi, value: int) -> None:
    """FillBox (trace_test.cpp:1768-1791): every voxel of the inclusive box, creating air bricks."""
    for z in range(lo[2], hi[2] + 1):
        for y in range(lo[1], hi[1] + 1):
            for x in range(lo[0], hi[0] + 1):
                key = (x >> 3, y >> 3, z >> 3)
                brick = world.setdefault(key, np.zeros(512, np.uint16))
                brick[(x & 7) + 8 * (y & 7) + 64 * (z & 7)] = value


def half(x) -> np.ndarray:
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def packed_materials(min_emission: float = 20.0):
    """(table [65536, 8], ext [65536, 4], light-code words [2048]) of the packed encoding's built-in
    materials, values rounded to halves as the hot table stores them."""
    codes = np.arange(65536, dtype=np.uint32)
    table = np.zeros((65536, 8), np.float32)
    ext = np.zeros((65536, 4), np.float32)
    opaque = (codes & 0x8000) != 0
    emissive = ((codes & 0x8000) == 0) & ((codes & 0x4000) != 0)
    rgb5 = np.stack([(codes >> 10) & 31, (codes >> 5) & 31, codes & 31], axis=1).astype(np.float32) / np.float32(31)
    rgb4 = np.stack([(codes >> 10) & 15, (codes >> 6) & 15, (codes >> 2) & 15], axis=1).astype(np.float32) / np.float32(15)
    strength = np.exp2(np.float32(3) * (codes & 3).astype(np.float32)).astype(np.float32)
    table[opaque, 0:3] = rgb5[opaque]
    table[emissive, 0:3] = rgb4[emissive]
    table[emissive, 4:7] = rgb4[emissive] * strength[emissive, None]
    ext[opaque | emissive, 0] = 1.5
    # the glass palette's first entries (vacuum, clear glass): albedo 1, smoothness 1
    for code, ior in ((0, 1.0), (1, 1.55)):
        table[code, 0:3] = 1.0
        table[code, 3] = 1.0
        ext[code, 0] = ior
    lum = (np.float32(0.2126) * table[:, 4] + np.float32(0.7152) * table[:, 5]) + np.float32(0.0722) * table[:, 6]
    light = lum >= np.float32(min_emission)
    words = np.zeros(2048, np.uint32)
    for b in range(32):
        words |= light.reshape(2048, 32)[:, b].astype(np.uint32) << np.uint32(b)
    return half(table), half(ext), words


def scene_from_world(world: dict, camera: dict, group_shift: int = 4, encoding: int = PACKED,
                     min_emission: float = 20.0, lights: bool = True, glass: bool = False) -> SceneData:
    """A SceneData of the world's bricks (keys in sorted order, as KeysOf walks the reference's map), the
    packed built-in materials and a grid that covers every key. lights False: no light codes (the plain
    path tracer, as the reference renders with the light NEE off)."""
    keys = sorted(world)
    table, ext, words = packed_materials(min_emission)
    arr = np.array(keys, np.int32)
    lo = arr.min(axis=0)
    dim = int(np.ceil((arr.max(axis=0) - lo + 1).max() / 8.0) * 8)
    payload = np.stack([world[k] for k in keys]).astype(np.uint16)
    light_words = words if lights else np.zeros(2048, np.uint32)
    codes = payload.astype(np.int64)
    lit = ((light_words[codes >> 5] >> (codes & 31).astype(np.uint32)) & 1).astype(bool)
    return SceneData(keys=arr, payload=payload, materials=table, light_codes=light_words,
                     grid_min=tuple(int(v) for v in lo), grid_dim=dim, group_shift=group_shift, camera=camera,
                     emissive_bricks=np.nonzero(lit.any(axis=1))[0], encoding=encoding,
                     materials_ext=ext if glass else None)


def _camera(position, right, up, forward, tan_half_fov_y, size):
    f32 = np.float32
    return {"position": np.array(position, f32), "forward": np.array(forward, f32), "right": np.array(right, f32),
            "up": np.array(up, f32), "tan_half_fov_y": f32(tan_half_fov_y), "aspect": f32(1.0),
            "width": int(size), "height": int(size)}


def look_plus_z(x: float, y: float, z: float, size: int = 256) -> dict:
    return _camera((x, y, z), (1, 0, 0), (0, 1, 0), (0, 0, 1), np.tan(np.float32(30.0 * 3.14159265 / 180.0)), size)


def look_down(size: int = 256) -> dict:
    """At (32, 12, 32), looking down at the floor, vertical field of view 30 degrees."""
    return _camera((32, 12, 32), (1, 0, 0), (0, 0, 1), (0, -1, 0),
                   np.tan(np.float32(15.0) * np.float32(3.14159265358979) / np.float32(180.0)), size)


def look_along(position, forward, tan_half_fov_y: float, size: int) -> dict:
    f = np.asarray(forward, np.float64)
    f = f / np.linalg.norm(f)
    r = np.array([f[2], 0.0, -f[0]])
    r /= np.sqrt(r[0] ** 2 + r[2] ** 2)
    u = np.array([f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2], f[0] * r[1] - f[1] * r[0]])
    return _camera(position, r, u, f, tan_half_fov_y, size)


RED = 0x8000 | (31 << 10)
GREY_FLOOR = 0x8000 | (15 << 10) | (15 << 5) | 15
WHITE_EMITTER = 0x4000 | (15 << 10) | (15 << 6) | (15 << 2)
WHITE_LAMP = WHITE_EMITTER | 3
CLEAR_GLASS = 1
CEILING_Y, ROOM_SIZE, CAMERA_HEIGHT = 24.0, 64.0, 12.0
LAMP_BOX_INTERIOR, LAMP_BOX_IMAGE = 24, 128


def red_brick() -> dict:
    return {(0, 0, 0): np.full(512, RED, np.uint16)}


def ceiling_room() -> dict:
    world = {}
    for z in range(8):
        for x in range(8):
            world[(x, -1, z)] = np.full(512, GREY_FLOOR, np.uint16)
            world[(x, 3, z)] = np.full(512, WHITE_EMITTER, np.uint16)
    return world


def lamp_box() -> dict:
    world: dict = {}
    lo, hi = -1, LAMP_BOX_INTERIOR
    fill_box(world, (lo, lo, lo), (hi, lo, hi), GREY_FLOOR)
    fill_box(world, (lo, hi, lo), (hi, hi, hi), GREY_FLOOR)
    fill_box(world, (lo, lo, lo), (lo, hi, hi), GREY_FLOOR)
    fill_box(world, (hi, lo, lo), (hi, hi, hi), GREY_FLOOR)
    fill_box(world, (lo, lo, lo), (hi, hi, lo), GREY_FLOOR)
    fill_box(world, (lo, lo, hi), (hi, hi, hi), GREY_FLOOR)
    fill_box(world, (3, hi, 3), (8, hi, 8), 0)
    fill_box(world, (3, hi, 3), (5, hi, 8), CLEAR_GLASS)
    fill_box(world, (15, 0, 9), (17, 13, 11), GREY_FLOOR)
    fill_box(world, (16, 14, 10), (16, 14, 10), WHITE_LAMP)
    fill_box(world, (18, hi, 18), (18, hi, 18), WHITE_LAMP)
    fill_box(world, (10, hi, 20), (10, hi, 20), WHITE_LAMP)
    fill_box(world, (hi, 6, 12), (hi, 6, 12), WHITE_LAMP)
    fill_box(world, (6, lo, 18), (6, lo, 18), WHITE_LAMP)
    return world


def lamp_box_camera() -> dict:
    return look_along((12.0, 14.0, 1.5), (0.0, -0.35, 1.0), np.tan(40.0 * np.pi / 180.0), LAMP_BOX_IMAGE)
