""".cxw world files: the reader (the library's cx_cxw_* entry points) and a writer for synthetic worlds.

CxwFile mirrors the reference's Rust reader (host-rs/caelix-cxw/src/lib.rs): open() parses the header,
entity table and sector indices; world / sectors; decode(sector) returns one sector's non-air bricks;
decode_all(threads) decodes every sector on host threads and returns all bricks in sector order. A
format error raises CxwError with the kind and fields of the reference's Error variant
(lib.rs:69-109). scene_bricks() turns the decoded bricks into the keys / payload arrays of SceneData,
so a world read from disk goes through the same GPU scene build as a synthetic one.

write() writes a .cxw file (format version 2..8) from Python sectors, for tests and benchmarks: raw
deflate payloads (zlib, no header) with the reference reader's layout (lib.rs:10-64, 314-656).
"""
from __future__ import annotations

import ctypes
import struct
import zlib
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import CX_OK, CxError, load

FILE_MAGIC = 0x534C5843        # "CXLS"
LEGACY_FILE_MAGIC = 0x534C5856  # "VXLS"
PAYLOAD_MAGIC = 0x32535856      # "VXS2"
CURRENT_VERSION = 8
BRICKS_IN_SECTOR = 4096
BLOCKS_IN_BRICK = 512

ERROR_KINDS = {0: "None", 1: "Io", 2: "BadMagic", 3: "UnsupportedVersion", 4: "BadFlags", 5: "Truncated",
               6: "BadRegionSize", 7: "BadSectorMagic", 8: "BadCounts", 9: "Inflate", 10: "DuplicateSector",
               11: "BadPayload"}


class CxCxwError(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("coord", ctypes.c_int32 * 3), ("offset", ctypes.c_uint64),
                ("expected", ctypes.c_uint64), ("actual", ctypes.c_uint64), ("what", ctypes.c_char * 64),
                ("message", ctypes.c_char * 192)]


class CxCxwWorld(ctypes.Structure):
    _fields_ = [("version", ctypes.c_uint16), ("has_player_pose", ctypes.c_uint8), ("pad", ctypes.c_uint8),
                ("pose_position", ctypes.c_float * 3), ("pose_rotation", ctypes.c_float * 4),
                ("entity_count", ctypes.c_uint32), ("sector_count", ctypes.c_uint64)]


class CxCxwEntity(ctypes.Structure):
    _fields_ = [("guid", ctypes.c_uint8 * 16), ("position", ctypes.c_float * 3), ("rotation", ctypes.c_float * 4),
                ("velocity", ctypes.c_float * 6), ("sector_count", ctypes.c_uint32),
                ("require_update_flags", ctypes.c_uint16), ("flags", ctypes.c_uint8), ("is_protected", ctypes.c_uint8)]


class CxCxwSector(ctypes.Structure):
    _fields_ = [("entity", ctypes.c_uint32), ("coord", ctypes.c_int32 * 3), ("region_offset", ctypes.c_uint64),
                ("region_size", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class CxwError(CxError):
    """A .cxw format error: kind is the reference's Error variant name; the fields are its fields."""

    def __init__(self, e: CxCxwError):
        self.kind = ERROR_KINDS.get(int(e.kind), str(e.kind))
        self.coord = [int(v) for v in e.coord]
        self.offset = int(e.offset)
        self.expected = int(e.expected)
        self.actual = int(e.actual)
        self.what = e.what.decode()
        super().__init__(f"{self.kind}: {e.message.decode()}")


@dataclass
class Entity:
    guid: bytes
    position: list
    rotation: list
    flags: int
    protected: bool
    velocity: list
    require_update_flags: int
    sector_count: int

    @property
    def has_body(self) -> bool:
        return self.flags & 1 != 0

    @property
    def is_static(self) -> bool:
        return self.flags & 2 != 0

    def guid_string(self) -> str:
        """The dashed UUID string of the 16 bytes read like .NET Guid(byte[]) (lib.rs:138-150)."""
        g = self.guid
        a, b, c = struct.unpack_from("<IHH", g)
        return f"{a:08x}-{b:04x}-{c:04x}-{g[8]:02x}{g[9]:02x}-" + "".join(f"{x:02x}" for x in g[10:16])


@dataclass
class World:
    version: int
    player_pose: tuple | None
    entities: list


@dataclass
class Sector:
    entity: int
    coord: tuple
    region_offset: int
    region_size: int


@dataclass
class DecodedSector:
    entity: int
    coord: tuple
    keys: np.ndarray       # int32 [n, 3] entity-local brick keys, ascending brick index
    blocks: np.ndarray     # uint16 [n, 512] raw values, voxel x + 8y + 64z
    summaries: np.ndarray  # uint32 [n] preview words
    voxels: int


class CxwFile:
    """An open .cxw file (memory-mapped by the library)."""

    def __init__(self, path):
        lib = self.lib = load()
        self._handle = ctypes.c_void_p()
        err = CxCxwError()
        code = lib.cx_cxw_open(str(path).encode(), ctypes.byref(self._handle), ctypes.byref(err))
        if code != CX_OK:
            self._handle = None
            raise CxwError(err) if err.kind != 0 else CxError(f"cx_cxw_open failed with code {code}")
        w = CxCxwWorld()
        lib.cx_cxw_world(self._handle, ctypes.byref(w))
        entities = []
        for i in range(w.entity_count):
            e = CxCxwEntity()
            lib.cx_cxw_entity(self._handle, i, ctypes.byref(e))
            entities.append(Entity(bytes(e.guid), list(e.position), list(e.rotation), int(e.flags),
                                   bool(e.is_protected), list(e.velocity), int(e.require_update_flags),
                                   int(e.sector_count)))
        pose = (list(w.pose_position), list(w.pose_rotation)) if w.has_player_pose else None
        self.world = World(int(w.version), pose, entities)
        self.sectors = []
        for i in range(w.sector_count):
            s = CxCxwSector()
            lib.cx_cxw_sector(self._handle, i, ctypes.byref(s))
            self.sectors.append(Sector(int(s.entity), tuple(s.coord), int(s.region_offset), int(s.region_size)))

    def close(self):
        if self._handle:
            self.lib.cx_cxw_close(self._handle)
            self._handle = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def decode(self, index: int) -> DecodedSector:
        keys = np.empty((BRICKS_IN_SECTOR, 3), np.int32)
        blocks = np.empty((BRICKS_IN_SECTOR, BLOCKS_IN_BRICK), np.uint16)
        summaries = np.empty(BRICKS_IN_SECTOR, np.uint32)
        count = ctypes.c_uint32()
        voxels = ctypes.c_uint64()
        err = CxCxwError()
        code = self.lib.cx_cxw_decode(self._handle, index, keys.ctypes.data, blocks.ctypes.data, summaries.ctypes.data,
                                      ctypes.byref(count), ctypes.byref(voxels), ctypes.byref(err))
        if code != CX_OK:
            raise CxwError(err) if err.kind != 0 else CxError(f"cx_cxw_decode failed with code {code}")
        n = int(count.value)
        s = self.sectors[index]
        return DecodedSector(s.entity, s.coord, keys[:n].copy(), blocks[:n].copy(), summaries[:n].copy(),
                             int(voxels.value))

    def decode_all(self, threads: int = 0) -> dict:
        """Every sector's bricks, concatenated in sector order: keys, blocks, summaries, entity."""
        bricks = ctypes.c_uint64()
        voxels = ctypes.c_uint64()
        err = CxCxwError()
        code = self.lib.cx_cxw_decode_all(self._handle, threads, ctypes.byref(bricks), ctypes.byref(voxels),
                                          ctypes.byref(err))
        if code != CX_OK:
            raise CxwError(err) if err.kind != 0 else CxError(f"cx_cxw_decode_all failed with code {code}")
        n = int(bricks.value)
        keys = np.empty((n, 3), np.int32)
        blocks = np.empty((n, BLOCKS_IN_BRICK), np.uint16)
        summaries = np.empty(n, np.uint32)
        entity = np.empty(n, np.uint32)
        code = self.lib.cx_cxw_copy_decoded(self._handle, keys.ctypes.data, blocks.ctypes.data, summaries.ctypes.data,
                                            entity.ctypes.data, n)
        if code != CX_OK:
            raise CxError(f"cx_cxw_copy_decoded failed with code {code}")
        return {"keys": keys, "blocks": blocks, "summaries": summaries, "entity": entity, "voxels": int(voxels.value)}


def scene_bricks(decoded: dict, entity: int = 0) -> tuple[np.ndarray, np.ndarray, tuple, int]:
    """keys, payload, grid_min, grid_dim for the scene build (SceneData fields) from one entity's
    bricks: the grid is the bounding box of the keys rounded up to whole 8^3-brick super cells."""
    sel = decoded["entity"] == entity
    keys = decoded["keys"][sel]
    payload = decoded["blocks"][sel]
    if len(keys) == 0:
        return keys, payload, (0, 0, 0), 8
    lo = keys.min(axis=0)
    extent = int((keys.max(axis=0) - lo).max()) + 1
    dim = max(8, (extent + 7) // 8 * 8)
    return keys, payload, tuple(int(v) for v in lo), dim


# ------------------------------------------------------------------------------------------ writer

@dataclass
class WSector:
    coord: tuple
    bricks: list                   # [(absolute brick index 0..4095, uint16[512])]
    previews: dict = field(default_factory=dict)  # brick index -> preview word (0 otherwise)
    free: int = 0                  # free brick ids, placed below the used ones
    physics_slot_first: bool = False
    payload_magic: int = PAYLOAD_MAGIC
    compressed_override: bytes | None = None


@dataclass
class WEntity:
    sectors: list
    guid: bytes = bytes(range(1, 17))
    position: tuple = (1.0, 2.0, 3.0)
    rotation: tuple = (0.0, 0.0, 0.0, 1.0)
    require: int = 0
    flags_byte: int = 2
    velocity: tuple = (0.0,) * 6
    protected: int = 0


def payload_bytes(s: WSector) -> bytes:
    """The inflated sector payload: used brick i gets id capacity - 1 - i; free ids hold garbage."""
    n = len(s.bricks)
    capacity = n + s.free
    indices = np.full(BRICKS_IN_SECTOR, -1, np.int16)
    for i, (bi, _) in enumerate(s.bricks):
        indices[bi] = capacity - 1 - i
    out = bytearray(struct.pack("<4I", s.payload_magic, capacity, n, s.free))
    out += indices.tobytes()
    out += np.arange(s.free, dtype=np.int16).tobytes()
    out += struct.pack("<H", 1) + bytes(BRICKS_IN_SECTOR * 2)
    out += struct.pack("<I", 2 if s.physics_slot_first else 1)
    if s.physics_slot_first:
        out += struct.pack("<BH", 1, 1) + bytes([0x7E]) * (capacity * BLOCKS_IN_BRICK)
    out += struct.pack("<BH", 0, 2)
    block = np.full((capacity, BLOCKS_IN_BRICK), 0xFFFF, np.uint16)
    for i, (_, values) in enumerate(s.bricks):
        block[capacity - 1 - i] = values
    out += block.tobytes()
    return bytes(out)


def deflate_raw(data: bytes, level: int = 6) -> bytes:
    c = zlib.compressobj(level, zlib.DEFLATED, -15)
    return c.compress(data) + c.flush()


def _record_bytes(version: int) -> int:
    n = 16 + 12 + 16 + 2
    n += 1 if version >= 3 else 0
    n += 24 if version >= 4 else 0
    n += 1 if version >= 5 else 0
    return n + 8 + 4


def _region(s: WSector, level: int) -> bytes:
    compressed = s.compressed_override if s.compressed_override is not None else deflate_raw(payload_bytes(s), level)
    preview = np.zeros(BRICKS_IN_SECTOR, np.uint32)
    for bi, word in s.previews.items():
        preview[bi] = word
    return preview.tobytes() + struct.pack("<I", len(compressed)) + compressed


def _table(entities: list, index_at: list, version: int, world: tuple) -> bytes:
    table = bytearray()
    if version >= 7:
        table += struct.pack("<B3f4f", world[0], *world[1], *world[2])
    table += struct.pack("<I", len(entities))
    for e, at in zip(entities, index_at):
        table += bytes(e.guid) + struct.pack("<3f4fH", *e.position, *e.rotation, e.require)
        if version >= 3:
            table += struct.pack("<B", e.flags_byte)
        if version >= 4:
            table += struct.pack("<6f", *e.velocity)
        if version >= 5:
            table += struct.pack("<B", e.protected)
        table += struct.pack("<QI", at, len(e.sectors))
    return bytes(table)


def write(path, entities: list, version: int = CURRENT_VERSION, magic: int = FILE_MAGIC, flags: int = 1,
          world: tuple = (0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0)), level: int = 6,
          table_first: bool = False) -> bytes:
    """Writes a .cxw file and returns its bytes. Default order: header, regions, sector indices, entity
    table. table_first: header, entity table, then per entity its sector index and its regions (the
    reader only follows offsets, so both orders read the same)."""
    regions = [[_region(s, level) for s in e.sectors] for e in entities]
    table_len = len(_table(entities, [0] * len(entities), version, world))
    out = bytearray()
    if table_first:
        body = bytearray()
        index_at = []
        for e, rs in zip(entities, regions):
            at = 64 + table_len + len(body)
            index_at.append(at)
            offset = at + 24 * len(rs)
            index = bytearray()
            for s, r in zip(e.sectors, rs):
                index += struct.pack("<3iQI", *s.coord, offset, len(r))
                offset += len(r)
            body += index + b"".join(rs)
        out += struct.pack("<IHHQ", magic, version, flags, 64).ljust(64, b"\0")
        out += _table(entities, index_at, version, world) + body
    else:
        blob = bytearray()
        entries = []
        for e, rs in zip(entities, regions):
            entries.append([])
            for s, r in zip(e.sectors, rs):
                entries[-1].append((s.coord, 64 + len(blob), len(r)))
                blob += r
        index = bytearray()
        index_at = []
        for es in entries:
            index_at.append(64 + len(blob) + len(index))
            for coord, offset, size in es:
                index += struct.pack("<3iQI", *coord, offset, size)
        table_offset = 64 + len(blob) + len(index)
        out += struct.pack("<IHHQ", magic, version, flags, table_offset).ljust(64, b"\0")
        out += blob + index + _table(entities, index_at, version, world)
    Path(path).write_bytes(bytes(out))
    return bytes(out)


def sectors_from_bricks(keys: np.ndarray, payload: np.ndarray) -> list:
    """Groups bricks (int32 [n, 3] keys, uint16 [n, 512]) into 16^3-brick sectors of one entity."""
    coords = np.floor_divide(keys, 16)
    local = keys - coords * 16
    bi = local[:, 0] + 16 * local[:, 1] + 256 * local[:, 2]
    order = np.lexsort((bi, coords[:, 2], coords[:, 1], coords[:, 0]))
    out = []
    current = None
    for i in order:
        c = tuple(int(v) for v in coords[i])
        if current is None or current.coord != c:
            current = WSector(c, [])
            out.append(current)
        current.bricks.append((int(bi[i]), payload[i]))
    return out


def load_scene(path, materials: np.ndarray, light_codes: np.ndarray, camera: dict, threads: int = 0,
               entity: int = 0, group_shift: int = 5):
    """SceneData (caelix_amd.scene) of one entity of a .cxw world: its non-air bricks on the bounding
    grid of its keys, with the material table, light-code bitset and camera given (a .cxw file
    holds neither). emissive_bricks = the bricks holding a block whose light-code bit is set."""
    from .scene import SceneData

    with CxwFile(path) as f:
        decoded = f.decode_all(threads)
    keys, payload, grid_min, dim = scene_bricks(decoded, entity)
    codes = np.asarray(light_codes, np.uint32)
    lit = ((codes[payload >> 5] >> (payload & 31).astype(np.uint32)) & 1).astype(bool) & (payload != 0)
    emissive = np.nonzero(lit.any(axis=1))[0]
    return SceneData(keys=np.ascontiguousarray(keys), payload=np.ascontiguousarray(payload), materials=materials,
                     light_codes=codes, grid_min=grid_min, grid_dim=dim, group_shift=group_shift, camera=camera,
                     emissive_bricks=emissive)
