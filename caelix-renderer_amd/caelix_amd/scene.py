"""Seeded synthetic voxel scenes for the trace workloads (host side, numpy).

terrain(world, seed, ...) builds a value-noise heightfield world of 8^3 bricks with uint8 material ids
stored as block-ID-encoded opaque blocks (0x8000 | id, record_layout.h:27-34):

  small   world 128^3: all 16^3 = 4,096 brick cells resident, about 30 % solid voxels, 512 emissive
          bricks with emission uniform in [0.1, 10], 256 x 256 pixels
  1080p   world 4096^3: 512^2 brick columns x 4 surface bricks = 1,048,576 resident bricks, 8-octave
          terrain, 65,536 emissive bricks with emission log-uniform in [0.01, 100], 1920 x 1080

Materials: ids 1..127 diffuse (no emission), ids 128..255 emissive (the emissive bricks' solid voxels
use one of them each); albedo and emission are float16 values (the reference's material table holds
halves, caelix_material.hlsl:90-113). The light threshold is the smallest emission luminance, so every
emissive material is a light code.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

OPAQUE = 0x8000


@dataclass
class SceneData:
    keys: np.ndarray         # int32 [n, 3] brick coordinates
    payload: np.ndarray      # uint16 [n, 512] block values, voxel x + 8y + 64z
    materials: np.ndarray    # float32 [65536, 8]
    light_codes: np.ndarray  # uint32 [2048]
    grid_min: tuple
    grid_dim: int
    group_shift: int
    camera: dict
    emissive_bricks: np.ndarray
    encoding: int = 1        # block values: 1 block-ID (opaque 0x8000, id 0x1FF, faces << 9), 0 packed
    materials_ext: np.ndarray | None = None  # float32 [65536, 4] (ior, extinction, 0, 0): the full material model
    subsurface: np.ndarray | None = None     # SUBSURFACE [65536]: the subsurface entries (CxSubsurface)


# CxSubsurface (include/caelix_amd/caelix_amd.h): one per material code, 32 bytes.
SUBSURFACE = np.dtype([("one_minus_albedo", "<f4", 3), ("anisotropy", "<f4"), ("probability", "<f4"),
                       ("model", "<u4"), ("steps", "<u4"), ("pad", "<u4")])
assert SUBSURFACE.itemsize == 32
SSS_MODEL_HG = 1


def subsurface_table(entries: dict) -> np.ndarray:
    """The subsurface table of {material code: (sss_albedo rgb, anisotropy, probability, steps[, model])},
    packed as the reference's host packs its GPU subsurface entry (docs/MATERIALS.md: halves of
    1 - sss_albedo, of g and of p, rounded to nearest even) and widened back to float."""
    t = np.zeros(65536, SUBSURFACE)
    half = lambda x: np.float32(np.float16(np.float32(x)))  # noqa: E731
    for code, e in entries.items():
        albedo, g, p, steps = e[:4]
        model = e[4] if len(e) > 4 else SSS_MODEL_HG
        t[code]["one_minus_albedo"] = [half(np.float32(1.0) - np.float32(a)) for a in albedo]
        t[code]["anisotropy"] = half(g)
        t[code]["probability"] = half(p)
        t[code]["model"] = model
        t[code]["steps"] = int(steps) & 0xFF
    return t


def value_noise(size: int, seed: int, octaves: int, base_period: int) -> np.ndarray:
    """Sum of `octaves` value-noise layers on a size x size grid, in [0, 1): lattice values at every
    `period` cells (period halves per octave), smoothstep-interpolated, amplitude halves per octave."""
    rng = np.random.default_rng(seed)
    total = np.zeros((size, size), np.float64)
    norm = 0.0
    amplitude = 1.0
    period = base_period
    coords = np.arange(size, dtype=np.float64)
    for _ in range(octaves):
        cells = size // period + 2
        lattice = rng.uniform(0.0, 1.0, size=(cells, cells))
        f = coords / period
        i = np.floor(f).astype(np.int64)
        t = f - i
        s = t * t * (3.0 - 2.0 * t)
        a = lattice[i[:, None], i[None, :]]
        b = lattice[i[:, None] + 1, i[None, :]]
        c = lattice[i[:, None], i[None, :] + 1]
        d = lattice[i[:, None] + 1, i[None, :] + 1]
        sx = s[:, None]
        sz = s[None, :]
        total += amplitude * ((a * (1 - sx) + b * sx) * (1 - sz) + (c * (1 - sx) + d * sx) * sz)
        norm += amplitude
        amplitude *= 0.5
        period = max(period // 2, 1)
    return total / norm


def make_materials(seed: int, emission: tuple[float, float], log_uniform: bool):
    rng = np.random.default_rng(seed)
    table = np.zeros((65536, 8), np.float32)
    ids = np.arange(1, 256)
    albedo = rng.uniform(0.2, 0.9, size=(255, 3)).astype(np.float16).astype(np.float32)
    table[OPAQUE | ids, 0:3] = albedo
    lo, hi = emission
    n_e = 128
    if log_uniform:
        e = np.exp(rng.uniform(np.log(lo), np.log(hi), size=n_e))
    else:
        e = rng.uniform(lo, hi, size=n_e)
    tint = rng.uniform(0.7, 1.0, size=(n_e, 3))
    em = (e[:, None] * tint).astype(np.float16).astype(np.float32)
    table[OPAQUE | np.arange(128, 256), 4:7] = em
    lum = em[:, 0] * np.float32(0.2126) + em[:, 1] * np.float32(0.7152) + em[:, 2] * np.float32(0.0722)
    threshold = lum.min()
    codes = np.zeros(65536, bool)
    all_lum = table[:, 4] * np.float32(0.2126) + table[:, 5] * np.float32(0.7152) + table[:, 6] * np.float32(0.0722)
    codes[(all_lum >= threshold) & (all_lum > 0)] = True
    # word i holds codes 32i .. 32i+31, bit b = code 32i + b
    words = np.zeros(2048, np.uint32)
    for b in range(32):
        words |= codes.reshape(2048, 32)[:, b].astype(np.uint32) << np.uint32(b)
    return table, words


def look_at(position, target, width, height, fov_y_deg=60.0):
    p = np.asarray(position, np.float64)
    f = np.asarray(target, np.float64) - p
    f /= np.linalg.norm(f)
    up0 = np.array([0.0, 1.0, 0.0])
    r = np.cross(f, up0)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    return {"position": p.astype(np.float32), "forward": f.astype(np.float32), "right": r.astype(np.float32),
            "up": u.astype(np.float32), "tan_half_fov_y": np.float32(np.tan(np.radians(fov_y_deg) / 2)),
            "aspect": np.float32(width / height), "width": int(width), "height": int(height)}


def terrain(world: int, seed: int, octaves: int, bricks_per_column: int | None, emissive: int,
            emission: tuple[float, float], log_uniform: bool, fill: float, relief: float,
            width: int, height: int, base_period: int | None = None, chunk: int = 1 << 16) -> SceneData:
    """A heightfield world of `world`^3 voxels. bricks_per_column None: every brick cell of the grid is
    resident; else that many bricks per column, from the column's highest surface brick down."""
    dim = world // 8
    base_period = base_period if base_period is not None else max(world // 4, 8)
    h = value_noise(world, seed, octaves, base_period)
    heights = np.clip(np.round(fill * world + (h - 0.5) * relief * world), 1, world - 1).astype(np.int32)
    # heights[x, z]
    col_max = heights.reshape(dim, 8, dim, 8).max(axis=(1, 3))  # [bx, bz]
    if bricks_per_column is None:
        bx, by, bz = np.meshgrid(np.arange(dim), np.arange(dim), np.arange(dim), indexing="ij")
        keys = np.stack([bx.ravel(), by.ravel(), bz.ravel()], axis=1).astype(np.int32)
    else:
        top = (col_max - 1) // 8
        bx, bz = np.meshgrid(np.arange(dim), np.arange(dim), indexing="ij")
        cols = []
        for k in range(bricks_per_column):
            cols.append(np.stack([bx.ravel(), np.maximum(top.ravel() - k, 0), bz.ravel()], axis=1))
        keys = np.concatenate(cols).astype(np.int32)
        keys = np.unique(keys, axis=0)
    rng = np.random.default_rng(seed + 1)
    n = len(keys)
    payload = np.zeros((n, 512), np.uint16)
    v = np.arange(8)
    for start in range(0, n, chunk):
        k = keys[start:start + chunk]
        xs = k[:, 0:1] * 8 + v[None, :]  # [m, 8]
        ys = k[:, 1:2] * 8 + v[None, :]
        zs = k[:, 2:3] * 8 + v[None, :]
        hz = heights[xs[:, None, :], zs[:, :, None]]  # [m, z, x]
        solid = ys[:, None, :, None] < hz[:, :, None, :]  # [m, z, y, x]
        depth = hz[:, :, None, :] - ys[:, None, :, None]
        mat = 1 + ((xs[:, None, None, :] * 7 + zs[:, :, None, None] * 13 + depth * 3) % 127)
        blocks = np.where(solid, OPAQUE | mat, 0).astype(np.uint16)
        payload[start:start + chunk] = blocks.reshape(len(k), 512)
    # Emissive bricks: resident bricks with solid voxels and an open top (the surface bricks).
    solid_count = (payload != 0).sum(axis=1)
    surface = np.nonzero((solid_count > 0) & (solid_count < 512))[0]
    chosen = rng.choice(surface, size=min(emissive, len(surface)), replace=False)
    emissive_ids = rng.integers(128, 256, size=len(chosen))
    for c, e in zip(chosen, emissive_ids):
        row = payload[c]
        row[row != 0] = OPAQUE | e
    table, light_codes = make_materials(seed + 2, emission, log_uniform)
    centre = world / 2
    peak = float(heights.max())
    camera = look_at((centre * 0.55, peak + 0.35 * world, centre * 0.35),
                     (centre, float(np.median(heights)), centre * 1.1), width, height)
    return SceneData(keys=keys, payload=payload, materials=table, light_codes=light_codes, grid_min=(0, 0, 0),
                     grid_dim=dim, group_shift=5, camera=camera, emissive_bricks=np.sort(chosen))


def small_scene(seed: int = 1) -> SceneData:
    """4,096 bricks in 128^3, ~30 % solid, 512 emissive bricks (emission U[0.1, 10]), 256 x 256."""
    return terrain(128, seed, octaves=5, bricks_per_column=None, emissive=512, emission=(0.1, 10.0),
                   log_uniform=False, fill=0.3, relief=0.6, width=256, height=256, base_period=16)


def hd_scene(seed: int = 2) -> SceneData:
    """1,048,576 bricks in 4096^3 (8-octave terrain), 65,536 emissive (log-U[0.01, 100]), 1920 x 1080."""
    return terrain(4096, seed, octaves=8, bricks_per_column=4, emissive=65536, emission=(0.01, 100.0),
                   log_uniform=True, fill=0.45, relief=0.35, width=1920, height=1080)


def furnace_box(size_bricks: int = 8, wall: int = 8, emission: float = 1.0, albedo: float = 0.5,
                width: int = 128, height: int = 128) -> SceneData:
    """A closed box room: every voxel outside [wall, world - wall)^3 is one opaque material with the
    given grey emission and albedo, the inside is air, the camera sits at the centre. Every direction
    from inside sees the same radiance, so the path tracer's expected radiance after B bounces has the
    closed form emission * (1 - albedo^B) / (1 - albedo)."""
    world = size_bricks * 8
    dim = size_bricks
    bx, by, bz = np.meshgrid(np.arange(dim), np.arange(dim), np.arange(dim), indexing="ij")
    keys = np.stack([bx.ravel(), by.ravel(), bz.ravel()], axis=1).astype(np.int32)
    block = OPAQUE | 200
    v = np.arange(512)
    lx, ly, lz = v & 7, (v >> 3) & 7, v >> 6
    x = keys[:, 0:1] * 8 + lx[None, :]
    y = keys[:, 1:2] * 8 + ly[None, :]
    z = keys[:, 2:3] * 8 + lz[None, :]
    inside = ((x >= wall) & (x < world - wall) & (y >= wall) & (y < world - wall) & (z >= wall) & (z < world - wall))
    payload = np.where(inside, 0, block).astype(np.uint16)
    table = np.zeros((65536, 8), np.float32)
    table[block, 0:3] = np.float32(albedo)
    table[block, 4:7] = np.float32(emission)
    light_codes = np.zeros(2048, np.uint32)
    light_codes[block >> 5] |= np.uint32(1) << np.uint32(block & 31)
    c = world / 2
    camera = look_at((c + 0.37, c + 0.21, c + 0.13), (c + 10.0, c - 3.0, c + 7.0), width, height, fov_y_deg=90.0)
    return SceneData(keys=keys, payload=payload, materials=table, light_codes=light_codes, grid_min=(0, 0, 0),
                     grid_dim=dim, group_shift=5, camera=camera, emissive_bricks=np.arange(len(keys)))


def sun_floor(size_bricks: int = 16, albedo: float = 123 / 255, lamp: str = "below", lamp_emission: float = 20.0,
              width: int = 64, height: int = 64, glass_above=None) -> SceneData:
    """A flat floor under an open sky: size_bricks^2 bricks in one layer whose voxels y in [4, 8) are one
    grey opaque material (top face at y = 8), plus one lamp: "below" a 2^3 emissive cube under the
    floor (voxels y 0..1 of the central floor brick; every face it opens lies below the floor top, so
    no point on the floor can receive its light), "above" a 4^3 emissive cube in the air over the
    floor's centre (voxels y 18..21). The camera looks (almost) straight down at the floor's centre
    from y = 120 with a 30 degree field of view, so every primary ray hits the floor or the lamp.
    glass_above = (tint, extinction): also a smooth glass slab (id 0x42, ior 1.5) over the whole floor,
    voxels y 40..43, resident air bricks below and above it (so both of its faces are interfaces), black
    opaque walls (albedo 0) on the world's border from the floor up to the slab (the gap between them
    is closed: grazing rays that the slab's underside reflects end there, and no shadow ray leaves
    under the slab), the vacuum entry for code 0 and the extended material table."""
    dim = size_bricks
    world = dim * 8
    bx, bz = np.meshgrid(np.arange(dim), np.arange(dim), indexing="ij")
    keys = np.stack([bx.ravel(), np.zeros(dim * dim, np.int64), bz.ravel()], axis=1).astype(np.int32)
    floor_block, lamp_block = OPAQUE | 1, OPAQUE | 200
    v = np.arange(512)
    lx, ly, lz = v & 7, (v >> 3) & 7, v >> 6
    payload = np.where(ly[None, :] >= 4, floor_block, 0).repeat(len(keys), axis=0).astype(np.uint16)
    centre = dim // 2
    cube = (lx >= 3) & (lx < 5) & (lz >= 3) & (lz < 5)
    centre_index = int(np.nonzero((keys[:, 0] == centre) & (keys[:, 2] == centre))[0][0])
    if lamp == "below":
        payload[centre_index, cube & (ly < 2)] = lamp_block
        emissive = np.array([centre_index])
    elif lamp == "above":
        lamp_payload = np.zeros((1, 512), np.uint16)
        cube4 = (lx >= 2) & (lx < 6) & (lz >= 2) & (lz < 6) & (ly >= 2) & (ly < 6)
        lamp_payload[0, cube4] = lamp_block
        keys = np.concatenate([keys, np.array([[centre, 2, centre]], np.int32)])
        payload = np.concatenate([payload, lamp_payload])
        emissive = np.array([len(keys) - 1])
    else:
        raise ValueError(f"unknown lamp placement {lamp!r}")
    table = np.zeros((65536, 8), np.float32)
    table[floor_block, 0:3] = np.float32(albedo)
    table[lamp_block, 0:3] = np.float32(0.5)
    table[lamp_block, 4:7] = np.float32(lamp_emission)
    light_codes = np.zeros(2048, np.uint32)
    light_codes[lamp_block >> 5] |= np.uint32(1) << np.uint32(lamp_block & 31)
    ext = None
    if glass_above is not None:
        tint, extinction = glass_above
        slab = np.where(ly < 4, 0x42, 0).astype(np.uint16)
        wall = OPAQUE | 2
        gx = bx.ravel()[:, None] * 8 + (v & 7)[None, :]
        gz = bz.ravel()[:, None] * 8 + (v >> 6)[None, :]
        border = (gx == 0) | (gx == world - 1) | (gz == 0) | (gz == world - 1)
        for layer in (1, 2, 3, 4, 5, 6):
            if layer == 5:
                blocks = np.tile(slab, (dim * dim, 1))
            elif layer == 6:
                blocks = np.zeros((dim * dim, 512), np.uint16)
            else:
                blocks = np.where(border, wall, 0).astype(np.uint16)
            keys = np.concatenate([keys, np.stack([bx.ravel(), np.full(dim * dim, layer), bz.ravel()], axis=1).astype(np.int32)])
            payload = np.concatenate([payload, blocks])
        vacuum(table)
        table[0x42, 0:3] = np.asarray(tint, np.float32)
        table[0x42, 3] = 1.0
        ext = np.zeros((65536, 4), np.float32)
        ext[:, 0] = 1.5
        ext[0x42, 1] = np.float32(extinction)
    c = world / 2
    camera = look_at((c + 0.013, 120.0, c + 0.021), (c, 8.0, c), width, height, fov_y_deg=30.0)
    return SceneData(keys=keys, payload=payload, materials=table, light_codes=light_codes, grid_min=(0, 0, 0),
                     grid_dim=dim, group_shift=5, camera=camera, emissive_bricks=emissive, materials_ext=ext)


def glass_world(encoding: int = 1) -> SceneData:
    """The reference finalize test's glass world (tests/finalize_test.cpp:264-290): bricks (0..2, 0, 0)
    of opaque floor 0x8123 and (0..2, 1, 0) holding a glass slab x 2..21, y 8..11, z 2..5 in air; the
    glass is transparent id 0x42 in either encoding."""
    glass, floor = 0x42, 0x8123
    keys, payload = [], []
    v = np.arange(512)
    lx, ly, lz = v & 7, (v >> 3) & 7, v >> 6
    for bx in range(3):
        keys.append((bx, 0, 0))
        payload.append(np.full(512, floor, np.uint16))
        gx, gy = bx * 8 + lx, 8 + ly
        inside = (gx >= 2) & (gx < 22) & (gy >= 8) & (gy < 12) & (lz >= 2) & (lz < 6)
        keys.append((bx, 1, 0))
        payload.append(np.where(inside, glass, 0).astype(np.uint16))
    table = np.zeros((65536, 8), np.float32)
    table[floor, 0:3] = 0.5
    table[glass, 0:3] = 0.9
    camera = look_at((12.0, 20.0, -10.0), (12.0, 9.0, 4.0), 32, 32)
    return SceneData(keys=np.array(keys, np.int32), payload=np.stack(payload), materials=table,
                     light_codes=np.zeros(2048, np.uint32), grid_min=(0, 0, 0), grid_dim=8, group_shift=4,
                     camera=camera, emissive_bricks=np.zeros(0, np.int64), encoding=encoding)


GLASS = 0x42  # the transparent id of the glass scenes


def vacuum(table: np.ndarray) -> None:
    """Material code 0 (air) as the reference's vacuum entry: albedo 1, smoothness 1 (a glass-to-air
    exit is a smooth interface the delta chain crosses), no emission."""
    table[0, 0:3] = 1.0
    table[0, 3] = 1.0


def glass_slab(size_bricks: int = 6, ior: float = 1.5, extinction: float = 0.3, tint=(0.9, 0.7, 0.5),
               smoothness: float = 1.0, floor_emission=(2.0, 2.0, 2.0), width: int = 64, height: int = 64) -> SceneData:
    """A glass slab on an emissive floor under an open sky: size_bricks^2 floor bricks (voxels y 0..7
    opaque 0x8123, albedo 0.5, emission floor_emission, not a light code, so no NEE) and above them glass
    voxels y 8..11 (transparent id 0x42: tint, smoothness, ior, extinction in the extended table; the
    other codes ior 1.5, extinction 0). The camera looks down at the slab from y = 60 with a 40 degree
    field of view, so every camera ray meets the slab's top face (y = 12) first."""
    dim = size_bricks
    bx, bz = np.meshgrid(np.arange(dim), np.arange(dim), indexing="ij")
    cols = np.stack([bx.ravel(), bz.ravel()], axis=1)
    floor = OPAQUE | 0x123
    v = np.arange(512)
    ly = (v >> 3) & 7
    keys = np.concatenate([np.stack([cols[:, 0], np.zeros(len(cols), np.int64), cols[:, 1]], axis=1),
                           np.stack([cols[:, 0], np.ones(len(cols), np.int64), cols[:, 1]], axis=1)]).astype(np.int32)
    payload = np.concatenate([np.full((len(cols), 512), floor, np.uint16),
                              np.tile(np.where(ly < 4, GLASS, 0).astype(np.uint16), (len(cols), 1))])
    table = np.zeros((65536, 8), np.float32)
    vacuum(table)
    table[floor, 0:3] = 0.5
    table[floor, 4:7] = np.asarray(floor_emission, np.float32)
    table[GLASS, 0:3] = np.asarray(tint, np.float32)
    table[GLASS, 3] = np.float32(smoothness)
    ext = np.zeros((65536, 4), np.float32)
    ext[:, 0] = 1.5
    ext[GLASS, 0] = np.float32(ior)
    ext[GLASS, 1] = np.float32(extinction)
    c = dim * 4.0
    camera = look_at((c + 3.1, 60.0, c - 2.3), (c, 12.0, c), width, height, fov_y_deg=40.0)
    return SceneData(keys=keys, payload=payload, materials=table, light_codes=np.zeros(2048, np.uint32),
                     grid_min=(0, 0, 0), grid_dim=max(8, (dim + 7) // 8 * 8), group_shift=4, camera=camera,
                     emissive_bricks=np.zeros(0, np.int64), materials_ext=ext)


def glass_terrain(seed: int = 1) -> SceneData:
    """The small terrain with the full material model: opaque materials 1..16 become transparent media
    (ids 1..16; 1..8 smooth, delta interfaces; 9..16 rough, shaded stochastically; tints, ior 1.2..1.8,
    extinction 0.02..0.4), materials 17..48 get a specular lobe (smoothness 0.3..1, metallic 0..1,
    ior 1.2..2); the emissive bricks stay lamps."""
    data = small_scene(seed)
    rng = np.random.default_rng(seed + 77)
    payload = np.asarray(data.payload).copy()
    ids = payload & 0x7FFF
    to_glass = ((payload & OPAQUE) != 0) & (ids >= 1) & (ids <= 16)
    payload[to_glass] = ids[to_glass]
    table = np.asarray(data.materials).copy()
    vacuum(table)
    ext = np.zeros((65536, 4), np.float32)
    ext[:, 0] = 1.5
    media = np.arange(1, 17)
    table[media, 0:3] = rng.uniform(0.4, 1.0, size=(16, 3)).astype(np.float16).astype(np.float32)
    table[media, 3] = np.where(media <= 8, 1.0, rng.uniform(0.2, 0.9, size=16)).astype(np.float32)
    ext[media, 0] = rng.uniform(1.2, 1.8, size=16).astype(np.float16).astype(np.float32)
    ext[media, 1] = rng.uniform(0.02, 0.4, size=16).astype(np.float16).astype(np.float32)
    shiny = OPAQUE | np.arange(17, 49)
    table[shiny, 3] = rng.uniform(0.3, 1.0, size=32).astype(np.float16).astype(np.float32)
    table[shiny, 7] = rng.uniform(0.0, 1.0, size=32).astype(np.float16).astype(np.float32)
    ext[shiny, 0] = rng.uniform(1.2, 2.0, size=32).astype(np.float16).astype(np.float32)
    return SceneData(keys=data.keys, payload=payload, materials=table, light_codes=data.light_codes,
                     grid_min=data.grid_min, grid_dim=data.grid_dim, group_shift=data.group_shift, camera=data.camera,
                     emissive_bricks=data.emissive_bricks, materials_ext=ext)


def diff_scene(seed: int = 5) -> SceneData:
    """The differentiable workload: the 1080p terrain scaled to 262,144 bricks (2048^3, 256^2 columns x
    4 bricks), 16,384 emissive bricks (the emitters), emission log-uniform in [0.01, 100], 512 x 512."""
    return terrain(2048, seed, octaves=8, bricks_per_column=4, emissive=16384, emission=(0.01, 100.0),
                   log_uniform=True, fill=0.45, relief=0.35, width=512, height=512)


def chunk_scene(seed: int, bricks: int, grid: int = 64, emissive_fraction: float = 0.05, width: int = 256,
                height: int = 256, materials=None) -> SceneData:
    """One chunk scene of the batched workload: a value-noise terrain in a grid^3-brick cell (grid a
    multiple of 8), exactly `bricks` resident bricks taken layer by layer from the surface down (the
    surface bricks of all columns first, a random subset of columns for a partial layer), about
    `emissive_fraction` of them emissive (at least one; chosen among the bricks with an open face),
    camera from the seed. `materials` (table,
    light codes) is shared by the scenes of a batch; by default the batch palette."""
    world = grid * 8
    rng = np.random.default_rng(seed)
    h = value_noise(world, seed, 5, max(world // 4, 8))
    heights = np.clip(np.round(0.45 * world + (h - 0.5) * 0.4 * world), 1, world - 1).astype(np.int32)
    col_max = heights.reshape(grid, 8, grid, 8).max(axis=(1, 3))
    top = (col_max - 1) // 8
    bx, bz = np.meshgrid(np.arange(grid), np.arange(grid), indexing="ij")
    cols = np.stack([bx.ravel(), bz.ravel()], axis=1)
    keys = []
    remaining = int(bricks)
    layer = 0
    while remaining > 0 and layer < grid:
        ys = top.ravel() - layer
        ok = np.nonzero(ys >= 0)[0]
        if len(ok) == 0:
            break
        take = ok if len(ok) <= remaining else rng.choice(ok, size=remaining, replace=False)
        keys.append(np.stack([cols[take, 0], ys[take], cols[take, 1]], axis=1))
        remaining -= len(take)
        layer += 1
    keys = np.concatenate(keys).astype(np.int32)
    order = np.lexsort((keys[:, 0], keys[:, 1], keys[:, 2]))
    keys = keys[order]
    n = len(keys)
    v = np.arange(8)
    xs = keys[:, 0:1] * 8 + v[None, :]
    ys = keys[:, 1:2] * 8 + v[None, :]
    zs = keys[:, 2:3] * 8 + v[None, :]
    hz = heights[xs[:, None, :], zs[:, :, None]]
    solid = ys[:, None, :, None] < hz[:, :, None, :]
    depth = hz[:, :, None, :] - ys[:, None, :, None]
    mat = 1 + ((xs[:, None, None, :] * 7 + zs[:, :, None, None] * 13 + depth * 3) % 127)
    payload = np.where(solid, OPAQUE | mat, 0).astype(np.uint16).reshape(n, 512)
    # Emitters among the bricks with an open face: a brick that holds both solid and air voxels has a
    # solid voxel next to an air voxel, so its emission reaches the scene (a fully solid brick may be
    # buried and give no light face).
    solid_count = (payload != 0).sum(axis=1)
    candidates = np.nonzero((solid_count > 0) & (solid_count < 512))[0]
    if len(candidates) == 0:
        candidates = np.nonzero(solid_count > 0)[0]
    emissive = max(1, int(round(emissive_fraction * n)))
    chosen = rng.choice(candidates, size=min(emissive, len(candidates)), replace=False)
    emissive_ids = rng.integers(128, 256, size=len(chosen))
    for c, e in zip(chosen, emissive_ids):
        row = payload[c]
        row[row != 0] = OPAQUE | e
    table, light_codes = materials if materials is not None else make_materials(4242, (0.01, 100.0), True)
    peak = float(heights.max())
    c = world / 2
    camera = look_at((c * 0.3 + rng.uniform(-8, 8), peak + 0.3 * world, c * 0.2 + rng.uniform(-8, 8)),
                     (c, float(np.median(heights)), c * 1.2), width, height)
    return SceneData(keys=keys, payload=payload, materials=table, light_codes=light_codes, grid_min=(0, 0, 0),
                     grid_dim=grid, group_shift=5, camera=camera, emissive_bricks=np.sort(chosen))


def chunk_batch(count: int = 256, seed: int = 9, low: int = 64, high: int = 65536, grid: int = 64,
                emissive_fraction: float = 0.05, width: int = 256, height: int = 256) -> list:
    """The batched workload: `count` chunk scenes with brick counts log-uniform in [low, high]."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(low), np.log(high), size=count)).astype(np.int64)
    palette = make_materials(4242, (0.01, 100.0), True)
    return [chunk_scene(seed * 1000 + i, int(sizes[i]), grid, emissive_fraction, width, height, palette)
            for i in range(count)]
