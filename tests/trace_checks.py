"""Independent checks for the trace side (numpy, float64): a dense voxel world, a brute-force voxel
walk, and recomputations of the finalize words and the brick light records straight from the
block payloads. None of these share code with the oracle or the kernels."""
from __future__ import annotations

import numpy as np

OPAQUE = 0x8000
FACE_AXIS = np.array([0, 0, 1, 1, 2, 2])
FACE_SIGN = np.array([1, -1, 1, -1, 1, -1])


def dense_blocks(data) -> np.ndarray:
    """blocks[x, y, z] (uint16) of the whole grid; bricks that are not resident are air."""
    dim = data.grid_dim
    world = np.zeros((dim * 8, dim * 8, dim * 8), np.uint16)
    g = np.asarray(data.keys, np.int64) - np.asarray(data.grid_min, np.int64)[None, :]
    v = np.arange(512)
    lx, ly, lz = v & 7, (v >> 3) & 7, v >> 6
    x = g[:, 0:1] * 8 + lx[None, :]
    y = g[:, 1:2] * 8 + ly[None, :]
    z = g[:, 2:3] * 8 + lz[None, :]
    world[x, y, z] = data.payload
    return world


def is_light(blocks: np.ndarray, light_codes: np.ndarray) -> np.ndarray:
    """Light voxels: the material code (an opaque value, a transparent id without face bits) is a light
    code (block-ID encoding)."""
    b = blocks.astype(np.int64)
    b = np.where((b & OPAQUE) != 0, b, b & 0x1FF)
    return ((light_codes[b >> 5] >> (b & 31).astype(np.uint32)) & 1).astype(bool)


def walk_first_solid(solid: np.ndarray, o, d, max_steps: int = 4096):
    """Amanatides-Woo in float64 from origin o along d through the unit voxel grid `solid`.
    Returns (voxel or None, t_entry, entry_axis, tie) where tie is the smallest gap between two
    consecutive plane crossings seen (a small gap means the float32 DDA may cut a corner)."""
    n = np.array(solid.shape)
    o = np.asarray(o, np.float64)
    d = np.asarray(d, np.float64)
    inv = np.where(np.abs(d) > 1e-30, 1.0 / np.where(d == 0, 1.0, d), np.inf)
    t_enter, t_exit = 0.0, np.inf
    for a in range(3):
        if d[a] == 0.0:
            if o[a] < 0 or o[a] >= n[a]:
                return None, 0.0, -1, np.inf
            continue
        t0 = (0.0 - o[a]) * inv[a]
        t1 = (n[a] - o[a]) * inv[a]
        t_enter = max(t_enter, min(t0, t1))
        t_exit = min(t_exit, max(t0, t1))
    if t_enter > t_exit:
        return None, 0.0, -1, np.inf
    p = o + d * t_enter
    cell = np.clip(np.floor(p).astype(np.int64), 0, n - 1)
    step = np.where(d >= 0, 1, -1)
    t_next = np.where(d != 0, ((cell + (step > 0)) - o) * inv, np.inf)
    t_delta = np.abs(inv)
    t = t_enter
    axis = -1
    tie = np.inf
    for _ in range(max_steps):
        if solid[cell[0], cell[1], cell[2]]:
            return cell.copy(), t, axis, tie
        order = np.argsort(t_next)
        tie = min(tie, t_next[order[1]] - t_next[order[0]])
        axis = int(order[0])
        t = t_next[axis]
        cell[axis] += step[axis]
        if cell[axis] < 0 or cell[axis] >= n[axis]:
            return None, t, axis, tie
        t_next[axis] += t_delta[axis]
    return None, t, axis, tie


ENCODINGS = {0: (0xC000, 0x00FF, 8), 1: (0x8000, 0x01FF, 9)}  # opaque mask, id mask, face shift


def finalize_payload(data, encoding: int = 1):
    """(finalized payload [n, 512] uint16, occupied [n, 512] bool) by the rules of the reference's
    finalize test (tests/finalize_test.cpp:74-116, Reference()) over each brick and its face neighbours
    (absent bricks are air): a transparent or air voxel takes the face bits of the neighbours that are
    not opaque and of another medium; a voxel is occupied when alive (fewer than six opaque neighbours)
    and not 0 and, as the kernel adds (caelix_finalize.hlsl:244-247), an interface (opaque or a face
    bit). That last term differs from the test's Reference() only for a transparent voxel with no face."""
    opaque_mask, id_mask, face_shift = ENCODINGS[encoding]
    raw = np.asarray(data.payload).astype(np.int64)
    n = len(raw)
    keys = np.asarray(data.keys, np.int64)
    index = {tuple(k): i for i, k in enumerate(keys.tolist())}
    cube = raw.reshape(n, 8, 8, 8)  # [brick, z, y, x]
    pad = np.zeros((n, 10, 10, 10), np.int64)
    pad[:, 1:9, 1:9, 1:9] = cube
    for axis, sign in ((0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1)):
        step = np.zeros(3, np.int64)
        step[axis] = sign
        nb = np.array([index.get(tuple(k), -1) for k in (keys + step[None, :]).tolist()], np.int64)
        ok = nb >= 0
        src = cube[np.where(ok, nb, 0)] * ok[:, None, None, None]
        dim = 3 - axis  # array axis of this world axis ([brick, z, y, x])
        take = 0 if sign > 0 else 7
        put = 9 if sign > 0 else 0
        sl = [slice(None), slice(1, 9), slice(1, 9), slice(1, 9)]
        sl_src = [slice(None), slice(None), slice(None), slice(None)]
        sl[dim] = put
        sl_src[dim] = take
        pad[tuple(sl)] = src[tuple(sl_src)]
    opaque = (cube & opaque_mask) != 0
    own = np.where(opaque, cube, cube & id_mask)
    opaque_count = np.zeros_like(own)
    faces = np.zeros_like(own)
    for d, (dx, dy, dz) in enumerate(((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))):
        nbv = pad[:, 1 + dz:9 + dz, 1 + dy:9 + dy, 1 + dx:9 + dx]
        n_opaque = (nbv & opaque_mask) != 0
        opaque_count += n_opaque
        faces |= np.where(~n_opaque & ((nbv & id_mask) != own), 1 << d, 0)
    result = np.where(opaque, own, own | (faces << face_shift))
    occupied = (opaque_count < 6) & (result != 0) & (opaque | (faces != 0))
    return result.reshape(n, 512).astype(np.uint16), occupied.reshape(n, 512)


def finalize_expected(data, encoding: int = 1):
    """(info words [n, 2] with the light count left 0, occupancy [n, 8] uint64, brick_bits, super_bits)
    of the finalized bricks (finalize_payload's occupied voxels)."""
    _, solid = finalize_payload(data, encoding)
    n = len(solid)
    v = np.arange(512)
    x, y, z = v & 7, (v >> 3) & 7, v >> 6
    coarse_of = (x >> 2) | ((y >> 2) << 1) | ((z >> 2) << 2)
    micro_of = (x & 3) | ((y & 3) << 2) | ((z & 3) << 4)
    occ = np.zeros((n, 8), np.uint64)
    for c in range(8):
        sel = coarse_of == c
        bits = (np.uint64(1) << micro_of[sel].astype(np.uint64))[None, :]
        occ[:, c] = np.bitwise_or.reduce(np.where(solid[:, sel], bits, np.uint64(0)), axis=1)
    coarse = np.zeros(n, np.uint32)
    for c in range(8):
        coarse |= (occ[:, c] != 0).astype(np.uint32) << np.uint32(c)
    keys = np.asarray(data.keys, np.int64)
    g = data.group_shift
    mask = (1 << g) - 1
    pos = (keys[:, 0] & mask) | ((keys[:, 1] & mask) << g) | ((keys[:, 2] & mask) << (2 * g))
    w0 = (pos & 0xFFFF).astype(np.uint32) | (coarse << np.uint32(16))
    any_ = solid.any(axis=1)
    big = 99
    lo = [np.where(solid, c[None, :], big).min(axis=1) for c in (x, y, z)]
    hi = [np.where(solid, c[None, :], -1).max(axis=1) for c in (x, y, z)]
    w1 = (lo[0] | (lo[1] << 3) | (lo[2] << 6) | (hi[0] << 9) | (hi[1] << 12) | (hi[2] << 15))
    w1 = np.where(any_, w1, 0).astype(np.uint32)
    dim = data.grid_dim
    gk = keys - np.asarray(data.grid_min, np.int64)[None, :]
    cells = dim ** 3
    brick_bits = np.zeros((cells + 31) // 32, np.uint32)
    super_bits = np.zeros((cells // 512 + 31) // 32 + 1, np.uint32)
    for i in np.nonzero(any_)[0]:
        c = (gk[i, 2] * dim + gk[i, 1]) * dim + gk[i, 0]
        brick_bits[c >> 5] |= np.uint32(1) << np.uint32(c & 31)
        sd = dim >> 3
        s = ((gk[i, 2] >> 3) * sd + (gk[i, 1] >> 3)) * sd + (gk[i, 0] >> 3)
        super_bits[s >> 5] |= np.uint32(1) << np.uint32(s & 31)
    return np.stack([w0, w1], axis=1), occ, brick_bits, super_bits


def open_light_faces(data) -> np.ndarray:
    """open[n, 512, 6] bool: voxel v of brick i is a light and its face f borders a voxel that is not
    opaque (air, or a brick that is not resident)."""
    world = dense_blocks(data)
    light = is_light(world, np.asarray(data.light_codes, np.uint32))
    opaque = (world & OPAQUE) != 0
    padded = np.zeros(tuple(s + 2 for s in opaque.shape), bool)
    padded[1:-1, 1:-1, 1:-1] = opaque
    g = np.asarray(data.keys, np.int64) - np.asarray(data.grid_min, np.int64)[None, :]
    v = np.arange(512)
    x = g[:, 0:1] * 8 + (v & 7)[None, :]
    y = g[:, 1:2] * 8 + ((v >> 3) & 7)[None, :]
    z = g[:, 2:3] * 8 + (v >> 6)[None, :]
    out = np.zeros((len(g), 512, 6), bool)
    lit = light[x, y, z]
    for f in range(6):
        dx = [0, 0, 0]
        dx[FACE_AXIS[f]] = FACE_SIGN[f]
        nb = padded[x + 1 + dx[0], y + 1 + dx[1], z + 1 + dx[2]]
        out[:, :, f] = lit & ~nb
    return out


def primary_dirs(cam):
    """The primary directions of a pinhole camera in float32, operation by operation as the kernels
    and the oracle compute them (CaelixPrimaryDirection)."""
    f32 = np.float32
    w, h = cam["width"], cam["height"]
    px, py = np.meshgrid(np.arange(w), np.arange(h))
    px, py = px.ravel(), py.ravel()
    ndc_x = (((px.astype(f32) + f32(0.5)) / f32(w)) * f32(2.0)) - f32(1.0)
    ndc_y = f32(1.0) - ((py.astype(f32) + f32(0.5)) / f32(h)) * f32(2.0)
    t, a = f32(cam["tan_half_fov_y"]), f32(cam["aspect"])
    right, up, fwd = (np.asarray(cam[n], f32) for n in ("right", "up", "forward"))
    sx = (ndc_x * t) * a
    sy = ndc_y * t
    v = (right[None, :] * sx[:, None] + up[None, :] * sy[:, None]) + fwd[None, :]
    dp = np.maximum(f32(1.175494351e-38), (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
    return v * (f32(1.0) / np.sqrt(dp))[:, None]


def walk_sparse(keys: np.ndarray, payload: np.ndarray, opaque_mask: int, origins: np.ndarray, dirs: np.ndarray,
                max_steps: int = 8192):
    """Amanatides-Woo in float64 for many rays at once through a sparse world of bricks (keys [n, 3],
    payload [n, 512]): the first voxel whose block has a bit of opaque_mask. Returns (voxel [m, 3]
    int64, t_entry, entry_axis, hit flag, tie), tie = the smallest gap between two consecutive plane
    crossings on the way (the float32 DDA may cut a corner where it is small)."""
    keys = np.asarray(keys, np.int64)
    lo_b = keys.min(axis=0)
    shape = keys.max(axis=0) - lo_b + 1
    index = np.full(tuple(shape), -1, np.int64)
    index[tuple((keys - lo_b).T)] = np.arange(len(keys))
    box_lo = (lo_b * 8).astype(np.float64)
    box_hi = ((lo_b + shape) * 8).astype(np.float64)
    o = np.asarray(origins, np.float64)
    d = np.asarray(dirs, np.float64)
    m = len(o)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.where(d != 0.0, 1.0 / d, np.inf)
        t0 = (box_lo[None, :] - o) * inv
        t1 = (box_hi[None, :] - o) * inv
    t0 = np.where(d == 0.0, np.where((o >= box_lo) & (o < box_hi), -np.inf, np.inf), t0)
    t1 = np.where(d == 0.0, np.where((o >= box_lo) & (o < box_hi), np.inf, -np.inf), t1)
    near = np.minimum(t0, t1)
    t_enter = np.maximum(near.max(axis=1), 0.0)
    t_exit = np.maximum(t0, t1).min(axis=1)
    active = t_enter <= t_exit
    p = o + d * t_enter[:, None]
    cell = np.clip(np.floor(p), box_lo, box_hi - 1).astype(np.int64)
    step = np.where(d >= 0, 1, -1)
    with np.errstate(invalid="ignore"):
        t_next = np.where(d != 0, ((cell + (step > 0)) - o) * inv, np.inf)
    t_delta = np.abs(inv)
    t = t_enter.copy()
    # a ray that starts outside the box enters through the face of its last slab
    axis = np.where(near.max(axis=1) > 0.0, near.argmax(axis=1), -1)
    tie = np.full(m, np.inf)
    hit = np.zeros(m, bool)
    lo_v, hi_v = (lo_b * 8), ((lo_b + shape) * 8)
    rows = np.arange(m)
    for _ in range(max_steps):
        idx = np.nonzero(active)[0]
        if idx.size == 0:
            break
        c = cell[idx]
        brick = index[tuple(((c >> 3) - lo_b).T)]
        local = (c[:, 0] & 7) + 8 * (c[:, 1] & 7) + 64 * (c[:, 2] & 7)
        block = np.where(brick >= 0, payload[np.maximum(brick, 0), local], 0)
        solid = (block & opaque_mask) != 0
        hit[idx[solid]] = True
        active[idx[solid]] = False
        idx = idx[~solid]
        tn = t_next[idx]
        order = np.argsort(tn, axis=1)
        first = order[:, 0]
        r = np.arange(len(idx))
        tie[idx] = np.minimum(tie[idx], tn[r, order[:, 1]] - tn[r, first])
        axis[idx] = first
        t[idx] = tn[r, first]
        cell[idx, first] += step[idx, first]
        t_next[idx, first] += t_delta[idx, first]
        out = (cell[idx, first] < lo_v[first]) | (cell[idx, first] >= hi_v[first])
        active[idx[out]] = False
    del rows
    return cell, t, axis, hit, tie
