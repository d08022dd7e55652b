"""The light tree's structure checked against the scene (the reference's CheckStructure,
tests/light_tree_test.cpp:641-793), computed from scratch with numpy: per group, its row (light bricks,
light faces, a block, flags cleared), its block (header record count; index entries in strictly
increasing Morton order of the brick position, dense id = the brick's record; records at their
offsets, pos, tight box of the light voxels, entry count, per face family the (voxel, block) entries in
voxel order, face sums of the emission luminance within 1e-5; the block's size exactly header + index
+ records), each brick's leaf in the tree of its entity's motion pointing back at (group, entry) with a
world power not below the local one; the trees' leaf counts. Works on the oracle's tables and on the
GPU's read back, in one normalised form (`tables`)."""
from __future__ import annotations

import numpy as np

from caelix_amd import layout

ROW = layout.ROW


def part1by2(v):
    v = np.asarray(v, np.uint64) & 0x3FF
    v = (v | (v << 16)) & 0x030000FF
    v = (v | (v << 8)) & 0x0300F00F
    v = (v | (v << 4)) & 0x030C30C3
    v = (v | (v << 2)) & 0x09249249
    return v


def morton_of_pos(pos, shift):
    pos = np.asarray(pos, np.uint64)
    m = (1 << shift) - 1
    return part1by2(pos & m) | (part1by2((pos >> shift) & m) << 1) | (part1by2((pos >> (2 * shift)) & m) << 2)


def record_units(entries: int) -> int:
    return (48 + 4 * entries + 15) // 16


def from_oracle(lt) -> dict:
    """Normalised tables of an oracle LightTables."""
    control = lt.control[0]
    return {"rows": lt.rows.copy(), "blocks": lt.blocks(),
            "leaves": [lt.leaves[t].view(layout.LEAF) for t in range(2)],
            "nodes": [lt.nodes[t] for t in range(2)],
            "headers": [control["headers"][t].view(layout.HEADER)[0] for t in range(2)],
            "control": lt.control.view(np.uint8).view(layout.CONTROL)[0]}


def from_gpu(tables: dict) -> dict:
    """Normalised tables of a GPU light tree read back (caelix_amd.light_tables)."""
    control = tables["control"].view(layout.CONTROL)[0]
    rows = tables["rows"].view(ROW)
    heap = tables["heap%d" % (int(control["current"]) & 1)].view(np.uint32)
    blocks = {}
    for g, r in enumerate(rows):
        if int(r["block_units"]):
            o = int(r["block_offset"]) * 4
            blocks[g] = heap[o:o + int(r["block_units"]) * 4]
    return {"rows": rows.copy(), "blocks": blocks, "leaves": [tables["leaves%d" % t].view(layout.LEAF) for t in range(2)],
            "nodes": [tables["nodes%d" % t] for t in range(2)],
            "headers": [control["headers"][t] for t in range(2)], "control": control}


def open_faces_sparse(keys, payload, light_codes, encoding: int = 1) -> dict:
    """{brick index: open[512, 6] bool} for the bricks with a light voxel: voxel v (x + 8y + 64z) is a light
    (its material code's bit in light_codes) and its face f borders a voxel that is not opaque (a
    neighbour brick through the keys; air where there is none)."""
    keys = np.asarray(keys, np.int64)
    payload = np.asarray(payload, np.uint16)
    codes = np.asarray(light_codes, np.uint32)
    opaque_mask = layout.PACKED_OPAQUE_MASK if encoding == layout.ENCODING_PACKED else layout.BLOCK_ID_OPAQUE_MASK
    id_mask = layout.PACKED_ID_MASK if encoding == layout.ENCODING_PACKED else layout.BLOCK_ID_ID_MASK
    code = np.where(payload & opaque_mask, payload, payload & id_mask).astype(np.int64)
    light = ((codes[code >> 5] >> (code & 31).astype(np.uint32)) & 1).astype(bool)
    index = {tuple(k): i for i, k in enumerate(keys.tolist())}
    out = {}
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for b in np.nonzero(light.any(axis=1))[0]:
        pad = np.zeros((10, 10, 10), bool)  # [z, y, x] opaque, one voxel of border
        pad[1:9, 1:9, 1:9] = (payload[b].reshape(8, 8, 8) & opaque_mask) != 0
        k = keys[b]
        for dx, dy, dz in dirs:
            n = index.get((int(k[0] + dx), int(k[1] + dy), int(k[2] + dz)))
            if n is None:
                continue
            o = (payload[n].reshape(8, 8, 8) & opaque_mask) != 0
            if dx:
                pad[1:9, 1:9, 9 if dx > 0 else 0] = o[:, :, 0 if dx > 0 else 7]
            elif dy:
                pad[1:9, 9 if dy > 0 else 0, 1:9] = o[:, 0 if dy > 0 else 7, :]
            else:
                pad[9 if dz > 0 else 0, 1:9, 1:9] = o[0 if dz > 0 else 7, :, :]
        open_ = np.zeros((8, 8, 8, 6), bool)
        lit = light[b].reshape(8, 8, 8)
        for f, (dx, dy, dz) in enumerate(dirs):
            nb = pad[1 + dz:9 + dz, 1 + dy:9 + dy, 1 + dx:9 + dx]
            open_[..., f] = lit & ~nb
        out[int(b)] = open_.reshape(512, 6)
    return out


def expected_bricks(data, group_of_brick, group_coords, record_ids=None):
    """Per group, per light brick record: its families [(voxel, block)], face sums (float64), group-local
    box (lo, hi), face count, pos. record_ids[i]: the record of data brick i (default i)."""
    shift = int(data.group_shift)
    faces = open_faces_sparse(data.keys, data.payload, data.light_codes, int(getattr(data, "encoding", 1)))
    e = np.asarray(data.materials, np.float32).reshape(65536, 8)
    lum = (e[:, 4] * np.float32(0.2126) + e[:, 5] * np.float32(0.7152) + e[:, 6] * np.float32(0.0722)).astype(np.float64)
    mask = (1 << shift) - 1
    out = {}
    for b, open_ in faces.items():
        count = int(open_.sum())
        if count == 0:
            continue
        g = int(group_of_brick[b])
        key = np.asarray(data.keys[b], np.int64)
        bxyz = key & mask
        fams = []
        sums = np.zeros(6)
        for f in range(6):
            vox = np.nonzero(open_[:, f])[0]
            fams.append([(int(v), int(data.payload[b, v])) for v in vox])
            sums[f] = lum[canonical_codes(data.payload[b, vox], data)].sum()
        vox = np.nonzero(open_.any(axis=1))[0]
        vxyz = np.stack([vox & 7, (vox >> 3) & 7, vox >> 6], axis=1)
        lo = bxyz * 8 + vxyz.min(axis=0)
        hi = bxyz * 8 + vxyz.max(axis=0) + 1
        pos = int(bxyz[0] | (bxyz[1] << shift) | (bxyz[2] << (2 * shift)))
        rid = int(b) if record_ids is None else int(record_ids[b])
        out.setdefault(g, {})[rid] = {"fams": fams, "sums": sums, "lo": lo, "hi": hi, "faces": count, "pos": pos}
    return out


def canonical_codes(blocks, data) -> np.ndarray:
    blocks = np.asarray(blocks, np.int64)
    if int(getattr(data, "encoding", 1)) == layout.ENCODING_PACKED:
        return np.where(blocks & layout.PACKED_OPAQUE_MASK, blocks, blocks & layout.PACKED_ID_MASK)
    return np.where(blocks & layout.BLOCK_ID_OPAQUE_MASK, blocks, blocks & layout.BLOCK_ID_ID_MASK)


def canonical(block: int, data) -> int:
    """BlockMaterialCode: an opaque value as it is, a transparent one without its face bits."""
    if int(getattr(data, "encoding", 1)) == layout.ENCODING_PACKED:
        return block if block & layout.PACKED_OPAQUE_MASK else block & layout.PACKED_ID_MASK
    return block if block & layout.BLOCK_ID_OPAQUE_MASK else block & layout.BLOCK_ID_ID_MASK


def pack(v) -> int:
    return int(v[0]) | (int(v[1]) << 10) | (int(v[2]) << 20)


def check_structure(data, group_of_brick, group_coords, tables: dict, group_entity=None, motion=None,
                    record_ids=None) -> dict:
    """CheckStructure over normalised tables; returns the totals (bricks, faces, per-tree bricks).
    group_of_brick[i], record_ids[i]: the group and the record of data brick i."""
    shift = int(data.group_shift)
    groups = len(group_coords)
    group_entity = np.zeros(groups, np.int64) if group_entity is None else np.asarray(group_entity)
    motion = np.zeros(int(group_entity.max()) + 1 if groups else 1, np.int64) if motion is None else np.asarray(motion)
    want = expected_bricks(data, group_of_brick, group_coords, record_ids)
    rows, blocks, leaves = tables["rows"], tables["blocks"], tables["leaves"]
    total_bricks = total_faces = 0
    tree_bricks = [0, 0]
    for g in range(groups):
        row = rows[g]
        bricks = want.get(g)
        if bricks is None:
            assert int(row["bricks"]) & 0xFFFF == 0 and int(row["light_faces"]) == 0 and int(row["block_units"]) == 0, g
            continue
        faces = sum(b["faces"] for b in bricks.values())
        total_bricks += len(bricks)
        total_faces += faces
        assert int(row["bricks"]) & 0xFFFF == len(bricks), g
        assert int(row["light_faces"]) == faces, g
        assert int(row["block_units"]) != 0 and int(row["cpu_flags"]) == 0 and int(row["gpu_flags"]) == 0, g
        block = blocks[g]
        count = int(block[0])
        assert count == len(bricks), g
        records_offset = 16 + 16 * count
        previous = -1
        record_bytes = 0
        tree = int(motion[group_entity[g]]) & 1
        for r in range(count):
            pos, dense, offset, bvh_leaf = (int(x) for x in block[4 + 4 * r:8 + 4 * r])
            assert dense in bricks, (g, r, dense)
            want_b = bricks[dense]
            assert pos == want_b["pos"], (g, r)
            m = int(morton_of_pos(pos, shift))
            assert m > previous, (g, r)
            previous = m
            at = offset * 4
            assert offset * 16 >= records_offset and at + 12 <= len(block), (g, r)
            rec = block[at:]
            assert int(rec[8]) == pos
            assert int(rec[6]) == pack(want_b["lo"]) and int(rec[7]) == pack(want_b["hi"]), (g, r)
            assert int(rec[9]) == want_b["faces"], (g, r)
            record_bytes += record_units(int(rec[9])) * 16
            first = 12
            face = rec[:6].view(np.float32)
            for f in range(6):
                word = int(rec[10]) if f < 3 else int(rec[11])
                fc = (word >> (10 * (f % 3))) & 0x3FF
                expect = want_b["fams"][f]
                assert fc == len(expect), (g, r, f)
                vals = rec[first:first + fc].astype(np.int64)
                got = [(int(v & 0x1FF), canonical(int(v >> 16), data)) for v in vals]
                assert got == [(v, canonical(b, data)) for v, b in expect], (g, r, f)
                assert abs(float(face[f]) - want_b["sums"][f]) <= 1e-5 * max(1.0, abs(want_b["sums"][f])), (g, r, f)
                first += fc
            assert bvh_leaf != layout.LT_NONE and bvh_leaf >> 31 == tree, (g, r)
            leaf = leaves[tree][bvh_leaf & 0x7FFFFFFF]
            assert int(leaf["group"]) == g and int(leaf["entry"]) == r, (g, r)
            local = float(np.sum(face.astype(np.float64)))
            world = float(np.sum(leaf["summary"]["face"].astype(np.float64)))
            assert world >= local * (1.0 - 1e-5), (g, r)
            tree_bricks[tree] += 1
        assert records_offset + record_bytes == len(block) * 4, g
    for t in range(2):
        assert int(tables["headers"][t]["leaf_count"]) == tree_bricks[t], t
    c = tables["control"]
    assert int(c["light_bricks"]) == total_bricks and int(c["light_faces"]) == total_faces
    assert int(c["light_groups"]) == len(want) and int(c["bricks_dropped"]) == 0
    return {"bricks": total_bricks, "faces": total_faces, "trees": tree_bricks}


def compare_tables(got: dict, want: dict) -> None:
    """Two normalised light trees bit-exact: rows, every group's block (header, index entries with their
    leaf words, records), both trees' headers, nodes and leaves, and the pass totals."""
    assert got["rows"].tobytes() == want["rows"].tobytes()
    assert sorted(got["blocks"]) == sorted(want["blocks"])
    for g, block in want["blocks"].items():
        assert np.array_equal(got["blocks"][g], block), g
    for t in range(2):
        h, w = got["headers"][t], want["headers"][t]
        assert h.tobytes() == w.tobytes(), t
        if int(w["root"]) != layout.LT_NONE:
            n = int(w["node_count"])
            assert got_nodes(got, t, n) == got_nodes(want, t, n), t
        leaves = int(w["leaf_count"])
        assert got["leaves"][t][:leaves].tobytes() == want["leaves"][t][:leaves].tobytes(), t
    for name in ("light_groups", "bricks_dropped", "light_bricks", "light_faces", "records_recomputed",
                 "records_copied"):
        assert int(got["control"][name]) == int(want["control"][name]), name


def got_nodes(tables: dict, tree: int, count: int) -> bytes:
    return tables["nodes"][tree][:count * 224].tobytes()
