"""CPU: the .cxw reader (cx_cxw_* through caelix_amd.cxw) on hand-built files, restating the cases of
the reference reader's tests (host-rs/caelix-cxw/tests/synthetic.rs:128-540): version gates and
migrations (v2..v8), payload details (free list, brick id remapping, all-air bricks, extra slots,
negative coords) and every error kind. The reference's fixture file (packed-materials.cxw, from
another repository) is not in the reference tree, so the files here are written by caelix_amd.cxw's
writer; the expected values are the reference tests' own literals. Pure host code: no GPU."""
from __future__ import annotations

import struct

import numpy as np
import pytest

from caelix_amd import cxw


def brick(seed: int, non_zero: int) -> np.ndarray:
    """synthetic.rs:248-257: values seed + i | 0x8000 for the first non_zero voxels, then air."""
    v = np.zeros(512, np.uint16)
    i = np.arange(non_zero)
    v[:non_zero] = ((seed + i) & 0xFFFF) | 0x8000
    return v


def preview_word(bi: int) -> int:
    return 0xA5A00000 | (bi << 8) | 0x0F


def sector(coord, bricks, **kw) -> cxw.WSector:
    return cxw.WSector(tuple(coord), bricks, previews={bi: preview_word(bi) for bi, _ in bricks}, **kw)


def simple_sector(coord) -> cxw.WSector:
    return sector(coord, [(0, brick(1, 10))])


@pytest.fixture
def write(tmp_path):
    counter = [0]

    def _write(entities, version=8, table_first=True, **kw):
        counter[0] += 1
        path = tmp_path / f"world{counter[0]}.cxw"
        data = cxw.write(path, entities, version=version, table_first=table_first, **kw)
        return path, data
    return _write


def open_bytes(tmp_path, data: bytes):
    path = tmp_path / "raw.cxw"
    path.write_bytes(data)
    return cxw.CxwFile(path)


def base():
    return [cxw.WEntity([simple_sector((0, 0, 0)), simple_sector((1, 0, 0))])]


@pytest.mark.parametrize("table_first", [True, False])
def test_v8_world_record_and_payload_details(write, table_first):
    b_far = brick(100, 512)
    b_some = brick(7, 3)
    s = sector((-2, 0, 3), [(15 + 256, b_far), (4 + 16 * 2 + 256 * 15, np.zeros(512, np.uint16)), (16 * 15, b_some)],
               free=2, physics_slot_first=True)
    e = cxw.WEntity([s, simple_sector((0, 0, 0))], flags_byte=0b1111_1101, velocity=(1, 2, 3, 4, 5, 6),
                    protected=1, require=0x0003)
    path, _ = write([e, cxw.WEntity([])], world=(0xFF, (10.0, 20.0, 30.0), (0.0, 0.6, 0.0, 0.8)),
                    table_first=table_first)
    f = cxw.CxwFile(path)
    w = f.world
    assert w.version == 8
    assert w.player_pose[0] == [10.0, 20.0, 30.0]
    assert np.allclose(w.player_pose[1], [0.0, 0.6, 0.0, 0.8])
    assert w.player_pose[1] == [float(np.float32(v)) for v in (0.0, 0.6, 0.0, 0.8)]
    assert len(w.entities) == 2
    e0 = w.entities[0]
    assert e0.flags == 1 and e0.has_body and not e0.is_static
    assert e0.velocity == [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]
    assert e0.protected
    assert e0.require_update_flags == 3  # raw value, no v<8 conversion
    assert e0.position == [1.0, 2.0, 3.0]
    assert e0.sector_count == 2 and w.entities[1].sector_count == 0
    assert e0.guid_string() == "04030201-0605-0807-090a-0b0c0d0e0f10"
    assert len(f.sectors) == 2
    d = f.decode(0)
    assert d.coord == (-2, 0, 3) and d.entity == 0
    # ascending absolute brick index: (0, 15, 0) = 240 before (15, 0, 1) = 271
    assert d.keys.tolist() == [[-32, 15, 48], [-17, 0, 49]]
    assert d.summaries.tolist() == [preview_word(240), preview_word(271)]
    assert np.array_equal(d.blocks[0], b_some) and np.array_equal(d.blocks[1], b_far)
    assert d.voxels == 3 + 512
    all_ = f.decode_all(threads=2)
    assert all_["keys"].tolist() == [[-32, 15, 48], [-17, 0, 49], [0, 0, 0]]
    assert all_["voxels"] == 3 + 512 + 10
    assert np.array_equal(all_["blocks"][:2], d.blocks)
    assert all_["summaries"][:2].tolist() == d.summaries.tolist()


def test_v7_without_pose(write):
    path, _ = write([cxw.WEntity([simple_sector((0, 0, 0))])], version=7, world=(0, (5.0,) * 3, (0.0, 0.0, 0.0, 1.0)))
    f = cxw.CxwFile(path)
    assert f.world.player_pose is None
    assert f.world.entities[0].position == [1.0, 2.0, 3.0]
    assert f.decode(0).voxels == 10


@pytest.mark.parametrize("version", [3, 4, 5])
def test_legacy_flag_migration_v3_to_v5(write, version):
    entities = [cxw.WEntity([simple_sector((state, 0, 0))], flags_byte=state, velocity=(0.5,) * 6, protected=1)
                for state in (0, 1, 2, 9)]
    path, _ = write(entities, version=version)
    f = cxw.CxwFile(path)
    assert f.world.player_pose is None
    # off -> Static, static -> HasBody|Static, dynamic -> HasBody, unknown -> Static
    assert [e.flags for e in f.world.entities] == [2, 3, 1, 2]
    for e in f.world.entities:
        assert e.velocity == ([0.5] * 6 if version >= 4 else [0.0] * 6)
        assert e.protected == (version >= 5)
    assert len(f.sectors) == 4
    for i, s in enumerate(f.sectors):
        assert f.decode(i).keys.tolist() == [[s.coord[0] * 16, 0, 0]]


def test_v2_has_no_flags_byte(write):
    path, _ = write([cxw.WEntity([simple_sector((0, -1, 0))])], version=2)
    f = cxw.CxwFile(path)
    e = f.world.entities[0]
    assert e.flags == 2 and not e.protected and e.velocity == [0.0] * 6
    assert f.decode(0).keys.tolist() == [[0, -16, 0]]


def test_legacy_magic_accepted(write):
    path, _ = write([cxw.WEntity([simple_sector((0, 0, 0))])], version=6, magic=cxw.LEGACY_FILE_MAGIC)
    assert cxw.CxwFile(path).world.version == 6


def expect(kind, fn):
    with pytest.raises(cxw.CxwError) as info:
        fn()
    assert info.value.kind == kind, str(info.value)
    return info.value


def test_bad_magic(write):
    path, _ = write(base(), magic=0x12345678)
    assert expect("BadMagic", lambda: cxw.CxwFile(path)).actual == 0x12345678


@pytest.mark.parametrize("flags", [0, 3])
def test_bad_flags(write, flags):
    path, _ = write(base(), flags=flags)
    assert expect("BadFlags", lambda: cxw.CxwFile(path)).actual == flags


def test_unsupported_version(write, tmp_path):
    _, data = write(base())
    data = bytearray(data)
    data[4:6] = struct.pack("<H", 99)
    assert expect("UnsupportedVersion", lambda: open_bytes(tmp_path, bytes(data))).actual == 99


def test_file_shorter_than_header(tmp_path):
    assert expect("Truncated", lambda: open_bytes(tmp_path, bytes([0x43, 0x58, 0x4C, 0x53]))).what == "header"


def test_truncated_region_payload(write, tmp_path):
    _, data = write(base())  # table first: the last region ends at the end of the file
    assert expect("Truncated", lambda: open_bytes(tmp_path, data[:-5])).what == "sector payload"


def test_truncated_region_preview(write, tmp_path):
    _, data = write(base())
    assert expect("Truncated", lambda: open_bytes(tmp_path, data[:-9000])).what == "sector region"


def test_truncated_entity_table(write, tmp_path):
    _, data = write(base())
    data = bytearray(data)
    data[8:16] = struct.pack("<Q", len(data) - 10)  # the world record runs past the end
    assert expect("Truncated", lambda: open_bytes(tmp_path, bytes(data))).what == "world record"


def test_bad_region_size(write, tmp_path):
    _, data = write(base())
    data = bytearray(data)
    entry = 64 + 29 + 4 + 84  # first sector index entry, right after the v8 table
    size_at = entry + 20
    size = struct.unpack_from("<I", data, size_at)[0]
    data[size_at:size_at + 4] = struct.pack("<I", size + 1)
    e = expect("BadRegionSize", lambda: open_bytes(tmp_path, bytes(data)))
    assert e.coord == [0, 0, 0] and e.actual == size + 1 and e.expected == size


def test_duplicate_sector(write):
    path, _ = write([cxw.WEntity([simple_sector((2, 3, 4)), simple_sector((2, 3, 4))])])
    assert expect("DuplicateSector", lambda: cxw.CxwFile(path)).coord == [2, 3, 4]
    # the same coord in two entities is fine
    path, _ = write([cxw.WEntity([simple_sector((2, 3, 4))]), cxw.WEntity([simple_sector((2, 3, 4))])])
    assert len(cxw.CxwFile(path).sectors) == 2


def test_bad_sector_magic_and_decode_all_error(write):
    bad = simple_sector((1, 1, 1))
    bad.payload_magic = 0xDEADBEEF
    sectors = [simple_sector((i, 0, 0)) for i in range(20)] + [bad]
    path, _ = write([cxw.WEntity(sectors)])
    f = cxw.CxwFile(path)
    assert expect("BadSectorMagic", lambda: f.decode(len(f.sectors) - 1)).actual == 0xDEADBEEF
    assert expect("BadSectorMagic", lambda: f.decode_all(threads=4)).actual == 0xDEADBEEF


def test_bad_counts(write):
    raw = bytearray(cxw.payload_bytes(simple_sector((0, 0, 0))))
    raw[8:12] = struct.pack("<I", 5)  # count 5 > capacity 1
    s = simple_sector((0, 0, 0))
    s.compressed_override = cxw.deflate_raw(bytes(raw))
    path, _ = write([cxw.WEntity([s])])
    f = cxw.CxwFile(path)
    e = expect("BadCounts", lambda: f.decode(0))
    assert (e.expected, e.actual, e.offset) == (1, 5, 0)  # capacity, count, free count


def test_corrupt_deflate(write):
    s = simple_sector((0, 0, 0))
    compressed = cxw.deflate_raw(cxw.payload_bytes(s))
    s.compressed_override = compressed[: len(compressed) // 2]  # stream ends early
    s2 = simple_sector((1, 0, 0))
    s2.compressed_override = bytes([0xFF] * 64)  # invalid block type
    path, _ = write([cxw.WEntity([s, s2])])
    f = cxw.CxwFile(path)
    for i in range(len(f.sectors)):
        expect("Inflate", lambda: f.decode(i))


def test_truncated_payload_slot_data(write):
    raw = cxw.payload_bytes(simple_sector((0, 0, 0)))[:-100]
    s = simple_sector((0, 0, 0))
    s.compressed_override = cxw.deflate_raw(raw)
    path, _ = write([cxw.WEntity([s])])
    f = cxw.CxwFile(path)
    assert expect("Truncated", lambda: f.decode(0)).what == "payload slot data"


def test_large_payload_grows_inflate_buffer(write):
    bricks = [(bi, brick(bi, 1 + bi % 512)) for bi in range(4096)]
    expected_voxels = sum(1 + bi % 512 for bi in range(4096))
    path, _ = write([cxw.WEntity([cxw.WSector((-1, -1, -1), bricks)])])
    f = cxw.CxwFile(path)
    d = f.decode(0)
    assert len(d.keys) == 4096 and d.voxels == expected_voxels
    assert d.keys[0].tolist() == [-16, -16, -16] and d.keys[4095].tolist() == [-1, -1, -1]
    assert np.array_equal(d.blocks[4095], brick(4095, 1 + 4095 % 512))


def test_payload_structure_errors(write):
    """The BadPayload reasons (lib.rs:558-640): slot ids, strides, brick ids, missing Block slot."""
    def with_payload(mutate):
        raw = bytearray(cxw.payload_bytes(simple_sector((0, 0, 0))))
        mutate(raw)
        s = simple_sector((0, 0, 0))
        s.compressed_override = cxw.deflate_raw(bytes(raw))
        path, _ = write([cxw.WEntity([s])])
        return cxw.CxwFile(path)
    slot_count_at = 16 + 4096 * 2 + 2 + 4096 * 2  # header, indices, empty free list, require flags
    f = with_payload(lambda r: r.__setitem__(slice(slot_count_at, slot_count_at + 4), struct.pack("<I", 17)))
    assert expect("BadPayload", lambda: f.decode(0)).what == "more than 16 slot records"
    f = with_payload(lambda r: r.__setitem__(slice(slot_count_at + 5, slot_count_at + 7), struct.pack("<H", 1)))
    assert expect("BadPayload", lambda: f.decode(0)).what in ("Block slot stride is not 2", "payload slot data")
    f = with_payload(lambda r: r.__setitem__(slice(slot_count_at + 5, slot_count_at + 7), struct.pack("<H", 0)))
    assert expect("BadPayload", lambda: f.decode(0)).what == "slot stride is 0"
    f = with_payload(lambda r: r.__setitem__(slice(16, 18), struct.pack("<h", 7)))  # id 7 >= capacity 1
    assert expect("BadPayload", lambda: f.decode(0)).what == "brick index table entry outside -1..capacity-1"
    f = with_payload(lambda r: r.__setitem__(slot_count_at + 4, 3))  # the only slot is not Block
    assert expect("BadPayload", lambda: f.decode(0)).what == "allocated bricks but no Block slot"


def test_terrain_round_trip_matches_scene_data(write):
    """A world written from a synthetic scene reads back as the same non-air bricks (the keys and
    payload the scene build takes; all-air bricks are dropped, as the format does), in any thread
    count."""
    from caelix_amd import scene

    data = scene.small_scene(3)
    sectors = cxw.sectors_from_bricks(data.keys, data.payload)
    path, _ = write([cxw.WEntity(sectors)], table_first=False)
    f = cxw.CxwFile(path)
    solid = (data.payload != 0).any(axis=1)
    want_keys = data.keys[solid]
    want_payload = data.payload[solid]
    order_want = np.lexsort(want_keys.T[::-1])
    for threads in (1, 3, 0):
        got = f.decode_all(threads=threads)
        keys, payload, grid_min, dim = cxw.scene_bricks(got)
        order_got = np.lexsort(keys.T[::-1])
        assert np.array_equal(keys[order_got], want_keys[order_want])
        assert np.array_equal(payload[order_got], want_payload[order_want])
        assert got["voxels"] == int((data.payload != 0).sum())
        assert grid_min == tuple(int(v) for v in want_keys.min(axis=0))
