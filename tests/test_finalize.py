"""The oracle's finalize against the reference finalize test's rules and scenario, without the GPU.

tests/trace_checks.finalize_payload restates Reference() of the reference's finalize test
(tests/finalize_test.cpp:74-157) with the kernel's interface term; the glass scenario
(finalize_test.cpp:264-335, GlassScenario) is restated with its literal checks, in both block
encodings: the air above the slab carries the -Y face bit and is occupied, the glass on the floor has
no face towards the opaque floor and one towards the air at -X, glass inside the slab has no faces,
air far from the glass stays 0 and unoccupied. Opaque terrain: a buried voxel (six opaque neighbours,
across brick boundaries too) is not occupied."""
from __future__ import annotations

import numpy as np
import pytest

import trace_checks as tc
from caelix_amd import scene

LOAD = 0.7


def build(oracle, data):
    return oracle.OracleScene.build(data, int(np.ceil(len(data.keys) / LOAD)))


def occupied_bit(occ_words, brick, x, y, z) -> bool:
    coarse = (x >> 2) | ((y >> 2) << 1) | ((z >> 2) << 2)
    micro = (x & 3) | ((y & 3) << 2) | ((z & 3) << 4)
    return bool((int(occ_words[brick * 8 + coarse]) >> micro) & 1)


@pytest.mark.parametrize("encoding", [0, 1])
def test_glass_scenario(oracle, encoding):
    data = scene.glass_world(encoding)
    ref = build(oracle, data)
    opaque_mask, id_mask, face_shift = tc.ENCODINGS[encoding]
    payload = ref.a["payload"].reshape(-1, 512)
    occ = ref.a["occupancy"]
    keys = [tuple(k) for k in np.asarray(data.keys).tolist()]
    b010, b110 = keys.index((0, 1, 0)), keys.index((1, 1, 0))

    def value(b, x, y, z):
        return int(payload[b, x + 8 * y + 64 * z])

    air = value(b010, 4, 4, 3)
    assert air & id_mask == 0 and (air >> face_shift) & 0x3F == 1 << 3
    assert occupied_bit(occ, b010, 4, 4, 3)
    bottom = value(b010, 2, 0, 3)
    assert bottom & id_mask == 0x42
    assert (bottom >> face_shift) & (1 << 3) == 0 and (bottom >> face_shift) & (1 << 1) != 0
    assert value(b110, 3, 1, 3) == 0x42
    assert value(b010, 0, 7, 7) == 0 and not occupied_bit(occ, b010, 0, 7, 7)
    # the kernel's interface term: glass inside the slab (no face) is not occupied
    assert not occupied_bit(occ, b110, 3, 1, 3)
    # every word as the restated rules give it
    want_payload, _ = tc.finalize_payload(data, encoding)
    assert np.array_equal(payload, want_payload)
    words, occ_want, brick_bits, super_bits = tc.finalize_expected(data, encoding)
    info = ref.a["info"].reshape(-1, 2)
    assert np.array_equal(info[:, 0], words[:, 0])
    assert np.array_equal(info[:, 1] & 0x3FFFF, words[:, 1])
    assert np.array_equal(occ.reshape(-1, 8), occ_want)
    assert np.array_equal(ref.a["brick_bits"][: len(brick_bits)], brick_bits)


def test_finalize_is_idempotent(oracle):
    """Face bits already in the payload are masked off: finalizing again changes nothing."""
    data = scene.glass_world(1)
    once = build(oracle, data)
    again = scene.glass_world(1)
    again.payload = once.a["payload"].reshape(-1, 512).copy()
    twice = build(oracle, again)
    for name in ("payload", "info", "occupancy", "brick_bits", "super_bits"):
        assert np.array_equal(once.a[name], twice.a[name]), name


def test_buried_voxels_not_occupied(oracle):
    """A solid 2 x 1 x 1 block of bricks: only its surface voxels are occupied; the voxels next to the
    shared brick face are buried through the neighbour brick."""
    keys = np.array([[0, 0, 0], [1, 0, 0]], np.int32)
    payload = np.full((2, 512), 0x8001, np.uint16)
    data = scene.SceneData(keys=keys, payload=payload, materials=np.zeros((65536, 8), np.float32),
                           light_codes=np.zeros(2048, np.uint32), grid_min=(0, 0, 0), grid_dim=8, group_shift=4,
                           camera=scene.look_at((-5, 4, 4), (4, 4, 4), 8, 8), emissive_bricks=np.zeros(0, np.int64))
    ref = build(oracle, data)
    occ = ref.a["occupancy"]
    count = sum(bin(int(w)).count("1") for w in occ)
    interior = np.zeros((16, 8, 8), bool)  # [x, y, z] over both bricks
    interior[1:-1, 1:-1, 1:-1] = True  # six solid neighbours: buried
    assert count == int((~interior).sum()) == 1024 - 14 * 6 * 6
    for b in range(2):
        for x in range(8):
            gx = b * 8 + x
            assert occupied_bit(occ, b, x, 3, 3) == (gx in (0, 15))
    # tight bounds of brick 0 still span the whole brick (its surface voxels reach every side)
    assert ref.a["info"].reshape(-1, 2)[0, 1] & 0x3FFFF == (0 | (0 << 3) | (0 << 6) | (7 << 9) | (7 << 12) | (7 << 15))


def test_terrain_finalize_matches_rules(oracle):
    data = scene.small_scene()
    ref = build(oracle, data)
    payload, occupied = tc.finalize_payload(data)
    assert np.array_equal(ref.a["payload"].reshape(-1, 512), payload)
    solid = np.asarray(data.payload) != 0
    assert (solid & ~occupied).sum() > 100000  # most terrain voxels are buried
    assert not (occupied & ~solid).any()  # opaque terrain: air never becomes an interface
