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


# ---- the reference finalize test's other scenarios (finalize_test.cpp:319-451), packed encoding ----

import ref_fixtures as rf  # noqa: E402


def random_bricks_world():
    """Scenario 1: RandomBlobs(300) and a RandomBrick per key from mt19937(20261014)."""
    rng = rf.Mt19937Stream(20261014)
    keys = rf.random_blobs(rng, 300)
    return {k: rf.random_brick(rng) for k in keys}, rng


def check_rules(o, data, encoding=0):
    payload, _ = tc.finalize_payload(data, encoding)
    assert np.array_equal(o.a["payload"].reshape(-1, 512), payload)
    words, occ, brick_bits, super_bits = tc.finalize_expected(data, encoding)
    info = o.a["info"].reshape(-1, 2)
    assert np.array_equal(info[:, 0], words[:, 0])
    assert np.array_equal(info[:, 1] & 0x3FFFF, words[:, 1])
    assert np.array_equal(o.a["occupancy"].reshape(-1, 8), occ)
    assert np.array_equal(o.a["brick_bits"][: len(brick_bits)], brick_bits)


def scenario_worlds():
    """(scenario 1 world, scenario 5 world: 20 bricks removed and 20 others replaced with new draws)."""
    world, rng = random_bricks_world()
    keys = sorted(world)
    order = np.random.default_rng(5).permutation(len(keys))
    removed = [keys[i] for i in order[:20]]
    replaced = [keys[i] for i in order[20:40]]
    updated = dict(world)
    for k in removed:
        del updated[k]
    for k in replaced:
        updated[k] = rf.random_brick(rng)
    return world, updated


def test_random_bricks_removal_and_replacement(oracle):
    world, updated = scenario_worlds()
    for w in (world, updated):
        data = rf.scene_from_world(w, rf.look_plus_z(0, 0, -100, 8), group_shift=3, lights=False)
        o = build(oracle, data)
        check_rules(o, data)
        # finalizing the finalized payload again changes nothing (the lossless rule)
        again = rf.scene_from_world(w, rf.look_plus_z(0, 0, -100, 8), group_shift=3, lights=False)
        again.payload = o.a["payload"].reshape(-1, 512).copy()
        o2 = build(oracle, again)
        for name in ("payload", "info", "occupancy", "brick_bits", "super_bits"):
            assert np.array_equal(o.a[name], o2.a[name]), name
    assert len(updated) == 280


def solid_cube_checks(info, occ_words, keys):
    index = {tuple(k): i for i, k in enumerate(np.asarray(keys).tolist())}
    centre = info[index[(16, 0, 32)]]
    assert centre[0] >> 16 == 0 and centre[1] == 0  # the interior brick: no voxel alive
    face = info[index[(15, 0, 32)]]
    assert face[1] & 0x3FFFF == (0 | (0 << 3) | (0 << 6) | (0 << 9) | (7 << 12) | (7 << 15))  # only x = 0 alive
    b = index[(15, 0, 32)]
    assert occupied_bit(occ_words, b, 0, 3, 3) and not occupied_bit(occ_words, b, 1, 3, 3)


def test_solid_cube_across_group_boundaries(oracle):
    world = rf.solid_cube_world(7)
    assert len({(x >> 4, y >> 4, z >> 4) for x, y, z in world}) == 8
    data = rf.scene_from_world(world, rf.look_plus_z(0, 0, -100, 8), group_shift=4, lights=False)
    o = build(oracle, data)
    check_rules(o, data)
    solid_cube_checks(o.a["info"].reshape(-1, 2), o.a["occupancy"], data.keys)


def neighbour_worlds():
    solid = np.full(512, 0x8001, np.uint16)
    return {(15, 0, 0): solid.copy()}, {(15, 0, 0): solid.copy(), (16, 0, 0): solid.copy()}


def neighbour_checks(alone_info, alone_occ, both_info, both_occ, both_keys):
    assert occupied_bit(alone_occ, 0, 7, 3, 3) and not occupied_bit(alone_occ, 0, 3, 3, 3)
    first = [tuple(k) for k in np.asarray(both_keys).tolist()].index((15, 0, 0))
    assert not occupied_bit(both_occ, first, 7, 3, 3)
    assert not np.array_equal(both_occ.reshape(-1, 8)[first], alone_occ.reshape(-1, 8)[0])
    assert (int(both_info[first, 1]) >> 9) & 7 == 7  # max x stays 7 (the edges of the x = 7 face)


def test_neighbour_across_a_group_boundary(oracle):
    alone, both = neighbour_worlds()
    cam = rf.look_plus_z(0, 0, -100, 8)
    d1 = rf.scene_from_world(alone, cam, group_shift=4, lights=False)
    d2 = rf.scene_from_world(both, cam, group_shift=4, lights=False)
    o1, o2 = build(oracle, d1), build(oracle, d2)
    check_rules(o2, d2)
    neighbour_checks(o1.a["info"].reshape(-1, 2), o1.a["occupancy"], o2.a["info"].reshape(-1, 2), o2.a["occupancy"],
                     d2.keys)


@pytest.mark.gpu
def test_reference_finalize_scenarios_gpu(cx, oracle):
    """Scenarios 1, 5 (removal, replacement), 2 (solid cube) and 6 (a neighbour in the next group) on the
    GPU: every record word, occupancy word, payload value and bitmap equals the oracle's."""
    cam = rf.look_plus_z(0, 0, -100, 8)
    world, updated = scenario_worlds()
    alone, both = neighbour_worlds()
    cases = [(world, 3), (updated, 3), (rf.solid_cube_world(7), 4), (alone, 4), (both, 4)]
    got = []
    for w, shift in cases:
        data = rf.scene_from_world(w, cam, group_shift=shift, lights=False)
        g = cx.TraceScene(data, load_factor=LOAD).host_arrays()
        o = build(oracle, data)
        for name in ("payload", "info", "occupancy", "brick_bits", "super_bits"):
            assert np.array_equal(g[name], o.a[name]), name
        got.append((g, data))
    g, data = got[2]
    solid_cube_checks(g["info"].reshape(-1, 2), g["occupancy"], data.keys)
    (g1, _), (g2, d2) = got[3], got[4]
    neighbour_checks(g1["info"].reshape(-1, 2), g1["occupancy"], g2["info"].reshape(-1, 2), g2["occupancy"], d2.keys)
