"""The light tree through its updates, as the reference's light tree test drives it
(tests/light_tree_test.cpp:1300-1460 "structure, sampling and updates", :1605-1700 record copies and one
window of uploads, :1700-1722 heap growth from a tiny heap), on its own worlds regenerated from their seeds
(tests/world_fixtures.py): a static entity (BorderWorld) and a rotated, moved dynamic one (SmallWorld).

After every step the tables pass CheckStructure (tests/lt_structure.py: rows, blocks, index entries,
records and leaves against a numpy recomputation of the light faces, both trees' leaf counts), the rows'
generations change exactly for the groups rebuilt (version) and released (life), and the statistics
count the records recomputed and copied as the reference's expects. The oracle (CPU) runs every step with
its own restatement of the light tree's CPU half (oracle.LightTreeModel); on the GPU the library's
adapter (cx_light_tree_*) runs the same steps and its tables equal the oracle's bit for bit, and the
light draws and point pdfs over the rotated frames equal the oracle's, with the draw's pdf equal to the
point pdf of the drawn face and the point pdfs of every light face summing to 1 per receiver."""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

import lt_structure as lts
import world_fixtures as wf
from caelix_amd import layout, light_codes_for
from caelix_amd.world import REGION, WorldState
from ref_fixtures import Mt19937Stream

MIN_EMISSION = 20.0
ENTITIES = 2
CAPACITY = 1024
GRID = 64
ROT = wf.axis_angle_rows((0.3, 0.8, 0.5), 37.0, (300.0, 50.0, -200.0))
MOVED = wf.axis_angle_rows((-0.6, 0.1, 0.9), 37.0, (-150.0, 20.0, 80.0))


def codes(min_emission: float) -> np.ndarray:
    return light_codes_for(wf.lamp_materials(), min_emission, 1)


class Pair:
    """The same edits on one or two worlds: the oracle's (a WorldState whose hooks go to the oracle's
    light tree model) and optionally the GPU's (LightWorld)."""

    def __init__(self, oracle, gpu: bool, heap_bytes: int = 0):
        self.materials = wf.lamp_materials()
        self.min_emission = MIN_EMISSION
        self.state = WorldState(4, CAPACITY, ENTITIES, GRID)
        self.ow = oracle.OracleWorld(self.materials, codes(MIN_EMISSION), 4, 1, CAPACITY, self.state.grid_dim, ENTITIES,
                                     heap_units=max(1, heap_bytes // 16) if heap_bytes else 1024)
        self.state.hook = self.ow.model
        self.gpu = None
        if gpu:
            from caelix_amd.world import LightWorld

            self.gpu = LightWorld(self.materials, 4, 1, CAPACITY, ENTITIES, GRID, MIN_EMISSION, heap_bytes)
        self.worlds = [self.state] + ([self.gpu] if self.gpu is not None else [])

    def do(self, name, *args):
        out = [getattr(w, name)(*args) for w in self.worlds]
        return out[0]

    def set_min_emission(self, value: float):
        self.min_emission = value
        self.ow.set_light_codes(codes(value))
        if self.gpu is not None:
            self.gpu.set_min_emission(value)

    def flush(self):
        records, fresh = self.state.take_dirty()
        self.ow.flush(self.state, records, fresh)
        if self.gpu is not None:
            self.gpu.flush()

    # ---- checks ---------------------------------------------------------------------------------
    def data(self):
        live = np.array(sorted(self.state.record_of.values()), np.int64)
        return SimpleNamespace(keys=self.state.keys[live], payload=self.state.payload[live], materials=self.materials,
                               light_codes=codes(self.min_emission), group_shift=4, encoding=1), live

    def check(self) -> dict:
        """CheckStructure on the oracle (and the GPU, with its tables equal to the oracle's)."""
        data, live = self.data()
        coords = self.state.group_coords()
        motion = self.state.entity_motion
        want = lts.from_oracle(self.ow.lt)
        totals = lts.check_structure(data, self.state.group_of_brick[live], coords, want, coords[:, 0], motion, live)
        if self.gpu is not None:
            got = lts.from_gpu(self.gpu.tables_read())
            lts.compare_tables(got, want)
            st = self.gpu.stats()
            assert st["light_bricks"] == totals["bricks"] and st["light_faces"] == totals["faces"]
            assert st["tree_leaves"] == totals["trees"]
        return totals

    def generations(self) -> np.ndarray:
        return self.ow.lt.rows["generation"].copy()

    def control(self):
        return self.ow.lt.control[0]


def changed(before, after, versions, lives):
    """The rows' generations (life << 16 | version) changed exactly in these groups."""
    n = max(len(before), len(after))
    b = np.zeros(n, np.int64)
    a = np.zeros(n, np.int64)
    b[:len(before)] = before
    a[:len(after)] = after
    for g in range(n):
        if g in lives:
            assert a[g] >> 16 == (b[g] >> 16) + 1 and a[g] & 0xFFFF == 0, g
        elif g in versions:
            assert a[g] >> 16 == b[g] >> 16 and (a[g] & 0xFFFF) == ((b[g] + 1) & 0xFFFF), g
        else:
            assert a[g] == b[g], g


def upload_world(pair, entity, world, keys=None):
    keys = list(world) if keys is None else keys
    pair.do("upload", entity, keys, [world[k] for k in keys])


def initial(pair):
    rng = Mt19937Stream(42)
    w0 = wf.border_world(rng)
    w1 = wf.small_world(rng)
    e0 = pair.do("add_entity", 0, None)
    e1 = pair.do("add_entity", 1, ROT)
    upload_world(pair, e0, w0)
    upload_world(pair, e1, w1)
    pair.flush()
    return rng, (e0, w0), (e1, w1)


def group_of(pair, entity, key):
    return pair.state.groups[(entity, key[0] >> 4, key[1] >> 4, key[2] >> 4)]


def run_updates(pair):
    """The reference's "updates" section; returns the per-step totals."""
    rng, (e0, w0), (e1, w1) = initial(pair)
    steps = [pair.check()]
    assert steps[0]["trees"][0] > 0 and steps[0]["trees"][1] > 0
    # light bricks added inside the cluster group: only that group is rebuilt
    key = (21, 5 + wf.Y_SHIFT, 37)
    cluster = group_of(pair, e0, key)
    before = pair.generations()
    w0[key] = wf.lamp_brick(rng, 20)
    upload_world(pair, e0, w0, [key])
    pair.flush()
    changed(before, pair.generations(), {cluster}, set())
    steps.append(pair.check())
    # light bricks across a group border
    keys = [(15, 2 + wf.Y_SHIFT, 31), (16, 2 + wf.Y_SHIFT, 31), (15, 2 + wf.Y_SHIFT, 32)]
    for k in keys:
        w0[k] = wf.lamp_brick(rng, 25)
    upload_world(pair, e0, w0, keys)
    pair.flush()
    steps.append(pair.check())
    # min_emission up (the warm lamps drop out), then down (the dim lamps join)
    pair.set_min_emission(24.0)
    pair.flush()
    steps.append(pair.check())
    pair.set_min_emission(10.0)
    pair.flush()
    steps.append(pair.check())
    assert steps[-1]["faces"] > steps[-2]["faces"] > 0
    # entity 1 moved and rotated: no group is rebuilt
    before = pair.generations()
    pair.do("set_transform", e1, MOVED)
    pair.flush()
    changed(before, pair.generations(), set(), set())
    steps.append(pair.check())
    # light bricks removed (inside the cluster and at the border)
    keys = [(22, 6 + wf.Y_SHIFT, 38), (16, 2 + wf.Y_SHIFT, 31)]
    for k in keys:
        w0.pop(k)
    pair.do("remove", e0, keys)
    pair.flush()
    steps.append(pair.check())
    # entity 1 removed: its groups are released
    before = pair.generations()
    released = {g for g, k in enumerate(pair.state.group_key) if k is not None and k[0] == e1}
    pair.do("remove_entity", e1)
    pair.flush()
    changed(before, pair.generations(), set(), released)
    steps.append(pair.check())
    assert steps[-1]["trees"][1] == 0
    # an edit recomputes only the changed brick records
    key = (21, 5 + wf.Y_SHIFT, 37)
    cluster = group_of(pair, e0, key)
    block_before = records_by_key(pair, cluster)
    w0[key] = wf.lamp_brick(rng, 30)
    upload_world(pair, e0, w0, [key])
    pair.flush()
    after = records_by_key(pair, cluster)
    touched = {key} | {(key[0] + d[0], key[1] + d[1], key[2] + d[2])
                       for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))}
    recomputed = copied = 0
    for k, words in after.items():
        if k in touched:
            recomputed += 1
        else:
            copied += 1
            assert np.array_equal(block_before[k], words), k
    c = pair.control()
    assert int(c["records_recomputed"]) == recomputed and int(c["records_copied"]) == copied
    assert copied > 40
    if pair.gpu is not None:
        st = pair.gpu.stats()
        assert st["records_recomputed"] == recomputed and st["records_copied"] == copied
    steps.append(pair.check())
    return steps


def records_by_key(pair, group) -> dict:
    """{brick key (entity-local): its record's words} of a group's block in the oracle's current heap."""
    block = pair.ow.lt.blocks()[group]
    key_of = {r: k[1:] for k, r in pair.state.record_of.items()}
    out = {}
    for r in range(int(block[0])):
        entry = block[4 + 4 * r:8 + 4 * r]
        rec = block[int(entry[2]) * 4:]
        units = lts.record_units(int(rec[9]))
        out[key_of[int(entry[1])]] = rec[:units * 4].copy()
    return out


def test_updates_oracle(oracle):
    pair = Pair(oracle, gpu=False)
    steps = run_updates(pair)
    assert len(steps) == 9


def test_isolated_brick_and_one_window_oracle(oracle):
    run_isolated(Pair(oracle, gpu=False))


def run_isolated(pair):
    """light_tree_test.cpp:1605-1700: one group with an isolated light brick (its only neighbour stone)
    and two other light bricks; the isolated brick turns dark (only its own old word 1 marks the group);
    a light brick with a new layout (1 recomputed, 1 copied); one window of uploads and a removal."""
    rng = Mt19937Stream(61)
    e = pair.do("add_entity", 0, None)
    stone = np.full(512, wf.STONE, np.uint16)
    isolated, lamp, other = (2, 2, 2), (7, 7, 7), (11, 3, 9)
    world = {isolated: wf.lamp_brick(rng, 20), (3, 2, 2): stone, lamp: wf.lamp_brick(rng, 20),
             other: wf.lamp_brick(rng, 20)}
    upload_world(pair, e, world)
    pair.flush()
    pair.check()
    group = group_of(pair, e, isolated)
    dark = np.where(np.arange(512) % 3 == 0, wf.STONE, 0).astype(np.uint16)
    world[isolated] = dark
    before = pair.generations()
    upload_world(pair, e, world, [isolated])
    pair.flush()
    changed(before, pair.generations(), {group}, set())
    assert pair.check()["bricks"] == 2
    world[lamp] = wf.lamp_brick(rng, 35)
    upload_world(pair, e, world, [lamp])
    pair.flush()
    c = pair.control()
    assert int(c["records_recomputed"]) == 1 and int(c["records_copied"]) == 1
    pair.check()
    # one window: a new brick uploaded twice, a resident brick replaced and then removed while a new
    # brick takes its record, a resident brick replaced twice
    fresh, reused = (5, 9, 1), (12, 12, 12)
    world[fresh] = wf.lamp_brick(rng, 20)
    upload_world(pair, e, world, [fresh])
    world[fresh] = wf.lamp_brick(rng, 40)
    upload_world(pair, e, world, [fresh])
    world[isolated] = wf.lamp_brick(rng, 25)
    upload_world(pair, e, world, [isolated])
    pair.do("remove", e, [isolated])
    world.pop(isolated)
    world[reused] = wf.lamp_brick(rng, 30)
    upload_world(pair, e, world, [reused])
    world[lamp] = wf.lamp_brick(rng, 15)
    upload_world(pair, e, world, [lamp])
    world[lamp] = wf.lamp_brick(rng, 45)
    upload_world(pair, e, world, [lamp])
    pair.flush()
    pair.check()


# ------------------------------------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_updates_gpu_equal_oracle(oracle):
    pair = Pair(oracle, gpu=True)
    run_updates(pair)


@pytest.mark.gpu
def test_isolated_brick_and_one_window_gpu(oracle):
    run_isolated(Pair(oracle, gpu=True))


@pytest.mark.gpu
def test_heap_growth_from_a_tiny_heap(oracle):
    """light_tree_test.cpp:1700-1722: a 4096-byte heap: the first passes fail, the heap grows, a pass
    commits; the tables then equal the oracle's."""
    pair = Pair(oracle, gpu=True, heap_bytes=4096)
    rng = Mt19937Stream(7)
    e = pair.do("add_entity", 0, None)
    upload_world(pair, e, wf.border_world(rng))
    pair.flush()
    st = pair.gpu.stats()
    assert st["failed_passes"] >= 1 and st["passes"] > st["failed_passes"]
    assert st["heap_bytes_capacity"] > 4096 and st["heap_bytes_used"] <= st["heap_bytes_capacity"]
    pair.check()


def light_face_queries(pair, receiver) -> np.ndarray:
    """A point pdf query (x, n, group, group-local voxel, face, block) for every light face."""
    data, live = pair.data()
    faces = lts.open_faces_sparse(data.keys, data.payload, data.light_codes)
    rows = []
    for b, open_ in faces.items():
        r = int(live[b])
        g = int(pair.state.group_of_brick[r])
        key = pair.state.keys[r].astype(np.int64)
        v, f = np.nonzero(open_)
        base = (key & 15) * 8
        vox = np.stack([base[0] + (v & 7), base[1] + ((v >> 3) & 7), base[2] + (v >> 6)], axis=1)
        q = np.zeros((len(v), 12), np.uint32)
        q[:, 0:6] = np.asarray(receiver, np.float32).view(np.uint32)
        q[:, 6] = g
        q[:, 7:10] = vox
        q[:, 10] = f
        q[:, 11] = data.payload[b, v]
        rows.append(q)
    return np.concatenate(rows)


@pytest.mark.gpu
def test_rotated_entities_sampling_gpu(cx, oracle):
    """Draws and point pdfs over both trees and the rotated frames: GPU = oracle bit for bit; the draw's
    pdf = the point pdf of the drawn face (1e-4); sum of the point pdfs of every light face = 1 (1e-4)
    per receiver; a drawn point lies on its face, moved to the world by its entity's frame."""
    pair = Pair(oracle, gpu=True)
    initial(pair)
    view = pair.gpu.view
    ov = pair.ow.lt.view(True)
    rng = np.random.default_rng(5)
    centres = [np.array([100.0, 200.0, 280.0]), ROT[:, 3] + ROT[:, :3] @ np.array([140.0, 135.0, 10.0])]
    recv = []
    for c in centres:
        p = c + rng.uniform(-150, 150, size=(64, 3))
        n = rng.normal(size=(64, 3))
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        n[::2] = 0.0
        recv.append(np.concatenate([p, n], axis=1))
    recv = np.concatenate(recv).astype(np.float32)
    rep = np.repeat(recv, 32, axis=0)
    seeds = rng.integers(0, 2 ** 32, size=len(rep), dtype=np.uint64).astype(np.uint32)
    got, ok = cx.light_sample(view, rep, seeds)
    want, ok_want = oracle.light_sample(ov, rep, seeds)
    assert np.array_equal(ok, ok_want.astype(ok.dtype)) and ok.mean() > 0.5
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    drawn = ok.astype(bool)
    w = got[drawn].view(np.uint32)
    q = np.zeros((int(drawn.sum()), 12), np.uint32)
    q[:, 0:6] = rep[drawn].view(np.uint32)
    q[:, 6] = w[:, 9]
    q[:, 7] = w[:, 11] & 0x3FF
    q[:, 8] = (w[:, 11] >> 10) & 0x3FF
    q[:, 9] = w[:, 11] >> 20
    q[:, 10] = w[:, 8]
    q[:, 11] = w[:, 7]
    pdf = cx.light_point_pdf(view, q)
    assert np.array_equal(pdf.view(np.uint32), oracle.light_point_pdf(ov, q).view(np.uint32))
    p_draw = got[drawn, 6].astype(np.float64)
    assert np.all(np.abs(pdf / p_draw - 1.0) < 1e-4)
    # the drawn point on its face: the group-local face rectangle moved by the entity's frame
    coords = pair.state.group_coords()
    ent = coords[q[:, 6], 0]
    origin = coords[q[:, 6], 1:].astype(np.float64) * 128.0
    local = origin + q[:, 7:10].astype(np.float64)
    face = q[:, 10]
    axis = face >> 1
    rows = pair.state.entity_rows[ent].astype(np.float64)
    world_pts = got[drawn, 0:3].astype(np.float64)
    back = np.einsum("nji,nj->ni", rows[:, :, :3], world_pts - rows[:, :, 3])  # R^T (p - t)
    rel = back - local
    plane = np.where(face & 1, 0.0, 1.0)
    tol = 2e-3 + 2e-5 * np.abs(world_pts).max(axis=1)
    assert np.all(np.abs(rel[np.arange(len(rel)), axis] - plane) < tol)
    assert np.all((rel > -tol[:, None]) & (rel < 1 + tol[:, None]))
    for r in recv[::16]:
        qa = light_face_queries(pair, r)
        total = cx.light_point_pdf(view, qa).astype(np.float64).sum()
        assert abs(total - 1.0) < 1e-4, total
