"""The region (box) receiver of the light tree, as the reference's light NEE test checks it
(tests/light_nee_test.cpp:440-550 "the box sampler supports every light face that can reach the box"), on
lamps of a static entity and of a rotated dynamic one (the light tree test's BorderWorld and SmallWorld,
tests/world_fixtures.py), on the oracle (CPU) and on the HIP path through the C ABI (gpu):

support            for receiver boxes (regions of 66 or 258 voxels around receivers near the lamps, the
                   same boxes shrunk to a point, boxes of 6 voxels centred on lamp faces and facing along
                   them), every light face that can send light to a receiver point of the box (it faces the
                   point and a corner of it lies in front of the point's surface; the box's 8 corners,
                   centre and 24 random points) is found with a leaf, and has a positive box pdf
                   (LightTreeBrickPdfBox) and target (LightTreeLeafImportanceBox)
normalisation      the box pdfs of every brick with a leaf sum to 1 per box (each choice of the walk is a
                   distribution over its slots; a dropped term or wrong slot breaks the sum)
draw and pdf       a brick drawn for a box (LightTreeSampleBrickBox) has the box pdf of that brick (1e-5
                   relative: the draw multiplies down, the pdf up) and the drawn importance is the brick's
                   target bit for bit when the draw did not fall back to the power; the draws' frequencies
                   over 2^17 draws match the pdfs (within 5 standard deviations per brick)
concentric boxes   a receiver box centred on a leaf's summary box sees every face (all emitter cosines 1, the
                   receiver term 1): its target is the closed form sum(face) / max(1, (|h_B| + |h_R|)^2)
                   (1e-5 relative), which fixes the box form's extent term
point form         a zero-size box gives the point pdf (LightTreeBrickPdf) within 1e-5
selection          LightTreeSelect with one tree: the other tree's bricks get pdf 0, the selected tree's
                   pdfs sum to 1, every draw lands in the selected tree
stale references   a brick reference of another generation or past the record count has pdf 0
GPU = oracle       draws, box and point pdfs and support results bit for bit
"""
from __future__ import annotations

import numpy as np
import pytest

import lt_structure as lts
from test_light_world import Pair, initial

EDGE_SMALL, EDGE_LARGE = 32.0, 128.0


class LampWorld:
    """The two-entity world after its first flush, its light faces in world space and its bricks with a
    leaf."""

    def __init__(self, oracle, gpu: bool):
        self.pair = Pair(oracle, gpu=gpu)
        initial(self.pair)
        st = self.pair.state
        data, live = self.pair.data()
        coords = st.group_coords()
        faces = lts.open_faces_sparse(data.keys, data.payload, data.light_codes)
        rows = []
        for b, open_ in sorted(faces.items()):
            r = int(live[b])
            g = int(st.group_of_brick[r])
            base = (st.keys[r].astype(np.int64) & 15) * 8
            v, f = np.nonzero(open_)
            vox = np.stack([base[0] + (v & 7), base[1] + ((v >> 3) & 7), base[2] + (v >> 6)], axis=1)
            rows.append(np.concatenate([np.full((len(v), 1), g), vox, f[:, None], data.payload[b, v][:, None]], axis=1))
        table = np.concatenate(rows).astype(np.int64)
        self.group, self.vig, self.face, self.block = table[:, 0], table[:, 1:4], table[:, 4], table[:, 5]
        ent = coords[self.group, 0]
        local = coords[self.group, 1:].astype(np.float64) * 128.0 + self.vig
        axis = self.face >> 1
        d = np.zeros((len(table), 3))
        d[np.arange(len(d)), axis] = np.where(self.face & 1, -1.0, 1.0)
        centre = local + 0.5 + 0.5 * d
        frames = st.entity_rows[ent].astype(np.float64)
        rot, t = frames[:, :, :3], frames[:, :, 3]
        self.entity = ent
        self.centre = np.einsum("nij,nj->ni", rot, centre) + t
        self.normal = np.einsum("nij,nj->ni", rot, d)
        corners = []
        for sb in (-0.5, 0.5):
            for sc in (-0.5, 0.5):
                c = centre.copy()
                c[np.arange(len(c)), (axis + 1) % 3] += sb
                c[np.arange(len(c)), (axis + 2) % 3] += sc
                corners.append(np.einsum("nij,nj->ni", rot, c) + t)
        self.corners = np.stack(corners, axis=1)  # [faces, 4, 3]
        self.leaves = []  # (tree, group, entry) of every leaf
        tables = lts.from_oracle(self.pair.ow.lt)
        for k in range(2):
            count = int(tables["headers"][k]["leaf_count"])
            lv = tables["leaves"][k][:count]
            self.leaves += [(k, int(a), int(b)) for a, b in zip(lv["group"], lv["entry"])]
        self.generation = self.pair.ow.lt.rows["generation"].copy()

    def face_of_brick(self) -> dict:
        """(group, index entry) -> a light face row of the brick, through the group blocks' index entries
        (pos = x | y << 4 | z << 8 of the brick in its group)."""
        if getattr(self, "_first", None) is None:
            blocks = lts.from_oracle(self.pair.ow.lt)["blocks"]
            self._first = {}
            for i in range(len(self.group)):
                g = int(self.group[i])
                b = self.vig[i] >> 3
                pos = int(b[0]) | (int(b[1]) << 4) | (int(b[2]) << 8)
                block = blocks[g]
                count = int(block[0])
                index = block[4:4 + 4 * count].reshape(count, 4)
                entry = int(np.nonzero(index[:, 0] == pos)[0][0])
                self._first.setdefault((g, entry), i)
        return self._first

    def oracle_view(self, select: int = 0):
        return self.pair.ow.lt.view(True, select)

    def boxes(self, rng, regions: int, face_stride: int) -> np.ndarray:
        """[boxes, 9] float64: lo, hi, n. Regions around receivers near the lamps of both entities (every
        fourth with the larger edge), half of them again with zero size, boxes on lamp faces."""
        out = []
        for i in range(regions):
            pick = rng.choice(np.nonzero(self.entity == (i & 1))[0])
            x = self.centre[pick] + rng.uniform(-40.0, 40.0, size=3)
            n = rng.normal(size=3)
            n /= np.linalg.norm(n)
            half = (EDGE_LARGE if i % 4 == 0 else EDGE_SMALL) + 1.0
            out.append(np.concatenate([x - half, x + half, n]))
        for i in range(0, regions, 2):
            c = 0.5 * (out[i][0:3] + out[i][3:6])
            out.append(np.concatenate([c, c, out[i][6:9]]))
        for i in range(0, len(self.centre), face_stride):
            out.append(np.concatenate([self.centre[i] - 3.0, self.centre[i] + 3.0, self.normal[i]]))
        return np.array(out)

    def reaching(self, box, rng) -> np.ndarray:
        """The faces that can send light to a receiver point of the box (Contributes,
        light_nee_test.cpp:251-266, at the 8 corners, the centre and 24 random points)."""
        lo, hi, n = box[0:3], box[3:6], box[6:9]
        pts = [np.array([hi[0] if k & 1 else lo[0], hi[1] if k & 2 else lo[1], hi[2] if k & 4 else lo[2]])
               for k in range(8)]
        pts.append(0.5 * (lo + hi))
        pts += [lo + rng.uniform(size=3) * (hi - lo) for _ in range(24)]
        p = np.array(pts)  # [33, 3]
        facing = np.einsum("fa,pfa->pf", self.normal, p[:, None, :] - self.centre[None]) > 1e-9
        front = (np.einsum("a,fca->fc", n, self.corners)[None] - (p @ n)[:, None, None] > 1e-9).any(axis=2)
        return np.nonzero((facing & front).any(axis=0))[0]

    def support_queries(self, boxes, rng) -> tuple[np.ndarray, np.ndarray]:
        qs, owner = [], []
        for bi, box in enumerate(boxes):
            faces = self.reaching(box, rng)
            q = np.zeros((len(faces), 16), np.uint32)
            f32 = box.astype(np.float32)
            front = (self.centre[faces] + 0.5 * self.normal[faces]).astype(np.float32)
            q[:, 0:3] = f32[0:3].view(np.uint32)
            q[:, 4:7] = f32[3:6].view(np.uint32)
            q[:, 8:11] = f32[6:9].view(np.uint32)
            q[:, 3] = front[:, 0].view(np.uint32)
            q[:, 7] = front[:, 1].view(np.uint32)
            q[:, 11] = front[:, 2].view(np.uint32)
            q[:, 12] = self.group[faces]
            q[:, 13] = self.vig[faces, 0] | (self.vig[faces, 1] << 10) | (self.vig[faces, 2] << 20)
            q[:, 14] = self.face[faces]
            q[:, 15] = self.block[faces]
            qs.append(q)
            owner.append(np.full(len(faces), bi))
        return np.concatenate(qs), np.concatenate(owner)

    def brick_queries(self, box) -> np.ndarray:
        """A brick pdf query (lo, hi, n, group, entry, generation) for every brick with a leaf."""
        q = np.zeros((len(self.leaves), 12), np.uint32)
        q[:, 0:9] = np.asarray(box, np.float32).view(np.uint32)
        lv = np.array(self.leaves, np.int64)
        q[:, 9] = lv[:, 1]
        q[:, 10] = lv[:, 2]
        q[:, 11] = self.generation[lv[:, 1]]
        return q


def check_support(res: np.ndarray) -> None:
    assert len(res) > 1000
    assert np.all(res[:, 0] == 1.0) and np.all(res[:, 1] == 1.0)
    assert np.all(res[:, 2] > 0.0) and np.all(res[:, 3] > 0.0)


def check_normalised(world, box_pdf, boxes) -> None:
    for box in boxes:
        total = box_pdf(world.brick_queries(box)).astype(np.float64).sum()
        assert abs(total - 1.0) < 1e-5, total


def draw_boxes(boxes, draws, seed):
    rep = np.repeat(boxes.astype(np.float32), draws, axis=0)
    seeds = np.random.default_rng(seed).integers(0, 2 ** 32, size=len(rep), dtype=np.uint64).astype(np.uint32)
    return rep, seeds


def check_draws(world, out, rep, box_pdf, box_target) -> None:
    """Every draw succeeds and names a brick with a leaf of its own; its pdf is the box pdf of the brick;
    its importance, when positive, is the brick's target."""
    assert np.all(out[:, 0] == 1)
    q = np.zeros((len(out), 12), np.uint32)
    q[:, 0:9] = rep.view(np.uint32)
    q[:, 9:12] = out[:, 1:4]
    pdf = out[:, 5].view(np.float32).astype(np.float64)
    want = box_pdf(q).astype(np.float64)
    assert np.all(want > 0) and np.all(np.abs(pdf / want - 1.0) < 1e-5)
    leaf_words = {(k << 31) | i for k in range(2) for i in range(sum(1 for l in world.leaves if l[0] == k))}
    assert set(out[:, 4].tolist()) <= leaf_words
    imp = out[:, 6].view(np.float32)
    positive = imp > 0
    assert positive.mean() > 0.9
    tgt = box_target(q[positive])
    assert np.array_equal(imp[positive].view(np.uint32), tgt.view(np.uint32))


def check_frequencies(world, out, box, box_pdf) -> None:
    n = len(out)
    pdf = box_pdf(world.brick_queries(box)).astype(np.float64)
    index = {(g, e): i for i, (_, g, e) in enumerate(world.leaves)}
    counts = np.zeros(len(world.leaves))
    for g, e in zip(out[:, 1].tolist(), out[:, 2].tolist()):
        counts[index[(g, e)]] += 1
    sigma = np.sqrt(n * pdf * (1 - pdf)) + 1.0
    assert np.all(np.abs(counts - n * pdf) < 5 * sigma)


def check_selection(world, box_pdf_for, draw_for, box) -> None:
    trees = np.array([k for k, _, _ in world.leaves])
    for select, tree in ((1, 0), (2, 1)):
        pdf = box_pdf_for(select)(world.brick_queries(box)).astype(np.float64)
        assert np.all(pdf[trees != tree] == 0.0)
        assert abs(pdf[trees == tree].sum() - 1.0) < 1e-5
        rep, seeds = draw_boxes(box[None], 512, 40 + select)
        out = draw_for(select)(rep, seeds)
        assert np.all(out[:, 0] == 1) and np.all(out[:, 4] >> 31 == tree)


def check_concentric(world, support, view, rng) -> None:
    """Targets of receiver boxes centred on each brick's leaf summary against the closed form."""
    tables = lts.from_oracle(world.pair.ow.lt)
    rows, want = [], []
    picks = [(k, i) for k in range(2) for i in range(int(tables["headers"][k]["leaf_count"]))][::3]
    for k, i in picks:
        g, e = (int(v) for v in tables["leaves"][k][i][["group", "entry"]].tolist())
        sm = tables["leaves"][k][i]["summary"]
        lo, hi = sm["lo"].astype(np.float64), sm["hi"].astype(np.float64)
        c = 0.5 * (lo + hi)
        h_r = rng.uniform(0.0, 20.0, size=3)
        n = rng.normal(size=3)
        n = n / np.linalg.norm(n) if len(rows) % 2 else np.zeros(3)
        h_b = 0.5 * (hi - lo)
        want.append(sm["face"].astype(np.float64).sum() / max(1.0, (np.linalg.norm(h_b) + np.linalg.norm(h_r)) ** 2))
        q = np.zeros(12, np.uint32)
        q[0:9] = np.concatenate([c - h_r, c + h_r, n]).astype(np.float32).view(np.uint32)
        q[9:12] = (g, e, world.generation[g])
        rows.append(q)
    got = _targets(support, view, world, np.array(rows)).astype(np.float64)
    want = np.array(want)
    assert len(got) > 20 and np.all(np.abs(got / want - 1.0) < 1e-5)


def stale_queries(world, box) -> np.ndarray:
    q = world.brick_queries(box)[:8].copy()
    q[0::2, 11] ^= 1  # another version of the row
    q[1::2, 10] = 100000  # past the record count
    return q


def point_vs_zero_box(world, box_pdf, point_pdf, rng) -> None:
    for _ in range(6):
        pick = rng.integers(len(world.centre))
        x = world.centre[pick] + rng.uniform(-30.0, 30.0, size=3)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        q = world.brick_queries(np.concatenate([x, x, n]))
        a = box_pdf(q).astype(np.float64)
        b = point_pdf(q).astype(np.float64)
        assert np.all(np.abs(a - b) <= 1e-5 * np.maximum(a, b) + 1e-12)
        assert abs(b.sum() - 1.0) < 1e-5


# ------------------------------------------------------------------------------------------- oracle


@pytest.fixture(scope="module")
def world_cpu(oracle):
    return LampWorld(oracle, gpu=False)


def test_box_support_oracle(oracle, world_cpu):
    rng = np.random.default_rng(13)
    boxes = world_cpu.boxes(rng, 24, 211)
    q, _ = world_cpu.support_queries(boxes, rng)
    check_support(oracle.light_box_support(world_cpu.oracle_view(), q))


def test_box_pdf_normalised_oracle(oracle, world_cpu):
    boxes = world_cpu.boxes(np.random.default_rng(3), 8, 997)
    check_normalised(world_cpu, lambda q: oracle.light_brick_pdf(world_cpu.oracle_view(), q, True), boxes)


def test_box_draws_oracle(oracle, world_cpu):
    view = world_cpu.oracle_view()
    boxes = world_cpu.boxes(np.random.default_rng(4), 8, 997)
    rep, seeds = draw_boxes(boxes, 64, 5)
    out = oracle.light_sample_brick_box(view, rep, seeds)
    target = lambda q: _targets(oracle.light_box_support, view, world_cpu, q)
    check_draws(world_cpu, out, rep, lambda q: oracle.light_brick_pdf(view, q, True), target)
    rep, seeds = draw_boxes(boxes[:1], 1 << 17, 6)
    check_frequencies(world_cpu, oracle.light_sample_brick_box(view, rep, seeds), boxes[0],
                      lambda q: oracle.light_brick_pdf(view, q, True))


def _targets(support, view, world, brick_q):
    """The support query's target for each (box, brick) of brick_q, through one light face of the brick."""
    q = np.zeros((len(brick_q), 16), np.uint32)
    q[:, 0:3] = brick_q[:, 0:3]
    q[:, 4:7] = brick_q[:, 3:6]
    q[:, 8:11] = brick_q[:, 6:9]
    first = world.face_of_brick()
    faces = np.array([first[(int(g), int(e))] for g, e in zip(brick_q[:, 9], brick_q[:, 10])])
    front = (world.centre[faces] + 0.5 * world.normal[faces]).astype(np.float32)
    q[:, 3] = front[:, 0].view(np.uint32)
    q[:, 7] = front[:, 1].view(np.uint32)
    q[:, 11] = front[:, 2].view(np.uint32)
    q[:, 12] = world.group[faces]
    q[:, 13] = world.vig[faces, 0] | (world.vig[faces, 1] << 10) | (world.vig[faces, 2] << 20)
    q[:, 14] = world.face[faces]
    q[:, 15] = world.block[faces]
    res = support(view, q)
    assert np.all(res[:, 0] == 1.0)
    return res[:, 3].copy()


def test_point_form_and_stale_references_oracle(oracle, world_cpu):
    view = world_cpu.oracle_view()
    point_vs_zero_box(world_cpu, lambda q: oracle.light_brick_pdf(view, q, True),
                      lambda q: oracle.light_brick_pdf(view, q, False), np.random.default_rng(8))
    box = world_cpu.boxes(np.random.default_rng(9), 1, 10 ** 9)[0]
    for box_flag in (True, False):
        assert np.all(oracle.light_brick_pdf(view, stale_queries(world_cpu, box), box_flag) == 0.0)


def test_concentric_targets_oracle(oracle, world_cpu):
    check_concentric(world_cpu, oracle.light_box_support, world_cpu.oracle_view(), np.random.default_rng(12))


def test_tree_selection_oracle(oracle, world_cpu):
    box = world_cpu.boxes(np.random.default_rng(10), 2, 10 ** 9)[1]
    check_selection(world_cpu, lambda s: (lambda q: oracle.light_brick_pdf(world_cpu.oracle_view(s), q, True)),
                    lambda s: (lambda r, sd: oracle.light_sample_brick_box(world_cpu.oracle_view(s), r, sd)), box)


# ------------------------------------------------------------------------------------------- GPU


@pytest.fixture(scope="module")
def world_gpu(oracle):
    return LampWorld(oracle, gpu=True)


def gpu_view(cx, world, select: int = 0):
    v = world.pair.gpu.view
    return cx.CxLightTreeView(v.descriptor, v.mis, select)


@pytest.mark.gpu
def test_box_support_gpu(cx, oracle, world_gpu):
    rng = np.random.default_rng(13)
    boxes = world_gpu.boxes(rng, 160, 17)
    q, _ = world_gpu.support_queries(boxes, rng)
    got = cx.light_box_support(gpu_view(cx, world_gpu), q)
    check_support(got)
    assert np.array_equal(got.view(np.uint32), oracle.light_box_support(world_gpu.oracle_view(), q).view(np.uint32))


@pytest.mark.gpu
def test_box_draws_and_pdfs_gpu(cx, oracle, world_gpu):
    view, ov = gpu_view(cx, world_gpu), world_gpu.oracle_view()
    boxes = world_gpu.boxes(np.random.default_rng(4), 32, 97)
    rep, seeds = draw_boxes(boxes, 256, 5)
    out = cx.light_sample_brick_box(view, rep, seeds)
    assert np.array_equal(out, oracle.light_sample_brick_box(ov, rep, seeds))
    box_pdf = lambda q: cx.light_brick_pdf(view, q, True)
    check_draws(world_gpu, out, rep, box_pdf, lambda q: _targets(cx.light_box_support, view, world_gpu, q))
    check_normalised(world_gpu, box_pdf, boxes[::4])
    for box_flag in (True, False):
        q = np.concatenate([world_gpu.brick_queries(b) for b in boxes[::8]])
        got = cx.light_brick_pdf(view, q, box_flag)
        assert np.array_equal(got.view(np.uint32), oracle.light_brick_pdf(ov, q, box_flag).view(np.uint32))
        assert np.all(cx.light_brick_pdf(view, stale_queries(world_gpu, boxes[0]), box_flag) == 0.0)
    point_vs_zero_box(world_gpu, box_pdf, lambda q: cx.light_brick_pdf(view, q, False), np.random.default_rng(8))
    check_concentric(world_gpu, cx.light_box_support, view, np.random.default_rng(12))
    rep, seeds = draw_boxes(boxes[:1], 1 << 17, 6)
    check_frequencies(world_gpu, cx.light_sample_brick_box(view, rep, seeds), boxes[0], box_pdf)


@pytest.mark.gpu
def test_tree_selection_gpu(cx, oracle, world_gpu):
    box = world_gpu.boxes(np.random.default_rng(10), 2, 10 ** 9)[1]
    check_selection(world_gpu, lambda s: (lambda q: cx.light_brick_pdf(gpu_view(cx, world_gpu, s), q, True)),
                    lambda s: (lambda r, sd: cx.light_sample_brick_box(gpu_view(cx, world_gpu, s), r, sd)), box)
    rep, seeds = draw_boxes(box[None], 256, 11)
    for s in (1, 2):
        assert np.array_equal(cx.light_sample_brick_box(gpu_view(cx, world_gpu, s), rep, seeds),
                              oracle.light_sample_brick_box(world_gpu.oracle_view(s), rep, seeds))
