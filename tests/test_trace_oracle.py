"""Pins the trace-side oracle (oracle/trace_oracle.cpp) without the GPU: finalize words and light
records against numpy recomputations from the block payloads, the world DDA against a brute-force
float64 voxel walk, the light-tree sample/pdf pair against the exhaustive normalisation and a
chi-square histogram (as tests/light_tree_test.cpp:888-1068 checks the reference's), and the whole
path tracer against the closed-form radiance of a closed emissive box (a furnace test)."""
from __future__ import annotations

import numpy as np
import pytest

import lt_structure as lts
import trace_checks as tc
from caelix_amd import layout, scene

SMALL = scene.small_scene()
LOAD = 0.7


def slot_count(data) -> int:
    return int(np.ceil(len(data.keys) / LOAD))


@pytest.fixture(scope="module")
def small(oracle):
    return oracle.OracleScene.build(SMALL, slot_count(SMALL))


def test_finalize_words_match_payload(small):
    words, occ, brick_bits, super_bits = tc.finalize_expected(SMALL)
    info = small.a["info"].reshape(-1, 2)
    assert np.array_equal(info[:, 0], words[:, 0])
    assert np.array_equal(info[:, 1] & 0x3FFFF, words[:, 1])
    assert np.array_equal(small.a["occupancy"].reshape(-1, 8), occ)
    assert np.array_equal(small.a["brick_bits"], brick_bits)
    assert np.array_equal(small.a["super_bits"][: len(super_bits)], super_bits)


def test_group_pass_structure_matches_open_faces(small):
    """CheckStructure (light_tree_test.cpp:641-793) on the oracle's group pass, gather and build; record
    word 1 holds the light face counts (stale bits cleared by the pass)."""
    open_ = tc.open_light_faces(SMALL)
    counts = open_.sum(axis=(1, 2))
    info = small.a["info"].reshape(-1, 2)
    assert np.array_equal(info[:, 1] >> 18, np.minimum(counts, 0xFFF))
    coords = small.lt.infos["coord"]
    totals = lts.check_structure(SMALL, small.a["group_of_brick"], coords, lts.from_oracle(small.lt))
    assert totals["bricks"] == len(SMALL.emissive_bricks) and totals["faces"] == counts.sum()
    assert totals["trees"] == [len(SMALL.emissive_bricks), 0]
    c = small.lt.control[0]
    assert int(c["records_recomputed"]) == totals["bricks"] and int(c["records_copied"]) == 0
    assert int(c["passes"]) == 1 and int(c["failed_passes"]) == 0 and int(c["trees_changed"]) == 1


def _air_origins(solid, rng, count, lo, hi):
    out = []
    while len(out) < count:
        p = rng.uniform(lo, hi, size=3)
        c = np.floor(p).astype(int)
        if not solid[c[0], c[1], c[2]]:
            out.append(p)
    return np.array(out)


def test_world_dda_matches_bruteforce_walk(small):
    solid = tc.dense_blocks(SMALL) != 0
    rng = np.random.default_rng(5)
    origins = _air_origins(solid, rng, 1500, 0.5, 127.5)
    dirs = rng.normal(size=(1500, 3))
    dirs[:100, 1] = -np.abs(dirs[:100, 1])
    dirs[100:120] = np.eye(3)[rng.integers(0, 3, 20)] * rng.choice([-1, 1], size=(20, 1))  # axis-aligned
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    o32 = origins.astype(np.float32)
    d32 = dirs.astype(np.float32)
    rays = np.concatenate([o32, d32, np.full((len(o32), 1), 1e30, np.float32)], axis=1)
    hits = small.trace_rays(rays)
    compared = ties = 0
    for i in range(len(rays)):
        voxel, t, axis, tie = tc.walk_first_solid(solid, o32[i].astype(np.float64), d32[i].astype(np.float64))
        if tie < 5e-4:
            ties += 1
            continue
        compared += 1
        h = hits[i]
        if voxel is None:
            assert h[1] == 0, i
            continue
        assert h[1] != 0, i
        v = (h[1] >> 6) & 0x1FF
        got = np.array([h[3] * 8 + (v & 7), h[4] * 8 + ((v >> 3) & 7), h[5] * 8 + (v >> 6)])
        assert np.array_equal(got, voxel), (i, got, voxel)
        t_got = h[0:1].view(np.float32)[0]
        assert abs(t_got - t) < 1e-3 * max(1.0, t), (i, t_got, t)
        flags = h[1] & 0x3F
        expect = 1 << (2 * axis + (1 if d32[i][axis] > 0 else 0))
        assert flags == expect, (i, flags, expect)
        key = tuple(int(k) for k in h[3:6])
        assert SMALL.payload[h[2], v] == h[1] >> 16
        assert tuple(SMALL.keys[h[2]]) == key
    assert compared > 0.9 * len(rays) and ties < 0.1 * len(rays)


def _receivers(solid, rng, count):
    pts = _air_origins(solid, rng, count, 1.0, 127.0)
    normals = np.zeros((count, 3))
    normals[: count // 2, 1] = 1.0
    a, b = count // 2, 3 * count // 4
    normals[a:b] = rng.normal(size=(b - a, 3))  # the rest: no normal (a point receiver)
    normals /= np.maximum(np.linalg.norm(normals, axis=1, keepdims=True), 1e-30)
    return np.concatenate([pts, normals], axis=1).astype(np.float32)


def _all_light_faces(data, small):
    """Queries (12 words) for every open light face of the scene, per receiver."""
    open_ = tc.open_light_faces(data)
    b, v, f = np.nonzero(open_)
    keys = data.keys[b].astype(np.int64)
    vox = keys * 8 + np.stack([v & 7, (v >> 3) & 7, v >> 6], axis=1)
    return b, vox, f, data.payload[b, v].astype(np.uint32)


def _queries(receiver, b, vox, f, block):
    q = np.zeros((len(b), 12), np.uint32)
    q[:, 0:6] = np.broadcast_to(receiver.view(np.uint32), (len(b), 6))
    q[:, 6] = b
    q[:, 7:10] = vox.astype(np.int32).view(np.uint32)
    q[:, 10] = f
    q[:, 11] = block
    return q


def test_light_face_pdf_sums_to_one(small):
    solid = tc.dense_blocks(SMALL) != 0
    rng = np.random.default_rng(9)
    recv = _receivers(solid, rng, 8)
    b, vox, f, block = _all_light_faces(SMALL, small)
    for r in recv:
        pdf = small.light_face_pdf(_queries(r, b, vox, f, block)).astype(np.float64)
        assert np.all(pdf >= 0)
        assert abs(pdf.sum() - 1.0) < 1e-4, pdf.sum()


def _sample_queries(data, small, receiver, samples):
    pos = samples[:, 0:3].astype(np.float64)
    normal = samples[:, 3:6].astype(np.float64)
    vox = np.floor(pos - 0.5 * normal).astype(np.int64)
    key = vox >> 3
    lookup = {tuple(k): i for i, k in enumerate(data.keys.tolist())}
    rec = np.array([lookup[tuple(k)] for k in key.tolist()], np.uint32)
    face = samples[:, 8].view(np.uint32)
    block = samples[:, 7].view(np.uint32)
    return _queries(receiver, rec, vox, face, block)


def test_light_sample_pdf_equals_face_pdf(small):
    solid = tc.dense_blocks(SMALL) != 0
    rng = np.random.default_rng(13)
    recv = _receivers(solid, rng, 6)
    for r in recv:
        seeds = rng.integers(0, 2 ** 32, size=4000, dtype=np.uint64).astype(np.uint32)
        out, ok = small.light_sample(np.repeat(r[None, :], len(seeds), axis=0), seeds)
        assert ok.all()
        pdf = small.light_face_pdf(_sample_queries(SMALL, small, r, out))
        np.testing.assert_allclose(out[:, 6], pdf, rtol=2e-5)
        # the drawn point lies on the drawn face of the drawn voxel
        pos = out[:, 0:3].astype(np.float64)
        n = out[:, 3:6].astype(np.float64)
        frac = pos - np.floor(pos - 0.5 * n)
        axis = np.argmax(np.abs(n), axis=1)
        on = frac[np.arange(len(n)), axis]
        assert np.all(np.where(n[np.arange(len(n)), axis] > 0, on == 1.0, on == 0.0))


def test_light_sample_histogram(small):
    """Chi-square of light-brick draw counts against the per-brick pdf (sum of its face pdfs)."""
    rng = np.random.default_rng(21)
    solid = tc.dense_blocks(SMALL) != 0
    r = _receivers(solid, rng, 1)[0]
    b, vox, f, block = _all_light_faces(SMALL, small)
    pdf = small.light_face_pdf(_queries(r, b, vox, f, block)).astype(np.float64)
    per_brick = np.bincount(b, weights=pdf, minlength=len(SMALL.keys))
    n = 200_000
    seeds = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    out, ok = small.light_sample(np.repeat(r[None, :], n, axis=0), seeds)
    assert ok.all()
    pos = out[:, 0:3].astype(np.float64)
    nrm = out[:, 3:6].astype(np.float64)
    key = np.floor(pos - 0.5 * nrm).astype(np.int64) >> 3
    lookup = {tuple(k): i for i, k in enumerate(SMALL.keys.tolist())}
    drawn = np.array([lookup[tuple(k)] for k in key.tolist()])
    observed = np.bincount(drawn, minlength=len(SMALL.keys)).astype(np.float64)
    expected = per_brick * n
    assert observed[expected == 0].sum() == 0
    big = expected >= 5
    chi2 = ((observed[big] - expected[big]) ** 2 / expected[big]).sum()
    dof = int(big.sum()) - 1
    assert chi2 < dof + 6 * np.sqrt(2 * dof), (chi2, dof)


@pytest.mark.parametrize("mis", [True, False])
@pytest.mark.parametrize("bounces", [1, 2, 3])
def test_path_tracer_furnace_closed_form(oracle, mis, bounces):
    """Closed box, every wall voxel emission E = 1 and albedo a = 0.5: expected radiance after B
    bounces is E (1 - a^B) / (1 - a), with NEE + MIS or NEE alone (MIS off drops the BSDF-hit
    emission of light faces the NEE can draw)."""
    data = scene.furnace_box(width=96, height=96)
    o = oracle.OracleScene.build(data, slot_count(data), mis=mis)
    spp = 4
    color, radiance, attrib, counters = o.trace_frame(data.camera, bounces, 17, spp)
    assert np.all(attrib != 0)
    per_pixel = radiance[:, 0] / spp
    assert np.array_equal(radiance[:, 0], radiance[:, 1])
    mean = per_pixel.mean()
    stderr = per_pixel.std() / np.sqrt(len(per_pixel))
    expected = (1 - 0.5 ** bounces) / (1 - 0.5)
    assert abs(mean - expected) < 5 * stderr + 1e-4, (mean, expected, stderr)
    # primary hits see the wall's emission and albedo
    np.testing.assert_allclose(color[:, 0], spp * 1.0 + 0.5 * radiance[:, 0], rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("q", [0.5, 0.25])
def test_bounce_nee_budget_keeps_the_furnace(oracle, q):
    """The bounce NEE budget (bounce_nee_probability q): a bounce vertex runs its NEE with probability q
    and scales it by 1 / q, so the furnace's closed form still holds; fewer shadow rays are traced."""
    data = scene.furnace_box(width=96, height=96)
    o = oracle.OracleScene.build(data, slot_count(data))
    spp, bounces = 4, 3
    _, radiance, _, counters = o.trace_frame(data.camera, bounces, 17, spp, bounce_nee_probability=q)
    _, _, _, full = o.trace_frame(data.camera, bounces, 17, spp)
    per_pixel = radiance[:, 0] / spp
    stderr = per_pixel.std() / np.sqrt(len(per_pixel))
    expected = (1 - 0.5 ** bounces) / (1 - 0.5)
    assert abs(per_pixel.mean() - expected) < 5 * stderr + 1e-4, (per_pixel.mean(), expected, stderr)
    assert counters[1] < full[1]
    # q = 1 (and 0, taken as 1) draws no number: the plain frame
    assert np.array_equal(o.trace_frame(data.camera, bounces, 17, 1, bounce_nee_probability=1.0)[1],
                          o.trace_frame(data.camera, bounces, 17, 1, bounce_nee_probability=0.0)[1])
