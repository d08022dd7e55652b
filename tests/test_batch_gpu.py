"""GPU checks of the batched form: chunk scenes of uneven brick counts in one brick store and one
(scene, x, y, z) hash table, their light trees built by one batched build and their frames traced as
one wavefront, give exactly the light trees and images of each scene built and traced alone."""
from __future__ import annotations

import numpy as np
import pytest

from caelix_amd import layout, scene

pytestmark = pytest.mark.gpu

SIZES = [64, 200, 1000, 3000, 70, 5000]


@pytest.fixture(scope="module")
def batch_and_plain(cx):
    from caelix_amd.batch import PlainScenes, TraceBatch

    palette = scene.make_materials(4242, (0.01, 100.0), True)
    datas = [scene.chunk_scene(31 + i, n, grid=16, width=64, height=48, materials=palette) for i, n in enumerate(SIZES)]
    print(f"batch of {len(datas)} chunk scenes, bricks {[len(d.keys) for d in datas]}", flush=True)
    batch = TraceBatch(datas)
    plain = PlainScenes(datas)
    return datas, batch, plain


def test_batched_trees_equal_single_trees(batch_and_plain):
    datas, batch, plain = batch_and_plain
    for i, (b, ts) in enumerate(zip(batch.builders, plain.scenes)):
        got = b.read()
        want = ts.tree_read()
        assert int(batch.light_bricks[i]) == ts.light_bricks
        assert got["header"].tobytes() == want["header"].tobytes(), i
        count = int(want["header"]["node_count"])
        assert got["nodes"][:count].tobytes() == want["nodes"][:count].tobytes(), i
        n = ts.light_bricks
        assert np.array_equal(got["leaves"]["parent"][:n], want["leaves"]["parent"][:n])
        assert got["leaves"]["summary"][:n].tobytes() == want["leaves"]["summary"][:n].tobytes()


@pytest.mark.parametrize("bounces,spp,outdoor", [(0, 1, False), (2, 1, False), (3, 2, False), (2, 1, True)])
def test_batched_frames_equal_single_frames(cx, batch_and_plain, bounces, spp, outdoor):
    """Also under the procedural sky, the sun and fog."""
    datas, batch, plain = batch_and_plain
    k = cx.sky_params(sun_dir=(0.3, 0.9, 0.2), sun_radiance=(40.0, 36.0, 30.0), sun_radius_deg=3.0,
                      zenith=(0.2, 0.3, 0.6), horizon=(0.6, 0.6, 0.7), ground=(0.1, 0.1, 0.1)) if outdoor else None
    fog = ((0.3, 0.3, 0.4), 0.01) if outdoor else None
    batch.counters.zero_()
    plain.counters.zero_()
    batch.build_trees()
    batch.trace(bounces, 17, spp, sky=k, fog=fog)
    plain.step(bounces, 17, spp, sky=k, fog=fog)
    color, radiance, attrib = batch.images()
    p = plain.pixels
    for i in range(len(datas)):
        assert np.array_equal(attrib[i].cpu().numpy(), plain.attrib[i * p:(i + 1) * p].cpu().numpy()), i
        assert np.array_equal(color[i].cpu().numpy().reshape(-1), plain.color[i * p * 4:(i + 1) * p * 4].cpu().numpy()), i
        assert np.array_equal(radiance[i].cpu().numpy().reshape(-1),
                              plain.radiance[i * p * 4:(i + 1) * p * 4].cpu().numpy()), i
    got = batch.counters.cpu().numpy()
    want = plain.counters.cpu().numpy()
    # rays, shadow rays, cells and steps are the same work; probes differ (one shared table)
    assert np.array_equal(got[[0, 1, 2, 4]], want[[0, 1, 2, 4]])


def test_batch_graph_replays_match_eager(batch_and_plain):
    """The whole batch step (trees + frame) captured in a hipGraph: every replay, with eager work
    and synchronisations between replays, gives the eager result."""
    import torch

    datas, batch, plain = batch_and_plain
    batch.build_trees()
    batch.trace(2, 5, 1)
    torch.cuda.synchronize()
    eager = [t.clone() for t in batch.images()]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        batch.build_trees()  # warm-up on the capture stream
        batch.trace(2, 5, 1)
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=side):
            batch.build_trees()
            batch.trace(2, 5, 1)
    torch.cuda.current_stream().wait_stream(side)
    for replay in range(3):
        batch.color.zero_()
        batch.attrib.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for a, b in zip(batch.images(), eager):
            assert torch.equal(a, b), replay


def test_tuning_changes_speed_not_results(batch_and_plain):
    from caelix_amd import autotune

    datas, batch, plain = batch_and_plain
    before = autotune.get_tuning()
    try:
        batch.build_trees()
        batch.trace(2, 23, 1)
        reference = [t.clone() for t in batch.images()]
        for chunk, per_cu in [(32, 0), (256, 2), (64, 4)]:
            autotune.set_tuning(chunk, per_cu)
            batch.trace(2, 23, 1)
            for a, b in zip(batch.images(), reference):
                assert np.array_equal(a.cpu().numpy(), b.cpu().numpy()), (chunk, per_cu)
        with pytest.raises(RuntimeError):
            autotune.set_tuning(0, 0)
        result = autotune.tune("test-batch", lambda: batch.trace(2, 23, 1), reps=1, use_cache=False)
        assert result["queue_chunk"] in autotune.QUEUE_CHUNKS and len(result["trials"]) == 12
    finally:
        autotune.set_tuning(before["queue_chunk"], before["workgroups_per_cu"])
