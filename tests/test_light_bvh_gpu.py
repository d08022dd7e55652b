"""GPU parity of the light-tree build (cx_light_bvh_build, and its 64-bit key form
cx_light_bvh_build_keys64) and the onesweep sorts (cx_radix_sort_pairs, cx_radix_sort_keys64)
against the oracle: bit-exact for Morton codes, sorted order, Karras topology, depths, 4-wide
indices, aggregates, nodes, leaf parents, header and state."""
from __future__ import annotations

import numpy as np
import pytest

import lt_checks
from caelix_amd import layout, synthetic

pytestmark = pytest.mark.gpu


def gpu_and_oracle(cx, oracle, leaves, capacity, node_capacity=None, found=None, builds=0, keys64=False):
    node_capacity = node_capacity if node_capacity is not None else layout.node_capacity_for(capacity)
    n = len(leaves)
    print(f"light bvh build: n={n} capacity={capacity} node_capacity={node_capacity} found={found} "
          f"keys64={keys64}", flush=True)
    builder = cx.LightBvhBuilder(capacity, node_capacity)
    builder.upload(leaves)
    if found is not None or builds:
        import torch
        st = cx.to_host(builder.state).view(layout.STATE).copy()
        st["leaves_found"] = n if found is None else found
        st["builds"] = builds
        builder.state.copy_(cx.to_device(st.view(np.uint8)))
        torch.cuda.synchronize()
    builder.build(keys64=keys64)
    gpu = builder.read()
    state = np.zeros(1, layout.STATE)
    state["building"] = 1
    state["leaves_found"] = n if found is None else found
    state["builds"] = builds
    ref = oracle.light_bvh_build(leaves.view(np.uint8), capacity, node_capacity, state.view(np.uint8))
    return gpu, ref


def compare(gpu, ref, capacity, n):
    g_scratch, r_scratch = gpu["scratch"], ref["scratch"]
    for name, offset, count in [
        ("keys", layout.SCRATCH_KEYS, n),
        ("values", layout.SCRATCH_VALUES, n),
        ("left", layout.SCRATCH_LEFT, max(n - 1, 0)),
        ("right", layout.SCRATCH_RIGHT, max(n - 1, 0)),
        ("parent", layout.SCRATCH_PARENT, max(n - 1, 0)),
        ("sorted_leaf_parent", layout.SCRATCH_SORTED_LEAF_PARENT, n if n >= 2 else 0),
        ("counter", layout.SCRATCH_COUNTER, capacity),
        ("wide", layout.SCRATCH_WIDE, capacity),
    ]:
        g = layout.scratch_u32(g_scratch, offset, capacity)[:count]
        r = layout.scratch_u32(r_scratch, offset, capacity)[:count]
        bad = np.nonzero(g != r)[0]
        assert bad.size == 0, f"{name}: {bad.size} mismatches, first at {bad[:5]}: gpu {g[bad[:5]]} oracle {r[bad[:5]]}"
    if n >= 2:
        g_agg = layout.scratch_aggregates(g_scratch, capacity)[: n - 1]
        r_agg = layout.scratch_aggregates(r_scratch, capacity)[: n - 1]
        assert g_agg.tobytes() == r_agg.tobytes(), "aggregates differ"
    g_header = gpu["header"]
    r_header = ref["header"].view(layout.HEADER)[0]
    assert g_header.tobytes() == r_header.tobytes(), f"header: gpu {g_header} oracle {r_header}"
    g_state = gpu["state"]
    r_state = ref["state"].view(layout.STATE)[0]
    for field in ("building", "leaves_found", "leaves_dropped", "builds", "node_overflow", "bounds"):
        assert np.array_equal(g_state[field], r_state[field]), field
    assert int(g_state["build_error"]) == 0, "a work list of the GPU build passed its bound"
    node_count = int(r_header["node_count"]) if int(r_state["node_overflow"]) == 0 else 0
    g_nodes = gpu["nodes"][:node_count]
    r_nodes = ref["nodes"].view(layout.NODE)[:node_count]
    for field in ("child_face", "child_lo", "child_hi", "child", "parent", "parent_slot"):
        if node_count == 0:
            break
        bad = np.nonzero(np.any((g_nodes[field] != r_nodes[field]).reshape(node_count, -1), axis=1))[0]
        assert bad.size == 0, f"nodes.{field}: {bad.size} mismatches, first {bad[:5]}"
    g_parent = gpu["leaves"]["parent"][:n]
    r_parent = ref["leaves"].view(layout.LEAF)["parent"][:n]
    assert np.array_equal(g_parent, r_parent), "leaf parents differ"


@pytest.mark.parametrize("n", [0, 1, 2, 3, 64, 1000, 4095, 4096, 4097, 70000])
def test_build_matches_oracle_small(cx, oracle, n):
    leaves = synthetic.light_leaves(n, 100 + n, extent=2048.0, clusters=8, sigma=16.0)
    capacity = max(64, (n + 63) // 64 * 64)
    gpu, ref = gpu_and_oracle(cx, oracle, leaves, capacity)
    compare(gpu, ref, capacity, n)
    if n > 0:
        h = gpu["header"]
        lt_checks.check_bvh(h, gpu["state"], gpu["nodes"][: int(h["node_count"])], gpu["leaves"][:n])


@pytest.mark.parametrize("n", [0, 1, 2, 3, 64, 1000, 4095, 4097, 70000])
def test_build_keys64_matches_oracle_small(cx, oracle, n):
    """The 64-bit key form (code << 32 | leaf sorted as one key) makes the oracle's tree too, with
    the same sorted codes and leaves in the scratch."""
    leaves = synthetic.light_leaves(n, 300 + n, extent=2048.0, clusters=8, sigma=16.0)
    capacity = max(64, (n + 63) // 64 * 64)
    gpu, ref = gpu_and_oracle(cx, oracle, leaves, capacity, keys64=True)
    compare(gpu, ref, capacity, n)


def test_build_capacity_above_count_and_dropped(cx, oracle):
    leaves = synthetic.light_leaves(5000, 7)
    gpu, ref = gpu_and_oracle(cx, oracle, leaves, capacity=8192)
    compare(gpu, ref, 8192, 5000)
    # found above the capacity: 64 leaves dropped (the reference's 4,194,240 cap in miniature)
    gpu, ref = gpu_and_oracle(cx, oracle, leaves[:4096], capacity=4096, found=4160)
    compare(gpu, ref, 4096, 4096)
    assert int(gpu["state"]["leaves_dropped"]) == 64


def test_build_node_overflow_and_rebuild_counts(cx, oracle):
    leaves = synthetic.light_leaves(3000, 8)
    gpu, ref = gpu_and_oracle(cx, oracle, leaves, capacity=3008, node_capacity=100, builds=5)
    compare(gpu, ref, 3008, 3000)
    assert int(gpu["state"]["node_overflow"]) == 1
    assert int(gpu["state"]["builds"]) == 6


def test_build_gate_closed_changes_nothing(cx):
    leaves = synthetic.light_leaves(1000, 9)
    builder = cx.LightBvhBuilder(1024)
    builder.upload(leaves, building=0)
    builder.build()
    out = builder.read()
    assert int(out["header"]["leaf_count"]) == 0
    assert int(out["state"]["builds"]) == 0


def test_repeated_builds_reuse_scratch(cx, oracle):
    builder = cx.LightBvhBuilder(16384)
    for seed, n in [(11, 16000), (12, 300), (13, 16384)]:
        leaves = synthetic.light_leaves(n, seed)
        builder.upload(leaves)
        builder.build()
        gpu = builder.read()
        state = np.zeros(1, layout.STATE)
        state["building"] = 1
        state["leaves_found"] = n
        ref = oracle.light_bvh_build(leaves.view(np.uint8), 16384, builder.node_capacity, state.view(np.uint8))
        assert gpu["header"].tobytes()[:60] == ref["header"].tobytes()[:60]
        n_nodes = int(gpu["header"]["node_count"])
        g_nodes = gpu["nodes"][:n_nodes]
        r_nodes = ref["nodes"].view(layout.NODE)[:n_nodes]
        assert np.array_equal(g_nodes["child"], r_nodes["child"])


def test_radix_sort_matches_std_sort_on_reference_inputs(cx):
    import torch

    for n, keys in synthetic.reference_sort_inputs().items():
        print(f"radix sort reference inputs n={n}", flush=True)
        values = np.arange(n, dtype=np.uint32)
        k = torch.from_numpy(keys.view(np.int32).copy()).cuda()
        v = torch.from_numpy(values.view(np.int32).copy()).cuda()
        cx.radix_sort_pairs(k, v)
        torch.cuda.synchronize()
        order = np.lexsort((values, keys))
        assert np.array_equal(k.cpu().numpy().view(np.uint32), keys[order]), n
        assert np.array_equal(v.cpu().numpy().view(np.uint32), values[order]), n


def test_radix_sort_full_32bit_keys(cx, oracle):
    import torch

    rng = np.random.default_rng(5)
    n = 1 << 20
    keys = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    keys[::7] = keys[0]  # many duplicates
    values = np.arange(n, dtype=np.uint32)
    k = torch.from_numpy(keys.view(np.int32).copy()).cuda()
    v = torch.from_numpy(values.view(np.int32).copy()).cuda()
    cx.radix_sort_pairs(k, v)
    torch.cuda.synchronize()
    rk, rv = oracle.radix_sort_pairs(keys, values)
    assert np.array_equal(k.cpu().numpy().view(np.uint32), rk)
    assert np.array_equal(v.cpu().numpy().view(np.uint32), rv)


def test_radix_sort_keys64_matches_stable_sort(cx):
    """cx_radix_sort_keys64 against numpy's stable sort of the same uint64 keys: random 63-bit
    keys with many duplicates, keys that differ only in the low word or only in the high word, and
    sizes around the 4096-key tile."""
    import torch

    rng = np.random.default_rng(64)
    cases = {}
    for n in [1, 2, 100, 4095, 4096, 4097, 1 << 20]:
        keys = rng.integers(0, 2**63, size=n, dtype=np.uint64)
        keys[::5] = keys[0]
        cases[f"random{n}"] = keys
    n = 300000
    cases["low_word"] = (np.uint64(7) << np.uint64(40)) | rng.integers(0, 2**32, size=n, dtype=np.uint64)
    cases["high_word"] = rng.integers(0, 2**31, size=n, dtype=np.uint64) << np.uint64(32)
    cases["code_leaf"] = (rng.integers(0, 2**30, size=n, dtype=np.uint64) << np.uint64(32)) | np.arange(
        n, dtype=np.uint64)
    for name, keys in cases.items():
        print(f"radix sort keys64 {name} n={len(keys)}", flush=True)
        k = torch.from_numpy(keys.view(np.int64).copy()).cuda()
        cx.radix_sort_keys64(k)
        torch.cuda.synchronize()
        assert np.array_equal(k.cpu().numpy().view(np.uint64), np.sort(keys, kind="stable")), name


def test_rebuild_stress_keys64_matches_u32_build(cx):
    """At the full 4,194,304-leaf rebuild size, the 64-bit key build writes exactly what the
    32-bit key build (checked against the oracle in the next test) writes."""
    n = 1 << 22
    leaves = synthetic.light_leaves(n, 2024, extent=8192.0, clusters=64, sigma=32.0)
    outs = []
    for keys64 in (False, True):
        print(f"light bvh stress build keys64={keys64}", flush=True)
        builder = cx.LightBvhBuilder(n)
        builder.upload(leaves)
        builder.build(keys64=keys64)
        outs.append(builder.read())
    a, b = outs
    assert a["header"].tobytes() == b["header"].tobytes()
    assert a["state"].tobytes() == b["state"].tobytes()
    count = int(a["header"]["node_count"])
    for field in ("child_face", "child_lo", "child_hi", "child", "parent", "parent_slot"):
        assert np.array_equal(a["nodes"][field][:count], b["nodes"][field][:count]), field
    assert a["leaves"].tobytes() == b["leaves"].tobytes()
    for name, offset, count in [("keys", layout.SCRATCH_KEYS, n), ("values", layout.SCRATCH_VALUES, n),
                                ("left", layout.SCRATCH_LEFT, n - 1), ("counter", layout.SCRATCH_COUNTER, n),
                                ("wide", layout.SCRATCH_WIDE, n)]:
        ga = layout.scratch_u32(a["scratch"], offset, n)[:count]
        gb = layout.scratch_u32(b["scratch"], offset, n)[:count]
        assert np.array_equal(ga, gb), name
    assert layout.scratch_aggregates(a["scratch"], n)[: n - 1].tobytes() == \
        layout.scratch_aggregates(b["scratch"], n)[: n - 1].tobytes()


def test_rebuild_stress_size_matches_oracle(cx, oracle):
    """The 4,194,304-leaf rebuild workload (half uniform in 8192^3, half in 64 clusters of sigma 32)
    against the oracle, at full size."""
    n = 1 << 22
    leaves = synthetic.light_leaves(n, 2024, extent=8192.0, clusters=64, sigma=32.0)
    gpu, ref = gpu_and_oracle(cx, oracle, leaves, capacity=n)
    compare(gpu, ref, n, n)
    # the reference's own leaf cap: 64 of the 4,194,304 leaves dropped
    capped = 65535 * 64
    gpu, ref = gpu_and_oracle(cx, oracle, leaves[:capped], capacity=capped, found=n)
    compare(gpu, ref, capped, capped)
    assert int(gpu["state"]["leaves_dropped"]) == 64


def test_batch_build_matches_oracle_per_tree(cx, oracle):
    """40 trees of uneven sizes in one batched build (cx_light_bvh_build_batch): each is the oracle's
    tree for its own leaves, and rebuilding the batch gives the same trees again."""
    rng = np.random.default_rng(77)
    sizes = [0, 1, 2, 3, 63, 64, 65, 4095, 4096, 4097, 9000] + list(rng.integers(5, 3000, size=29))
    builders, leaves_of = [], []
    for k, n in enumerate(sizes):
        leaves = synthetic.light_leaves(int(n), 500 + k, extent=1024.0, clusters=4, sigma=20.0)
        capacity = max(64, (int(n) + 63) // 64 * 64)
        b = cx.LightBvhBuilder(capacity)
        b.upload(leaves)
        builders.append(b)
        leaves_of.append((leaves, capacity))
    batch = cx.LightBvhBatch(builders)
    print(f"batch build: {len(builders)} trees, largest capacity {batch.max_capacity}", flush=True)
    for rebuild in range(2):
        batch.build()
        for b, (leaves, capacity) in zip(builders, leaves_of):
            n = len(leaves)
            state = np.zeros(1, layout.STATE)
            state["building"] = 1
            state["leaves_found"] = n
            state["builds"] = rebuild
            ref = oracle.light_bvh_build(leaves.view(np.uint8), capacity, layout.node_capacity_for(capacity),
                                         state.view(np.uint8))
            compare(b.read(), ref, capacity, n)


def test_pool_stress_spills_and_matches_oracle(cx, oracle):
    """One level-2 block whose 31 sub-block boundaries are each crossed by 25 nested nodes: 1,613 pool
    arrivals against 1,536 LDS pool entries (the CPU demand test pins the count), so the upper-level
    build spills records to global memory. The tree is the oracle's, not an empty overflow."""
    leaves = synthetic.pool_stress_leaves(3)
    n = len(leaves)
    capacity = (n + 63) // 64 * 64
    gpu, ref = gpu_and_oracle(cx, oracle, leaves, capacity)
    compare(gpu, ref, capacity, n)
    assert int(gpu["header"]["leaf_count"]) == n and int(gpu["state"]["node_overflow"]) == 0
    h = gpu["header"]
    lt_checks.check_bvh(h, gpu["state"], gpu["nodes"][: int(h["node_count"])], gpu["leaves"][:n])


@pytest.mark.parametrize("kind", ["equal", "caterpillar", "cluster"])
def test_degenerate_inputs_at_rebuild_size_match_oracle(cx, oracle, kind):
    """4,194,304 leaves at the edges of the key distribution (all centres equal; a 30-deep spine over a
    deep equal-key group, 51 levels; one cluster of sigma 0.5): bit-exact with the oracle, non-empty,
    no overflow."""
    n = 1 << 22
    leaves = synthetic.degenerate_leaves(kind, n, 11)
    gpu, ref = gpu_and_oracle(cx, oracle, leaves, capacity=n)
    compare(gpu, ref, n, n)
    assert int(gpu["header"]["leaf_count"]) == n and int(gpu["header"]["root"]) == 0


@pytest.mark.parametrize("bits,tile", [(6, 4096), (6, 8192), (8, 4096)])
def test_sort_tuning_gives_the_same_tree(cx, oracle, bits, tile):
    """Every digit width and tile of the build's sort (cx_light_bvh_set_tuning) sorts stably, so the
    trees are the oracle's: small and odd sizes, a batch of uneven trees, and the 4,194,304-leaf rebuild
    equal to the default setting's bytes."""
    from caelix_amd import autotune

    default = autotune.get_build_tuning()
    assert default == {"digit_bits": 8, "tile": 8192}
    try:
        autotune.set_build_tuning(bits, tile)
        print(f"sort tuning {bits} bits, tile {tile}", flush=True)
        for n in (2, 4097, 70000):
            leaves = synthetic.light_leaves(n, 900 + n, extent=4096.0, clusters=8, sigma=16.0)
            capacity = (n + 63) // 64 * 64
            gpu, ref = gpu_and_oracle(cx, oracle, leaves, capacity)
            compare(gpu, ref, capacity, n)
        sizes = [1, 64, 4095, 9000, 300]
        builders, leaves_of = [], []
        for k, n in enumerate(sizes):
            leaves = synthetic.light_leaves(n, 700 + k, extent=1024.0, clusters=4, sigma=20.0)
            capacity = max(64, (n + 63) // 64 * 64)
            b = cx.LightBvhBuilder(capacity)
            b.upload(leaves)
            builders.append(b)
            leaves_of.append((leaves, capacity))
        cx.LightBvhBatch(builders).build()
        for b, (leaves, capacity) in zip(builders, leaves_of):
            state = np.zeros(1, layout.STATE)
            state["building"] = 1
            state["leaves_found"] = len(leaves)
            ref = oracle.light_bvh_build(leaves.view(np.uint8), capacity, layout.node_capacity_for(capacity),
                                         state.view(np.uint8))
            compare(b.read(), ref, capacity, len(leaves))
        n = 1 << 22
        leaves = synthetic.light_leaves(n, 2024, extent=8192.0, clusters=64, sigma=32.0)
        outs = []
        for setting in ((bits, tile), (8, 8192)):
            autotune.set_build_tuning(*setting)
            builder = cx.LightBvhBuilder(n)
            builder.upload(leaves)
            builder.build()
            outs.append(builder.read())
        a, b = outs
        assert a["header"].tobytes() == b["header"].tobytes() and a["leaves"].tobytes() == b["leaves"].tobytes()
        count = int(a["header"]["node_count"])
        assert a["nodes"][:count].tobytes() == b["nodes"][:count].tobytes()
        for offset in (layout.SCRATCH_KEYS, layout.SCRATCH_VALUES, layout.SCRATCH_LEFT, layout.SCRATCH_WIDE):
            assert np.array_equal(layout.scratch_u32(a["scratch"], offset, n)[: n - 1],
                                  layout.scratch_u32(b["scratch"], offset, n)[: n - 1])
    finally:
        autotune.set_build_tuning(default["digit_bits"], default["tile"])


def test_sort_tuning_rejects_other_settings(cx):
    from caelix_amd import autotune

    for bits, tile in ((7, 8192), (8, 2048), (10, 8192)):
        with pytest.raises(cx.CxError):
            autotune.set_build_tuning(bits, tile)
    assert autotune.get_build_tuning() == {"digit_bits": 8, "tile": 8192}
