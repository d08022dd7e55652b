"""CPU: pins the light-tree build oracle (oracle/light_bvh_oracle.cpp) against checks that do not
restate its formulas: brute-force Morton interleave, numpy lexsort (= std::sort of (key, value)
pairs, the reference test's own oracle), a top-down radix-tree construction, float64 range sums,
a parent-walk depth, and the reference test's CheckBvh traversal."""
from __future__ import annotations

import numpy as np
import pytest

import lt_checks
from caelix_amd import layout, synthetic

LEAF_BIT = layout.LT_CHILD_LEAF_BIT
NONE = layout.LT_NONE


def run_oracle(oracle, leaves, capacity=None, node_capacity=None, found=None, building=1, builds=0):
    n = len(leaves)
    capacity = capacity if capacity is not None else max(64, (n + 63) // 64 * 64)
    node_capacity = node_capacity if node_capacity is not None else layout.node_capacity_for(capacity)
    state = np.zeros(1, layout.STATE)
    state["building"] = building
    state["leaves_found"] = n if found is None else found
    state["builds"] = builds
    out = oracle.light_bvh_build(leaves.view(np.uint8), capacity, node_capacity, state.view(np.uint8))
    out["header_rec"] = out["header"].view(layout.HEADER)[0]
    out["state_rec"] = out["state"].view(layout.STATE)[0]
    out["nodes_rec"] = out["nodes"].view(layout.NODE)
    out["leaves_rec"] = out["leaves"].view(layout.LEAF)
    out["capacity"] = capacity
    return out


def test_morton_matches_bruteforce_interleave(oracle):
    rng = np.random.default_rng(1)
    cells = rng.integers(0, 1024, size=(2000, 3)).astype(np.uint32)
    cells[:8] = [[0, 0, 0], [1023, 1023, 1023], [1, 0, 0], [0, 1, 0], [0, 0, 1], [512, 0, 0], [0, 512, 0], [0, 0, 512]]
    expected = lt_checks.morton_bruteforce(cells[:, 0], cells[:, 1], cells[:, 2])
    got = np.array([oracle.morton3(int(x), int(y), int(z)) for x, y, z in cells], np.uint32)
    assert np.array_equal(got, expected)
    assert oracle.morton3(1, 0, 0) == 1 and oracle.morton3(0, 1, 0) == 2 and oracle.morton3(0, 0, 1) == 4


def test_oracle_radix_sort_matches_std_sort_on_reference_inputs(oracle):
    for n, keys in synthetic.reference_sort_inputs().items():
        values = np.arange(n, dtype=np.uint32)
        k, v = oracle.radix_sort_pairs(keys, values)
        order = np.lexsort((values, keys))
        assert np.array_equal(k, keys[order]), n
        assert np.array_equal(v, values[order]), n


@pytest.mark.parametrize("n,seed", [(2, 1), (3, 2), (17, 3), (1000, 4), (5000, 5)])
def test_oracle_build_matches_independent_checks(oracle, n, seed):
    leaves = synthetic.light_leaves(n, seed, extent=512.0, clusters=4, sigma=6.0)
    out = run_oracle(oracle, leaves)
    cap = out["capacity"]
    scratch = out["scratch"]
    header = out["header_rec"]
    assert int(header["leaf_count"]) == n

    # bounds of the centres
    centre = 0.5 * (leaves["summary"]["lo"] + leaves["summary"]["hi"])
    assert np.array_equal(out["state_rec"]["bounds"][:3], centre.min(axis=0))
    assert np.array_equal(out["state_rec"]["bounds"][3:], centre.max(axis=0))

    # Morton keys from float32 quantisation, interleaved bit by bit
    lo = out["state_rec"]["bounds"][:3]
    extent = np.maximum(out["state_rec"]["bounds"][3:] - lo, np.float32(1e-6))
    q = np.clip((centre - lo) / extent * np.float32(1024.0), 0.0, 1023.0).astype(np.uint32)
    # the quantisation against float64 arithmetic (the mathematics the float32 formula rounds): equal
    # cells except where the float64 position lies within 1e-3 of a cell boundary
    x64 = (centre.astype(np.float64) - lo.astype(np.float64)) / extent.astype(np.float64) * 1024.0
    q64 = np.clip(np.floor(x64), 0, 1023).astype(np.uint32)
    near_edge = np.abs(x64 - np.round(x64)) < 1e-3
    assert np.all((q == q64) | near_edge) and (q != q64).mean() < 0.01
    codes = lt_checks.morton_bruteforce(q[:, 0], q[:, 1], q[:, 2])

    # sorted pairs = std::sort of (code, leaf index)
    order = np.lexsort((np.arange(n), codes))
    keys0 = layout.scratch_u32(scratch, layout.SCRATCH_KEYS, cap)[:n]
    values0 = layout.scratch_u32(scratch, layout.SCRATCH_VALUES, cap)[:n]
    assert np.array_equal(keys0, codes[order])
    assert np.array_equal(values0, order.astype(np.uint32))

    # Karras topology = top-down radix tree
    left, right, parent, leaf_parent = lt_checks.top_down_radix_tree(codes[order])
    assert np.array_equal(layout.scratch_u32(scratch, layout.SCRATCH_LEFT, cap)[: n - 1], left)
    assert np.array_equal(layout.scratch_u32(scratch, layout.SCRATCH_RIGHT, cap)[: n - 1], right)
    assert np.array_equal(layout.scratch_u32(scratch, layout.SCRATCH_PARENT, cap)[: n - 1], parent)
    assert np.array_equal(layout.scratch_u32(scratch, layout.SCRATCH_SORTED_LEAF_PARENT, cap)[:n], leaf_parent)

    # depths by parent walk; 4-wide indices = exclusive scan of even-depth flags over the capacity
    depth = lt_checks.depths_by_walk(parent)
    assert np.array_equal(layout.scratch_u32(scratch, layout.SCRATCH_COUNTER, cap)[: n - 1], depth)
    flags = np.zeros(cap, np.uint32)
    flags[: n - 1] = (depth % 2 == 0)
    wide = np.concatenate([[0], np.cumsum(flags)[:-1]]).astype(np.uint32)
    assert np.array_equal(layout.scratch_u32(scratch, layout.SCRATCH_WIDE, cap), wide)
    assert int(header["node_count"]) == int(flags.sum())

    # aggregates: float64 range sums and exact boxes
    sorted_summaries = leaves["summary"][order]
    first, last, sums = lt_checks.range_sums(sorted_summaries, left, right)
    agg = layout.scratch_aggregates(scratch, cap)[: n - 1]
    assert np.allclose(agg["face"].astype(np.float64), sums, rtol=1e-5, atol=1e-4)
    for i in range(n - 1):
        a, b = first[i], last[i]
        assert np.array_equal(agg["lo"][i], sorted_summaries["lo"][a:b + 1].min(axis=0))
        assert np.array_equal(agg["hi"][i], sorted_summaries["hi"][a:b + 1].max(axis=0))

    # the reference test's structure check
    nodes = out["nodes_rec"][: int(header["node_count"])]
    lt_checks.check_bvh(header, out["state_rec"], nodes, out["leaves_rec"][:n])


def test_oracle_empty_single_and_gate(oracle):
    empty = run_oracle(oracle, synthetic.light_leaves(0, 1))
    h = empty["header_rec"]
    assert int(h["root"]) == layout.LT_NONE and int(h["leaf_count"]) == 0 and int(h["node_count"]) == 0
    assert int(empty["state_rec"]["builds"]) == 1

    one = run_oracle(oracle, synthetic.light_leaves(1, 2))
    h = one["header_rec"]
    assert int(h["root"]) == 0 and int(h["leaf_count"]) == 1 and int(h["node_count"]) == 1
    node = one["nodes_rec"][0]
    assert list(node["child"]) == [layout.LT_CHILD_LEAF_BIT, layout.LT_NONE, layout.LT_NONE, layout.LT_NONE]
    assert int(one["leaves_rec"][0]["parent"]) == 0
    lt_checks.check_bvh(h, one["state_rec"], one["nodes_rec"][:1], one["leaves_rec"][:1])

    gated = run_oracle(oracle, synthetic.light_leaves(100, 3), building=0, builds=7)
    assert int(gated["state_rec"]["builds"]) == 7
    assert int(gated["header_rec"]["leaf_count"]) == 0


def test_oracle_dropped_leaves_and_node_overflow(oracle):
    leaves = synthetic.light_leaves(200, 4)
    dropped = run_oracle(oracle, leaves[:128], capacity=128, found=200)
    assert int(dropped["header_rec"]["leaf_count"]) == 128
    assert int(dropped["state_rec"]["leaves_dropped"]) == 72
    overflow = run_oracle(oracle, leaves, capacity=256, node_capacity=8)
    assert int(overflow["state_rec"]["node_overflow"]) == 1
    assert int(overflow["header_rec"]["leaf_count"]) == 0
    assert int(overflow["header_rec"]["root"]) == layout.LT_NONE
    assert int(overflow["header_rec"]["node_count"]) > 8


# The work-list bounds of the nested-block GPU build (caelix-renderer_amd/csrc/light_bvh.hip): open
# elements per block, gap-hash entries and pool arrivals per upper-level block. They follow from at most
# 60 internal nodes on a root-to-leaf path (the deltas 2..31 of distinct 30-bit codes and 34..63 of equal
# codes strictly increase down a path).
OPEN_CAP, HASH_ENTRIES, POOL_LDS, ARRIVAL_BOUND, DELTA_LEVELS = 256, 2048, 1536, 3840, 60


def demand_of(out):
    cap, n = out["capacity"], int(out["header_rec"]["leaf_count"])
    left = layout.scratch_u32(out["scratch"], layout.SCRATCH_LEFT, cap)[: n - 1]
    right = layout.scratch_u32(out["scratch"], layout.SCRATCH_RIGHT, cap)[: n - 1]
    depth = layout.scratch_u32(out["scratch"], layout.SCRATCH_COUNTER, cap)[: n - 1]
    return lt_checks.block_demand(left, right, n), int(depth.max())


@pytest.mark.parametrize("kind", ["equal", "caterpillar", "cluster"])
def test_degenerate_inputs_stay_inside_the_work_list_bounds(oracle, kind):
    """Equal centres, a 30-deep spine over a deep equal-key group, and one tight cluster: the oracle's
    tree (pinned by the checks above) asks the GPU build's work lists for less than their bounds."""
    n = 1 << 20
    out = run_oracle(oracle, synthetic.degenerate_leaves(kind, n, 11))
    assert int(out["header_rec"]["leaf_count"]) == n
    demand, depth = demand_of(out)
    assert depth + 1 <= DELTA_LEVELS
    if kind == "caterpillar":
        assert depth >= 45  # the spine and the equal-key group below it
    for row in demand:
        assert row["open"] <= 2 * DELTA_LEVELS <= OPEN_CAP, row
        if row["level"] >= 2:
            assert row["hash"] <= HASH_ENTRIES and row["arrivals"] <= ARRIVAL_BOUND, row


def test_pool_stress_input_needs_the_spill(oracle):
    """The pool stress leaves cross all 31 sub-block boundaries of one level-2 block with 25 nested
    nodes each: more pool arrivals than the block's LDS pool holds (the GPU build spills the rest), and
    still inside the arrival bound. The construction is checked against the top-down radix tree too."""
    leaves = synthetic.pool_stress_leaves(3)
    out = run_oracle(oracle, leaves)
    n = len(leaves)
    assert int(out["header_rec"]["leaf_count"]) == n
    demand, depth = demand_of(out)
    level2 = demand[1]
    assert level2["level"] == 2
    assert POOL_LDS < level2["arrivals"] <= ARRIVAL_BOUND, level2
    assert level2["hash"] == 807  # 31 boundaries x 25 nested nodes + the 32 nodes over the parts
    cap = out["capacity"]
    keys = layout.scratch_u32(out["scratch"], layout.SCRATCH_KEYS, cap)[:n]
    left, right, _, _ = lt_checks.top_down_radix_tree(keys)
    assert np.array_equal(left, layout.scratch_u32(out["scratch"], layout.SCRATCH_LEFT, cap)[: n - 1])
    assert np.array_equal(right, layout.scratch_u32(out["scratch"], layout.SCRATCH_RIGHT, cap)[: n - 1])


@pytest.mark.parametrize("n,seed", [(3000, 21), (777, 22)])
def test_wide_slot_order_from_the_top_down_tree(oracle, n, seed):
    """Every 4-wide node as caelix_light_bvh_wide.hlsl:10-100 makes it, recomputed from the top-down
    radix tree (not the oracle's): node w = the exclusive scan of the even-depth flags at binary node i;
    its slots are i's children in order, an internal child replaced by its two children (left first); a
    leaf slot holds kLtChildLeafBit | the leaf's index, an internal one its 4-wide index; leaf parents
    are w << 2 | slot, child nodes' parent links (w, slot)."""
    leaves = synthetic.light_leaves(n, seed, extent=512.0, clusters=3, sigma=2.0)  # many equal codes
    out = run_oracle(oracle, leaves)
    cap = out["capacity"]
    keys = layout.scratch_u32(out["scratch"], layout.SCRATCH_KEYS, cap)[:n]
    values = layout.scratch_u32(out["scratch"], layout.SCRATCH_VALUES, cap)[:n]
    assert len(np.unique(keys)) < n  # equal codes present
    left, right, parent, _ = lt_checks.top_down_radix_tree(keys)
    depth = lt_checks.depths_by_walk(parent)
    even = (depth & 1) == 0
    wide = np.concatenate([[0], np.cumsum(even)[:-1]]).astype(np.int64)
    nodes = out["nodes_rec"]
    leaf_parent = out["leaves_rec"]["parent"]
    count = int(out["header_rec"]["node_count"])
    assert count == int(even.sum())
    for i in np.nonzero(even)[0]:
        w = int(wide[i])
        slots = []
        for c in (int(left[i]), int(right[i])):
            slots += [c] if c & LEAF_BIT else [int(left[c]), int(right[c])]
        words = [LEAF_BIT | int(values[s & ~LEAF_BIT]) if s & LEAF_BIT else int(wide[s]) for s in slots]
        words += [NONE] * (4 - len(words))
        assert list(nodes[w]["child"]) == words, (i, w)
        for s, word in enumerate(words[: len(slots)]):
            if word & LEAF_BIT:
                assert int(leaf_parent[word & ~LEAF_BIT]) == (w << 2) | s
            else:
                assert (int(nodes[word]["parent"]), int(nodes[word]["parent_slot"])) == (w, s)
    assert int(nodes[0]["parent"]) == NONE
