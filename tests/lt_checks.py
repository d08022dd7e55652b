"""Independent checks of light-tree builds (used on the oracle's output and on the GPU's).

top_down_radix_tree   the binary radix tree over 64-bit keys (code << 32 | sorted position), built
                      top-down by splitting every range at the highest differing bit; node numbering
                      as Karras 2012 assigns it (root 0; the child covering [a, g] is node g, the one
                      covering [g + 1, b] is node g + 1; a one-leaf range is a leaf). A different
                      algorithm from the bottom-up per-node search of the reference's kernel.
check_bvh             restates the reference test's CheckBvh (tests/light_tree_test.cpp:508-639):
                      every leaf reached once from the root, parent links, inline leaf slots equal
                      to the leaf bytes, node slot sums (1e-5 rel, 1e-4 abs) and box containment,
                      header summary.
"""
from __future__ import annotations

import numpy as np

from caelix_amd import layout

LEAF_BIT = layout.LT_CHILD_LEAF_BIT
NONE = layout.LT_NONE


def morton_bruteforce(x: np.ndarray, y: np.ndarray, z: np.ndarray) -> np.ndarray:
    """Bit b of x -> bit 3b, of y -> 3b + 1, of z -> 3b + 2, one bit at a time."""
    out = np.zeros(len(x), np.uint64)
    for b in range(10):
        out |= ((x.astype(np.uint64) >> b) & 1) << (3 * b)
        out |= ((y.astype(np.uint64) >> b) & 1) << (3 * b + 1)
        out |= ((z.astype(np.uint64) >> b) & 1) << (3 * b + 2)
    return out.astype(np.uint32)


def top_down_radix_tree(sorted_codes: np.ndarray):
    """Returns (left, right, parent, sorted_leaf_parent) arrays for n >= 2 sorted 32-bit codes."""
    n = len(sorted_codes)
    keys = (sorted_codes.astype(np.uint64) << np.uint64(32)) | np.arange(n, dtype=np.uint64)
    left = np.full(n - 1, NONE, np.uint32)
    right = np.full(n - 1, NONE, np.uint32)
    parent = np.full(n - 1, NONE, np.uint32)
    leaf_parent = np.full(n, NONE, np.uint32)
    stack = [(0, n - 1, 0)]  # (first, last, node index)
    while stack:
        a, b, node = stack.pop()
        x = int(keys[a]) ^ int(keys[b])
        bit = x.bit_length() - 1  # highest differing bit
        prefix = int(keys[a]) >> (bit + 1) << (bit + 1)
        threshold = prefix | (1 << bit)
        g = int(np.searchsorted(keys[a:b + 1], np.uint64(threshold), side="left")) + a - 1
        if g == a:
            left[node] = LEAF_BIT | g
            leaf_parent[g] = node
        else:
            left[node] = g
            parent[g] = node
            stack.append((a, g, g))
        if g + 1 == b:
            right[node] = LEAF_BIT | (g + 1)
            leaf_parent[g + 1] = node
        else:
            right[node] = g + 1
            parent[g + 1] = node
            stack.append((g + 1, b, g + 1))
    return left, right, parent, leaf_parent


def depths_by_walk(parent: np.ndarray) -> np.ndarray:
    depth = np.zeros(len(parent), np.uint32)
    for i in range(len(parent)):
        d = 0
        p = parent[i]
        while p != NONE:
            d += 1
            p = parent[p]
        depth[i] = d
    return depth


def near(a, b, rel, abs_tol=1e-6):
    return abs(a - b) <= abs_tol + rel * max(abs(a), abs(b))


def check_bvh(header, state, nodes: np.ndarray, leaves: np.ndarray) -> None:
    """CheckBvh of the reference test, on numpy records (header HEADER, state STATE, nodes NODE[]
    holding node_count entries, leaves LEAF[] holding leaf_count entries)."""
    assert int(state["leaves_dropped"]) == 0
    assert int(state["node_overflow"]) == 0
    leaf_count = int(header["leaf_count"])
    if leaf_count == 0:
        assert int(header["root"]) == NONE
        return
    assert int(header["root"]) == 0
    assert len(nodes) >= 1
    assert int(nodes[0]["parent"]) == NONE

    def node_sum(n):
        face = np.zeros(6, np.float64)
        lo = np.full(3, np.inf, np.float32)
        hi = np.full(3, -np.inf, np.float32)
        for slot in range(4):
            if int(nodes[n]["child"][slot]) == NONE:
                continue
            f = nodes[n]["child_face"][slot].astype(np.float64)
            face += f
            if f.sum() > 0.0:
                lo = np.minimum(lo, nodes[n]["child_lo"][slot])
                hi = np.maximum(hi, nodes[n]["child_hi"][slot])
        return face, lo, hi

    leaf_seen = np.zeros(len(leaves), np.int32)
    node_seen = np.zeros(len(nodes), np.int32)
    stack = [0]
    while stack:
        n = stack.pop()
        assert n < len(nodes)
        assert node_seen[n] == 0
        node_seen[n] = 1
        used = 0
        for slot in range(4):
            child = int(nodes[n]["child"][slot])
            if child == NONE:
                continue
            used += 1
            if child & LEAF_BIT:
                leaf = child & ~LEAF_BIT
                assert leaf < len(leaves)
                assert leaf_seen[leaf] == 0
                leaf_seen[leaf] = 1
                assert int(leaves[leaf]["parent"]) == (n << 2) | slot
                s = leaves[leaf]["summary"]
                assert nodes[n]["child_face"][slot].tobytes() == s["face"].tobytes()
                assert nodes[n]["child_lo"][slot].tobytes() == s["lo"].tobytes()
                assert nodes[n]["child_hi"][slot].tobytes() == s["hi"].tobytes()
            else:
                assert child < len(nodes)
                assert int(nodes[child]["parent"]) == n
                assert int(nodes[child]["parent_slot"]) == slot
                face, lo, hi = node_sum(child)
                for f in range(6):
                    assert near(float(nodes[n]["child_face"][slot][f]), face[f], 1e-5, 1e-4)
                if face.sum() > 0.0:
                    assert np.all(nodes[n]["child_lo"][slot] <= lo)
                    assert np.all(nodes[n]["child_hi"][slot] >= hi)
                stack.append(child)
        assert used >= (1 if leaf_count == 1 else 2)
    assert np.all(leaf_seen == 1)
    assert np.all(node_seen == 1)
    face, lo, hi = node_sum(0)
    for f in range(6):
        assert near(float(header["summary"]["face"][f]), face[f], 1e-5, 1e-4)
    assert np.all(header["summary"]["lo"] <= lo)
    assert np.all(header["summary"]["hi"] >= hi)


def range_sums(sorted_leaf_summaries: np.ndarray, left: np.ndarray, right: np.ndarray):
    """Float64 face sums and float32 boxes of every binary node over its leaf range (recomputed from
    the topology, independent of the aggregation order)."""
    n = len(sorted_leaf_summaries)
    first = np.zeros(n - 1, np.int64)
    last = np.zeros(n - 1, np.int64)
    # ranges from the topology: post-order over the tree
    order = []
    stack = [0]
    while stack:
        i = stack.pop()
        order.append(i)
        for c in (int(left[i]), int(right[i])):
            if not c & LEAF_BIT:
                stack.append(c)
    for i in reversed(order):
        lc, rc = int(left[i]), int(right[i])
        first[i] = (lc & ~LEAF_BIT) if lc & LEAF_BIT else first[lc]
        last[i] = (rc & ~LEAF_BIT) if rc & LEAF_BIT else last[rc]
    faces = sorted_leaf_summaries["face"].astype(np.float64)
    prefix = np.vstack([np.zeros((1, 6)), np.cumsum(faces, axis=0)])
    sums = prefix[last + 1] - prefix[first]
    return first, last, sums


def leaf_ranges(left: np.ndarray, right: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """first[i], last[i]: the sorted-leaf range of binary node i, by pointer jumping down the leftmost
    and rightmost child chains (at most one round per tree level)."""
    def chase(child_of):
        cur = child_of.astype(np.int64)
        inner = (cur & LEAF_BIT) == 0
        while inner.any():
            cur[inner] = child_of[cur[inner]]
            inner = (cur & LEAF_BIT) == 0
        return cur & ~LEAF_BIT
    return chase(left), chase(right)


def block_demand(left: np.ndarray, right: np.ndarray, n: int, tile: int = 512, fanout: int = 32) -> list[dict]:
    """What the nested-block tree build asks of its work lists, per level, from a finished binary tree
    (Karras numbering; n >= 2 leaves): the largest per-block count of open elements (range inside the
    block, parent range not), of gap-hash entries at levels >= 2 (nodes built there, i.e. inside the
    block but inside none of its sub-blocks, plus open elements whose parent's split gap lies inside the
    block and so park) and of pool arrivals (two per built node plus one per parked element)."""
    first, last = leaf_ranges(left, right)
    internal = np.arange(n - 1, dtype=np.int64)
    # Every element: internal nodes 0..n-2, then leaves (sorted positions) as n-1+p.
    e_first = np.concatenate([first, np.arange(n, dtype=np.int64)])
    e_last = np.concatenate([last, np.arange(n, dtype=np.int64)])
    parent = np.full(2 * n - 1, -1, np.int64)
    for child_of in (left, right):
        c = child_of.astype(np.int64)
        is_leaf = (c & LEAF_BIT) != 0
        parent[np.where(is_leaf, n - 1 + (c & ~LEAF_BIT), c)] = internal
    has_parent = parent >= 0
    p = np.where(has_parent, parent, 0)
    p_first, p_last = e_first[p], e_last[p]
    lc = left.astype(np.int64)[p]
    split = np.where((lc & LEAF_BIT) != 0, lc & ~LEAF_BIT, last[np.minimum(lc & ~LEAF_BIT, n - 2)])
    out = []
    size = tile
    level = 1
    while True:
        sub = size // fanout
        inside = (e_first // size) == (e_last // size)
        parent_inside = has_parent & ((p_first // size) == (p_last // size))
        is_open = inside & has_parent & ~parent_inside
        block = e_first // size
        blocks = (n + size - 1) // size
        opened = np.bincount(block[is_open], minlength=blocks)
        row = {"level": level, "tile": size, "blocks": blocks, "open": int(opened.max())}
        if level >= 2:
            built = inside[: n - 1] & ((first // sub) != (last // sub))
            nodes = np.bincount(block[: n - 1][built], minlength=blocks)
            parks = is_open & ((split // size) == block) & (((split + 1) // size) == block)
            parked = np.bincount(block[parks], minlength=blocks)
            row["hash"] = int((nodes + parked).max())
            row["arrivals"] = int((2 * nodes + parked).max())
        out.append(row)
        if size >= n:
            break
        size *= fanout
        level += 1
    return out
