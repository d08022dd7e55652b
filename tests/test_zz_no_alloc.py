"""GPU: cx_light_bvh_build_static (the build on the library's static scratch, for callers that cannot
allocate) against cx_light_bvh_build on the same leaves, through the C ABI: two calls in a row, the
second with a smaller tree, every output bit-exact; sizes above the static scratch are refused."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from caelix_amd import layout, synthetic

pytestmark = pytest.mark.gpu


def static_build(cx, leaves: np.ndarray, capacity: int) -> dict:
    """The static-scratch build over caller-owned leaves, nodes, header and state (no scratch)."""
    import torch

    node_capacity = layout.node_capacity_for(capacity)
    n = len(leaves)
    leaves_dev = cx.device_bytes(capacity * 64)
    leaves_dev.zero_()
    leaves_dev[: n * 64].copy_(cx.to_device(leaves.view(np.uint8)))
    nodes = cx.device_bytes(node_capacity * 224)
    nodes.zero_()
    header = cx.device_bytes(64)
    header.zero_()
    state = np.zeros(1, layout.STATE)
    state["building"] = 1
    state["leaves_found"] = n
    state_dev = cx.to_device(state.view(np.uint8))
    desc = cx.CxLightBvhBuild(leaves_dev.data_ptr(), nodes.data_ptr(), header.data_ptr(), state_dev.data_ptr(), None,
                              capacity, node_capacity)
    cx.check(cx.load().cx_light_bvh_build_static(ctypes.byref(desc), cx.stream_handle()), "cx_light_bvh_build_static")
    torch.cuda.synchronize()
    return {
        "header": cx.to_host(header).view(layout.HEADER)[0].copy(),
        "state": cx.to_host(state_dev).view(layout.STATE)[0].copy(),
        "nodes": cx.to_host(nodes).view(layout.NODE).copy(),
        "leaves": cx.to_host(leaves_dev).view(layout.LEAF).copy(),
    }


def plain_build(cx, leaves: np.ndarray, capacity: int) -> dict:
    builder = cx.LightBvhBuilder(capacity)
    builder.upload(leaves)
    builder.build()
    return builder.read()


def test_static_build_equals_plain_build_twice(cx):
    for n, seed in [(60000, 41), (5000, 42)]:
        capacity = (n + 63) // 64 * 64
        print(f"static-scratch build n={n} capacity={capacity}", flush=True)
        leaves = synthetic.light_leaves(n, seed, extent=4096.0, clusters=16, sigma=24.0)
        a = static_build(cx, leaves, capacity)
        b = plain_build(cx, leaves, capacity)
        assert a["header"].tobytes() == b["header"].tobytes(), n
        assert a["state"].tobytes() == b["state"].tobytes(), n
        count = int(b["header"]["node_count"])
        assert int(b["header"]["leaf_count"]) == n and count > 0
        assert a["nodes"][:count].tobytes() == b["nodes"][:count].tobytes(), n
        assert a["leaves"][:n].tobytes() == b["leaves"][:n].tobytes(), n


def test_static_build_refuses_what_it_cannot_hold(cx):
    lib = cx.load()
    desc = cx.CxLightBvhBuild(1, 1, 1, 1, None, 65600, 41064)  # above the 65,536-leaf static scratch
    assert lib.cx_light_bvh_build_static(ctypes.byref(desc), None) == 1
    desc = cx.CxLightBvhBuild(1, 1, 1, 1, 1, 4096, 2624)  # a caller scratch: not this entry point's
    assert lib.cx_light_bvh_build_static(ctypes.byref(desc), None) == 1
    assert lib.cx_light_bvh_build_static(None, None) == 1
