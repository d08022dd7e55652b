"""The batched form: many chunk scenes built and traced together.

TraceBatch puts the bricks of every scene in one store (records scene-major), keys them by
(scene, x, y, z) in one bulk hash table, finalizes them and runs the light pass once, builds the
per-scene light trees with one batched build (cx_light_bvh_build_batch, each tree over its scene's
slice of the light leaves) and traces every scene's frame as one wavefront (cx_trace_batch). Both
steps record only kernel launches, so a whole batch step can be captured in a hipGraph and replayed.

PlainScenes is the launch-per-scene form of the same work (one TraceScene, one build and one
cx_trace_frame per scene), the baseline the batched form is compared with.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import (CxCamera, CxLightScene, CxLightTreeView, CxScene, CxTraceOutput, CxTraceParams, LightBvhBatch,
               LightBvhBuilder, TraceScene, _torch, apply_sky, camera_struct, check, device_bytes, layout, load,
               stream_handle, to_device, to_host)


class TraceBatch:
    def __init__(self, scenes, load_factor: float = 0.7, mis: bool = True):
        torch = _torch()
        lib = self.lib = load()
        self.datas = list(scenes)
        S = self.count = len(self.datas)
        grid = int(self.datas[0].grid_dim)
        shift = int(self.datas[0].group_shift)
        for d in self.datas:
            if int(d.grid_dim) != grid or int(d.group_shift) != shift or tuple(d.grid_min) != (0, 0, 0):
                raise ValueError("the scenes of a batch share grid_dim, group_shift and grid_min (0, 0, 0)")
        cam0 = self.datas[0].camera
        self.width, self.height = int(cam0["width"]), int(cam0["height"])
        counts = np.array([len(d.keys) for d in self.datas], np.int64)
        self.brick_offset = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        n = self.bricks = int(self.brick_offset[-1])
        keys4 = np.empty((n, 4), np.int32)
        for i, d in enumerate(self.datas):
            a, b = self.brick_offset[i], self.brick_offset[i + 1]
            keys4[a:b, 0] = i
            keys4[a:b, 1:] = d.keys
        payload = np.concatenate([d.payload for d in self.datas])
        dev = "cuda"
        self.keys = torch.from_numpy(keys4).to(dev)
        self.payload = torch.from_numpy(payload.view(np.int16)).to(dev)
        self.materials = torch.from_numpy(np.ascontiguousarray(self.datas[0].materials, np.float32)).to(dev)
        self.light_codes = torch.from_numpy(np.ascontiguousarray(self.datas[0].light_codes).view(np.int32)).to(dev)
        # one bulk table keyed by (scene, x, y, z)
        self.slot_count = max(int(np.ceil(n / load_factor)), n + 1)
        self.slots = torch.empty(self.slot_count * 4, dtype=torch.int32, device=dev)
        self.hash_info = torch.zeros(4, dtype=torch.int32, device=dev)
        hs = device_bytes(int(lib.cx_hash_build_scratch_bytes(n)), dev)
        check(lib.cx_hash_build_keys(self.keys.data_ptr(), 4, n, self.slots.data_ptr(), self.slot_count,
                                     hs.data_ptr(), self.hash_info.data_ptr(), stream_handle()), "cx_hash_build_keys")
        cells = grid ** 3
        self.info = torch.zeros(n * 2, dtype=torch.int32, device=dev)
        self.occupancy = torch.zeros(n * 8, dtype=torch.int64, device=dev)
        self.brick_bits = torch.zeros((S * cells + 31) // 32, dtype=torch.int32, device=dev)
        self.super_bits = torch.zeros((S * (cells // 512) + 31) // 32 + 1, dtype=torch.int32, device=dev)
        self.occupied_box = torch.zeros(S * 6, dtype=torch.int32, device=dev)
        sc = self.scene = CxScene()
        sc.keys = self.keys.data_ptr()
        sc.info = self.info.data_ptr()
        sc.occupancy = self.occupancy.data_ptr()
        sc.payload = self.payload.data_ptr()
        sc.slots = self.slots.data_ptr()
        sc.slot_count = self.slot_count
        sc.bricks = n
        sc.hash_info = self.hash_info.data_ptr()
        sc.brick_bits = self.brick_bits.data_ptr()
        sc.super_bits = self.super_bits.data_ptr()
        sc.grid_min[:] = [0, 0, 0]
        sc.grid_dim = grid
        sc.group_shift = shift
        sc.encoding = 1
        sc.materials = self.materials.data_ptr()
        sc.light_codes = self.light_codes.data_ptr()
        sc.emission_scale = 1.0
        sc.scene_count = S
        sc.occupied_box = self.occupied_box.data_ptr()
        h = stream_handle()
        check(lib.cx_scene_finalize(ctypes.byref(sc), h), "cx_scene_finalize")
        # the light pass over all scenes: ordinals and leaves scene-major
        self.light_scratch = device_bytes(int(lib.cx_scene_light_scratch_bytes(n)), dev)
        self.record_offset = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
        self.ordinal_of_brick = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
        self.totals = torch.zeros(4, dtype=torch.int32, device=dev)
        self.lt_state = torch.zeros(16, dtype=torch.int32, device=dev)
        leaf_capacity = max(64, (n + 63) // 64 * 64)
        self.leaves = device_bytes((leaf_capacity + 64 * S) * 64, dev)
        heap_units = max(64, n * 8)
        while True:
            self.heap = torch.zeros(heap_units * 4, dtype=torch.int32, device=dev)
            ls = self.light = CxLightScene()
            ls.heap = self.heap.data_ptr()
            ls.heap_units = heap_units
            ls.leaf_capacity = leaf_capacity
            ls.record_offset = self.record_offset.data_ptr()
            ls.ordinal_of_brick = self.ordinal_of_brick.data_ptr()
            ls.leaves = self.leaves.data_ptr()
            ls.state = self.lt_state.data_ptr()
            ls.totals = self.totals.data_ptr()
            ls.scratch = self.light_scratch.data_ptr()
            check(lib.cx_scene_light_records(ctypes.byref(sc), ctypes.byref(ls), h), "cx_scene_light_records")
            totals = self.totals.cpu().numpy().view(np.uint32)
            if totals[3] == 0:
                break
            heap_units = int(totals[1]) + 64
        # per-scene leaf slices: a scene's light bricks are the consecutive ordinals of its records
        light = (self.info.view(-1, 2)[:, 1].cpu().numpy().view(np.uint32) >> 18) > 0
        per_scene = np.array([light[self.brick_offset[i]:self.brick_offset[i + 1]].sum() for i in range(S)], np.int64)
        self.leaf_offset = np.concatenate([[0], np.cumsum(per_scene)]).astype(np.int64)
        self.light_bricks = per_scene
        self.builders = []
        # per-scene views read their tree's header in place: headers[scene] = (tree header, empty tree)
        self.headers = torch.zeros(S * 32, dtype=torch.int32, device=dev)
        empty = np.zeros(1, layout.HEADER)
        empty["root"] = layout.LT_NONE
        empty_words = torch.from_numpy(empty.view(np.int32).copy()).to(dev)
        for i in range(S):
            self.headers[i * 32 + 16:i * 32 + 32].copy_(empty_words)
        header_bytes = self.headers.view(torch.uint8)
        for i in range(S):
            cap = max(64, (int(per_scene[i]) + 63) // 64 * 64)
            view = self.leaves[int(self.leaf_offset[i]) * 64:(int(self.leaf_offset[i]) + cap) * 64]
            b = LightBvhBuilder(cap, leaves=view, header=header_bytes[i * 128:i * 128 + 64])
            st = np.zeros(1, layout.STATE)
            st["building"] = 1
            st["leaves_found"] = int(per_scene[i])
            b.state.copy_(to_device(st.view(np.uint8), b.state.device))
            self.builders.append(b)
        self.tree_batch = LightBvhBatch(self.builders)
        self.empty_nodes = torch.zeros(56, dtype=torch.int32, device=dev)
        self.empty_leaves = torch.zeros(16, dtype=torch.int32, device=dev)
        views = (CxLightTreeView * S)()
        cams = (CxCamera * S)()
        for i, d in enumerate(self.datas):
            v = views[i]
            v.headers = self.headers.data_ptr() + i * 128
            v.nodes[0] = self.builders[i].nodes.data_ptr()
            v.nodes[1] = self.empty_nodes.data_ptr()
            v.leaves[0] = self.builders[i].leaves.data_ptr()
            v.leaves[1] = self.empty_leaves.data_ptr()
            v.heap = self.heap.data_ptr()
            v.record_offset = self.record_offset.data_ptr()
            v.ordinal_of_brick = self.ordinal_of_brick.data_ptr()
            v.group_shift = shift
            v.mis = 1 if mis else 0
            v.leaf_base = int(self.leaf_offset[i])
            cams[i] = camera_struct(d.camera)
        self._views, self._cams = views, cams
        self.device_views = device_bytes(int(lib.cx_trace_batch_bytes(S)), dev)
        check(lib.cx_trace_batch_prepare(views, cams, S, self.device_views.data_ptr(), h), "cx_trace_batch_prepare")
        pixels = self.width * self.height
        self.color = torch.empty(S * pixels * 4, dtype=torch.float32, device=dev)
        self.radiance = torch.empty(S * pixels * 4, dtype=torch.float32, device=dev)
        self.attrib = torch.empty(S * pixels, dtype=torch.int32, device=dev)
        self.counters = torch.zeros(8, dtype=torch.int64, device=dev)
        self.workspace = device_bytes(int(lib.cx_trace_batch_workspace_bytes(self.width, self.height, S)), dev)
        self.build_trees()
        torch.cuda.synchronize()

    def build_trees(self, stream=None) -> None:
        """Rebuild every scene's light tree: one batched build (the headers land in the views)."""
        self.tree_batch.build(stream)

    def trace(self, bounces: int, seed: int, spp: int = 1, flags: int = 0, stream=None, sky: dict | None = None,
              fog=None):
        params = CxTraceParams(bounces, seed, spp, flags)
        apply_sky(params, sky, fog)
        out = CxTraceOutput(self.color.data_ptr(), self.radiance.data_ptr(), self.attrib.data_ptr(),
                            self.counters.data_ptr(), self.workspace.data_ptr(), self.workspace.numel())
        check(self.lib.cx_trace_batch(ctypes.byref(self.scene), self.device_views.data_ptr(), self.count, self.width,
                                      self.height, ctypes.byref(params), ctypes.byref(out), stream_handle(stream)),
              "cx_trace_batch")

    def images(self):
        """(color, radiance, attrib) per scene: [S, H*W, 4], [S, H*W, 4], [S, H*W]."""
        pixels = self.width * self.height
        return (self.color.view(self.count, pixels, 4), self.radiance.view(self.count, pixels, 4),
                self.attrib.view(self.count, pixels))


class PlainScenes:
    """The launch-per-scene form: one TraceScene per chunk scene; a step rebuilds each scene's light
    tree and traces its frame, one scene after the other."""

    def __init__(self, scenes, load_factor: float = 0.7):
        torch = _torch()
        self.scenes = [TraceScene(d, load_factor=load_factor) for d in scenes]
        cam = scenes[0].camera
        pixels = int(cam["width"]) * int(cam["height"])
        self.cameras = [d.camera for d in scenes]
        self.color = torch.empty(len(scenes) * pixels * 4, dtype=torch.float32, device="cuda")
        self.radiance = torch.empty(len(scenes) * pixels * 4, dtype=torch.float32, device="cuda")
        self.attrib = torch.empty(len(scenes) * pixels, dtype=torch.int32, device="cuda")
        self.counters = torch.zeros(8, dtype=torch.int64, device="cuda")
        self.pixels = pixels

    def step(self, bounces: int, seed: int, spp: int = 1, stream=None, sky: dict | None = None, fog=None) -> None:
        p = self.pixels
        for i, (ts, cam) in enumerate(zip(self.scenes, self.cameras)):
            ts.tree.build(stream)
            ts.trace_frame(cam, bounces, seed, spp, color=self.color[i * p * 4:(i + 1) * p * 4],
                           radiance=self.radiance[i * p * 4:(i + 1) * p * 4], attrib=self.attrib[i * p:(i + 1) * p],
                           counters=self.counters, stream=stream, sky=sky, fog=fog)
