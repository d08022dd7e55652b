"""The batched form: many chunk scenes built and traced together.

TraceBatch puts the bricks of every scene in one store (records scene-major), keys them by
(scene, x, y, z) in one bulk hash table, and runs one light tree's tables over all of them: each scene
is an entity (identity frame, static) whose groups are consecutive (scene-major group numbering); one
finalize with the light state, one group pass, and one gather that numbers each entity's leaves from
its own base (cx_light_gather_entities), so each scene's leaves form a tree of its own. The per-scene
trees are built with one batched build (cx_light_bvh_build_batch), each scene gets a descriptor of its
own (its control with its tree header, its nodes and its slice of the leaves; the shared heap, rows,
infos and codes), and every scene's frame is traced as one wavefront (cx_trace_batch). Both steps record
only kernel launches, so a whole batch step can be captured in a hipGraph and replayed.

PlainScenes is the launch-per-scene form of the same work (one TraceScene, one build and one
cx_trace_frame per scene), the baseline the batched form is compared with.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import (CxCamera, CxLightPass, CxLightTreeView, CxScene, CxTraceOutput, CxTraceParams, LightBvhBatch,
               LightBvhBuilder, TraceScene, _torch, apply_sky, camera_struct, check, device_bytes, entity_infos,
               group_infos, groups_of, layout, light_codes_for, load, min_emission_for, stream_handle, to_device,
               to_host)


def _tensor(array: np.ndarray, dev="cuda"):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(array).view(np.uint8).copy()).to(dev)


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
        materials = np.ascontiguousarray(self.datas[0].materials, np.float32)
        self.materials = torch.from_numpy(materials).to(dev)
        encoding = 1
        codes = light_codes_for(materials, min_emission_for(self.datas[0]), encoding)
        self.light_codes = torch.from_numpy(codes.view(np.int32)).to(dev)
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
        # the light tree's tables: an entity per scene, groups scene-major
        coords, gob = groups_of(keys4, shift)
        G = self.groups = len(coords)
        self.group_of_brick = torch.from_numpy(gob.view(np.int32)).to(dev)
        self.group_infos = _tensor(group_infos(coords, coords[:, 0]))
        self.entity_infos = _tensor(entity_infos(S))
        self.motion = torch.zeros(S, dtype=torch.int32, device=dev)
        control = np.zeros(1, layout.CONTROL)
        control["headers"]["root"] = layout.LT_NONE
        self.control = _tensor(control)
        self.rows = torch.zeros(G * 8, dtype=torch.int32, device=dev)
        self.staged = torch.full((G,), layout.CPU_REBUILD | layout.CPU_FRESH, dtype=torch.int32, device=dev)
        self.empty_nodes = torch.zeros(56, dtype=torch.int32, device=dev)
        self.empty_leaves = torch.zeros(16, dtype=torch.int32, device=dev)
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
        sc.encoding = encoding
        sc.materials = self.materials.data_ptr()
        sc.light_codes = self.light_codes.data_ptr()
        sc.emission_scale = 1.0
        sc.scene_count = S
        sc.occupied_box = self.occupied_box.data_ptr()
        sc.group_of_brick = self.group_of_brick.data_ptr()
        sc.light_group_rows = self.rows.data_ptr()
        h = stream_handle()
        check(lib.cx_scene_finalize(ctypes.byref(sc), h), "cx_scene_finalize")
        # the group pass over every scene, the heap grown until it fits (the flags stay for a retry)
        self.pass_scratch = device_bytes(int(lib.cx_light_pass_scratch_bytes(n, G)), dev)
        heap_units = max(1024, n * 8)
        while True:
            self.heaps = [torch.zeros(heap_units * 4, dtype=torch.int32, device=dev) for _ in range(2)]
            self.descriptor = _tensor(self._descriptor(self.control.data_ptr(), self.empty_nodes.data_ptr(),
                                                       self.empty_leaves.data_ptr(), heap_units, shift, encoding))
            self.light_pass = CxLightPass(self.descriptor.data_ptr(), heap_units, self.staged.data_ptr(),
                                          self.pass_scratch.data_ptr(), G, 0)
            check(lib.cx_light_pass(ctypes.byref(sc), ctypes.byref(self.light_pass), h), "cx_light_pass")
            c = to_host(self.control).view(layout.CONTROL)[0]
            if int(c["pass_result"]) == layout.PASS_OK:
                break
            heap_units = int(c["needed_units"]) + 1024
        self.current = int(c["current"]) & 1
        rows = to_host(self.rows).view(layout.ROW)
        per_group = (rows["bricks"] & 0xFFFF).astype(np.int64)
        per_scene = np.bincount(coords[:, 0], weights=per_group, minlength=S).astype(np.int64)
        self.leaf_offset = np.concatenate([[0], np.cumsum(per_scene)]).astype(np.int64)
        self.light_bricks = per_scene
        # one gather, leaves numbered per scene from its base
        total = int(self.leaf_offset[-1])
        leaf_capacity = max(64, (total + 63) // 64 * 64)
        self.leaves = device_bytes((leaf_capacity + 64 * S) * 64, dev)
        gather_state = np.zeros(1, layout.STATE)
        gather_state["building"] = 1
        self.gather_state = _tensor(gather_state)
        self.leaf_base = torch.from_numpy(self.leaf_offset[:-1].astype(np.int32)).to(dev)
        check(lib.cx_light_gather_entities(ctypes.byref(self.light_pass), 0, self.leaves.data_ptr(), leaf_capacity,
                                           self.gather_state.data_ptr(), self.leaf_base.data_ptr(), h),
              "cx_light_gather_entities")
        # per scene: its control (the pass's current buffer, its tree header, tree 1 empty), its tree
        scene_control = np.zeros(S, layout.CONTROL)
        scene_control["current"] = self.current
        scene_control["headers"]["root"] = layout.LT_NONE
        self.scene_controls = _tensor(scene_control)
        self.builders = []
        for i in range(S):
            cap = max(64, (int(per_scene[i]) + 63) // 64 * 64)
            view = self.leaves[int(self.leaf_offset[i]) * 64:(int(self.leaf_offset[i]) + cap) * 64]
            header = self.scene_controls[i * 320 + 64:i * 320 + 128]
            b = LightBvhBuilder(cap, leaves=view, header=header)
            st = np.zeros(1, layout.STATE)
            st["building"] = 1
            st["leaves_found"] = int(per_scene[i])
            b.state.copy_(to_device(st.view(np.uint8), b.state.device))
            self.builders.append(b)
        self.tree_batch = LightBvhBatch(self.builders)
        descs = np.zeros(S, layout.DESCRIPTOR)
        for i in range(S):
            d = self._descriptor(self.scene_controls.data_ptr() + i * 320, self.empty_nodes.data_ptr(),
                                 self.empty_leaves.data_ptr(), heap_units, shift, encoding)[0]
            d["nodes"][0] = self.builders[i].nodes.data_ptr()
            d["leaves"][0] = self.leaves.data_ptr() + int(self.leaf_offset[i]) * 64
            descs[i] = d
        self.scene_descriptors = _tensor(descs)
        views = (CxLightTreeView * S)()
        cams = (CxCamera * S)()
        for i, d in enumerate(self.datas):
            views[i].descriptor = self.scene_descriptors.data_ptr() + i * 128
            views[i].mis = 1 if mis else 0
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

    def _descriptor(self, control: int, empty_nodes: int, empty_leaves: int, heap_units: int, shift: int,
                    encoding: int) -> np.ndarray:
        d = np.zeros(1, layout.DESCRIPTOR)
        d["control"] = control
        d["heap"] = [self.heaps[0].data_ptr(), self.heaps[1].data_ptr()]
        d["group_rows"] = self.rows.data_ptr()
        d["group_infos"] = self.group_infos.data_ptr()
        d["entity_infos"] = self.entity_infos.data_ptr()
        d["entity_motion"] = self.motion.data_ptr()
        d["nodes"] = [empty_nodes, empty_nodes]
        d["leaves"] = [self.leaves.data_ptr() if hasattr(self, "leaves") else empty_leaves, empty_leaves]
        d["light_codes"] = self.light_codes.data_ptr()
        d["tree_headers"] = control + 64
        d["group_shift"] = shift
        d["encoding"] = encoding
        d["group_count"] = self.groups
        return d

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
            ts.lights.rebuild(0, stream)
            ts.trace_frame(cam, bounces, seed, spp, color=self.color[i * p * 4:(i + 1) * p * 4],
                           radiance=self.radiance[i * p * 4:(i + 1) * p * 4], attrib=self.attrib[i * p:(i + 1) * p],
                           counters=self.counters, stream=stream, sky=sky, fog=fog)
