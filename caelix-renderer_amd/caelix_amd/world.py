"""An editable scene of entities with a light tree: the scene side of the light tree's seam.

The reference's Scene owns bricks per entity (uploads, removals), render groups of (1 << group_shift)^3
bricks per entity, entities with a frame and a motion, and calls the light tree's hooks where they
change (lib/src/scene/scene_bricks.cpp:428, scene_entities.cpp:426-745, scene_flush.cpp:551-891):
OnEntityAdded / Removed / Moved / Motion, OnGroupCreated when a group gets its first brick,
OnGroupBricksRemoved when it loses bricks, OnGroupReleased when it loses its last one; and per flush the
light tree's steps around the finalize of the uploaded bricks and their face neighbours. LightWorld does
that over this library: a brick store of stable records (a removed brick's record is free and may be
taken again), one hash table over the records in use, group indices reused after release (a new life of
the index), entity infos (frame rows), and the adapter (cx_light_tree_*).

Entities share one key space in the store: entity e's brick (x, y, z) is stored at key
(x + e * REGION, y, z), which keeps the entities' bricks apart (no face neighbours across entities); its
group keeps the entity-local coordinate (GroupInfo::coord) and the entity's frame maps entity-local
voxels to the world, as the reference's. Brick coordinates must lie in [0, REGION) along x and in
[0, grid) along y and z.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import (CxLightSceneTables, CxLightTreeView, CxScene, LightTree, _torch, check, entity_infos, layout, load,
               stream_handle)

REGION = 64  # bricks along x per entity region (a multiple of every group size)


class WorldState:
    """The host bookkeeping of an editable scene (no device): stable brick records, group indices with
    their lives, entities with frames and motions; every change calls `hook(name, *args)` with the
    light tree hook the reference's Scene calls there (LightTree::On*, light_tree.h:60-71)."""

    def __init__(self, group_shift: int = 4, capacity: int = 4096, entities: int = 4, grid: int = 64, hook=None):
        self.shift = int(group_shift)
        self.capacity = int(capacity)
        self.max_entities = int(entities)
        self.grid = int(grid)
        if REGION % (1 << self.shift) or self.grid % 8:
            raise ValueError("the region must hold whole groups and the grid be a multiple of 8")
        n = self.capacity
        self.hook = hook if hook is not None else (lambda name, *args: None)
        self.keys = np.zeros((n, 3), np.int32)
        self.keys[:, 0] = -1 - np.arange(n)  # free records: keys no lookup asks for
        self.keys[:, 1:] = -1
        self.payload = np.zeros((n, 512), np.uint16)
        self.group_of_brick = np.full(n, layout.LT_NONE, np.uint32)
        self.free_records = list(range(n - 1, -1, -1))
        self.record_of = {}  # (entity, x, y, z) -> record
        self.groups = {}  # (entity, gx, gy, gz) -> group index
        self.group_key = []  # group index -> (entity, gx, gy, gz), or None when free
        self.group_bricks = []
        self.free_groups = []
        self.entity_rows = np.tile(np.eye(3, 4, dtype=np.float32), (self.max_entities, 1, 1))
        self.entity_motion = np.zeros(self.max_entities, np.uint32)
        self.entity_alive = np.zeros(self.max_entities, bool)
        self.dirty = set()  # records to finalize at the next flush
        self.fresh_records = set()  # records taken or freed since the last flush: word 1 starts at 0
        self.grid_dim = max(self.grid, (REGION * self.max_entities + 7) // 8 * 8)

    # ---- entities -------------------------------------------------------------------------------
    def add_entity(self, motion: int = 0, rows=None) -> int:
        e = int(np.nonzero(~self.entity_alive)[0][0])
        self.entity_alive[e] = True
        self.entity_motion[e] = motion
        self.entity_rows[e] = np.eye(3, 4, dtype=np.float32) if rows is None else np.asarray(rows, np.float32)
        self.hook("on_entity_added", e, int(motion))
        return e

    def set_transform(self, entity: int, rows) -> None:
        self.entity_rows[entity] = np.asarray(rows, np.float32)
        self.hook("on_entity_moved", entity)

    def set_motion(self, entity: int, motion: int) -> None:
        self.entity_motion[entity] = motion
        self.hook("on_entity_motion", entity, int(motion))

    def remove_entity(self, entity: int) -> None:
        keys = [k[1:] for k in self.record_of if k[0] == entity]
        self.remove(entity, keys)
        self.hook("on_entity_removed", entity)
        self.entity_alive[entity] = False

    # ---- bricks ---------------------------------------------------------------------------------
    def _neighbours(self, entity: int, key) -> list[int]:
        out = []
        for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            r = self.record_of.get((entity, key[0] + d[0], key[1] + d[1], key[2] + d[2]))
            if r is not None:
                out.append(r)
        return out

    def upload(self, entity: int, keys, payloads) -> None:
        """New bricks or new contents of resident ones (a resident record keeps its word 1)."""
        for key, payload in zip(keys, payloads):
            key = tuple(int(v) for v in key)
            if not (0 <= key[0] < REGION and 0 <= key[1] < self.grid and 0 <= key[2] < self.grid):
                raise ValueError(f"brick {key} outside the entity's region")
            r = self.record_of.get((entity,) + key)
            if r is None:
                r = self.free_records.pop()
                self.record_of[(entity,) + key] = r
                self.keys[r] = (key[0] + entity * REGION, key[1], key[2])
                gk = (entity, key[0] >> self.shift, key[1] >> self.shift, key[2] >> self.shift)
                g = self.groups.get(gk)
                if g is None:
                    g = self.free_groups.pop() if self.free_groups else len(self.group_key)
                    if g == len(self.group_key):
                        self.group_key.append(None)
                        self.group_bricks.append(0)
                    self.groups[gk] = g
                    self.group_key[g] = gk
                    self.group_bricks[g] = 0
                    self.hook("on_group_created", g)
                self.group_bricks[g] += 1
                self.group_of_brick[r] = g
                self.fresh_records.add(r)
            self.payload[r] = np.asarray(payload, np.uint16)
            self.dirty.add(r)
            self.dirty.update(self._neighbours(entity, key))

    def remove(self, entity: int, keys) -> None:
        for key in keys:
            key = tuple(int(v) for v in key)
            r = self.record_of.pop((entity,) + key)
            g = int(self.group_of_brick[r])
            self.group_bricks[g] -= 1
            self.group_of_brick[r] = layout.LT_NONE
            self.payload[r] = 0
            self.keys[r] = (-1 - r, -1, -1)
            self.free_records.append(r)
            self.fresh_records.add(r)
            self.dirty.discard(r)
            self.dirty.update(self._neighbours(entity, key))
            if self.group_bricks[g] == 0:
                gk = self.group_key[g]
                del self.groups[gk]
                self.group_key[g] = None
                self.free_groups.append(g)
                self.hook("on_group_released", g, int(self.entity_motion[entity]))
            else:
                self.hook("on_group_bricks_removed", g)

    def group_coords(self) -> np.ndarray:
        """[groups, 4] (entity, gx, gy, gz) in entity-local group coordinates (zeros for a free index)."""
        out = np.zeros((len(self.group_key), 4), np.int32)
        for g, k in enumerate(self.group_key):
            if k is not None:
                out[g] = k
        return out

    def take_dirty(self):
        """(records to finalize, records whose word 1 starts at 0), both cleared."""
        records = sorted(self.dirty)
        fresh = sorted(self.fresh_records)
        self.dirty = set()
        self.fresh_records = set()
        return records, fresh


class LightWorld(WorldState):
    """WorldState on the GPU: the brick store, the hash table, the finalize and the light tree adapter."""

    def __init__(self, materials: np.ndarray, group_shift: int = 4, encoding: int = 1, capacity: int = 4096,
                 entities: int = 4, grid: int = 64, min_emission: float = 0.0, heap_bytes: int = 0, mis: bool = True):
        super().__init__(group_shift, capacity, entities, grid)
        torch = _torch()
        self.lib = load()
        self.encoding = int(encoding)
        self.materials = np.ascontiguousarray(materials, np.float32).reshape(65536, 8)
        self.lights = LightTree(self.shift, self.encoding, min_emission, heap_bytes)
        self.hook = self.lights.hook
        dev = "cuda"
        n = self.capacity
        self.d_keys = torch.zeros(n * 3, dtype=torch.int32, device=dev)
        self.d_payload = torch.zeros(n * 512, dtype=torch.int16, device=dev)
        self.d_info = torch.zeros(n * 2, dtype=torch.int32, device=dev)
        self.d_occupancy = torch.zeros(n * 8, dtype=torch.int64, device=dev)
        self.d_group_of_brick = torch.zeros(n, dtype=torch.int32, device=dev)
        self.slot_count = int(np.ceil(n / 0.7)) + 1
        self.d_slots = torch.zeros(self.slot_count * 4, dtype=torch.int32, device=dev)
        self.d_hash_info = torch.zeros(4, dtype=torch.int32, device=dev)
        self.d_hash_scratch = torch.zeros(int(self.lib.cx_hash_build_scratch_bytes(n)), dtype=torch.uint8, device=dev)
        dim = self.grid_dim
        self.d_brick_bits = torch.zeros((dim ** 3 + 31) // 32, dtype=torch.int32, device=dev)
        self.d_super_bits = torch.zeros(((dim // 8) ** 3 + 31) // 32 + 1, dtype=torch.int32, device=dev)
        self.d_materials = torch.from_numpy(self.materials.reshape(-1).copy()).to(dev)
        self.d_group_infos = torch.zeros(32, dtype=torch.uint8, device=dev)
        self.d_entity_infos = torch.zeros(self.max_entities * 128, dtype=torch.uint8, device=dev)
        sc = self.scene = CxScene()
        sc.keys = self.d_keys.data_ptr()
        sc.info = self.d_info.data_ptr()
        sc.occupancy = self.d_occupancy.data_ptr()
        sc.payload = self.d_payload.data_ptr()
        sc.slots = self.d_slots.data_ptr()
        sc.slot_count = self.slot_count
        sc.bricks = n
        sc.hash_info = self.d_hash_info.data_ptr()
        sc.brick_bits = self.d_brick_bits.data_ptr()
        sc.super_bits = self.d_super_bits.data_ptr()
        sc.grid_min[:] = [0, 0, 0]
        sc.grid_dim = dim
        sc.group_shift = self.shift
        sc.encoding = self.encoding
        sc.materials = self.d_materials.data_ptr()
        sc.emission_scale = 1.0
        sc.group_of_brick = self.d_group_of_brick.data_ptr()
        self.view = CxLightTreeView()
        self.view.mis = 1 if mis else 0
        self.full_finalize = True

    # ---- flush ----------------------------------------------------------------------------------
    def group_coords(self) -> np.ndarray:
        out = np.zeros((len(self.group_key), 4), np.int32)
        for g, k in enumerate(self.group_key):
            if k is not None:
                out[g] = k
        return out

    def set_min_emission(self, min_emission: float) -> None:
        """The light threshold (caelix_scene_set_light_options): a new code set finalizes every record."""
        self.lights.hook("set_min_emission", ctypes.c_float(min_emission))

    def group_infos(self) -> np.ndarray:
        from . import group_infos as make

        c = self.group_coords()
        return make(c[:, 1:], c[:, 0])

    def set_min_emission(self, min_emission: float) -> None:
        """The light threshold (caelix_scene_set_light_options): a new code set finalizes every record."""
        self.lights.hook("set_min_emission", ctypes.c_float(min_emission))

    def group_infos(self) -> np.ndarray:
        from . import group_infos as make

        c = self.group_coords()
        return make(c[:, 1:], c[:, 0])

    def flush(self, stream=None) -> int:
        """Uploads the changed records, finalizes them and their face neighbours (every record on the first
        flush), and runs the light tree's flushes until nothing is pending; returns the flushes."""
        torch = _torch()
        lib = self.lib
        h = stream_handle(stream)
        records, fresh = self.take_dirty()
        changed = sorted(set(records) | set(fresh))
        if self.full_finalize:
            changed = list(range(self.capacity))
        if changed:
            idx = torch.from_numpy(np.array(changed, np.int64)).cuda()
            self.d_keys.view(-1, 3)[idx] = torch.from_numpy(self.keys[changed]).cuda()
            self.d_payload.view(-1, 512)[idx] = torch.from_numpy(self.payload[changed].view(np.int16)).cuda()
            self.d_group_of_brick[idx] = torch.from_numpy(self.group_of_brick[changed].view(np.int32)).cuda()
        if fresh:
            self.d_info.view(-1, 2)[torch.from_numpy(np.array(fresh, np.int64)).cuda(), 1] = 0
        check(lib.cx_hash_build_bricks(self.d_keys.data_ptr(), self.capacity, self.d_slots.data_ptr(), self.slot_count,
                                       self.d_hash_scratch.data_ptr(), self.d_hash_info.data_ptr(), h),
              "cx_hash_build_bricks")
        infos = self.group_infos()
        if len(infos):
            self.d_group_infos = torch.from_numpy(infos.view(np.uint8).copy()).cuda()
        self.d_entity_infos.copy_(torch.from_numpy(entity_infos(self.max_entities, self.entity_rows).view(np.uint8)))
        self.tables = CxLightSceneTables(self.d_group_infos.data_ptr(), self.d_entity_infos.data_ptr(),
                                         len(self.group_key), self.max_entities, self.capacity)
        lt = self.lights
        flushes = 0
        for i in range(16):
            lt.frame += 1
            requeue = ctypes.c_uint32(0)
            check(lib.cx_light_tree_prepare(lt.handle, self.materials.ctypes.data, ctypes.byref(self.tables),
                                            ctypes.byref(requeue)), "cx_light_tree_prepare")
            full = self.full_finalize or requeue.value != 0
            finalize = i == 0 and (full or bool(records))
            staged = ctypes.c_uint32(0)
            check(lib.cx_light_tree_stage(lt.handle, 1 if finalize else 0, ctypes.byref(staged)), "cx_light_tree_stage")
            check(lib.cx_light_tree_record_transfers(lt.handle, h), "cx_light_tree_record_transfers")
            self.scene.light_codes = lt.light_codes
            self.scene.light_group_rows = lt.group_rows
            if finalize and full:
                check(lib.cx_scene_finalize(ctypes.byref(self.scene), h), "cx_scene_finalize")
                self.full_finalize = False
            elif finalize:
                rec = torch.from_numpy(np.array(records, np.int32)).cuda()
                check(lib.cx_scene_finalize_bricks(ctypes.byref(self.scene), rec.data_ptr(), len(records), h),
                      "cx_scene_finalize_bricks")
                torch.cuda.synchronize()
            check(lib.cx_light_tree_record(lt.handle, ctypes.byref(self.scene), lt.frame, h), "cx_light_tree_record")
            check(lib.cx_light_tree_end_flush(lt.handle, lt.frame), "cx_light_tree_end_flush")
            lt.recycle()
            flushes += 1
            if not lt.pending():
                break
        self.view.descriptor = lt.descriptor
        return flushes

    def stats(self) -> dict:
        return self.lights.stats()

    def tables_read(self) -> dict:
        from . import light_tables

        _torch().cuda.synchronize()
        return light_tables(self.lights.descriptor, len(self.group_key), self.max_entities)
