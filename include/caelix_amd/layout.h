/*
 * GPU data layouts of the Caelix hot path on MI355X: the light tree (descriptor, control, headers,
 * build states, leaves, 4-wide nodes, build scratch), the 64-bit-key hash table (16-byte key slots,
 * payload header words, counters) and the brick record (word 0/1 info, micro occupancy, payload).
 *
 * One header serves C, C++ host code and HIP device code. Every struct is byte-identical to the
 * reference's layout (same field order, offsets and sizes; the static asserts below pin them):
 *   light tree      lib/src/light/light_tree_layout.h:13-439
 *   hash table      lib/src/rhi/hash_table_layout.h:13-142
 *   brick record    lib/src/scene/record_layout.h:8-95, shaders/common/caelix_record.hlsl:14-60
 *   group / entity  lib/src/scene/gpu_structs.h:45-108
 * The names follow the reference so that a reader of either side finds the same fields.
 */
#ifndef CAELIX_AMD_LAYOUT_H
#define CAELIX_AMD_LAYOUT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define CX_CONST(type, name, value) constexpr type name = (value)
#define CX_STATIC_ASSERT(cond) static_assert(cond, #cond)
#else
#define CX_CONST(type, name, value) enum { name = (value) }
#define CX_STATIC_ASSERT(cond) _Static_assert(cond, #cond)
#endif

/* ------------------------------------------------------------------------------------------
 * Light tree (lib/src/light/light_tree_layout.h)
 * ------------------------------------------------------------------------------------------ */
#ifdef __cplusplus
constexpr uint32_t kLtNone = 0xFFFFFFFFu;
constexpr uint32_t kLtChildLeafBit = 0x80000000u;
#else
#define kLtNone 0xFFFFFFFFu
#define kLtChildLeafBit 0x80000000u
#endif
CX_CONST(uint32_t, kLtTrees, 2);
CX_CONST(uint32_t, kLtMaxLightBricksPerGroup, 4096);
CX_CONST(uint32_t, kLtRecordFaceCountShift, 18);
CX_CONST(uint32_t, kLtRecordFaceCountMask, 0xFFF);
CX_CONST(uint32_t, kLtRecordStaleBit, 1u << 30);

/* The descriptor (128 bytes): every kernel that uses the tree receives only its address. */
typedef struct LightTreeDescriptor {
    uint64_t control;
    uint64_t heap[2];
    uint64_t group_rows;
    uint64_t group_infos;
    uint64_t entity_infos;
    uint64_t entity_motion;
    uint64_t grids;
    uint64_t nodes[2];
    uint64_t leaves[2];
    uint64_t light_codes;
    uint64_t tree_headers;
    uint32_t group_shift;
    uint32_t encoding;
    uint32_t pad;
    uint32_t group_count;
} LightTreeDescriptor;
CX_STATIC_ASSERT(offsetof(LightTreeDescriptor, nodes) == 64);
CX_STATIC_ASSERT(offsetof(LightTreeDescriptor, leaves) == 80);
CX_STATIC_ASSERT(offsetof(LightTreeDescriptor, tree_headers) == 104);
CX_STATIC_ASSERT(offsetof(LightTreeDescriptor, group_count) == 124);
CX_STATIC_ASSERT(sizeof(LightTreeDescriptor) == 128);

/* A summary: face sums and a box (48 bytes). Leaves, tree headers and binary aggregates start
 * with it. */
typedef struct LightSummary {
    float face[6];
    float lo[3];
    float hi[3];
} LightSummary;
CX_STATIC_ASSERT(offsetof(LightSummary, lo) == 24);
CX_STATIC_ASSERT(offsetof(LightSummary, hi) == 36);
CX_STATIC_ASSERT(sizeof(LightSummary) == 48);

/* The root summary of one BVH (64 bytes). */
typedef struct LightTreeHeader {
    LightSummary summary;
    uint32_t root; /* node index; kLtNone = empty tree */
    uint32_t leaf_count;
    uint32_t node_count;
    uint32_t pad;
} LightTreeHeader;
CX_STATIC_ASSERT(offsetof(LightTreeHeader, root) == 48);
CX_STATIC_ASSERT(offsetof(LightTreeHeader, leaf_count) == 52);
CX_STATIC_ASSERT(offsetof(LightTreeHeader, node_count) == 56);
CX_STATIC_ASSERT(sizeof(LightTreeHeader) == 64);

/* The build state of one BVH (64 bytes). */
typedef struct LightTreeState {
    uint32_t building;       /* 1 while this flush rebuilds the tree (gates every build kernel) */
    uint32_t leaves_found;   /* light bricks of the tree at the last build */
    uint32_t leaves_dropped; /* of those, above the leaf capacity */
    uint32_t builds;         /* builds so far */
    uint32_t node_overflow;  /* 1 when the last build needed more nodes than the capacity */
    /* Not in the reference (its first pad word): nonzero when the last build exceeded a bound of
     * this build's own work lists (never expected: each is sized from the tree's depth bound). The
     * tree is then left empty; unlike node_overflow, a larger node capacity does not help. */
    uint32_t build_error;
    uint32_t pad0[2];
    float bounds[6]; /* lo[3], hi[3] of the leaf box centres */
    uint32_t pad1[2];
} LightTreeState;
CX_STATIC_ASSERT(offsetof(LightTreeState, node_overflow) == 16);
CX_STATIC_ASSERT(offsetof(LightTreeState, bounds) == 32);
CX_STATIC_ASSERT(sizeof(LightTreeState) == 64);

/* The control buffer (320 bytes): the group pass state, then two tree headers and states. */
typedef struct LightControl {
    uint32_t current;
    uint32_t pass_result;
    uint32_t needed_units;
    uint32_t pad;
    uint32_t light_groups;
    uint32_t bricks_dropped;
    uint64_t light_bricks;
    uint64_t light_faces;
    uint32_t passes;
    uint32_t failed_passes;
    uint32_t pass_flags;
    uint32_t records_recomputed;
    uint32_t records_copied;
    uint32_t trees_changed;
    LightTreeHeader headers[2];
    LightTreeState states[2];
} LightControl;
CX_STATIC_ASSERT(offsetof(LightControl, light_bricks) == 24);
CX_STATIC_ASSERT(offsetof(LightControl, trees_changed) == 60);
CX_STATIC_ASSERT(offsetof(LightControl, headers) == 64);
CX_STATIC_ASSERT(offsetof(LightControl, states) == 192);
CX_STATIC_ASSERT(sizeof(LightControl) == 320);

CX_CONST(uint32_t, kLtPassFailed, 0);
CX_CONST(uint32_t, kLtPassOk, 1);
CX_CONST(uint32_t, kLtPassNoChange, 2);

/* LightControl::pass_flags (light_tree_layout.h:208-212): a group was flagged or released; bit 1 + t: one of
 * them belongs to tree t. */
CX_CONST(uint32_t, kLtPassFlagChanged, 1);
CX_CONST(uint32_t, kLtPassFlagTreeShift, 1);

/* The light row of a group, by group index (32 bytes; light_tree_layout.h:214-250). The host never
 * writes a row: its per-group flags go through the staged flags (kLtCpu*), merged by the group pass. */
typedef struct GroupLightRow {
    uint32_t block_offset; /* units of 16 bytes in the current heap buffer */
    uint32_t block_units;  /* 0 = no block */
    uint32_t generation;   /* life << 16 | version */
    uint32_t gpu_flags;    /* kLtGpuLightChanged */
    uint32_t cpu_flags;    /* the merged staged flags (kLtCpu*) */
    uint32_t bricks;       /* light bricks | dropped << 16 */
    uint32_t light_faces;
    uint32_t pad;
} GroupLightRow;
CX_STATIC_ASSERT(offsetof(GroupLightRow, cpu_flags) == 16);
CX_STATIC_ASSERT(sizeof(GroupLightRow) == 32);
CX_CONST(uint32_t, kLtRowBricksMask, 0xFFFF);
CX_CONST(uint32_t, kLtRowDroppedShift, 16);
CX_CONST(uint32_t, kLtGenerationLifeShift, 16);
CX_CONST(uint32_t, kLtGenerationVersionMask, 0xFFFF);
CX_CONST(uint32_t, kLtGpuLightChanged, 1);
CX_CONST(uint32_t, kLtCpuRebuild, 1);  /* rebuild the block (bricks removed) */
CX_CONST(uint32_t, kLtCpuReleased, 2); /* the group index is free: clear the row */
CX_CONST(uint32_t, kLtCpuFresh, 4);    /* a new life of the group index */

/* Group light block in a heap buffer (units of 16 bytes): header, index entries (Morton order of the
 * brick position in the group), records (light_tree_layout.h:252-334). */
CX_CONST(uint32_t, kLtUnitBytes, 16);
CX_CONST(uint32_t, kLtBlockIndexOffset, 16);
typedef struct LightBlockHeader {
    uint32_t record_count; /* index entries = brick light records of the block */
    uint32_t pad[3];
} LightBlockHeader;
CX_STATIC_ASSERT(sizeof(LightBlockHeader) == 16);
CX_CONST(uint32_t, kLtBvhLeafTreeShift, 31);
CX_CONST(uint32_t, kLtBvhLeafIndexMask, 0x7FFFFFFF);
typedef struct LightIndexEntry {
    uint32_t pos;      /* group-local brick position (x | y << shift | z << 2 shift) */
    uint32_t dense_id; /* the brick's id in the group (this build: its record in the brick store) */
    uint32_t record_offset; /* units from the block start */
    uint32_t bvh_leaf;      /* tree << 31 | leaf index; kLtNone = none */
} LightIndexEntry;
CX_STATIC_ASSERT(sizeof(LightIndexEntry) == 16);

/* Header of a brick light record; uint32 entries[entry_count] follow, zero-padded to a unit. */
typedef struct BrickLightRecordHeader {
    float face[6];
    uint32_t box_lo; /* group-local voxels, inclusive (x | y << 10 | z << 20) */
    uint32_t box_hi; /* exclusive */
    uint32_t pos;
    uint32_t entry_count;
    uint32_t counts_lo; /* face counts of +X, -X, +Y (10 bits each) */
    uint32_t counts_hi; /* face counts of -Y, +Z, -Z */
} BrickLightRecordHeader;
CX_STATIC_ASSERT(sizeof(BrickLightRecordHeader) == 48);
CX_CONST(uint32_t, kLtFamilyCountBits, 10);
CX_CONST(uint32_t, kLtFamilyCountMask, 0x3FF);
CX_CONST(uint32_t, kLtFaceEntryVoxelMask, 0x1FF);
CX_CONST(uint32_t, kLtFaceEntryBlockShift, 16);
CX_CONST(uint32_t, kLtVoxelBits, 10);
CX_CONST(uint32_t, kLtVoxelMask, 0x3FF);

/* BVH leaf: one light brick in world space (64 bytes). */
typedef struct LightBvhLeaf {
    LightSummary summary;
    uint32_t group;
    uint32_t entry;
    uint32_t pad;
    uint32_t parent; /* node << 2 | slot */
} LightBvhLeaf;
CX_STATIC_ASSERT(offsetof(LightBvhLeaf, group) == 48);
CX_STATIC_ASSERT(offsetof(LightBvhLeaf, entry) == 52);
CX_STATIC_ASSERT(offsetof(LightBvhLeaf, parent) == 60);
CX_STATIC_ASSERT(sizeof(LightBvhLeaf) == 64);
CX_CONST(uint32_t, kLtParentSlotBits, 2);
CX_CONST(uint32_t, kLtParentSlotMask, 3);

/* 4-wide BVH node, children inline (224 bytes). Slots are filled from 0. */
typedef struct LightBvhNode {
    float child_face[4][6];
    float child_lo[4][3];
    float child_hi[4][3];
    uint32_t child[4]; /* kLtChildLeafBit | leaf, else node index; kLtNone = empty slot */
    uint32_t parent;   /* kLtNone = root */
    uint32_t parent_slot;
    uint32_t pad[2];
} LightBvhNode;
CX_STATIC_ASSERT(offsetof(LightBvhNode, child_lo) == 96);
CX_STATIC_ASSERT(offsetof(LightBvhNode, child_hi) == 144);
CX_STATIC_ASSERT(offsetof(LightBvhNode, child) == 192);
CX_STATIC_ASSERT(offsetof(LightBvhNode, parent) == 208);
CX_STATIC_ASSERT(offsetof(LightBvhNode, parent_slot) == 212);
CX_STATIC_ASSERT(sizeof(LightBvhNode) == 224);

/* Build scratch for a leaf capacity C (a multiple of 64). Array X holds C elements and starts at
 * byte kLtScratchX * C, as in the reference (light_tree_layout.h:392-423): sort keys/values in two
 * ping-pong buffers (sorted result in buffer 0), binary children, parents, arrival counters then
 * depths, even-depth flags then 4-wide indices, binary aggregates. The reference's per-block digit
 * histogram region is replaced here by the onesweep sort's tile state (kLtScratchSortState), whose
 * size is fixed by the tile count. */
CX_CONST(uint32_t, kLtScratchKeys, 0);
CX_CONST(uint32_t, kLtScratchValues, 4);
CX_CONST(uint32_t, kLtScratchSortBufferStride, 8);
CX_CONST(uint32_t, kLtScratchLeft, 16);
CX_CONST(uint32_t, kLtScratchRight, 20);
CX_CONST(uint32_t, kLtScratchParent, 24);
CX_CONST(uint32_t, kLtScratchSortedLeafParent, 28);
CX_CONST(uint32_t, kLtScratchCounter, 32);
CX_CONST(uint32_t, kLtScratchWide, 36);
CX_CONST(uint32_t, kLtScratchAggregates, 40);
CX_CONST(uint32_t, kLtScratchSortState, 88);

/* Node capacity for a leaf capacity (lib/src/light/light_tree.cpp:25-28). */
static inline uint32_t LightNodeCapacityFor(uint32_t leaf_capacity) { return leaf_capacity / 8 * 5 + 64; }
/* The reference caps a tree at 65535 * 64 leaves (lib/src/light/light_tree_record.cpp:24-28). */
CX_CONST(uint32_t, kLtReferenceMaxLeaves, 65535u * 64u);

/* ------------------------------------------------------------------------------------------
 * Hash table (lib/src/rhi/hash_table_layout.h)
 * ------------------------------------------------------------------------------------------ */
#ifdef __cplusplus
constexpr uint32_t kHashNone = 0xFFFFFFFFu;
#else
#define kHashNone 0xFFFFFFFFu
#endif
CX_CONST(uint32_t, kHashProbes, 8);
CX_CONST(uint32_t, kHashKeySlotsPerPayload, 4);

/* Key slot (16 bytes): key_lo 0 = empty; key_lo >= 1 and key_hi != 0 for a real key; key_lo set
 * and key_hi 0 = an insert in progress. */
typedef struct HashKeySlot {
    uint32_t key_lo;
    uint32_t key_hi;
    uint32_t payload;
    uint32_t pad;
} HashKeySlot;
CX_STATIC_ASSERT(sizeof(HashKeySlot) == 16);

CX_CONST(uint32_t, kHashPayloadState, 0);
CX_CONST(uint32_t, kHashPayloadTouched, 4);
CX_CONST(uint32_t, kHashPayloadKey, 8);
CX_CONST(uint32_t, kHashMaxKeyWords, 8);
CX_CONST(uint32_t, kHashStateFree, 0);
CX_CONST(uint32_t, kHashStateUninitialised, 1);
CX_CONST(uint32_t, kHashStateReady, 2);
CX_CONST(uint32_t, kHashStackTop, 0);
CX_CONST(uint32_t, kHashStackEntries, 4);

typedef struct HashCounters {
    uint32_t live;
    uint32_t active;
    uint32_t inserted;
    uint32_t pending;
    uint32_t failed_slot;
    uint32_t failed_payload;
    uint32_t evicted;
    uint32_t rebuild_dropped;
    uint32_t user[8];
} HashCounters;
CX_STATIC_ASSERT(sizeof(HashCounters) == 64);

/* Lookup results (hash_table_layout.h:136-142). */
CX_CONST(uint32_t, kHashHit, 0);
CX_CONST(uint32_t, kHashInserted, 1);
CX_CONST(uint32_t, kHashPending, 2);
CX_CONST(uint32_t, kHashFullWindow, 3);
CX_CONST(uint32_t, kHashNoPayload, 4);
CX_CONST(uint32_t, kHashEmptyEntry, 5);
CX_CONST(uint32_t, kHashMissing, 6);

/* ------------------------------------------------------------------------------------------
 * Brick record (lib/src/scene/record_layout.h, shaders/common/caelix_record.hlsl:14-60).
 * The reference stores a 1096-byte record per brick. The MI355X build keeps the same words but
 * splits them into three arrays indexed by the record (= payload) index, so that the traversal
 * reads only what it needs, coalesced:
 *   info[record]       uint32 x2: word 0 (pos | coarse << 16), word 1 (tight bounds | light)
 *   occupancy[record]  uint64 x8: micro occupancy of coarse cell c (the reference's words 2+2c,
 *                      3+2c as one little-endian 64-bit word)
 *   payload[record]    uint16 x512: block at voxel x + 8y + 64z (the little-endian image of the
 *                      reference's payload words 18..273)
 * ------------------------------------------------------------------------------------------ */
CX_CONST(uint32_t, kRecordInfoWords, 2);
CX_CONST(uint32_t, kRecordOccupancyWords, 16);
CX_CONST(uint32_t, kRecordPayloadWord, 18);
CX_CONST(uint32_t, kRecordWords, 274);
CX_CONST(uint32_t, kRecordPosMask, 0xFFFF);
CX_CONST(uint32_t, kRecordCoarseShift, 16);
CX_CONST(uint32_t, kRecordCoarseMask, 0xFF);
CX_CONST(uint32_t, kRecordBoundsAxisBits, 3);
CX_CONST(uint32_t, kRecordBoundsAxisMask, 7);
CX_CONST(uint32_t, kRecordBoundsMaxShift, 9);
CX_CONST(uint32_t, kRecordBoundsMask, 0x3FFFF);
CX_CONST(uint32_t, kBrickEdge, 8);
CX_CONST(uint32_t, kBrickVoxels, 512);

/* Block encodings (record_layout.h:26-34). */
CX_CONST(uint32_t, kEncodingPacked, 0);
CX_CONST(uint32_t, kEncodingBlockId, 1);
CX_CONST(uint32_t, kPackedOpaqueMask, 0xC000);
CX_CONST(uint32_t, kPackedIdMask, 0xFF);
CX_CONST(uint32_t, kPackedFaceShift, 8);
CX_CONST(uint32_t, kBlockIdOpaqueMask, 0x8000);
CX_CONST(uint32_t, kBlockIdIdMask, 0x1FF);
CX_CONST(uint32_t, kBlockIdFaceShift, 9);

/* GroupInfo / EntityInfo (gpu_structs.h:45-108), used by the leaf gather. */
typedef struct GroupInfo {
    int32_t coord[3];
    uint32_t entity;
    uint32_t slot_row_base;
    uint32_t reverse_base;
    uint32_t aabb_row_base;
    uint32_t high_water;
} GroupInfo;
CX_STATIC_ASSERT(sizeof(GroupInfo) == 32);

typedef struct EntityInfo {
    int32_t grid_min[3];
    uint32_t grid_base;
    int32_t grid_dims[3];
    uint32_t life;
    float cur_rows[3][4];
    float prev_rows[3][4];
} EntityInfo;
CX_STATIC_ASSERT(offsetof(EntityInfo, cur_rows) == 32);
CX_STATIC_ASSERT(sizeof(EntityInfo) == 128);

#endif /* CAELIX_AMD_LAYOUT_H */
