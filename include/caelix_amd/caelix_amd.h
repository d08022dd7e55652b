/*
 * C ABI of the MI355X-native Caelix hot path: light-tree build, light-tree sample/pdf, hash-table
 * brick lookup, brick DDA + NEE trace. Plain pointers and sizes; device pointers are HIP device
 * addresses; `stream` is a hipStream_t (NULL = the default stream). Every entry point is
 * exception-free and returns a CaelixResult code (include/caelix/caelix.h:115-122 in the
 * reference); argument errors log once through the log callback and change nothing, as the
 * reference's scene entry points do (lib/src/scene/scene_light.cpp:18-38,
 * lib/src/abi/caelix.cpp:83-111).
 *
 * Work is stream-ordered and asynchronous: an entry point only enqueues kernels, like the
 * reference's library only records into the host's command buffer (caelix.h:15-18). The caller owns
 * every buffer. Calls on one stream are externally synchronised (caelix.h:19-23).
 */
#ifndef CAELIX_AMD_H
#define CAELIX_AMD_H

#include <stdint.h>

#include "caelix_amd/layout.h"

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define CX_API __attribute__((visibility("default")))
#else
#define CX_API
#endif

/* Result codes: the values of the reference's CaelixResult (caelix.h:115-122); CX_ERROR_HIP takes
 * the place of CAELIX_ERROR_VULKAN. */
typedef enum CxResult {
    CX_OK = 0,
    CX_ERROR_INVALID_ARGUMENT = 1,
    CX_ERROR_UNSUPPORTED = 2,
    CX_ERROR_HIP = 3,
    CX_ERROR_OUT_OF_MEMORY = 4
} CxResult;

/* Log callback (caelix.h:128-139): level 0 debug .. 3 error. NULL disables logging. */
typedef void (*CxLogCallback)(int level, const char* message, void* user_data);
CX_API void cx_set_log_callback(CxLogCallback callback, void* user_data);

/* Library version string and the number of gfx950 kernels built in. */
CX_API const char* cx_version_string(void);

/* ------------------------------------------------------------------------------------------
 * Light-tree build (the BVH half of LightTree::RecordTree, lib/src/light/light_tree_record.cpp:
 * 230-309, after the gather): bounds -> Morton -> radix sort -> Karras -> aggregate -> depth ->
 * scan -> 4-wide collapse -> finish, one tree per call.
 *
 * Inputs in device memory, in the reference's formats (light_tree_layout.h):
 *   leaves[leaf_capacity]  LightBvhLeaf; the first state->leaves_found hold the gathered leaves
 *   state                  LightTreeState; `building` gates the build as caelix_light_bvh_begin
 *                          sets it (0: every kernel returns at once), `leaves_found` as the
 *                          gather's scan wrote it
 * Outputs: header (LightTreeHeader), state (leaves_dropped, bounds, node_overflow, build_error, builds),
 * nodes[node_capacity] (LightBvhNode), leaves[].parent, and the scratch arrays in the reference's
 * scratch layout (sorted keys/values in sort buffer 0, left/right/parent/sorted_leaf_parent,
 * depths in `counter`, scanned 4-wide indices in `wide`, binary aggregates).
 * ------------------------------------------------------------------------------------------ */
typedef struct CxLightBvhBuild {
    LightBvhLeaf* leaves;
    LightBvhNode* nodes;
    LightTreeHeader* header;
    LightTreeState* state;
    void* scratch;          /* cx_light_bvh_scratch_bytes(leaf_capacity) bytes */
    uint32_t leaf_capacity; /* a multiple of 64, >= 64 */
    uint32_t node_capacity;
} CxLightBvhBuild;

/* Scratch bytes for a leaf capacity: the reference's arrays (88 bytes per leaf,
 * LightBvhScratchBytes, light_tree_layout.h:434-439) plus the onesweep sort's tile state. */
CX_API uint64_t cx_light_bvh_scratch_bytes(uint32_t leaf_capacity);
/* Prepares freshly allocated scratch (zeroes the build's control words). Call once per
 * allocation, before the first build; builds leave the control words ready for the next one. */
CX_API int cx_light_bvh_scratch_init(void* scratch, uint32_t leaf_capacity, void* stream);
/* caelix_light_bvh_begin (shaders/light/bvh/caelix_light_bvh_begin.hlsl:9-17): sets
 * state->building from `reason` and the committed group pass recorded in `control`
 * (LightControl, may be NULL when pass_recorded is 0). */
CX_API int cx_light_bvh_begin(LightControl* control, LightTreeState* state, uint32_t tree, uint32_t reason,
                              uint32_t pass_recorded, void* stream);
/* The build (light_tree_record.cpp:271-308). */
CX_API int cx_light_bvh_build(const CxLightBvhBuild* build, void* stream);

/* The build without a caller scratch buffer, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in the
 * reference. `build->scratch` must be NULL: the build uses the library's static scratch of 11 MiB per
 * GPU (held until the process ends), so leaf_capacity is at most 65,536; a larger one is rejected with
 * CX_ERROR_INVALID_ARGUMENT. It allocates nothing and launches the kernels of cx_light_bvh_build (one
 * more to clear the control words), with the same results. Every call on one GPU shares that scratch:
 * calls that overlap on one GPU (other streams, other threads) race; the scratch keeps the last
 * call's contents. Stage timing (cx_profile_enable) does not time it. */
CX_API int cx_light_bvh_build_static(const CxLightBvhBuild* build, void* stream);

/* The batched form of the build: `count` builds (each its own leaves, nodes, header, state,
 * scratch and capacity, as for cx_light_bvh_build) in one sequence of launches whose grid covers
 * every build (grid y = build), so many small trees (chunk scenes) cost one build's launches.
 * cx_light_bvh_batch_prepare copies the build descriptions to `descriptors` (device memory of
 * cx_light_bvh_batch_bytes(count) bytes); cx_light_bvh_build_batch records only kernel launches (it
 * can be captured in a hipGraph) for `max_leaf_capacity` = the largest capacity of the builds. Each
 * tree is the one cx_light_bvh_build makes. No counterpart in the reference (one RecordTree per
 * tree, light_tree_record.cpp:230-309). */
CX_API uint64_t cx_light_bvh_batch_bytes(uint32_t count);
CX_API int cx_light_bvh_batch_prepare(const CxLightBvhBuild* builds, uint32_t count, void* descriptors, void* stream);
CX_API int cx_light_bvh_build_batch(const void* descriptors, uint32_t count, uint32_t max_leaf_capacity, void* stream);

/* Tunables of the 32-bit key builds' sort (cx_light_bvh_build, _static, _batch and the light tree
 * adapter; process-wide, read when a build is recorded): digit_bits = 6 (5 passes over the 30-bit
 * Morton codes) or 8 (4 passes, the default), tile = 4096 or 8192 (the default) elements per
 * workgroup. The order and the tree are the same for every setting (the passes are stable); only the
 * speed changes. The Python package's autotuner picks them per workload. No counterpart in the
 * reference (its sort is fixed). */
typedef struct CxLightBvhTuning {
    uint32_t digit_bits;
    uint32_t tile;
} CxLightBvhTuning;
CX_API int cx_light_bvh_set_tuning(const CxLightBvhTuning* tuning);
CX_API void cx_light_bvh_get_tuning(CxLightBvhTuning* tuning);

/* The build with 64-bit sort keys: code << 32 | leaf index, sorted as one key by
 * cx_radix_sort_keys64's passes (7 x 9 bits) in place of the (code, leaf) pair sort. The order, and
 * so every output (tree, nodes, leaves, header, scratch arrays), is the one cx_light_bvh_build makes.
 * No counterpart in the reference, whose keys are 32-bit (light_tree_record.cpp:271-308); it is the
 * wide-key form for callers whose keys do not fit 32 bits. */
CX_API int cx_light_bvh_build_keys64(const CxLightBvhBuild* build, void* stream);

/* Stable LSD radix sort of (uint32 key, uint32 value) pairs (the reference's
 * caelix_light_bvh_sort_histogram + caelix_light_bvh_scan + caelix_light_bvh_sort_scatter passes,
 * shaders/light/bvh/caelix_light_bvh_sort.hlsl:1-38), as 4 onesweep passes of 8 bits. The count is
 * read from device memory (count_address, like LightSortPush); gate_address (nullable) skips the
 * sort when the u32 there is 0. Result in keys0/values0; keys1/values1 are the ping-pong buffers.
 * `sort_state` holds cx_radix_sort_state_bytes(capacity) bytes; capacity >= the count. */
CX_API uint64_t cx_radix_sort_state_bytes(uint32_t capacity);
CX_API int cx_radix_sort_pairs(uint32_t* keys0, uint32_t* values0, uint32_t* keys1, uint32_t* values1,
                               const uint32_t* count_address, const uint32_t* gate_address, uint32_t capacity,
                               void* sort_state, void* stream);
/* Stable LSD radix sort of uint64 keys alone (bits 0..62: 7 onesweep passes of 9 bits; keys with
 * bit 63 set sort as if it were clear), count and gate as for cx_radix_sort_pairs. The result is in
 * keys1 (odd pass count); keys0 is overwritten. `sort_state` holds
 * cx_radix_sort_keys64_state_bytes(capacity) bytes. No counterpart in the reference (32-bit keys,
 * caelix_light_bvh_sort.hlsl:1-38). */
CX_API uint64_t cx_radix_sort_keys64_state_bytes(uint32_t capacity);
CX_API int cx_radix_sort_keys64(uint64_t* keys0, uint64_t* keys1, const uint32_t* count_address,
                                const uint32_t* gate_address, uint32_t capacity, void* sort_state, void* stream);

/* ------------------------------------------------------------------------------------------
 * Hash table (lib/src/rhi/hash_table.{h,cpp}, shaders/common/caelix_hash_table.hlsl,
 * shaders/hash/ (every kernel)). The table maps 64-bit key fingerprints of 1..8 key words to payload indices
 * in 16-byte HashKeySlot slots. The home slot of a key is key_lo mod slot_count (the reference's
 * key_lo & key_mask for a power-of-two slot count); a lookup probes probe_limit slots from home
 * (kHashProbes = 8 in the reference's form).
 * ------------------------------------------------------------------------------------------ */
typedef struct CxHashTable {
    void* keys;     /* HashKeySlot[slot_count] */
    void* stack;    /* free stack: int top, uint entries[payload capacity] (lookups that insert) */
    void* counters; /* HashCounters (lookups that insert, maintenance) */
    void* payloads; /* the user's pool: payload i at payloads + i * payload_stride */
    uint32_t slot_count;
    uint32_t payload_stride;
    uint32_t probe_limit;
    uint32_t pad;
} CxHashTable;

/* HashFingerprint (caelix_hash_table.hlsl:113-128) of `count` keys of `word_count` words each:
 * out[2i] = key_lo, out[2i+1] = key_hi. */
CX_API int cx_hash_fingerprint(const uint32_t* words, uint32_t word_count, uint32_t count, uint32_t* out,
                               void* stream);
/* One thread per key: HashLookup (insert 0) or HashLookupOrInsert (insert 1) with the frame counter
 * (caelix_hash_table.hlsl:217-230); results[i] = kHash* code, payloads[i] = payload or kHashNone.
 * Inserts count into the counters (HashCountResult). */
CX_API int cx_hash_lookup(const CxHashTable* table, const uint32_t* words, uint32_t word_count, uint32_t count,
                          uint32_t frame, uint32_t insert, uint32_t* results, uint32_t* payloads, void* stream);
/* HashTable::RecordMaintenance (hash_table.cpp:79-130): [init on the first call] begin, key fill,
 * maintain (evict payloads with frame - touched > max_age, rebuild the keys of the live ones, active
 * list), finish (dispatch = {x, y, 1, active count}). `active` holds payload_capacity u32,
 * `dispatch` 4 u32. */
CX_API int cx_hash_maintain(const CxHashTable* table, uint32_t payload_capacity, uint32_t key_words, uint32_t frame,
                            uint32_t max_age, uint32_t items_per_group, uint32_t initialise, uint32_t* active,
                            uint32_t* dispatch, void* stream);

/* Bulk brick table: every key (3 int32 brick coordinates) of `keys_xyz` inserted at once into
 * `slots` (any slot_count > count, e.g. count / 0.7 for load 0.7), payload = the key's index. The
 * layout is deterministic: ordered linear probing, as a sequential insert in (home slot, key index)
 * order places the keys. info[0] = probe limit (longest displacement + 1), info[1] = longest
 * displacement, info[2] = keys wrapped past the last slot. No counterpart in the reference, whose
 * keys are inserted one lookup at a time (HashLookupOrInsert); the slot format and probing are its. */
CX_API uint64_t cx_hash_build_scratch_bytes(uint32_t count);
/* The same bulk build over keys of `word_count` uint32 words each (1..8; the fingerprint of
 * caelix_hash_table.hlsl:113-128): the batched scenes key their bricks by (scene, x, y, z). */
CX_API int cx_hash_build_keys(const uint32_t* words, uint32_t word_count, uint32_t count, HashKeySlot* slots,
                              uint32_t slot_count, void* scratch, uint32_t* info, void* stream);
CX_API int cx_hash_build_bricks(const int32_t* keys_xyz, uint32_t count, HashKeySlot* slots, uint32_t slot_count,
                                void* scratch, uint32_t* info, void* stream);
/* Lookups in a bulk-built table: payloads[i] = the brick's index or kHashNone; probes[i] (nullable) =
 * slots read. */
CX_API int cx_hash_get_bricks(const HashKeySlot* slots, uint32_t slot_count, const uint32_t* info,
                              const int32_t* keys_xyz, uint32_t count, uint32_t* payloads, uint32_t* probes,
                              void* stream);

/* ------------------------------------------------------------------------------------------
 * Scene: resident bricks in the split record format of layout.h (info words, micro occupancy,
 * payload; record index = the brick's index in `keys`), found through a bulk-built hash table, and
 * bitmaps of resident bricks / 8^3-brick super cells over a grid of grid_dim^3 bricks from
 * grid_min. One entity at the identity frame; groups of (1 << group_shift)^3 bricks. Materials: 2
 * float4 per 16-bit block value, (albedo rgb, smoothness) and (emission rgb, metallic) (the first 16
 * bytes of the reference's FP16 hot entry, caelix_material.hlsl:1-20, 84-100, as float), looked up by
 * material code (a transparent value without its face bits); materials_ext optionally the rest,
 * (ior, extinction, 0, 0); light_codes: 65536 bits, the block values whose emission luminance
 * reaches the light threshold.
 * ------------------------------------------------------------------------------------------ */
typedef struct CxScene {
    const int32_t* keys;      /* [bricks][3] */
    uint32_t* info;           /* [bricks][2] record words 0 and 1 (written by cx_scene_finalize) */
    void* occupancy;          /* [bricks][8] uint64 micro occupancy (written by cx_scene_finalize) */
    uint16_t* payload;        /* [bricks][512] block values, voxel x + 8y + 64z, 16-byte aligned;
                               * cx_scene_finalize writes the face bits of transparent and air voxels */
    const HashKeySlot* slots; /* cx_hash_build_bricks over keys */
    uint32_t slot_count;
    uint32_t bricks;
    const uint32_t* hash_info; /* cx_hash_build_bricks info (probe limit) */
    uint32_t* brick_bits;      /* grid_dim^3 bits */
    uint32_t* super_bits;      /* (grid_dim / 8)^3 bits */
    int32_t grid_min[3];
    uint32_t grid_dim;         /* a multiple of 8 */
    uint32_t group_shift;
    uint32_t encoding;         /* kEncodingPacked or kEncodingBlockId */
    const float* materials;    /* [65536][8] */
    const uint32_t* light_codes;
    float emission_scale;
    uint32_t pad;
    /* [bricks] per-brick emission scale (the emitters' parameters of the differentiable form), or NULL
     * for 1. Multiplies the emission the path tracer adds; the light tree samples with the unscaled
     * emission. */
    const float* brick_emission_scale;
    /* A batch of chunk scenes when > 1: keys are [bricks][4] = (scene, x, y, z), the hash table is
     * keyed by those 4 words (cx_hash_build_keys), the bitmaps hold scene_count grids of grid_dim^3
     * (and (grid_dim / 8)^3) bits, scene-major. 0 or 1: one scene, keys [bricks][3]. */
    uint32_t scene_count;
    uint32_t pad2;
    /* [max(scene_count, 1)][6], or NULL: per scene the box of super cells (lo x, y, z, hi x, y, z,
     * inclusive) that hold a brick with a voxel, written by cx_scene_finalize. The traversal then ends
     * a ray where it leaves that box widened by one super cell (no resident brick lies outside; hits
     * and counts are unchanged). NULL: every ray is walked to the end of the grid. */
    int32_t* occupied_box;
    /* [65536][4] (ior, extinction, 0, 0) per material code, or NULL. Set: cx_trace_frame shades by the
     * reference's full rule (ShadeVoxelHit, caelix_shading.hlsl:267-307): the specular lobe of opaque
     * materials (smoothness, metallic, ior), transparent blocks as interfaces between media (reflect or
     * refract by Fresnel) whose extinction attenuates the segments inside them, no NEE at specular
     * vertices. NULL: every block is opaque Lambertian (the benchmarked subset). */
    const float* materials_ext;
    /* [bricks] the brick's group index (its GroupInfo and GroupLightRow), written by the light tree's
     * host side: the light pass groups the light bricks by it and the trace finds a hit light face
     * through it (LightTreeFindLightFace). */
    const uint32_t* group_of_brick;
    /* [groups] GroupLightRow, or NULL (FinalizePush::light_group_rows, caelix_finalize.hlsl:52): set, the
     * finalize keeps the light state of record word 1 (bits 18..30): the open faces of the brick's light
     * voxels (kLtRecordFaceCountShift) and the stale bit, set with kLtGpuLightChanged in the group's row
     * when the old or the new count is not 0 (caelix_finalize.hlsl:22-29, 326-340). New bricks start with
     * word 1 = 0. NULL: those bits are 0. */
    GroupLightRow* light_group_rows;
    /* [65536] CxSubsurface per material code, or NULL (no SSS material): read by cx_trace_frame with
     * CX_TRACE_SUBSURFACE. */
    const struct CxSubsurface* subsurface;
} CxScene;

/* The subsurface entry of a material code (the reference's CaelixMaterial sss_* fields as its GPU
 * subsurface entry holds them, docs/MATERIALS.md "Subsurface entry" w0..w3, the halves widened to
 * float): an SSS voxel is an opaque block whose code has model CX_SSS_MODEL_HG, probability > 0 and
 * steps > 0; the walk's per-step albedo is 1 - one_minus_albedo. 32 bytes. */
typedef struct CxSubsurface {
    float one_minus_albedo[3]; /* 1 - sss_albedo (the entry's w0, w1 halves) */
    float anisotropy;          /* sss_anisotropy g in [-1, 1] (the walk clamps it to [-0.99, 0.99]) */
    float probability;         /* sss_probability p in [0, 1] */
    uint32_t model;            /* CX_SSS_MODEL_NONE or CX_SSS_MODEL_HG; other values act as none */
    uint32_t steps;            /* sss_steps, 0..255 (0: never entered) */
    uint32_t pad;
} CxSubsurface;
#define CX_SSS_MODEL_NONE 0u
#define CX_SSS_MODEL_HG 1u

/* The record producer (shaders/scene/caelix_finalize.hlsl:112-361) over the bricks and their face
 * neighbours (found through the hash table: slots and hash_info are read): face bits of transparent and
 * air voxels written into the payload, word 0 (group position, coarse byte), word 1 bounds and the
 * micro occupancy of the occupied voxels (alive: fewer than six opaque neighbours; not 0; opaque or a
 * face bit); and the resident bitmaps; with light_group_rows, the light state of word 1. Running it
 * again changes nothing but the stale bits and row flags of bricks with light faces. */
CX_API int cx_scene_finalize(const CxScene* scene, void* stream);
/* The finalize of the listed records only (the reference finalizes the records of a flush's uploads and
 * their face neighbours, caelix_finalize.hlsl:1-29): the same words, the resident bitmaps and the box only
 * gain bits (a brick emptied keeps its bits until the next cx_scene_finalize). A record whose
 * group_of_brick is kLtNone is a free record: it gets no light state. No counterpart name in the
 * reference (its finalize takes the list in its push constants). */
CX_API int cx_scene_finalize_bricks(const CxScene* scene, const uint32_t* records, uint32_t count, void* stream);

/* The group pass (shaders/light/records/caelix_light_tree_{size,scan,build,commit}.hlsl), over the tables
 * the descriptor names (control, heap[2], group rows, group infos, entity infos and motion words, light
 * codes), as lib/src/light/light_tree_record.cpp:198-228 records it. The staged flags are merged into the
 * rows; a group is released (kLtCpuReleased), rebuilt (kLtCpuRebuild, or the finalize's
 * kLtGpuLightChanged) or copied. When a group is released or rebuilt, every group's light block is
 * written into the heap buffer that is not current, if all blocks fit heap_units: a copied group's block
 * as it is; a rebuilt group's header, index entries in Morton order of the brick position in the group
 * and brick light records of its light bricks (record word 1 counts light faces; at most 4096 per group,
 * the lowest Morton positions; the rest counted as dropped), each record copied from the old block when
 * the brick is not stale and had an entry there, else recomputed; then rows (generation: the next version
 * of a rebuilt group, the next life of a released one), statistics and the current buffer are committed
 * (kLtPassOk), the stale bits of the rebuilt groups' bricks cleared, and trees_changed names the trees of
 * the changed groups' entities. A pass that does not fit changes nothing but pass_result (kLtPassFailed),
 * needed_units and the counters; the flags stay for the next pass. Bricks are the scene's, their groups
 * group_of_brick (the descriptor's group_shift must be the scene's). Groups: at most 2^32 / 8^group_shift. */
typedef struct CxLightPass {
    LightTreeDescriptor* descriptor; /* device address; the descriptor and its tables in device memory */
    uint64_t heap_units;             /* capacity of each heap buffer, in units of 16 bytes */
    uint32_t* staged_flags;          /* [group_count] kLtCpu* flags staged by the host; cleared by the pass */
    void* scratch;                   /* cx_light_pass_scratch_bytes(bricks, group_count) bytes */
    uint32_t group_count;
    uint32_t pad;
} CxLightPass;
CX_API uint64_t cx_light_pass_scratch_bytes(uint32_t bricks, uint32_t groups);
CX_API int cx_light_pass(const CxScene* scene, const CxLightPass* pass, void* stream);

/* The gather of tree `tree` (caelix_light_gather_count.hlsl, caelix_light_bvh_scan.hlsl,
 * caelix_light_gather_emit.hlsl), gated by its state's `building` (cx_light_bvh_begin): per group whose
 * entity has the tree's motion, its index entries become leaves at the group's offset (an exclusive scan
 * of the counts over the group indices; state->leaves_found = the total), each the record's box and face
 * sums moved to world space by the entity's frame (LtWorldSummary: the AABB of the rotated box, projected
 * face sums), group and entry = the brick's group and index entry; the index entry's bvh_leaf =
 * tree << 31 | leaf, or kLtNone past leaf_capacity. Then cx_light_bvh_build builds the tree over the
 * leaves. `pass` names the descriptor and the scratch of the last cx_light_pass. */
CX_API int cx_light_gather(const CxLightPass* pass, uint32_t tree, LightBvhLeaf* leaves, uint32_t leaf_capacity,
                           LightTreeState* state, void* stream);
/* The gather for many trees at once, one per entity (chunk scenes: an entity each, its groups
 * consecutive): as cx_light_gather, but the bvh_leaf words count leaves from entity_leaf_base[entity] (the
 * leaf index of the entity's first leaf), so that each entity's leaves form a tree of their own over the
 * slice leaves + entity_leaf_base[entity]. No counterpart in the reference (two trees per descriptor). */
CX_API int cx_light_gather_entities(const CxLightPass* pass, uint32_t tree, LightBvhLeaf* leaves,
                                    uint32_t leaf_capacity, LightTreeState* state, const uint32_t* entity_leaf_base,
                                    void* stream);

/* The light tree as the trace and the light probes read it: the device address of its 128-byte
 * descriptor (LightTreeDescriptor, the reference's LightTreeOpen reads it the same way:
 * shaders/light/caelix_light_tree.hlsl:239-287), whether the NEE and the BSDF hits weigh each other
 * by the balance heuristic (the reference's light NEE without CAELIX_LIGHT_NEE_FLAG_NO_MIS), and the
 * trees the draws and pdfs use (LightTreeSelect, :266-272: kLtSelect* 1 static, 2 dynamic, 3 both; 0
 * is taken as both, LightTreeOpen's default). */
typedef struct CxLightTreeView {
    const LightTreeDescriptor* descriptor;
    uint32_t mis;
    uint32_t select;
} CxLightTreeView;

/* Pinhole camera (caelix_frame.hlsl:164-171): unit forward/right/up. */
typedef struct CxCamera {
    float position[3];
    float forward[3];
    float right[3];
    float up[3];
    float tan_half_fov_y;
    float aspect;
    uint32_t width;
    uint32_t height;
} CxCamera;

struct CxLightHint;

struct CxSkyCubemap;

typedef struct CxTraceParams {
    uint32_t bounces;    /* bounce_count */
    uint32_t frame_seed;
    uint32_t spp;        /* samples per pixel; sample s uses frame_seed + s */
    uint32_t flags;      /* CX_TRACE_* */
    /* Sky and sun, read with CX_TRACE_SKY (the frame constants of shaders/common/caelix_frame.hlsl:
     * sun_dir, sun_radiance, sun_cos_radius, sun_solid_angle, sky_zenith/horizon/ground). */
    float sun_dir[3];        /* unit vector toward the sun */
    float sun_solid_angle;   /* 2 pi (1 - cos r) of the sun cone (computed in double by the host) */
    float sun_radiance[3];   /* radiance inside the sun disk, >= 0; all 0: no sun (its NEE off) */
    float sun_cos_radius;    /* cos r */
    float sky_zenith[3];
    float pad0;
    float sky_horizon[3];
    float pad1;
    float sky_ground[3];
    float pad2;
    /* Fog of the air (the frame constant fog_color: rgb radiance, w density; caelix_shading.hlsl:
     * 150-175): fog_density > 0 turns it on. Every segment in the air that ends on a surface (the
     * camera ray, every bounce segment) keeps exp(-density t) of its throughput and gains
     * fog_color (1 - exp(-density t)); misses and shadow rays see no fog. Not supported by
     * cx_trace_frame_adjoint. */
    float fog_color[3];
    float fog_density;
    /* The delta chain of the camera ray (bounce_count_transparent, 0..15; caelix_path_trace.hlsl:
     * 595-678): it crosses up to this many smooth transparent interfaces (smoothness >= 0.99) by
     * reflection or refraction; the first other hit is the primary surface, everything it sends times
     * the chain's transmission, depth the chain's length. Needs CxScene::materials_ext. */
    uint32_t bounces_transparent;
    /* The bounce NEE budget (bounce_nee_probability, caelix_path_trace.hlsl:909-944): a bounce vertex
     * runs its NEE with this probability (one number of the sun stream) and scales it by its inverse;
     * when it does not, its segment still carries the state of a vertex that ran it. 0 or 1: always
     * (no number drawn). In [0, 1]. Not supported by cx_trace_frame_adjoint below 1. */
    float bounce_nee_probability;
    /* The light NEE's visibility hint (below); NULL: off (the choice weights unchanged). */
    const struct CxLightHint* hint;
    /* With CX_TRACE_SKY: the host's cubemap in place of the procedural sky (below); NULL: procedural. */
    const struct CxSkyCubemap* sky_cubemap;
} CxTraceParams;

/* The host's sky cubemap (CaelixRenderOptions::sky, caelix.h:1009-1017) in device memory: mip_levels
 * mips (1 .. log2(size) + 1), each 6 faces (+X, -X, +Y, -Y, +Z, -Z) of (size >> m)^2 texels, rows from
 * the top of the face, rgba float32 (a ignored), mip-major. The frame samples it as a linear
 * mipmapped sampler does (the Vulkan face selection; bilinear inside a face, clamped to its edge
 * texels; linear between mips): primary misses and the delta chain's sky end at mip 0, a specular
 * bounce that misses at mip roughness * (mip_levels - 1), a diffuse bounce that misses its irradiance
 * SH / pi at the vertex normal (caelix_shading.hlsl:373-447); each frame first projects the SH
 * (caelix_sky_sh.hlsl) from the mip whose face is about 32 texels. The sun (disk and NEE) stays as
 * CxTraceParams gives it. */
typedef struct CxSkyCubemap {
    const float* texels;
    uint32_t size;       /* face width = height of mip 0 */
    uint32_t mip_levels;
} CxSkyCubemap;

/* The light NEE's visibility hint per pixel patch (kLightNeeHint; caelix_path_trace.hlsl:229-236,
 * 338-344, 453-580): the primary vertex multiplies its sun and lamp choice weights by max(0.05, hint)
 * of its patch in the previous frame, read at the previous pixel of its primary surface (the previous
 * camera's nearest pixel; its own patch without a surface, a previous camera, or when the point was
 * behind it or outside its image); after the frame each patch moves its hint by 0.2 toward the fraction
 * of its primary shadow rays (sun, lamp) that reached their light, and keeps a channel without one.
 * Only the choice changes: the estimate stays unbiased. Words: sun | lamp << 8, each k / 255. With
 * spp > 1 every sample reads `previous` and the update pools the samples' shadow rays (the value read
 * is the patch's first pixel's, sample 0). Not supported by cx_trace_frame_adjoint or the batch. */
typedef struct CxLightHint {
    const uint16_t* previous; /* [(height + patch - 1) >> patch_shift][(width + patch - 1) >> patch_shift] */
    uint16_t* current;        /* the same shape, every patch written; not `previous` */
    uint16_t* read;           /* nullable: [height][width], the word each primary vertex read (sample 0) */
    const CxCamera* previous_camera; /* NULL: none; else the same image size as the camera */
    uint32_t patch_shift;     /* 0, 1 or 2: patches of 1, 2 or 4 pixels square */
    uint32_t reset;           /* nonzero: no history (kLightNeeHintReset): every vertex reads 0x8080 */
} CxLightHint;

/* The procedural sky on misses and the sun (caelix_shading.hlsl:364-403, 453-485;
 * caelix_path_trace.hlsl:276-290, 356-413, 811-846): a primary miss shows SkyRadiance (the sky plus
 * the sun disk; the color output, radiance 0); a bounce ray that misses adds throughput x
 * ProceduralSky, plus the sun disk when the sun's NEE is off; with the sun's NEE every NEE vertex
 * draws a sun candidate next to the lamp candidate and keeps one in proportion to its unoccluded
 * luminance (one-sample RIS, one shadow ray). Without it the sky is black and there is no sun. Not
 * supported by cx_trace_frame_adjoint (CX_ERROR_UNSUPPORTED). */
#define CX_TRACE_SKY 2u

/* With CX_TRACE_SKY: no sun NEE (the reference's kRenderFlagSunNee off); bounce rays that reach the
 * sun disk see its radiance instead. */
#define CX_TRACE_NO_SUN_NEE 4u

/* Transmissive shadows at the primary vertex (kRenderFlagShadowTransmission; caelix_scene_trace.hlsl:
 * 401-477, caelix_path_trace.hlsl:795-930): its shadow ray starts in the vertex's medium and goes
 * straight through transparent voxels (at most 8 hops), attenuated by each medium's extinction (no
 * fog), stopped by an opaque voxel or a transmittance below 1e-3; that NEE claims the light it saw:
 * after a transparent vertex, up to the first opaque one, segments see no sun disk (with sun NEE) and
 * count a lamp the tree can sample beyond the near field as 0. Implies the full material model. */
#define CX_TRACE_SHADOW_TRANSMISSION 8u

/* The delta checkerboard (kRenderFlagDeltaCheckerboard; caelix_path_trace.hlsl:608-660, the split mode
 * of caelix_shading.hlsl:178-243): the delta chain's first interface takes the branch of the pixel's
 * parity (px + py even refracts, odd reflects) with the weight 2 (1 - F) or 2 F and draws no random
 * number; total internal reflection still reflects with weight 1. */
#define CX_TRACE_DELTA_CHECKERBOARD 16u

/* Subsurface scattering (kRenderFlagSubsurface; shaders/render/caelix_subsurface.hlsl, docs/SUBSURFACE.md):
 * with CxScene::subsurface set, a path vertex on an SSS voxel (the primary surface, or a bounce hit that
 * has a next segment) draws u from the BSDF stream: u >= p keeps the surface shade; else a
 * Henyey-Greenstein lattice walk through the SSS voxels (entry jitter on the face, per step the voxel's
 * albedo and a new direction, then the face neighbour along its dominant axis: an SSS voxel is entered,
 * another opaque voxel is a wall, anything else is the exit) replaces the vertex by a diffuse vertex on
 * the exit face with the walk's weight as its albedo; a walk that reaches the entry material's step cap
 * ends the path (dead). The primary vertex's albedo output is then lerp(albedo, sss_albedo, p) and the
 * branch weight over it its throughput. Implies the full material model. */
#define CX_TRACE_SUBSURFACE 32u

/* No Russian roulette: paths run all `bounces` (the roulette's random number is still drawn, so the
 * other decisions stay the same). The estimator is then a smooth function of the albedos and the
 * emission scales at fixed random numbers, which central finite differences can check. */
#define CX_TRACE_NO_RUSSIAN_ROULETTE 1u

/* Per-pixel outputs (nullable except counters): color = emission + albedo * radiance (rgba, a =
 * primary hit t), radiance = TracePath's demodulated radiance, accumulated over the samples;
 * counters: [0] rays (primary + bounce), [1] shadow rays, [2] brick cells visited, [3] hash probes,
 * [4] brick DDA steps (u64 each). */
typedef struct CxTraceOutput {
    float* color;
    float* radiance;
    uint32_t* attrib; /* primary hit attribute (0 = miss) */
    uint64_t* counters;
    void* workspace;          /* device memory of cx_trace_workspace_bytes(width, height) bytes */
    uint64_t workspace_bytes; /* its size */
} CxTraceOutput;

/* Device workspace of cx_trace_frame (and cx_trace_frame_adjoint) for a width x height image: the
 * path state, the path queues, the shadow rays and the adjoint's gradient rows of the wavefront
 * (about 230 bytes per pixel, 8x8-tile padded; 0.44 GiB at 1920 x 1080). */
CX_API uint64_t cx_trace_workspace_bytes(uint32_t width, uint32_t height);

/* One frame: primary ray, brick DDA, light NEE + MIS, `bounces` bounces (at most 255), for every
 * pixel (shaders/render/caelix_path_trace.hlsl:1050-, TracePath :737-978), opaque Lambertian subset,
 * as a wavefront of trace / shade / shadow launches over compacted path queues. */
CX_API int cx_trace_frame(const CxScene* scene, const CxLightTreeView* lights, const CxCamera* camera,
                          const CxTraceParams* params, const CxTraceOutput* out, void* stream);

/* The adjoint of cx_trace_frame (a path-replay adjoint, the differentiable form): for the image
 * I = color / spp of the same scene, camera and params, and d_color = dL/dI (rgba per pixel, a
 * ignored), adds dL/d(albedo) and dL/d(emission scale) to the gradient buffers. The replay traces
 * the same rays and draws the same lights from the same random streams as the forward frame (the
 * decisions are recomputed, not stored; the path state between the wavefront's launches is the
 * per-bounce checkpoint); sampling pdfs, MIS weights and the roulette's normalisation are held
 * constant (detached). Albedo parameters are the material entries of the opaque blocks 0x8000 | id,
 * id < 256: grad_albedo[id] (rgb + pad, 256 float4); emission scales are per brick record:
 * grad_scale[record] (scene->bricks floats). The caller zeroes both; the adds are float atomics.
 * No counterpart in the reference. */
typedef struct CxTraceGradients {
    const float* d_color; /* [height][width][4] */
    float* grad_albedo;   /* [256][4] */
    float* grad_scale;    /* [bricks] */
} CxTraceGradients;
CX_API int cx_trace_frame_adjoint(const CxScene* scene, const CxLightTreeView* lights, const CxCamera* camera,
                                  const CxTraceParams* params, const CxTraceGradients* grads,
                                  const CxTraceOutput* out, void* stream);

/* Tunables of cx_trace_frame / cx_trace_batch (process-wide; set between frames): queue_chunk = the
 * ray-queue depth a wave claims per refill of the persistent traversal (default 64), and
 * workgroups_per_cu = its persistent workgroups per CU (0 = as many as fit, the default). The
 * autotuner of the Python package picks them per workload. They change speed only, never results. */
typedef struct CxTraceTuning {
    uint32_t queue_chunk;
    uint32_t workgroups_per_cu;
} CxTraceTuning;
CX_API int cx_trace_set_tuning(const CxTraceTuning* tuning);
CX_API void cx_trace_get_tuning(CxTraceTuning* tuning);

/* The batched form of the frame: `scenes` chunk scenes of one batched CxScene (scene_count =
 * scenes), each with its own light tree view and camera (one image size), traced as one wavefront:
 * the paths of every scene go through the same launches (scene s owns paths [s * P, (s + 1) * P),
 * P = the 8x8-tile-padded pixel count), so the whole batch costs one frame's launches. Outputs are
 * the per-scene images, scene-major (width * height pixels each). cx_trace_batch_prepare copies the
 * views and cameras to `device_views` (cx_trace_batch_bytes(scenes) bytes of device memory);
 * cx_trace_batch records only kernel launches (it can be captured in a hipGraph). The workspace is
 * cx_trace_batch_workspace_bytes(width, height, scenes) bytes. Each scene's image is the one
 * cx_trace_frame makes of it alone. No counterpart in the reference. */
CX_API uint64_t cx_trace_batch_bytes(uint32_t scenes);
CX_API uint64_t cx_trace_batch_workspace_bytes(uint32_t width, uint32_t height, uint32_t scenes);
CX_API int cx_trace_batch_prepare(const CxLightTreeView* lights, const CxCamera* cameras, uint32_t scenes,
                                  void* device_views, void* stream);
CX_API int cx_trace_batch(const CxScene* scene, const void* device_views, uint32_t scenes, uint32_t width,
                          uint32_t height, const CxTraceParams* params, const CxTraceOutput* out, void* stream);

/* Rays through the world (TraceScene / TraceShadow semantics): rays[i] = origin(3), direction(3),
 * t_max; hits[i] = {t (float bits), attrib, record, brick x, y, z, cells, probes, steps}. */
CX_API int cx_trace_rays(const CxScene* scene, const float* rays, uint32_t count, uint32_t* hits, void* stream);

/* LightTreeSample (caelix_light_tree.hlsl:749-767) for receivers[i] = x(3), n(3) with PCG state
 * seeds[i]; samples[i] = {position(3), normal(3), pdf, block, face, group, entry, voxel in group} (12
 * words: floats, then block, face, the brick (group, index entry) and the group-local voxel
 * x | y << 10 | z << 20 as bits); ok[i] = 1 when a point was drawn. */
CX_API int cx_light_sample(const CxLightTreeView* lights, const float* receivers, const uint32_t* seeds,
                           uint32_t count, float* samples, uint32_t* ok, void* stream);
/* LightTreePointPdf (caelix_light_tree.hlsl:936-949): the area pdf with which LightTreeSample draws
 * the point on a light face, for a receiver: queries[i] = {x(3), n(3) as floats; group, group-local
 * voxel x, y, z, face, raw block value as u32} (12 words); pdf[i] (0 when the face is not a light face
 * of the tree or its brick has no leaf). */
CX_API int cx_light_point_pdf(const CxLightTreeView* lights, const uint32_t* queries, uint32_t count, float* pdf,
                              void* stream);

/* The region receiver's draw (LightTreeSampleBrickBox, caelix_light_tree.hlsl:787-818): a brick drawn for
 * every receiver point of a box with a normal, by the box importance (LightTreeImportanceBox, :162-166) at
 * every choice. boxes[i] = lo(3), hi(3), n(3) (9 floats; n = 0: no surface), seeds[i] the PCG state;
 * out[i] = {ok, group, entry, generation, leaf (tree << 31 | index), pdf_brick (float bits), importance
 * (float bits: the leaf summary's box importance, 0 when the draw fell back to the power), 0}. */
CX_API int cx_light_sample_brick_box(const CxLightTreeView* lights, const float* boxes, const uint32_t* seeds,
                                     uint32_t count, uint32_t* out, void* stream);
/* LightTreeBrickPdf (box = 0, :769-777) or LightTreeBrickPdfBox (box = 1, :822-831) of a brick reference:
 * queries[i] = lo(3), hi(3), n(3) as floats, then group, entry, generation (12 words; a point receiver is
 * lo); pdf[i] = the probability of drawing that brick, 0 when the reference is stale (LightBrickRefValid,
 * :640-646) or the brick has no leaf. */
CX_API int cx_light_brick_pdf(const CxLightTreeView* lights, const uint32_t* queries, uint32_t count, uint32_t box,
                              float* pdf, void* stream);
/* The support query of the reference's light NEE probe (tests/shaders/caelix_light_nee_probe.hlsl mode 0):
 * queries[i] = lo(3), x.x, hi(3), x.y, n(3), x.z as floats, then group, voxel in group (x | y << 10 |
 * z << 20), face, block (16 words; x a point in front of the face); out[i] = (found, has leaf, box pdf,
 * target) as floats: LightTreeFindLightFace from x, LightTreeBrickPdfBox and LightTreeLeafImportanceBox
 * of the box. */
CX_API int cx_light_box_support(const CxLightTreeView* lights, const uint32_t* queries, uint32_t count, float* out,
                                void* stream);

/* The cubemap sky as the frame reads it: queries[i] = direction(3), lod; out[i] = rgb (SampleLevel
 * of caelix_shading.hlsl:375-386 with a linear mipmapped sampler). */
CX_API int cx_sky_cube_sample(const CxSkyCubemap* cube, const float* queries, uint32_t count, float* out,
                              void* stream);
/* The irradiance SH of the cubemap that each frame projects first (shaders/render/caelix_sky_sh.hlsl):
 * sh = 9 float4 (rgb, 0) of device memory, E(n) = sum sh[i] Y_i(n) (caelix_shading.hlsl:346-419). */
CX_API int cx_sky_sh(const CxSkyCubemap* cube, float* sh, void* stream);

/* LightNeeMisWeight (shaders/render/caelix_light_nee.hlsl:20-34), the balance heuristic both MIS sides
 * share, per query: queries[i] = {p_area, d, cos_l, receiver normal (3), unit direction (3)} (9
 * floats); weights[i] = {NEE side, BSDF side}. A non-finite or degenerate solid-angle density counts
 * as infinite (NEE 1, BSDF 0); both densities 0 give NEE 0, BSDF 1. */
CX_API int cx_light_mis_weight(const float* queries, uint32_t count, float* weights, void* stream);

/* ------------------------------------------------------------------------------------------
 * The light tree's host side (the reference's LightTree, lib/src/light/light_tree.h:37-101,
 * light_tree.cpp, light_tree_record.cpp): the CPU state mirrored from the scene (entity motions,
 * group lives and the flags staged for the group pass), the light code set, and the device tables
 * it owns: control, group rows, staged flags, entity motion words, the two heap buffers, leaves and
 * nodes of the static and the dynamic tree, the build scratch, the 128-byte descriptor. The scene
 * calls the hooks when its entities and groups change and, per flush, the steps in this order:
 *
 *   cx_light_tree_prepare           CPU work and every allocation; *requeue = 1 when the set of light
 *                                   codes changed (the scene finalizes every resident brick again)
 *   cx_light_tree_stage             the staged group flags and entity motion words, for the next
 *                                   transfers; *staged = 0 when nothing of the tree is recorded
 *   cx_light_tree_record_transfers  growth copies, zero fills of new tables, the descriptor, the code
 *                                   set and the staged words (PrepareTransfers + RecordTransfers)
 *   cx_scene_finalize               the scene's finalize with light_group_rows = the tree's rows
 *   cx_light_tree_record            the group pass (cx_light_pass), then per tree that changed the
 *                                   begin, gather and build (cx_light_bvh_begin, cx_light_gather,
 *                                   cx_light_bvh_build), and a copy of the control to the host
 *   cx_light_tree_end_flush         closes the flush
 *
 * cx_light_tree_recycle(safe_frame) reads the copies of frames <= safe_frame (the caller knows those
 * frames are done on the GPU), frees replaced buffers and schedules heap and BVH growth: a pass that
 * did not fit the heap is retried at the next flush after the heap grew (x2, or 1.5 x the need), a
 * tree whose leaves passed 3/4 of its capacity (or dropped leaves) grows and is rebuilt, a node
 * overflow grows the node capacity to the leaf capacity. Calls are externally synchronised; all work
 * is enqueued on the stream given to the recording steps. No GPU memory is allocated outside prepare
 * and record_transfers (record allocates the pass scratch for a larger brick count).
 * ------------------------------------------------------------------------------------------ */

/* CaelixLightOptions (caelix.h:629-641): min_emission 0 = 20; heap_bytes 0 = 16 MiB, rounded up to 16
 * bytes, at most 2^32 x 16 bytes. `enabled` is not read here (a disabled tree is one not created). */
typedef struct CxLightOptions {
    uint32_t enabled;
    float min_emission;
    uint64_t heap_bytes;
} CxLightOptions;

/* CaelixLightStats (caelix.h:648-680), 104 bytes: the newest light update read back. */
typedef struct CxLightStats {
    uint32_t enabled;
    uint32_t light_groups;
    uint64_t light_bricks;
    uint64_t light_faces;
    uint64_t bricks_dropped;
    uint64_t heap_bytes_used;
    uint64_t heap_bytes_capacity;
    uint32_t passes;
    uint32_t failed_passes;
    uint32_t records_recomputed;
    uint32_t records_copied;
    float gpu_ms;
    uint32_t pad;
    uint32_t tree_leaves[2];
    uint32_t tree_nodes[2];
    uint32_t tree_builds[2];
    uint32_t tree_leaves_dropped[2];
} CxLightStats;

/* What the tree reads of the scene at a flush (LightSceneTables, light_tree.h:22-35): the device
 * tables of groups and entities (GroupInfo, EntityInfo) and their capacities, and the brick capacity
 * (the group pass's scratch). */
typedef struct CxLightSceneTables {
    const GroupInfo* group_infos;
    const EntityInfo* entity_infos;
    uint32_t group_capacity;  /* group indices in use or free: [0, group_capacity) */
    uint32_t entity_capacity; /* entity ids: [0, entity_capacity) */
    uint32_t brick_capacity;  /* bricks of the CxScene the pass reads */
    uint32_t pad;
} CxLightSceneTables;

typedef struct CxLightTree CxLightTree; /* opaque */

/* Entity motions (CAELIX_ENTITY_STATIC / DYNAMIC): the tree that holds an entity's lights. */
#define CX_ENTITY_STATIC 0u
#define CX_ENTITY_DYNAMIC 1u

/* LightTree::LightTree (light_tree.cpp:41-86): group_shift 1..7, encoding kEncoding*. */
CX_API int cx_light_tree_create(uint32_t group_shift, uint32_t encoding, const CxLightOptions* options,
                                CxLightTree** out);
/* Waits for the device and frees everything. */
CX_API void cx_light_tree_destroy(CxLightTree* tree);
/* The scene hooks (light_tree.cpp:113-178). */
CX_API int cx_light_tree_on_entity_added(CxLightTree* tree, uint32_t entity, uint32_t motion);
CX_API int cx_light_tree_on_entity_removed(CxLightTree* tree, uint32_t entity);
CX_API int cx_light_tree_on_entity_moved(CxLightTree* tree, uint32_t entity);
CX_API int cx_light_tree_on_entity_motion(CxLightTree* tree, uint32_t entity, uint32_t motion);
CX_API int cx_light_tree_on_group_created(CxLightTree* tree, uint32_t group);
CX_API int cx_light_tree_on_group_released(CxLightTree* tree, uint32_t group, uint32_t motion);
CX_API int cx_light_tree_on_group_bricks_removed(CxLightTree* tree, uint32_t group);
CX_API int cx_light_tree_on_materials_changed(CxLightTree* tree);
CX_API int cx_light_tree_set_min_emission(CxLightTree* tree, float min_emission);
/* The flush steps (light_tree.cpp:198-346, light_tree_record.cpp:102-404). `materials` is the host
 * copy of the scene's table ([65536][8] floats, as CxScene::materials): a code is a light code when the
 * luminance of its canonical code's emission (BlockMaterialCode) reaches min_emission. */
CX_API int cx_light_tree_prepare(CxLightTree* tree, const float* materials, const CxLightSceneTables* tables,
                                 uint32_t* requeue);
CX_API int cx_light_tree_stage(CxLightTree* tree, uint32_t finalize_ran, uint32_t* staged);
CX_API int cx_light_tree_record_transfers(CxLightTree* tree, void* stream);
CX_API int cx_light_tree_record(CxLightTree* tree, const CxScene* scene, uint64_t frame, void* stream);
CX_API int cx_light_tree_end_flush(CxLightTree* tree, uint64_t frame);
/* The gather and build of tree `which` now, from the current blocks (RecordTree with a reason,
 * light_tree_record.cpp:230-309), outside a flush: only kernel launches (it can be captured in a
 * hipGraph), for a caller that rebuilds a tree every frame. After a flush that recorded the tree. No
 * counterpart in the reference's C ABI. */
CX_API int cx_light_tree_rebuild(CxLightTree* tree, uint32_t which, void* stream);
CX_API int cx_light_tree_recycle(CxLightTree* tree, uint64_t safe_frame);
/* LightTree::GetStats (light_tree_record.cpp:406-428); a NULL tree reports enabled 0. */
CX_API void cx_light_tree_get_stats(const CxLightTree* tree, CxLightStats* out);
/* 1 while the tree has work for a later flush or results not yet read back: a pass to retry, growth
 * to make, a tree to rebuild, flags staged by the hooks, a readback not yet recycled. No counterpart in
 * the reference (its flushes run every frame); a caller that flushes only on change loops on it. */
CX_API uint32_t cx_light_tree_pending(const CxLightTree* tree);
/* Device addresses: the descriptor (caelix_scene_get_light_tree, caelix.h:682-688; valid for work
 * enqueued after the next record_transfers), the light code set and the group rows (the finalize's
 * light_group_rows). */
CX_API const LightTreeDescriptor* cx_light_tree_descriptor(const CxLightTree* tree);
CX_API const uint32_t* cx_light_tree_light_codes(const CxLightTree* tree);
CX_API GroupLightRow* cx_light_tree_group_rows(const CxLightTree* tree);

/* ------------------------------------------------------------------------------------------
 * .cxw world files (format version 8 and older): the reader of the bricks the scene build takes.
 * Replaces the reference's Rust reader host-rs/caelix-cxw/src/lib.rs:1-656 (CxwFile::open,
 * ::world, ::sectors, ::decode, ::decode_all; Error variants lib.rs:69-109; read rules of
 * parse_index lib.rs:314-413, region lib.rs:415-452, decode_payload lib.rs:515-656). Host code:
 * the file is memory-mapped, sectors are inflated with zlib (raw deflate) on host threads.
 * Format errors return CX_ERROR_INVALID_ARGUMENT and fill *error (kind = a CxCxwErrorKind, with
 * the fields of the matching reference variant); a NULL error pointer is allowed.
 * ------------------------------------------------------------------------------------------ */
typedef struct CxCxw CxCxw; /* an open file (opaque) */

typedef enum CxCxwErrorKind {
    CX_CXW_NONE = 0,
    CX_CXW_IO = 1,
    CX_CXW_BAD_MAGIC = 2,           /* actual = the magic */
    CX_CXW_UNSUPPORTED_VERSION = 3, /* actual = the version */
    CX_CXW_BAD_FLAGS = 4,           /* actual = the flags */
    CX_CXW_TRUNCATED = 5,           /* what, offset (a payload offset when what starts "payload") */
    CX_CXW_BAD_REGION_SIZE = 6,     /* coord, expected, actual */
    CX_CXW_BAD_SECTOR_MAGIC = 7,    /* actual = the magic */
    CX_CXW_BAD_COUNTS = 8,          /* expected = capacity, actual = count, offset = free count */
    CX_CXW_INFLATE = 9,
    CX_CXW_DUPLICATE_SECTOR = 10,   /* coord */
    CX_CXW_BAD_PAYLOAD = 11         /* coord, what = the reason */
} CxCxwErrorKind;

typedef struct CxCxwError {
    int32_t kind;
    int32_t coord[3];
    uint64_t offset;
    uint64_t expected;
    uint64_t actual;
    char what[64];
    char message[192];
} CxCxwError;

typedef struct CxCxwWorld {
    uint16_t version;
    uint8_t has_player_pose; /* version 7+ world record with a pose */
    uint8_t pad;
    float pose_position[3];
    float pose_rotation[4]; /* quaternion x, y, z, w */
    uint32_t entity_count;
    uint64_t sector_count;  /* every sector of every entity */
} CxCxwWorld;

typedef struct CxCxwEntity { /* EntityInfo (lib.rs:111-133), after the version migrations */
    uint8_t guid[16];
    float position[3];
    float rotation[4];
    float velocity[6];      /* linear xyz, angular xyz; zero before version 4 */
    uint32_t sector_count;
    uint16_t require_update_flags; /* raw, as stored */
    uint8_t flags;          /* bit 0 HasBody, bit 1 Static */
    uint8_t is_protected;
} CxCxwEntity;

typedef struct CxCxwSector { /* SectorRef (lib.rs:185-192) */
    uint32_t entity;
    int32_t coord[3];
    uint64_t region_offset;
    uint32_t region_size;
    uint32_t pad;
} CxCxwSector;

CX_API int cx_cxw_open(const char* path, CxCxw** out, CxCxwError* error);
CX_API void cx_cxw_close(CxCxw* file);
CX_API int cx_cxw_world(const CxCxw* file, CxCxwWorld* out);
CX_API int cx_cxw_entity(const CxCxw* file, uint32_t index, CxCxwEntity* out);
/* Sectors in entity order, then sector index order. */
CX_API int cx_cxw_sector(const CxCxw* file, uint64_t index, CxCxwSector* out);
/* One sector's non-air bricks in ascending brick index order: keys[3 * i] = entity-local brick key
 * (sector coord * 16 + brick in sector), blocks[512 * i + x + 8y + 64z] = raw value, summaries[i] =
 * the preview word; buffers of 4096 bricks. A brick whose 512 values are all 0 is dropped. */
CX_API int cx_cxw_decode(const CxCxw* file, uint64_t sector, int32_t* keys, uint16_t* blocks, uint32_t* summaries,
                         uint32_t* brick_count, uint64_t* voxels, CxCxwError* error);
/* Decodes every sector on `threads` host threads (0 = one per core) and keeps the results in the
 * handle; stops at the first error. Then cx_cxw_copy_decoded writes all bricks in sector order
 * (entity_of_brick and summaries may be NULL). */
CX_API int cx_cxw_decode_all(CxCxw* file, uint32_t threads, uint64_t* brick_count, uint64_t* voxels,
                             CxCxwError* error);
CX_API int cx_cxw_copy_decoded(const CxCxw* file, int32_t* keys, uint16_t* blocks, uint32_t* summaries,
                               uint32_t* entity_of_brick, uint64_t capacity_bricks);

/* ------------------------------------------------------------------------------------------
 * Stage timing: when enabled, entry points bracket every kernel with HIP events on the stream it
 * runs on. cx_profile_read synchronises those events and returns per-stage sums (milliseconds)
 * and launch counts since the last reset; names are static strings.
 * ------------------------------------------------------------------------------------------ */
CX_API void cx_profile_enable(int enabled);
CX_API void cx_profile_reset(void);
CX_API int cx_profile_read(uint32_t max_stages, const char** names, double* total_ms, uint32_t* launches,
                           uint32_t* stage_count);

#ifdef __cplusplus
}
#endif

#endif /* CAELIX_AMD_H */
