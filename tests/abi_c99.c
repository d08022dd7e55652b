/* Compile check: the public header is valid C99 with no warnings (gcc -std=c99 -pedantic -Wall -Wextra
   -Werror). Touches every entry point of the ABI and checks the byte layouts a C host binds (C99 has no
   static_assert: a negative array size fails the compile). Not run. */
#include "caelix_amd/caelix_amd.h"

#include <stddef.h>

typedef void (*AnyFn)(void);

size_t cx_abi_c99_touch(void);

/* The reference's layouts (light_tree_layout.h, caelix.h:629-680). */
typedef char descriptor_is_128_bytes[(sizeof(LightTreeDescriptor) == 128) ? 1 : -1];
typedef char control_is_320_bytes[(sizeof(LightControl) == 320) ? 1 : -1];
typedef char control_headers_at_64[(offsetof(LightControl, headers) == 64) ? 1 : -1];
typedef char control_states_at_192[(offsetof(LightControl, states) == 192) ? 1 : -1];
typedef char row_is_32_bytes[(sizeof(GroupLightRow) == 32) ? 1 : -1];
typedef char index_entry_is_16_bytes[(sizeof(LightIndexEntry) == 16) ? 1 : -1];
typedef char record_header_is_48_bytes[(sizeof(BrickLightRecordHeader) == 48) ? 1 : -1];
typedef char leaf_is_64_bytes[(sizeof(LightBvhLeaf) == 64) ? 1 : -1];
typedef char node_is_224_bytes[(sizeof(LightBvhNode) == 224) ? 1 : -1];
typedef char group_info_is_32_bytes[(sizeof(GroupInfo) == 32) ? 1 : -1];
typedef char entity_info_is_128_bytes[(sizeof(EntityInfo) == 128) ? 1 : -1];
typedef char light_stats_is_104_bytes[(sizeof(CxLightStats) == 104) ? 1 : -1];
typedef char light_options_is_16_bytes[(sizeof(CxLightOptions) == 16) ? 1 : -1];
typedef char tree_view_is_16_bytes[(sizeof(CxLightTreeView) == 16) ? 1 : -1];
typedef char trace_params_is_136_bytes[(sizeof(CxTraceParams) == 136) ? 1 : -1];
typedef char sky_cubemap_is_16_bytes[(sizeof(CxSkyCubemap) == 16) ? 1 : -1];
typedef char trace_params_hint_at_120[(offsetof(CxTraceParams, hint) == 120) ? 1 : -1];
typedef char light_hint_is_40_bytes[(sizeof(CxLightHint) == 40) ? 1 : -1];
typedef char bvh_tuning_is_8_bytes[(sizeof(CxLightBvhTuning) == 8) ? 1 : -1];
typedef char hash_slot_is_16_bytes[(sizeof(HashKeySlot) == 16) ? 1 : -1];

size_t cx_abi_c99_touch(void)
{
    CxLightOptions options = {1u, 20.0f, 0u};
    CxLightStats stats;
    CxLightSceneTables tables = {NULL, NULL, 0u, 0u, 0u, 0u};
    CxLightTreeView view = {NULL, 1u, 0u};
    CxLightPass light_pass = {NULL, 0u, NULL, NULL, 0u, 0u};
    CxLightBvhBuild build = {NULL, NULL, NULL, NULL, NULL, 64u, 104u};
    AnyFn table[] = {
        (AnyFn)cx_cxw_close,
        (AnyFn)cx_cxw_copy_decoded,
        (AnyFn)cx_cxw_decode,
        (AnyFn)cx_cxw_decode_all,
        (AnyFn)cx_cxw_entity,
        (AnyFn)cx_cxw_open,
        (AnyFn)cx_cxw_sector,
        (AnyFn)cx_cxw_world,
        (AnyFn)cx_hash_build_bricks,
        (AnyFn)cx_hash_build_keys,
        (AnyFn)cx_hash_build_scratch_bytes,
        (AnyFn)cx_hash_fingerprint,
        (AnyFn)cx_hash_get_bricks,
        (AnyFn)cx_hash_lookup,
        (AnyFn)cx_hash_maintain,
        (AnyFn)cx_light_bvh_batch_bytes,
        (AnyFn)cx_light_bvh_batch_prepare,
        (AnyFn)cx_light_bvh_begin,
        (AnyFn)cx_light_bvh_build,
        (AnyFn)cx_light_bvh_build_batch,
        (AnyFn)cx_light_bvh_build_keys64,
        (AnyFn)cx_light_bvh_build_static,
        (AnyFn)cx_light_bvh_get_tuning,
        (AnyFn)cx_light_bvh_scratch_bytes,
        (AnyFn)cx_light_bvh_scratch_init,
        (AnyFn)cx_light_bvh_set_tuning,
        (AnyFn)cx_light_box_support,
        (AnyFn)cx_light_brick_pdf,
        (AnyFn)cx_light_gather,
        (AnyFn)cx_light_gather_entities,
        (AnyFn)cx_light_mis_weight,
        (AnyFn)cx_light_pass,
        (AnyFn)cx_light_pass_scratch_bytes,
        (AnyFn)cx_light_point_pdf,
        (AnyFn)cx_light_sample,
        (AnyFn)cx_light_sample_brick_box,
        (AnyFn)cx_light_tree_create,
        (AnyFn)cx_light_tree_descriptor,
        (AnyFn)cx_light_tree_destroy,
        (AnyFn)cx_light_tree_end_flush,
        (AnyFn)cx_light_tree_get_stats,
        (AnyFn)cx_light_tree_group_rows,
        (AnyFn)cx_light_tree_light_codes,
        (AnyFn)cx_light_tree_on_entity_added,
        (AnyFn)cx_light_tree_on_entity_motion,
        (AnyFn)cx_light_tree_on_entity_moved,
        (AnyFn)cx_light_tree_on_entity_removed,
        (AnyFn)cx_light_tree_on_group_bricks_removed,
        (AnyFn)cx_light_tree_on_group_created,
        (AnyFn)cx_light_tree_on_group_released,
        (AnyFn)cx_light_tree_on_materials_changed,
        (AnyFn)cx_light_tree_pending,
        (AnyFn)cx_light_tree_prepare,
        (AnyFn)cx_light_tree_rebuild,
        (AnyFn)cx_light_tree_record,
        (AnyFn)cx_light_tree_record_transfers,
        (AnyFn)cx_light_tree_recycle,
        (AnyFn)cx_light_tree_set_min_emission,
        (AnyFn)cx_light_tree_stage,
        (AnyFn)cx_profile_enable,
        (AnyFn)cx_profile_read,
        (AnyFn)cx_profile_reset,
        (AnyFn)cx_radix_sort_keys64,
        (AnyFn)cx_radix_sort_keys64_state_bytes,
        (AnyFn)cx_radix_sort_pairs,
        (AnyFn)cx_radix_sort_state_bytes,
        (AnyFn)cx_scene_finalize,
        (AnyFn)cx_scene_finalize_bricks,
        (AnyFn)cx_set_log_callback,
        (AnyFn)cx_sky_cube_sample,
        (AnyFn)cx_sky_sh,
        (AnyFn)cx_trace_batch,
        (AnyFn)cx_trace_batch_bytes,
        (AnyFn)cx_trace_batch_prepare,
        (AnyFn)cx_trace_batch_workspace_bytes,
        (AnyFn)cx_trace_frame,
        (AnyFn)cx_trace_frame_adjoint,
        (AnyFn)cx_trace_get_tuning,
        (AnyFn)cx_trace_rays,
        (AnyFn)cx_trace_set_tuning,
        (AnyFn)cx_trace_workspace_bytes,
        (AnyFn)cx_version_string,
    };
    size_t total = sizeof(table) / sizeof(table[0]);
    stats.enabled = 0u;
    total += sizeof(options) + sizeof(stats) + sizeof(tables) + sizeof(view) + sizeof(light_pass) + sizeof(build);
    return total + stats.enabled;
}
