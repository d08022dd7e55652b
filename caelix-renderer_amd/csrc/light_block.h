// A group's light block in a heap buffer (16-byte units): the LightBlockHeader, the index entries in
// Morton order of the brick position in the group, the brick light records (light_tree_layout.h:252-334,
// shaders/common/caelix_light_tree_layout.hlsl:336-401). Shared by the group pass and the light probes.
#pragma once
#include "device_common.h"

namespace cx {

// LtMorton of a group-local brick position pos = x | y << shift | z << 2 shift
// (caelix_light_tree_layout.hlsl:258-261, BrickOfPos caelix_record.hlsl:61-66).
CX_DEV uint32_t MortonOfPos(uint32_t pos, uint32_t shift)
{
    const uint32_t mask = (1u << shift) - 1u;
    return Morton3(pos & mask, (pos >> shift) & mask, (pos >> (2 * shift)) & mask);
}

CX_DEV const LightIndexEntry* IndexEntry(const uint32_t* block, uint32_t entry)
{
    return reinterpret_cast<const LightIndexEntry*>(block + 4) + entry;
}

// LtFindIndexEntry (caelix_light_tree_layout.hlsl:336-360): the index entry of the brick whose Morton
// key is `key` among the block's `record_count` entries (binary search), or kLtNone.
CX_DEV uint32_t FindIndexEntry(const uint32_t* block, uint32_t record_count, uint32_t key, uint32_t shift)
{
    uint32_t lo = 0;
    uint32_t hi = record_count;
    while (lo < hi)
    {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t mid_key = MortonOfPos(IndexEntry(block, mid)->pos, shift);
        if (mid_key == key)
        {
            return mid;
        }
        if (mid_key < key)
        {
            lo = mid + 1;
        }
        else
        {
            hi = mid;
        }
    }
    return kLtNone;
}

// LtRecordUnits (caelix_light_tree_group.hlsl:209-212): units of a brick light record with `entries`
// face entries (48-byte header + 4 bytes each).
CX_DEV uint32_t RecordUnits(uint32_t entries)
{
    return (48u + 4u * entries + kLtUnitBytes - 1u) / kLtUnitBytes;
}

}  // namespace cx
