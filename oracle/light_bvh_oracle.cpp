// ORACLE — test infrastructure only. A sequential CPU restatement of the reference's light BVH build
// (shaders/light/bvh/*.hlsl), used by tests/ and by bench.py's CPU baseline as the checker. The
// product path never links, loads or calls this file.
//
// Compiled with -ffp-contract=off so that every float operation is a single IEEE-754 binary32
// operation in source order, as written in the HLSL.
//
// Pinning (what checks this restatement independently of itself, in tests/):
//   bounds       numpy min/max of the centres
//   morton       numpy bit interleave of the quantised cells, brute force per bit
//   sort         std::sort / numpy lexsort of (key, value) pairs on the reference test's inputs
//                (tests/light_tree_test.cpp:1460-1515: mt19937(3), n in {1, 2, 1000, 100000})
//   karras       a top-down recursive binary radix tree (highest differing bit split) with the same
//                node numbering, built in Python
//   aggregate    float64 sums and box containment (light_tree_test.cpp:508-639 CheckBvh)
//   depth/wide   parent walk in Python, CheckBvh traversal (every leaf once, parent links, inline
//                slot summaries equal to the leaf bytes)
#include "oracle_common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

struct Summary
{
    float face[6];
    float lo[3];
    float hi[3];
};

Summary LoadSummary(const LightSummary& s)
{
    Summary r;
    std::memcpy(&r, &s, sizeof(r));
    return r;
}

float Power(const Summary& s)
{
    // LtPower: ((((f0 + f1) + f2) + f3) + f4) + f5 (caelix_light_tree_layout.hlsl:271-274)
    return s.face[0] + s.face[1] + s.face[2] + s.face[3] + s.face[4] + s.face[5];
}

Summary EmptySummary()
{
    Summary s;
    std::memset(&s, 0, sizeof(s));
    return s;
}

// LtUnion (caelix_light_tree_layout.hlsl:315-334): a side with power <= 0 is skipped.
Summary Union(const Summary& a, const Summary& b)
{
    if (Power(a) <= 0.0f)
    {
        return b;
    }
    if (Power(b) <= 0.0f)
    {
        return a;
    }
    Summary s;
    for (int f = 0; f < 6; ++f)
    {
        s.face[f] = a.face[f] + b.face[f];
    }
    for (int k = 0; k < 3; ++k)
    {
        s.lo[k] = std::fmin(a.lo[k], b.lo[k]);
        s.hi[k] = std::fmax(a.hi[k], b.hi[k]);
    }
    return s;
}

// LtPart1By2 / LtMorton (caelix_light_tree_layout.hlsl:247-261).
uint32_t Part1By2(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

uint32_t Clz(uint32_t x)
{
    return x == 0 ? 32u : static_cast<uint32_t>(__builtin_clz(x));
}

struct Arrays
{
    uint32_t capacity;
    uint8_t* base;
    uint32_t* u32(uint32_t offset) const
    {
        return reinterpret_cast<uint32_t*>(base + uint64_t{offset} * capacity);
    }
    Summary* aggregates() const
    {
        return reinterpret_cast<Summary*>(base + uint64_t{kLtScratchAggregates} * capacity);
    }
};

// LtDelta (caelix_light_bvh_karras.hlsl:18-31).
int Delta(const uint32_t* keys, int i, int j, int n, uint32_t key_i)
{
    if (j < 0 || j >= n)
    {
        return -1;
    }
    const uint32_t key_j = keys[j];
    if (key_i == key_j)
    {
        return 32 + static_cast<int>(Clz(static_cast<uint32_t>(i) ^ static_cast<uint32_t>(j)));
    }
    return static_cast<int>(Clz(key_i ^ key_j));
}

Summary ChildSummary(const Arrays& a, const LightBvhLeaf* leaves, uint32_t child)
{
    if ((child & kLtChildLeafBit) != 0)
    {
        const uint32_t leaf = a.u32(kLtScratchValues)[child & ~kLtChildLeafBit];
        return LoadSummary(leaves[leaf].summary);
    }
    return a.aggregates()[child];
}

}  // namespace

extern "C" {

// Morton code of a cell with 10-bit coordinates.
uint32_t oracle_morton3(uint32_t x, uint32_t y, uint32_t z)
{
    return Part1By2(x) | (Part1By2(y) << 1) | (Part1By2(z) << 2);
}

// caelix_light_bvh_bounds.hlsl:10-53 — leaf count, dropped count, bounds of the leaf box centres.
void oracle_lt_bounds(const LightBvhLeaf* leaves, uint32_t leaf_capacity, LightTreeHeader* header,
                      LightTreeState* state)
{
    const uint32_t found = state->leaves_found;
    const uint32_t n = std::min(found, leaf_capacity);
    header->leaf_count = n;
    state->leaves_dropped = found - n;
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f};
    float hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (uint32_t i = 0; i < n; ++i)
    {
        for (int k = 0; k < 3; ++k)
        {
            const float c = 0.5f * (leaves[i].summary.lo[k] + leaves[i].summary.hi[k]);
            lo[k] = std::fmin(lo[k], c);
            hi[k] = std::fmax(hi[k], c);
        }
    }
    for (int k = 0; k < 3; ++k)
    {
        state->bounds[k] = lo[k];
        state->bounds[3 + k] = hi[k];
    }
}

// caelix_light_bvh_morton.hlsl:5-28 — sort input and cleared counters.
void oracle_lt_morton(const LightBvhLeaf* leaves, uint32_t leaf_capacity, const LightTreeHeader* header,
                      const LightTreeState* state, uint8_t* scratch)
{
    const Arrays a{leaf_capacity, scratch};
    const uint32_t n = header->leaf_count;
    float lo[3];
    float extent[3];
    for (int k = 0; k < 3; ++k)
    {
        lo[k] = state->bounds[k];
        extent[k] = std::fmax(state->bounds[3 + k] - lo[k], 1e-6f);
    }
    for (uint32_t i = 0; i < leaf_capacity; ++i)
    {
        a.u32(kLtScratchCounter)[i] = 0;
        if (i >= n)
        {
            continue;
        }
        uint32_t cell[3];
        for (int k = 0; k < 3; ++k)
        {
            const float centre = 0.5f * (leaves[i].summary.lo[k] + leaves[i].summary.hi[k]);
            float q = (centre - lo[k]) / extent[k] * 1024.0f;
            q = std::fmin(std::fmax(q, 0.0f), 1023.0f);  // clamp(x, 0, 1023)
            cell[k] = static_cast<uint32_t>(q);
        }
        a.u32(kLtScratchKeys)[i] = oracle_morton3(cell[0], cell[1], cell[2]);
        a.u32(kLtScratchValues)[i] = i;
    }
}

// The stable LSD radix sort of caelix_light_bvh_sort_{histogram,scatter}.hlsl: 4 bits per pass, 8
// passes, ping-pong between the two sort buffers; the result is in buffer 0. Restated per pass as a
// counting sort (histogram, exclusive scan, stable scatter in input order).
void oracle_radix_sort_pairs(uint32_t* keys0, uint32_t* values0, uint32_t* keys1, uint32_t* values1,
                             uint32_t n)
{
    uint32_t* keys[2] = {keys0, keys1};
    uint32_t* values[2] = {values0, values1};
    for (uint32_t pass = 0; pass < 8; ++pass)
    {
        const uint32_t in = pass & 1u;
        const uint32_t shift = 4 * pass;
        uint32_t count[16] = {};
        for (uint32_t i = 0; i < n; ++i)
        {
            ++count[(keys[in][i] >> shift) & 15u];
        }
        uint32_t offset[16];
        uint32_t sum = 0;
        for (int d = 0; d < 16; ++d)
        {
            offset[d] = sum;
            sum += count[d];
        }
        for (uint32_t i = 0; i < n; ++i)
        {
            const uint32_t d = (keys[in][i] >> shift) & 15u;
            const uint32_t dst = offset[d]++;
            keys[1 - in][dst] = keys[in][i];
            values[1 - in][dst] = values[in][i];
        }
    }
}

// caelix_light_bvh_karras.hlsl:33-96 — the binary radix tree over the sorted keys.
void oracle_lt_karras(uint32_t leaf_capacity, const LightTreeHeader* header, uint8_t* scratch)
{
    const Arrays a{leaf_capacity, scratch};
    const int n = static_cast<int>(header->leaf_count);
    const uint32_t* keys = a.u32(kLtScratchKeys);
    for (int i = 0; i + 1 < n; ++i)
    {
        if (i == 0)
        {
            a.u32(kLtScratchParent)[0] = kLtNone;
        }
        const uint32_t key_i = keys[i];
        const int d = Delta(keys, i, i + 1, n, key_i) - Delta(keys, i, i - 1, n, key_i) >= 0 ? 1 : -1;
        const int delta_min = Delta(keys, i, i - d, n, key_i);
        int length_max = 2;
        while (Delta(keys, i, i + length_max * d, n, key_i) > delta_min)
        {
            length_max *= 2;
        }
        int length = 0;
        for (int step = length_max / 2; step >= 1; step /= 2)
        {
            if (Delta(keys, i, i + (length + step) * d, n, key_i) > delta_min)
            {
                length += step;
            }
        }
        const int j = i + length * d;
        const int delta_node = Delta(keys, i, j, n, key_i);
        int split = 0;
        int divisor = 2;
        int step2 = (length + divisor - 1) / divisor;
        while (true)
        {
            if (Delta(keys, i, i + (split + step2) * d, n, key_i) > delta_node)
            {
                split += step2;
            }
            if (step2 == 1)
            {
                break;
            }
            divisor *= 2;
            step2 = (length + divisor - 1) / divisor;
        }
        const int gamma = i + split * d + std::min(d, 0);
        const uint32_t left = std::min(i, j) == gamma ? (kLtChildLeafBit | uint32_t(gamma)) : uint32_t(gamma);
        const uint32_t right =
            std::max(i, j) == gamma + 1 ? (kLtChildLeafBit | uint32_t(gamma + 1)) : uint32_t(gamma + 1);
        a.u32(kLtScratchLeft)[i] = left;
        a.u32(kLtScratchRight)[i] = right;
        a.u32((left & kLtChildLeafBit) ? kLtScratchSortedLeafParent : kLtScratchParent)[left & ~kLtChildLeafBit] =
            uint32_t(i);
        a.u32((right & kLtChildLeafBit) ? kLtScratchSortedLeafParent : kLtScratchParent)[right & ~kLtChildLeafBit] =
            uint32_t(i);
    }
}

// caelix_light_bvh_aggregate.hlsl:30-55 — bottom-up aggregation, one walk per sorted leaf; the
// second arrival at a node unions its children (left, then right).
void oracle_lt_aggregate(const LightBvhLeaf* leaves, uint32_t leaf_capacity, const LightTreeHeader* header,
                         uint8_t* scratch)
{
    const Arrays a{leaf_capacity, scratch};
    const uint32_t n = header->leaf_count;
    if (n < 2)
    {
        return;
    }
    for (uint32_t j = 0; j < n; ++j)
    {
        uint32_t node = a.u32(kLtScratchSortedLeafParent)[j];
        for (uint32_t depth = 0; depth < 128 && node != kLtNone; ++depth)
        {
            if (a.u32(kLtScratchCounter)[node]++ == 0)
            {
                break;
            }
            a.aggregates()[node] = Union(ChildSummary(a, leaves, a.u32(kLtScratchLeft)[node]),
                                         ChildSummary(a, leaves, a.u32(kLtScratchRight)[node]));
            node = a.u32(kLtScratchParent)[node];
        }
    }
}

// caelix_light_bvh_depth.hlsl:7-30 — depth by parent walk; even-depth flags.
void oracle_lt_depth(uint32_t leaf_capacity, const LightTreeHeader* header, uint8_t* scratch)
{
    const Arrays a{leaf_capacity, scratch};
    const uint32_t n = header->leaf_count;
    for (uint32_t i = 0; i < leaf_capacity; ++i)
    {
        uint32_t flag = 0;
        if (i + 1 < n)
        {
            uint32_t depth = 0;
            uint32_t node = a.u32(kLtScratchParent)[i];
            for (uint32_t guard = 0; guard < 128 && node != kLtNone; ++guard)
            {
                ++depth;
                node = a.u32(kLtScratchParent)[node];
            }
            a.u32(kLtScratchCounter)[i] = depth;
            flag = (depth & 1u) == 0 ? 1u : 0u;
        }
        a.u32(kLtScratchWide)[i] = flag;
    }
}

// caelix_light_bvh_scan.hlsl:22-58 — exclusive prefix sum in place; returns the total.
uint32_t oracle_exclusive_scan(uint32_t* data, uint32_t count)
{
    uint32_t carry = 0;
    for (uint32_t i = 0; i < count; ++i)
    {
        const uint32_t v = data[i];
        data[i] = carry;
        carry += v;
    }
    return carry;
}

// caelix_light_bvh_wide.hlsl:10-100 — collapse to 4-wide nodes.
void oracle_lt_wide(LightBvhLeaf* leaves, uint32_t leaf_capacity, uint32_t node_capacity,
                    const LightTreeHeader* header, uint8_t* scratch, LightBvhNode* nodes)
{
    const Arrays a{leaf_capacity, scratch};
    const uint32_t n = header->leaf_count;
    for (uint32_t i = 0; i + 1 < n; ++i)
    {
        if ((a.u32(kLtScratchCounter)[i] & 1u) != 0)
        {
            continue;
        }
        const uint32_t w = a.u32(kLtScratchWide)[i];
        if (w >= node_capacity)
        {
            continue;
        }
        uint32_t slots[4] = {kLtNone, kLtNone, kLtNone, kLtNone};
        uint32_t count = 0;
        const uint32_t children[2] = {a.u32(kLtScratchLeft)[i], a.u32(kLtScratchRight)[i]};
        for (uint32_t c = 0; c < 2; ++c)
        {
            if ((children[c] & kLtChildLeafBit) != 0)
            {
                slots[count++] = children[c];
            }
            else
            {
                slots[count++] = a.u32(kLtScratchLeft)[children[c]];
                slots[count++] = a.u32(kLtScratchRight)[children[c]];
            }
        }
        LightBvhNode& node = nodes[w];
        for (uint32_t s = 0; s < 4; ++s)
        {
            Summary summary = EmptySummary();
            uint32_t word = kLtNone;
            if (s < count)
            {
                const uint32_t slot = slots[s];
                summary = ChildSummary(a, leaves, slot);
                if ((slot & kLtChildLeafBit) != 0)
                {
                    const uint32_t leaf = a.u32(kLtScratchValues)[slot & ~kLtChildLeafBit];
                    word = kLtChildLeafBit | leaf;
                    leaves[leaf].parent = (w << kLtParentSlotBits) | s;
                }
                else
                {
                    const uint32_t child_node = a.u32(kLtScratchWide)[slot];
                    word = child_node;
                    if (child_node < node_capacity)
                    {
                        nodes[child_node].parent = w;
                        nodes[child_node].parent_slot = s;
                    }
                }
            }
            std::memcpy(node.child_face[s], summary.face, sizeof(summary.face));
            std::memcpy(node.child_lo[s], summary.lo, sizeof(summary.lo));
            std::memcpy(node.child_hi[s], summary.hi, sizeof(summary.hi));
            node.child[s] = word;
        }
        if (i == 0)
        {
            node.parent = kLtNone;
            node.parent_slot = 0;
        }
    }
}

// caelix_light_bvh_finish.hlsl:8-63 — the tree header.
void oracle_lt_finish(LightBvhLeaf* leaves, uint32_t leaf_capacity, uint32_t node_capacity,
                      LightTreeHeader* header, LightTreeState* state, const uint8_t* scratch,
                      LightBvhNode* nodes)
{
    const Arrays a{leaf_capacity, const_cast<uint8_t*>(scratch)};
    const uint32_t n = header->leaf_count;
    Summary root = EmptySummary();
    uint32_t root_index = kLtNone;
    uint32_t node_count = 0;
    uint32_t overflow = 0;
    if (n == 1)
    {
        root = LoadSummary(leaves[0].summary);
        LightBvhNode& node = nodes[0];
        std::memset(node.child_face, 0, sizeof(node.child_face));
        std::memset(node.child_lo, 0, sizeof(node.child_lo));
        std::memset(node.child_hi, 0, sizeof(node.child_hi));
        std::memcpy(node.child_face[0], root.face, sizeof(root.face));
        std::memcpy(node.child_lo[0], root.lo, sizeof(root.lo));
        std::memcpy(node.child_hi[0], root.hi, sizeof(root.hi));
        node.child[0] = kLtChildLeafBit | 0u;
        node.child[1] = node.child[2] = node.child[3] = kLtNone;
        node.parent = kLtNone;
        node.parent_slot = 0;
        leaves[0].parent = 0;
        root_index = 0;
        node_count = 1;
    }
    else if (n >= 2)
    {
        node_count = header->node_count;
        if (node_count > node_capacity)
        {
            overflow = 1;
        }
        else
        {
            root = a.aggregates()[0];
            root_index = 0;
        }
    }
    std::memcpy(&header->summary, &root, sizeof(root));
    header->root = root_index;
    header->leaf_count = overflow != 0 ? 0u : n;
    header->node_count = node_count;
    header->pad = 0;
    state->node_overflow = overflow;
    state->builds += 1;
}

// The whole build of one tree from its leaves (caelix_light_tree_record.cpp:271-308 after the
// gather): bounds, Morton, sort, Karras, aggregate, depth, scan, wide, finish. Gated on
// state->building like every reference kernel. `scratch` holds kLtScratchSortState * capacity
// bytes (the oracle needs no sort tile state).
void oracle_lt_build(LightBvhLeaf* leaves, uint32_t leaf_capacity, uint32_t node_capacity,
                     LightTreeHeader* header, LightTreeState* state, uint8_t* scratch, LightBvhNode* nodes)
{
    if (state->building == 0)
    {
        return;
    }
    const Arrays a{leaf_capacity, scratch};
    oracle_lt_bounds(leaves, leaf_capacity, header, state);
    oracle_lt_morton(leaves, leaf_capacity, header, state, scratch);
    const uint32_t n = header->leaf_count;
    oracle_radix_sort_pairs(a.u32(kLtScratchKeys), a.u32(kLtScratchValues),
                            a.u32(kLtScratchKeys + kLtScratchSortBufferStride),
                            a.u32(kLtScratchValues + kLtScratchSortBufferStride), n);
    oracle_lt_karras(leaf_capacity, header, scratch);
    oracle_lt_aggregate(leaves, leaf_capacity, header, scratch);
    oracle_lt_depth(leaf_capacity, header, scratch);
    header->node_count = oracle_exclusive_scan(a.u32(kLtScratchWide), leaf_capacity);
    oracle_lt_wide(leaves, leaf_capacity, node_capacity, header, scratch, nodes);
    oracle_lt_finish(leaves, leaf_capacity, node_capacity, header, state, scratch, nodes);
}

}  // extern "C"
