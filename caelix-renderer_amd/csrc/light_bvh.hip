// Light BVH build for gfx950 (the build half of LightTree::RecordTree,
// lib/src/light/light_tree_record.cpp:271-308; kernels shaders/light/bvh/*.hlsl).
//
// Launch sequence (one stream, 9 kernels + 4 sort passes):
//   bounds     grid-stride min/max of the leaf box centres; writes the centres (12 B per leaf, read
//              again by the Morton kernel instead of the 64-B leaves); the last workgroup reduces the
//              partials, writes leaf_count / leaves_dropped / bounds and resets the build's counters
//   morton     30-bit key + leaf index per leaf, the four 8-bit digit histograms of the onesweep
//              sort, the cleared arrival counters and gap-cover array, the tile status rows
//   sort x4    onesweep passes (radix_sort.hip); sorted pairs end in sort buffer 0
//   karras     binary radix tree (Karras 2012) as the reference builds it, plus +1/-1 at the two ends
//              of every node's leaf range in the gap-cover array and the node's split gap
//   aggregate  bottom-up union with agent-scope acq_rel arrival counters
//   cover      inclusive scan of the gap-cover array (decoupled look-back): cover[g] = the number of
//              binary nodes whose range spans the gap between sorted leaves g and g + 1
//   flags      depth(i) = cover[split(i)] - 1 (the nodes spanning node i's split gap are i and its
//              ancestors), even-depth flags, their exclusive scan = 4-wide node indices, and the
//              header (the reference's caelix_light_bvh_finish) in the last tile
//   wide       4-wide nodes from the even-depth binary nodes
// The reference computes the depth by walking every node up to the root (O(n * depth) dependent
// loads, caelix_light_bvh_depth.hlsl:7-30); the gap cover gives the same depths with two atomics per
// node and one scan.
#include "device_common.h"
#include "radix_sort.h"

#include <vector>
#include "runtime.h"

namespace cx {
namespace {

constexpr uint32_t kBvhThreads = 256;
constexpr uint32_t kBoundsMaxBlocks = 1024;
constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kScanItems = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;

// Scan tile status (64-bit): flag in bits 32..33, signed 32-bit value in bits 0..31.
constexpr uint64_t kScanAggregate = 1ull << 32;
constexpr uint64_t kScanPrefix = 2ull << 32;
constexpr uint64_t kScanFlagMask = 3ull << 32;

// Build control words (u32) at the start of the sort-state region of the scratch.
enum : uint32_t
{
    kCtlUnused = 0,
    kCtlLeafCount = 1,
    kCtlCoverTiles = 2,
    kCtlFlagsTiles = 3,
    kCtlWideNodes = 4,  // 4-wide nodes of the build (set by the flags scan)
    kCtlWords = 64,
};

// The bottom-up tree build runs in levels of nested leaf blocks: a level-1 workgroup owns 512
// sorted leaves (and works in LDS), a level-L workgroup the 32 level-(L-1) blocks below it.
constexpr uint32_t kTreeThreads = 512;  // level 1: one lane per leaf
constexpr uint32_t kTreeTile = 512;
constexpr uint32_t kTreeFanout = 32;
constexpr uint32_t kTopsCap = 256;

__host__ __device__ inline uint64_t TreeTile(uint32_t level)
{
    uint64_t t = kTreeTile;
    for (uint32_t l = 1; l < level; ++l)
    {
        t *= kTreeFanout;
    }
    return t;
}
__host__ __device__ inline uint32_t TreeLevels(uint32_t capacity)
{
    uint32_t level = 1;
    while (TreeTile(level) < capacity)
    {
        ++level;
    }
    return level;
}
__host__ __device__ inline uint32_t TreeBlocks(uint32_t capacity, uint32_t level)
{
    return uint32_t((uint64_t(capacity) + TreeTile(level) - 1) / TreeTile(level));
}
__host__ __device__ inline uint64_t TopsBytes(uint32_t capacity, uint32_t level)
{
    const uint64_t blocks = TreeBlocks(capacity, level);
    return ((blocks * kTopsCap * 8 + blocks * 4) + 255) & ~uint64_t(255);
}

struct Scratch
{
    uint8_t* base;
    uint32_t capacity;
    __host__ __device__ uint32_t* u32(uint32_t offset) const { return reinterpret_cast<uint32_t*>(base + uint64_t(offset) * capacity); }
    __host__ __device__ uint32_t* keys(uint32_t buffer) const { return u32(kLtScratchKeys + buffer * kLtScratchSortBufferStride); }
    __host__ __device__ uint32_t* values(uint32_t buffer) const { return u32(kLtScratchValues + buffer * kLtScratchSortBufferStride); }
    __host__ __device__ float* aggregates() const { return reinterpret_cast<float*>(base + uint64_t(kLtScratchAggregates) * capacity); }
    // Free arrays reused by this build: the leaf centres live in the aggregates array until the
    // aggregation; the gap-cover diffs in `wide` until the flags scan; after the sort, sort buffer 1
    // holds the cover counts (keys) and every node's split gap (values), and the flags scan replaces
    // the consumed split gaps by the binary node of every 4-wide node (wide_source).
    __host__ __device__ float* centres() const { return aggregates(); }
    __host__ __device__ int32_t* cover_diff() const { return reinterpret_cast<int32_t*>(u32(kLtScratchWide)); }
    __host__ __device__ int32_t* cover() const { return reinterpret_cast<int32_t*>(keys(1)); }
    __host__ __device__ uint32_t* split() const { return values(1); }
    __host__ __device__ uint32_t* wide_source() const { return values(1); }
    __host__ __device__ uint8_t* state_base() const { return base + uint64_t(kLtScratchSortState) * capacity; }
    __host__ __device__ uint32_t* ctl() const { return reinterpret_cast<uint32_t*>(state_base()); }
    __host__ __device__ float* bounds_partials() const { return reinterpret_cast<float*>(state_base() + kCtlWords * 4); }
    __host__ __device__ uint64_t* cover_status() const
    {
        return reinterpret_cast<uint64_t*>(state_base() + kCtlWords * 4 + kBoundsMaxBlocks * 32);
    }
    __host__ __device__ uint32_t scan_tiles() const { return DivUp(capacity, kScanTile); }
    __host__ __device__ uint64_t* flags_status() const { return cover_status() + scan_tiles(); }
    __host__ __device__ uint8_t* sort_base() const
    {
        const uint64_t offset = kCtlWords * 4 + kBoundsMaxBlocks * 32 + uint64_t(2) * scan_tiles() * 8;
        return state_base() + ((offset + 255) & ~uint64_t(255));
    }
    __host__ __device__ SortState sort() const { return SortStateAt(sort_base(), capacity); }
    // The 64-bit key variant's sort: its state at the same place (the region holds the larger of the
    // two layouts), its keys (Morton code << 32 | leaf) over sort buffer `buffer` (keys and values
    // side by side, 8 bytes per leaf).
    __host__ __device__ SortState sort64() const { return SortStateOf<SortKeys64>(sort_base(), capacity); }
    __host__ __device__ uint64_t* keys64(uint32_t buffer) const
    {
        return reinterpret_cast<uint64_t*>(u32(kLtScratchKeys + buffer * kLtScratchSortBufferStride));
    }
    // Tops of the tree levels: per level-L workgroup, up to kTopsCap (first, last) leaf ranges of
    // finished elements whose parent gap is on the workgroup's boundary, and their count.
    __host__ __device__ uint8_t* tops_base() const
    {
        const uint64_t state = SortStateBytesOf<SortPairs32>(capacity) > SortStateBytesOf<SortKeys64>(capacity)
                                   ? SortStateBytesOf<SortPairs32>(capacity)
                                   : SortStateBytesOf<SortKeys64>(capacity);
        const uint64_t offset = uint64_t(sort_base() - base) + state;
        return base + ((offset + 255) & ~uint64_t(255));
    }
    __host__ __device__ uint2* tops(uint32_t level) const
    {
        uint8_t* p = tops_base();
        for (uint32_t l = 1; l < level; ++l)
        {
            p += TopsBytes(capacity, l);
        }
        return reinterpret_cast<uint2*>(p);
    }
    __host__ __device__ uint32_t* tops_count(uint32_t level) const
    {
        return reinterpret_cast<uint32_t*>(tops(level) + uint64_t(TreeBlocks(capacity, level)) * kTopsCap);
    }
    static uint64_t Bytes(uint32_t capacity)
    {
        const Scratch s{nullptr, capacity};
        uint8_t* p = s.tops_base();
        for (uint32_t l = 1; l <= TreeLevels(capacity); ++l)
        {
            p += TopsBytes(capacity, l);
        }
        return uint64_t(p - static_cast<uint8_t*>(nullptr));
    }
};

struct Build
{
    LightBvhLeaf* leaves;
    LightBvhNode* nodes;
    LightTreeHeader* header;
    LightTreeState* state;
    Scratch scratch;
    uint32_t leaf_capacity;
    uint32_t node_capacity;
};

CX_DEV bool Building(const Build& b)
{
    return b.state->building != 0;
}

// The builds a launch works on: one (the plain entry point) or one per scene of a batch, selected by
// blockIdx.y. Every kernel sizes its own part of the grid from its build's capacity, so a batch
// launch covers the largest build and the workgroups beyond a smaller build's grid return at once.
struct OneBuild
{
    Build b;
    CX_DEV const Build& Get() const { return b; }
};
struct BuildArray
{
    const Build* builds;
    CX_DEV const Build& Get() const { return builds[blockIdx.y]; }
};

__host__ __device__ inline uint32_t BoundBlocks(uint32_t capacity)
{
    return SortTiles(capacity) < kBoundsMaxBlocks ? SortTiles(capacity) : kBoundsMaxBlocks;
}

struct Summary
{
    float face[6];
    float lo[3];
    float hi[3];
};

CX_DEV Summary LoadSummary(const float* p)
{
    const float4 a = reinterpret_cast<const float4*>(p)[0];
    const float4 b = reinterpret_cast<const float4*>(p)[1];
    const float4 c = reinterpret_cast<const float4*>(p)[2];
    Summary s;
    s.face[0] = a.x;
    s.face[1] = a.y;
    s.face[2] = a.z;
    s.face[3] = a.w;
    s.face[4] = b.x;
    s.face[5] = b.y;
    s.lo[0] = b.z;
    s.lo[1] = b.w;
    s.lo[2] = c.x;
    s.hi[0] = c.y;
    s.hi[1] = c.z;
    s.hi[2] = c.w;
    return s;
}

// Coherent load of a binary aggregate written in the same launch (agent-scope word loads).
CX_DEV Summary LoadSummaryCoherent(const float* p)
{
    const uint64_t* q = reinterpret_cast<const uint64_t*>(p);
    uint64_t w[6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
    {
        w[i] = LoadRelaxed(q + i);
    }
    Summary s;
    __builtin_memcpy(&s, w, sizeof(s));
    return s;
}

CX_DEV void StoreSummaryCoherent(float* p, const Summary& s)
{
    uint64_t w[6];
    __builtin_memcpy(w, &s, sizeof(s));
    uint64_t* q = reinterpret_cast<uint64_t*>(p);
#pragma unroll
    for (int i = 0; i < 6; ++i)
    {
        StoreRelaxed(q + i, w[i]);
    }
}

CX_DEV void StoreSummary(float* p, const Summary& s)
{
    float4* q = reinterpret_cast<float4*>(p);
    q[0] = make_float4(s.face[0], s.face[1], s.face[2], s.face[3]);
    q[1] = make_float4(s.face[4], s.face[5], s.lo[0], s.lo[1]);
    q[2] = make_float4(s.lo[2], s.hi[0], s.hi[1], s.hi[2]);
}

CX_DEV Summary EmptySummary()
{
    Summary s;
#pragma unroll
    for (int f = 0; f < 6; ++f)
    {
        s.face[f] = 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        s.lo[k] = 0.0f;
        s.hi[k] = 0.0f;
    }
    return s;
}

CX_DEV float Power(const Summary& s)
{
    return s.face[0] + s.face[1] + s.face[2] + s.face[3] + s.face[4] + s.face[5];
}

// LtUnion (shaders/common/caelix_light_tree_layout.hlsl:315-334).
CX_DEV Summary Union(const Summary& a, const Summary& b)
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
#pragma unroll
    for (int f = 0; f < 6; ++f)
    {
        s.face[f] = a.face[f] + b.face[f];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        s.lo[k] = fminf(a.lo[k], b.lo[k]);
        s.hi[k] = fmaxf(a.hi[k], b.hi[k]);
    }
    return s;
}

// LtBinaryChildSummary (caelix_light_bvh_common.hlsl:94-121).
CX_DEV Summary ChildSummary(const Build& b, uint32_t child, bool coherent)
{
    if ((child & kLtChildLeafBit) != 0)
    {
        const uint32_t leaf = b.scratch.values(0)[child & ~kLtChildLeafBit];
        return LoadSummary(b.leaves[leaf].summary.face);
    }
    const float* p = b.scratch.aggregates() + uint64_t(child) * 12;
    return coherent ? LoadSummaryCoherent(p) : LoadSummary(p);
}

// ------------------------------------------------------------------------------------------------
// bounds (caelix_light_bvh_bounds.hlsl:10-53), first half: per-workgroup min/max of the leaf box
// centres (the partials) and the centres themselves. Workgroup 0 also resets the build's counters
// and the sort histograms for the kernels that follow. The Morton kernel reduces the partials, so no
// workgroup of this kernel waits on another.
template <int kWaves>
CX_DEV void MinMax3(float (&lo)[3], float (&hi)[3], float* lds, uint32_t wave)  // lds: 6 x kWaves floats
{
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset >>= 1)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], offset, kWave));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], offset, kWave));
        }
    }
    if (Lane() == 0)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            lds[k * kWaves + wave] = lo[k];
            lds[(3 + k) * kWaves + wave] = hi[k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        float l = lds[k * kWaves];
        float h = lds[(3 + k) * kWaves];
#pragma unroll
        for (int w = 1; w < kWaves; ++w)
        {
            l = fminf(l, lds[k * kWaves + w]);
            h = fmaxf(h, lds[(3 + k) * kWaves + w]);
        }
        lo[k] = l;
        hi[k] = h;
    }
}

template <class Set, class Keys>
__global__ void __launch_bounds__(kBvhThreads) BoundsKernel(Set set)
{
    const Build& b = set.Get();
    const uint32_t grid = BoundBlocks(b.leaf_capacity);
    if (blockIdx.x >= grid || !Building(b))
    {
        return;
    }
    __shared__ float lds[6 * (kBvhThreads / kWave)];
    const uint32_t t = threadIdx.x;
    const uint32_t n = min(b.state->leaves_found, b.leaf_capacity);
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f};
    float hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    float* centres = b.scratch.centres();
    for (uint32_t i = blockIdx.x * kBvhThreads + t; i < n; i += grid * kBvhThreads)
    {
        const float* leaf = b.leaves[i].summary.lo;
        const float2 l01 = *reinterpret_cast<const float2*>(leaf);       // lo.x, lo.y
        const float2 l2h0 = *reinterpret_cast<const float2*>(leaf + 2);  // lo.z, hi.x
        const float2 h12 = *reinterpret_cast<const float2*>(leaf + 4);   // hi.y, hi.z
        const float c0 = 0.5f * (l01.x + l2h0.y);
        const float c1 = 0.5f * (l01.y + h12.x);
        const float c2 = 0.5f * (l2h0.x + h12.y);
        centres[3 * uint64_t(i)] = c0;
        centres[3 * uint64_t(i) + 1] = c1;
        centres[3 * uint64_t(i) + 2] = c2;
        lo[0] = fminf(lo[0], c0);
        lo[1] = fminf(lo[1], c1);
        lo[2] = fminf(lo[2], c2);
        hi[0] = fmaxf(hi[0], c0);
        hi[1] = fmaxf(hi[1], c1);
        hi[2] = fmaxf(hi[2], c2);
    }
    MinMax3<kBvhThreads / kWave>(lo, hi, lds, t / kWave);
    float* partials = b.scratch.bounds_partials();
    if (t < 6)
    {
        partials[blockIdx.x * 8 + t] = t < 3 ? lo[t] : hi[t - 3];
    }
    if (blockIdx.x == 0)
    {
        uint32_t* ctl = b.scratch.ctl();
        if (t == 0)
        {
            ctl[kCtlCoverTiles] = 0;
            ctl[kCtlFlagsTiles] = 0;
        }
        const SortState s = SortStateOf<Keys>(b.scratch.sort_base(), b.leaf_capacity);
        for (uint32_t i = t; i < kSortCtlBytes / 4; i += kBvhThreads)
        {
            s.ctl[i] = 0;
        }
        for (uint32_t i = t; i < SortHistBytesOf<Keys>() / 4; i += kBvhThreads)
        {
            s.hist[i] = 0;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// morton (caelix_light_bvh_morton.hlsl:5-28) + the sort's digit histograms and status clear.
// One workgroup per 8192 leaves. Every workgroup first reduces the bounds partials (the second half
// of caelix_light_bvh_bounds); workgroup 0 writes leaf_count, leaves_dropped and bounds. The keys:
// (code, leaf) pairs in sort buffer 0 (SortPairs32), or code << 32 | leaf in sort buffer 1
// (SortKeys64, whose tiles are half as long: two status rows per workgroup).
template <class Set, class Keys>
__global__ void __launch_bounds__(kSortThreads) MortonKernel(Set set)
{
    const Build& b = set.Get();
    const uint32_t bound_blocks = BoundBlocks(b.leaf_capacity);
    if (blockIdx.x >= SortTiles(b.leaf_capacity) || !Building(b))
    {
        return;
    }
    constexpr uint32_t kHistWords = Keys::kPasses * Keys::kRadix;
    constexpr uint32_t kStatusRows = kSortTile / (Keys::kThreads * Keys::kItems);  // key tiles per workgroup
    __shared__ uint32_t lds_hist[kHistWords];
    __shared__ float lds_minmax[6 * (kSortThreads / kWave)];
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < kHistWords; i += kSortThreads)
    {
        lds_hist[i] = 0;
    }
    const SortState s = SortStateOf<Keys>(b.scratch.sort_base(), b.leaf_capacity);
#pragma unroll
    for (uint32_t r = 0; r < kStatusRows; ++r)
    {
        SortClearTileStatusOf<Keys>(s, blockIdx.x * kStatusRows + r);
    }
    if (t == 0)
    {
        const uint32_t tile = blockIdx.x;
        if (tile < b.scratch.scan_tiles())
        {
            b.scratch.cover_status()[tile] = 0;
            b.scratch.flags_status()[tile] = 0;
        }
    }
    const uint32_t found = b.state->leaves_found;
    const uint32_t n = min(found, b.leaf_capacity);
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f};
    float hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    const float* partials = b.scratch.bounds_partials();
    for (uint32_t g = t; g < bound_blocks; g += kSortThreads)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = fminf(lo[k], partials[g * 8 + k]);
            hi[k] = fmaxf(hi[k], partials[g * 8 + 3 + k]);
        }
    }
    MinMax3<kSortThreads / kWave>(lo, hi, lds_minmax, t / kWave);
    if (blockIdx.x == 0 && t == 0)
    {
        b.header->leaf_count = n;
        b.state->leaves_dropped = found - n;
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            b.state->bounds[k] = lo[k];
            b.state->bounds[3 + k] = hi[k];
        }
        b.scratch.ctl()[kCtlLeafCount] = n;
    }
    const float lo0 = lo[0];
    const float lo1 = lo[1];
    const float lo2 = lo[2];
    const float e0 = fmaxf(hi[0] - lo0, 1e-6f);
    const float e1 = fmaxf(hi[1] - lo1, 1e-6f);
    const float e2 = fmaxf(hi[2] - lo2, 1e-6f);
    const float* centres = b.scratch.centres();
    uint32_t* keys = b.scratch.keys(0);
    uint32_t* values = b.scratch.values(0);
    uint32_t* counter = b.scratch.u32(kLtScratchCounter);
    int32_t* diff = b.scratch.cover_diff();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll 4
    for (int k = 0; k < kSortItems; ++k)
    {
        const uint32_t i = base + k * kSortThreads + t;
        if (i >= b.leaf_capacity)
        {
            break;
        }
        counter[i] = i + 1 < n ? kLtNone : 0u;  // gap slots of the tree build; 0 past the last gap
        diff[i] = 0;
        if (i < n)
        {
            const float c0 = centres[3 * uint64_t(i)];
            const float c1 = centres[3 * uint64_t(i) + 1];
            const float c2 = centres[3 * uint64_t(i) + 2];
            // clamp((c - lo) / extent * 1024, 0, 1023) with an IEEE division, truncated to uint.
            const uint32_t x = uint32_t(fminf(fmaxf(__fdiv_rn(c0 - lo0, e0) * 1024.0f, 0.0f), 1023.0f));
            const uint32_t y = uint32_t(fminf(fmaxf(__fdiv_rn(c1 - lo1, e1) * 1024.0f, 0.0f), 1023.0f));
            const uint32_t z = uint32_t(fminf(fmaxf(__fdiv_rn(c2 - lo2, e2) * 1024.0f, 0.0f), 1023.0f));
            const uint32_t key = Morton3(x, y, z);
            if constexpr (Keys::kValues)
            {
                keys[i] = key;
                values[i] = i;
                SortCountKeyOf<Keys>(lds_hist, key);
            }
            else
            {
                const uint64_t wide_key = (uint64_t(key) << 32) | i;
                b.scratch.keys64(1)[i] = wide_key;
                SortCountKeyOf<Keys>(lds_hist, wide_key);
            }
        }
    }
    __syncthreads();
    SortFlushHistogramOf<Keys>(lds_hist, s.hist);
}

// ------------------------------------------------------------------------------------------------
// The binary radix tree and its aggregation in one bottom-up pass.
//
// The tree is the reference's (caelix_light_bvh_karras.hlsl:8-96): the binary radix tree over the
// sorted keys, equal keys ordered by sorted position, so the key of sorted position i is
// (code << 32 | i) and delta(g) = the common prefix length of positions g and g + 1 (32 + clz of
// the positions for equal codes). Instead of a binary search per node, every element (a leaf, or a
// node with leaf range [a, b]) climbs to its parent: the parent's split is the adjacent gap with the
// longer common prefix, b if delta(b) > delta(a - 1), else a - 1 (Apetrei 2014). Karras numbers a
// node by the range end that is not its parent's split: a left child [a, g] is node g, a right child
// [g + 1, b] is node g + 1, the root is node 0. The first element to arrive at a gap parks its far
// range end in slot[gap] and stops; the second builds the parent: range, children, parent links, and
// the union of the two summaries (left, then right: caelix_light_bvh_aggregate.hlsl:30-55), then
// climbs on. Slot arrivals are workgroup-scope atomics: a level-L workgroup only arrives at gaps
// inside its own block of sorted leaves, so both arrivals of a gap happen in one workgroup of one
// launch, or the earlier one in an earlier launch. An element whose parent gap is on the block's
// boundary is handed to the next level as a "top". Also written per node: its split gap and +1/-1 at
// its range ends in the gap-cover array (for the depths).

CX_DEV uint32_t Clz(uint32_t x)
{
    return x == 0 ? 32u : uint32_t(__clz(int(x)));
}

// delta(g) of the gap between sorted positions g and g + 1; -1 outside [0, n - 1).
CX_DEV int GapDelta(const uint32_t* keys, int g, int n)
{
    if (g < 0 || g + 1 >= n)
    {
        return -1;
    }
    const uint32_t k0 = keys[g];
    const uint32_t k1 = keys[g + 1];
    if (k0 == k1)
    {
        return 32 + int(Clz(uint32_t(g) ^ uint32_t(g + 1)));
    }
    return int(Clz(k0 ^ k1));
}

CX_DEV Summary ElementSummary(const Build& b, uint32_t element)
{
    if ((element & kLtChildLeafBit) != 0)
    {
        const uint32_t leaf = b.scratch.values(0)[element & ~kLtChildLeafBit];
        return LoadSummary(b.leaves[leaf].summary.face);
    }
    const uint64_t* q = reinterpret_cast<const uint64_t*>(b.scratch.aggregates() + uint64_t(element) * 12);
    uint64_t w[6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
    {
        w[i] = __hip_atomic_load(q + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    Summary s;
    __builtin_memcpy(&s, w, sizeof(s));
    return s;
}

CX_DEV void StoreAggregate(const Build& b, uint32_t node, const Summary& s)
{
    uint64_t w[6];
    __builtin_memcpy(w, &s, sizeof(s));
    uint64_t* q = reinterpret_cast<uint64_t*>(b.scratch.aggregates() + uint64_t(node) * 12);
#pragma unroll
    for (int i = 0; i < 6; ++i)
    {
        __hip_atomic_store(q + i, w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// An element climbing at level L > 1: leaf range [a, last] (a leaf when a == last), the deltas of
// its two boundary gaps (d_left = delta(a - 1), d_right = delta(last)) and its summary, all carried
// by the lane, so a climb step only fetches what belongs to the sibling.
struct Climb
{
    uint32_t a;
    uint32_t last;
    int d_left;
    int d_right;
    Summary own;
};

// The element id of [a, last] given its boundary deltas: a leaf, else the internal node numbered
// by its parent gap (a left child carries the index of its right end).
CX_DEV uint32_t ElementOf(uint32_t a, uint32_t last, int d_left, int d_right)
{
    if (a == last)
    {
        return kLtChildLeafBit | a;
    }
    return d_right > d_left ? last : a;
}

// One climb step of a finished element inside the block [block_lo, block_hi). Returns true when the
// element's parent is built (the lane then holds the parent and climbs on); false when the element
// arrived first at its gap, is the root, or has its parent gap on the block boundary (then it is
// appended to the block's tops). A higher level uses the global gap slots, where an earlier level
// may have parked a value.
CX_DEV bool ClimbStep(const Build& b, int n, Climb& e, uint64_t block_lo, uint64_t block_hi, uint2* tops,
                      uint32_t* lds_top_count)
{
    const uint32_t* keys = b.scratch.keys(0);
    if (e.a == 0 && e.last == uint32_t(n - 1))
    {
        return false;  // the root (node 0) is finished
    }
    const bool is_left_child = e.d_right > e.d_left;
    const uint32_t gap = is_left_child ? e.last : e.a - 1;
    if (uint64_t(gap) < block_lo || uint64_t(gap) + 1 >= block_hi)
    {
        const uint32_t k = atomicAdd(lds_top_count, 1u);
        if (k < kTopsCap)
        {
            tops[k] = make_uint2(e.a, e.last);
        }
        return false;
    }
    const uint32_t far_end = is_left_child ? e.a : e.last;
    const uint32_t other = __hip_atomic_exchange(b.scratch.u32(kLtScratchCounter) + gap, far_end, __ATOMIC_ACQ_REL,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
    if (other == kLtNone)
    {
        return false;  // the sibling finishes the parent
    }
    // The parent: range [pa, pb], split `gap`; the sibling's summary and its far boundary delta are
    // the only loads.
    const uint32_t pa = is_left_child ? e.a : other;
    const uint32_t pb = is_left_child ? other : e.last;
    const uint32_t left = pa == gap ? (kLtChildLeafBit | gap) : gap;
    const uint32_t right = pb == gap + 1 ? (kLtChildLeafBit | (gap + 1)) : gap + 1;
    const int d_far = is_left_child ? GapDelta(keys, int(pb), n) : GapDelta(keys, int(pa) - 1, n);
    const Summary sibling = ElementSummary(b, is_left_child ? right : left);
    const int d_left = is_left_child ? e.d_left : d_far;
    const int d_right = is_left_child ? d_far : e.d_right;
    uint32_t* parent_out = b.scratch.u32(kLtScratchParent);
    uint32_t* leaf_parent_out = b.scratch.u32(kLtScratchSortedLeafParent);
    uint32_t node;
    if (pa == 0 && pb == uint32_t(n - 1))
    {
        node = 0;
        parent_out[0] = kLtNone;
    }
    else
    {
        node = d_right > d_left ? pb : pa;
    }
    b.scratch.u32(kLtScratchLeft)[node] = left;
    b.scratch.u32(kLtScratchRight)[node] = right;
    ((left & kLtChildLeafBit) != 0 ? leaf_parent_out : parent_out)[left & ~kLtChildLeafBit] = node;
    ((right & kLtChildLeafBit) != 0 ? leaf_parent_out : parent_out)[right & ~kLtChildLeafBit] = node;
    b.scratch.split()[node] = gap;
    // Range ends lie in this workgroup's block: no other workgroup of this launch touches them.
    int32_t* diff = b.scratch.cover_diff();
    __hip_atomic_fetch_add(&diff[pa], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&diff[pb], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    e.own = is_left_child ? Union(e.own, sibling) : Union(sibling, e.own);
    StoreAggregate(b, node, e.own);
    e.a = pa;
    e.last = pb;
    e.d_left = d_left;
    e.d_right = d_right;
    return true;
}

// ---- level 1 in LDS ------------------------------------------------------------------------------
// A level-1 workgroup keeps everything of its block of kTreeTile sorted leaves in LDS: the keys of
// the block and its two neighbours, the leaf summaries (gathered once, in sorted order), the gap
// slots and cover diffs, and the outputs of the nodes it builds (children, split, parent links,
// aggregates), which it writes out with coalesced stores at the end. The climb is ClimbStep's.
constexpr uint32_t kUnset = 0xFFFFFFFEu;  // a staged output the block did not write

struct Level1Lds
{
    int32_t delta[kTreeTile + 1];  // delta of the gaps [block_lo - 1, block_hi - 1]
    float leaf[kTreeTile][12];     // summaries of the block's sorted leaves
    float agg[kTreeTile][12];      // aggregates of the nodes built here (node - block_lo)
    uint32_t slot[kTreeTile];
    int32_t diff[kTreeTile];
    uint32_t left[kTreeTile];
    uint32_t right[kTreeTile];
    uint32_t split[kTreeTile];
    uint32_t parent[kTreeTile];       // parent of internal node block_lo + k
    uint32_t leaf_parent[kTreeTile];  // parent of sorted leaf block_lo + k
};

CX_DEV int GapDeltaLds(const Level1Lds& l, int g, int lo)
{
    return l.delta[g - (lo - 1)];
}

CX_DEV Summary LdsSummary(const float (&v)[12])
{
    Summary s;
#pragma unroll
    for (int f = 0; f < 6; ++f)
    {
        s.face[f] = v[f];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        s.lo[k] = v[6 + k];
        s.hi[k] = v[9 + k];
    }
    return s;
}

CX_DEV void LdsStore(float (&v)[12], const Summary& s)
{
#pragma unroll
    for (int f = 0; f < 6; ++f)
    {
        v[f] = s.face[f];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        v[6 + k] = s.lo[k];
        v[9 + k] = s.hi[k];
    }
}

CX_DEV Summary Level1Summary(const Level1Lds& l, uint32_t element, uint32_t lo)
{
    return (element & kLtChildLeafBit) != 0 ? LdsSummary(l.leaf[(element & ~kLtChildLeafBit) - lo])
                                            : LdsSummary(l.agg[element - lo]);
}

// ClimbStep for level 1: every arrival is in LDS; the climbing lane carries its element's summary.
CX_DEV bool ClimbStepLds(Level1Lds& l, int n, Climb& e, uint32_t lo, uint32_t hi, uint2* tops,
                         uint32_t* lds_top_count)
{
    const uint32_t a = e.a;
    const uint32_t last = e.last;
    if (a == 0 && last == uint32_t(n - 1))
    {
        return false;  // the root (node 0) is finished
    }
    const bool is_left_child = e.d_right > e.d_left;
    const uint32_t gap = is_left_child ? last : a - 1;
    if (gap < lo || gap + 1 >= hi || (a == 0 && !is_left_child))
    {
        const uint32_t k = atomicAdd(lds_top_count, 1u);
        if (k < kTopsCap)
        {
            tops[k] = make_uint2(a, last);
        }
        return false;
    }
    const uint32_t far_end = is_left_child ? a : last;
    const uint32_t other =
        __hip_atomic_exchange(&l.slot[gap - lo], far_end, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (other == kLtNone)
    {
        return false;  // the sibling finishes the parent
    }
    const uint32_t pa = is_left_child ? a : other;
    const uint32_t pb = is_left_child ? other : last;
    const uint32_t left = pa == gap ? (kLtChildLeafBit | gap) : gap;
    const uint32_t right = pb == gap + 1 ? (kLtChildLeafBit | (gap + 1)) : gap + 1;
    // The sibling's far boundary delta and summary are the only reads.
    const int d_far = is_left_child ? GapDeltaLds(l, int(pb), int(lo)) : GapDeltaLds(l, int(pa) - 1, int(lo));
    const Summary sibling = Level1Summary(l, is_left_child ? right : left, lo);
    const int d_left = is_left_child ? e.d_left : d_far;
    const int d_right = is_left_child ? d_far : e.d_right;
    uint32_t node;
    if (pa == 0 && pb == uint32_t(n - 1))
    {
        node = 0;
        l.parent[0] = kLtNone;  // the block holds the whole tree (lo = 0)
    }
    else
    {
        node = d_right > d_left ? pb : pa;
    }
    l.left[node - lo] = left;
    l.right[node - lo] = right;
    l.split[node - lo] = gap;
    ((left & kLtChildLeafBit) != 0 ? l.leaf_parent : l.parent)[(left & ~kLtChildLeafBit) - lo] = node;
    ((right & kLtChildLeafBit) != 0 ? l.leaf_parent : l.parent)[(right & ~kLtChildLeafBit) - lo] = node;
    atomicAdd(&l.diff[pa - lo], 1);
    atomicAdd(&l.diff[pb - lo], -1);
    e.own = is_left_child ? Union(e.own, sibling) : Union(sibling, e.own);
    LdsStore(l.agg[node - lo], e.own);
    e.a = pa;
    e.last = pb;
    e.d_left = d_left;
    e.d_right = d_right;
    return true;
}

CX_DEV void TreeLevel1(const Build& b, int n, uint32_t lo, uint32_t hi, uint2* tops, uint32_t* lds_top_count,
                       Level1Lds& l)
{
    const uint32_t t = threadIdx.x;
    const uint32_t* keys = b.scratch.keys(0);
    const uint32_t* values = b.scratch.values(0);
    // delta(g) of the gaps g = lo - 1 .. hi - 1 (-1 outside [0, n - 1)), as GapDelta.
    for (uint32_t k = t; k < kTreeTile + 1; k += blockDim.x)
    {
        const int64_t g = int64_t(lo) - 1 + k;
        int d = -1;
        if (g >= 0 && g + 1 < n)
        {
            const uint32_t k0 = keys[g];
            const uint32_t k1 = keys[g + 1];
            d = k0 == k1 ? 32 + int(Clz(uint32_t(g) ^ uint32_t(g + 1))) : int(Clz(k0 ^ k1));
        }
        l.delta[k] = d;
    }
    for (uint32_t k = t; k < kTreeTile; k += blockDim.x)
    {
        l.slot[k] = kLtNone;
        l.diff[k] = 0;
        l.left[k] = kUnset;
        l.parent[k] = kUnset;
        l.leaf_parent[k] = kUnset;
    }
    // The leaf summaries in sorted order: 3 x 16-byte loads per leaf, all issued up front.
    for (uint32_t k = t; k < kTreeTile; k += blockDim.x)
    {
        if (lo + k < uint32_t(n))
        {
            const float4* src = reinterpret_cast<const float4*>(b.leaves[values[lo + k]].summary.face);
            const float4 p = src[0];
            const float4 q = src[1];
            const float4 r = src[2];
            float* d = l.leaf[k];
            d[0] = p.x, d[1] = p.y, d[2] = p.z, d[3] = p.w;
            d[4] = q.x, d[5] = q.y, d[6] = q.z, d[7] = q.w;
            d[8] = r.x, d[9] = r.y, d[10] = r.z, d[11] = r.w;
        }
    }
    __syncthreads();
    uint32_t next = t;
    Climb e;
    bool live = false;
    while (true)
    {
        if (!live)
        {
            if (next >= kTreeTile || lo + next >= uint32_t(n))
            {
                break;
            }
            e.a = lo + next;
            e.last = e.a;
            e.d_left = l.delta[next];       // gap e.a - 1
            e.d_right = l.delta[next + 1];  // gap e.a
            e.own = LdsSummary(l.leaf[next]);
            next += blockDim.x;
        }
        live = ClimbStepLds(l, n, e, lo, hi, tops, lds_top_count);
    }
    __syncthreads();
    // Coalesced write-out of the block.
    uint32_t* slots = b.scratch.u32(kLtScratchCounter);
    int32_t* diff = b.scratch.cover_diff();
    uint32_t* left_out = b.scratch.u32(kLtScratchLeft);
    uint32_t* right_out = b.scratch.u32(kLtScratchRight);
    uint32_t* split_out = b.scratch.split();
    uint32_t* parent_out = b.scratch.u32(kLtScratchParent);
    uint32_t* leaf_parent_out = b.scratch.u32(kLtScratchSortedLeafParent);
    float* agg_out = b.scratch.aggregates();
    for (uint32_t k = t; k < kTreeTile; k += blockDim.x)
    {
        const uint32_t g = lo + k;
        if (g + 1 < uint32_t(n))
        {
            slots[g] = l.slot[k];
        }
        if (g < uint32_t(n))
        {
            diff[g] = l.diff[k];
        }
        if (l.left[k] != kUnset)
        {
            left_out[g] = l.left[k];
            right_out[g] = l.right[k];
            split_out[g] = l.split[k];
        }
        if (l.parent[k] != kUnset)
        {
            parent_out[g] = l.parent[k];
        }
        if (l.leaf_parent[k] != kUnset)
        {
            leaf_parent_out[g] = l.leaf_parent[k];
        }
    }
    // aggregates: 12 floats per node, as float4 triples over consecutive threads
    for (uint32_t q = t; q < kTreeTile * 3; q += blockDim.x)
    {
        const uint32_t k = q / 3;
        if (l.left[k] != kUnset)
        {
            const float* v = &l.agg[k][(q % 3) * 4];
            reinterpret_cast<float4*>(agg_out + uint64_t(lo + k) * 12)[q % 3] = make_float4(v[0], v[1], v[2], v[3]);
        }
    }
}

// Level 1: the workgroup's own 512 sorted leaves. Level L > 1: the tops of its 32 sub-blocks.
template <class Set>
__global__ void __launch_bounds__(1024) TreeLevelKernel(Set set, uint32_t level)
{
    const Build& b = set.Get();
    if (level > TreeLevels(b.leaf_capacity) || blockIdx.x >= TreeBlocks(b.leaf_capacity, level) || !Building(b))
    {
        return;
    }
    __shared__ uint32_t lds_top_count;
    __shared__ uint32_t lds_offsets[kTreeFanout + 1];
    __shared__ Level1Lds lds_level1;
    const uint32_t t = threadIdx.x;
    const int n = int(b.scratch.ctl()[kCtlLeafCount]);
    const uint64_t tile = TreeTile(level);
    const uint64_t block_lo = uint64_t(blockIdx.x) * tile;
    const uint64_t block_hi = block_lo + tile;
    uint2* tops = b.scratch.tops(level) + uint64_t(blockIdx.x) * kTopsCap;
    if (t == 0)
    {
        lds_top_count = 0;
    }
    __syncthreads();
    if (n >= 2 && block_lo < uint64_t(n))
    {
        if (level == 1)
        {
            TreeLevel1(b, n, uint32_t(block_lo), uint32_t(block_hi), tops, &lds_top_count, lds_level1);
        }
        else
        {
            const uint32_t child_first = blockIdx.x * kTreeFanout;
            const uint32_t child_blocks = TreeBlocks(b.leaf_capacity, level - 1);
            const uint32_t* counts = b.scratch.tops_count(level - 1);
            if (t == 0)
            {
                uint32_t sum = 0;
                for (uint32_t c = 0; c < kTreeFanout; ++c)
                {
                    lds_offsets[c] = sum;
                    const uint32_t child = child_first + c;
                    sum += child < child_blocks ? min(counts[child], kTopsCap) : 0u;
                }
                lds_offsets[kTreeFanout] = sum;
            }
            __syncthreads();
            const uint2* child_tops = b.scratch.tops(level - 1);
            const uint32_t total = lds_offsets[kTreeFanout];
            uint32_t next = t;
            Climb e;
            bool live = false;
            while (true)
            {
                if (!live)
                {
                    if (next >= total)
                    {
                        break;
                    }
                    uint32_t c = 0;
                    while (lds_offsets[c + 1] <= next)
                    {
                        ++c;
                    }
                    const uint2 top = child_tops[uint64_t(child_first + c) * kTopsCap + (next - lds_offsets[c])];
                    const uint32_t* keys = b.scratch.keys(0);
                    e.a = top.x;
                    e.last = top.y;
                    e.d_left = GapDelta(keys, int(e.a) - 1, n);
                    e.d_right = GapDelta(keys, int(e.last), n);
                    e.own = ElementSummary(b, ElementOf(e.a, e.last, e.d_left, e.d_right));
                    next += blockDim.x;
                }
                live = ClimbStep(b, n, e, block_lo, block_hi, tops, &lds_top_count);
            }
        }
    }
    __syncthreads();
    if (t == 0)
    {
        b.scratch.tops_count(level)[blockIdx.x] = lds_top_count;
        if (lds_top_count > kTopsCap)
        {
            b.state->pad0[0] = 1;  // never expected: tops are bounded by twice the tree depth
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Single-pass scans with decoupled look-back over 4096-element tiles (64-bit tile status).
// Returns the exclusive prefix of the tile (sum of all earlier tiles), publishing `tile_sum`.
CX_DEV int32_t ScanLookBack(uint64_t* status, uint32_t tile, int32_t tile_sum, int32_t* lds_prefix)
{
    if (threadIdx.x < kWave)
    {
        const uint32_t lane = threadIdx.x;
        int32_t prefix = 0;
        if (tile == 0)
        {
            if (lane == 0)
            {
                StoreRelaxed(&status[0], kScanPrefix | uint32_t(tile_sum));
            }
        }
        else
        {
            if (lane == 0)
            {
                StoreRelaxed(&status[tile], kScanAggregate | uint32_t(tile_sum));
            }
            // The wave reads 64 predecessors per round trip: lane j reads tile p - j. Aggregates up
            // to the first prefix are summed; a not yet published tile ends the round early.
            int64_t p = int64_t(tile) - 1;
            while (true)
            {
                const int64_t q = p - int64_t(lane);
                const uint64_t word = q >= 0 ? LoadRelaxed(&status[q]) : kScanPrefix;
                const uint64_t flag = word & kScanFlagMask;
                const uint64_t waiting = Ballot(flag == 0);
                const uint64_t done = Ballot(flag == kScanPrefix);
                const uint64_t stop = waiting | done;
                const uint32_t first = stop != 0 ? uint32_t(__ffsll(static_cast<unsigned long long>(stop)) - 1) : kWave;
                const bool finished = first < kWave && ((done >> first) & 1u) != 0;
                const bool take = lane < first || (finished && lane == first);
                prefix += WaveSum(take ? int32_t(uint32_t(word)) : 0);
                if (finished)
                {
                    break;
                }
                p -= first;
                if (first < kWave)
                {
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            if (lane == 0)
            {
                StoreRelaxed(&status[tile], kScanPrefix | uint32_t(prefix + tile_sum));
            }
        }
        if (lane == 0)
        {
            *lds_prefix = prefix;
        }
    }
    __syncthreads();
    return *lds_prefix;
}

// Tile scan. Item k of thread t is tile element k * kScanThreads + t (striped, so every load and
// store of a wave is one contiguous run); the scan transposes through LDS to thread-blocked order
// (thread t: elements [t * 16, t * 16 + 16)), scans there and transposes back. One pad word per 16
// keeps both orders free of bank conflicts. Converts `v` into exclusive prefixes within the tile and
// returns the tile total.
constexpr uint32_t kScanLdsWords = kScanTile + kScanTile / 16;

CX_DEV uint32_t ScanPad(uint32_t j)
{
    return j + (j >> 4);
}

CX_DEV int32_t TileExclusiveScan(int32_t (&v)[kScanItems], int32_t* lds_items, int32_t* lds_waves)
{
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k)
    {
        lds_items[ScanPad(k * kScanThreads + t)] = v[k];
    }
    __syncthreads();
    int32_t w[kScanItems];
    int32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k)
    {
        const int32_t x = lds_items[ScanPad(t * kScanItems + k)];
        w[k] = sum;
        sum += x;
    }
    int32_t total;
    const int32_t before = BlockExclusiveSum<kScanThreads>(sum, lds_waves, total);
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k)
    {
        lds_items[ScanPad(t * kScanItems + k)] = w[k] + before;  // this thread's own slots
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k)
    {
        v[k] = lds_items[ScanPad(k * kScanThreads + t)];
    }
    return total;
}

// cover[g] = sum of diff[0 .. g] for the gaps g < n - 1.
template <class Set>
__global__ void __launch_bounds__(kScanThreads) CoverScanKernel(Set set)
{
    const Build& b = set.Get();
    if (blockIdx.x >= b.scratch.scan_tiles() || !Building(b))
    {
        return;
    }
    __shared__ uint32_t lds_tile;
    __shared__ int32_t lds_waves[kScanThreads / kWave];
    __shared__ int32_t lds_prefix;
    __shared__ int32_t lds_items[kScanLdsWords];
    uint32_t* ctl = b.scratch.ctl();
    if (threadIdx.x == 0)
    {
        lds_tile = AtomicAddRelaxed(&ctl[kCtlCoverTiles], 1u);
    }
    __syncthreads();
    const uint32_t tile = lds_tile;
    const uint32_t n = ctl[kCtlLeafCount];
    const uint32_t gaps = n > 0 ? n - 1 : 0;
    const uint32_t first = tile * kScanTile;
    if (first >= gaps)
    {
        return;
    }
    const int32_t* diff = b.scratch.cover_diff();
    int32_t* cover = b.scratch.cover();
    int32_t v[kScanItems];
    const uint32_t base = first + threadIdx.x;  // item k: base + k * kScanThreads
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k)
    {
        const uint32_t g = base + k * kScanThreads;
        v[k] = g < gaps ? diff[g] : 0;
    }
    int32_t own[kScanItems];
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k)
    {
        own[k] = v[k];
    }
    const int32_t total = TileExclusiveScan(v, lds_items, lds_waves);
    const int32_t prefix = ScanLookBack(b.scratch.cover_status(), tile, total, &lds_prefix);
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k)
    {
        const uint32_t g = base + k * kScanThreads;
        if (g < gaps)
        {
            cover[g] = prefix + v[k] + own[k];
        }
    }
}

// Depth, even-depth flags and their exclusive scan (the reference's depth kernel + scan over the
// leaf capacity, caelix_light_bvh_depth.hlsl:7-30), then the header in the last tile
// (caelix_light_bvh_finish.hlsl:8-63).
template <class Set>
__global__ void __launch_bounds__(kScanThreads) FlagsScanKernel(Set set)
{
    const Build& b = set.Get();
    if (blockIdx.x >= b.scratch.scan_tiles() || !Building(b))
    {
        return;
    }
    __shared__ uint32_t lds_tile;
    __shared__ int32_t lds_waves[kScanThreads / kWave];
    __shared__ int32_t lds_prefix;
    __shared__ int32_t lds_items[kScanLdsWords];
    uint32_t* ctl = b.scratch.ctl();
    if (threadIdx.x == 0)
    {
        lds_tile = AtomicAddRelaxed(&ctl[kCtlFlagsTiles], 1u);
    }
    __syncthreads();
    const uint32_t tile = lds_tile;
    const uint32_t n = ctl[kCtlLeafCount];
    const uint32_t capacity = b.leaf_capacity;
    const uint32_t first = tile * kScanTile;
    const int32_t* cover = b.scratch.cover();
    const uint32_t* split = b.scratch.split();
    uint32_t* counter = b.scratch.u32(kLtScratchCounter);
    uint32_t* wide = b.scratch.u32(kLtScratchWide);
    // All loads of the thread's items first (splits, then the covers they index), stores after: a
    // store between them would order every later load behind it.
    const uint32_t base = first + threadIdx.x;  // item k: base + k * kScanThreads
    uint32_t at[kScanItems];
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k)
    {
        const uint32_t i = base + k * kScanThreads;
        at[k] = i + 1 < n ? split[i] : 0u;  // gap 0 stands in for items past the gaps
    }
    uint32_t depth[kScanItems];
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k)
    {
        const int32_t c = cover[at[k]];
        depth[k] = base + k * kScanThreads + 1 < n ? uint32_t(c - 1) : 1u;
    }
    int32_t v[kScanItems];
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k)
    {
        v[k] = (depth[k] & 1u) == 0 ? 1 : 0;
    }
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k)
    {
        const uint32_t i = base + k * kScanThreads;
        if (i + 1 < n)
        {
            counter[i] = depth[k];
        }
    }
    const int32_t total = TileExclusiveScan(v, lds_items, lds_waves);
    const int32_t prefix = ScanLookBack(b.scratch.flags_status(), tile, total, &lds_prefix);
    // The binary node of each 4-wide node. Slot w <= item index, and every earlier tile has read its
    // split gaps before publishing, so these stores only replace consumed split gaps.
    uint32_t* source = b.scratch.wide_source();
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k)
    {
        const uint32_t i = base + k * kScanThreads;
        if (i < capacity)
        {
            wide[i] = uint32_t(prefix + v[k]);
        }
        if (i + 1 < n && (depth[k] & 1u) == 0)
        {
            source[prefix + v[k]] = i;
        }
    }
    if (tile + 1 != b.scratch.scan_tiles() || threadIdx.x != 0)
    {
        return;
    }
    // Finish: the header. n = 1: one node with the leaf in slot 0. n >= 2: the binary root's summary;
    // more 4-wide nodes than the capacity leave the tree empty with the overflow flag.
    const uint32_t node_total = uint32_t(prefix + total);
    ctl[kCtlWideNodes] = n >= 2 ? node_total : 0u;
    Summary root = EmptySummary();
    uint32_t root_index = kLtNone;
    uint32_t node_count = 0;
    uint32_t overflow = 0;
    if (n == 1)
    {
        LightBvhLeaf* leaf = b.leaves;
        root = LoadSummary(leaf->summary.face);
        LightBvhNode* node = b.nodes;
        uint4* zero = reinterpret_cast<uint4*>(node);
        for (int q = 0; q < 12; ++q)
        {
            zero[q] = make_uint4(0, 0, 0, 0);
        }
        for (int f = 0; f < 6; ++f)
        {
            node->child_face[0][f] = root.face[f];
        }
        for (int k = 0; k < 3; ++k)
        {
            node->child_lo[0][k] = root.lo[k];
            node->child_hi[0][k] = root.hi[k];
        }
        node->child[0] = kLtChildLeafBit | 0u;
        node->child[1] = kLtNone;
        node->child[2] = kLtNone;
        node->child[3] = kLtNone;
        node->parent = kLtNone;
        node->parent_slot = 0;
        leaf->parent = 0;
        root_index = 0;
        node_count = 1;
    }
    else if (n >= 2)
    {
        node_count = node_total;
        if (node_count > b.node_capacity)
        {
            overflow = 1;
        }
        else
        {
            root = LoadSummary(b.scratch.aggregates());
            root_index = 0;
        }
    }
    StoreSummary(b.header->summary.face, root);
    b.header->root = root_index;
    b.header->leaf_count = overflow != 0 ? 0u : n;
    b.header->node_count = node_count;
    b.header->pad = 0;
    b.state->node_overflow = overflow;
    b.state->builds = b.state->builds + 1;
}

// ------------------------------------------------------------------------------------------------
// wide (caelix_light_bvh_wide.hlsl:10-100), one lane per 4-wide node: the reference's lanes are the
// binary nodes, of which the odd-depth half returns at once; here the flags scan lists the
// even-depth ones, so every lane builds a node.
template <class Set>
__global__ void __launch_bounds__(kBvhThreads) WideKernel(Set set)
{
    const Build& b = set.Get();
    const uint32_t w = blockIdx.x * kBvhThreads + threadIdx.x;
    if (w + 1 >= b.leaf_capacity || !Building(b))
    {
        return;
    }
    if (w >= b.scratch.ctl()[kCtlWideNodes] || w >= b.node_capacity)
    {
        return;
    }
    const uint32_t i = b.scratch.wide_source()[w];  // the binary node (an even-depth one)
    const uint32_t* wide = b.scratch.u32(kLtScratchWide);
    const uint32_t* left = b.scratch.u32(kLtScratchLeft);
    const uint32_t* right = b.scratch.u32(kLtScratchRight);
    const uint32_t* values = b.scratch.values(0);
    // Every load is issued before the first store, in four dependent rounds: the children, the
    // grandchildren, the leaf ids / wide indices of the slots, the slot summaries.
    const uint32_t children[2] = {left[i], right[i]};
    uint32_t grand[4];
#pragma unroll
    for (int c = 0; c < 2; ++c)
    {
        const bool inner = (children[c] & kLtChildLeafBit) == 0;
        const uint32_t at = inner ? children[c] : i;  // node i stands in for a leaf child
        const uint32_t l = left[at];
        const uint32_t r = right[at];
        grand[2 * c] = inner ? l : children[c];
        grand[2 * c + 1] = inner ? r : kLtNone;
    }
    // The slots in order: a leaf child takes one, an internal child its two children.
    const bool inner0 = (children[0] & kLtChildLeafBit) == 0;
    const bool inner1 = (children[1] & kLtChildLeafBit) == 0;
    const uint32_t count = 2u + uint32_t(inner0) + uint32_t(inner1);
    const uint32_t slots[4] = {grand[0], inner0 ? grand[1] : grand[2], inner0 ? grand[2] : grand[3],
                               inner0 ? grand[3] : kLtNone};
    // Leaf ids and wide indices, loaded unconditionally (index 0 stands in for the other kind).
    uint32_t child_words[4] = {kLtNone, kLtNone, kLtNone, kLtNone};
#pragma unroll
    for (uint32_t s = 0; s < 4; ++s)
    {
        const bool used = s < count;
        const bool leaf = (slots[s] & kLtChildLeafBit) != 0;
        const uint32_t leaf_id = values[used && leaf ? slots[s] & ~kLtChildLeafBit : 0u];
        const uint32_t wide_id = wide[used && !leaf ? slots[s] : 0u];
        if (used)
        {
            child_words[s] = leaf ? kLtChildLeafBit | leaf_id : wide_id;
        }
    }
    float faces[24];
    float bounds[24];
#pragma unroll
    for (uint32_t s = 0; s < 4; ++s)
    {
        const bool used = s < count;
        const bool leaf = (slots[s] & kLtChildLeafBit) != 0;
        const float* p = used && leaf ? b.leaves[child_words[s] & ~kLtChildLeafBit].summary.face
                                      : b.scratch.aggregates() + (used ? uint64_t(slots[s]) * 12 : 0u);
        Summary summary = LoadSummary(p);
        if (!used)
        {
            summary = EmptySummary();
        }
#pragma unroll
        for (int f = 0; f < 6; ++f)
        {
            faces[6 * s + f] = summary.face[f];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            bounds[3 * s + k] = summary.lo[k];
            bounds[12 + 3 * s + k] = summary.hi[k];
        }
    }
    LightBvhNode* node = b.nodes + w;
#pragma unroll
    for (uint32_t s = 0; s < 4; ++s)
    {
        if (s < count)
        {
            if ((slots[s] & kLtChildLeafBit) != 0)
            {
                b.leaves[child_words[s] & ~kLtChildLeafBit].parent = (w << kLtParentSlotBits) | s;
            }
            else if (child_words[s] < b.node_capacity)
            {
                *reinterpret_cast<uint2*>(&b.nodes[child_words[s]].parent) = make_uint2(w, s);
            }
        }
    }
    float4* face_out = reinterpret_cast<float4*>(node->child_face[0]);
    float4* bound_out = reinterpret_cast<float4*>(node->child_lo[0]);
#pragma unroll
    for (int q = 0; q < 6; ++q)
    {
        face_out[q] = make_float4(faces[4 * q], faces[4 * q + 1], faces[4 * q + 2], faces[4 * q + 3]);
        bound_out[q] = make_float4(bounds[4 * q], bounds[4 * q + 1], bounds[4 * q + 2], bounds[4 * q + 3]);
    }
    *reinterpret_cast<uint4*>(node->child) = make_uint4(child_words[0], child_words[1], child_words[2], child_words[3]);
    if (i == 0)
    {
        *reinterpret_cast<uint2*>(&node->parent) = make_uint2(kLtNone, 0u);
    }
}

// caelix_light_bvh_begin.hlsl:9-17
// One onesweep pass of the build's sort (keys / values ping-pong in the scratch).
template <class Set>
__global__ void __launch_bounds__(kSortThreads, kSortWavesPerSimd) BuildSortPassKernel(Set set, uint32_t pass)
{
    const Build& b = set.Get();
    if (blockIdx.x >= SortTiles(b.leaf_capacity))
    {
        return;
    }
    const uint32_t in = pass & 1u;
    SortPassTile(b.scratch.keys(in), b.scratch.values(in), b.scratch.keys(1 - in), b.scratch.values(1 - in),
                 &b.header->leaf_count, &b.state->building, b.scratch.sort(), pass);
}

// One pass of the 64-bit key variant's sort: the keys (code << 32 | leaf) start in sort buffer 1,
// pass p reads buffer 1 - p % 2 and writes the other; the last pass writes the sorted codes and
// leaves as the tree kernels read them (keys(0), values(0)).
template <class Set>
__global__ void __launch_bounds__(SortKeys64::kThreads, kSortWavesPerSimd) BuildSort64PassKernel(Set set,
                                                                                                   uint32_t pass)
{
    const Build& b = set.Get();
    if (blockIdx.x >= SortTilesOf<SortKeys64>(b.leaf_capacity))
    {
        return;
    }
    const uint32_t in = 1u - (pass & 1u);
    const bool last = pass + 1 == uint32_t(SortKeys64::kPasses);
    const SortSink<SortKeys64> out{last ? nullptr : b.scratch.keys64(1u - in), nullptr, b.scratch.keys(0),
                                   b.scratch.values(0)};
    SortPassTileOf<SortKeys64>(b.scratch.keys64(in), nullptr, out, &b.header->leaf_count, &b.state->building,
                               b.scratch.sort64(), pass);
}

// The build sequence (light_tree_record.cpp:271-308) for a set of builds; `capacity` = the largest
// leaf capacity of the set, `count` = the number of builds (grid y). Keys = SortPairs32 (the
// reference's uint32 Morton keys with leaf values) or SortKeys64 (code << 32 | leaf, sorted as one
// 64-bit key); both give the same order, hence the same tree.
template <class Keys, class Set>
void RecordBuild(const Set& set, uint32_t capacity, uint32_t count, hipStream_t s)
{
    const uint32_t per_item_blocks = DivUp(capacity, kBvhThreads);
    const uint32_t tiles = SortTiles(capacity);
    {
        StageScope scope(s, "lt_bounds");
        BoundsKernel<Set, Keys><<<dim3(BoundBlocks(capacity), count), kBvhThreads, 0, s>>>(set);
    }
    {
        StageScope scope(s, "lt_morton");
        MortonKernel<Set, Keys><<<dim3(tiles, count), kSortThreads, 0, s>>>(set);
    }
    if constexpr (Keys::kValues)
    {
        for (uint32_t pass = 0; pass < kSortPasses; ++pass)
        {
            StageScope scope(s, "sort_pass");
            BuildSortPassKernel<<<dim3(tiles, count), kSortThreads, 0, s>>>(set, pass);
        }
    }
    else
    {
        for (uint32_t pass = 0; pass < uint32_t(SortKeys64::kPasses); ++pass)
        {
            StageScope scope(s, "sort64_pass");
            BuildSort64PassKernel<<<dim3(SortTilesOf<SortKeys64>(capacity), count), SortKeys64::kThreads, 0, s>>>(
                set, pass);
        }
    }
    static const char* const kTreeStage[] = {"lt_tree_l1", "lt_tree_l2", "lt_tree_l3", "lt_tree_l4", "lt_tree_l5",
                                             "lt_tree_l6"};
    for (uint32_t level = 1; level <= TreeLevels(capacity); ++level)
    {
        StageScope scope(s, kTreeStage[std::min(level, 6u) - 1]);
        const uint32_t threads = level == 1 ? kTreeThreads : 1024u;
        TreeLevelKernel<<<dim3(TreeBlocks(capacity, level), count), threads, 0, s>>>(set, level);
    }
    const uint32_t scan_tiles = DivUp(capacity, kScanTile);
    {
        StageScope scope(s, "lt_cover_scan");
        CoverScanKernel<<<dim3(scan_tiles, count), kScanThreads, 0, s>>>(set);
    }
    {
        StageScope scope(s, "lt_flags_scan");
        FlagsScanKernel<<<dim3(scan_tiles, count), kScanThreads, 0, s>>>(set);
    }
    {
        StageScope scope(s, "lt_wide");
        WideKernel<<<dim3(per_item_blocks, count), kBvhThreads, 0, s>>>(set);
    }
}

__global__ void BeginKernel(const LightControl* control, LightTreeState* state, uint32_t tree, uint32_t reason,
                            uint32_t pass_recorded)
{
    const bool committed = pass_recorded != 0 && control != nullptr && control->pass_result == kLtPassOk &&
                           ((control->trees_changed >> tree) & 1u) != 0;
    state->building = (reason != 0 || committed) ? 1u : 0u;
}

__global__ void ScratchInitKernel(uint32_t* ctl)
{
    if (threadIdx.x < kCtlWords)
    {
        ctl[threadIdx.x] = 0;
    }
}

}  // namespace
}  // namespace cx

namespace cx {
// Validates a build description and converts it (argument errors log and return, changing nothing).
int CheckBuild(const CxLightBvhBuild* build, const char* what, Build& b)
{
    if (build == nullptr || build->leaves == nullptr || build->nodes == nullptr || build->header == nullptr ||
        build->state == nullptr || build->scratch == nullptr)
    {
        return ArgError("%s: a buffer is NULL", what);
    }
    if (build->leaf_capacity == 0 || build->leaf_capacity % 64 != 0 || build->leaf_capacity >= (1u << 30))
    {
        return ArgError("%s: leaf capacity %u is not a positive multiple of 64 below 2^30", what, build->leaf_capacity);
    }
    if (build->node_capacity == 0)
    {
        return ArgError("%s: node capacity 0", what);
    }
    b.leaves = build->leaves;
    b.nodes = build->nodes;
    b.header = build->header;
    b.state = build->state;
    b.scratch = Scratch{static_cast<uint8_t*>(build->scratch), build->leaf_capacity};
    b.leaf_capacity = build->leaf_capacity;
    b.node_capacity = build->node_capacity;
    return CX_OK;
}
}  // namespace cx

extern "C" {

CX_API uint64_t cx_light_bvh_scratch_bytes(uint32_t leaf_capacity)
{
    return cx::Scratch::Bytes(leaf_capacity);
}

CX_API int cx_light_bvh_scratch_init(void* scratch, uint32_t leaf_capacity, void* stream)
{
    if (scratch == nullptr || leaf_capacity == 0 || leaf_capacity % 64 != 0)
    {
        return cx::ArgError("cx_light_bvh_scratch_init: NULL scratch or a leaf capacity that is not a multiple of 64");
    }
    const cx::Scratch s{static_cast<uint8_t*>(scratch), leaf_capacity};
    cx::ScratchInitKernel<<<1, 64, 0, static_cast<hipStream_t>(stream)>>>(s.ctl());
    return cx::LaunchResult("cx_light_bvh_scratch_init");
}

CX_API int cx_light_bvh_begin(LightControl* control, LightTreeState* state, uint32_t tree, uint32_t reason,
                              uint32_t pass_recorded, void* stream)
{
    if (state == nullptr || tree >= kLtTrees || (pass_recorded != 0 && control == nullptr))
    {
        return cx::ArgError("cx_light_bvh_begin: NULL state, tree >= 2, or a recorded pass without control");
    }
    cx::BeginKernel<<<1, 1, 0, static_cast<hipStream_t>(stream)>>>(control, state, tree, reason, pass_recorded);
    return cx::LaunchResult("cx_light_bvh_begin");
}

CX_API int cx_light_bvh_build(const CxLightBvhBuild* build, void* stream)
{
    cx::Build b;
    if (int r = cx::CheckBuild(build, "cx_light_bvh_build", b); r != CX_OK)
    {
        return r;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::RecordBuild<cx::SortPairs32>(cx::OneBuild{b}, build->leaf_capacity, 1, s);
    return cx::LaunchResult("cx_light_bvh_build");
}

CX_API int cx_light_bvh_build_keys64(const CxLightBvhBuild* build, void* stream)
{
    cx::Build b;
    if (int r = cx::CheckBuild(build, "cx_light_bvh_build_keys64", b); r != CX_OK)
    {
        return r;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::RecordBuild<cx::SortKeys64>(cx::OneBuild{b}, build->leaf_capacity, 1, s);
    return cx::LaunchResult("cx_light_bvh_build_keys64");
}

CX_API uint64_t cx_light_bvh_batch_bytes(uint32_t count)
{
    return uint64_t(count) * sizeof(cx::Build);
}

CX_API int cx_light_bvh_batch_prepare(const CxLightBvhBuild* builds, uint32_t count, void* descriptors, void* stream)
{
    if (builds == nullptr || descriptors == nullptr || count == 0)
    {
        return cx::ArgError("cx_light_bvh_batch_prepare: no builds or no descriptor buffer");
    }
    std::vector<cx::Build> host(count);
    for (uint32_t i = 0; i < count; ++i)
    {
        if (int r = cx::CheckBuild(&builds[i], "cx_light_bvh_batch_prepare", host[i]); r != CX_OK)
        {
            return r;
        }
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    return cx::HipResult(hipMemcpyAsync(descriptors, host.data(), host.size() * sizeof(cx::Build), hipMemcpyHostToDevice, s),
                         "cx_light_bvh_batch_prepare");
}

CX_API int cx_light_bvh_build_batch(const void* descriptors, uint32_t count, uint32_t max_leaf_capacity, void* stream)
{
    if (descriptors == nullptr || count == 0 || count > 65535)
    {
        return cx::ArgError("cx_light_bvh_build_batch: no descriptors or %u builds (1..65535)", count);
    }
    if (max_leaf_capacity == 0 || max_leaf_capacity % 64 != 0 || max_leaf_capacity >= (1u << 30))
    {
        return cx::ArgError("cx_light_bvh_build_batch: largest leaf capacity %u invalid", max_leaf_capacity);
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::RecordBuild<cx::SortPairs32>(cx::BuildArray{static_cast<const cx::Build*>(descriptors)}, max_leaf_capacity,
                                     count, s);
    return cx::LaunchResult("cx_light_bvh_build_batch");
}

}  // extern "C"
