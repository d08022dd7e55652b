// Light BVH build for gfx950 (the build half of LightTree::RecordTree,
// lib/src/light/light_tree_record.cpp:271-308; kernels shaders/light/bvh/*.hlsl).
//
// Launch sequence (one stream; 4 sort passes, or 7 with 64-bit keys):
//   bounds     grid-stride min/max of the leaf box centres into per-workgroup partials; writes the
//              centres (12 B per leaf, read again by the Morton kernel instead of the 64-B leaves);
//              workgroup 0 resets the build's counters and the sort histograms
//   morton     reduces the partials (leaf_count, leaves_dropped, bounds), 30-bit key + leaf index per
//              leaf, the onesweep sort's digit histograms and tile status rows
//   sort       onesweep passes (radix_sort.h); sorted pairs end in sort buffer 0
//   tree L1    per block of 512 sorted leaves, in LDS: the binary radix tree as the reference builds
//              it (Karras 2012 numbering) and its bottom-up aggregation, by one climb from the leaves;
//              per node its children, parent links, +1/-1 at the two ends of its range in the
//              gap-cover array, aggregate and 4-wide slot list; the block's open elements (whose
//              parent lies outside it) go to its open list
//   tree L2+   per block of 32 sub-blocks, in LDS: the same climb from the sub-blocks' open elements
//   cover      inclusive scan of the gap-cover array (decoupled look-back): cover[g] = the number of
//              binary nodes whose range spans the gap between sorted leaves g and g + 1
//   flags      depth(i) = cover[split(i)] - 1, split(i) = node i's left child index (the nodes spanning
//              node i's split gap are i and its ancestors), even-depth flags, their exclusive scan =
//              4-wide node indices (where it steps: the 4-wide nodes), and the
//              header (the reference's caelix_light_bvh_finish) in the last tile
//   wide       4-wide nodes from the even-depth binary nodes
// The reference computes the depth by walking every node up to the root (O(n * depth) dependent
// loads, caelix_light_bvh_depth.hlsl:7-30); the gap cover gives the same depths with two atomics per
// node and one scan.
#include "device_common.h"
#include "radix_sort.h"

#include <mutex>
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
    kCtlTreeOverflow = 5,  // nonzero: an open list or an upper level's LDS tables overflowed
    kCtlWords = 64,
};

// The bottom-up tree build runs in levels of nested leaf blocks: a level-1 workgroup owns 512
// sorted leaves (and works in LDS), a level-L workgroup the 32 level-(L-1) blocks below it.
constexpr uint32_t kTreeThreads = 512;  // level 1: one lane per leaf
constexpr uint32_t kTreeTile = 512;
constexpr uint32_t kTreeFanout = 32;
constexpr uint32_t kOpenWords = 18;  // one open element: range, boundary deltas, child words, summary
// The bounds every work list below is sized from. Along a root-to-leaf path the deltas of the nodes
// strictly increase; a delta is clz(k0 ^ k1) in 2..31 for two distinct 30-bit codes, or 32 + clz(i ^ j)
// in 34..63 for equal codes at sorted positions below 2^30. So no path holds more than kDeltaLevels
// internal nodes, and:
//  - the elements of a block whose parent lies outside it hang off the nodes that cross one of the
//    block's two ends, at most one per such node: at most 2 * kDeltaLevels open elements;
//  - the nodes an upper-level block builds cross one of the 31 boundaries between its sub-blocks; the
//    nodes that cross one boundary lie on one path: at most 31 * kDeltaLevels of them; each is
//    reached by its two children, and each open element that parks takes one more arrival.
constexpr uint32_t kDeltaLevels = 60;
constexpr uint32_t kOpenCap = 256;   // open elements per block: >= 2 * kDeltaLevels
constexpr uint32_t kMaxClimb = 128;  // climb steps per element and level: > kDeltaLevels
constexpr uint32_t kUpperBuiltBound = (kTreeFanout - 1) * kDeltaLevels;
constexpr uint32_t kUpperArrivalBound = 2 * kUpperBuiltBound + 2 * kDeltaLevels;  // 3840
static_assert(kOpenCap >= 2 * kDeltaLevels && kMaxClimb > kDeltaLevels, "work lists below the depth bound");

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
__host__ __device__ inline uint64_t OpenBytes(uint32_t capacity, uint32_t level)
{
    const uint64_t blocks = TreeBlocks(capacity, level);
    return ((blocks * kOpenCap * kOpenWords * 4 + blocks * 4) + 255) & ~uint64_t(255);
}

// An upper-level block keeps its first kPoolEntries element records in LDS and the rest, up to the
// arrival bound, in its slice of a global spill area (shared by the levels, which run one after
// another; sized for the level-2 grid, the largest).
constexpr uint32_t kPoolEntries = 1536;
constexpr uint32_t kSpillEntries = kUpperArrivalBound - kPoolEntries;
__host__ __device__ inline uint64_t SpillBytes(uint32_t capacity)
{
    return TreeLevels(capacity) >= 2 ? uint64_t(TreeBlocks(capacity, 2)) * kSpillEntries * kOpenWords * 4 : 0u;
}

// The sort layouts of the build: the 64-bit key variant and the tunable 32-bit key layouts (digit
// bits x tile elements, cx_light_bvh_set_tuning).
using SortMorton6x4096 = SortMorton<6, 8>;
using SortMorton6x8192 = SortMorton<6, 16>;
using SortMorton8x4096 = SortMorton<8, 8>;
using SortMorton8x8192 = SortMorton<8, 16>;

// The sort state region: the largest of the layouts the build may use.
__host__ __device__ inline uint64_t BuildSortStateBytes(uint32_t capacity)
{
    const uint64_t sizes[] = {SortStateBytesOf<SortKeys64>(capacity), SortStateBytesOf<SortMorton6x4096>(capacity),
                              SortStateBytesOf<SortMorton6x8192>(capacity), SortStateBytesOf<SortMorton8x4096>(capacity),
                              SortStateBytesOf<SortMorton8x8192>(capacity)};
    uint64_t m = 0;
    for (uint64_t v : sizes)
    {
        m = v > m ? v : m;
    }
    return m;
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
    // holds the cover counts (keys).
    __host__ __device__ float* centres() const { return aggregates(); }
    __host__ __device__ int32_t* cover_diff() const { return reinterpret_cast<int32_t*>(u32(kLtScratchWide)); }
    __host__ __device__ int32_t* cover() const { return reinterpret_cast<int32_t*>(keys(1)); }
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
    // The wide collapse's input, written while the binary tree is built: per binary node the 4 slots
    // of the 4-wide node it becomes if its depth is even (`quad`: the child words of its grandchildren,
    // a leaf child as kLtChildLeafBit | leaf id in place of its absent children, kLtNone padding).
    // 16 bytes per leaf of capacity, at a fixed offset (no per-level layout walk).
    __host__ __device__ uint8_t* wide_inputs_base() const
    {
        const uint64_t offset = uint64_t(sort_base() - base) + BuildSortStateBytes(capacity);
        return base + ((offset + 255) & ~uint64_t(255));
    }
    __host__ __device__ uint4* quad() const { return reinterpret_cast<uint4*>(wide_inputs_base()); }
    // Open lists of the tree levels: per level-L block, up to kOpenCap elements whose parent the
    // block did not build (kOpenWords words each), and their count.
    __host__ __device__ uint8_t* open_base() const { return reinterpret_cast<uint8_t*>(quad() + capacity); }
    __host__ __device__ uint32_t* open(uint32_t level) const
    {
        uint8_t* p = open_base();
        for (uint32_t l = 1; l < level; ++l)
        {
            p += OpenBytes(capacity, l);
        }
        return reinterpret_cast<uint32_t*>(p);
    }
    __host__ __device__ uint32_t* open_count(uint32_t level) const
    {
        return open(level) + uint64_t(TreeBlocks(capacity, level)) * kOpenCap * kOpenWords;
    }
    // The upper levels' spill area (after the open lists of every level).
    __host__ __device__ uint32_t* spill() const { return open(TreeLevels(capacity) + 1); }
    static uint64_t Bytes(uint32_t capacity)
    {
        const Scratch s{nullptr, capacity};
        return uint64_t(reinterpret_cast<uint8_t*>(s.spill()) - static_cast<uint8_t*>(nullptr)) + SpillBytes(capacity);
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
        // selects, not lo[t] / hi[t - 3]: a per-lane index would put both arrays in scratch memory
        const float a = t == 0 ? lo[0] : (t == 1 ? lo[1] : lo[2]);
        const float z = t == 3 ? hi[0] : (t == 4 ? hi[1] : hi[2]);
        partials[blockIdx.x * 8 + t] = t < 3 ? a : z;
    }
    if (blockIdx.x == 0)
    {
        uint32_t* ctl = b.scratch.ctl();
        if (t == 0)
        {
            ctl[kCtlCoverTiles] = 0;
            ctl[kCtlFlagsTiles] = 0;
            ctl[kCtlTreeOverflow] = 0;
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
// (code, leaf) pairs in sort buffer passes % 2 (SortMorton), or code << 32 | leaf in sort buffer 1
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
    const float* centres = b.scratch.centres();
    const uint32_t base = blockIdx.x * kSortTile;
    // The centres of the lane's items, all loaded before any store (a store between them would
    // order the later loads behind it), and before the partials are reduced.
    float cv[kSortItems][3];
#pragma unroll
    for (int k = 0; k < kSortItems; ++k)
    {
        const uint32_t i = min(base + k * kSortThreads + t, n - 1);  // past n: unused
        cv[k][0] = n > 0 ? centres[3 * uint64_t(i)] : 0.0f;
        cv[k][1] = n > 0 ? centres[3 * uint64_t(i) + 1] : 0.0f;
        cv[k][2] = n > 0 ? centres[3 * uint64_t(i) + 2] : 0.0f;
    }
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
    // the pairs start in the buffer that an even number of passes leaves them in buffer 0 from
    uint32_t* keys = b.scratch.keys(uint32_t(Keys::kPasses) & 1u);
    uint32_t* values = b.scratch.values(uint32_t(Keys::kPasses) & 1u);
    uint32_t* counter = b.scratch.u32(kLtScratchCounter);
    int32_t* diff = b.scratch.cover_diff();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k)
    {
        const uint32_t i = base + k * kSortThreads + t;
        if (i >= b.leaf_capacity)
        {
            break;
        }
        // The depths replace counter[i] for i < n - 1 (flags scan) and level 1 writes diff[i] for
        // i < n; past them the arrays are zero.
        if (i + 1 >= n)
        {
            counter[i] = 0u;
            diff[i] = 0;
        }
        if (i < n)
        {
            const float c0 = cv[k][0];
            const float c1 = cv[k][1];
            const float c2 = cv[k][2];
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
// [g + 1, b] is node g + 1, the root is node 0. The first element to arrive at a gap parks and stops;
// the second builds the parent: range, children, parent links, and the union of the two summaries
// (left, then right: caelix_light_bvh_aggregate.hlsl:30-55), then climbs on. Also written per node:
// +1/-1 at its range ends in the gap-cover array (for the depths) and its 4-wide slot list (for the
// wide collapse).
//
// The climb runs in levels of nested blocks of sorted leaves, each level-L workgroup in LDS: every
// arrival at a gap inside the block happens in the block's workgroup, so the arrivals, parked
// elements and exchanges are LDS operations and a climb step makes no global round trip. An element
// whose parent is not built inside its block (its parent gap is on the block boundary, or it parked
// and its sibling reaches outside the block) is "open": the workgroup writes it, whole (range,
// boundary deltas, child words, summary), to its block's open list, and the level above starts from
// the open lists of its 32 sub-blocks. Kernel boundaries order the levels.

CX_DEV uint32_t Clz(uint32_t x)
{
    return x == 0 ? 32u : uint32_t(__clz(int(x)));
}

// An element climbing: leaf range [a, last] (a leaf when a == last), the deltas of its two boundary
// gaps (d_left = delta(a - 1), d_right = delta(last)), its summary, and what it contributes to its
// parent's 4-wide slots: a leaf its own child word (and kLtNone), an internal node its two child
// words. 72 bytes, also the format of an open-list record (kOpenWords words).
struct Climb
{
    uint32_t a;
    uint32_t last;
    int d_left;
    int d_right;
    uint32_t kid0;
    uint32_t kid1;
    Summary own;
};
static_assert(sizeof(Climb) == kOpenWords * 4, "open record layout");

CX_DEV void StoreClimb(uint32_t* dst, const Climb& e)  // 8-byte aligned
{
    uint64_t w[kOpenWords / 2];
    __builtin_memcpy(w, &e, sizeof(e));
#pragma unroll
    for (uint32_t i = 0; i < kOpenWords / 2; ++i)
    {
        reinterpret_cast<uint64_t*>(dst)[i] = w[i];
    }
}

CX_DEV Climb LoadClimb(const uint32_t* src)
{
    uint64_t w[kOpenWords / 2];
#pragma unroll
    for (uint32_t i = 0; i < kOpenWords / 2; ++i)
    {
        w[i] = reinterpret_cast<const uint64_t*>(src)[i];
    }
    Climb e;
    __builtin_memcpy(&e, w, sizeof(e));
    return e;
}

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

// The 4 slots of the 4-wide node made from a binary node whose children contribute (l0, l1) and
// (r0, r1), in order; kLtNone pads.
CX_DEV uint4 QuadOf(uint32_t l0, uint32_t l1, uint32_t r0, uint32_t r1)
{
    return l1 == kLtNone ? make_uint4(l0, r0, r1, kLtNone) : make_uint4(l0, l1, r0, r1);
}

// The parent of the left child `l` and the right child `r` split at `gap`: range, boundary deltas,
// child words and summary (left, then right: caelix_light_bvh_aggregate.hlsl:30-55).
CX_DEV uint32_t WordOf(const Climb& e)
{
    return e.a == e.last ? e.kid0 : ElementOf(e.a, e.last, e.d_left, e.d_right);
}

CX_DEV Climb ParentOf(const Climb& l, const Climb& r)
{
    Climb p;
    p.a = l.a;
    p.last = r.last;
    p.d_left = l.d_left;
    p.d_right = r.d_right;
    p.kid0 = WordOf(l);
    p.kid1 = WordOf(r);
    p.own = Union(l.own, r.own);
    return p;
}

// The node index of a built element (the root is node 0).
CX_DEV uint32_t NodeOf(const Climb& p, int n)
{
    return p.a == 0 && p.last == uint32_t(n - 1) ? 0u : ElementOf(p.a, p.last, p.d_left, p.d_right);
}

// The outputs of the node `p` built from `l` and `r` at `gap` that go to global memory: the leaf
// children's parent words, aggregate and 4-wide slot list; the internal children's parent
// words and the cover diffs go through the callbacks (LDS at level 1). Returns the node index and the
// reference-format child words (kLtChildLeafBit | sorted position for a leaf child).
template <class InternalParent, class CoverAdd>
CX_DEV uint32_t WriteNode(const Build& b, int n, const Climb& l, const Climb& r, const Climb& p, uint32_t gap,
                          InternalParent&& internal_parent, CoverAdd&& cover_add, uint32_t& left, uint32_t& right)
{
    const uint32_t node = NodeOf(p, n);
    if (node == 0)
    {
        internal_parent(0u, kLtNone);
    }
    left = l.a == l.last ? (kLtChildLeafBit | gap) : gap;
    right = r.a == r.last ? (kLtChildLeafBit | (gap + 1)) : gap + 1;
    uint32_t* leaf_parent_out = b.scratch.u32(kLtScratchSortedLeafParent);
    if ((left & kLtChildLeafBit) != 0)
    {
        leaf_parent_out[gap] = node;
    }
    else
    {
        internal_parent(left, node);
    }
    if ((right & kLtChildLeafBit) != 0)
    {
        leaf_parent_out[gap + 1] = node;
    }
    else
    {
        internal_parent(right, node);
    }
    cover_add(p.a, p.last);
    StoreSummary(b.scratch.aggregates() + uint64_t(node) * 12, p.own);
    b.scratch.quad()[node] = QuadOf(l.kid0, l.kid1, r.kid0, r.kid1);
    return node;
}

// How a climb step ended.
enum : uint32_t
{
    kClimbOn = 0,   // the parent is built; the lane holds it and climbs on
    kClimbRoot,     // the root is finished
    kClimbOpen,     // the parent gap is on the block boundary
    kClimbParked,   // arrived first at its parent gap
};

CX_DEV bool IsRoot(const Climb& e, int n)
{
    return e.a == 0 && e.last == uint32_t(n - 1);
}

// The parent gap of a non-root element and whether it lies inside the block [lo, hi).
CX_DEV uint32_t ParentGap(const Climb& e, bool& is_left_child)
{
    is_left_child = e.d_right > e.d_left;
    return is_left_child ? e.last : e.a - 1;
}

CX_DEV bool GapInside(uint32_t gap, uint64_t lo, uint64_t hi)
{
    return gap != kLtNone && uint64_t(gap) >= lo && uint64_t(gap) + 1 < hi;
}

// Appends an open element to the block's open list.
CX_DEV void EmitOpen(const Build& b, uint32_t* list, uint32_t* lds_count, const Climb& e)
{
    const uint32_t k = atomicAdd(lds_count, 1u);
    if (k < kOpenCap)
    {
        StoreClimb(list + uint64_t(k) * kOpenWords, e);
    }
    else
    {
        AtomicAddRelaxed(&b.scratch.ctl()[kCtlTreeOverflow], 1u);
    }
}

// ---- level 1 ------------------------------------------------------------------------------------
// A level-1 workgroup (one lane per leaf) keeps its block of kTreeTile sorted leaves in LDS: the gap
// deltas of the block and its two neighbours, the leaf ids, the gap slots and cover diffs, the
// children of the nodes it builds (written out with coalesced stores at the end), and one parked
// summary per element position. An element that arrives first at its gap parks its summary at its
// position p (leaf p or node p) for the sibling to read. Leaf p and node p never need the slot at
// once: node p is built only after leaf p's parent (node p is that parent or an ancestor of it), and
// leaf p's parked summary is read before its parent is built.
//
// The climb runs in stages of kClimbStage steps. Half the elements stop at every step, so after a
// few steps most lanes of a wave idle while the wave still issues every instruction of a step; at
// the end of a stage the elements still climbing park their summary and queue their range, and the
// next stage hands them to the lowest lanes. Any element's record can be rebuilt from LDS (range,
// the delta table, the parked summary, the children), so the open elements are found after the climb
// from the slots (a parked element whose sibling did not arrive) and the two boundary tops, and are
// written in position order. 39.4 KB of LDS: four workgroups per CU.
constexpr uint32_t kUnset = 0xFFFFFFFEu;     // a staged output the block did not write
constexpr uint32_t kResolved = 0xFFFFFFFDu;  // a gap slot whose parent is built
constexpr uint32_t kClimbStage = kMaxClimb;

struct Level1Lds
{
    float park[kTreeTile][12];     // summary of the parked element at position lo + k
    uint32_t slot[kTreeTile];      // far range end of the first arrival at gap lo + k
    int32_t diff[kTreeTile];
    uint32_t left[kTreeTile];
    uint32_t right[kTreeTile];
    uint32_t leaf_id[kTreeTile];   // leaf index (sorted value) of sorted leaf block_lo + k
    uint2 queue[kTreeTile];        // ranges of the elements climbing on; then the open elements by start
    int8_t delta[kTreeTile + 4];   // delta of the gaps [block_lo - 1, block_hi - 1] (-1 .. 63)
    uint2 top[2];                  // open elements whose parent gap is block_lo - 1 / block_hi - 1
    uint32_t queue_count[2];
    uint32_t wave_open[kTreeTile / kWave];
};

CX_DEV Summary LdsSummary(const float (&v)[12])
{
    const float4 a = *reinterpret_cast<const float4*>(&v[0]);
    const float4 b = *reinterpret_cast<const float4*>(&v[4]);
    const float4 c = *reinterpret_cast<const float4*>(&v[8]);
    Summary s;
    s.face[0] = a.x, s.face[1] = a.y, s.face[2] = a.z, s.face[3] = a.w;
    s.face[4] = b.x, s.face[5] = b.y, s.lo[0] = b.z, s.lo[1] = b.w;
    s.lo[2] = c.x, s.hi[0] = c.y, s.hi[1] = c.z, s.hi[2] = c.w;
    return s;
}

CX_DEV void LdsStore(float (&v)[12], const Summary& s)
{
    *reinterpret_cast<float4*>(&v[0]) = make_float4(s.face[0], s.face[1], s.face[2], s.face[3]);
    *reinterpret_cast<float4*>(&v[4]) = make_float4(s.face[4], s.face[5], s.lo[0], s.lo[1]);
    *reinterpret_cast<float4*>(&v[8]) = make_float4(s.lo[2], s.hi[0], s.hi[1], s.hi[2]);
}

// delta(g) from the table (g in [lo - 1, hi - 1]).
CX_DEV int DeltaLds(const Level1Lds& l, uint32_t g, uint32_t lo)
{
    return int(l.delta[g + 1 - lo]);
}

// The child word of a reference-format child (kLtChildLeafBit | sorted position, or a node index).
CX_DEV uint32_t ChildWordLds(const Level1Lds& l, uint32_t child, uint32_t lo)
{
    return (child & kLtChildLeafBit) != 0 ? kLtChildLeafBit | l.leaf_id[(child & ~kLtChildLeafBit) - lo] : child;
}

// The record of the element [a, last] of the block, rebuilt from LDS (its summary parked).
CX_DEV Climb ElementLds(const Level1Lds& l, uint32_t a, uint32_t last, uint32_t lo)
{
    Climb e;
    e.a = a;
    e.last = last;
    e.d_left = DeltaLds(l, a - 1, lo);
    e.d_right = DeltaLds(l, last, lo);
    const uint32_t id = ElementOf(a, last, e.d_left, e.d_right);
    const uint32_t at = (id & ~kLtChildLeafBit) - lo;
    e.own = LdsSummary(l.park[at]);
    if ((id & kLtChildLeafBit) != 0)
    {
        e.kid0 = kLtChildLeafBit | l.leaf_id[at];
        e.kid1 = kLtNone;
    }
    else
    {
        e.kid0 = ChildWordLds(l, l.left[at], lo);
        e.kid1 = ChildWordLds(l, l.right[at], lo);
    }
    return e;
}

CX_DEV void ParkLds(Level1Lds& l, const Climb& e, uint32_t lo)
{
    const uint32_t id = ElementOf(e.a, e.last, e.d_left, e.d_right);
    LdsStore(l.park[(id & ~kLtChildLeafBit) - lo], e.own);
}

// One level-1 climb step.
CX_DEV uint32_t ClimbStepLds(Level1Lds& l, int n, Climb& e, uint32_t lo, uint32_t hi)
{
    if (IsRoot(e, n))
    {
        ParkLds(l, e, lo);  // node 0's aggregate, for the write-out
        return kClimbRoot;
    }
    bool is_left_child;
    const uint32_t gap = ParentGap(e, is_left_child);
    if (!GapInside(gap, lo, hi))
    {
        ParkLds(l, e, lo);
        l.top[is_left_child ? 1 : 0] = make_uint2(e.a, e.last);
        return kClimbOpen;
    }
    // Park the summary at the element's position before arriving; harmless when the sibling is
    // already there. The release and acquire fences order the LDS accesses only (the parked summary,
    // children and leaf ids all live in LDS), so the exchange does not wait for global stores.
    ParkLds(l, e, lo);
    const uint32_t far_end = is_left_child ? e.a : e.last;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    const uint32_t other =
        __hip_atomic_exchange(&l.slot[gap - lo], far_end, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (other == kLtNone)
    {
        return kClimbParked;  // the sibling finishes the parent
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    l.slot[gap - lo] = kResolved;
    const Climb s = is_left_child ? ElementLds(l, gap + 1, other, lo) : ElementLds(l, other, gap, lo);
    const Climb& lc = is_left_child ? e : s;
    const Climb& rc = is_left_child ? s : e;
    const Climb parent = ParentOf(lc, rc);
    // The node's children in LDS (indexed by the node, which lies in the block) and its cover diffs;
    // everything else about it is written out after the climb.
    const uint32_t node = NodeOf(parent, n);
    l.left[node - lo] = lc.a == lc.last ? (kLtChildLeafBit | gap) : gap;
    l.right[node - lo] = rc.a == rc.last ? (kLtChildLeafBit | (gap + 1)) : gap + 1;
    atomicAdd(&l.diff[parent.a - lo], 1);
    atomicAdd(&l.diff[parent.last - lo], -1);
    e = parent;
    return kClimbOn;
}

CX_DEV void TreeLevel1(const Build& b, int n, uint32_t lo, uint32_t hi, uint32_t* open, uint32_t* open_count,
                       Level1Lds& l)
{
    const uint32_t t = threadIdx.x;
    const uint32_t* keys = b.scratch.keys(0);
    const uint32_t* values = b.scratch.values(0);
    // delta(g) of the gaps g = lo - 1 .. hi - 1 (-1 outside [0, n - 1)).
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
        l.delta[k] = int8_t(d);
    }
    // One leaf per lane (blockDim.x == kTreeTile): its id and summary.
    const uint32_t leaves = min(uint32_t(n) - lo, kTreeTile);
    const bool has_leaf = t < leaves;
    const uint32_t leaf = has_leaf ? values[lo + t] : 0u;
    Climb e;
    if (has_leaf)
    {
        e.own = LoadSummary(b.leaves[leaf].summary.face);
    }
    l.slot[t] = kLtNone;
    l.diff[t] = 0;
    l.left[t] = kUnset;
    l.leaf_id[t] = leaf;
    if (t < 2)
    {
        l.top[t] = make_uint2(kLtNone, kLtNone);
        l.queue_count[t] = 0;
    }
    __syncthreads();
    bool climbing = has_leaf;
    if (has_leaf)
    {
        e.a = lo + t;
        e.last = e.a;
        e.d_left = DeltaLds(l, e.a - 1, lo);
        e.d_right = DeltaLds(l, e.a, lo);
        e.kid0 = kLtChildLeafBit | leaf;
        e.kid1 = kLtNone;
    }
    for (uint32_t stage = 0; stage < kMaxClimb / kClimbStage; ++stage)
    {
        uint32_t* count = &l.queue_count[stage & 1u];
        if (climbing)
        {
            uint32_t end = kClimbOn;
            for (uint32_t k = 0; k < kClimbStage && end == kClimbOn; ++k)
            {
                end = ClimbStepLds(l, n, e, lo, hi);
            }
            if (end == kClimbOn)
            {
                ParkLds(l, e, lo);
                l.queue[atomicAdd(count, 1u)] = make_uint2(e.a, e.last);
            }
        }
        __syncthreads();
        const uint32_t survivors = *count;
        if (survivors == 0)
        {
            break;
        }
        if (stage + 1 == kMaxClimb / kClimbStage)
        {
            if (t == 0)
            {
                AtomicAddRelaxed(&b.scratch.ctl()[kCtlTreeOverflow], 1u);  // never expected
            }
            break;
        }
        climbing = t < survivors;
        if (climbing)
        {
            const uint2 range = l.queue[t];
            e = ElementLds(l, range.x, range.y, lo);
        }
        if (t == 0)
        {
            l.queue_count[(stage + 1) & 1u] = 0;
        }
        __syncthreads();
    }
    // Open elements: parked at a gap the block did not resolve, or the tops. queue[k] = the open
    // element starting at position lo + k.
    l.queue[t] = make_uint2(kLtNone, kLtNone);
    __syncthreads();
    if (lo + t + 1 < uint32_t(n) && t + 1 < kTreeTile)
    {
        const uint32_t g = lo + t;
        const uint32_t v = l.slot[t];
        if (v != kLtNone && v != kResolved)
        {
            const uint2 range = v <= g ? make_uint2(v, g) : make_uint2(g + 1, v);
            l.queue[range.x - lo] = range;
        }
    }
    if (t < 2 && l.top[t].x != kLtNone)
    {
        l.queue[l.top[t].x - lo] = l.top[t];
    }
    __syncthreads();
    const uint2 mine = l.queue[t];
    const bool is_open = mine.x != kLtNone;
    const uint64_t ballot = Ballot(is_open);
    const uint32_t wave = t / kWave;
    if (Lane() == 0)
    {
        l.wave_open[wave] = uint32_t(__popcll(ballot));
    }
    __syncthreads();
    uint32_t rank = uint32_t(__popcll(ballot & ((uint64_t(1) << Lane()) - 1)));
    for (uint32_t w = 0; w < wave; ++w)
    {
        rank += l.wave_open[w];
    }
    if (is_open && rank < kOpenCap)
    {
        StoreClimb(open + uint64_t(rank) * kOpenWords, ElementLds(l, mine.x, mine.y, lo));
    }
    if (t == 0)
    {
        uint32_t total = 0;
        for (uint32_t w = 0; w < kTreeTile / kWave; ++w)
        {
            total += l.wave_open[w];
        }
        if (total > kOpenCap)
        {
            AtomicAddRelaxed(&b.scratch.ctl()[kCtlTreeOverflow], 1u);
        }
        *open_count = min(total, kOpenCap);
    }
    // Write-out of the nodes the block built, node lo + t by lane t: children (the left child's
    // index is the split gap: a leaf at the split, or a left child numbered by its range end), the
    // 4-wide slot list, the parent words of the children and the aggregate (parked at the node's
    // position: node p is the last element to use position p).
    int32_t* diff = b.scratch.cover_diff();
    const uint32_t g = lo + t;
    if (g < uint32_t(n))
    {
        diff[g] = l.diff[t];
    }
    const uint32_t left = l.left[t];
    const bool built = left != kUnset;
    if (built)
    {
        const uint32_t right = l.right[t];
        b.scratch.u32(kLtScratchLeft)[g] = left;
        b.scratch.u32(kLtScratchRight)[g] = right;
        const uint32_t kid0 = ChildWordLds(l, left, lo);
        const uint32_t kid1 = ChildWordLds(l, right, lo);
        uint32_t l0 = kid0;
        uint32_t l1 = kLtNone;
        uint32_t r0 = kid1;
        uint32_t r1 = kLtNone;
        uint32_t* parent_out = b.scratch.u32(kLtScratchParent);
        uint32_t* leaf_parent_out = b.scratch.u32(kLtScratchSortedLeafParent);
        if ((left & kLtChildLeafBit) == 0)
        {
            l0 = ChildWordLds(l, l.left[left - lo], lo);
            l1 = ChildWordLds(l, l.right[left - lo], lo);
            parent_out[left] = g;
        }
        else
        {
            leaf_parent_out[left & ~kLtChildLeafBit] = g;
        }
        if ((right & kLtChildLeafBit) == 0)
        {
            r0 = ChildWordLds(l, l.left[right - lo], lo);
            r1 = ChildWordLds(l, l.right[right - lo], lo);
            parent_out[right] = g;
        }
        else
        {
            leaf_parent_out[right & ~kLtChildLeafBit] = g;
        }
        b.scratch.quad()[g] = QuadOf(l0, l1, r0, r1);
        if (g == 0)
        {
            parent_out[0] = kLtNone;  // the root (the block holds the whole tree)
        }
    }
    // Aggregates, as float4 triples over consecutive lanes.
    float* agg_out = b.scratch.aggregates();
    for (uint32_t q = t; q < kTreeTile * 3; q += blockDim.x)
    {
        const uint32_t k = q / 3;
        if (l.left[k] != kUnset)
        {
            const float* v = &l.park[k][(q % 3) * 4];
            reinterpret_cast<float4*>(agg_out + uint64_t(lo + k) * 12)[q % 3] = make_float4(v[0], v[1], v[2], v[3]);
        }
    }
}

template <class Set>
__global__ void __launch_bounds__(kTreeThreads) TreeLevel1Kernel(Set set)
{
    const Build& b = set.Get();
    if (blockIdx.x >= TreeBlocks(b.leaf_capacity, 1) || !Building(b))
    {
        return;
    }
    __shared__ Level1Lds lds;
    const int n = int(b.scratch.ctl()[kCtlLeafCount]);
    const uint32_t lo = blockIdx.x * kTreeTile;
    uint32_t* count = b.scratch.open_count(1) + blockIdx.x;
    if (n >= 2 && lo < uint32_t(n))
    {
        TreeLevel1(b, n, lo, lo + kTreeTile, b.scratch.open(1) + uint64_t(blockIdx.x) * kOpenCap * kOpenWords,
                   count, lds);
    }
    else if (threadIdx.x == 0)
    {
        *count = 0;
    }
}


// ---- levels 2 and up ----------------------------------------------------------------------------
// A level-L workgroup starts from the open elements of its 32 sub-blocks. Arrivals go through an LDS
// hash table keyed by gap (the first arrival of a gap claims the entry; linear probing) whose value
// is the parked element's far range end and pool index; the parked element's record is in the LDS
// pool. Every arrival writes its record to a fresh pool entry before exchanging (so the sibling can
// read it once the exchange publishes it); the second arrival builds the parent. The hash holds one
// entry per gap an element arrives at: at most kUpperBuiltBound built nodes plus 2 * kDeltaLevels
// parked open elements. Pool records past kPoolEntries go to the block's spill slice (global memory,
// so their exchange fences cover global memory as well).
constexpr uint32_t kUpperThreads = 1024;
constexpr uint32_t kHashEntries = 2048;  // power of two
static_assert(kHashEntries >= kUpperBuiltBound + 2 * kDeltaLevels, "gap hash below the depth bound");
constexpr uint32_t kHashEmpty = 0xFFFFFFFFu;
constexpr uint64_t kArrivalNone = ~uint64_t(0);      // no arrival yet; else the first's pool slot
constexpr uint64_t kArrivalResolved = uint64_t(1) << 62;  // | left child's slot << 16 | right child's slot

struct UpperLds
{
    uint32_t key[kHashEntries];
    uint64_t arrival[kHashEntries];
    uint32_t pool[kPoolEntries][kOpenWords];
    uint32_t offsets[kTreeFanout + 1];
    uint32_t pool_count;
    uint32_t open_count;
};

CX_DEV void StorePool(UpperLds& l, uint32_t* spill, uint32_t slot, const Climb& e)
{
    if (slot < kPoolEntries)
    {
        StoreClimb(l.pool[slot], e);
    }
    else
    {
        StoreClimb(spill + uint64_t(slot - kPoolEntries) * kOpenWords, e);
    }
}

CX_DEV Climb LoadPool(const UpperLds& l, const uint32_t* spill, uint32_t slot)
{
    return slot < kPoolEntries ? LoadClimb(l.pool[slot]) : LoadClimb(spill + uint64_t(slot - kPoolEntries) * kOpenWords);
}

CX_DEV uint32_t HashFindOrInsert(UpperLds& l, uint32_t gap)
{
    uint32_t h = (gap * 0x9E3779B1u) >> (32 - 11);
    static_assert(kHashEntries == 1u << 11, "hash bits");
    for (uint32_t probe = 0; probe < kHashEntries; ++probe)
    {
        const uint32_t k = atomicCAS(&l.key[h], kHashEmpty, gap);
        if (k == kHashEmpty || k == gap)
        {
            return h;
        }
        h = (h + 1) & (kHashEntries - 1);
    }
    return kHashEmpty;
}

template <class Set>
__global__ void __launch_bounds__(kUpperThreads) TreeUpperKernel(Set set, uint32_t level)
{
    const Build& b = set.Get();
    if (level > TreeLevels(b.leaf_capacity) || blockIdx.x >= TreeBlocks(b.leaf_capacity, level) || !Building(b))
    {
        return;
    }
    __shared__ UpperLds l;
    const uint32_t t = threadIdx.x;
    const int n = int(b.scratch.ctl()[kCtlLeafCount]);
    const uint64_t tile = TreeTile(level);
    const uint64_t block_lo = uint64_t(blockIdx.x) * tile;
    const uint64_t block_hi = block_lo + tile;
    for (uint32_t k = t; k < kHashEntries; k += kUpperThreads)
    {
        l.key[k] = kHashEmpty;
        l.arrival[k] = kArrivalNone;
    }
    const uint32_t child_first = blockIdx.x * kTreeFanout;
    const uint32_t child_blocks = TreeBlocks(b.leaf_capacity, level - 1);
    if (t < kWave)
    {
        // Offsets of the sub-blocks' open lists: one count per lane, a wave prefix sum.
        const uint32_t child = child_first + t;
        const uint32_t count = t < kTreeFanout && child < child_blocks ? b.scratch.open_count(level - 1)[child] : 0u;
        uint32_t inclusive = count;
#pragma unroll
        for (int offset = 1; offset < kWave; offset <<= 1)
        {
            const uint32_t up = __shfl_up(inclusive, offset, kWave);
            inclusive += Lane() >= uint32_t(offset) ? up : 0u;
        }
        if (t <= kTreeFanout)
        {
            l.offsets[t] = inclusive - count;
        }
        if (t == 0)
        {
            l.pool_count = 0;
            l.open_count = 0;
        }
    }
    __syncthreads();
    uint32_t* open = b.scratch.open(level) + uint64_t(blockIdx.x) * kOpenCap * kOpenWords;
    uint32_t* spill = b.scratch.spill() + uint64_t(blockIdx.x) * kSpillEntries * kOpenWords;
    if (n >= 2 && block_lo < uint64_t(n))
    {
        const uint32_t total = l.offsets[kTreeFanout];
        const uint32_t* child_open = b.scratch.open(level - 1);
        for (uint32_t next = t; next < total; next += kUpperThreads)
        {
            uint32_t c = 0;
            while (l.offsets[c + 1] <= next)
            {
                ++c;
            }
            Climb e = LoadClimb(child_open + (uint64_t(child_first + c) * kOpenCap + (next - l.offsets[c])) * kOpenWords);
            // The climb touches LDS only; the nodes it builds are written out after it.
            for (uint32_t step = 0;; ++step)
            {
                if (IsRoot(e, n))
                {
                    break;
                }
                if (step == kMaxClimb)
                {
                    AtomicAddRelaxed(&b.scratch.ctl()[kCtlTreeOverflow], 1u);  // never expected
                    break;
                }
                bool is_left_child;
                const uint32_t gap = ParentGap(e, is_left_child);
                if (!GapInside(gap, block_lo, block_hi))
                {
                    EmitOpen(b, open, &l.open_count, e);
                    break;
                }
                const uint32_t entry = HashFindOrInsert(l, gap);
                const uint32_t slot = atomicAdd(&l.pool_count, 1u);
                if (entry == kHashEmpty || slot >= kUpperArrivalBound)
                {
                    AtomicAddRelaxed(&b.scratch.ctl()[kCtlTreeOverflow], 1u);  // never expected
                    break;
                }
                StorePool(l, spill, slot, e);
                if (slot < kPoolEntries)
                {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
                }
                else
                {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                }
                const uint64_t other = __hip_atomic_exchange(&l.arrival[entry], uint64_t(slot), __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_WORKGROUP);
                if (other == kArrivalNone)
                {
                    break;  // parked: the sibling finishes the parent
                }
                // Resolved: the pool slots of the left and the right child, for the write-out.
                const uint32_t first = uint32_t(other);
                if (first < kPoolEntries)
                {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
                }
                else
                {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                }
                const uint32_t left_slot = is_left_child ? slot : first;
                const uint32_t right_slot = is_left_child ? first : slot;
                l.arrival[entry] = kArrivalResolved | (uint64_t(left_slot) << 16) | right_slot;
                const Climb sibling = LoadPool(l, spill, first);
                e = is_left_child ? ParentOf(e, sibling) : ParentOf(sibling, e);
            }
        }
    }
    __syncthreads();
    // Write-out: every node built here (from its children's pool records, in parallel), and the
    // parked elements whose sibling did not arrive, which stay open.
    uint32_t* left_out = b.scratch.u32(kLtScratchLeft);
    uint32_t* right_out = b.scratch.u32(kLtScratchRight);
    uint32_t* parent_out = b.scratch.u32(kLtScratchParent);
    int32_t* diff = b.scratch.cover_diff();
    for (uint32_t k = t; k < kHashEntries; k += kUpperThreads)
    {
        const uint64_t v = l.arrival[k];
        if (v == kArrivalNone)
        {
            continue;
        }
        if ((v & kArrivalResolved) == 0)
        {
            EmitOpen(b, open, &l.open_count, LoadPool(l, spill, uint32_t(v)));
            continue;
        }
        const Climb lc = LoadPool(l, spill, uint32_t(v >> 16) & 0xFFFFu);
        const Climb rc = LoadPool(l, spill, uint32_t(v) & 0xFFFFu);
        const Climb parent = ParentOf(lc, rc);
        uint32_t left;
        uint32_t right;
        const uint32_t node = WriteNode(
            b, n, lc, rc, parent, l.key[k], [&](uint32_t child, uint32_t nd) { parent_out[child] = nd; },
            [&](uint32_t pa, uint32_t pb) {
                __hip_atomic_fetch_add(&diff[pa], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(&diff[pb], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            },
            left, right);
        left_out[node] = left;
        right_out[node] = right;
    }
    __syncthreads();
    if (t == 0)
    {
        b.scratch.open_count(level)[blockIdx.x] = min(l.open_count, kOpenCap);
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
    const uint32_t* left = b.scratch.u32(kLtScratchLeft);
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
        // node i's split gap: its left child's index (kLtChildLeafBit | gap for a leaf, else node gap)
        at[k] = i + 1 < n ? left[i] & ~kLtChildLeafBit : 0u;  // gap 0 stands in for items past the gaps
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
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k)
    {
        const uint32_t i = base + k * kScanThreads;
        if (i < capacity)
        {
            wide[i] = uint32_t(prefix + v[k]);
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
    uint32_t error = 0;
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
        // More 4-wide nodes than the capacity (the reference's overflow: the host grows the node
        // buffer and builds again), or a tree level past one of its work-list bounds (never
        // expected: they are sized from the depth bound): the tree is left empty, with the flag
        // that says which.
        overflow = node_count > b.node_capacity ? 1u : 0u;
        error = ctl[kCtlTreeOverflow] != 0 ? 1u : 0u;
        if (overflow == 0 && error == 0)
        {
            root = LoadSummary(b.scratch.aggregates());
            root_index = 0;
        }
    }
    StoreSummary(b.header->summary.face, root);
    b.header->root = root_index;
    b.header->leaf_count = overflow != 0 || error != 0 ? 0u : n;
    b.header->node_count = node_count;
    b.header->pad = 0;
    b.state->node_overflow = overflow;
    b.state->build_error = error;
    b.state->builds = b.state->builds + 1;
}

// ------------------------------------------------------------------------------------------------
// wide (caelix_light_bvh_wide.hlsl:10-100), chunked: a workgroup owns kWideChunk consecutive binary nodes, whose 4-wide nodes are the
// consecutive indices [wide[c0], wide[c1]) (the exclusive scan of the even-depth flags). It stages the
// chunk's slot lists, 4-wide indices and binary parents in LDS (coalesced), then builds each of its
// 4-wide nodes in one lane: the slot summaries (a leaf's record, or the binary aggregate of an
// even-depth grandchild) and child words, and the node's own parent word (its binary grandparent's
// 4-wide index and the slot that holds it there), so every 224-byte record is written whole by the
// lane that owns it and no lane writes into another chunk's records. Lookups outside the chunk go to
// global memory. Each wave assembles its 64 records in its LDS slice and writes them out as one
// contiguous run. Leaf parents are scattered into the leaf records. 4-wide nodes past the node capacity
// are not built (the overflow leaves the tree empty).
constexpr uint32_t kWideChunk = 512;
constexpr uint32_t kWideChunkThreads = 256;

template <class Set>
__global__ void __launch_bounds__(kWideChunkThreads) WideChunkKernel(Set set)
{
    const Build& b = set.Get();
    if (!Building(b))
    {
        return;
    }
    const uint32_t* ctl = b.scratch.ctl();
    const uint32_t n = ctl[kCtlLeafCount];
    const uint32_t total = ctl[kCtlWideNodes];
    const uint32_t c0 = blockIdx.x * kWideChunk;
    if (total == 0 || uint64_t(c0) + 1 >= n)
    {
        return;
    }
    const uint32_t c1 = min(c0 + kWideChunk, n - 1);  // binary nodes [c0, c1)
    __shared__ uint4 lds_quad[kWideChunk];
    __shared__ uint32_t lds_wide[kWideChunk + 1];
    __shared__ uint32_t lds_parent[kWideChunk];
    __shared__ uint32_t lds_node[kWideChunk];  // binary node of each 4-wide node of the chunk
    __shared__ float4 lds_rec[kWideChunkThreads / kWave][kWave * 14];  // per wave: 64 records of 14 chunks
    const uint32_t* wide = b.scratch.u32(kLtScratchWide);
    const uint32_t* parent = b.scratch.u32(kLtScratchParent);
    const uint4* quad = b.scratch.quad();
    const uint32_t w0 = wide[c0];
    for (uint32_t k = threadIdx.x; k < c1 - c0; k += kWideChunkThreads)
    {
        const uint32_t i = c0 + k;
        lds_quad[k] = quad[i];
        lds_wide[k] = wide[i];
        lds_parent[k] = parent[i];
    }
    const uint32_t w1 = c1 < n - 1 ? wide[c1] : total;
    if (threadIdx.x == 0)
    {
        lds_wide[c1 - c0] = w1;
    }
    __syncthreads();
    // Node i has an even depth (is a 4-wide node) where the exclusive scan of the even-depth flags steps.
    for (uint32_t k = threadIdx.x; k < c1 - c0; k += kWideChunkThreads)
    {
        const uint32_t wi = lds_wide[k];
        if (lds_wide[k + 1] != wi)
        {
            lds_node[wi - w0] = c0 + k;
        }
    }
    __syncthreads();
    auto in_chunk = [&](uint32_t node) { return node >= c0 && node < c1; };
    auto wide_of = [&](uint32_t node) { return in_chunk(node) ? lds_wide[node - c0] : wide[node]; };
    auto quad_of = [&](uint32_t node) { return in_chunk(node) ? lds_quad[node - c0] : quad[node]; };
    auto parent_of = [&](uint32_t node) { return in_chunk(node) ? lds_parent[node - c0] : parent[node]; };
    const uint32_t built = min(w1, b.node_capacity);
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t lane = threadIdx.x % kWave;
    for (uint32_t base = w0; base < built; base += kWideChunkThreads)  // uniform: every thread runs each round
    {
        const uint32_t wave_first = base + wave * kWave;
        const uint32_t w = wave_first + lane;
        if (w < built)
        {
            const uint32_t j = lds_node[w - w0];
            const uint4 q = lds_quad[j - c0];
            const uint32_t slot_word[4] = {q.x, q.y, q.z, q.w};
            // All loads first: per slot the 4-wide index (internal) and the summary; then the parent word.
            uint32_t child_word[4];
            Summary summary[4];
#pragma unroll
            for (uint32_t s = 0; s < 4; ++s)
            {
                const uint32_t word = slot_word[s];
                const bool used = word != kLtNone;
                const bool leaf = used && (word & kLtChildLeafBit) != 0;
                child_word[s] = used && !leaf ? wide_of(word) : word;
                const float* src = leaf ? b.leaves[word & ~kLtChildLeafBit].summary.face
                                        : b.scratch.aggregates() + (used ? uint64_t(word) * 12 : 0u);
                summary[s] = used ? LoadSummary(src) : EmptySummary();
            }
            uint32_t up = kLtNone;
            uint32_t up_slot = 0;
            if (j != 0)
            {
                const uint32_t g = parent_of(parent_of(j));
                up = wide_of(g);
                const uint4 gq = quad_of(g);
                up_slot = gq.x == j ? 0u : gq.y == j ? 1u : gq.z == j ? 2u : 3u;
            }
            float4* dst = &lds_rec[wave][lane * 14];
#pragma unroll
            for (uint32_t s = 0; s < 4; s += 2)
            {
                // child_face[s], child_face[s + 1]: 12 floats = 3 chunks
                const Summary& x = summary[s];
                const Summary& y = summary[s + 1];
                dst[3 * (s / 2) + 0] = make_float4(x.face[0], x.face[1], x.face[2], x.face[3]);
                dst[3 * (s / 2) + 1] = make_float4(x.face[4], x.face[5], y.face[0], y.face[1]);
                dst[3 * (s / 2) + 2] = make_float4(y.face[2], y.face[3], y.face[4], y.face[5]);
            }
            // child_lo[4][3] (chunks 6..8), child_hi[4][3] (chunks 9..11)
            dst[6] = make_float4(summary[0].lo[0], summary[0].lo[1], summary[0].lo[2], summary[1].lo[0]);
            dst[7] = make_float4(summary[1].lo[1], summary[1].lo[2], summary[2].lo[0], summary[2].lo[1]);
            dst[8] = make_float4(summary[2].lo[2], summary[3].lo[0], summary[3].lo[1], summary[3].lo[2]);
            dst[9] = make_float4(summary[0].hi[0], summary[0].hi[1], summary[0].hi[2], summary[1].hi[0]);
            dst[10] = make_float4(summary[1].hi[1], summary[1].hi[2], summary[2].hi[0], summary[2].hi[1]);
            dst[11] = make_float4(summary[2].hi[2], summary[3].hi[0], summary[3].hi[1], summary[3].hi[2]);
            reinterpret_cast<uint4*>(dst)[12] = make_uint4(child_word[0], child_word[1], child_word[2], child_word[3]);
            reinterpret_cast<uint4*>(dst)[13] = make_uint4(up, up_slot, 0u, 0u);
#pragma unroll
            for (uint32_t s = 0; s < 4; ++s)
            {
                const uint32_t word = child_word[s];
                if (word != kLtNone && (word & kLtChildLeafBit) != 0)
                {
                    b.leaves[word & ~kLtChildLeafBit].parent = (w << kLtParentSlotBits) | s;
                }
            }
        }
        __syncthreads();
        if (wave_first < built)
        {
            const uint32_t records = min(kWave, built - wave_first);
            float4* out = reinterpret_cast<float4*>(b.nodes + wave_first);
            for (uint32_t c = lane; c < records * 14; c += kWave)
            {
                out[c] = lds_rec[wave][c];
            }
        }
        __syncthreads();
    }
}

// caelix_light_bvh_begin.hlsl:9-17
// One onesweep pass of the build's sort (keys / values ping-pong in the scratch).
template <class Set, class L>
__global__ void __launch_bounds__(L::kThreads, kSortWavesPerSimd) BuildSortPassKernel(Set set, uint32_t pass)
{
    const Build& b = set.Get();
    if (blockIdx.x >= SortTilesOf<L>(b.leaf_capacity))
    {
        return;
    }
    // pass p reads buffer (passes + p) % 2: the last pass writes buffer 0
    const uint32_t in = (uint32_t(L::kPasses) + pass) & 1u;
    const SortSink<L> out{b.scratch.keys(1 - in), b.scratch.values(1 - in), nullptr, nullptr};
    SortPassTileOf<L>(b.scratch.keys(in), b.scratch.values(in), out, &b.header->leaf_count, &b.state->building,
                      SortStateOf<L>(b.scratch.sort_base(), b.leaf_capacity), pass);
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
// leaf capacity of the set, `count` = the number of builds (grid y). Keys = a SortMorton layout (the
// reference's uint32 Morton keys with leaf values) or SortKeys64 (code << 32 | leaf, sorted as one
// 64-bit key); all give the same order, hence the same tree.
template <class Keys, class Set>
void RecordBuild(const Set& set, uint32_t capacity, uint32_t count, hipStream_t s, bool timed = true)
{
    const uint32_t tiles = SortTiles(capacity);
    {
        StageScope scope(s, "lt_bounds", timed);
        BoundsKernel<Set, Keys><<<dim3(BoundBlocks(capacity), count), kBvhThreads, 0, s>>>(set);
    }
    {
        StageScope scope(s, "lt_morton", timed);
        MortonKernel<Set, Keys><<<dim3(tiles, count), kSortThreads, 0, s>>>(set);
    }
    if constexpr (Keys::kValues)
    {
        for (uint32_t pass = 0; pass < uint32_t(Keys::kPasses); ++pass)
        {
            StageScope scope(s, "sort_pass", timed);
            BuildSortPassKernel<Set, Keys><<<dim3(SortTilesOf<Keys>(capacity), count), Keys::kThreads, 0, s>>>(set, pass);
        }
    }
    else
    {
        for (uint32_t pass = 0; pass < uint32_t(SortKeys64::kPasses); ++pass)
        {
            StageScope scope(s, "sort64_pass", timed);
            BuildSort64PassKernel<<<dim3(SortTilesOf<SortKeys64>(capacity), count), SortKeys64::kThreads, 0, s>>>(
                set, pass);
        }
    }
    static const char* const kTreeStage[] = {"lt_tree_l1", "lt_tree_l2", "lt_tree_l3", "lt_tree_l4", "lt_tree_l5",
                                             "lt_tree_l6"};
    {
        StageScope scope(s, kTreeStage[0], timed);
        TreeLevel1Kernel<<<dim3(TreeBlocks(capacity, 1), count), kTreeThreads, 0, s>>>(set);
    }
    for (uint32_t level = 2; level <= TreeLevels(capacity); ++level)
    {
        StageScope scope(s, kTreeStage[std::min(level, 6u) - 1], timed);
        TreeUpperKernel<<<dim3(TreeBlocks(capacity, level), count), kUpperThreads, 0, s>>>(set, level);
    }
    const uint32_t scan_tiles = DivUp(capacity, kScanTile);
    {
        StageScope scope(s, "lt_cover_scan", timed);
        CoverScanKernel<<<dim3(scan_tiles, count), kScanThreads, 0, s>>>(set);
    }
    {
        StageScope scope(s, "lt_flags_scan", timed);
        FlagsScanKernel<<<dim3(scan_tiles, count), kScanThreads, 0, s>>>(set);
    }
    {
        StageScope scope(s, "lt_wide", timed);
        WideChunkKernel<<<dim3(DivUp(capacity, kWideChunk), count), kWideChunkThreads, 0, s>>>(set);
    }
}

// The tunables of the 32-bit key build's sort (cx_light_bvh_set_tuning): digit bits 6 or 8, tiles of
// 4096 or 8192 elements. They change speed only: every layout sorts stably, so the order and the tree
// are the same.
CxLightBvhTuning g_bvh_tuning{8u, 8192u};

template <class Set>
void RecordBuild32(const Set& set, uint32_t capacity, uint32_t count, hipStream_t s, bool timed = true)
{
    const bool small = g_bvh_tuning.tile == 4096u;
    if (g_bvh_tuning.digit_bits == 6u)
    {
        small ? RecordBuild<SortMorton6x4096>(set, capacity, count, s, timed)
              : RecordBuild<SortMorton6x8192>(set, capacity, count, s, timed);
    }
    else
    {
        small ? RecordBuild<SortMorton8x4096>(set, capacity, count, s, timed)
              : RecordBuild<SortMorton8x8192>(set, capacity, count, s, timed);
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
// The build scratch of cx_light_bvh_build_static: one per GPU the library runs on, held until the
// process ends, shared by every call on that GPU. Its host-side name is not a device address; the
// address comes from hipGetSymbolAddress, per device.
constexpr uint32_t kStaticLeafCapacity = 65536;
constexpr uint64_t kStaticScratchBytes = 11ull << 20;
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

constexpr int kStaticMaxDevices = 64;
std::mutex g_static_mutex;
void* g_static_scratch[kStaticMaxDevices];

// The device address of the current device's static scratch (looked up on the first call per device).
int StaticScratch(void** out, const char* what)
{
    int device = 0;
    if (hipError_t e = hipGetDevice(&device); e != hipSuccess)
    {
        return HipResult(e, what);
    }
    if (device < 0 || device >= kStaticMaxDevices)
    {
        return ArgError("%s: device %d above the %d devices the static scratch serves", what, device, kStaticMaxDevices);
    }
    std::lock_guard<std::mutex> lock(g_static_mutex);
    if (g_static_scratch[device] == nullptr)
    {
        void* address = nullptr;
        if (hipError_t e = hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)); e != hipSuccess)
        {
            return HipResult(e, what);
        }
        g_static_scratch[device] = address;
    }
    *out = g_static_scratch[device];
    return CX_OK;
}

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

CX_API int cx_light_bvh_set_tuning(const CxLightBvhTuning* tuning)
{
    if (tuning == nullptr || (tuning->digit_bits != 6u && tuning->digit_bits != 8u) ||
        (tuning->tile != 4096u && tuning->tile != 8192u))
    {
        return cx::ArgError("cx_light_bvh_set_tuning: digit bits must be 6 or 8 and the tile 4096 or 8192");
    }
    cx::g_bvh_tuning = *tuning;
    return CX_OK;
}

CX_API void cx_light_bvh_get_tuning(CxLightBvhTuning* tuning)
{
    if (tuning != nullptr)
    {
        *tuning = cx::g_bvh_tuning;
    }
}

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
    cx::RecordBuild32(cx::OneBuild{b}, build->leaf_capacity, 1, s);
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

CX_API int cx_light_bvh_build_static(const CxLightBvhBuild* build, void* stream)
{
    const char* what = "cx_light_bvh_build_static";
    if (build == nullptr || build->scratch != nullptr)
    {
        return cx::ArgError("%s: no build, or a scratch buffer (this entry point uses its own; pass NULL)", what);
    }
    if (build->leaf_capacity > cx::kStaticLeafCapacity ||
        cx::Scratch::Bytes(build->leaf_capacity) > cx::kStaticScratchBytes)
    {
        return cx::ArgError("%s: leaf capacity %u above the static scratch's %u leaves", what, build->leaf_capacity,
                            cx::kStaticLeafCapacity);
    }
    void* static_scratch = nullptr;
    if (int r = cx::StaticScratch(&static_scratch, what); r != CX_OK)
    {
        return r;
    }
    CxLightBvhBuild desc = *build;
    desc.scratch = static_scratch;
    cx::Build b;
    if (int r = cx::CheckBuild(&desc, what, b); r != CX_OK)
    {
        return r;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // The control words sit at an offset that depends on the capacity, over whatever the previous call
    // left there: clear them with a kernel (no memset node if the caller captures the stream).
    cx::ScratchInitKernel<<<1, 64, 0, s>>>(b.scratch.ctl());
    cx::RecordBuild32(cx::OneBuild{b}, build->leaf_capacity, 1, s, false);
    return cx::LaunchResult(what);
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
    cx::RecordBuild32(cx::BuildArray{static_cast<const cx::Build*>(descriptors)}, max_leaf_capacity,
                                     count, s);
    return cx::LaunchResult("cx_light_bvh_build_batch");
}

}  // extern "C"
