// The light tree's host side on HIP (the reference's LightTree: lib/src/light/light_tree.h:37-101,
// light_tree.cpp, light_tree_record.cpp): the CPU state the scene hooks keep (entity motions, group
// lives, the groups whose flags are staged, the reasons to rebuild a tree), the light code set, the
// device tables, and the flush: staging of the flags and motion words, growth copies and zero fills,
// the group pass and per tree the gather and build, the copy of the control back to the host, and at
// recycle the growth of the heap and the trees. The reference records into a Vulkan command buffer and
// frees replaced buffers once their frame is safe; here every step is enqueued on a HIP stream, small
// uploads go through pinned host arenas (two, used by alternate flushes; an arena is reused once the
// event after its copies has completed), and replaced buffers are freed at recycle.
#include <algorithm>
#include <array>
#include <cstring>
#include <vector>

#include "runtime.h"

namespace cx {
namespace {

constexpr uint64_t kGroupsInitial = 1024;
constexpr uint64_t kEntitiesInitial = 64;
constexpr uint32_t kLeavesInitial = 1024;        // per tree; a multiple of 64
constexpr uint64_t kMaxHeapUnits = 1ull << 32;   // kLtMaxHeapUnits: 16-byte units per heap buffer
constexpr uint64_t kMaxLeaves = 1ull << 26;      // leaf capacity limit of one tree
constexpr uint32_t kCodeWords = 65536 / 32;      // one bit per material code
constexpr uint64_t kDefaultHeapBytes = 16ull << 20;
constexpr float kDefaultMinEmission = 20.0f;

// 4-wide nodes for a leaf capacity (light_tree.cpp:21-29).
uint32_t NodeCapacityFor(uint32_t leaf_capacity)
{
    return leaf_capacity / 8 * 5 + 64;
}

__global__ void __launch_bounds__(256) ZeroKernel(uint32_t* p, uint64_t words)
{
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += uint64_t(gridDim.x) * blockDim.x)
    {
        p[i] = 0u;
    }
}

void Zero(void* p, uint64_t bytes, hipStream_t s)
{
    const uint64_t words = bytes / 4;
    if (words > 0)
    {
        ZeroKernel<<<uint32_t(std::min<uint64_t>(DivUp(words, 256), 4096)), 256, 0, s>>>(static_cast<uint32_t*>(p), words);
    }
}

// BlockMaterialCode (caelix_record.hlsl:209) on the host: an opaque value as it is, a transparent one
// without its face bits.
uint32_t CanonicalCode(uint32_t code, uint32_t encoding)
{
    const bool packed = encoding == kEncodingPacked;
    return (code & (packed ? kPackedOpaqueMask : kBlockIdOpaqueMask)) != 0 ? code
                                                                            : (code & (packed ? kPackedIdMask : kBlockIdIdMask));
}

// A device buffer that grows (the reference's GrowableBuffer): Reserve allocates a larger buffer and
// keeps the old one as the growth source until the transfers take it.
struct Growable
{
    void* ptr = nullptr;
    uint64_t bytes = 0;
    void* source = nullptr;  // the replaced buffer, its contents not yet copied
    uint64_t source_bytes = 0;

    int Reserve(uint64_t need, const char* what)
    {
        if (need <= bytes)
        {
            return CX_OK;
        }
        const uint64_t grown = std::max(need, bytes * 2);
        void* p = nullptr;
        if (int r = HipResult(hipMalloc(&p, grown), what); r != CX_OK)
        {
            return r;
        }
        if (source == nullptr)
        {
            source = ptr;
            source_bytes = bytes;
        }
        else
        {
            // a second growth before the transfers: the intermediate buffer holds nothing yet
            (void)hipFree(ptr);
        }
        ptr = p;
        bytes = grown;
        return CX_OK;
    }
};

struct Growth
{
    void* source;  // null: fill only
    void* destination;
    uint64_t copy_bytes;
    uint64_t fill_from;  // zero [fill_from, fill_to) after the copy
    uint64_t fill_to;
};

struct Deferred
{
    uint64_t frame;  // 0: set at end_flush
    void* ptr;
};

struct Arena
{
    uint8_t* host = nullptr;  // pinned
    uint64_t bytes = 0;
    uint64_t used = 0;
    hipEvent_t done = nullptr;
    bool in_flight = false;
    struct Copy
    {
        uint64_t offset;
        void* destination;
        uint64_t bytes;
    };
    std::vector<Copy> copies;
};

struct ReadbackSlot
{
    LightControl* host = nullptr;  // pinned
    hipEvent_t done = nullptr;
    hipEvent_t start = nullptr;  // the light update's GPU time: start .. done
    uint64_t frame = 0;
    uint64_t heap_epoch = 0;
    std::array<uint64_t, kLtTrees> tree_epoch{};
    bool pending = false;
};

}  // namespace
}  // namespace cx

struct CxLightTree
{
    struct EntityState
    {
        uint32_t motion = CX_ENTITY_STATIC;
        bool motion_dirty = false;
    };
    struct GroupState
    {
        bool fresh = false;
        bool listed = false;
        bool alive = false;
    };

    uint32_t group_shift = 4;
    uint32_t encoding = 0;
    float min_emission = 20.0f;

    std::vector<EntityState> entities;
    std::vector<GroupState> groups;
    std::vector<uint32_t> staged_groups;
    bool materials_changed = true;
    bool pass_pending = false;
    std::array<bool, kLtTrees> tree_reason{};
    CxLightSceneTables tables{};
    uint32_t pass_group_count = 0;

    std::vector<uint32_t> codes;
    bool codes_upload = false;
    bool codes_valid = false;
    uint32_t* codes_device = nullptr;

    cx::Growable rows, staged_flags, entity_motion, pass_scratch;
    LightControl* control = nullptr;
    std::array<void*, 2> heaps{};
    uint64_t heap_units = 0;
    uint64_t heap_epoch = 0;
    uint64_t grow_request_units = 0;
    bool heap_limit_logged = false;
    bool fills_pending = true;

    std::array<cx::Growable, kLtTrees> leaves, nodes, bvh_scratch;
    std::array<uint32_t, kLtTrees> leaf_capacity{};
    std::array<uint32_t, kLtTrees> node_capacity{};
    std::array<uint32_t, kLtTrees> leaf_request{};
    std::array<uint32_t, kLtTrees> node_request{};
    std::array<uint64_t, kLtTrees> tree_epoch{};
    std::array<bool, kLtTrees> scratch_init{};

    LightTreeDescriptor* descriptor_device = nullptr;
    LightTreeDescriptor descriptor{};
    bool descriptor_upload = true;

    std::vector<cx::Growth> growth;
    bool record_pass = false;
    std::array<bool, kLtTrees> record_tree{};

    std::array<cx::Arena, 2> arenas;
    int arena = -1;  // the arena of the flush being built, -1 = none yet
    uint32_t next_arena = 0;
    std::vector<cx::ReadbackSlot> readback;
    LightControl last_control{};
    uint64_t last_control_frame = 0;
    float gpu_ms = 0.0f;
    std::vector<cx::Deferred> deferred;
    uint64_t safe_frame = 0;

    EntityState& Entity(uint32_t e)
    {
        if (e >= entities.size())
        {
            entities.resize(size_t(e) + 1);
        }
        return entities[e];
    }
    GroupState& Group(uint32_t g)
    {
        if (g >= groups.size())
        {
            groups.resize(size_t(g) + 1);
        }
        return groups[g];
    }
    void ListGroup(uint32_t g)
    {
        GroupState& s = Group(g);
        if (!s.listed)
        {
            s.listed = true;
            staged_groups.push_back(g);
        }
    }
    void Defer(void* p)
    {
        if (p != nullptr)
        {
            deferred.push_back(cx::Deferred{0, p});
        }
    }
    // The growth copy of a Growable's replaced buffer into the new one, then zeros up to zero_tail_to.
    void TakeGrowth(cx::Growable& b, uint64_t zero_tail_to)
    {
        if (b.source == nullptr)
        {
            return;
        }
        const uint64_t old = b.source_bytes;
        growth.push_back(cx::Growth{b.source, b.ptr, old, zero_tail_to > old ? old : 0, zero_tail_to > old ? zero_tail_to : 0});
        b.source = nullptr;
        b.source_bytes = 0;
    }
    // A replaced buffer that is not copied (the new one is filled whole or written before it is read).
    void DeferSource(cx::Growable& b)
    {
        Defer(b.source);
        b.source = nullptr;
        b.source_bytes = 0;
    }
    cx::Arena& OpenArena()
    {
        if (arena < 0)
        {
            arena = int(next_arena);
            next_arena ^= 1u;
            cx::Arena& a = arenas[arena];
            if (a.in_flight)
            {
                (void)hipEventSynchronize(a.done);
                a.in_flight = false;
            }
            a.used = 0;
            a.copies.clear();
        }
        return arenas[arena];
    }
    // Room for `bytes` in the flush's arena (16-byte aligned), growing it; nullptr on failure.
    uint8_t* Stage(uint64_t bytes, void* destination)
    {
        cx::Arena& a = OpenArena();
        const uint64_t offset = (a.used + 15) & ~uint64_t(15);
        if (offset + bytes > a.bytes)
        {
            // grow: the copies already recorded in this arena have not been enqueued yet
            const uint64_t grown = std::max<uint64_t>(2 * (offset + bytes), 64 * 1024);
            uint8_t* p = nullptr;
            if (hipHostMalloc(reinterpret_cast<void**>(&p), grown, hipHostMallocDefault) != hipSuccess)
            {
                return nullptr;
            }
            if (a.host != nullptr)
            {
                std::memcpy(p, a.host, a.used);
                (void)hipHostFree(a.host);
            }
            a.host = p;
            a.bytes = grown;
        }
        a.used = offset + bytes;
        a.copies.push_back(cx::Arena::Copy{offset, destination, bytes});
        return a.host + offset;
    }
    LightTreeDescriptor BuildDescriptor() const
    {
        LightTreeDescriptor d{};
        d.control = reinterpret_cast<uint64_t>(control);
        d.heap[0] = reinterpret_cast<uint64_t>(heaps[0]);
        d.heap[1] = reinterpret_cast<uint64_t>(heaps[1]);
        d.group_rows = reinterpret_cast<uint64_t>(rows.ptr);
        d.group_infos = reinterpret_cast<uint64_t>(tables.group_infos);
        d.entity_infos = reinterpret_cast<uint64_t>(tables.entity_infos);
        d.entity_motion = reinterpret_cast<uint64_t>(entity_motion.ptr);
        d.grids = 0;
        for (uint32_t t = 0; t < kLtTrees; ++t)
        {
            d.nodes[t] = reinterpret_cast<uint64_t>(nodes[t].ptr);
            d.leaves[t] = reinterpret_cast<uint64_t>(leaves[t].ptr);
        }
        d.light_codes = reinterpret_cast<uint64_t>(codes_device);
        d.tree_headers = reinterpret_cast<uint64_t>(control) + offsetof(LightControl, headers);
        d.group_shift = group_shift;
        d.encoding = encoding;
        d.group_count = tables.group_capacity;
        return d;
    }
    void GrowHeap(uint64_t needed)
    {
        uint64_t units = std::max(heap_units * 2, needed + needed / 2);
        units = std::min(units, cx::kMaxHeapUnits);
        if (units <= heap_units)
        {
            if (!heap_limit_logged)
            {
                cx::Log(3, "light tree: the heap needs %llu bytes, above the limit of %llu bytes per buffer",
                        static_cast<unsigned long long>(needed * kLtUnitBytes),
                        static_cast<unsigned long long>(cx::kMaxHeapUnits * kLtUnitBytes));
                heap_limit_logged = true;
            }
            return;
        }
        std::array<void*, 2> grown{};
        for (void*& h : grown)
        {
            if (hipMalloc(&h, units * kLtUnitBytes) != hipSuccess)
            {
                for (void* g : grown)
                {
                    (void)hipFree(g);
                }
                cx::Log(3, "light tree: the heap cannot grow to 2 x %llu bytes",
                        static_cast<unsigned long long>(units * kLtUnitBytes));
                return;
            }
        }
        cx::Log(1, "light tree: the heap grows from 2 x %llu to 2 x %llu bytes (%llu needed)",
                static_cast<unsigned long long>(heap_units * kLtUnitBytes),
                static_cast<unsigned long long>(units * kLtUnitBytes),
                static_cast<unsigned long long>(needed * kLtUnitBytes));
        // which buffer is current is known on the GPU only: both are copied, at the same offsets
        for (int i = 0; i < 2; ++i)
        {
            growth.push_back(cx::Growth{heaps[i], grown[i], heap_units * kLtUnitBytes, 0, 0});
            heaps[i] = grown[i];
        }
        heap_units = units;
        ++heap_epoch;
    }
    void RebuildCodes(const float* materials, bool& changed)
    {
        std::vector<uint32_t> next(cx::kCodeWords, 0u);
        for (uint32_t code = 0; code < 65536; ++code)
        {
            const float* e = materials + uint64_t(cx::CanonicalCode(code, encoding)) * 8 + 4;
            const float luminance = 0.2126f * e[0] + 0.7152f * e[1] + 0.0722f * e[2];
            if (luminance >= min_emission)
            {
                next[code >> 5] |= 1u << (code & 31u);
            }
        }
        changed = !codes_valid || next != codes;
        codes = std::move(next);
        codes_valid = true;
    }
};

namespace cx {
namespace {

int Prepare(CxLightTree* t, const float* materials, const CxLightSceneTables& tables, uint32_t* requeue)
{
    // allocations first: a failure leaves the CPU state as it was
    const uint64_t groups = std::max<uint64_t>(1, tables.group_capacity);
    const uint64_t entities = std::max<uint64_t>(1, tables.entity_capacity);
    if (int r = t->rows.Reserve(std::max(groups, kGroupsInitial) * sizeof(GroupLightRow), "light tree rows"); r != CX_OK)
    {
        return r;
    }
    if (int r = t->staged_flags.Reserve(std::max(groups, kGroupsInitial) * 4, "light tree staged flags"); r != CX_OK)
    {
        return r;
    }
    if (int r = t->entity_motion.Reserve(std::max(entities, kEntitiesInitial) * 4, "light tree motion words"); r != CX_OK)
    {
        return r;
    }
    if (t->grow_request_units != 0)
    {
        t->GrowHeap(t->grow_request_units);
        t->grow_request_units = 0;
    }
    for (uint32_t tree = 0; tree < kLtTrees; ++tree)
    {
        if (t->leaf_request[tree] > t->leaf_capacity[tree] || t->node_request[tree] > t->node_capacity[tree])
        {
            const uint32_t want = std::max(t->leaf_capacity[tree], (t->leaf_request[tree] + 63) / 64 * 64);
            const uint32_t n = std::max({t->node_capacity[tree], NodeCapacityFor(want), std::min(t->node_request[tree], want)});
            if (int r = t->leaves[tree].Reserve(uint64_t(want) * sizeof(LightBvhLeaf), "light tree leaves"); r != CX_OK)
            {
                return r;
            }
            if (int r = t->nodes[tree].Reserve(uint64_t(n) * sizeof(LightBvhNode), "light tree nodes"); r != CX_OK)
            {
                return r;
            }
            if (int r = t->bvh_scratch[tree].Reserve(cx_light_bvh_scratch_bytes(want), "light tree build scratch");
                r != CX_OK)
            {
                return r;
            }
            Log(1, "light tree: BVH %u grows from %u to %u leaves, %u to %u nodes", tree, t->leaf_capacity[tree], want,
                t->node_capacity[tree], n);
            t->leaf_capacity[tree] = want;
            t->node_capacity[tree] = n;
            t->scratch_init[tree] = false;
            ++t->tree_epoch[tree];
            t->tree_reason[tree] = true;
        }
        t->leaf_request[tree] = 0;
        t->node_request[tree] = 0;
    }
    bool has_free = false;
    for (const ReadbackSlot& slot : t->readback)
    {
        has_free = has_free || !slot.pending;
    }
    if (!has_free)
    {
        ReadbackSlot slot;
        if (int r = HipResult(hipHostMalloc(reinterpret_cast<void**>(&slot.host), sizeof(LightControl), hipHostMallocDefault),
                              "light tree readback");
            r != CX_OK)
        {
            return r;
        }
        if (int r = HipResult(hipEventCreate(&slot.done), "light tree readback event"); r != CX_OK)
        {
            (void)hipHostFree(slot.host);
            return r;
        }
        if (int r = HipResult(hipEventCreate(&slot.start), "light tree timer event"); r != CX_OK)
        {
            (void)hipEventDestroy(slot.done);
            (void)hipHostFree(slot.host);
            return r;
        }
        t->readback.push_back(slot);
    }
    uint32_t again = 0;
    if (t->materials_changed)
    {
        bool changed = false;
        t->RebuildCodes(materials, changed);
        t->materials_changed = false;
        if (changed)
        {
            t->codes_upload = true;
            again = 1;
        }
    }
    t->tables = tables;
    if (requeue != nullptr)
    {
        *requeue = again;
    }
    return CX_OK;
}

int StageFlags(CxLightTree* t, bool finalize_ran, uint32_t* staged)
{
    t->record_pass = false;
    t->record_tree = {};
    if (finalize_ran || !t->staged_groups.empty())
    {
        t->pass_pending = true;
    }
    uint32_t lo = 0, hi = 0;
    if (!t->staged_groups.empty())
    {
        lo = *std::min_element(t->staged_groups.begin(), t->staged_groups.end());
        hi = *std::max_element(t->staged_groups.begin(), t->staged_groups.end()) + 1;
    }
    std::vector<uint32_t> dirty;
    for (uint32_t e = 0; e < t->entities.size(); ++e)
    {
        if (t->entities[e].motion_dirty)
        {
            dirty.push_back(e);
        }
    }
    const bool pass = t->pass_pending && t->tables.group_capacity > 0;
    const bool trees = t->tree_reason[0] || t->tree_reason[1] || pass;
    if (hi == lo && dirty.empty() && !trees)
    {
        *staged = 0;
        return CX_OK;
    }
    if (uint64_t(hi) * 4 > t->staged_flags.bytes || (!dirty.empty() && uint64_t(dirty.back() + 1) * 4 > t->entity_motion.bytes))
    {
        return ArgError("cx_light_tree_stage: a hooked group or entity is beyond the capacities given to prepare");
    }
    if (hi > lo)
    {
        uint32_t* flags = reinterpret_cast<uint32_t*>(
            t->Stage(uint64_t(hi - lo) * 4, static_cast<uint32_t*>(t->staged_flags.ptr) + lo));
        if (flags == nullptr)
        {
            Log(0, "light tree: no pinned staging memory; the update waits for the next flush");
            *staged = 0;
            return CX_OK;
        }
        std::fill(flags, flags + (hi - lo), 0u);
        for (const uint32_t g : t->staged_groups)
        {
            CxLightTree::GroupState& s = t->groups[g];
            flags[g - lo] = s.alive ? (kLtCpuRebuild | (s.fresh ? kLtCpuFresh : 0u)) : kLtCpuReleased;
            s.fresh = false;
            s.listed = false;
        }
        t->staged_groups.clear();
    }
    for (const uint32_t e : dirty)
    {
        uint8_t* word = t->Stage(4, static_cast<uint32_t*>(t->entity_motion.ptr) + e);
        if (word == nullptr)
        {
            *staged = 0;
            return CX_OK;
        }
        std::memcpy(word, &t->entities[e].motion, 4);
        t->entities[e].motion_dirty = false;
    }
    t->record_pass = pass;
    t->pass_group_count = t->tables.group_capacity;
    t->pass_pending = false;
    for (uint32_t tree = 0; tree < kLtTrees; ++tree)
    {
        t->record_tree[tree] = t->tree_reason[tree] || t->record_pass;
    }
    *staged = 1;
    return CX_OK;
}

int RecordTransfers(CxLightTree* t, hipStream_t s)
{
    if (t->fills_pending)
    {
        // a new tree: the zero fills cover these buffers whole
        t->DeferSource(t->rows);
        t->DeferSource(t->staged_flags);
        t->DeferSource(t->entity_motion);
    }
    else
    {
        t->TakeGrowth(t->rows, t->rows.bytes);
        if (t->staged_flags.source != nullptr)
        {
            // the staged flags are all zero between flushes: the new buffer is zero-filled
            t->DeferSource(t->staged_flags);
            t->growth.push_back(Growth{nullptr, t->staged_flags.ptr, 0, 0, t->staged_flags.bytes});
        }
        t->TakeGrowth(t->entity_motion, t->entity_motion.bytes);
    }
    for (uint32_t tree = 0; tree < kLtTrees; ++tree)
    {
        // the old tree stays valid until its rebuild
        t->TakeGrowth(t->leaves[tree], 0);
        t->TakeGrowth(t->nodes[tree], 0);
        t->DeferSource(t->bvh_scratch[tree]);
    }
    t->DeferSource(t->pass_scratch);
    const LightTreeDescriptor d = t->BuildDescriptor();
    if (std::memcmp(&d, &t->descriptor, sizeof(d)) != 0)
    {
        t->descriptor = d;
        t->descriptor_upload = true;
    }
    const bool work = t->fills_pending || t->descriptor_upload || t->codes_upload || !t->growth.empty() || t->arena >= 0;
    if (!work)
    {
        return CX_OK;
    }
    for (const Growth& g : t->growth)
    {
        if (g.source != nullptr && g.copy_bytes > 0)
        {
            if (int r = HipResult(hipMemcpyAsync(g.destination, g.source, g.copy_bytes, hipMemcpyDeviceToDevice, s),
                                  "light tree growth copy");
                r != CX_OK)
            {
                return r;
            }
        }
        if (g.fill_to > g.fill_from)
        {
            Zero(static_cast<uint8_t*>(g.destination) + g.fill_from, g.fill_to - g.fill_from, s);
        }
    }
    if (t->fills_pending)
    {
        // zero control, both trees empty (root none) until their first build
        LightControl* c = reinterpret_cast<LightControl*>(t->Stage(sizeof(LightControl), t->control));
        if (c == nullptr)
        {
            return HipResult(hipErrorOutOfMemory, "light tree staging");
        }
        std::memset(c, 0, sizeof(LightControl));
        for (LightTreeHeader& h : c->headers)
        {
            h.root = kLtNone;
        }
        Zero(t->rows.ptr, t->rows.bytes, s);
        Zero(t->staged_flags.ptr, t->staged_flags.bytes, s);
        Zero(t->entity_motion.ptr, t->entity_motion.bytes, s);
        t->fills_pending = false;
    }
    if (t->descriptor_upload)
    {
        uint8_t* p = t->Stage(sizeof(LightTreeDescriptor), t->descriptor_device);
        if (p == nullptr)
        {
            return HipResult(hipErrorOutOfMemory, "light tree staging");
        }
        std::memcpy(p, &t->descriptor, sizeof(LightTreeDescriptor));
        t->descriptor_upload = false;
    }
    if (t->codes_upload)
    {
        uint8_t* p = t->Stage(kCodeWords * 4, t->codes_device);
        if (p == nullptr)
        {
            return HipResult(hipErrorOutOfMemory, "light tree staging");
        }
        std::memcpy(p, t->codes.data(), kCodeWords * 4);
        t->codes_upload = false;
    }
    // the arena's copies: the control and descriptor first were staged after the flags of Stage, but
    // every destination is written once, so their order does not matter; they follow the fills
    if (t->arena >= 0)
    {
        Arena& a = t->arenas[t->arena];
        for (const Arena::Copy& c : a.copies)
        {
            if (int r = HipResult(hipMemcpyAsync(c.destination, a.host + c.offset, c.bytes, hipMemcpyHostToDevice, s),
                                  "light tree upload");
                r != CX_OK)
            {
                return r;
            }
        }
        if (int r = HipResult(hipEventRecord(a.done, s), "light tree upload event"); r != CX_OK)
        {
            return r;
        }
        a.in_flight = true;
        a.copies.clear();
        t->arena = -1;
    }
    return LaunchResult("cx_light_tree_record_transfers");
}

CxLightPass PassOf(const CxLightTree* t)
{
    return CxLightPass{t->descriptor_device, t->heap_units, static_cast<uint32_t*>(t->staged_flags.ptr), t->pass_scratch.ptr,
                       std::max(t->pass_group_count, 1u), 0};
}

// LightTree::RecordTree (light_tree_record.cpp:230-309): begin (the GPU decides whether the tree is
// rebuilt), gather, build.
int RecordTree(CxLightTree* t, uint32_t tree, uint32_t reason, uint32_t pass_recorded, hipStream_t s)
{
    const CxLightPass pass = PassOf(t);
    LightTreeState* state = &t->control->states[tree];
    if (!t->scratch_init[tree])
    {
        if (int r = cx_light_bvh_scratch_init(t->bvh_scratch[tree].ptr, t->leaf_capacity[tree], s); r != CX_OK)
        {
            return r;
        }
        t->scratch_init[tree] = true;
    }
    if (int r = cx_light_bvh_begin(t->control, state, tree, reason, pass_recorded, s); r != CX_OK)
    {
        return r;
    }
    if (int r = cx_light_gather(&pass, tree, static_cast<LightBvhLeaf*>(t->leaves[tree].ptr), t->leaf_capacity[tree],
                                state, s);
        r != CX_OK)
    {
        return r;
    }
    CxLightBvhBuild b{static_cast<LightBvhLeaf*>(t->leaves[tree].ptr), static_cast<LightBvhNode*>(t->nodes[tree].ptr),
                      &t->control->headers[tree], state, t->bvh_scratch[tree].ptr, t->leaf_capacity[tree],
                      t->node_capacity[tree]};
    return cx_light_bvh_build(&b, s);
}

int Record(CxLightTree* t, const CxScene* scene, uint64_t frame, hipStream_t s)
{
    if (!t->record_pass && !t->record_tree[0] && !t->record_tree[1])
    {
        return CX_OK;
    }
    if (scene == nullptr)
    {
        return ArgError("cx_light_tree_record: the scene is NULL");
    }
    const uint64_t pass_bytes = cx_light_pass_scratch_bytes(std::max(scene->bricks, t->tables.brick_capacity),
                                                            std::max(t->pass_group_count, 1u));
    if (pass_bytes > t->pass_scratch.bytes)
    {
        void* old = t->pass_scratch.ptr;
        t->pass_scratch.ptr = nullptr;
        t->pass_scratch.bytes = 0;
        if (int r = t->pass_scratch.Reserve(pass_bytes, "light pass scratch"); r != CX_OK)
        {
            t->pass_scratch.ptr = old;
            return r;
        }
        t->pass_scratch.source = nullptr;
        t->Defer(old);
    }
    ReadbackSlot* slot = nullptr;
    for (ReadbackSlot& r : t->readback)
    {
        if (!r.pending)
        {
            slot = &r;
            break;
        }
    }
    if (slot != nullptr)
    {
        (void)hipEventRecord(slot->start, s);
    }
    const CxLightPass pass = PassOf(t);
    if (t->record_pass)
    {
        if (int r = cx_light_pass(scene, &pass, s); r != CX_OK)
        {
            return r;
        }
    }
    for (uint32_t tree = 0; tree < kLtTrees; ++tree)
    {
        if (t->record_tree[tree])
        {
            if (int r = RecordTree(t, tree, t->tree_reason[tree] ? 1u : 0u, t->record_pass ? 1u : 0u, s); r != CX_OK)
            {
                return r;
            }
            t->tree_reason[tree] = false;
        }
    }
    if (slot != nullptr)
    {
        if (int r = HipResult(hipMemcpyAsync(slot->host, t->control, sizeof(LightControl), hipMemcpyDeviceToHost, s),
                              "light tree readback");
            r != CX_OK)
        {
            return r;
        }
        if (int r = HipResult(hipEventRecord(slot->done, s), "light tree readback event"); r != CX_OK)
        {
            return r;
        }
        slot->frame = frame;
        slot->heap_epoch = t->heap_epoch;
        slot->tree_epoch = t->tree_epoch;
        slot->pending = true;
    }
    return CX_OK;
}

void EndFlush(CxLightTree* t, uint64_t frame)
{
    for (Growth& g : t->growth)
    {
        if (g.source != nullptr)
        {
            t->deferred.push_back(Deferred{frame, g.source});
        }
    }
    t->growth.clear();
    for (Deferred& d : t->deferred)
    {
        d.frame = d.frame == 0 ? frame : d.frame;
    }
    t->record_pass = false;
    t->record_tree = {};
}

int Recycle(CxLightTree* t, uint64_t safe)
{
    t->safe_frame = std::max(t->safe_frame, safe);
    std::vector<Deferred> keep;
    for (const Deferred& d : t->deferred)
    {
        if (d.frame != 0 && d.frame <= t->safe_frame)
        {
            (void)hipFree(d.ptr);
        }
        else
        {
            keep.push_back(d);
        }
    }
    t->deferred.swap(keep);
    for (ReadbackSlot& slot : t->readback)
    {
        if (!slot.pending || slot.frame > t->safe_frame)
        {
            continue;
        }
        if (int r = HipResult(hipEventSynchronize(slot.done), "light tree readback"); r != CX_OK)
        {
            return r;
        }
        slot.pending = false;
        LightControl c;
        std::memcpy(&c, slot.host, sizeof(c));
        if (slot.frame >= t->last_control_frame)
        {
            t->last_control = c;
            t->last_control_frame = slot.frame;
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, slot.start, slot.done) == hipSuccess)
            {
                t->gpu_ms = ms;
            }
        }
        if (c.pass_result == kLtPassFailed)
        {
            // retry: the flags are still set on the GPU and the rows unchanged; grow first unless the
            // heap has grown since that pass
            t->pass_pending = true;
            if (slot.heap_epoch == t->heap_epoch)
            {
                Log(0, "light tree: the group pass of frame %llu needs %llu heap bytes of %llu: retried after growth",
                    static_cast<unsigned long long>(slot.frame),
                    static_cast<unsigned long long>(uint64_t(c.needed_units) * kLtUnitBytes),
                    static_cast<unsigned long long>(t->heap_units * kLtUnitBytes));
                t->grow_request_units =
                    std::max<uint64_t>(t->grow_request_units, std::max<uint64_t>(c.needed_units, t->heap_units + 1));
            }
        }
        else if (c.pass_result == kLtPassOk && slot.heap_epoch == t->heap_epoch &&
                 uint64_t(c.needed_units) * 4 > t->heap_units * 3)
        {
            t->grow_request_units = std::max<uint64_t>(t->grow_request_units, c.needed_units);
        }
        for (uint32_t tree = 0; tree < kLtTrees; ++tree)
        {
            if (slot.tree_epoch[tree] != t->tree_epoch[tree])
            {
                continue;  // about a capacity that has grown since
            }
            const LightTreeState& st = c.states[tree];
            const uint64_t found = st.leaves_found;
            if (found * 4 > uint64_t(t->leaf_capacity[tree]) * 3 || st.leaves_dropped != 0)
            {
                const uint64_t wanted = std::max<uint64_t>(uint64_t(t->leaf_capacity[tree]) * 3 / 2, found + found / 2);
                const uint32_t request = uint32_t(std::min<uint64_t>(wanted, kMaxLeaves));
                t->leaf_request[tree] = std::max(t->leaf_request[tree], request);
            }
            if (st.node_overflow != 0)
            {
                t->node_request[tree] = t->leaf_capacity[tree];
            }
        }
    }
    return CX_OK;
}

bool ValidMotion(uint32_t motion)
{
    return motion == CX_ENTITY_STATIC || motion == CX_ENTITY_DYNAMIC;
}

}  // namespace
}  // namespace cx

extern "C" {

CX_API int cx_light_tree_create(uint32_t group_shift, uint32_t encoding, const CxLightOptions* options,
                                CxLightTree** out)
{
    if (out == nullptr || group_shift == 0 || group_shift > 7 || encoding > 1)
    {
        return cx::ArgError("cx_light_tree_create: out NULL, group_shift not in 1..7 or encoding not 0/1");
    }
    float min_emission = options != nullptr ? options->min_emission : 0.0f;
    uint64_t heap_bytes = options != nullptr ? options->heap_bytes : 0;
    if (!(min_emission >= 0.0f) || !(min_emission <= 3.4e38f) || heap_bytes > cx::kMaxHeapUnits * kLtUnitBytes)
    {
        return cx::ArgError("cx_light_tree_create: min_emission must be finite and >= 0, heap_bytes <= 2^32 x 16");
    }
    min_emission = min_emission == 0.0f ? cx::kDefaultMinEmission : min_emission;
    heap_bytes = heap_bytes == 0 ? cx::kDefaultHeapBytes : heap_bytes;
    CxLightTree* t = new (std::nothrow) CxLightTree();
    if (t == nullptr)
    {
        return CX_ERROR_OUT_OF_MEMORY;
    }
    t->group_shift = group_shift;
    t->encoding = encoding;
    t->min_emission = min_emission;
    t->codes.assign(cx::kCodeWords, 0u);
    t->heap_units = (heap_bytes + kLtUnitBytes - 1) / kLtUnitBytes;
    int r = cx::HipResult(hipMalloc(reinterpret_cast<void**>(&t->codes_device), cx::kCodeWords * 4), "light codes");
    r = r == CX_OK ? cx::HipResult(hipMalloc(reinterpret_cast<void**>(&t->control), sizeof(LightControl)), "control") : r;
    r = r == CX_OK ? cx::HipResult(hipMalloc(&t->heaps[0], t->heap_units * kLtUnitBytes), "heap") : r;
    r = r == CX_OK ? cx::HipResult(hipMalloc(&t->heaps[1], t->heap_units * kLtUnitBytes), "heap") : r;
    r = r == CX_OK ? cx::HipResult(hipMalloc(reinterpret_cast<void**>(&t->descriptor_device), sizeof(LightTreeDescriptor)),
                                   "descriptor")
                   : r;
    r = r == CX_OK ? t->rows.Reserve(cx::kGroupsInitial * sizeof(GroupLightRow), "rows") : r;
    r = r == CX_OK ? t->staged_flags.Reserve(cx::kGroupsInitial * 4, "staged flags") : r;
    r = r == CX_OK ? t->entity_motion.Reserve(cx::kEntitiesInitial * 4, "motion words") : r;
    for (uint32_t tree = 0; tree < kLtTrees && r == CX_OK; ++tree)
    {
        t->leaf_capacity[tree] = cx::kLeavesInitial;
        t->node_capacity[tree] = cx::NodeCapacityFor(cx::kLeavesInitial);
        r = t->leaves[tree].Reserve(uint64_t(cx::kLeavesInitial) * sizeof(LightBvhLeaf), "leaves");
        r = r == CX_OK ? t->nodes[tree].Reserve(uint64_t(t->node_capacity[tree]) * sizeof(LightBvhNode), "nodes") : r;
        r = r == CX_OK ? t->bvh_scratch[tree].Reserve(cx_light_bvh_scratch_bytes(cx::kLeavesInitial), "scratch") : r;
    }
    for (int i = 0; i < 2 && r == CX_OK; ++i)
    {
        r = cx::HipResult(hipEventCreateWithFlags(&t->arenas[i].done, hipEventDisableTiming), "staging event");
    }
    if (r != CX_OK)
    {
        cx_light_tree_destroy(t);
        return r;
    }
    cx::Log(1, "light tree created: min_emission %g, heap 2 x %llu bytes", double(min_emission),
            static_cast<unsigned long long>(t->heap_units * kLtUnitBytes));
    *out = t;
    return CX_OK;
}

CX_API void cx_light_tree_destroy(CxLightTree* t)
{
    if (t == nullptr)
    {
        return;
    }
    (void)hipDeviceSynchronize();
    (void)hipFree(t->codes_device);
    (void)hipFree(t->control);
    (void)hipFree(t->heaps[0]);
    (void)hipFree(t->heaps[1]);
    (void)hipFree(t->descriptor_device);
    for (cx::Growable* g : {&t->rows, &t->staged_flags, &t->entity_motion, &t->pass_scratch})
    {
        (void)hipFree(g->ptr);
        (void)hipFree(g->source);
    }
    for (uint32_t tree = 0; tree < kLtTrees; ++tree)
    {
        for (cx::Growable* g : {&t->leaves[tree], &t->nodes[tree], &t->bvh_scratch[tree]})
        {
            (void)hipFree(g->ptr);
            (void)hipFree(g->source);
        }
    }
    for (const cx::Growth& g : t->growth)
    {
        (void)hipFree(g.source);
    }
    for (const cx::Deferred& d : t->deferred)
    {
        (void)hipFree(d.ptr);
    }
    for (cx::Arena& a : t->arenas)
    {
        (void)hipHostFree(a.host);
        if (a.done != nullptr)
        {
            (void)hipEventDestroy(a.done);
        }
    }
    for (cx::ReadbackSlot& s : t->readback)
    {
        (void)hipHostFree(s.host);
        (void)hipEventDestroy(s.done);
        (void)hipEventDestroy(s.start);
    }
    delete t;
}

CX_API int cx_light_tree_on_entity_added(CxLightTree* t, uint32_t entity, uint32_t motion)
{
    if (t == nullptr || !cx::ValidMotion(motion))
    {
        return cx::ArgError("cx_light_tree_on_entity_added: tree NULL or motion not static/dynamic");
    }
    CxLightTree::EntityState& s = t->Entity(entity);
    s.motion = motion;
    s.motion_dirty = true;
    return CX_OK;
}

CX_API int cx_light_tree_on_entity_removed(CxLightTree* t, uint32_t entity)
{
    if (t == nullptr)
    {
        return cx::ArgError("cx_light_tree_on_entity_removed: tree NULL");
    }
    // its groups are released (a group pass follows); its tree loses the leaves
    t->tree_reason[t->Entity(entity).motion] = true;
    return CX_OK;
}

CX_API int cx_light_tree_on_entity_moved(CxLightTree* t, uint32_t entity)
{
    if (t == nullptr)
    {
        return cx::ArgError("cx_light_tree_on_entity_moved: tree NULL");
    }
    t->tree_reason[t->Entity(entity).motion] = true;
    return CX_OK;
}

CX_API int cx_light_tree_on_entity_motion(CxLightTree* t, uint32_t entity, uint32_t motion)
{
    if (t == nullptr || !cx::ValidMotion(motion))
    {
        return cx::ArgError("cx_light_tree_on_entity_motion: tree NULL or motion not static/dynamic");
    }
    CxLightTree::EntityState& s = t->Entity(entity);
    if (s.motion != motion)
    {
        s.motion = motion;
        s.motion_dirty = true;
        t->tree_reason[0] = true;
        t->tree_reason[1] = true;
    }
    return CX_OK;
}

CX_API int cx_light_tree_on_group_created(CxLightTree* t, uint32_t group)
{
    if (t == nullptr)
    {
        return cx::ArgError("cx_light_tree_on_group_created: tree NULL");
    }
    CxLightTree::GroupState& s = t->Group(group);
    s.alive = true;
    s.fresh = true;
    t->ListGroup(group);
    return CX_OK;
}

CX_API int cx_light_tree_on_group_released(CxLightTree* t, uint32_t group, uint32_t motion)
{
    if (t == nullptr || !cx::ValidMotion(motion))
    {
        return cx::ArgError("cx_light_tree_on_group_released: tree NULL or motion not static/dynamic");
    }
    t->Group(group).alive = false;
    t->ListGroup(group);
    t->tree_reason[motion] = true;
    return CX_OK;
}

CX_API int cx_light_tree_on_group_bricks_removed(CxLightTree* t, uint32_t group)
{
    if (t == nullptr)
    {
        return cx::ArgError("cx_light_tree_on_group_bricks_removed: tree NULL");
    }
    t->ListGroup(group);
    return CX_OK;
}

CX_API int cx_light_tree_on_materials_changed(CxLightTree* t)
{
    if (t == nullptr)
    {
        return cx::ArgError("cx_light_tree_on_materials_changed: tree NULL");
    }
    t->materials_changed = true;
    return CX_OK;
}

CX_API int cx_light_tree_set_min_emission(CxLightTree* t, float min_emission)
{
    if (t == nullptr || !(min_emission >= 0.0f) || !(min_emission <= 3.4e38f))
    {
        return cx::ArgError("cx_light_tree_set_min_emission: tree NULL or min_emission not finite and >= 0");
    }
    min_emission = min_emission == 0.0f ? cx::kDefaultMinEmission : min_emission;
    if (min_emission != t->min_emission)
    {
        t->min_emission = min_emission;
        t->materials_changed = true;
    }
    return CX_OK;
}

CX_API int cx_light_tree_prepare(CxLightTree* t, const float* materials, const CxLightSceneTables* tables,
                                 uint32_t* requeue)
{
    if (t == nullptr || tables == nullptr || (materials == nullptr && t->materials_changed))
    {
        return cx::ArgError("cx_light_tree_prepare: tree, tables or (with changed materials) materials NULL");
    }
    if (tables->group_capacity > 0 && (tables->group_infos == nullptr || tables->entity_infos == nullptr))
    {
        return cx::ArgError("cx_light_tree_prepare: groups without group or entity infos");
    }
    if ((uint64_t(tables->group_capacity) << (3 * t->group_shift)) > 0xFFFFFFFFull)
    {
        return cx::ArgError("cx_light_tree_prepare: %u groups above 2^32 / 8^%u", tables->group_capacity, t->group_shift);
    }
    return cx::Prepare(t, materials, *tables, requeue);
}

CX_API int cx_light_tree_stage(CxLightTree* t, uint32_t finalize_ran, uint32_t* staged)
{
    if (t == nullptr || staged == nullptr)
    {
        return cx::ArgError("cx_light_tree_stage: tree or staged NULL");
    }
    return cx::StageFlags(t, finalize_ran != 0, staged);
}

CX_API int cx_light_tree_record_transfers(CxLightTree* t, void* stream)
{
    if (t == nullptr)
    {
        return cx::ArgError("cx_light_tree_record_transfers: tree NULL");
    }
    return cx::RecordTransfers(t, static_cast<hipStream_t>(stream));
}

CX_API int cx_light_tree_record(CxLightTree* t, const CxScene* scene, uint64_t frame, void* stream)
{
    if (t == nullptr)
    {
        return cx::ArgError("cx_light_tree_record: tree NULL");
    }
    return cx::Record(t, scene, frame, static_cast<hipStream_t>(stream));
}

CX_API int cx_light_tree_rebuild(CxLightTree* t, uint32_t tree, void* stream)
{
    if (t == nullptr || tree >= kLtTrees || t->pass_scratch.ptr == nullptr || !t->scratch_init[tree])
    {
        return cx::ArgError("cx_light_tree_rebuild: tree NULL, tree index >= 2, or no flush recorded the tree yet");
    }
    return cx::RecordTree(t, tree, 1u, 0u, static_cast<hipStream_t>(stream));
}

CX_API int cx_light_tree_end_flush(CxLightTree* t, uint64_t frame)
{
    if (t == nullptr)
    {
        return cx::ArgError("cx_light_tree_end_flush: tree NULL");
    }
    cx::EndFlush(t, frame);
    return CX_OK;
}

CX_API int cx_light_tree_recycle(CxLightTree* t, uint64_t safe_frame)
{
    if (t == nullptr)
    {
        return cx::ArgError("cx_light_tree_recycle: tree NULL");
    }
    return cx::Recycle(t, safe_frame);
}

CX_API void cx_light_tree_get_stats(const CxLightTree* t, CxLightStats* out)
{
    if (out == nullptr)
    {
        return;
    }
    *out = CxLightStats{};
    if (t == nullptr)
    {
        return;
    }
    const LightControl& c = t->last_control;
    out->enabled = 1;
    out->light_groups = c.light_groups;
    out->light_bricks = c.light_bricks;
    out->light_faces = c.light_faces;
    out->bricks_dropped = c.bricks_dropped;
    out->heap_bytes_used = uint64_t(c.needed_units) * kLtUnitBytes;
    out->heap_bytes_capacity = t->heap_units * kLtUnitBytes;
    out->passes = c.passes;
    out->failed_passes = c.failed_passes;
    out->records_recomputed = c.records_recomputed;
    out->records_copied = c.records_copied;
    out->gpu_ms = t->gpu_ms;
    for (uint32_t tree = 0; tree < kLtTrees; ++tree)
    {
        out->tree_leaves[tree] = c.headers[tree].leaf_count;
        out->tree_nodes[tree] = c.headers[tree].node_count;
        out->tree_builds[tree] = c.states[tree].builds;
        out->tree_leaves_dropped[tree] = c.states[tree].leaves_dropped;
    }
}

CX_API uint32_t cx_light_tree_pending(const CxLightTree* t)
{
    if (t == nullptr)
    {
        return 0;
    }
    bool pending = t->pass_pending || !t->staged_groups.empty() || t->tree_reason[0] || t->tree_reason[1] ||
                   t->grow_request_units != 0 || t->materials_changed;
    for (uint32_t tree = 0; tree < kLtTrees; ++tree)
    {
        pending = pending || t->leaf_request[tree] > t->leaf_capacity[tree] || t->node_request[tree] > t->node_capacity[tree];
    }
    for (const cx::ReadbackSlot& slot : t->readback)
    {
        pending = pending || slot.pending;
    }
    return pending ? 1u : 0u;
}

CX_API const LightTreeDescriptor* cx_light_tree_descriptor(const CxLightTree* t)
{
    return t != nullptr ? t->descriptor_device : nullptr;
}

CX_API const uint32_t* cx_light_tree_light_codes(const CxLightTree* t)
{
    return t != nullptr ? t->codes_device : nullptr;
}

CX_API GroupLightRow* cx_light_tree_group_rows(const CxLightTree* t)
{
    return t != nullptr ? static_cast<GroupLightRow*>(t->rows.ptr) : nullptr;
}

}  // extern "C"
