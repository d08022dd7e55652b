// The light tree's device API for gfx950 kernels (shaders/light/caelix_light_tree.hlsl): importance
// of a summary for a point or a box receiver, the weighted reservoir pick, the walk down the 4-wide nodes
// to a leaf, the in-brick draw of a light face, the pdf walk up the parent links, brick references and
// the region (box) sampler. Layouts are the
// reference's (LightTreeHeader, LightBvhNode, LightBvhLeaf, brick light records).
//
// A leaf names its brick as (group, index entry); the group's row gives its light block in the current
// heap buffer, whose index entry gives the record (caelix_light_tree.hlsl:575-747). Receivers are taken
// into the group's local voxel frame through its entity's frame (R^T (x - t) - origin) and drawn points
// back to the world (R local + t).
#pragma once
#include "light_block.h"
#include "trace_device.h"

namespace cx {

struct Summ
{
    float face[6];
    V3 lo;
    V3 hi;
};

CX_DEV float Power6(const float* f)
{
    return f[0] + f[1] + f[2] + f[3] + f[4] + f[5];
}

CX_DEV Summ NodeSlot(const LightBvhNode* node, uint32_t s)
{
    Summ r;
#pragma unroll
    for (int f = 0; f < 6; ++f)
    {
        r.face[f] = node->child_face[s][f];
    }
    r.lo = Make(node->child_lo[s][0], node->child_lo[s][1], node->child_lo[s][2]);
    r.hi = Make(node->child_hi[s][0], node->child_hi[s][1], node->child_hi[s][2]);
    return r;
}

CX_DEV Summ HeaderSummary(const LightTreeHeader* h)
{
    Summ r;
#pragma unroll
    for (int f = 0; f < 6; ++f)
    {
        r.face[f] = h->summary.face[f];
    }
    r.lo = Make(h->summary.lo[0], h->summary.lo[1], h->summary.lo[2]);
    r.hi = Make(h->summary.hi[0], h->summary.hi[1], h->summary.hi[2]);
    return r;
}

// LtEmitterCosines (:85-99) over D = [d_lo, d_hi].
CX_DEV void EmitterCosines(V3 d_lo, V3 d_hi, float c[6])
{
    const V3 closest = Make(fminf(fmaxf(0.0f, d_lo.x), d_hi.x), fminf(fmaxf(0.0f, d_lo.y), d_hi.y),
                            fminf(fmaxf(0.0f, d_lo.z), d_hi.z));
    const V3 c2 = MulV(closest, closest);
    const float all2 = c2.x + c2.y + c2.z;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        const float others = all2 - Get(c2, a);
        const float dh = Get(d_hi, a);
        const float dl = Get(d_lo, a);
        c[2 * a] = dh > 0.0f ? Div(dh, Sqrt(dh * dh + others)) : 0.0f;
        c[2 * a + 1] = dl < 0.0f ? Div(-dl, Sqrt(dl * dl + others)) : 0.0f;
    }
}

// A receiver of the walk (LtReceiver, :368-395): the box [lo, hi] (a point: lo = hi) and the unit normal
// n (0: no surface). Every walk function takes kPoint as a template constant: the point form's code is
// the box form with the box terms removed.
struct Receiver
{
    V3 lo;
    V3 hi;
    V3 n;
};

CX_DEV Receiver PointReceiver(V3 x, V3 n)
{
    return Receiver{x, x, n};
}

// LtImportance (:118-155): the importance of a summary (box [lo, hi], face sums) for the receiver: the
// emitter cosines over D = [r_lo - hi, r_hi - lo], the receiver term cr = min(1, m / dmin), the
// distance term max(|c_R - c_B|^2, max(1, extent^2)).
template <bool kPoint>
CX_DEV float ImportanceR(const Receiver& r, V3 lo, V3 hi, const float face[6])
{
    float c[6];
    EmitterCosines(Sub(r.lo, hi), Sub(r.hi, lo), c);
    float e = 0.0f;
#pragma unroll
    for (int f = 0; f < 6; ++f)
    {
        e += face[f] * c[f];
    }
    if (e <= 0.0f)
    {
        return 0.0f;
    }
    const V3 centre = Mul(Add(lo, hi), 0.5f);
    const V3 half_size = Mul(Sub(hi, lo), 0.5f);
    const V3 r_centre = kPoint ? r.lo : Mul(Add(r.lo, r.hi), 0.5f);
    const V3 n = r.n;
    float cr = 1.0f;
    if (n.x != 0.0f || n.y != 0.0f || n.z != 0.0f)
    {
        const V3 abs_n = Make(fabsf(n.x), fabsf(n.y), fabsf(n.z));
        float m = Dot(n, Sub(centre, r_centre)) + Dot(abs_n, half_size);
        if (!kPoint)
        {
            m += Dot(abs_n, Mul(Sub(r.hi, r.lo), 0.5f));
        }
        if (m <= 0.0f)
        {
            return 0.0f;
        }
        const V3 gap = Make(fmaxf(fmaxf(lo.x - r.hi.x, r.lo.x - hi.x), 0.0f), fmaxf(fmaxf(lo.y - r.hi.y, r.lo.y - hi.y), 0.0f),
                            fmaxf(fmaxf(lo.z - r.hi.z, r.lo.z - hi.z), 0.0f));
        const float dmin = Sqrt(Dot(gap, gap));
        cr = dmin > 0.0f ? fminf(1.0f, Div(m, dmin)) : 1.0f;
    }
    const V3 to_centre = Sub(r_centre, centre);
    float extent2 = Dot(half_size, half_size);
    if (!kPoint)
    {
        const V3 r_half = Mul(Sub(r.hi, r.lo), 0.5f);
        const float extent = Sqrt(Dot(half_size, half_size)) + Sqrt(Dot(r_half, r_half));
        extent2 = extent * extent;
    }
    const float d2 = fmaxf(Dot(to_centre, to_centre), fmaxf(1.0f, extent2));
    return Div(e * cr, d2);
}

// LightTreeImportance (:157-160): the point form.
CX_DEV float Importance(V3 x, V3 n, V3 lo, V3 hi, const float face[6])
{
    return ImportanceR<true>(PointReceiver(x, n), lo, hi, face);
}

struct Pick
{
    float sum_i, sum_p, weight_i, weight_p;
    uint32_t index_i, index_p;
};
CX_DEV Pick PickBegin()
{
    return Pick{0.0f, 0.0f, 0.0f, 0.0f, kLtNone, kLtNone};
}
// LtPickAdd (:187-205)
CX_DEV void PickAdd(Pick& p, float importance, float power, uint32_t index, float u)
{
    if (importance > 0.0f)
    {
        p.sum_i += importance;
        if (u * p.sum_i < importance)
        {
            p.weight_i = importance;
            p.index_i = index;
        }
    }
    if (power > 0.0f)
    {
        p.sum_p += power;
        if (u * p.sum_p < power)
        {
            p.weight_p = power;
            p.index_p = index;
        }
    }
}
CX_DEV uint32_t PickEnd(const Pick& p, float& probability)
{
    if (p.sum_i > 0.0f)
    {
        probability = Div(p.weight_i, p.sum_i);
        return p.index_i;
    }
    probability = p.sum_p > 0.0f ? Div(p.weight_p, p.sum_p) : 0.0f;
    return p.sum_p > 0.0f ? p.index_p : kLtNone;
}
CX_DEV float PickProbability(float sum_i, float sum_p, float importance, float power)
{
    if (sum_i > 0.0f)
    {
        return Div(importance, sum_i);
    }
    return sum_p > 0.0f ? Div(power, sum_p) : 0.0f;
}

CX_DEV uint32_t ChildCount(const uint32_t* child)
{
    uint32_t n = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s)
    {
        n += child[s] != kLtNone ? 1u : 0u;
    }
    return n;
}

// LtTreeSummary (:289-304): the root summary of tree k, power 0 (root kLtNone) for an empty tree or one
// LightTreeSelect left out.
CX_DEV Summ TreeSummary(const LightView& L, uint32_t k, uint32_t& root)
{
    root = ((L.trees >> k) & 1u) != 0 ? L.Header(k)->root : kLtNone;
    if (root == kLtNone)
    {
        Summ e;
#pragma unroll
        for (int f = 0; f < 6; ++f)
        {
            e.face[f] = 0.0f;
        }
        e.lo = Make(0.0f, 0.0f, 0.0f);
        e.hi = Make(0.0f, 0.0f, 0.0f);
        return e;
    }
    return HeaderSummary(L.Header(k));
}

// LtChooseTreeR (:417-437): a tree drawn by its root's importance (kLtNone when both are empty); pdf
// times its probability.
template <bool kPoint>
CX_DEV uint32_t ChooseTreeR(const LightView& L, const Receiver& r, uint32_t& rng, float& pdf, uint32_t& root)
{
    Pick pick = PickBegin();
    uint32_t roots[2];
    for (uint32_t k = 0; k < 2; ++k)
    {
        const Summ s = TreeSummary(L, k, roots[k]);
        PickAdd(pick, ImportanceR<kPoint>(r, s.lo, s.hi, s.face), Power6(s.face), k, LightRandomFloat(rng));
    }
    float probability;
    const uint32_t k = PickEnd(pick, probability);
    root = k != kLtNone ? roots[k] : kLtNone;
    if (k != kLtNone)
    {
        pdf *= probability;
    }
    return k;
}

// LtChooseLeafR (:466-512): a leaf of tree k drawn from its root (kLtNone when the draw fails); pdf
// times each choice's probability; leaf_importance = the importance of the drawn leaf's summary (the
// slot the last choice took; 0 when it fell back to the power).
template <bool kPoint>
CX_DEV uint32_t ChooseLeafR(const LightView& L, uint32_t k, uint32_t root, const Receiver& r, uint32_t& rng, float& pdf,
                            float& leaf_importance)
{
    leaf_importance = 0.0f;
    uint32_t node = root;
    for (uint32_t depth = 0; depth < 64; ++depth)
    {
        const LightBvhNode* address = L.Nodes(k) + node;
        uint32_t children[4];
        *reinterpret_cast<uint4*>(children) = *reinterpret_cast<const uint4*>(address->child);
        uint32_t chosen;
        if (ChildCount(children) == 1)
        {
            chosen = children[0];
            if (!kPoint && (chosen & kLtChildLeafBit) != 0)
            {
                const Summ slot = NodeSlot(address, 0);
                leaf_importance = ImportanceR<kPoint>(r, slot.lo, slot.hi, slot.face);
            }
        }
        else
        {
            Pick p = PickBegin();
            for (uint32_t s = 0; s < 4; ++s)
            {
                if (children[s] == kLtNone)
                {
                    continue;
                }
                const Summ slot = NodeSlot(address, s);
                PickAdd(p, ImportanceR<kPoint>(r, slot.lo, slot.hi, slot.face), Power6(slot.face), s, LightRandomFloat(rng));
            }
            float prob;
            const uint32_t s = PickEnd(p, prob);
            if (s == kLtNone)
            {
                return kLtNone;
            }
            pdf *= prob;
            chosen = children[s];
            leaf_importance = p.sum_i > 0.0f ? p.weight_i : 0.0f;
        }
        if ((chosen & kLtChildLeafBit) != 0)
        {
            return chosen & ~kLtChildLeafBit;
        }
        node = chosen;
    }
    return kLtNone;
}

// The point form of the two draws (LtChooseTree + LtChooseLeaf, :619-629): a leaf, its tree.
CX_DEV uint32_t ChooseLeaf(const LightView& L, V3 x, V3 n, uint32_t& rng, float& pdf, uint32_t& tree)
{
    const Receiver r = PointReceiver(x, n);
    uint32_t root;
    tree = ChooseTreeR<true>(L, r, rng, pdf, root);
    if (tree == kLtNone)
    {
        return kLtNone;
    }
    float leaf_importance;
    return ChooseLeafR<true>(L, tree, root, r, rng, pdf, leaf_importance);
}

// LtSlotProbabilityR (:515-546)
template <bool kPoint>
CX_DEV float SlotProbabilityR(const LightBvhNode* address, const Receiver& r, uint32_t slot)
{
    uint32_t children[4];
    *reinterpret_cast<uint4*>(children) = *reinterpret_cast<const uint4*>(address->child);
    if (ChildCount(children) == 1)
    {
        return 1.0f;
    }
    float sum_i = 0.0f;
    float sum_p = 0.0f;
    float importance_t = 0.0f;
    float power_t = 0.0f;
    for (uint32_t s = 0; s < 4; ++s)
    {
        if (children[s] == kLtNone)
        {
            continue;
        }
        const Summ summary = NodeSlot(address, s);
        const float importance = ImportanceR<kPoint>(r, summary.lo, summary.hi, summary.face);
        const float power = Power6(summary.face);
        sum_i += importance > 0.0f ? importance : 0.0f;
        sum_p += power > 0.0f ? power : 0.0f;
        if (s == slot)
        {
            importance_t = importance;
            power_t = power;
        }
    }
    return PickProbability(sum_i, sum_p, importance_t, power_t);
}

// LtTreeProbabilityR (:439-462)
template <bool kPoint>
CX_DEV float TreeProbabilityR(const LightView& L, const Receiver& r, uint32_t target)
{
    float sum_i = 0.0f;
    float sum_p = 0.0f;
    float importance_t = 0.0f;
    float power_t = 0.0f;
    for (uint32_t k = 0; k < 2; ++k)
    {
        uint32_t root;
        const Summ s = TreeSummary(L, k, root);
        const float importance = ImportanceR<kPoint>(r, s.lo, s.hi, s.face);
        const float power = Power6(s.face);
        sum_i += importance > 0.0f ? importance : 0.0f;
        sum_p += power > 0.0f ? power : 0.0f;
        if (k == target)
        {
            importance_t = importance;
            power_t = power;
        }
    }
    return PickProbability(sum_i, sum_p, importance_t, power_t);
}

// LtLeafProbabilityR (:549-572): the probability of reaching `leaf` of tree k, 0 for an empty or
// unselected tree.
template <bool kPoint>
CX_DEV float LeafProbabilityR(const LightView& L, uint32_t k, uint32_t leaf, const Receiver& r)
{
    uint32_t root;
    TreeSummary(L, k, root);
    const uint32_t parent = L.Leaves(k)[leaf].parent;
    if (root == kLtNone || parent == kLtNone)
    {
        return 0.0f;
    }
    uint32_t node = parent >> kLtParentSlotBits;
    uint32_t slot = parent & kLtParentSlotMask;
    float probability = 1.0f;
    for (uint32_t depth = 0; depth < 64; ++depth)
    {
        const LightBvhNode* address = L.Nodes(k) + node;
        probability *= SlotProbabilityR<kPoint>(address, r, slot);
        const uint32_t up = address->parent;
        if (up == kLtNone)
        {
            return probability * TreeProbabilityR<kPoint>(L, r, k);
        }
        slot = address->parent_slot;
        node = up;
    }
    return 0.0f;
}

CX_DEV float LeafProbability(const LightView& L, uint32_t k, uint32_t leaf, V3 x, V3 n)
{
    return LeafProbabilityR<true>(L, k, leaf, PointReceiver(x, n));
}

// A brick light record (BrickLightRecordHeader + entries, light_tree_layout.h:296-334).
struct Record
{
    float face[6];
    uint32_t lo[3];
    uint32_t hi[3];
    uint32_t pos;
    uint32_t entry_count;
    uint32_t counts_lo;
    uint32_t counts_hi;
    const uint32_t* entries;
};

// A group's light block in the current heap buffer, or null when the group has none (LtGroupRow,
// caelix_light_tree.hlsl:339-343: row words block_offset and block_units).
CX_DEV const uint32_t* GroupBlock(const LightView& L, uint32_t group)
{
    const uint2 row = *reinterpret_cast<const uint2*>(&L.rows[group].block_offset);
    return row.y != 0u ? L.heap + uint64_t(row.x) * 4u : nullptr;
}

// LtLoadEntryRecord (caelix_light_tree_layout.hlsl:377-401): the record of index entry `entry`.
CX_DEV Record LoadRecord(const uint32_t* block, uint32_t entry)
{
    const uint32_t* r = block + uint64_t(IndexEntry(block, entry)->record_offset) * 4u;
    const uint4 a = *reinterpret_cast<const uint4*>(r);
    const uint4 b = *reinterpret_cast<const uint4*>(r + 4);
    const uint4 c = *reinterpret_cast<const uint4*>(r + 8);
    Record rec;
    rec.face[0] = __uint_as_float(a.x);
    rec.face[1] = __uint_as_float(a.y);
    rec.face[2] = __uint_as_float(a.z);
    rec.face[3] = __uint_as_float(a.w);
    rec.face[4] = __uint_as_float(b.x);
    rec.face[5] = __uint_as_float(b.y);
    rec.lo[0] = b.z & kLtVoxelMask;
    rec.lo[1] = (b.z >> kLtVoxelBits) & kLtVoxelMask;
    rec.lo[2] = (b.z >> (2 * kLtVoxelBits)) & kLtVoxelMask;
    rec.hi[0] = b.w & kLtVoxelMask;
    rec.hi[1] = (b.w >> kLtVoxelBits) & kLtVoxelMask;
    rec.hi[2] = (b.w >> (2 * kLtVoxelBits)) & kLtVoxelMask;
    rec.pos = c.x;
    rec.entry_count = c.y;
    rec.counts_lo = c.z;
    rec.counts_hi = c.w;
    rec.entries = r + 12;
    return rec;
}

CX_DEV uint32_t FamilyCount(const Record& r, uint32_t f)
{
    const uint32_t word = f < 3 ? r.counts_lo : r.counts_hi;
    return (word >> (kLtFamilyCountBits * (f % 3))) & kLtFamilyCountMask;
}

// LtFamilyWeights (:331-356) in the group-local frame.
CX_DEV void FamilyWeights(const Record& r, V3 x, float w[6], float& sum)
{
    float c[6];
    const V3 lo = Make(float(r.lo[0]), float(r.lo[1]), float(r.lo[2]));
    const V3 hi = Make(float(r.hi[0]), float(r.hi[1]), float(r.hi[2]));
    EmitterCosines(Sub(x, hi), Sub(x, lo), c);
    sum = 0.0f;
#pragma unroll
    for (int f = 0; f < 6; ++f)
    {
        w[f] = r.face[f] * c[f];
        sum += w[f];
    }
    if (sum <= 0.0f)
    {
        sum = 0.0f;
#pragma unroll
        for (int g = 0; g < 6; ++g)
        {
            w[g] = r.face[g];
            sum += w[g];
        }
    }
}

// An entity's current frame: object-to-world rotation rows and translation (LtEntityFrame,
// caelix_light_tree_layout.hlsl:412-425).
struct Frame
{
    float r[3][3];
    V3 t;
};

// LtGroupFrame (caelix_light_tree.hlsl:346-353) + LtGroupOrigin (caelix_light_tree_layout.hlsl:428-433):
// the group's entity-local voxel origin, and its entity's frame.
CX_DEV V3 GroupFrame(const LightView& L, uint32_t group, Frame& frame)
{
    const int4 w0 = *reinterpret_cast<const int4*>(L.group_infos[group].coord);
    const int32_t edge = int32_t(8u << L.group_shift);
    const EntityInfo* e = L.entity_infos + uint32_t(w0.w);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const float4 row = *reinterpret_cast<const float4*>(e->cur_rows[i]);
        frame.r[i][0] = row.x;
        frame.r[i][1] = row.y;
        frame.r[i][2] = row.z;
        Set(frame.t, i, row.w);
    }
    return Make(float(w0.x * edge), float(w0.y * edge), float(w0.z * edge));
}

// R^T (x - t) - origin: a world point in the group's local voxel frame.
CX_DEV V3 ToGroup(const Frame& f, V3 x, V3 origin)
{
    const V3 d = Sub(x, f.t);
    const V3 local = Make(d.x * f.r[0][0] + d.y * f.r[1][0] + d.z * f.r[2][0],
                          d.x * f.r[0][1] + d.y * f.r[1][1] + d.z * f.r[2][1],
                          d.x * f.r[0][2] + d.y * f.r[1][2] + d.z * f.r[2][2]);
    return Sub(local, origin);
}

// R v
CX_DEV V3 Rotate(const Frame& f, V3 v)
{
    return Make(f.r[0][0] * v.x + f.r[0][1] * v.y + f.r[0][2] * v.z,
                f.r[1][0] * v.x + f.r[1][1] * v.y + f.r[1][2] * v.z,
                f.r[2][0] * v.x + f.r[2][1] * v.y + f.r[2][2] * v.z);
}

// LtBrickLeaf (caelix_light_tree.hlsl:594-607): the brick's leaf word (tree << 31 | leaf) when that leaf
// points back at (group, entry), else kLtNone.
CX_DEV uint32_t BrickLeaf(const LightView& L, const uint32_t* block, uint32_t group, uint32_t entry)
{
    const uint32_t word = IndexEntry(block, entry)->bvh_leaf;
    if (word == kLtNone)
    {
        return kLtNone;
    }
    const LightBvhLeaf& leaf = L.Leaves(word >> kLtBvhLeafTreeShift)[word & kLtBvhLeafIndexMask];
    const uint2 owner = *reinterpret_cast<const uint2*>(&leaf.group);
    return owner.x == group && owner.y == entry ? word : kLtNone;
}

struct LightPoint
{
    V3 position;
    V3 normal;
    uint32_t block;
    uint32_t voxel;   // the voxel in its brick (x + 8y + 64z)
    uint32_t face;
    float pdf;
    uint32_t group;   // the brick: group, index entry
    uint32_t entry;
    uint32_t record;  // the brick's record in the brick store (the index entry's dense_id)
    uint32_t voxel_in_group;  // group-local voxel, x | y << 10 | z << 20
};

CX_DEV V3 FaceDirection(uint32_t f)
{
    const float sign = (f & 1) != 0 ? -1.0f : 1.0f;
    const uint32_t axis = f >> 1;
    return Make(axis == 0 ? sign : 0.0f, axis == 1 ? sign : 0.0f, axis == 2 ? sign : 0.0f);
}

// LightTreeSample (:749-767) = LightTreeSampleBrick (:647-670) + LightBrickSamplePoint (:672-747).
CX_DEV bool LightTreeSample(const LightView& L, V3 x, V3 n, uint32_t& rng, LightPoint& p)
{
    p.position = Make(0.0f, 0.0f, 0.0f);
    p.normal = Make(0.0f, 0.0f, 0.0f);
    p.block = 0;
    p.voxel = 0;
    p.face = 0;
    p.pdf = 0.0f;
    p.group = kLtNone;
    p.entry = kLtNone;
    p.record = kLtNone;
    p.voxel_in_group = 0;
    float pdf_brick = 1.0f;
    uint32_t tree;
    const uint32_t leaf = ChooseLeaf(L, x, n, rng, pdf_brick, tree);
    if (leaf == kLtNone)
    {
        return false;
    }
    const uint2 owner = *reinterpret_cast<const uint2*>(&L.Leaves(tree)[leaf].group);
    p.group = owner.x;
    p.entry = owner.y;
    const uint32_t* block = GroupBlock(L, owner.x);
    if (block == nullptr)
    {
        return false;
    }
    const Record record = LoadRecord(block, owner.y);
    p.record = IndexEntry(block, owner.y)->dense_id;
    Frame frame;
    const V3 origin = GroupFrame(L, owner.x, frame);
    const V3 x_group = ToGroup(frame, x, origin);
    float w[6];
    float sum;
    FamilyWeights(record, x_group, w, sum);
    if (sum <= 0.0f)
    {
        return false;
    }
    const float u = LightRandomFloat(rng) * sum;
    uint32_t f = kLtNone;
    uint32_t last = 0;
    float cumulative = 0.0f;
    for (uint32_t g = 0; g < 6; ++g)
    {
        if (w[g] > 0.0f)
        {
            last = g;
            cumulative += w[g];
            if (f == kLtNone && u < cumulative)
            {
                f = g;
            }
        }
    }
    f = f == kLtNone ? last : f;
    uint32_t entry_first = 0;
    for (uint32_t h = 0; h < f; ++h)
    {
        entry_first += FamilyCount(record, h);
    }
    const uint32_t count = FamilyCount(record, f);
    if (count == 0 || w[f] <= 0.0f)
    {
        return false;
    }
    const uint32_t entry_index = min(uint32_t(LightRandomFloat(rng) * float(count)), count - 1u);
    const uint32_t entry = record.entries[entry_first + entry_index];
    const uint32_t vi = entry & kLtFaceEntryVoxelMask;
    const uint32_t mask = (1u << L.group_shift) - 1u;
    const uint32_t vx = (record.pos & mask) * 8 + (vi & 7);
    const uint32_t vy = ((record.pos >> L.group_shift) & mask) * 8 + ((vi >> 3) & 7);
    const uint32_t vz = (record.pos >> (2 * L.group_shift)) * 8 + (vi >> 6);
    V3 local = Add(origin, Make(float(vx), float(vy), float(vz)));
    const float a = LightRandomFloat(rng);
    const float b = LightRandomFloat(rng);
    const int axis = int(f >> 1);
    Set(local, axis, Get(local, axis) + ((f & 1) != 0 ? 0.0f : 1.0f));
    Set(local, (axis + 1) % 3, Get(local, (axis + 1) % 3) + a);
    Set(local, (axis + 2) % 3, Get(local, (axis + 2) % 3) + b);
    p.position = Add(Rotate(frame, local), frame.t);
    p.normal = Rotate(frame, FaceDirection(f));
    p.block = entry >> kLtFaceEntryBlockShift;
    p.voxel = vi;
    p.face = f;
    p.voxel_in_group = vx | (vy << kLtVoxelBits) | (vz << (2 * kLtVoxelBits));
    p.pdf = Div(Div(w[f], sum), float(count)) * pdf_brick;
    return true;
}

// A brick of the tree (LightBrickRef): its group, index entry and the row's generation when it was
// named; a reference of an older generation is stale.
struct BrickRef
{
    uint32_t group;
    uint32_t entry;
    uint32_t generation;
};

// LightTreeFindLightFace (:853-932): whether face `face` of the group-local voxel `v` of `group`, with the
// raw block value `block_value`, is a light face of the tree (a light material code, a brick with an
// index entry, the voxel in that face family of its record). Then `brick` names the brick, `leaf` is its
// leaf word (kLtNone: no draw reaches it) and in_brick = P(face | brick, x) / face count.
CX_DEV bool FindLightFace(const LightView& L, V3 x, uint32_t group, const uint32_t v[3], uint32_t face,
                          uint32_t block_value, BrickRef& brick, uint32_t& leaf, float& in_brick)
{
    brick = BrickRef{group, kLtNone, 0u};
    leaf = kLtNone;
    in_brick = 0.0f;
    // BlockMaterialCode (caelix_record.hlsl:209): an opaque value as it is, a transparent one without its
    // face bits
    const uint32_t raw = block_value & 0xFFFFu;
    const bool packed = L.encoding == kEncodingPacked;
    const uint32_t code = (raw & (packed ? kPackedOpaqueMask : kBlockIdOpaqueMask)) != 0u
                              ? raw
                              : (raw & (packed ? kPackedIdMask : kBlockIdIdMask));
    if (((L.light_codes[code >> 5] >> (code & 31u)) & 1u) == 0u || face > 5u)
    {
        return false;
    }
    const uint32_t* block = GroupBlock(L, group);
    if (block == nullptr)
    {
        return false;
    }
    const uint32_t key = Morton3(v[0] >> 3, v[1] >> 3, v[2] >> 3);
    const uint32_t entry = FindIndexEntry(block, block[0], key, L.group_shift);
    if (entry == kLtNone)
    {
        return false;
    }
    brick.entry = entry;
    brick.generation = L.rows[group].generation;
    const Record rec = LoadRecord(block, entry);
    Frame frame;
    const V3 origin = GroupFrame(L, group, frame);
    float w[6];
    float sum;
    FamilyWeights(rec, ToGroup(frame, x, origin), w, sum);
    const uint32_t count = FamilyCount(rec, face);
    if (sum <= 0.0f || w[face] <= 0.0f || count == 0u)
    {
        return false;
    }
    uint32_t first = 0;
    for (uint32_t g = 0; g < face; ++g)
    {
        first += FamilyCount(rec, g);
    }
    const uint32_t vi = (v[0] & 7u) | ((v[1] & 7u) << 3) | ((v[2] & 7u) << 6);
    uint32_t a = 0;
    uint32_t b = count;
    bool hit = false;
    while (a < b)
    {
        const uint32_t mid = (a + b) >> 1;
        const uint32_t ev = rec.entries[first + mid] & kLtFaceEntryVoxelMask;
        if (ev == vi)
        {
            hit = true;
            break;
        }
        if (ev < vi)
        {
            a = mid + 1;
        }
        else
        {
            b = mid;
        }
    }
    if (!hit)
    {
        return false;
    }
    in_brick = Div(Div(w[face], sum), float(count));
    leaf = BrickLeaf(L, block, group, entry);
    return true;
}

// LightTreePointPdf (:936-949): the area pdf with which LightTreeSample draws the point on face `face` of
// the group-local voxel `v` of `group`, for the receiver (x, n); 0 when it is not a light face the tree
// can draw. `found` = a light face of the tree whose brick has a leaf (LampHitWeight's condition,
// caelix_path_trace.hlsl:438-443).
CX_DEV float LightPointPdf(const LightView& L, V3 x, V3 n, uint32_t group, const uint32_t v[3], uint32_t face,
                           uint32_t block_value, bool& found)
{
    BrickRef brick;
    uint32_t leaf;
    float in_brick;
    found = FindLightFace(L, x, group, v, face, block_value, brick, leaf, in_brick) && leaf != kLtNone;
    if (!found)
    {
        return 0.0f;  // no draw reaches it: LampHitWeight gives the full emission
    }
    return LeafProbability(L, leaf >> kLtBvhLeafTreeShift, leaf & kLtBvhLeafIndexMask, x, n) * in_brick;
}

// LightBrickRefValid (:640-646): the row still has a block of the reference's generation and the entry
// lies in it.
CX_DEV bool BrickRefValid(const LightView& L, const BrickRef& b)
{
    const uint4 row = *reinterpret_cast<const uint4*>(&L.rows[b.group].block_offset);
    return row.y != 0u && row.z == b.generation && b.entry < L.heap[uint64_t(row.x) * 4];
}

// LtBrickProbabilityR (:575-592) behind LightTreeBrickPdf / LightTreeBrickPdfBox (:769-777, :822-831): the
// probability that the brick draw takes `brick`, 0 when the reference is stale or the brick has no leaf.
template <bool kPoint>
CX_DEV float BrickPdfR(const LightView& L, const Receiver& r, const BrickRef& brick)
{
    if (!BrickRefValid(L, brick))
    {
        return 0.0f;
    }
    const uint32_t* block = GroupBlock(L, brick.group);
    const uint32_t leaf = BrickLeaf(L, block, brick.group, brick.entry);
    if (leaf == kLtNone)
    {
        return 0.0f;
    }
    return LeafProbabilityR<kPoint>(L, leaf >> kLtBvhLeafTreeShift, leaf & kLtBvhLeafIndexMask, r);
}

// LightTreeSampleBrickBox (:787-818): a brick drawn for every receiver point of the box with the normal n:
// pdf_brick includes the tree choice, leaf = tree << 31 | leaf index, importance = the box importance of
// the leaf's summary (0 when the draw fell back to the power).
CX_DEV bool SampleBrickBox(const LightView& L, V3 box_lo, V3 box_hi, V3 n, uint32_t& rng, BrickRef& brick, float& pdf_brick,
                           uint32_t& leaf, float& importance)
{
    brick = BrickRef{kLtNone, kLtNone, 0u};
    pdf_brick = 1.0f;
    leaf = kLtNone;
    importance = 0.0f;
    const Receiver r{box_lo, box_hi, n};
    uint32_t root;
    const uint32_t k = ChooseTreeR<false>(L, r, rng, pdf_brick, root);
    if (k == kLtNone)
    {
        return false;
    }
    const uint32_t index = ChooseLeafR<false>(L, k, root, r, rng, pdf_brick, importance);
    if (index == kLtNone)
    {
        importance = 0.0f;
        return false;
    }
    const uint2 owner = *reinterpret_cast<const uint2*>(&L.Leaves(k)[index].group);
    brick = BrickRef{owner.x, owner.y, L.rows[owner.x].generation};
    leaf = (k << kLtBvhLeafTreeShift) | index;
    return true;
}

// LightTreeLeafImportanceBox (:837-850): the box importance of a valid brick's leaf summary (0 when the
// brick has no leaf of its own).
CX_DEV float LeafImportanceBox(const LightView& L, const BrickRef& brick, V3 box_lo, V3 box_hi, V3 n)
{
    const uint32_t* block = GroupBlock(L, brick.group);
    const uint32_t leaf = block != nullptr ? BrickLeaf(L, block, brick.group, brick.entry) : kLtNone;
    if (leaf == kLtNone)
    {
        return 0.0f;
    }
    const LightBvhLeaf& l = L.Leaves(leaf >> kLtBvhLeafTreeShift)[leaf & kLtBvhLeafIndexMask];
    const V3 lo = Make(l.summary.lo[0], l.summary.lo[1], l.summary.lo[2]);
    const V3 hi = Make(l.summary.hi[0], l.summary.hi[1], l.summary.hi[2]);
    return ImportanceR<false>(Receiver{box_lo, box_hi, n}, lo, hi, l.summary.face);
}

// The light face pdf at a hit of the world traversal: brick `record` of the brick store and the world
// voxel `voxel`. The traversal's bricks belong to entities with the identity frame, so the group-local
// voxel is the world voxel less the group's origin.
CX_DEV float LightFacePdf(const SceneView& s, const LightView& L, V3 x, V3 n, uint32_t record, const int32_t voxel[3],
                          uint32_t face, uint32_t block, bool& found)
{
    const uint32_t group = s.group_of_brick[record];
    const int4 w0 = *reinterpret_cast<const int4*>(L.group_infos[group].coord);
    const int32_t edge = int32_t(8u << L.group_shift);
    const uint32_t v[3] = {uint32_t(voxel[0] - w0.x * edge), uint32_t(voxel[1] - w0.y * edge),
                           uint32_t(voxel[2] - w0.z * edge)};
    return LightPointPdf(L, x, n, group, v, face, block, found);
}

// LightNeeMisWeight (caelix_light_nee.hlsl:20-34).
CX_DEV float MisWeight(float p_area, V3 n_r, V3 w, float d, float cos_l, bool nee_side)
{
    const float p_sa = cos_l > 0.0f ? Div(p_area * (d * d), cos_l) : 0.0f;
    const float p_bsdf = fmaxf(0.0f, Dot(n_r, w)) * (1.0f / 3.1415926535f);
    if (!(cos_l > 0.0f) || !(p_sa <= 3.0e38f))
    {
        return nee_side ? 1.0f : 0.0f;
    }
    const float sum = p_sa + p_bsdf;
    if (!(sum > 0.0f))
    {
        return nee_side ? 0.0f : 1.0f;
    }
    return Div(nee_side ? p_sa : p_bsdf, sum);
}

}  // namespace cx
