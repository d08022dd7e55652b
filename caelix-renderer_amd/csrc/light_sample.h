// The light tree's device API for gfx950 kernels (shaders/light/caelix_light_tree.hlsl): importance
// of a summary for a point receiver, the weighted reservoir pick, the walk down the 4-wide nodes to a
// leaf, the in-brick draw of a light face, and the pdf walk up the parent links. Layouts are the
// reference's (LightTreeHeader, LightBvhNode, LightBvhLeaf, brick light records).
//
// The leaf -> brick light record indirection of this build: LightBvhLeaf::entry is the brick's light
// ordinal, whose record starts at heap unit record_offset[ordinal]; a brick record's ordinal is
// ordinal_of_brick[record]. (The reference goes leaf -> (group, index entry) -> group row -> block ->
// record, caelix_light_tree.hlsl:600-680; the records themselves are the same.) Entities are the
// identity frame.
#pragma once
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

// LtImportance for a point receiver x with normal n (:118-155, is_point).
CX_DEV float Importance(V3 x, V3 n, V3 lo, V3 hi, const float face[6])
{
    float c[6];
    EmitterCosines(Sub(x, hi), Sub(x, lo), c);
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
    float cr = 1.0f;
    if (n.x != 0.0f || n.y != 0.0f || n.z != 0.0f)
    {
        const float m = Dot(n, Sub(centre, x)) + Dot(Make(fabsf(n.x), fabsf(n.y), fabsf(n.z)), half_size);
        if (m <= 0.0f)
        {
            return 0.0f;
        }
        const V3 gap = Make(fmaxf(fmaxf(lo.x - x.x, x.x - hi.x), 0.0f), fmaxf(fmaxf(lo.y - x.y, x.y - hi.y), 0.0f),
                            fmaxf(fmaxf(lo.z - x.z, x.z - hi.z), 0.0f));
        const float dmin = Sqrt(Dot(gap, gap));
        cr = dmin > 0.0f ? fminf(1.0f, Div(m, dmin)) : 1.0f;
    }
    const V3 to_centre = Sub(x, centre);
    const float extent2 = Dot(half_size, half_size);
    const float d2 = fmaxf(Dot(to_centre, to_centre), fmaxf(1.0f, extent2));
    return Div(e * cr, d2);
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

// LtTreeSummary (:289-304): the root summary of tree k, power 0 for an empty tree.
CX_DEV Summ TreeSummary(const LightView& L, uint32_t k, uint32_t& root)
{
    root = L.headers[k].root;
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
    return HeaderSummary(&L.headers[k]);
}

// LtChooseTreeR (:417-437) + LtChooseLeafR (:466-512), point receiver.
CX_DEV uint32_t ChooseLeaf(const LightView& L, V3 x, V3 n, uint32_t& rng, float& pdf, uint32_t& tree)
{
    Pick pick = PickBegin();
    uint32_t roots[2];
    for (uint32_t k = 0; k < 2; ++k)
    {
        const Summ s = TreeSummary(L, k, roots[k]);
        PickAdd(pick, Importance(x, n, s.lo, s.hi, s.face), Power6(s.face), k, LightRandomFloat(rng));
    }
    float probability;
    const uint32_t k = PickEnd(pick, probability);
    tree = k;
    if (k == kLtNone)
    {
        return kLtNone;
    }
    pdf *= probability;
    uint32_t node = roots[k];
    for (uint32_t depth = 0; depth < 64; ++depth)
    {
        const LightBvhNode* address = L.nodes[k] + node;
        uint32_t children[4];
        *reinterpret_cast<uint4*>(children) = *reinterpret_cast<const uint4*>(address->child);
        uint32_t chosen;
        if (ChildCount(children) == 1)
        {
            chosen = children[0];
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
                p.sum_i += 0.0f;
                PickAdd(p, Importance(x, n, slot.lo, slot.hi, slot.face), Power6(slot.face), s, LightRandomFloat(rng));
            }
            float prob;
            const uint32_t s = PickEnd(p, prob);
            if (s == kLtNone)
            {
                return kLtNone;
            }
            pdf *= prob;
            chosen = children[s];
        }
        if ((chosen & kLtChildLeafBit) != 0)
        {
            return chosen & ~kLtChildLeafBit;
        }
        node = chosen;
    }
    return kLtNone;
}

// LtSlotProbabilityR (:515-546)
CX_DEV float SlotProbability(const LightBvhNode* address, V3 x, V3 n, uint32_t slot)
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
        const float importance = Importance(x, n, summary.lo, summary.hi, summary.face);
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
CX_DEV float TreeProbability(const LightView& L, V3 x, V3 n, uint32_t target)
{
    float sum_i = 0.0f;
    float sum_p = 0.0f;
    float importance_t = 0.0f;
    float power_t = 0.0f;
    for (uint32_t k = 0; k < 2; ++k)
    {
        uint32_t root;
        const Summ s = TreeSummary(L, k, root);
        const float importance = Importance(x, n, s.lo, s.hi, s.face);
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

// LtLeafProbabilityR (:549-572): the probability of reaching `leaf` of tree k.
CX_DEV float LeafProbability(const LightView& L, uint32_t k, uint32_t leaf, V3 x, V3 n)
{
    if (L.headers[k].root == kLtNone)
    {
        return 0.0f;
    }
    const uint32_t parent = L.leaves[k][leaf].parent;
    if (parent == kLtNone)
    {
        return 0.0f;
    }
    uint32_t node = parent >> kLtParentSlotBits;
    uint32_t slot = parent & kLtParentSlotMask;
    float probability = 1.0f;
    for (uint32_t depth = 0; depth < 64; ++depth)
    {
        const LightBvhNode* address = L.nodes[k] + node;
        probability *= SlotProbability(address, x, n, slot);
        const uint32_t up = address->parent;
        if (up == kLtNone)
        {
            return probability * TreeProbability(L, x, n, k);
        }
        slot = address->parent_slot;
        node = up;
    }
    return 0.0f;
}

// A brick light record (the reference's BrickLightRecordHeader + entries).
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

CX_DEV Record LoadRecord(const LightView& L, uint32_t ordinal)
{
    const uint32_t* r = L.heap + uint64_t(L.record_offset[ordinal]) * 4;
    Record rec;
#pragma unroll
    for (int f = 0; f < 6; ++f)
    {
        rec.face[f] = __uint_as_float(r[f]);
    }
    const uint32_t box_lo = r[6];
    const uint32_t box_hi = r[7];
    rec.lo[0] = box_lo & kLtVoxelMask;
    rec.lo[1] = (box_lo >> kLtVoxelBits) & kLtVoxelMask;
    rec.lo[2] = (box_lo >> (2 * kLtVoxelBits)) & kLtVoxelMask;
    rec.hi[0] = box_hi & kLtVoxelMask;
    rec.hi[1] = (box_hi >> kLtVoxelBits) & kLtVoxelMask;
    rec.hi[2] = (box_hi >> (2 * kLtVoxelBits)) & kLtVoxelMask;
    rec.pos = r[8];
    rec.entry_count = r[9];
    rec.counts_lo = r[10];
    rec.counts_hi = r[11];
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

CX_DEV V3 GroupOrigin(const SceneView& s, const LightView& L, int32_t bx, int32_t by, int32_t bz)
{
    const int32_t g = int32_t(L.group_shift);
    return Make(float((bx >> g) << g) * 8.0f, float((by >> g) << g) * 8.0f, float((bz >> g) << g) * 8.0f);
}

struct LightPoint
{
    V3 position;
    V3 normal;
    uint32_t block;
    uint32_t face;
    float pdf;
    uint32_t record;  // the brick record of the light
};

CX_DEV V3 FaceDirection(uint32_t f)
{
    const float sign = (f & 1) != 0 ? -1.0f : 1.0f;
    const uint32_t axis = f >> 1;
    return Make(axis == 0 ? sign : 0.0f, axis == 1 ? sign : 0.0f, axis == 2 ? sign : 0.0f);
}

// LightTreeSample (:749-767) = LightTreeSampleBrick + LightBrickSamplePoint (:647-747).
CX_DEV bool LightTreeSample(const SceneView& s, const LightView& L, V3 x, V3 n, uint32_t& rng, LightPoint& p)
{
    p.position = Make(0.0f, 0.0f, 0.0f);
    p.normal = Make(0.0f, 0.0f, 0.0f);
    p.block = 0;
    p.face = 0;
    p.pdf = 0.0f;
    p.record = kLtNone;
    float pdf_brick = 1.0f;
    uint32_t tree;
    const uint32_t leaf = ChooseLeaf(L, x, n, rng, pdf_brick, tree);
    if (leaf == kLtNone)
    {
        return false;
    }
    const LightBvhLeaf& lf = L.leaves[tree][leaf];
    const uint32_t brick_record = lf.group;
    const uint32_t ordinal = lf.entry;
    const Record record = LoadRecord(L, ordinal);
    const int32_t* key = BrickXyz(s, brick_record);
    const V3 origin = GroupOrigin(s, L, key[0], key[1], key[2]);
    const V3 x_group = Sub(x, origin);
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
    const uint32_t bx = record.pos & mask;
    const uint32_t by = (record.pos >> L.group_shift) & mask;
    const uint32_t bz = record.pos >> (2 * L.group_shift);
    V3 local = Add(origin, Make(float(bx * 8 + (vi & 7)), float(by * 8 + ((vi >> 3) & 7)), float(bz * 8 + (vi >> 6))));
    const float a = LightRandomFloat(rng);
    const float b = LightRandomFloat(rng);
    const int axis = int(f >> 1);
    Set(local, axis, Get(local, axis) + ((f & 1) != 0 ? 0.0f : 1.0f));
    Set(local, (axis + 1) % 3, Get(local, (axis + 1) % 3) + a);
    Set(local, (axis + 2) % 3, Get(local, (axis + 2) % 3) + b);
    p.position = local;
    p.normal = FaceDirection(f);
    p.block = entry >> kLtFaceEntryBlockShift;
    p.face = f;
    p.record = brick_record;
    p.pdf = Div(Div(w[f], sum), float(count)) * pdf_brick;
    return true;
}

// LightTreeFindLightFace (:853-932) + LightTreeBrickPdf: the area pdf with which LightTreeSample
// draws the point on face `face` of voxel `voxel` (world voxel coordinates) of brick `record`, for the
// receiver (x, n); 0 when it is not a light face the tree can draw. `found` = a light face.
CX_DEV float LightFacePdf(const SceneView& s, const LightView& L, V3 x, V3 n, uint32_t record, const int32_t voxel[3],
                          uint32_t face, uint32_t block, bool& found)
{
    found = false;
    const uint32_t code = block & 0xFFFFu;
    if (((s.light_codes[code >> 5] >> (code & 31u)) & 1u) == 0u || face > 5u)
    {
        return 0.0f;
    }
    const uint32_t ordinal = L.ordinal_of_brick[record];
    if (ordinal == kLtNone)
    {
        return 0.0f;
    }
    const Record rec = LoadRecord(L, ordinal);
    const int32_t* key = BrickXyz(s, record);
    const V3 origin = GroupOrigin(s, L, key[0], key[1], key[2]);
    float w[6];
    float sum;
    FamilyWeights(rec, Sub(x, origin), w, sum);
    const uint32_t count = FamilyCount(rec, face);
    if (sum <= 0.0f || w[face] <= 0.0f || count == 0u)
    {
        return 0.0f;
    }
    uint32_t first = 0;
    for (uint32_t g = 0; g < face; ++g)
    {
        first += FamilyCount(rec, g);
    }
    const uint32_t vi = uint32_t((voxel[0] & 7) | ((voxel[1] & 7) << 3) | ((voxel[2] & 7) << 6));
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
        return 0.0f;
    }
    found = true;
    const float in_brick = Div(Div(w[face], sum), float(count));
    // The brick's leaf: tree 0 (the static tree), leaf = the light ordinal - leaf_base when it is
    // below the count.
    const uint32_t leaf = ordinal - L.leaf_base;
    if (leaf >= L.headers[0].leaf_count)
    {
        found = false;  // no leaf: LampHitWeight gives the full emission
        return 0.0f;
    }
    return LeafProbability(L, 0, leaf, x, n) * in_brick;
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
