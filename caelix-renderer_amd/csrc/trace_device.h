// Device code of the trace for gfx950: the brick DDA (the reference's
// shaders/render/caelix_brick_trace.hlsl:203-451), the world traversal that finds bricks through the
// hash table, the light tree's sample and pdf walks (shaders/light/caelix_light_tree.hlsl), the MIS
// weight (shaders/render/caelix_light_nee.hlsl), the random streams (shaders/render/caelix_random.hlsl)
// and the opaque-diffuse subset of the path tracer (shaders/render/caelix_path_trace.hlsl:276-452,
// 737-978).
//
// Floating point: IEEE single operations in source order (no contraction; IEEE division and a
// correctly rounded square root),
// so the CPU restatement reproduces every decision. sin/cos are the Cephes single-precision
// polynomials (S. Moshier, Cephes Math Library, sinf.c / cosf.c) with Cody-Waite reduction, not the
// device library's.
#pragma once
#include "caelix_amd/caelix_amd.h"
#include "device_common.h"
#include "hash_device.h"

namespace cx {

// ---- scene -----------------------------------------------------------------------------------------
struct SceneView
{
    const int32_t* keys;        // [bricks][3] brick coordinates
    const uint2* info;          // [bricks] record words 0 and 1
    const uint64_t* occupancy;  // [bricks][8] micro occupancy of the coarse cells
    const uint16_t* payload;    // [bricks][512] block values, voxel x + 8y + 64z
    const HashKeySlot* slots;
    uint32_t slot_count;
    uint32_t pow2;
    const uint32_t* hash_info;  // [0] probe limit
    const uint32_t* brick_bits;  // grid_dim^3 bits: brick cell resident
    const uint32_t* super_bits;  // (grid_dim / 8)^3 bits: any resident brick in the 8^3 super cell
    const int32_t* box;          // per scene: super-cell box of the resident bricks (lo, hi), or null
    int32_t grid_min[3];         // brick coordinate of grid cell 0
    uint32_t grid_dim;           // bricks per axis, a multiple of 8
    const float4* materials;     // 2 per 16-bit block value: (albedo, smoothness), (emission, metallic)
    const float4* materials_ext; // null, or 1 per block value: (ior, extinction, 0, 0)
    const uint32_t* light_codes; // 65536 bits: the block value is a light material
    float emission_scale;
    uint32_t encoding;
    const float* brick_scale;    // [bricks] emission scale per brick, or null (1)
    uint32_t key_words;          // 3: keys (x, y, z); 4: (scene, x, y, z), a batch of chunk scenes
    uint64_t scene_cells;        // brick-bitmap bits per scene (grid_dim^3)
    uint64_t scene_supers;       // super-bitmap bits per scene ((grid_dim / 8)^3)
    const uint32_t* group_of_brick;  // [bricks] group index (the light tree's group rows and infos)
    const CxSubsurface* subsurface;  // [65536] subsurface entries, or null (no SSS material)
};

// The (x, y, z) brick coordinates of a record.
CX_DEV const int32_t* BrickXyz(const SceneView& s, uint32_t record)
{
    return s.keys + uint64_t(record) * s.key_words + (s.key_words - 3);
}

CX_DEV float BrickScale(const SceneView& s, uint32_t record)
{
    return s.brick_scale != nullptr ? s.brick_scale[record] : 1.0f;
}

// The light tree as the trace kernels read it, opened from its descriptor (LightTreeOpen /
// LightTreeView, caelix_light_tree.hlsl:239-287): the tree headers in the control buffer, the nodes and
// leaves of both trees, the current heap buffer, the group rows, group and entity infos.
struct LightView
{
    const LightTreeHeader* headers;  // [2]
    const LightBvhNode* nodes[2];
    const LightBvhLeaf* leaves[2];
    const uint32_t* heap;            // the current buffer: group light blocks, 16-byte units
    const GroupLightRow* rows;
    const GroupInfo* group_infos;
    const EntityInfo* entity_infos;
    const uint32_t* light_codes;     // 65536 bits: the block value is a light material
    uint32_t group_shift;            // a group is (1 << group_shift)^3 bricks
    uint32_t encoding;               // the block value encoding (kEncoding*)
    uint32_t mis;
    uint32_t trees;                  // LightTreeSelect: the trees draws and pdfs use (bit k = tree k)

    // Tree k's arrays, picked by a select: an index into the arrays themselves with a per-lane k puts
    // the whole view in per-lane scratch memory.
    CX_DEV const LightBvhNode* Nodes(uint32_t k) const { return k == 0u ? nodes[0] : nodes[1]; }
    CX_DEV const LightBvhLeaf* Leaves(uint32_t k) const { return k == 0u ? leaves[0] : leaves[1]; }
    CX_DEV const LightTreeHeader* Header(uint32_t k) const { return headers + k; }
};

// What a kernel receives of a light tree: its descriptor's device address, the MIS switch and the
// selected trees (kLtSelect*: 1 static, 2 dynamic, 3 both; 0 is taken as both).
struct LightRef
{
    const LightTreeDescriptor* descriptor;
    uint32_t mis;
    uint32_t select;
};

CX_DEV LightView OpenLights(const LightRef& r)
{
    const LightTreeDescriptor* d = r.descriptor;
    LightView l;
    l.headers = reinterpret_cast<const LightTreeHeader*>(d->tree_headers);
    l.nodes[0] = reinterpret_cast<const LightBvhNode*>(d->nodes[0]);
    l.nodes[1] = reinterpret_cast<const LightBvhNode*>(d->nodes[1]);
    l.leaves[0] = reinterpret_cast<const LightBvhLeaf*>(d->leaves[0]);
    l.leaves[1] = reinterpret_cast<const LightBvhLeaf*>(d->leaves[1]);
    const uint32_t current = reinterpret_cast<const LightControl*>(d->control)->current & 1u;
    l.heap = reinterpret_cast<const uint32_t*>(current != 0 ? d->heap[1] : d->heap[0]);
    l.rows = reinterpret_cast<const GroupLightRow*>(d->group_rows);
    l.group_infos = reinterpret_cast<const GroupInfo*>(d->group_infos);
    l.entity_infos = reinterpret_cast<const EntityInfo*>(d->entity_infos);
    l.light_codes = reinterpret_cast<const uint32_t*>(d->light_codes);
    l.group_shift = d->group_shift;
    l.encoding = d->encoding;
    l.mis = r.mis;
    l.trees = r.select != 0u ? (r.select & 3u) : 3u;
    return l;
}

struct CameraView
{
    float position[3];
    float forward[3];
    float right[3];
    float up[3];
    float tan_half_fov_y;
    float aspect;
    uint32_t width;
    uint32_t height;
};

constexpr float kPi = 3.1415926535f;
constexpr float kTwoPi = 6.283185307f;
constexpr float kRayTMax = 1e30f;
constexpr float kPushOff = 0.002f;
constexpr float kNearField = 8.0f;
constexpr uint32_t kBrickMaxSteps = 22;

CX_DEV float Div(float a, float b) { return __fdiv_rn(a, b); }

// Correctly rounded single-precision square root. The hardware v_sqrt_f32 is within 1 ulp (and
// __fsqrt_rn lowers to it alone on gfx950); the exact answer is s or one of its neighbours, and the
// signs of the fma residuals x - s_down * s and x - s_up * s say which: if x <= s_down * s the true
// root lies below the midpoint of s_down and s, if x > s_up * s above the midpoint of s and s_up.
// Inputs below 2^-96 are scaled by 2^32 first (the root by 2^-16 after) so that the residuals stay
// normal. Zero, infinity and NaN pass through (their residuals are NaN or not of the tested sign).
CX_DEV float Sqrt(float a)
{
    const bool tiny = a < 0x1p-96f;
    const float x = tiny ? a * 0x1p+32f : a;
    float s = __builtin_amdgcn_sqrtf(x);
    const float down = __uint_as_float(__float_as_uint(s) - 1u);
    const float up = __uint_as_float(__float_as_uint(s) + 1u);
    const float r_down = __builtin_fmaf(-down, s, x);
    const float r_up = __builtin_fmaf(-up, s, x);
    s = r_down <= 0.0f ? down : s;
    s = r_up > 0.0f ? up : s;
    return tiny ? s * 0x1p-16f : s;
}

struct V3
{
    float x, y, z;
};
CX_DEV V3 Make(float x, float y, float z) { return V3{x, y, z}; }
CX_DEV V3 Add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
CX_DEV V3 Sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
CX_DEV V3 Mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
CX_DEV V3 MulV(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
CX_DEV float Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CX_DEV float Get(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
CX_DEV void Set(V3& v, int i, float s)
{
    if (i == 0)
    {
        v.x = s;
    }
    else if (i == 1)
    {
        v.y = s;
    }
    else
    {
        v.z = s;
    }
}
// v * (1 / sqrt(max(FLT_MIN, dot(v, v)))) (SafeNormalize, caelix_shading.hlsl:110-114, with an IEEE
// reciprocal square root).
// a ? x : y per component (a select of whole V3 values can become a per-lane scratch copy)
CX_DEV V3 Select(bool a, V3 x, V3 y)
{
    return Make(a ? x.x : y.x, a ? x.y : y.y, a ? x.z : y.z);
}
CX_DEV V3 SafeNormalize(V3 v)
{
    const float dp3 = fmaxf(1.175494351e-38f, Dot(v, v));
    return Mul(v, Div(1.0f, Sqrt(dp3)));
}

// ---- Cephes sinf / cosf ------------------------------------------------------------------------
CX_DEV float PolySin(float x, float z)
{
    return ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
}
CX_DEV float PolyCos(float z)
{
    return ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z - 0.5f * z +
           1.0f;
}
// sin and cos of a >= 0 (a < 2^24).
CX_DEV void SinCos(float a, float& s, float& c)
{
    int j = int(a * 1.27323954473516f);  // 4 / pi
    float y = float(j);
    if ((j & 1) != 0)
    {
        j += 1;
        y += 1.0f;
    }
    j &= 7;
    const float x = ((a - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
    const float z = x * x;
    const float ps = PolySin(x, z);
    const float pc = PolyCos(z);
    // sin
    int js = j;
    float sign_s = 1.0f;
    if (js > 3)
    {
        sign_s = -1.0f;
        js -= 4;
    }
    s = (js == 1 || js == 2) ? pc : ps;
    s = sign_s < 0.0f ? -s : s;
    // cos
    int jc = j;
    float sign_c = 1.0f;
    if (jc > 3)
    {
        jc -= 4;
        sign_c = -sign_c;
    }
    if (jc > 1)
    {
        sign_c = -sign_c;
    }
    c = (jc == 1 || jc == 2) ? ps : pc;
    c = sign_c < 0.0f ? -c : c;
}

// ---- Cephes expf -------------------------------------------------------------------------------
// exp(x) (the transmissions of fog and media): x = n ln 2 + r with n = floor(x log2 e + 0.5), r by
// the two-part ln 2, the degree-5 polynomial of Cephes expf (S. Moshier, Cephes Math Library), times
// 2^n from its exponent bits; 0 below -87, where the result would leave the normal range, and +inf
// above 88. exp(0) is exactly 1.
CX_DEV float ExpCephes(float x)
{
    if (!(x >= -87.0f))
    {
        return 0.0f;
    }
    if (x > 88.0f)
    {
        return __int_as_float(0x7F800000);
    }
    const float fx = floorf(x * 1.44269504088896341f + 0.5f);
    float r = x - fx * 0.693359375f;
    r = r - fx * -2.12194440e-4f;
    const float z = r * r;
    const float p = ((((1.9875691500e-4f * r + 1.3981999507e-3f) * r + 8.3334519073e-3f) * r + 4.1665795894e-2f) * r +
                     1.6666665459e-1f) * r + 5.0000001201e-1f;
    const float e = p * z + r + 1.0f;
    return e * __int_as_float((int(fx) + 127) << 23);
}

// ---- random streams (caelix_random.hlsl) ---------------------------------------------------------
CX_DEV uint32_t PcgHash(uint32_t value)
{
    const uint32_t state = value * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
CX_DEV uint32_t InitRandom(uint32_t px, uint32_t py, uint32_t seed)
{
    const uint32_t state = PcgHash(px ^ (py << 16) ^ (seed * 0x9E3779B9u));
    return state != 0u ? state : 0x9E3779B9u;
}
CX_DEV uint32_t RandomU32(uint32_t& state)
{
    state ^= state << 13;
    state ^= state >> 17;
    state ^= state << 5;
    return state;
}
CX_DEV float RandomFloat01(uint32_t& state)
{
    return float(RandomU32(state) >> 8) * (1.0f / 16777216.0f);
}
CX_DEV V3 RandomUnitVector(uint32_t& state)
{
    const float u = RandomFloat01(state);
    const float v = RandomFloat01(state);
    const float a = u * kTwoPi;
    const float y = v * 2.0f - 1.0f;
    const float sy = Sqrt(fmaxf(1.0f - y * y, 0.0f));
    float s, c;
    SinCos(a, s, c);
    return Make(s * sy, y, c * sy);
}
// The light tree's PCG stream (caelix_light_tree.hlsl:65-76).
CX_DEV uint32_t LightRandom(uint32_t& rng)
{
    rng = rng * 747796405u + 2891336453u;
    const uint32_t word = ((rng >> ((rng >> 28) + 4u)) ^ rng) * 277803737u;
    return (word >> 22) ^ word;
}
CX_DEV float LightRandomFloat(uint32_t& rng)
{
    return float(LightRandom(rng) >> 8) * (1.0f / 16777216.0f);
}

// ---- brick DDA (caelix_brick_trace.hlsl) --------------------------------------------------------
struct BrickHit
{
    float t;
    uint32_t attrib;  // block << 16 | cell << 6 | face flags; 0 = miss
    uint32_t steps;   // DDA steps taken
};

CX_DEV float SafeInvAxis(float d)
{
    if (fabsf(d) >= 1e-20f)
    {
        return Div(1.0f, d);
    }
    return d < 0.0f ? Div(-1.0f, 1e-20f) : Div(1.0f, 1e-20f);
}

CX_DEV uint32_t ReadBlock(const uint16_t* payload, int x, int y, int z)
{
    return payload[x + 8 * y + 64 * z];
}

CX_DEV bool IsOpaque(uint32_t block, uint32_t encoding)
{
    return (block & (encoding == kEncodingPacked ? kPackedOpaqueMask : kBlockIdOpaqueMask)) != 0;
}
CX_DEV uint32_t FaceBits(uint32_t block, uint32_t encoding)
{
    return encoding == kEncodingPacked ? ((block >> kPackedFaceShift) & 0x3Fu) : ((block >> kBlockIdFaceShift) & 0x3Fu);
}

// CaelixTraceBrickRay (:299-368). `inv` = SafeInvAxis of each component of dir (a function of the
// ray only, so callers compute it once per ray).
CX_DEV BrickHit TraceBrickRay(const uint64_t* occupancy, const uint16_t* payload, uint32_t encoding, V3 entry,
                              V3 dir, V3 inv, float entry_t, uint32_t entry_normal_coarse)
{
    const V3 boundary = Make(dir.x >= 0.0f ? 1.0f : 0.0f, dir.y >= 0.0f ? 1.0f : 0.0f, dir.z >= 0.0f ? 1.0f : 0.0f);
    const V3 t_start = MulV(Sub(boundary, entry), inv);
    const float hi = 8.0f - 0.0001f;
    const V3 origin = Make(fminf(fmaxf(entry.x, 0.0f), hi), fminf(fmaxf(entry.y, 0.0f), hi), fminf(fmaxf(entry.z, 0.0f), hi));
    int cx_ = int(floorf(origin.x));
    int cy = int(floorf(origin.y));
    int cz = int(floorf(origin.z));
    int px = cx_, py = cy, pz = cz;
    BrickHit hit{0.0f, 0u, 0u};
    for (uint32_t step = 0; step < kBrickMaxSteps; ++step)
    {
        hit.steps = step + 1;
        if (uint32_t(cx_) >= 8u || uint32_t(cy) >= 8u || uint32_t(cz) >= 8u)
        {
            break;
        }
        const uint32_t coarse_bit = uint32_t((cx_ >> 2) | ((cy >> 2) << 1) | ((cz >> 2) << 2));
        bool advance_from_jump = false;
        if ((entry_normal_coarse & (1u << coarse_bit)) == 0u)
        {
            // VRTJump: the far corner of the empty coarse cell, then one step.
            cx_ = dir.x < 0.0f ? (cx_ & ~3) : (cx_ | 3);
            cy = dir.y < 0.0f ? (cy & ~3) : (cy | 3);
            cz = dir.z < 0.0f ? (cz & ~3) : (cz | 3);
            advance_from_jump = true;
        }
        else
        {
            const uint64_t occ = occupancy[coarse_bit];
            const uint32_t micro = uint32_t((cx_ & 3) | ((cy & 3) << 2) | ((cz & 3) << 4));
            if (((occ >> micro) & 1ull) != 0ull)
            {
                const uint32_t block = ReadBlock(payload, cx_, cy, cz);
                float T = 0.0f;
                uint32_t flags;
                if (step == 0u)
                {
                    flags = entry_normal_coarse >> 26;
                }
                else
                {
                    // CaelixBrickRayNormalFlags from the previous cell.
                    const V3 side = Add(t_start, MulV(Make(float(px), float(py), float(pz)), inv));
                    T = fminf(fminf(side.x, side.y), side.z);
                    const float lim = T + 0.0001f;
                    const uint32_t x_use = side.x <= lim ? 0xFFFFFFFFu : 0u;
                    const uint32_t y_use = (side.y <= lim ? 0xFFFFFFFFu : 0u) & ~x_use;
                    const uint32_t z_use = ~(x_use | y_use);
                    const uint32_t xf = dir.x > 0.0f ? 0x02u : 0x01u;
                    const uint32_t yf = dir.y > 0.0f ? 0x08u : 0x04u;
                    const uint32_t zf = dir.z > 0.0f ? 0x20u : 0x10u;
                    flags = (xf & x_use) | (yf & y_use) | (zf & z_use);
                }
                bool terminate = IsOpaque(block, encoding);
                if (!terminate && !(step == 0u && entry_t == 0.0f))
                {
                    terminate = (FaceBits(block, encoding) & flags) != 0u;
                }
                if (terminate)
                {
                    hit.attrib = (block << 16) | (uint32_t(cx_ | (cy << 3) | (cz << 6)) << 6) | flags;
                    hit.t = entry_t + T;
                    return hit;
                }
            }
        }
        if (!advance_from_jump)
        {
            px = cx_;
            py = cy;
            pz = cz;
        }
        else
        {
            px = cx_;
            py = cy;
            pz = cz;
        }
        // VRTStep
        const V3 side = Add(t_start, MulV(Make(float(cx_), float(cy), float(cz)), inv));
        const float tmin = fminf(fminf(side.x, side.y), side.z) + 0.0001f;
        const V3 pos = Add(origin, Mul(dir, tmin));
        cx_ = int(floorf(pos.x));
        cy = int(floorf(pos.y));
        cz = int(floorf(pos.z));
    }
    hit.attrib = 0u;
    hit.t = 0.0f;
    return hit;
}

// The two reciprocals of a ray direction the brick test uses: 1 / d (the slab test,
// caelix_brick_trace.hlsl:376-451, unguarded) and SafeInvAxis (the brick DDA).
struct RayInv
{
    V3 raw;
    V3 safe;
};
CX_DEV RayInv InverseOf(V3 dir)
{
    RayInv r;
    r.raw = Make(Div(1.0f, dir.x), Div(1.0f, dir.y), Div(1.0f, dir.z));
    r.safe = Make(SafeInvAxis(dir.x), SafeInvAxis(dir.y), SafeInvAxis(dir.z));
    return r;
}

// CaelixTraceBrickPrimitiveCore (:376-451): slab test against the tight bounds, entry face, DDA.
CX_DEV BrickHit TraceBrick(const SceneView& s, uint32_t record, V3 o, V3 dir, const RayInv& ri, float t_current,
                           V3 brick_origin)
{
    const uint2 info = s.info[record];
    const uint32_t tb = info.y;
    const V3 tmin_box = Make(float(tb & 7u), float((tb >> 3) & 7u), float((tb >> 6) & 7u));
    const V3 tmax_box = Make(float((tb >> 9) & 7u) + 1.0f, float((tb >> 12) & 7u) + 1.0f, float((tb >> 15) & 7u) + 1.0f);
    const V3 aabb_min = Add(brick_origin, tmin_box);
    const V3 aabb_max = Add(brick_origin, tmax_box);
    const V3 t0 = MulV(Sub(aabb_min, o), ri.raw);
    const V3 t1 = MulV(Sub(aabb_max, o), ri.raw);
    const V3 tn = Make(fminf(t0.x, t1.x), fminf(t0.y, t1.y), fminf(t0.z, t1.z));
    const V3 tf = Make(fmaxf(t0.x, t1.x), fmaxf(t0.y, t1.y), fmaxf(t0.z, t1.z));
    const float largest_tmin = fmaxf(fmaxf(tn.x, tn.y), tn.z);
    const float smallest_tmax = fminf(fminf(tf.x, tf.y), tf.z);
    BrickHit miss{0.0f, 0u, 0u};
    if (largest_tmin > smallest_tmax || smallest_tmax < 0.0f || largest_tmin > t_current)
    {
        return miss;
    }
    const float t = fmaxf(0.0f, largest_tmin);
    const V3 entry = Sub(Add(o, Mul(dir, t)), brick_origin);
    uint32_t flags;
    if (largest_tmin == tn.x)
    {
        flags = dir.x > 0.0f ? 0x02u : 0x01u;
    }
    else if (largest_tmin == tn.y)
    {
        flags = dir.y > 0.0f ? 0x08u : 0x04u;
    }
    else
    {
        flags = dir.z > 0.0f ? 0x20u : 0x10u;
    }
    const uint32_t coarse = (info.x >> kRecordCoarseShift) & kRecordCoarseMask;
    return TraceBrickRay(s.occupancy + uint64_t(record) * 8, s.payload + uint64_t(record) * 512, s.encoding, entry,
                         dir, ri.safe, t, (flags << 26) | coarse);
}

// ---- world traversal ------------------------------------------------------------------------------
struct WorldHit
{
    float t;
    uint32_t attrib;
    uint32_t record;   // brick record of the hit, kHashNone for a miss
    int32_t brick[3];  // brick coordinate of the hit
    uint32_t cells;    // brick cells visited
    uint32_t probes;   // hash slots read
    uint32_t steps;    // brick DDA steps
};

CX_DEV bool BitSet(const uint32_t* bits, uint64_t index)
{
    return ((bits[index >> 5] >> (index & 31u)) & 1u) != 0u;
}

// The hash key of brick (x, y, z) of a scene: the 3 coordinates, or (scene, x, y, z) in a batch.
CX_DEV uint2 SceneBrickFingerprint(const SceneView& s, uint32_t scene, int32_t x, int32_t y, int32_t z)
{
    if (s.key_words == 4)
    {
        const uint32_t w[4] = {scene, uint32_t(x), uint32_t(y), uint32_t(z)};
        return Fingerprint(w, 4);
    }
    return BrickFingerprint(x, y, z);
}

// The ray [o, o + t d], t in [0, t_max], through the brick grid: a DDA over the 8^3-brick super
// cells and, inside a resident super cell, over its bricks; a resident brick is found through the
// hash table and traced with the brick DDA. The first hit is the closest (bricks are disjoint and
// visited in ray order). `any_hit`: a shadow ray (any hit with t < t_max).
// Ends the traversal where the ray leaves the box of super cells holding resident bricks, widened by
// one super cell within the grid; t_exit < t_enter when the ray misses it. No resident brick lies
// outside the box, and its planes are super-cell planes computed as the super-cell exits are
// ((k * 64 + grid origin - o) * inv), so every resident super cell is reached exactly as before.
CX_DEV void ClipToOccupied(const SceneView& s, uint32_t scene, const float (&ov)[3], const float (&inv)[3],
                           float t_enter, float& t_exit)
{
    const int32_t* b = s.box + uint64_t(scene) * 6;
    const int super_dim = int(s.grid_dim >> 3);
    float box_enter = t_enter;
    float box_exit = t_exit;
    bool empty = false;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        const int lo = max(b[a], 1) - 1;
        const int hi = min(b[3 + a], super_dim - 2) + 1;
        empty = empty || b[a] > b[3 + a];
        const float g0 = float(s.grid_min[a] * 8);
        const float t0 = (float(lo * 64) + g0 - ov[a]) * inv[a];
        const float t1 = (float((hi + 1) * 64) + g0 - ov[a]) * inv[a];
        box_enter = fmaxf(box_enter, fminf(t0, t1));
        box_exit = fminf(box_exit, fmaxf(t0, t1));
    }
    t_exit = empty || box_enter > box_exit ? -1.0f : box_exit;
}

CX_DEV WorldHit TraceWorld(const SceneView& s, V3 o, V3 dir, float t_max, uint32_t scene = 0)
{
    WorldHit result{0.0f, 0u, kHashNone, {0, 0, 0}, 0u, 0u, 0u};
    const RayInv ri = InverseOf(dir);
    const float g0[3] = {float(s.grid_min[0] * 8), float(s.grid_min[1] * 8), float(s.grid_min[2] * 8)};
    const float extent = float(s.grid_dim * 8);
    const float ov[3] = {o.x, o.y, o.z};
    const float dv[3] = {dir.x, dir.y, dir.z};
    float inv[3];
    int step[3];
    float t_enter = 0.0f;
    float t_exit = t_max;
    for (int a = 0; a < 3; ++a)
    {
        inv[a] = SafeInvAxis(dv[a]);
        step[a] = dv[a] >= 0.0f ? 1 : -1;
        const float t0 = (g0[a] - ov[a]) * inv[a];
        const float t1 = (g0[a] + extent - ov[a]) * inv[a];
        t_enter = fmaxf(t_enter, fminf(t0, t1));
        t_exit = fminf(t_exit, fmaxf(t0, t1));
    }
    if (s.box != nullptr && t_enter <= t_exit)
    {
        ClipToOccupied(s, scene, ov, inv, t_enter, t_exit);
    }
    if (t_enter > t_exit)
    {
        return result;
    }
    const int super_dim = int(s.grid_dim >> 3);
    // Super cell of the entry point.
    int sc[3];
    for (int a = 0; a < 3; ++a)
    {
        const float p = ov[a] + dv[a] * t_enter;
        int c = int(floorf((p - g0[a]) * (1.0f / 64.0f)));
        sc[a] = min(max(c, 0), super_dim - 1);
    }
    float t_cell = t_enter;
    while (true)
    {
        // Exit of the current super cell (plane crossings (k * 64 + g0 - o) * inv).
        float t_next[3];
        for (int a = 0; a < 3; ++a)
        {
            const int k = sc[a] + (step[a] > 0 ? 1 : 0);
            t_next[a] = (float(k * 64) + g0[a] - ov[a]) * inv[a];
        }
        const float t_leave = fminf(fminf(t_next[0], t_next[1]), t_next[2]);
        const uint64_t super_index = scene * s.scene_supers + (uint64_t(sc[2]) * super_dim + sc[1]) * super_dim + sc[0];
        if (BitSet(s.super_bits, super_index))
        {
            // The bricks of this super cell along the ray, from the entry t_cell.
            int bc[3];
            for (int a = 0; a < 3; ++a)
            {
                const float p = ov[a] + dv[a] * t_cell;
                int c = int(floorf((p - g0[a]) * 0.125f));
                bc[a] = min(max(c, sc[a] * 8), sc[a] * 8 + 7);
            }
            while (true)
            {
                result.cells += 1;
                const uint64_t brick_index = scene * s.scene_cells + (uint64_t(bc[2]) * s.grid_dim + bc[1]) * s.grid_dim + bc[0];
                if (BitSet(s.brick_bits, brick_index))
                {
                    const int32_t key[3] = {s.grid_min[0] + bc[0], s.grid_min[1] + bc[1], s.grid_min[2] + bc[2]};
                    uint32_t probes = 0;
                    const uint32_t record = HashGet(s.slots, s.slot_count, s.pow2, s.hash_info[0],
                                                    SceneBrickFingerprint(s, scene, key[0], key[1], key[2]), &probes);
                    result.probes += probes;
                    if (record != kHashNone)
                    {
                        const V3 origin = Make(float(key[0] * 8), float(key[1] * 8), float(key[2] * 8));
                        const BrickHit h = TraceBrick(s, record, o, dir, ri, t_max, origin);
                        result.steps += h.steps;
                        if (h.attrib != 0u && h.t < t_max)
                        {
                            result.t = h.t;
                            result.attrib = h.attrib;
                            result.record = record;
                            result.brick[0] = key[0];
                            result.brick[1] = key[1];
                            result.brick[2] = key[2];
                            return result;
                        }
                    }
                }
                // Next brick cell.
                float tb[3];
                for (int a = 0; a < 3; ++a)
                {
                    const int k = bc[a] + (step[a] > 0 ? 1 : 0);
                    tb[a] = (float(k * 8) + g0[a] - ov[a]) * inv[a];
                }
                int axis = 0;
                if (tb[1] < tb[axis])
                {
                    axis = 1;
                }
                if (tb[2] < tb[axis])
                {
                    axis = 2;
                }
                if (tb[axis] > t_max)
                {
                    return result;
                }
                bc[axis] += step[axis];
                if (bc[axis] < sc[axis] * 8 || bc[axis] > sc[axis] * 8 + 7)
                {
                    break;  // left the super cell
                }
            }
        }
        int axis = 0;
        if (t_next[1] < t_next[axis])
        {
            axis = 1;
        }
        if (t_next[2] < t_next[axis])
        {
            axis = 2;
        }
        if (t_leave > t_exit)
        {
            return result;
        }
        sc[axis] += step[axis];
        if (sc[axis] < 0 || sc[axis] >= super_dim)
        {
            return result;
        }
        t_cell = t_next[axis];
    }
}

// Work counters of a lane (summed over a wave, then added to the frame's counters).
struct PathCounters
{
    uint32_t rays;        // primary + bounce rays traced
    uint32_t shadow_rays;
    uint32_t cells;
    uint32_t probes;
    uint32_t steps;
};

// ---- resumable world traversal ------------------------------------------------------------------
// TraceWorld as a state machine that advances one brick cell (or one super cell) per step, so that a
// persistent kernel can give a lane a new ray as soon as its ray ends instead of leaving it idle
// until the longest ray of its wave ends. Every float value is computed by the same expression as in
// TraceWorld, so both give the same hit and the same cell / probe / step counts.
constexpr uint32_t kRayDone = 0;     // finished
constexpr uint32_t kRaySuper = 1;    // at the top of the super-cell loop for sc
constexpr uint32_t kRayBricks = 2;   // at brick cell bc inside the resident super cell sc
constexpr uint32_t kRayAdvance = 3;  // leaving super cell sc
constexpr uint32_t kRayBrick = 4;    // at the resident brick cell bc, its brick not yet traced

struct WorldRay
{
    V3 o;
    V3 dir;
    RayInv ri;
    float t_max;
    float t_exit;
    float t_cell;
    int sc0, sc1, sc2;
    int bc0, bc1, bc2;
    uint32_t phase;
    uint32_t scene;
};

CX_DEV float GridOrigin(const SceneView& s, int a)
{
    return float(s.grid_min[a] * 8);
}

CX_DEV int StepOf(float d)
{
    return d >= 0.0f ? 1 : -1;
}

// Starts a ray; phase kRayDone when it misses the grid.
CX_DEV void WorldRayBegin(const SceneView& s, V3 o, V3 dir, float t_max, uint32_t scene, WorldRay& r)
{
    r.scene = scene;
    r.o = o;
    r.dir = dir;
    r.ri = InverseOf(dir);
    r.t_max = t_max;
    const float extent = float(s.grid_dim * 8);
    float t_enter = 0.0f;
    float t_exit = t_max;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        const float g0 = GridOrigin(s, a);
        const float inv = Get(r.ri.safe, a);
        const float t0 = (g0 - Get(o, a)) * inv;
        const float t1 = (g0 + extent - Get(o, a)) * inv;
        t_enter = fmaxf(t_enter, fminf(t0, t1));
        t_exit = fminf(t_exit, fmaxf(t0, t1));
    }
    if (s.box != nullptr && t_enter <= t_exit)
    {
        const float ov[3] = {o.x, o.y, o.z};
        const float inv[3] = {r.ri.safe.x, r.ri.safe.y, r.ri.safe.z};
        ClipToOccupied(s, scene, ov, inv, t_enter, t_exit);
    }
    r.t_exit = t_exit;
    r.t_cell = t_enter;
    if (t_enter > t_exit)
    {
        r.phase = kRayDone;
        return;
    }
    const int super_dim = int(s.grid_dim >> 3);
    int sc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        const float p = Get(o, a) + Get(dir, a) * t_enter;
        const int c = int(floorf((p - GridOrigin(s, a)) * (1.0f / 64.0f)));
        sc[a] = min(max(c, 0), super_dim - 1);
    }
    r.sc0 = sc[0];
    r.sc1 = sc[1];
    r.sc2 = sc[2];
    r.phase = kRaySuper;
}

// The ray's move past brick cell bc: to the next brick cell of the super cell (phase kRayBricks), out
// of the super cell (kRayAdvance) or, past t_max, to the end (kRayDone).
CX_DEV void WorldRayNextCell(const SceneView& s, WorldRay& r)
{
    const int st0 = StepOf(r.dir.x), st1 = StepOf(r.dir.y), st2 = StepOf(r.dir.z);
    const float tb0 = (float((r.bc0 + (st0 > 0 ? 1 : 0)) * 8) + GridOrigin(s, 0) - r.o.x) * r.ri.safe.x;
    const float tb1 = (float((r.bc1 + (st1 > 0 ? 1 : 0)) * 8) + GridOrigin(s, 1) - r.o.y) * r.ri.safe.y;
    const float tb2 = (float((r.bc2 + (st2 > 0 ? 1 : 0)) * 8) + GridOrigin(s, 2) - r.o.z) * r.ri.safe.z;
    int axis = 0;
    float t_axis = tb0;
    if (tb1 < t_axis)
    {
        axis = 1;
        t_axis = tb1;
    }
    if (tb2 < t_axis)
    {
        axis = 2;
        t_axis = tb2;
    }
    if (t_axis > r.t_max)
    {
        r.phase = kRayDone;
        return;
    }
    bool left;
    if (axis == 0)
    {
        r.bc0 += st0;
        left = r.bc0 < r.sc0 * 8 || r.bc0 > r.sc0 * 8 + 7;
    }
    else if (axis == 1)
    {
        r.bc1 += st1;
        left = r.bc1 < r.sc1 * 8 || r.bc1 > r.sc1 * 8 + 7;
    }
    else
    {
        r.bc2 += st2;
        left = r.bc2 < r.sc2 * 8 || r.bc2 > r.sc2 * 8 + 7;
    }
    r.phase = left ? kRayAdvance : kRayBricks;
}

// One cheap step of the ray: the super-cell test, the bitmap test of one brick cell, or the super-cell
// advance. A resident brick cell stops the ray in phase kRayBrick (WorldRayBrick traces it); a ray that
// leaves the grid or passes t_max ends (phase kRayDone: a miss). The counters count this step's work.
CX_DEV void WorldRayCellStep(const SceneView& s, WorldRay& r, PathCounters& c)
{
    const int super_dim = int(s.grid_dim >> 3);
    if (r.phase == kRaySuper)
    {
        const uint64_t super_index =
            r.scene * s.scene_supers + (uint64_t(r.sc2) * super_dim + r.sc1) * super_dim + r.sc0;
        if (BitSet(s.super_bits, super_index))
        {
            const int sc[3] = {r.sc0, r.sc1, r.sc2};
            int bc[3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
            {
                const float p = Get(r.o, a) + Get(r.dir, a) * r.t_cell;
                const int cc = int(floorf((p - GridOrigin(s, a)) * 0.125f));
                bc[a] = min(max(cc, sc[a] * 8), sc[a] * 8 + 7);
            }
            r.bc0 = bc[0];
            r.bc1 = bc[1];
            r.bc2 = bc[2];
            r.phase = kRayBricks;
        }
        else
        {
            r.phase = kRayAdvance;
        }
    }
    if (r.phase == kRayBricks)
    {
        c.cells += 1;
        const uint64_t brick_index =
            r.scene * s.scene_cells + (uint64_t(r.bc2) * s.grid_dim + r.bc1) * s.grid_dim + r.bc0;
        if (BitSet(s.brick_bits, brick_index))
        {
            r.phase = kRayBrick;
            return;
        }
        WorldRayNextCell(s, r);
        if (r.phase != kRayAdvance)
        {
            return;
        }
    }
    if (r.phase == kRayAdvance)
    {
        const int st0 = StepOf(r.dir.x), st1 = StepOf(r.dir.y), st2 = StepOf(r.dir.z);
        const float tn0 = (float((r.sc0 + (st0 > 0 ? 1 : 0)) * 64) + GridOrigin(s, 0) - r.o.x) * r.ri.safe.x;
        const float tn1 = (float((r.sc1 + (st1 > 0 ? 1 : 0)) * 64) + GridOrigin(s, 1) - r.o.y) * r.ri.safe.y;
        const float tn2 = (float((r.sc2 + (st2 > 0 ? 1 : 0)) * 64) + GridOrigin(s, 2) - r.o.z) * r.ri.safe.z;
        const float t_leave = fminf(fminf(tn0, tn1), tn2);
        int axis = 0;
        float t_axis = tn0;
        if (tn1 < t_axis)
        {
            axis = 1;
            t_axis = tn1;
        }
        if (tn2 < t_axis)
        {
            axis = 2;
            t_axis = tn2;
        }
        if (t_leave > r.t_exit)
        {
            r.phase = kRayDone;
            return;
        }
        int moved;
        if (axis == 0)
        {
            moved = r.sc0 += st0;
        }
        else if (axis == 1)
        {
            moved = r.sc1 += st1;
        }
        else
        {
            moved = r.sc2 += st2;
        }
        if (moved < 0 || moved >= super_dim)
        {
            r.phase = kRayDone;
            return;
        }
        r.t_cell = t_axis;
        r.phase = kRaySuper;
    }
}

// The resident brick cell bc of a ray in phase kRayBrick: the hash lookup and, for a brick in the table,
// the brick DDA. A hit ends the ray (phase kRayDone, `hit` set); otherwise the ray moves to the next
// cell. The counters count the probes and DDA steps.
CX_DEV void WorldRayBrick(const SceneView& s, WorldRay& r, WorldHit& hit, PathCounters& c)
{
    const int32_t key0 = s.grid_min[0] + r.bc0;
    const int32_t key1 = s.grid_min[1] + r.bc1;
    const int32_t key2 = s.grid_min[2] + r.bc2;
    uint32_t probes = 0;
    const uint32_t record = HashGet(s.slots, s.slot_count, s.pow2, s.hash_info[0],
                                    SceneBrickFingerprint(s, r.scene, key0, key1, key2), &probes);
    c.probes += probes;
    if (record != kHashNone)
    {
        const V3 origin = Make(float(key0 * 8), float(key1 * 8), float(key2 * 8));
        const BrickHit h = TraceBrick(s, record, r.o, r.dir, r.ri, r.t_max, origin);
        c.steps += h.steps;
        if (h.attrib != 0u && h.t < r.t_max)
        {
            hit.t = h.t;
            hit.attrib = h.attrib;
            hit.record = record;
            hit.brick[0] = key0;
            hit.brick[1] = key1;
            hit.brick[2] = key2;
            r.phase = kRayDone;
            return;
        }
    }
    WorldRayNextCell(s, r);
}

// How a wave advances its rays: the lanes waiting at a resident brick (phase kRayBrick) trace their
// bricks together once at least half the busy lanes wait, or when no busy lane can take a cheap step;
// otherwise the other busy lanes take up to kCellSteps cheap steps each. The hash lookup and brick DDA
// are the costly part of a ray's traversal; run only when most lanes need them, they keep most of the
// wave's lanes busy instead of a few. Each ray still visits its cells in order, so hits and counters
// are those of TraceWorld. Returns true for the lanes whose ray ended in this call (hit or miss in `hit`).
constexpr uint32_t kCellSteps = 4;
CX_DEV bool WorldRayAdvance(const SceneView& s, WorldRay& r, bool busy, WorldHit& hit, PathCounters& c)
{
    hit = WorldHit{0.0f, 0u, kHashNone, {0, 0, 0}, 0u, 0u, 0u};
    const uint64_t busy_lanes = Ballot(busy);
    const uint64_t waiting = Ballot(busy && r.phase == kRayBrick);
    const bool bricks = waiting == busy_lanes || 2u * PopC(waiting) >= PopC(busy_lanes);
    if (!busy)
    {
        return false;
    }
    if (bricks)
    {
        if (r.phase == kRayBrick)
        {
            WorldRayBrick(s, r, hit, c);
        }
    }
    else if (r.phase != kRayBrick)
    {
        for (uint32_t k = 0; k < kCellSteps; ++k)
        {
            WorldRayCellStep(s, r, c);
            if (r.phase == kRayBrick || r.phase == kRayDone)
            {
                break;
            }
        }
    }
    return r.phase == kRayDone;
}

}  // namespace cx
