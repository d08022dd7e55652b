// ORACLE — test infrastructure only. Sequential CPU restatement of the trace side of the hot path:
// the record producer (shaders/scene/caelix_finalize.hlsl:112-361), the light pass
// (shaders/light/records/caelix_light_tree_build.hlsl:93-283, caelix_light_gather_emit.hlsl:36-86),
// the brick DDA (shaders/render/caelix_brick_trace.hlsl:203-451), the world traversal through the hash
// table, the light tree's sample and pdf (shaders/light/caelix_light_tree.hlsl:85-777), the MIS weight
// (shaders/render/caelix_light_nee.hlsl:20-34) and the path tracer
// (shaders/render/caelix_path_trace.hlsl:276-452, 595-978): opaque Lambertian voxels, and with the
// extended material table the full material model (shaders/render/caelix_shading.hlsl:116-307: media,
// glass, specular lobes, the delta chain, transmissive shadows of caelix_scene_trace.hlsl:401-477),
// with the subsurface table the Henyey-Greenstein lattice walk (shaders/render/caelix_subsurface.hlsl);
// the procedural sky and sun, the fog of the air.
//
// Every float operation is one IEEE binary32 operation in source order (-ffp-contract=off), as in the
// gfx950 kernels, so decisions (DDA steps, draws, Fresnel choices, roulette) match; the path tracer
// accumulates radiance in float64. sin/cos and exp: the Cephes single-precision polynomials
// (S. Moshier, Cephes Math Library).
//
// Pinning (tests/):
//   brick/world DDA   brute force: the first opaque voxel along the ray by exact rational stepping
//                     on axis-aligned and grazing-free random rays; step counts bounded (<= 22);
//                     bricks found by hash agree with a dict of the scene
//   finalize          numpy restatement of the reference finalize test's Reference() (face bits,
//                     alive rule, occupancy, coarse byte, bounds) and its glass scenario's checks
//   light records     numpy recomputation of open light faces, families, counts, box
//   light sampling    sum of pdf over all light faces = 1 per receiver (exhaustive), draw pdf = face
//                     pdf, histogram of draws vs pdf (chi-square), as the reference test
//                     (tests/light_tree_test.cpp:888-1068)
//   path tracer       closed forms: the furnace box, the sun-lit floor (L sin^2 r), constant and graded
//                     skies, the fogged camera segment and the isothermal fogged furnace, the glass
//                     slab (Fresnel share, Snell path, Beer transmission, depth), sun through a
//                     floating slab with transmissive shadows; RIS unbiasedness (sun + lamp); the
//                     subsurface walk: the HG cosine distribution (mean g, KS), the reference's furnace
//                     (a pixel is exactly 1 or 0, the step cap) and lifecycle checks, a one-step closed
//                     form (start voxel, entry jitter, dominant axis) and the guide albedo / branch
//                     weights (tests/test_sss_oracle.py); beyond
//                     these, radiance at matched streams on terrain scenes is checked against this
//                     oracle only
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include <omp.h>

#include "caelix_amd/caelix_amd.h"
#include "oracle_common.h"

extern "C" uint32_t oracle_morton3(uint32_t x, uint32_t y, uint32_t z);

namespace {

uint32_t Morton3(uint32_t x, uint32_t y, uint32_t z)
{
    return oracle_morton3(x, y, z);
}

struct V
{
    float x, y, z;
};
V Mk(float x, float y, float z) { return V{x, y, z}; }
V Add(V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; }
V Sub(V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; }
V Mul(V a, float s) { return V{a.x * s, a.y * s, a.z * s}; }
V MulV(V a, V b) { return V{a.x * b.x, a.y * b.y, a.z * b.z}; }
float Dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
float Comp(V v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
void SetComp(V& v, int i, float s)
{
    (i == 0 ? v.x : (i == 1 ? v.y : v.z)) = s;
}
V SafeNormalize(V v)
{
    const float dp3 = std::fmax(1.175494351e-38f, Dot(v, v));
    return Mul(v, 1.0f / std::sqrt(dp3));
}

constexpr float kPi = 3.1415926535f;
constexpr float kTwoPi = 6.283185307f;
constexpr float kRayTMax = 1e30f;
constexpr float kPushOff = 0.002f;
constexpr uint32_t kNoneU = 0xFFFFFFFFu;

void SinCos(float a, float& s, float& c)
{
    int j = int(a * 1.27323954473516f);
    float y = float(j);
    if ((j & 1) != 0)
    {
        j += 1;
        y += 1.0f;
    }
    j &= 7;
    const float x = ((a - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
    const float z = x * x;
    const float ps = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
    const float pc = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z - 0.5f * z + 1.0f;
    int js = j;
    bool neg_s = false;
    if (js > 3)
    {
        neg_s = true;
        js -= 4;
    }
    s = (js == 1 || js == 2) ? pc : ps;
    if (neg_s)
    {
        s = -s;
    }
    int jc = j;
    bool neg_c = false;
    if (jc > 3)
    {
        jc -= 4;
        neg_c = !neg_c;
    }
    if (jc > 1)
    {
        neg_c = !neg_c;
    }
    c = (jc == 1 || jc == 2) ? ps : pc;
    if (neg_c)
    {
        c = -c;
    }
}

uint32_t PcgHash(uint32_t value)
{
    const uint32_t state = value * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
uint32_t InitRandom(uint32_t px, uint32_t py, uint32_t seed)
{
    const uint32_t state = PcgHash(px ^ (py << 16) ^ (seed * 0x9E3779B9u));
    return state != 0u ? state : 0x9E3779B9u;
}
uint32_t RandomU32(uint32_t& state)
{
    state ^= state << 13;
    state ^= state >> 17;
    state ^= state << 5;
    return state;
}
float RandomFloat01(uint32_t& state)
{
    return float(RandomU32(state) >> 8) * (1.0f / 16777216.0f);
}
V RandomUnitVector(uint32_t& state)
{
    const float u = RandomFloat01(state);
    const float v = RandomFloat01(state);
    const float a = u * kTwoPi;
    const float y = v * 2.0f - 1.0f;
    const float sy = std::sqrt(std::fmax(1.0f - y * y, 0.0f));
    float s, c;
    SinCos(a, s, c);
    return Mk(s * sy, y, c * sy);
}
uint32_t LightRandom(uint32_t& rng)
{
    rng = rng * 747796405u + 2891336453u;
    const uint32_t word = ((rng >> ((rng >> 28) + 4u)) ^ rng) * 277803737u;
    return (word >> 22) ^ word;
}
float LightRandomFloat(uint32_t& rng)
{
    return float(LightRandom(rng) >> 8) * (1.0f / 16777216.0f);
}

// ---- hash ----
uint64_t SplitMix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
void BrickKey(int32_t x, int32_t y, int32_t z, uint32_t& lo, uint32_t& hi)
{
    uint64_t h = SplitMix64(0 ^ ((uint64_t(uint32_t(y)) << 32) | uint32_t(x)));
    h = SplitMix64(h ^ uint64_t(uint32_t(z)));
    lo = std::max(uint32_t(h), 1u);
    hi = uint32_t(h >> 32) != 0 ? uint32_t(h >> 32) : 1u;
}
// Demand bytes: every global load the kernels make, counted at its size (no cache model), per kind
// of work: 0 primary-ray traversal, 1 bounce-ray traversal, 2 shadow-ray traversal, 3 shading at the
// primary vertex (shade + NEE draw), 4 shading at bounce vertices (shade + MIS pdf + NEE draw). Used
// to state the algorithmic bytes of each stage of the trace; one set per OpenMP thread.
thread_local uint64_t g_demand[5] = {};
thread_local int g_kind = 0;

uint32_t HashGet(const CxScene& s, int32_t x, int32_t y, int32_t z, uint32_t& probes)
{
    uint32_t lo, hi;
    BrickKey(x, y, z, lo, hi);
    const bool pow2 = (s.slot_count & (s.slot_count - 1)) == 0;
    uint32_t slot = pow2 ? (lo & (s.slot_count - 1)) : (lo % s.slot_count);
    const uint32_t limit = s.hash_info[0];
    for (uint32_t p = 0; p < limit; ++p)
    {
        const HashKeySlot& k = s.slots[slot];
        g_demand[g_kind] += 16;
        if (k.key_lo == 0)
        {
            probes = p + 1;
            return kNoneU;
        }
        if (k.key_lo == lo && k.key_hi == hi)
        {
            probes = p + 1;
            return k.payload;
        }
        slot = slot + 1 == s.slot_count ? 0 : slot + 1;
    }
    probes = limit;
    return kNoneU;
}

bool Opaque(uint32_t block, uint32_t encoding)
{
    return (block & (encoding == kEncodingPacked ? kPackedOpaqueMask : kBlockIdOpaqueMask)) != 0;
}
uint32_t MaterialCode(uint32_t block, uint32_t encoding)
{
    return Opaque(block, encoding) ? block : (block & (encoding == kEncodingPacked ? kPackedIdMask : kBlockIdIdMask));
}
uint32_t FaceBits(uint32_t block, uint32_t encoding)
{
    return encoding == kEncodingPacked ? ((block >> kPackedFaceShift) & 0x3Fu) : ((block >> kBlockIdFaceShift) & 0x3Fu);
}

// ---- brick DDA ----
struct BrickHit
{
    float t;
    uint32_t attrib;
    uint32_t steps;
};

float SafeInvAxis(float d)
{
    if (std::fabs(d) >= 1e-20f)
    {
        return 1.0f / d;
    }
    return d < 0.0f ? -1.0f / 1e-20f : 1.0f / 1e-20f;
}

BrickHit TraceBrickRay(const uint64_t* occ, const uint16_t* payload, uint32_t encoding, V entry, V dir, float entry_t,
                       uint32_t entry_normal_coarse)
{
    const V inv = Mk(SafeInvAxis(dir.x), SafeInvAxis(dir.y), SafeInvAxis(dir.z));
    const V boundary = Mk(dir.x >= 0.0f ? 1.0f : 0.0f, dir.y >= 0.0f ? 1.0f : 0.0f, dir.z >= 0.0f ? 1.0f : 0.0f);
    const V t_start = MulV(Sub(boundary, entry), inv);
    const float hi = 8.0f - 0.0001f;
    const V origin = Mk(std::fmin(std::fmax(entry.x, 0.0f), hi), std::fmin(std::fmax(entry.y, 0.0f), hi),
                        std::fmin(std::fmax(entry.z, 0.0f), hi));
    int cell[3] = {int(std::floor(origin.x)), int(std::floor(origin.y)), int(std::floor(origin.z))};
    int prev[3] = {cell[0], cell[1], cell[2]};
    BrickHit hit{0.0f, 0u, 0u};
    for (uint32_t step = 0; step < 22; ++step)
    {
        hit.steps = step + 1;
        if (uint32_t(cell[0]) >= 8u || uint32_t(cell[1]) >= 8u || uint32_t(cell[2]) >= 8u)
        {
            break;
        }
        const uint32_t coarse_bit = uint32_t((cell[0] >> 2) | ((cell[1] >> 2) << 1) | ((cell[2] >> 2) << 2));
        if ((entry_normal_coarse & (1u << coarse_bit)) == 0u)
        {
            // VRTJump (caelix_brick_trace.hlsl:286-289), then VRTStep.
            cell[0] = dir.x < 0.0f ? (cell[0] & ~3) : (cell[0] | 3);
            cell[1] = dir.y < 0.0f ? (cell[1] & ~3) : (cell[1] | 3);
            cell[2] = dir.z < 0.0f ? (cell[2] & ~3) : (cell[2] | 3);
        }
        else
        {
            const uint32_t micro = uint32_t((cell[0] & 3) | ((cell[1] & 3) << 2) | ((cell[2] & 3) << 4));
            g_demand[g_kind] += 8;
            if (((occ[coarse_bit] >> micro) & 1ull) != 0ull)
            {
                const uint32_t block = payload[cell[0] + 8 * cell[1] + 64 * cell[2]];
                g_demand[g_kind] += 2;
                float T = 0.0f;
                uint32_t flags;
                if (step == 0)
                {
                    flags = entry_normal_coarse >> 26;
                }
                else
                {
                    const V side = Add(t_start, MulV(Mk(float(prev[0]), float(prev[1]), float(prev[2])), inv));
                    T = std::fmin(std::fmin(side.x, side.y), side.z);
                    const float lim = T + 0.0001f;
                    const uint32_t x_use = side.x <= lim ? 0xFFFFFFFFu : 0u;
                    const uint32_t y_use = (side.y <= lim ? 0xFFFFFFFFu : 0u) & ~x_use;
                    const uint32_t z_use = ~(x_use | y_use);
                    flags = ((dir.x > 0.0f ? 0x02u : 0x01u) & x_use) | ((dir.y > 0.0f ? 0x08u : 0x04u) & y_use) |
                            ((dir.z > 0.0f ? 0x20u : 0x10u) & z_use);
                }
                bool terminate = Opaque(block, encoding);
                if (!terminate && !(step == 0 && entry_t == 0.0f))
                {
                    terminate = (FaceBits(block, encoding) & flags) != 0u;
                }
                if (terminate)
                {
                    hit.attrib = (block << 16) | (uint32_t(cell[0] | (cell[1] << 3) | (cell[2] << 6)) << 6) | flags;
                    hit.t = entry_t + T;
                    return hit;
                }
            }
        }
        prev[0] = cell[0];
        prev[1] = cell[1];
        prev[2] = cell[2];
        const V side = Add(t_start, MulV(Mk(float(cell[0]), float(cell[1]), float(cell[2])), inv));
        const float tmin = std::fmin(std::fmin(side.x, side.y), side.z) + 0.0001f;
        const V pos = Add(origin, Mul(dir, tmin));
        cell[0] = int(std::floor(pos.x));
        cell[1] = int(std::floor(pos.y));
        cell[2] = int(std::floor(pos.z));
    }
    hit.attrib = 0u;
    hit.t = 0.0f;
    return hit;
}

BrickHit TraceBrick(const CxScene& s, uint32_t record, V o, V dir, float t_current, V brick_origin)
{
    const uint32_t tb = s.info[2 * uint64_t(record) + 1];
    g_demand[g_kind] += 8;
    const V tmin_box = Mk(float(tb & 7u), float((tb >> 3) & 7u), float((tb >> 6) & 7u));
    const V tmax_box = Mk(float((tb >> 9) & 7u) + 1.0f, float((tb >> 12) & 7u) + 1.0f, float((tb >> 15) & 7u) + 1.0f);
    const V aabb_min = Add(brick_origin, tmin_box);
    const V aabb_max = Add(brick_origin, tmax_box);
    const V inv = Mk(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
    const V t0 = MulV(Sub(aabb_min, o), inv);
    const V t1 = MulV(Sub(aabb_max, o), inv);
    const V tn = Mk(std::fmin(t0.x, t1.x), std::fmin(t0.y, t1.y), std::fmin(t0.z, t1.z));
    const V tf = Mk(std::fmax(t0.x, t1.x), std::fmax(t0.y, t1.y), std::fmax(t0.z, t1.z));
    const float largest_tmin = std::fmax(std::fmax(tn.x, tn.y), tn.z);
    const float smallest_tmax = std::fmin(std::fmin(tf.x, tf.y), tf.z);
    if (largest_tmin > smallest_tmax || smallest_tmax < 0.0f || largest_tmin > t_current)
    {
        return BrickHit{0.0f, 0u, 0u};
    }
    const float t = std::fmax(0.0f, largest_tmin);
    const V entry = Sub(Add(o, Mul(dir, t)), brick_origin);
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
    const uint32_t coarse = (s.info[2 * uint64_t(record)] >> kRecordCoarseShift) & kRecordCoarseMask;
    const uint64_t* occ = static_cast<const uint64_t*>(s.occupancy) + uint64_t(record) * 8;
    return TraceBrickRay(occ, s.payload + uint64_t(record) * 512, s.encoding, entry, dir, t, (flags << 26) | coarse);
}

struct WorldHit
{
    float t;
    uint32_t attrib;
    uint32_t record;
    int32_t brick[3];
    uint32_t cells, probes, steps;
};

bool Bit(const uint32_t* bits, uint64_t i)
{
    g_demand[g_kind] += 4;
    return ((bits[i >> 5] >> (i & 31u)) & 1u) != 0u;
}

// Super-cell DDA, then brick DDA inside resident super cells; plane crossings (k * size + g0 - o) * inv.
WorldHit TraceWorld(const CxScene& s, V o, V dir, float t_max)
{
    WorldHit r{0.0f, 0u, kNoneU, {0, 0, 0}, 0, 0, 0};
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
        t_enter = std::fmax(t_enter, std::fmin(t0, t1));
        t_exit = std::fmin(t_exit, std::fmax(t0, t1));
    }
    if (t_enter > t_exit)
    {
        return r;
    }
    const int sd = int(s.grid_dim >> 3);
    int sc[3];
    for (int a = 0; a < 3; ++a)
    {
        const float p = ov[a] + dv[a] * t_enter;
        sc[a] = std::min(std::max(int(std::floor((p - g0[a]) * (1.0f / 64.0f))), 0), sd - 1);
    }
    float t_cell = t_enter;
    while (true)
    {
        float t_next[3];
        for (int a = 0; a < 3; ++a)
        {
            const int k = sc[a] + (step[a] > 0 ? 1 : 0);
            t_next[a] = (float(k * 64) + g0[a] - ov[a]) * inv[a];
        }
        const float t_leave = std::fmin(std::fmin(t_next[0], t_next[1]), t_next[2]);
        if (Bit(s.super_bits, (uint64_t(sc[2]) * sd + sc[1]) * sd + sc[0]))
        {
            int bc[3];
            for (int a = 0; a < 3; ++a)
            {
                const float p = ov[a] + dv[a] * t_cell;
                bc[a] = std::min(std::max(int(std::floor((p - g0[a]) * 0.125f)), sc[a] * 8), sc[a] * 8 + 7);
            }
            while (true)
            {
                r.cells += 1;
                if (Bit(s.brick_bits, (uint64_t(bc[2]) * s.grid_dim + bc[1]) * s.grid_dim + bc[0]))
                {
                    const int32_t key[3] = {s.grid_min[0] + bc[0], s.grid_min[1] + bc[1], s.grid_min[2] + bc[2]};
                    uint32_t probes = 0;
                    const uint32_t record = HashGet(s, key[0], key[1], key[2], probes);
                    r.probes += probes;
                    if (record != kNoneU)
                    {
                        const BrickHit h = TraceBrick(s, record, o, dir, t_max,
                                                      Mk(float(key[0] * 8), float(key[1] * 8), float(key[2] * 8)));
                        r.steps += h.steps;
                        if (h.attrib != 0u && h.t < t_max)
                        {
                            r.t = h.t;
                            r.attrib = h.attrib;
                            r.record = record;
                            r.brick[0] = key[0];
                            r.brick[1] = key[1];
                            r.brick[2] = key[2];
                            return r;
                        }
                    }
                }
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
                    return r;
                }
                bc[axis] += step[axis];
                if (bc[axis] < sc[axis] * 8 || bc[axis] > sc[axis] * 8 + 7)
                {
                    break;
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
            return r;
        }
        sc[axis] += step[axis];
        if (sc[axis] < 0 || sc[axis] >= sd)
        {
            return r;
        }
        t_cell = t_next[axis];
    }
}

// ---- light tree ----
struct Summ
{
    float face[6];
    V lo, hi;
};
float Power6(const float* f)
{
    return f[0] + f[1] + f[2] + f[3] + f[4] + f[5];
}
Summ SlotOf(const LightBvhNode& n, uint32_t s)
{
    g_demand[g_kind] += 48;
    Summ r;
    for (int f = 0; f < 6; ++f)
    {
        r.face[f] = n.child_face[s][f];
    }
    r.lo = Mk(n.child_lo[s][0], n.child_lo[s][1], n.child_lo[s][2]);
    r.hi = Mk(n.child_hi[s][0], n.child_hi[s][1], n.child_hi[s][2]);
    return r;
}
void EmitterCosines(V d_lo, V d_hi, float c[6])
{
    const V closest = Mk(std::fmin(std::fmax(0.0f, d_lo.x), d_hi.x), std::fmin(std::fmax(0.0f, d_lo.y), d_hi.y),
                         std::fmin(std::fmax(0.0f, d_lo.z), d_hi.z));
    const V c2 = MulV(closest, closest);
    const float all2 = c2.x + c2.y + c2.z;
    for (int a = 0; a < 3; ++a)
    {
        const float others = all2 - Comp(c2, a);
        const float dh = Comp(d_hi, a);
        const float dl = Comp(d_lo, a);
        c[2 * a] = dh > 0.0f ? dh / std::sqrt(dh * dh + others) : 0.0f;
        c[2 * a + 1] = dl < 0.0f ? -dl / std::sqrt(dl * dl + others) : 0.0f;
    }
}
// A receiver: the box [lo, hi] (lo = hi for a point) with the unit normal n (LtReceiver,
// caelix_light_tree.hlsl:368-395); `point` selects the point form of LtImportance.
struct Recv
{
    V lo, hi, n;
    bool point;
};
Recv PointRecv(V x, V n)
{
    return Recv{x, x, n, true};
}
// LtImportance (caelix_light_tree.hlsl:105-155): emitter cosines over D = [r_lo - hi, r_hi - lo], the
// receiver bound cr = min(1, m / dmin) with m = n.(c_B - c_R) + |n|.(h_B + h_R), the distance term
// max(|c_R - c_B|^2, max(1, extent^2)), extent^2 = |h_B|^2 for a point, (|h_B| + |h_R|)^2 for a box.
float ImportanceR(const Recv& r, V lo, V hi, const float face[6])
{
    float c[6];
    EmitterCosines(Sub(r.lo, hi), Sub(r.hi, lo), c);
    float e = 0.0f;
    for (int f = 0; f < 6; ++f)
    {
        e += face[f] * c[f];
    }
    if (e <= 0.0f)
    {
        return 0.0f;
    }
    const V centre = Mul(Add(lo, hi), 0.5f);
    const V half = Mul(Sub(hi, lo), 0.5f);
    const V rc = r.point ? r.lo : Mul(Add(r.lo, r.hi), 0.5f);
    const V rh = Mul(Sub(r.hi, r.lo), 0.5f);
    const V n = r.n;
    float cr = 1.0f;
    if (n.x != 0.0f || n.y != 0.0f || n.z != 0.0f)
    {
        const V an = Mk(std::fabs(n.x), std::fabs(n.y), std::fabs(n.z));
        float m = Dot(n, Sub(centre, rc)) + Dot(an, half);
        if (!r.point)
        {
            m += Dot(an, rh);
        }
        if (m <= 0.0f)
        {
            return 0.0f;
        }
        const V gap = Mk(std::fmax(std::fmax(lo.x - r.hi.x, r.lo.x - hi.x), 0.0f),
                         std::fmax(std::fmax(lo.y - r.hi.y, r.lo.y - hi.y), 0.0f),
                         std::fmax(std::fmax(lo.z - r.hi.z, r.lo.z - hi.z), 0.0f));
        const float dmin = std::sqrt(Dot(gap, gap));
        cr = dmin > 0.0f ? std::fmin(1.0f, m / dmin) : 1.0f;
    }
    const V tc = Sub(rc, centre);
    float ext2 = Dot(half, half);
    if (!r.point)
    {
        const float ext = std::sqrt(Dot(half, half)) + std::sqrt(Dot(rh, rh));
        ext2 = ext * ext;
    }
    const float d2 = std::fmax(Dot(tc, tc), std::fmax(1.0f, ext2));
    return (e * cr) / d2;
}
float Importance(V x, V n, V lo, V hi, const float face[6])
{
    return ImportanceR(PointRecv(x, n), lo, hi, face);
}
struct Pick
{
    float sum_i = 0.0f, sum_p = 0.0f, weight_i = 0.0f, weight_p = 0.0f;
    uint32_t index_i = kNoneU, index_p = kNoneU;
};
void PickAdd(Pick& p, float importance, float power, uint32_t index, float u)
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
uint32_t PickEnd(const Pick& p, float& probability)
{
    if (p.sum_i > 0.0f)
    {
        probability = p.weight_i / p.sum_i;
        return p.index_i;
    }
    probability = p.sum_p > 0.0f ? p.weight_p / p.sum_p : 0.0f;
    return p.sum_p > 0.0f ? p.index_p : kNoneU;
}
float PickProbability(float sum_i, float sum_p, float importance, float power)
{
    if (sum_i > 0.0f)
    {
        return importance / sum_i;
    }
    return sum_p > 0.0f ? power / sum_p : 0.0f;
}
// The light tree opened from its descriptor (LightTreeOpen / LightTreeView, caelix_light_tree.hlsl:
// 239-287): the oracle's descriptor holds host addresses of the same tables.
struct Lights
{
    const LightTreeHeader* headers;
    const LightBvhNode* nodes[2];
    const LightBvhLeaf* leaves[2];
    const uint32_t* heap;  // the current buffer
    const GroupLightRow* rows;
    const GroupInfo* group_infos;
    const EntityInfo* entity_infos;
    const uint32_t* light_codes;
    uint32_t group_shift;
    uint32_t encoding;
    uint32_t mis;
    uint32_t trees;  // LightTreeSelect's mask: bit k = tree k takes part
};
Lights OpenLights(const CxLightTreeView& v)
{
    const LightTreeDescriptor* d = v.descriptor;
    Lights l;
    l.headers = reinterpret_cast<const LightTreeHeader*>(d->tree_headers);
    for (int k = 0; k < 2; ++k)
    {
        l.nodes[k] = reinterpret_cast<const LightBvhNode*>(d->nodes[k]);
        l.leaves[k] = reinterpret_cast<const LightBvhLeaf*>(d->leaves[k]);
    }
    const uint32_t current = reinterpret_cast<const LightControl*>(d->control)->current & 1u;
    l.heap = reinterpret_cast<const uint32_t*>(d->heap[current]);
    l.rows = reinterpret_cast<const GroupLightRow*>(d->group_rows);
    l.group_infos = reinterpret_cast<const GroupInfo*>(d->group_infos);
    l.entity_infos = reinterpret_cast<const EntityInfo*>(d->entity_infos);
    l.light_codes = reinterpret_cast<const uint32_t*>(d->light_codes);
    l.group_shift = d->group_shift;
    l.encoding = d->encoding;
    l.mis = v.mis;
    l.trees = v.select != 0u ? (v.select & 3u) : 3u;
    return l;
}
Summ TreeSummary(const Lights& L, uint32_t k, uint32_t& root)
{
    root = ((L.trees >> k) & 1u) != 0u ? L.headers[k].root : kNoneU;
    g_demand[g_kind] += 64;
    Summ e;
    if (root == kNoneU)
    {
        std::memset(&e, 0, sizeof(e));
        return e;
    }
    for (int f = 0; f < 6; ++f)
    {
        e.face[f] = L.headers[k].summary.face[f];
    }
    e.lo = Mk(L.headers[k].summary.lo[0], L.headers[k].summary.lo[1], L.headers[k].summary.lo[2]);
    e.hi = Mk(L.headers[k].summary.hi[0], L.headers[k].summary.hi[1], L.headers[k].summary.hi[2]);
    return e;
}
uint32_t ChildCount(const uint32_t* c)
{
    g_demand[g_kind] += 16;
    return uint32_t(c[0] != kNoneU) + uint32_t(c[1] != kNoneU) + uint32_t(c[2] != kNoneU) + uint32_t(c[3] != kNoneU);
}
// LtChooseTreeR (caelix_light_tree.hlsl:417-437) and LtChooseLeafR (:466-512): a tree, then a leaf of
// it, each choice a weighted reservoir over importance with power as the fallback; pdf times every
// choice's probability. leaf_importance = the importance of the summary of the drawn leaf's slot (0 when
// the last choice fell back to the power; for the point form it is not asked for).
uint32_t ChooseLeafR(const Lights& L, const Recv& r, uint32_t& rng, float& pdf, uint32_t& tree, float& leaf_importance)
{
    leaf_importance = 0.0f;
    Pick pick;
    uint32_t roots[2];
    for (uint32_t k = 0; k < 2; ++k)
    {
        const Summ s = TreeSummary(L, k, roots[k]);
        PickAdd(pick, ImportanceR(r, s.lo, s.hi, s.face), Power6(s.face), k, LightRandomFloat(rng));
    }
    float probability;
    const uint32_t k = PickEnd(pick, probability);
    tree = k;
    if (k == kNoneU)
    {
        return kNoneU;
    }
    pdf *= probability;
    uint32_t node = roots[k];
    for (uint32_t depth = 0; depth < 64; ++depth)
    {
        const LightBvhNode& nd = L.nodes[k][node];
        uint32_t chosen;
        if (ChildCount(nd.child) == 1)
        {
            chosen = nd.child[0];
            if (!r.point && (chosen & kLtChildLeafBit) != 0)
            {
                const Summ slot = SlotOf(nd, 0);
                leaf_importance = ImportanceR(r, slot.lo, slot.hi, slot.face);
            }
        }
        else
        {
            Pick p;
            for (uint32_t s = 0; s < 4; ++s)
            {
                if (nd.child[s] == kNoneU)
                {
                    continue;
                }
                const Summ slot = SlotOf(nd, s);
                PickAdd(p, ImportanceR(r, slot.lo, slot.hi, slot.face), Power6(slot.face), s, LightRandomFloat(rng));
            }
            float prob;
            const uint32_t s = PickEnd(p, prob);
            if (s == kNoneU)
            {
                return kNoneU;
            }
            pdf *= prob;
            chosen = nd.child[s];
            leaf_importance = p.sum_i > 0.0f ? p.weight_i : 0.0f;
        }
        if ((chosen & kLtChildLeafBit) != 0)
        {
            return chosen & ~kLtChildLeafBit;
        }
        node = chosen;
    }
    return kNoneU;
}
uint32_t ChooseLeaf(const Lights& L, V x, V n, uint32_t& rng, float& pdf, uint32_t& tree)
{
    float unused;
    return ChooseLeafR(L, PointRecv(x, n), rng, pdf, tree, unused);
}
// LtSlotProbabilityR (:515-546): the reservoir's probability of `slot` as a closed form.
float SlotProbabilityR(const LightBvhNode& nd, const Recv& r, uint32_t slot)
{
    if (ChildCount(nd.child) == 1)
    {
        return 1.0f;
    }
    float sum_i = 0.0f, sum_p = 0.0f, it = 0.0f, pt = 0.0f;
    for (uint32_t s = 0; s < 4; ++s)
    {
        if (nd.child[s] == kNoneU)
        {
            continue;
        }
        const Summ sm = SlotOf(nd, s);
        const float i = ImportanceR(r, sm.lo, sm.hi, sm.face);
        const float p = Power6(sm.face);
        sum_i += i > 0.0f ? i : 0.0f;
        sum_p += p > 0.0f ? p : 0.0f;
        if (s == slot)
        {
            it = i;
            pt = p;
        }
    }
    return PickProbability(sum_i, sum_p, it, pt);
}
// LtTreeProbabilityR (:439-462)
float TreeProbabilityR(const Lights& L, const Recv& r, uint32_t target)
{
    float sum_i = 0.0f, sum_p = 0.0f, it = 0.0f, pt = 0.0f;
    for (uint32_t k = 0; k < 2; ++k)
    {
        uint32_t root;
        const Summ s = TreeSummary(L, k, root);
        const float i = ImportanceR(r, s.lo, s.hi, s.face);
        const float p = Power6(s.face);
        sum_i += i > 0.0f ? i : 0.0f;
        sum_p += p > 0.0f ? p : 0.0f;
        if (k == target)
        {
            it = i;
            pt = p;
        }
    }
    return PickProbability(sum_i, sum_p, it, pt);
}
// LtLeafProbabilityR (:549-572): the product of the slot probabilities up the parent links, times the
// tree's; 0 for an empty or unselected tree.
float LeafProbabilityR(const Lights& L, uint32_t k, uint32_t leaf, const Recv& r)
{
    uint32_t root;
    TreeSummary(L, k, root);
    if (root == kNoneU)
    {
        return 0.0f;
    }
    const uint32_t parent = L.leaves[k][leaf].parent;
    g_demand[g_kind] += 4;
    if (parent == kNoneU)
    {
        return 0.0f;
    }
    uint32_t node = parent >> 2;
    uint32_t slot = parent & 3u;
    float probability = 1.0f;
    for (uint32_t depth = 0; depth < 64; ++depth)
    {
        const LightBvhNode& nd = L.nodes[k][node];
        probability *= SlotProbabilityR(nd, r, slot);
        if (nd.parent == kNoneU)
        {
            return probability * TreeProbabilityR(L, r, k);
        }
        slot = nd.parent_slot;
        node = nd.parent;
        g_demand[g_kind] += 8;
    }
    return 0.0f;
}
float LeafProbability(const Lights& L, uint32_t k, uint32_t leaf, V x, V n)
{
    return LeafProbabilityR(L, k, leaf, PointRecv(x, n));
}
struct Rec
{
    float face[6];
    uint32_t lo[3], hi[3];
    uint32_t pos, entry_count, counts_lo, counts_hi;
    const uint32_t* entries;
};
// A group's light block in the current heap buffer, or null (LtGroupRow, caelix_light_tree.hlsl:339-343).
const uint32_t* GroupBlock(const Lights& L, uint32_t group)
{
    g_demand[g_kind] += 8;
    const GroupLightRow& row = L.rows[group];
    return row.block_units != 0u ? L.heap + uint64_t(row.block_offset) * 4 : nullptr;
}
const LightIndexEntry& IndexEntryOf(const uint32_t* block, uint32_t entry)
{
    return reinterpret_cast<const LightIndexEntry*>(block + 4)[entry];
}
// LtLoadEntryRecord (caelix_light_tree_layout.hlsl:377-401).
Rec LoadRec(const uint32_t* block, uint32_t entry)
{
    const uint32_t* r = block + uint64_t(IndexEntryOf(block, entry).record_offset) * 4;
    g_demand[g_kind] += 4 + 48;
    Rec rec;
    std::memcpy(rec.face, r, 24);
    for (int a = 0; a < 3; ++a)
    {
        rec.lo[a] = (r[6] >> (10 * a)) & 0x3FFu;
        rec.hi[a] = (r[7] >> (10 * a)) & 0x3FFu;
    }
    rec.pos = r[8];
    rec.entry_count = r[9];
    rec.counts_lo = r[10];
    rec.counts_hi = r[11];
    rec.entries = r + 12;
    return rec;
}
uint32_t FamilyCount(const Rec& r, uint32_t f)
{
    return ((f < 3 ? r.counts_lo : r.counts_hi) >> (10 * (f % 3))) & 0x3FFu;
}
void FamilyWeights(const Rec& r, V x, float w[6], float& sum)
{
    float c[6];
    EmitterCosines(Sub(x, Mk(float(r.hi[0]), float(r.hi[1]), float(r.hi[2]))),
                   Sub(x, Mk(float(r.lo[0]), float(r.lo[1]), float(r.lo[2]))), c);
    sum = 0.0f;
    for (int f = 0; f < 6; ++f)
    {
        w[f] = r.face[f] * c[f];
        sum += w[f];
    }
    if (sum <= 0.0f)
    {
        sum = 0.0f;
        for (int g = 0; g < 6; ++g)
        {
            w[g] = r.face[g];
            sum += w[g];
        }
    }
}
// LtEntityFrame + LtGroupOrigin (caelix_light_tree_layout.hlsl:412-433): the group's entity-local voxel
// origin (its coordinate times 8 << shift, in integers), its entity's rotation rows and translation.
struct Frame
{
    float r[3][3];
    V t;
};
V GroupFrame(const Lights& L, uint32_t group, Frame& f)
{
    g_demand[g_kind] += 16 + 48;
    const GroupInfo& info = L.group_infos[group];
    const EntityInfo& e = L.entity_infos[info.entity];
    for (int i = 0; i < 3; ++i)
    {
        for (int j = 0; j < 3; ++j)
        {
            f.r[i][j] = e.cur_rows[i][j];
        }
    }
    f.t = Mk(e.cur_rows[0][3], e.cur_rows[1][3], e.cur_rows[2][3]);
    const int32_t edge = int32_t(8u << L.group_shift);
    return Mk(float(info.coord[0] * edge), float(info.coord[1] * edge), float(info.coord[2] * edge));
}
// mul(x - t, R) - origin = R^T (x - t) - origin (caelix_light_tree.hlsl:691).
V ToGroup(const Frame& f, V x, V origin)
{
    const V d = Sub(x, f.t);
    const V l = Mk(d.x * f.r[0][0] + d.y * f.r[1][0] + d.z * f.r[2][0], d.x * f.r[0][1] + d.y * f.r[1][1] + d.z * f.r[2][1],
                   d.x * f.r[0][2] + d.y * f.r[1][2] + d.z * f.r[2][2]);
    return Sub(l, origin);
}
// mul(R, v)
V Rotate(const Frame& f, V v)
{
    return Mk(f.r[0][0] * v.x + f.r[0][1] * v.y + f.r[0][2] * v.z, f.r[1][0] * v.x + f.r[1][1] * v.y + f.r[1][2] * v.z,
              f.r[2][0] * v.x + f.r[2][1] * v.y + f.r[2][2] * v.z);
}
uint32_t MortonOfPos(uint32_t pos, uint32_t shift)
{
    const uint32_t mask = (1u << shift) - 1u;
    return Morton3(pos & mask, (pos >> shift) & mask, (pos >> (2 * shift)) & mask);
}
// LtFindIndexEntry (caelix_light_tree_layout.hlsl:336-360), restated as a linear scan for the entry
// with the key (the entries are in strictly increasing Morton order, so a binary search finds the same).
uint32_t FindIndexEntry(const uint32_t* block, uint32_t record_count, uint32_t key, uint32_t shift)
{
    for (uint32_t e = 0; e < record_count; ++e)
    {
        g_demand[g_kind] += 4;
        if (MortonOfPos(IndexEntryOf(block, e).pos, shift) == key)
        {
            return e;
        }
    }
    return kNoneU;
}
// LtBrickLeaf (caelix_light_tree.hlsl:594-607).
uint32_t BrickLeaf(const Lights& L, const uint32_t* block, uint32_t group, uint32_t entry)
{
    const uint32_t word = IndexEntryOf(block, entry).bvh_leaf;
    g_demand[g_kind] += 4;
    if (word == kNoneU)
    {
        return kNoneU;
    }
    const LightBvhLeaf& leaf = L.leaves[word >> 31][word & 0x7FFFFFFFu];
    g_demand[g_kind] += 8;
    return leaf.group == group && leaf.entry == entry ? word : kNoneU;
}
V FaceDirection(uint32_t f)
{
    const float sign = (f & 1) != 0 ? -1.0f : 1.0f;
    const uint32_t axis = f >> 1;
    return Mk(axis == 0 ? sign : 0.0f, axis == 1 ? sign : 0.0f, axis == 2 ? sign : 0.0f);
}
struct LightPoint
{
    V position, normal;
    uint32_t block, voxel, face;
    float pdf;
    uint32_t group, entry;
    uint32_t record;          // the index entry's dense_id: the brick's record in the brick store
    uint32_t voxel_in_group;  // x | y << 10 | z << 20
};
// LightTreeSample (caelix_light_tree.hlsl:749-767): LightTreeSampleBrick (:647-670), then
// LightBrickSamplePoint (:672-747) in the group's local frame, the point moved back to the world.
bool LightTreeSample(const Lights& L, V x, V n, uint32_t& rng, LightPoint& p)
{
    p = LightPoint{Mk(0, 0, 0), Mk(0, 0, 0), 0u, 0u, 0u, 0.0f, kNoneU, kNoneU, kNoneU, 0u};
    float pdf_brick = 1.0f;
    uint32_t tree;
    const uint32_t leaf = ChooseLeaf(L, x, n, rng, pdf_brick, tree);
    if (leaf == kNoneU)
    {
        return false;
    }
    const LightBvhLeaf& lf = L.leaves[tree][leaf];
    g_demand[g_kind] += 8;
    p.group = lf.group;
    p.entry = lf.entry;
    const uint32_t* block = GroupBlock(L, lf.group);
    if (block == nullptr)
    {
        return false;
    }
    const Rec record = LoadRec(block, lf.entry);
    p.record = IndexEntryOf(block, lf.entry).dense_id;
    Frame frame;
    const V origin = GroupFrame(L, lf.group, frame);
    float w[6];
    float sum;
    FamilyWeights(record, ToGroup(frame, x, origin), w, sum);
    if (sum <= 0.0f)
    {
        return false;
    }
    const float u = LightRandomFloat(rng) * sum;
    uint32_t f = kNoneU;
    uint32_t last = 0;
    float cumulative = 0.0f;
    for (uint32_t g = 0; g < 6; ++g)
    {
        if (w[g] > 0.0f)
        {
            last = g;
            cumulative += w[g];
            if (f == kNoneU && u < cumulative)
            {
                f = g;
            }
        }
    }
    f = f == kNoneU ? last : f;
    uint32_t first = 0;
    for (uint32_t h = 0; h < f; ++h)
    {
        first += FamilyCount(record, h);
    }
    const uint32_t count = FamilyCount(record, f);
    if (count == 0 || w[f] <= 0.0f)
    {
        return false;
    }
    const uint32_t index = std::min(uint32_t(LightRandomFloat(rng) * float(count)), count - 1u);
    const uint32_t entry = record.entries[first + index];
    g_demand[g_kind] += 4;
    const uint32_t vi = entry & 0x1FFu;
    const uint32_t mask = (1u << L.group_shift) - 1u;
    const uint32_t v[3] = {(record.pos & mask) * 8 + (vi & 7), ((record.pos >> L.group_shift) & mask) * 8 + ((vi >> 3) & 7),
                           (record.pos >> (2 * L.group_shift)) * 8 + (vi >> 6)};
    V local = Add(origin, Mk(float(v[0]), float(v[1]), float(v[2])));
    const float a = LightRandomFloat(rng);
    const float c = LightRandomFloat(rng);
    const int axis = int(f >> 1);
    SetComp(local, axis, Comp(local, axis) + ((f & 1) != 0 ? 0.0f : 1.0f));
    SetComp(local, (axis + 1) % 3, Comp(local, (axis + 1) % 3) + a);
    SetComp(local, (axis + 2) % 3, Comp(local, (axis + 2) % 3) + c);
    p.position = Add(Rotate(frame, local), frame.t);
    p.normal = Rotate(frame, FaceDirection(f));
    p.block = entry >> 16;
    p.voxel = vi;
    p.face = f;
    p.voxel_in_group = v[0] | (v[1] << 10) | (v[2] << 20);
    p.pdf = ((w[f] / sum) / float(count)) * pdf_brick;
    return true;
}
// A brick of the tree by (group, index entry) and the row generation it was named at (LightBrickRef).
struct BrickRef
{
    uint32_t group, entry, generation;
};
// LightTreeFindLightFace (caelix_light_tree.hlsl:853-932): face `face` of the group-local voxel v of
// `group` with the raw block value is a light face of the tree when its material code is a light, the
// group has a block with the voxel's brick in its index and the voxel is in that face family of the
// record. Then brick, leaf (LtBrickLeaf; kNoneU when no draw reaches the brick) and in_brick =
// P(face | brick, x) / family count for the receiver x.
bool FindLightFace(const Lights& L, V x, uint32_t group, const uint32_t v[3], uint32_t face, uint32_t block_value,
                   BrickRef& brick, uint32_t& leaf, float& in_brick)
{
    brick = BrickRef{group, kNoneU, 0u};
    leaf = kNoneU;
    in_brick = 0.0f;
    const uint32_t code = MaterialCode(block_value & 0xFFFFu, L.encoding);
    g_demand[g_kind] += 4;
    if (((L.light_codes[code >> 5] >> (code & 31u)) & 1u) == 0u || face > 5u)
    {
        return false;
    }
    const uint32_t* block = GroupBlock(L, group);
    if (block == nullptr)
    {
        return false;
    }
    const uint32_t entry = FindIndexEntry(block, block[0], Morton3(v[0] >> 3, v[1] >> 3, v[2] >> 3), L.group_shift);
    if (entry == kNoneU)
    {
        return false;
    }
    brick.entry = entry;
    brick.generation = L.rows[group].generation;
    const Rec rec = LoadRec(block, entry);
    Frame frame;
    const V origin = GroupFrame(L, group, frame);
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
    bool hit = false;
    for (uint32_t e = 0; e < count; ++e)  // linear scan: the entries are sorted, any search finds it
    {
        g_demand[g_kind] += 4;
        if ((rec.entries[first + e] & 0x1FFu) == vi)
        {
            hit = true;
            break;
        }
    }
    if (!hit)
    {
        return false;
    }
    in_brick = (w[face] / sum) / float(count);
    leaf = BrickLeaf(L, block, group, entry);
    return true;
}
// LightTreePointPdf (caelix_light_tree.hlsl:936-949): the area pdf of the draw of that face for the
// receiver (x, n); `found` = a light face whose brick has a leaf (LampHitWeight's condition,
// caelix_path_trace.hlsl:438-443).
float LightPointPdf(const Lights& L, V x, V n, uint32_t group, const uint32_t v[3], uint32_t face, uint32_t block_value,
                    bool& found)
{
    BrickRef brick;
    uint32_t leaf;
    float in_brick;
    found = FindLightFace(L, x, group, v, face, block_value, brick, leaf, in_brick) && leaf != kNoneU;
    if (!found)
    {
        return 0.0f;
    }
    return LeafProbability(L, leaf >> 31, leaf & 0x7FFFFFFFu, x, n) * in_brick;
}
// LightBrickRefValid (caelix_light_tree.hlsl:640-646): the row has a block, of the reference's
// generation, and the entry is below its record count.
bool BrickRefValid(const Lights& L, const BrickRef& b)
{
    const GroupLightRow& row = L.rows[b.group];
    return row.block_units != 0u && row.generation == b.generation && b.entry < L.heap[uint64_t(row.block_offset) * 4];
}
// LightTreeBrickPdf / LightTreeBrickPdfBox (:769-777, :822-831) over LtBrickProbabilityR (:575-592).
float BrickPdfR(const Lights& L, const Recv& r, const BrickRef& brick)
{
    if (!BrickRefValid(L, brick))
    {
        return 0.0f;
    }
    const uint32_t leaf = BrickLeaf(L, GroupBlock(L, brick.group), brick.group, brick.entry);
    return leaf == kNoneU ? 0.0f : LeafProbabilityR(L, leaf >> 31, leaf & 0x7FFFFFFFu, r);
}
// LightTreeSampleBrickBox (:787-818): the box-receiver draw of a brick; leaf = tree << 31 | index.
bool SampleBrickBox(const Lights& L, const Recv& r, uint32_t& rng, BrickRef& brick, float& pdf_brick, uint32_t& leaf,
                    float& importance)
{
    brick = BrickRef{kNoneU, kNoneU, 0u};
    pdf_brick = 1.0f;
    leaf = kNoneU;
    uint32_t tree;
    const uint32_t index = ChooseLeafR(L, r, rng, pdf_brick, tree, importance);
    if (index == kNoneU)
    {
        importance = 0.0f;
        return false;
    }
    const LightBvhLeaf& lf = L.leaves[tree][index];
    brick = BrickRef{lf.group, lf.entry, L.rows[lf.group].generation};
    leaf = (tree << 31) | index;
    return true;
}
// LightTreeLeafImportanceBox (:837-850): the box importance of the brick's own leaf summary.
float LeafImportanceBox(const Lights& L, const BrickRef& brick, const Recv& r)
{
    const uint32_t* block = GroupBlock(L, brick.group);
    const uint32_t leaf = block != nullptr ? BrickLeaf(L, block, brick.group, brick.entry) : kNoneU;
    if (leaf == kNoneU)
    {
        return 0.0f;
    }
    const LightBvhLeaf& lf = L.leaves[leaf >> 31][leaf & 0x7FFFFFFFu];
    return ImportanceR(r, Mk(lf.summary.lo[0], lf.summary.lo[1], lf.summary.lo[2]),
                       Mk(lf.summary.hi[0], lf.summary.hi[1], lf.summary.hi[2]), lf.summary.face);
}
// A hit of the world traversal (identity entity frames): the group of the brick record, the voxel less
// the group origin.
float LightFacePdf(const CxScene& s, const Lights& L, V x, V n, uint32_t record, const int32_t voxel[3], uint32_t face,
                   uint32_t block, bool& found)
{
    const uint32_t group = s.group_of_brick[record];
    const GroupInfo& info = L.group_infos[group];
    const int32_t edge = int32_t(8u << L.group_shift);
    const uint32_t v[3] = {uint32_t(voxel[0] - info.coord[0] * edge), uint32_t(voxel[1] - info.coord[1] * edge),
                           uint32_t(voxel[2] - info.coord[2] * edge)};
    g_demand[g_kind] += 4 + 16;
    return LightPointPdf(L, x, n, group, v, face, block, found);
}
float MisWeight(float p_area, V n_r, V w, float d, float cos_l, bool nee_side)
{
    const float p_sa = cos_l > 0.0f ? (p_area * (d * d)) / cos_l : 0.0f;
    const float p_bsdf = std::fmax(0.0f, Dot(n_r, w)) * (1.0f / 3.1415926535f);
    if (!(cos_l > 0.0f) || !(p_sa <= 3.0e38f))
    {
        return nee_side ? 1.0f : 0.0f;
    }
    const float sum = p_sa + p_bsdf;
    if (!(sum > 0.0f))
    {
        return nee_side ? 0.0f : 1.0f;
    }
    return (nee_side ? p_sa : p_bsdf) / sum;
}

// ---- path tracer ----
V MatAlbedo(const CxScene& s, uint32_t b)
{
    const float* m = s.materials + uint64_t(b) * 8;
    return Mk(m[0], m[1], m[2]);
}
V MatEmission(const CxScene& s, uint32_t b)
{
    const float* m = s.materials + uint64_t(b) * 8 + 4;
    return Mul(Mk(m[0], m[1], m[2]), s.emission_scale);
}
float Luminance(V c)
{
    return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f;
}
uint32_t LowBit(uint32_t v)
{
    return uint32_t(__builtin_ctz(v));
}
struct Counters
{
    uint64_t rays = 0, shadow = 0, cells = 0, probes = 0, steps = 0;
};
WorldHit Trace(const CxScene& s, V o, V d, float t_max, Counters& c)
{
    const WorldHit h = TraceWorld(s, o, d, t_max);
    c.cells += h.cells;
    c.probes += h.probes;
    c.steps += h.steps;
    return h;
}
struct Shade
{
    V albedo, emission, origin, direction;
};
Shade ShadeHit(const CxScene& s, uint32_t block, V hp, V normal, uint32_t& rng)
{
    Shade sh;
    g_demand[g_kind] += 32;
    sh.emission = MatEmission(s, block);
    sh.albedo = MatAlbedo(s, block);
    (void)RandomFloat01(rng);
    V d = SafeNormalize(Add(normal, RandomUnitVector(rng)));
    d = Dot(d, d) > 0.0f ? d : normal;
    sh.origin = Add(hp, Mul(normal, kPushOff));
    sh.direction = d;
    return sh;
}
float BrickScale(const CxScene& s, uint32_t record)
{
    return s.brick_emission_scale != nullptr ? s.brick_emission_scale[record] : 1.0f;
}

// exp(x): the Cephes expf polynomial (S. Moshier, Cephes Math Library) with the two-part ln 2 and
// 2^n from exponent bits, in the kernels' operation order; 0 below -87, +inf above 88.
float ExpCephes(float x)
{
    if (!(x >= -87.0f))
    {
        return 0.0f;
    }
    if (x > 88.0f)
    {
        return INFINITY;
    }
    const float fx = std::floor(x * 1.44269504088896341f + 0.5f);
    float r = x - fx * 0.693359375f;
    r = r - fx * -2.12194440e-4f;
    const float z = r * r;
    const float p = ((((1.9875691500e-4f * r + 1.3981999507e-3f) * r + 8.3334519073e-3f) * r + 4.1665795894e-2f) * r +
                     1.6666665459e-1f) * r + 5.0000001201e-1f;
    const float e = p * z + r + 1.0f;
    const uint32_t bits = uint32_t(int(fx) + 127) << 23;
    float scale;
    std::memcpy(&scale, &bits, 4);
    return e * scale;
}

// TraverseSegment (shaders/render/caelix_shading.hlsl:150-175) in the air with fog: the segment of
// length t keeps exp(-density t) of the throughput and adds fog (1 - exp(-density t)) times the
// throughput before it.
void FogSegment(V fog, float density, float t, V& throughput, V& radiance)
{
    const float transmission = ExpCephes((-density) * t);
    radiance = Add(radiance, Mul(MulV(throughput, fog), 1.0f - transmission));
    throughput = Mul(throughput, transmission);
}

// The procedural sky and the sun of CX_TRACE_SKY (the frame constants sky_zenith/horizon/ground,
// sun_dir, sun_radiance, sun_cos_radius, sun_solid_angle of shaders/render/caelix_frame.hlsl).
struct Sky
{
    V fog;                     // the fog of the air (fog_color rgb, density; density 0: none)
    float fog_density = 0.0f;
    bool glass = false;        // the full material model (materials_ext set, or a delta chain)
    uint32_t delta_budget = 0; // bounce_count_transparent
    bool transmissive = false; // kRenderFlagShadowTransmission at the primary vertex
    bool checkerboard = false; // kRenderFlagDeltaCheckerboard: the chain's first interface by pixel parity
    float nee_q = 1.0f;        // bounce_nee_probability (0 is taken as 1)
    bool subsurface = false;   // kRenderFlagSubsurface with a subsurface table: the walk at SSS vertices
    bool on = false;
    bool sun = false;  // the sun's NEE: `on`, kRenderFlagSunNee, a sun radiance above 0 (caelix_path_trace.hlsl:742)
    V sun_dir, sun_radiance, zenith, horizon, ground;
    float sun_solid_angle = 0.0f, sun_cos_radius = 1.0f;
    // the host's cubemap (CxSkyCubemap; null: the procedural sky) and its irradiance SH
    const float* cube = nullptr;
    uint32_t cube_size = 0, cube_mips = 0;
    float sh[9][3] = {};
};
Sky SkyOf(const CxTraceParams& p)
{
    Sky k;
    k.sun_dir = k.sun_radiance = k.zenith = k.horizon = k.ground = Mk(0, 0, 0);
    k.fog = Mk(p.fog_color[0], p.fog_color[1], p.fog_color[2]);
    k.fog_density = p.fog_density;
    k.delta_budget = p.bounces_transparent;
    k.transmissive = (p.flags & CX_TRACE_SHADOW_TRANSMISSION) != 0u;
    k.checkerboard = (p.flags & CX_TRACE_DELTA_CHECKERBOARD) != 0u;
    k.nee_q = p.bounce_nee_probability > 0.0f ? p.bounce_nee_probability : 1.0f;
    if ((p.flags & CX_TRACE_SKY) == 0u)
    {
        return k;
    }
    k.on = true;
    k.sun = (p.flags & CX_TRACE_NO_SUN_NEE) == 0u &&
            (p.sun_radiance[0] > 0.0f || p.sun_radiance[1] > 0.0f || p.sun_radiance[2] > 0.0f);
    k.sun_dir = Mk(p.sun_dir[0], p.sun_dir[1], p.sun_dir[2]);
    k.sun_radiance = Mk(p.sun_radiance[0], p.sun_radiance[1], p.sun_radiance[2]);
    k.zenith = Mk(p.sky_zenith[0], p.sky_zenith[1], p.sky_zenith[2]);
    k.horizon = Mk(p.sky_horizon[0], p.sky_horizon[1], p.sky_horizon[2]);
    k.ground = Mk(p.sky_ground[0], p.sky_ground[1], p.sky_ground[2]);
    k.sun_solid_angle = p.sun_solid_angle;
    k.sun_cos_radius = p.sun_cos_radius;
    if (p.sky_cubemap != nullptr)
    {
        k.cube = p.sky_cubemap->texels;
        k.cube_size = p.sky_cubemap->size;
        k.cube_mips = p.sky_cubemap->mip_levels;
    }
    return k;
}
float Saturate(float x)
{
    return std::fmin(std::fmax(x, 0.0f), 1.0f);
}
// lerp(a, b, t) = a + t (b - a), per component.
V Lerp(V a, V b, float t)
{
    return Add(a, Mul(Sub(b, a), t));
}
// ProceduralSky (shaders/render/caelix_shading.hlsl:364-371).
V ProceduralSky(const Sky& k, V d)
{
    if (d.y >= 0.0f)
    {
        return Lerp(k.horizon, k.zenith, std::sqrt(Saturate(d.y)));
    }
    return Lerp(k.horizon, k.ground, std::sqrt(Saturate(-d.y)));
}
// SunDisk (caelix_shading.hlsl:394-398).
V SunDisk(const Sky& k, V d)
{
    return Dot(d, k.sun_dir) > k.sun_cos_radius ? k.sun_radiance : Mk(0, 0, 0);
}
// The cubemap as a linear mipmapped sampler reads it (SampleLevel, caelix_shading.hlsl:375-386): the
// face of the largest-magnitude axis (x, y, z on ties), face coordinates s, t = (sc / |ma| + 1) / 2 for
// +X (-z, -y), -X (z, -y), +Y (x, z), -Y (x, -z), +Z (x, -y), -Z (-x, -y) (the inverse of the reference
// test's CubeTexelDirection); bilinear at texel (s size - 0.5, t size - 0.5) inside the face, the
// neighbours clamped to its edge; linear between the two mips around the clamped lod.
uint32_t CubeFaceOf(V d, float& s, float& t)
{
    const float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    uint32_t face;
    float sc, tc, ma;
    if (ax >= ay && ax >= az)
    {
        face = d.x >= 0.0f ? 0u : 1u;
        sc = d.x >= 0.0f ? -d.z : d.z;
        tc = -d.y;
        ma = ax;
    }
    else if (ay >= az)
    {
        face = d.y >= 0.0f ? 2u : 3u;
        sc = d.x;
        tc = d.y >= 0.0f ? d.z : -d.z;
        ma = ay;
    }
    else
    {
        face = d.z >= 0.0f ? 4u : 5u;
        sc = d.z >= 0.0f ? d.x : -d.x;
        tc = -d.y;
        ma = az;
    }
    s = (sc / ma + 1.0f) * 0.5f;
    t = (tc / ma + 1.0f) * 0.5f;
    return face;
}
V CubeTexel(const Sky& k, uint32_t m, uint32_t face, int x, int y)
{
    uint64_t offset = 0;
    for (uint32_t i = 0; i < m; ++i)
    {
        const uint64_t e = std::max(k.cube_size >> i, 1u);
        offset += 6 * e * e;
    }
    const uint64_t e = std::max(k.cube_size >> m, 1u);
    const float* p = k.cube + (offset + uint64_t(face) * e * e + uint64_t(y) * e + uint64_t(x)) * 4;
    return Mk(p[0], p[1], p[2]);
}
V CubeBilinear(const Sky& k, uint32_t m, uint32_t face, float s, float t)
{
    const int e = int(std::max(k.cube_size >> m, 1u));
    const float u = s * float(e) - 0.5f;
    const float v = t * float(e) - 0.5f;
    const float fu = std::floor(u), fv = std::floor(v);
    const float a = u - fu, b = v - fv;
    auto clampi = [&](int i) { return std::min(std::max(i, 0), e - 1); };
    const int x0 = clampi(int(fu)), x1 = clampi(int(fu) + 1), y0 = clampi(int(fv)), y1 = clampi(int(fv) + 1);
    const V top = Lerp(CubeTexel(k, m, face, x0, y0), CubeTexel(k, m, face, x1, y0), a);
    const V bottom = Lerp(CubeTexel(k, m, face, x0, y1), CubeTexel(k, m, face, x1, y1), a);
    return Lerp(top, bottom, b);
}
V CubeSample(const Sky& k, V d, float lod)
{
    float s, t;
    const uint32_t face = CubeFaceOf(d, s, t);
    const float l = std::fmin(std::fmax(lod, 0.0f), float(k.cube_mips - 1));
    const uint32_t m0 = uint32_t(std::floor(l));
    const uint32_t m1 = std::min(m0 + 1u, k.cube_mips - 1u);
    const float f = l - float(m0);
    const V c0 = CubeBilinear(k, m0, face, s, t);
    return f == 0.0f ? c0 : Lerp(c0, CubeBilinear(k, m1, face, s, t), f);
}
// CaelixShBasis (caelix_shading.hlsl:346-359).
void ShBasis(V d, float basis[9])
{
    basis[0] = 0.282095f;
    basis[1] = 0.488603f * d.y;
    basis[2] = 0.488603f * d.z;
    basis[3] = 0.488603f * d.x;
    basis[4] = 1.092548f * d.x * d.y;
    basis[5] = 1.092548f * d.y * d.z;
    basis[6] = 0.315392f * (3.0f * d.z * d.z - 1.0f);
    basis[7] = 1.092548f * d.x * d.z;
    basis[8] = 0.546274f * (d.x * d.x - d.y * d.y);
}
// The sky SH projection (shaders/render/caelix_sky_sh.hlsl): 6 x 32 x 32 texel-centre directions of the
// reference's CubeTexelDirection at mip `lod`, solid angles (2 / 32)^2 / (1 + s^2 + t^2)^1.5,
// renormalized by 4 pi / their sum, bands times pi, 2 pi / 3, pi / 4. The sums run as the kernel adds
// them: per thread of 256 over its directions in order, then a tree over the threads (strides
// 128 .. 1).
void ProjectSh(Sky& k, float lod)
{
    constexpr uint32_t kThreads = 256, kFace = 32;
    std::vector<float> sums(kThreads * 28, 0.0f);  // per thread: 9 x 3 sums, the weight
    const float texel = 2.0f / float(kFace);
    for (uint32_t th = 0; th < kThreads; ++th)
    {
        float* acc = &sums[th * 28];
        for (uint32_t i = th; i < 6 * kFace * kFace; i += kThreads)
        {
            const uint32_t face = i / (kFace * kFace);
            const uint32_t x = i % kFace, y = (i / kFace) % kFace;
            const float s = (float(x) + 0.5f) * texel - 1.0f;
            const float t = (float(y) + 0.5f) * texel - 1.0f;
            V v;
            switch (face)
            {
                case 0: v = Mk(1.0f, -t, -s); break;
                case 1: v = Mk(-1.0f, -t, s); break;
                case 2: v = Mk(s, 1.0f, t); break;
                case 3: v = Mk(s, -1.0f, -t); break;
                case 4: v = Mk(s, -t, 1.0f); break;
                default: v = Mk(-s, -t, -1.0f); break;
            }
            const V dir = Mul(v, 1.0f / std::sqrt(Dot(v, v)));
            const float r2 = 1.0f + s * s + t * t;
            const float weight = (texel * texel) / (r2 * std::sqrt(r2));
            const V radiance = CubeSample(k, dir, lod);
            float basis[9];
            ShBasis(dir, basis);
            for (int b = 0; b < 9; ++b)
            {
                const float w = basis[b] * weight;
                acc[3 * b] += radiance.x * w;
                acc[3 * b + 1] += radiance.y * w;
                acc[3 * b + 2] += radiance.z * w;
            }
            acc[27] += weight;
        }
    }
    auto tree = [&](int column) {
        std::vector<float> v(kThreads);
        for (uint32_t th = 0; th < kThreads; ++th)
        {
            v[th] = sums[th * 28 + column];
        }
        for (uint32_t stride = kThreads / 2; stride > 0; stride >>= 1)
        {
            for (uint32_t th = 0; th < stride; ++th)
            {
                v[th] = v[th] + v[th + stride];
            }
        }
        return v[0];
    };
    const float normalization = (4.0f * kPi) / tree(27);
    for (int b = 0; b < 9; ++b)
    {
        const float band = b == 0 ? kPi : (b < 4 ? 2.0f * kPi / 3.0f : 0.25f * kPi);
        for (int ch = 0; ch < 3; ++ch)
        {
            k.sh[b][ch] = tree(3 * b + ch) * (band * normalization);
        }
    }
}
// SkyIrradianceOverPi (caelix_shading.hlsl:406-419).
V SkyIrradianceOverPi(const Sky& k, V n)
{
    float basis[9];
    ShBasis(n, basis);
    V sum = Mk(0, 0, 0);
    for (int i = 0; i < 9; ++i)
    {
        sum = Add(sum, Mul(Mk(k.sh[i][0], k.sh[i][1], k.sh[i][2]), basis[i]));
    }
    return Mul(Mk(std::fmax(sum.x, 0.0f), std::fmax(sum.y, 0.0f), std::fmax(sum.z, 0.0f)), 1.0f / kPi);
}
// SkyRadianceNoSunLod (caelix_shading.hlsl:375-386).
V SkyNoSun(const Sky& k, V d, float lod)
{
    return k.cube != nullptr ? CubeSample(k, d, lod) : ProceduralSky(k, d);
}
// SkyRadiance (caelix_shading.hlsl:400-404), the primary miss (caelix_composite.hlsl:130-133).
V SkyRadiance(const Sky& k, V d)
{
    return Add(SkyNoSun(k, d, 0.0f), SunDisk(k, d));
}
// SkyDiffuseMiss (caelix_shading.hlsl:428-437): the cubemap's SH irradiance / pi at the normal, else the
// procedural sky in d; the sun disk when the segment sees it.
V SkyDiffuseMiss(const Sky& k, V n, V d, bool sees_sun)
{
    return Add(k.cube != nullptr ? SkyIrradianceOverPi(k, n) : ProceduralSky(k, d), sees_sun ? SunDisk(k, d) : Mk(0, 0, 0));
}
// SkySpecularMiss (caelix_shading.hlsl:443-447).
V SkySpecularMiss(const Sky& k, V d, float roughness, bool sees_sun)
{
    return Add(SkyNoSun(k, d, roughness * float(k.cube_mips > 0 ? k.cube_mips - 1 : 0)),
               sees_sun ? SunDisk(k, d) : Mk(0, 0, 0));
}
// The opaque path's bounce miss: a bounce segment leaves a vertex that ran the NEE, so
// segment_sees_sun = !sun (caelix_path_trace.hlsl:809, 903).
V SkyMiss(const Sky& k, V n, V d)
{
    return SkyDiffuseMiss(k, n, d, !k.sun);
}
// SampleSunDirection (caelix_shading.hlsl:467-485): uniform in the cone of solid angle
// sun_solid_angle around sun_dir in the branchless basis of Duff et al. 2017.
V SampleSunDirection(const Sky& k, uint32_t& rng)
{
    const float one_minus_cos_r = k.sun_solid_angle * (1.0f / kTwoPi);
    const float a = RandomFloat01(rng) * one_minus_cos_r;
    const float phi = RandomFloat01(rng) * kTwoPi;
    const float cos_theta = 1.0f - a;
    const float sin_theta = std::sqrt(std::fmax(a * (2.0f - a), 0.0f));
    const V w = k.sun_dir;
    const float sign_z = w.z >= 0.0f ? 1.0f : -1.0f;
    const float p = -1.0f / (sign_z + w.z);
    const float q = w.x * w.y * p;
    const V t = Mk(1.0f + sign_z * w.x * w.x * p, sign_z * q, -sign_z * w.x);
    const V b = Mk(q, sign_z + w.y * w.y * p, -w.y);
    float sp, cp;
    SinCos(phi, sp, cp);
    return SafeNormalize(Add(Add(Mul(t, cp * sin_theta), Mul(b, sp * sin_theta)), Mul(w, cos_theta)));
}
// SunCandidate (caelix_path_trace.hlsl:276-290).
V SunCandidate(const Sky& k, V normal, uint32_t& rng, V& direction)
{
    direction = SampleSunDirection(k, rng);
    const float cos_l = Dot(normal, direction);
    if (!(cos_l > 0.0f))
    {
        return Mk(0, 0, 0);
    }
    return Mul(k.sun_radiance, k.sun_solid_angle * cos_l * (1.0f / kPi));
}

// ---- the full material model (shaders/render/caelix_shading.hlsl:116-307) ----------------------
struct Material
{
    V albedo, emission;
    float smoothness, metallic, ior, extinction;
};
// CaelixGetMaterial of a table index (caelix_material.hlsl:84-100); without the extended table ior
// 1.5 and extinction 0.
Material GetMaterial(const CxScene& s, uint32_t code)
{
    const float* m = s.materials + uint64_t(code) * 8;
    Material r;
    r.albedo = Mk(m[0], m[1], m[2]);
    r.smoothness = m[3];
    r.emission = Mk(m[4], m[5], m[6]);
    r.metallic = m[7];
    r.ior = 1.5f;
    r.extinction = 0.0f;
    if (s.materials_ext != nullptr)
    {
        r.ior = s.materials_ext[uint64_t(code) * 4];
        r.extinction = s.materials_ext[uint64_t(code) * 4 + 1];
    }
    return r;
}
V Reflect(V i, V n)
{
    return Sub(i, Mul(n, 2.0f * Dot(n, i)));
}
V Refract(V i, V n, float eta)
{
    const float c = Dot(n, i);
    const float k = 1.0f - eta * eta * (1.0f - c * c);
    if (k < 0.0f)
    {
        return Mk(0, 0, 0);
    }
    return Sub(Mul(i, eta), Mul(n, eta * c + std::sqrt(k)));
}
float Schlick(float r0, float cos_x)
{
    const float x = 1.0f - cos_x;
    const float xx = x * x;
    return r0 + (1.0f - r0) * xx * xx * x;
}
// FresnelReflectAmountOpaque / FresnelReflectAmountTransparent (caelix_shading.hlsl:116-148).
float FresnelOpaque(float n1, float n2, V incident, V normal)
{
    float r0 = (n1 - n2) / (n1 + n2);
    r0 *= r0;
    return Schlick(r0, -Dot(normal, incident));
}
float FresnelTransparent(float n1, float n2, V incident, V normal)
{
    float r0 = (n1 - n2) / (n1 + n2);
    r0 *= r0;
    float cos_x = -Dot(normal, incident);
    if (n1 > n2)
    {
        const float n = n1 / n2;
        const float sin_t2 = n * n * (1.0f - cos_x * cos_x);
        if (sin_t2 >= 1.0f)
        {
            return 1.0f;
        }
        cos_x = std::sqrt(1.0f - sin_t2);
    }
    return Schlick(r0, cos_x);
}
struct Interface
{
    V normal, direction;
    bool refracted;
    uint32_t medium;
    float weight;
    bool split_applied;
};
// CrossInterface (caelix_shading.hlsl:203-243): split < 0 random, 0 / 1 the checkerboard's reflect /
// refract with weight 2 F / 2 (1 - F).
Interface CrossInterface(const CxScene& s, uint32_t source, uint32_t destination, float destination_ior, V dir, V normal,
                         uint32_t& rng, int split = -1)
{
    Interface it;
    it.normal = SafeNormalize(normal);
    if (Dot(dir, it.normal) > 0.0f)
    {
        it.normal = Mk(-it.normal.x, -it.normal.y, -it.normal.z);
    }
    const float source_ior = source != 0u ? GetMaterial(s, source).ior : 1.0f;
    const float target_ior = destination != 0u ? destination_ior : 1.0f;
    const float eta = source_ior / target_ior;
    const V reflection = Reflect(dir, it.normal);
    const V refraction = Refract(dir, it.normal, eta);
    const bool can_refract = Dot(refraction, refraction) > 0.000001f;
    const float fresnel = can_refract ? FresnelTransparent(source_ior, target_ior, dir, it.normal) : 1.0f;
    it.refracted = false;
    it.weight = 1.0f;
    it.split_applied = false;
    if (can_refract)
    {
        if (split < 0)
        {
            it.refracted = RandomFloat01(rng) >= fresnel;
        }
        else
        {
            it.refracted = split == 1;
            it.weight = 2.0f * (it.refracted ? 1.0f - fresnel : fresnel);
            it.split_applied = true;
        }
    }
    it.direction = SafeNormalize(it.refracted ? refraction : reflection);
    it.medium = it.refracted ? destination : source;
    return it;
}
V LerpV(V a, V b, float t)
{
    return Add(a, Mul(Sub(b, a), t));
}
struct ShadeFull
{
    V albedo, emission, origin, direction;
    bool specular;
    float roughness;  // the lobe's: 1 - smoothness, 0 at a transparent block (caelix_shading.hlsl:252)
};
// ShadeVoxelHit (caelix_shading.hlsl:267-307); lerp(a, b, t) = a + t (b - a).
ShadeFull ShadeVoxel(const CxScene& s, uint32_t block, V hp, V dir, V normal, uint32_t& rng, uint32_t& medium)
{
    g_demand[g_kind] += 32;
    const uint32_t code = MaterialCode(block, s.encoding);
    const Material m = GetMaterial(s, code);
    ShadeFull sh;
    sh.emission = Mul(m.emission, s.emission_scale);
    if (!Opaque(block, s.encoding))
    {
        const Interface it = CrossInterface(s, medium, code, m.ior, dir, normal, rng);
        sh.albedo = Mk(1, 1, 1);
        sh.origin = Add(hp, Mul(it.normal, it.refracted ? -kPushOff : kPushOff));
        sh.direction = it.direction;
        sh.specular = true;
        sh.roughness = 0.0f;
        medium = it.medium;
        return sh;
    }
    sh.albedo = m.albedo;
    const float fresnel = FresnelOpaque(1.0f, m.ior, dir, normal);
    const float chance = m.metallic + fresnel * m.smoothness * (1.0f - m.metallic);
    const float do_specular = RandomFloat01(rng) < chance ? 1.0f : 0.0f;
    V diffuse = SafeNormalize(Add(normal, RandomUnitVector(rng)));
    diffuse = Dot(diffuse, diffuse) > 0.0f ? diffuse : normal;
    const V specular = SafeNormalize(LerpV(diffuse, Reflect(dir, normal), m.smoothness));
    sh.origin = Add(hp, Mul(normal, kPushOff));
    sh.direction = LerpV(diffuse, specular, do_specular);
    sh.specular = do_specular > 0.5f;
    sh.roughness = 1.0f - m.smoothness;
    return sh;
}
// ---- subsurface scattering (shaders/render/caelix_subsurface.hlsl, docs/SUBSURFACE.md) ---------------
// Pinned by the reference's furnace and lifecycle checks, the HG cosine distribution, a one-step closed
// form over the start voxel and the entry jitter, and the branch weights (tests/test_sss_oracle.py).
enum SssState : uint32_t
{
    kSssNone = 0,  // not an SSS vertex (or the last vertex)
    kSssBsdf = 1,  // the draw kept the surface BSDF
    kSssExit = 2,  // the walk left: the vertex is its exit
    kSssDead = 3,  // the walk reached the step cap
};

// SubsurfaceVoxel: opaque, its code's entry HG with p > 0 and steps > 0; albedo 1 - one_minus_albedo.
bool SssVoxel(const CxScene& s, uint32_t block, V& albedo)
{
    if (!Opaque(block, s.encoding))
    {
        return false;
    }
    const CxSubsurface& e = s.subsurface[MaterialCode(block, s.encoding)];
    if (e.model != CX_SSS_MODEL_HG || !(e.probability > 0.0f) || (e.steps & 0xFFu) == 0u)
    {
        return false;
    }
    albedo = Mk(1.0f - e.one_minus_albedo[0], 1.0f - e.one_minus_albedo[1], 1.0f - e.one_minus_albedo[2]);
    return true;
}

// SampleHenyeyGreenstein (caelix_subsurface.hlsl:62-85), the basis of OrthonormalBasis
// (caelix_shading.hlsl:454-461).
V SampleHg(V w, float g, uint32_t& rng)
{
    const float u1 = RandomFloat01(rng);
    const float u2 = RandomFloat01(rng);
    float ct;
    if (std::fabs(g) < 1e-3f)
    {
        ct = 1.0f - 2.0f * u1;
    }
    else
    {
        const float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * u1);
        ct = (1.0f + g * g - sq * sq) / (2.0f * g);
    }
    ct = std::fmin(std::fmax(ct, -1.0f), 1.0f);
    const float st = std::sqrt(std::fmax(1.0f - ct * ct, 0.0f));
    const float phi = kTwoPi * u2;
    const float sign_z = w.z >= 0.0f ? 1.0f : -1.0f;
    const float pz = -1.0f / (sign_z + w.z);
    const float q = w.x * w.y * pz;
    const V t = Mk(1.0f + sign_z * w.x * w.x * pz, sign_z * q, -sign_z * w.x);
    const V b = Mk(q, sign_z + w.y * w.y * pz, -w.y);
    float sp, cp;
    SinCos(phi, sp, cp);
    return Add(Add(Mul(t, cp * st), Mul(b, sp * st)), Mul(w, ct));
}

// The voxel under the walk: world voxel coordinates and its brick record (kNoneU: not resident, air).
struct Cursor
{
    int32_t v[3];
    uint32_t record;
};

int32_t BrickOf(int32_t v)
{
    return v >= 0 ? v / 8 : -((7 - v) / 8);
}

uint32_t CursorValue(const CxScene& s, const Cursor& c)
{
    if (c.record == kNoneU)
    {
        return 0u;
    }
    const int32_t lx = c.v[0] - 8 * BrickOf(c.v[0]);
    const int32_t ly = c.v[1] - 8 * BrickOf(c.v[1]);
    const int32_t lz = c.v[2] - 8 * BrickOf(c.v[2]);
    return s.payload[uint64_t(c.record) * 512 + uint32_t(lx + 8 * ly + 64 * lz)];
}

Cursor CursorMoved(const CxScene& s, Cursor c, const int32_t d[3])
{
    Cursor n = c;
    bool other = false;
    for (int a = 0; a < 3; ++a)
    {
        n.v[a] = c.v[a] + d[a];
        other = other || BrickOf(n.v[a]) != BrickOf(c.v[a]);
    }
    if (other)
    {
        uint32_t probes;
        n.record = HashGet(s, BrickOf(n.v[0]), BrickOf(n.v[1]), BrickOf(n.v[2]), probes);
    }
    return n;
}

// SubsurfaceWalk (caelix_subsurface.hlsl:95-190) from the hit (record, attrib, point hp) of a ray along
// `dir`: the entry jitter, then per step albedo, an HG direction and the dominant-axis neighbour (an SSS
// voxel is entered, another opaque voxel is a wall, anything else is the exit face). False at the cap.
bool SssWalk(const CxScene& s, uint32_t record, uint32_t attrib, V hp, V dir, float g, uint32_t steps, V entry_albedo,
             uint32_t& rng, V& origin, V& normal, V& direction, V& weight)
{
    origin = normal = direction = weight = Mk(0, 0, 0);
    const uint32_t cell = (attrib >> 6) & 0x1FFu;
    Cursor c;
    c.v[0] = s.keys[3 * uint64_t(record)] * 8 + int32_t(cell & 7u);
    c.v[1] = s.keys[3 * uint64_t(record) + 1] * 8 + int32_t((cell >> 3) & 7u);
    c.v[2] = s.keys[3 * uint64_t(record) + 2] * 8 + int32_t(cell >> 6);
    c.record = record;
    V a = entry_albedo;
    {
        const V fn = FaceDirection(LowBit(attrib & 0x3Fu));
        const float u1 = RandomFloat01(rng) - 0.5f;
        const float u2 = RandomFloat01(rng) - 0.5f;
        const V off = fn.x != 0.0f ? Mk(0, u1, u2) : (fn.y != 0.0f ? Mk(u1, 0, u2) : Mk(u1, u2, 0));
        const V target = Add(Sub(hp, Mul(fn, 0.5f)), off);
        int32_t d[3];
        for (int k = 0; k < 3; ++k)
        {
            const int32_t tv = int32_t(std::floor(Comp(target, k)));
            d[k] = Comp(fn, k) == 0.0f ? std::min(std::max(tv - c.v[k], -1), 1) : 0;
        }
        if (d[0] != 0 || d[1] != 0 || d[2] != 0)
        {
            const Cursor start = CursorMoved(s, c, d);
            V sa;
            if (SssVoxel(s, CursorValue(s, start), sa))
            {
                c = start;
                a = sa;
            }
        }
    }
    V w = Mk(1, 1, 1);
    V dd = dir;
    for (uint32_t i = 0; i < steps; ++i)
    {
        w = MulV(w, a);
        dd = SampleHg(dd, g, rng);
        const float ax = std::fabs(dd.x), ay = std::fabs(dd.y), az = std::fabs(dd.z);
        int32_t d[3] = {0, 0, 0};
        if (ax >= ay && ax >= az)
        {
            d[0] = dd.x >= 0.0f ? 1 : -1;
        }
        else if (ay >= az)
        {
            d[1] = dd.y >= 0.0f ? 1 : -1;
        }
        else
        {
            d[2] = dd.z >= 0.0f ? 1 : -1;
        }
        const Cursor next = CursorMoved(s, c, d);
        const uint32_t nb = CursorValue(s, next);
        if (Opaque(nb, s.encoding))
        {
            V na;
            if (SssVoxel(s, nb, na))
            {
                c = next;
                a = na;
            }
            continue;
        }
        const V n = Mk(float(d[0]), float(d[1]), float(d[2]));
        const float u1 = RandomFloat01(rng) - 0.5f;
        const float u2 = RandomFloat01(rng) - 0.5f;
        const V off = d[0] != 0 ? Mk(0, u1, u2) : (d[1] != 0 ? Mk(u1, 0, u2) : Mk(u1, u2, 0));
        const V exit_point = Add(Add(Add(Mk(float(c.v[0]), float(c.v[1]), float(c.v[2])), Mk(0.5f, 0.5f, 0.5f)),
                                     Mul(n, 0.5f)), off);
        normal = n;
        origin = Add(exit_point, Mul(normal, kPushOff));
        direction = SafeNormalize(Add(normal, RandomUnitVector(rng)));
        direction = Dot(direction, direction) > 0.0f ? direction : normal;
        weight = w;
        return true;
    }
    return false;
}

// SubsurfaceVertex (caelix_subsurface.hlsl:192-250).
uint32_t SssVertex(const CxScene& s, uint32_t record, uint32_t attrib, V hp, V dir, uint32_t& rng, V& albedo,
                   V& origin, V& direction, bool& specular, float& roughness, V& vertex_normal, V& guide, bool last)
{
    guide = albedo;
    const uint32_t block = (attrib >> 16) & 0xFFFFu;
    if (last || !Opaque(block, s.encoding))
    {
        return kSssNone;
    }
    const CxSubsurface& e = s.subsurface[MaterialCode(block, s.encoding)];
    const float p = e.probability;
    if (e.model != CX_SSS_MODEL_HG || !(p > 0.0f))
    {
        return kSssNone;
    }
    const uint32_t steps = e.steps & 0xFFu;
    if (steps == 0u)
    {
        return kSssNone;
    }
    const V entry_albedo = Mk(1.0f - e.one_minus_albedo[0], 1.0f - e.one_minus_albedo[1], 1.0f - e.one_minus_albedo[2]);
    guide = Lerp(albedo, entry_albedo, p);
    if (!(RandomFloat01(rng) < p))
    {
        return kSssBsdf;
    }
    const float g = std::fmin(std::fmax(e.anisotropy, -0.99f), 0.99f);
    V o, n, d, wgt;
    if (!SssWalk(s, record, attrib, hp, dir, g, steps, entry_albedo, rng, o, n, d, wgt))
    {
        albedo = Mk(0, 0, 0);
        return kSssDead;
    }
    albedo = wgt;
    origin = o;
    direction = d;
    specular = false;
    roughness = 1.0f;
    vertex_normal = n;
    return kSssExit;
}

// TraverseSegment (caelix_shading.hlsl:150-175): fog in the air, exp(-(1 - albedo) t extinction)
// per channel inside a medium.
void MediumSegment(const CxScene& s, const Sky& k, uint32_t medium, float t, V& throughput, V& radiance)
{
    if (medium == 0u)
    {
        if (k.fog_density > 0.0f)
        {
            FogSegment(k.fog, k.fog_density, t, throughput, radiance);
        }
        return;
    }
    const Material m = GetMaterial(s, medium);
    const V transmission = Mk(ExpCephes((-(1.0f - m.albedo.x)) * t * m.extinction),
                              ExpCephes((-(1.0f - m.albedo.y)) * t * m.extinction),
                              ExpCephes((-(1.0f - m.albedo.z)) * t * m.extinction));
    throughput = MulV(throughput, transmission);
}

// TraceShadowTransmissive (caelix_scene_trace.hlsl:248-266, 401-477) without subsurface.
V ShadowGlassTransmittance(const CxScene& s, uint32_t medium, float t)
{
    if (medium == 0u)
    {
        return Mk(1, 1, 1);
    }
    const Material m = GetMaterial(s, medium);
    return Mk(ExpCephes((-(1.0f - m.albedo.x)) * t * m.extinction), ExpCephes((-(1.0f - m.albedo.y)) * t * m.extinction),
              ExpCephes((-(1.0f - m.albedo.z)) * t * m.extinction));
}
bool ShadowOccluded(V t)
{
    return std::fmax(t.x, std::fmax(t.y, t.z)) < 1e-3f;
}
V ShadowTransmissive(const CxScene& s, V origin, V dir, float t_max, uint32_t medium, Counters& c)
{
    V transmittance = Mk(1, 1, 1);
    uint32_t cur = medium;
    float remaining = t_max;
    for (uint32_t hop = 0; hop < 8u; ++hop)
    {
        const WorldHit h = Trace(s, origin, dir, remaining, c);
        if (h.attrib == 0u)
        {
            if (cur != 0u && remaining < 1e29f)
            {
                transmittance = MulV(transmittance, ShadowGlassTransmittance(s, cur, remaining));
            }
            return ShadowOccluded(transmittance) ? Mk(0, 0, 0) : transmittance;
        }
        transmittance = MulV(transmittance, ShadowGlassTransmittance(s, cur, h.t));
        if (ShadowOccluded(transmittance))
        {
            return Mk(0, 0, 0);
        }
        remaining -= h.t;
        const uint32_t block = (h.attrib >> 16) & 0xFFFFu;
        if (Opaque(block, s.encoding))
        {
            return Mk(0, 0, 0);
        }
        cur = MaterialCode(block, s.encoding);
        const V offset = Mul(FaceDirection(LowBit(h.attrib & 0x3Fu)), -kPushOff);
        origin = Add(Add(origin, Mul(dir, h.t)), offset);
        remaining -= Dot(offset, dir);
    }
    return Mk(0, 0, 0);
}

// What the adjoint needs of one NEE: the radiance it adds without the light brick's emission scale,
// that brick, and whether it was added (unoccluded).
struct NeeTerm
{
    bool visible = false;
    V unscaled;
    uint32_t light = kNoneU;
};

// NextEventAtVertex (caelix_path_trace.hlsl:356-413): the lamp candidate and, with the sun's NEE,
// the sun candidate and the choice between them (one-sample RIS), then one shadow ray.
// The visibility hint of a primary vertex (kLightNeeHint, caelix_path_trace.hlsl:229-236, 338-344,
// 453-580): the word read, and which shadow ray the vertex traced (1 sun, 2 lamp) and whether it reached
// its light.
struct HintIo
{
    const CxLightHint* params = nullptr;  // the frame's hint images (null: no hint)
    uint32_t word = 0xFFFFu;  // kHintWordOff: the weights unchanged
    uint32_t traced = 0u;
    bool visible = false;
};
V NextEvent(const CxScene& s, const Lights& L, const Sky& k, V x_r, V n_r, uint32_t& light_rng,
            uint32_t& nee_rng, Counters& c, NeeTerm& term, bool transmissive = false, uint32_t medium = 0u,
            HintIo* hint = nullptr)
{
    // UnpackHint (:465-468); the choice weights times max(kHintFloor = 0.05, hint)
    const uint32_t word = hint != nullptr ? hint->word : 0xFFFFu;
    const float hint_sun = float(word & 0xFFu) * (1.0f / 255.0f);
    const float hint_lamp = float((word >> 8) & 0xFFu) * (1.0f / 255.0f);
    term = NeeTerm{};
    term.unscaled = Mk(0, 0, 0);
    const uint32_t seed = RandomU32(light_rng);
    V contribution = Mk(0, 0, 0);
    V direction = Mk(0, 0, 0);
    float t_max = 0.0f;
    uint32_t rng = seed;
    LightPoint p;
    if (LightTreeSample(L, x_r, n_r, rng, p) && p.pdf > 0.0f)
    {
        const V d = Sub(p.position, x_r);
        const float d2 = Dot(d, d);
        const float distance = std::sqrt(d2);
        if (distance >= 8.0f)
        {
            const V w = Mk(d.x / distance, d.y / distance, d.z / distance);
            const float cos_r = Dot(n_r, w);
            const float cos_l = -Dot(p.normal, w);
            if (cos_r > 0.0f && cos_l > 0.0f)
            {
                float scale = ((cos_r * (1.0f / kPi)) * (cos_l / d2)) / p.pdf;
                if (L.mis != 0u)
                {
                    scale *= MisWeight(p.pdf, n_r, w, distance, cos_l, true);
                }
                direction = w;
                t_max = distance - 2.0f * kPushOff;
                term.unscaled = Mul(MatEmission(s, p.block), scale);
                term.light = p.record;
                contribution = Mul(Mul(MatEmission(s, p.block), BrickScale(s, p.record)), scale);
                g_demand[g_kind] += 16;
            }
        }
    }
    const float w_lamp = Luminance(contribution) * std::fmax(0.05f, hint_lamp);
    const uint32_t sun_start = nee_rng;
    float w_sun = 0.0f;
    float u_choice = 0.0f;
    if (k.sun)
    {
        V sun_direction;
        w_sun = Luminance(SunCandidate(k, n_r, nee_rng, sun_direction)) * std::fmax(0.05f, hint_sun);
        u_choice = RandomFloat01(nee_rng);
    }
    const float w_sum = w_sun + w_lamp;
    V radiance = Mk(0, 0, 0);
    if (w_sum > 0.0f && u_choice * w_sum < w_sun)
    {
        // The sun candidate, drawn again from its saved state; its shadow ray runs to kRayTMax.
        c.shadow += 1;
        uint32_t replay = sun_start;
        const V sun = SunCandidate(k, n_r, replay, direction);
        term = NeeTerm{};
        term.unscaled = Mk(0, 0, 0);
        const int kind = g_kind;
        g_kind = 2;
        bool reached = false;
        if (transmissive)
        {
            const V tr = ShadowTransmissive(s, x_r, direction, kRayTMax, medium, c);
            radiance = MulV(Mul(sun, w_sum / w_sun), tr);
            reached = tr.x > 0.0f || tr.y > 0.0f || tr.z > 0.0f;
        }
        else
        {
            const WorldHit h = Trace(s, x_r, direction, kRayTMax, c);
            if (h.attrib == 0u)
            {
                radiance = Mul(sun, w_sum / w_sun);
                reached = true;
            }
        }
        g_kind = kind;
        if (hint != nullptr)
        {
            hint->traced = 1u;
            hint->visible = reached;
        }
    }
    else if (w_sum > 0.0f)
    {
        c.shadow += 1;
        if (hint != nullptr)
        {
            hint->traced = 2u;
        }
        if (t_max > 0.0f)
        {
            const int kind = g_kind;
            g_kind = 2;
            bool reached = false;
            if (transmissive)
            {
                const V tr = ShadowTransmissive(s, x_r, direction, t_max, medium, c);
                radiance = MulV(Mul(contribution, w_sum / w_lamp), tr);
                reached = tr.x > 0.0f || tr.y > 0.0f || tr.z > 0.0f;
            }
            else
            {
                const WorldHit h = Trace(s, x_r, direction, t_max, c);
                if (h.attrib == 0u)
                {
                    radiance = Mul(contribution, w_sum / w_lamp);
                    term.visible = true;
                    reached = true;
                }
            }
            g_kind = kind;
            if (hint != nullptr)
            {
                hint->visible = reached;
            }
        }
    }
    return radiance;
}
float LampHitWeight(const CxScene& s, const Lights& L, uint32_t segment_light, const WorldHit& hit, V origin,
                    V direction, V segment_normal)
{
    if (segment_light == 0u || hit.t < 8.0f)
    {
        return 1.0f;
    }
    const V n = FaceDirection(LowBit(hit.attrib & 0x3Fu));
    const V hp = Add(origin, Mul(direction, hit.t));
    const V inside = Sub(hp, Mul(n, 0.25f));
    const int32_t voxel[3] = {int32_t(std::floor(inside.x)), int32_t(std::floor(inside.y)), int32_t(std::floor(inside.z))};
    bool found;
    const float p_area = LightFacePdf(s, L, origin, segment_normal, hit.record, voxel, LowBit(hit.attrib & 0x3Fu),
                                      (hit.attrib >> 16) & 0xFFFFu, found);
    if (!found)
    {
        return 1.0f;
    }
    if ((segment_light & 2u) == 0u)
    {
        return 0.0f;
    }
    return MisWeight(p_area, segment_normal, direction, hit.t, -Dot(n, direction), false);
}

struct DV
{
    double x = 0, y = 0, z = 0;
};
void AddD(DV& a, V b)
{
    a.x += double(b.x);
    a.y += double(b.y);
    a.z += double(b.z);
}

// The gradient accumulators of the adjoint (float64): dL/dI / spp of the pixel, and the sums.
struct Adjoint
{
    double dcol[3];
    double* grad_albedo;  // [256][3]
    double* grad_scale;   // [bricks]
};

// A term T = base * prod_j albedo(ids[j]) (per channel, float64): adds adj * dT/d(albedo of vertex j)
// for every vertex and returns the product of the albedos.
void AdjointAlbedoTerms(const CxScene& s, Adjoint& g, const double adj[3], const double base[3],
                        const std::vector<uint32_t>& ids, double product[3])
{
    std::vector<std::array<double, 3>> a(ids.size());
    for (int c = 0; c < 3; ++c)
    {
        product[c] = 1.0;
    }
    for (size_t j = 0; j < ids.size(); ++j)
    {
        const V al = MatAlbedo(s, 0x8000u | ids[j]);
        a[j] = {double(al.x), double(al.y), double(al.z)};
        for (int c = 0; c < 3; ++c)
        {
            product[c] *= a[j][c];
        }
    }
    for (size_t j = 0; j < ids.size(); ++j)
    {
        for (int c = 0; c < 3; ++c)
        {
            double others = base[c];
            for (size_t i = 0; i < ids.size(); ++i)
            {
                if (i != j)
                {
                    others *= a[i][c];
                }
            }
            g.grad_albedo[ids[j] * 3 + c] += adj[c] * others;
        }
    }
}

// One pixel sample; radiance accumulated in float64 (the float32 terms are the kernel's). With
// `adjoint`, also the path-replay gradients of the sample (sampling pdfs, MIS weights and the
// roulette's normalisation held constant).
// PreviousPixel (caelix_path_trace.hlsl:154-172): where the previous camera saw the world point p, in
// pixels of the current image; false when p was not in front of it.
bool PreviousPixel(const CxCamera& prev, uint32_t width, uint32_t height, V p, float& x, float& y)
{
    const V rel = Sub(p, Mk(prev.position[0], prev.position[1], prev.position[2]));
    const float z = Dot(rel, Mk(prev.forward[0], prev.forward[1], prev.forward[2]));
    x = y = 0.0f;
    if (!(z > 1e-6f))
    {
        return false;
    }
    const float nx = Dot(rel, Mk(prev.right[0], prev.right[1], prev.right[2])) / ((z * prev.tan_half_fov_y) * prev.aspect);
    const float ny = Dot(rel, Mk(prev.up[0], prev.up[1], prev.up[2])) / (z * prev.tan_half_fov_y);
    x = ((nx + 1.0f) * 0.5f) * float(width);
    y = ((1.0f - ny) * 0.5f) * float(height);
    return std::fabs(x) < 1e30f && std::fabs(y) < 1e30f;
}
// ReadLightHint (caelix_path_trace.hlsl:479-498): 0x8080 with the reset; else the previous image at the
// patch of the previous pixel of the primary surface point (when it has one, there is a previous
// camera, and the point was in front of it and inside its image), else at the pixel's own patch.
uint32_t ReadHint(const CxLightHint& h, const CxCamera& cam, uint32_t px, uint32_t py, bool surface, V p)
{
    if (h.reset != 0u)
    {
        return 0x8080u;
    }
    const uint32_t columns = (cam.width + (1u << h.patch_shift) - 1u) >> h.patch_shift;
    uint32_t qx = px >> h.patch_shift, qy = py >> h.patch_shift;
    float x, y;
    if (surface && h.previous_camera != nullptr && PreviousPixel(*h.previous_camera, cam.width, cam.height, p, x, y) &&
        x >= 0.0f && y >= 0.0f && x < float(cam.width) && y < float(cam.height))
    {
        qx = uint32_t(x) >> h.patch_shift;
        qy = uint32_t(y) >> h.patch_shift;
    }
    return h.previous[uint64_t(qy) * columns + qx];
}
// One channel of WriteLightHint (:554-580): lerp(read, visible / traced, kHintAlpha = 0.2), rounded to
// k / 255 as PackHint rounds (:459-463); unchanged without a traced ray.
uint32_t HintChannel(uint32_t k, uint32_t traced, uint32_t visible)
{
    if (traced == 0u)
    {
        return k;
    }
    const float h = float(k) * (1.0f / 255.0f);
    const float cur = h + 0.2f * (float(visible) / float(traced) - h);
    return uint32_t(std::fmin(std::fmax(cur, 0.0f), 1.0f) * 255.0f + 0.5f);
}
void TracePixel(const CxScene& s, const Lights& L, const CxCamera& cam, uint32_t bounces, uint32_t seed,
                uint32_t flags, const Sky& k, uint32_t px, uint32_t py, DV& radiance, V& albedo, V& emission,
                float& t_hit, uint32_t& attrib, Counters& c, Adjoint* adjoint = nullptr, HintIo* hint = nullptr)
{
    const bool roulette = (flags & CX_TRACE_NO_RUSSIAN_ROULETTE) == 0u;
    const float sx = float(px) + 0.5f;
    const float sy = float(py) + 0.5f;
    const float ndc_x = (sx / float(cam.width)) * 2.0f - 1.0f;
    const float ndc_y = 1.0f - (sy / float(cam.height)) * 2.0f;
    const V right = Mk(cam.right[0], cam.right[1], cam.right[2]);
    const V up = Mk(cam.up[0], cam.up[1], cam.up[2]);
    const V fwd = Mk(cam.forward[0], cam.forward[1], cam.forward[2]);
    const V dir = SafeNormalize(Add(Add(Mul(right, ndc_x * cam.tan_half_fov_y * cam.aspect), Mul(up, ndc_y * cam.tan_half_fov_y)), fwd));
    const V origin = Mk(cam.position[0], cam.position[1], cam.position[2]);
    albedo = Mk(0, 0, 0);
    emission = Mk(0, 0, 0);
    t_hit = 0.0f;
    attrib = 0u;
    c.rays += 1;
    g_kind = 0;
    const WorldHit primary = Trace(s, origin, dir, kRayTMax, c);
    g_kind = 3;
    if (hint != nullptr && hint->params != nullptr)
    {
        hint->word = ReadHint(*hint->params, cam, px, py, primary.attrib != 0u, Add(origin, Mul(dir, primary.t)));
    }
    if (primary.attrib == 0u)
    {
        if (k.on)
        {
            emission = SkyRadiance(k, dir);
        }
        return;
    }
    t_hit = primary.t;
    attrib = primary.attrib;
    uint32_t rng = InitRandom(px, py, seed);
    uint32_t light_rng = InitRandom(px, py, seed ^ 0x7F4A7C15u);
    uint32_t nee_rng = InitRandom(px, py, seed ^ 0x68E31DA4u);
    const uint32_t mis_bit = L.mis != 0u ? 2u : 0u;
    uint32_t segment_light = 0u;
    const V primary_normal = FaceDirection(LowBit(primary.attrib & 0x3Fu));
    const Shade ps = ShadeHit(s, (primary.attrib >> 16) & 0xFFFFu, Add(origin, Mul(dir, primary.t)), primary_normal, rng);
    albedo = ps.albedo;
    emission = Mul(ps.emission, BrickScale(s, primary.record));
    if (k.fog_density > 0.0f)
    {
        // The camera segment: its transmission multiplies what the surface sends, its fog emission
        // is deterministic (caelix_path_trace.hlsl:1273-1283).
        V chain = Mk(1, 1, 1), inscatter = Mk(0, 0, 0);
        FogSegment(k.fog, k.fog_density, primary.t, chain, inscatter);
        albedo = MulV(chain, ps.albedo);
        emission = Add(MulV(chain, emission), inscatter);
    }
    double adj[3] = {0, 0, 0};
    double rr = 1.0;
    std::vector<uint32_t> ids;
    if (adjoint != nullptr)
    {
        const double e[3] = {ps.emission.x, ps.emission.y, ps.emission.z};
        const double a[3] = {ps.albedo.x, ps.albedo.y, ps.albedo.z};
        double v = 0.0;
        for (int k = 0; k < 3; ++k)
        {
            v += adjoint->dcol[k] * e[k];
            adj[k] = adjoint->dcol[k] * a[k];
        }
        adjoint->grad_scale[primary.record] += v;
    }
    // An unoccluded NEE term: unscaled * scale(light) * rr * prod(albedos of ids).
    auto nee_adjoint = [&](const NeeTerm& term) {
        if (adjoint == nullptr || !term.visible)
        {
            return;
        }
        const double se = BrickScale(s, term.light);
        const double u[3] = {double(term.unscaled.x) * rr, double(term.unscaled.y) * rr, double(term.unscaled.z) * rr};
        const double base[3] = {u[0] * se, u[1] * se, u[2] * se};
        double product[3];
        AdjointAlbedoTerms(s, *adjoint, adj, base, ids, product);
        double v = 0.0;
        for (int k = 0; k < 3; ++k)
        {
            v += adj[k] * u[k] * product[k];
        }
        adjoint->grad_scale[term.light] += v;
    };
    V throughput = Mk(1, 1, 1);
    V segment_normal = primary_normal;
    if (bounces >= 1u)
    {
        NeeTerm term;
        AddD(radiance, MulV(NextEvent(s, L, k, ps.origin, primary_normal, light_rng, nee_rng, c, term, false, 0u,
                                      hint != nullptr && hint->params != nullptr ? hint : nullptr),
                            throughput));
        nee_adjoint(term);
        segment_light = 1u | mis_bit;
    }
    V bo = ps.origin;
    V bd = ps.direction;
    uint32_t bounce_index = 1;
    while (bounce_index <= bounces)
    {
        c.rays += 1;
        g_kind = 1;
        const WorldHit hit = Trace(s, bo, bd, kRayTMax, c);
        g_kind = 4;
        if (hit.attrib == 0u)
        {
            if (k.on)
            {
                AddD(radiance, MulV(SkyMiss(k, segment_normal, bd), throughput));
            }
            break;
        }
        if (k.fog_density > 0.0f)
        {
            // The segment's fog, added in float32 as the kernels add it to their float32 radiance.
            V fog_add = Mk(0, 0, 0);
            FogSegment(k.fog, k.fog_density, hit.t, throughput, fog_add);
            AddD(radiance, fog_add);
        }
        const V vn = FaceDirection(LowBit(hit.attrib & 0x3Fu));
        const uint32_t block = (hit.attrib >> 16) & 0xFFFFu;
        const Shade sh = ShadeHit(s, block, Add(bo, Mul(bd, hit.t)), vn, rng);
        const V scaled_emission = Mul(sh.emission, BrickScale(s, hit.record));
        bounce_index += 1;
        float ew = 1.0f;
        if (segment_light != 0u && (sh.emission.x > 0.0f || sh.emission.y > 0.0f || sh.emission.z > 0.0f))
        {
            ew = LampHitWeight(s, L, segment_light, hit, bo, bd, segment_normal);
        }
        AddD(radiance, MulV(scaled_emission, Mul(throughput, ew)));
        if (adjoint != nullptr && (sh.emission.x > 0.0f || sh.emission.y > 0.0f || sh.emission.z > 0.0f))
        {
            const double w = double(ew) * rr;
            const double base[3] = {scaled_emission.x * w, scaled_emission.y * w, scaled_emission.z * w};
            double product[3];
            AdjointAlbedoTerms(s, *adjoint, adj, base, ids, product);
            double v = 0.0;
            const double e[3] = {sh.emission.x, sh.emission.y, sh.emission.z};
            for (int k = 0; k < 3; ++k)
            {
                v += adj[k] * e[k] * w * product[k];
            }
            adjoint->grad_scale[hit.record] += v;
        }
        throughput = MulV(throughput, Mk(sh.albedo.x / 1.0f, sh.albedo.y / 1.0f, sh.albedo.z / 1.0f));
        if (ids.size() < 3)
        {
            ids.push_back(block & 0xFFu);
        }
        segment_light = 0u;
        if (bounce_index <= bounces)
        {
            segment_light = 1u | mis_bit;
            // The bounce NEE budget (caelix_path_trace.hlsl:909-944): with probability q, scaled by 1 / q.
            if (k.nee_q >= 1.0f || RandomFloat01(nee_rng) < k.nee_q)
            {
                NeeTerm term;
                AddD(radiance, MulV(NextEvent(s, L, k, sh.origin, vn, light_rng, nee_rng, c, term),
                                    Mul(throughput, 1.0f / k.nee_q)));
                nee_adjoint(term);
            }
        }
        const float cp = std::fmax(throughput.x, std::fmax(throughput.y, throughput.z));
        if (cp <= 0.0f)
        {
            break;
        }
        const float u = RandomFloat01(rng);
        if (roulette)
        {
            if (cp < u)
            {
                break;
            }
            const float inv_cp = 1.0f / cp;
            throughput = Mul(throughput, inv_cp);
            rr *= double(inv_cp);
        }
        bo = sh.origin;
        bd = sh.direction;
        segment_normal = vn;
    }
    if (adjoint != nullptr)
    {
        const uint32_t id = (primary.attrib >> 16) & 0xFFu;
        adjoint->grad_albedo[id * 3 + 0] += adjoint->dcol[0] * radiance.x;
        adjoint->grad_albedo[id * 3 + 1] += adjoint->dcol[1] * radiance.y;
        adjoint->grad_albedo[id * 3 + 2] += adjoint->dcol[2] * radiance.z;
    }
}

// TracePath in the full material model (caelix_path_trace.hlsl:595-678, 737-978): the delta chain of
// the camera ray (RunDeltaChain: smooth transparent interfaces crossed up to the budget, every segment
// through its medium), ShadeVoxelHit at the surface that ended it (albedo and emission times the
// chain's transmission, plus the fog along it), then the bounce loop with media, specular lobes (no
// NEE at a specular vertex; its segment sees the sun disk and counts a lamp in full) and the
// procedural sky on misses. A chain that leaves the scene shows the sky in its last direction times
// its transmission. t_hit = the chain's length, attrib = the last surface or interface it met.
void TracePixelGlass(const CxScene& s, const Lights& L, const CxCamera& cam, uint32_t bounces, uint32_t seed,
                     uint32_t flags, const Sky& k, uint32_t px, uint32_t py, DV& radiance, V& albedo, V& emission,
                     float& t_hit, uint32_t& attrib, Counters& c, HintIo* hint = nullptr)
{
    const bool roulette = (flags & CX_TRACE_NO_RUSSIAN_ROULETTE) == 0u;
    const float sx = float(px) + 0.5f;
    const float sy = float(py) + 0.5f;
    const float ndc_x = (sx / float(cam.width)) * 2.0f - 1.0f;
    const float ndc_y = 1.0f - (sy / float(cam.height)) * 2.0f;
    const V right = Mk(cam.right[0], cam.right[1], cam.right[2]);
    const V up = Mk(cam.up[0], cam.up[1], cam.up[2]);
    const V fwd = Mk(cam.forward[0], cam.forward[1], cam.forward[2]);
    const V dir = SafeNormalize(Add(Add(Mul(right, ndc_x * cam.tan_half_fov_y * cam.aspect), Mul(up, ndc_y * cam.tan_half_fov_y)), fwd));
    const V origin = Mk(cam.position[0], cam.position[1], cam.position[2]);
    albedo = Mk(0, 0, 0);
    emission = Mk(0, 0, 0);
    t_hit = 0.0f;
    attrib = 0u;
    c.rays += 1;
    g_kind = 0;
    const WorldHit primary = Trace(s, origin, dir, kRayTMax, c);
    g_kind = 3;
    const bool hinted = hint != nullptr && hint->params != nullptr;
    if (primary.attrib == 0u)
    {
        if (hinted)
        {
            hint->word = ReadHint(*hint->params, cam, px, py, false, origin);
        }
        if (k.on)
        {
            emission = SkyRadiance(k, dir);
        }
        return;
    }
    uint32_t rng = InitRandom(px, py, seed);
    uint32_t light_rng = InitRandom(px, py, seed ^ 0x7F4A7C15u);
    uint32_t nee_rng = InitRandom(px, py, seed ^ 0x68E31DA4u);
    const uint32_t mis_bit = L.mis != 0u ? 2u : 0u;
    WorldHit surface = primary;
    uint32_t geometry = primary.attrib;
    V ro = origin, rd = dir;
    V chain = Mk(1, 1, 1), inscatter = Mk(0, 0, 0);
    float t_virtual = 0.0f;
    uint32_t medium = 0u, interfaces = 0u;
    bool surface_hit = true;
    for (;;)
    {
        MediumSegment(s, k, medium, surface.t, chain, inscatter);
        t_virtual += surface.t;
        geometry = surface.attrib;
        if (interfaces >= k.delta_budget)
        {
            break;
        }
        const uint32_t block = (surface.attrib >> 16) & 0xFFFFu;
        if (Opaque(block, s.encoding))
        {
            break;
        }
        const uint32_t code = MaterialCode(block, s.encoding);
        const Material m = GetMaterial(s, code);
        if (m.smoothness < 0.99f)
        {
            break;
        }
        const V hp = Add(ro, Mul(rd, surface.t));
        const int split = k.checkerboard && interfaces == 0u ? (((px + py) & 1u) == 0u ? 1 : 0) : -1;
        const Interface it =
            CrossInterface(s, medium, code, m.ior, rd, FaceDirection(LowBit(surface.attrib & 0x3Fu)), rng, split);
        if (it.split_applied)
        {
            chain = Mul(chain, it.weight);
        }
        interfaces += 1u;
        medium = it.medium;
        ro = Add(hp, Mul(it.normal, it.refracted ? -kPushOff : kPushOff));
        rd = it.direction;
        c.rays += 1;
        g_kind = 0;
        surface = Trace(s, ro, rd, kRayTMax, c);
        g_kind = 3;
        if (surface.attrib == 0u)
        {
            surface_hit = false;
            break;
        }
    }
    t_hit = t_virtual;
    attrib = geometry;
    if (hinted)
    {
        hint->word = ReadHint(*hint->params, cam, px, py, surface_hit, Add(ro, Mul(rd, surface.t)));
    }
    if (!surface_hit)
    {
        const V sky = k.on ? SkyRadiance(k, rd) : Mk(0, 0, 0);
        emission = Add(MulV(chain, sky), inscatter);
        return;
    }
    V pn = FaceDirection(LowBit(surface.attrib & 0x3Fu));
    const V php = Add(ro, Mul(rd, surface.t));
    ShadeFull ps = ShadeVoxel(s, (surface.attrib >> 16) & 0xFFFFu, php, rd, pn, rng, medium);
    // Subsurface at the primary vertex (caelix_path_trace.hlsl:757-781): throughput = branch weight /
    // guide albedo, the albedo output the guide; a dead walk ends the path after the emission.
    V throughput = Mk(1, 1, 1);
    uint32_t walk = kSssNone;
    if (k.subsurface)
    {
        V guide;
        walk = SssVertex(s, surface.record, surface.attrib, php, rd, rng, ps.albedo, ps.origin, ps.direction, ps.specular,
                         ps.roughness, pn, guide, false);
        if (walk != kSssNone)
        {
            throughput = Mk(ps.albedo.x / std::fmax(guide.x, 0.001f), ps.albedo.y / std::fmax(guide.y, 0.001f),
                            ps.albedo.z / std::fmax(guide.z, 0.001f));
            ps.albedo = guide;
            if (walk == kSssExit)
            {
                medium = 0u;
            }
        }
    }
    albedo = MulV(chain, ps.albedo);
    emission = Add(MulV(chain, Mul(ps.emission, BrickScale(s, surface.record))), inscatter);
    if (walk == kSssDead)
    {
        return;
    }
    V segment_normal = pn;
    float segment_rough = ps.specular ? ps.roughness : -1.0f;  // the segment's lobe; -1: diffuse
    uint32_t segment_light = 0u;
    bool sees_sun = true;
    bool claim = false;
    if (bounces >= 1u && !ps.specular)
    {
        NeeTerm term;
        AddD(radiance, MulV(NextEvent(s, L, k, ps.origin, pn, light_rng, nee_rng, c, term, k.transmissive, medium,
                                      hinted ? hint : nullptr),
                            throughput));
        segment_light = 1u | mis_bit;
        sees_sun = !k.sun;
        claim = k.transmissive;
    }
    V bo = ps.origin;
    V bd = ps.direction;
    uint32_t bounce_index = 1;
    while (bounce_index <= bounces)
    {
        c.rays += 1;
        g_kind = 1;
        const WorldHit hit = Trace(s, bo, bd, kRayTMax, c);
        g_kind = 4;
        if (hit.attrib == 0u)
        {
            if (k.on)
            {
                AddD(radiance, MulV(segment_rough < 0.0f ? SkyDiffuseMiss(k, segment_normal, bd, sees_sun)
                                                         : SkySpecularMiss(k, bd, segment_rough, sees_sun),
                                    throughput));
            }
            break;
        }
        V segment_add = Mk(0, 0, 0);
        MediumSegment(s, k, medium, hit.t, throughput, segment_add);
        AddD(radiance, segment_add);
        V vn = FaceDirection(LowBit(hit.attrib & 0x3Fu));
        const uint32_t block = (hit.attrib >> 16) & 0xFFFFu;
        const V hp = Add(bo, Mul(bd, hit.t));
        ShadeFull sh = ShadeVoxel(s, block, hp, bd, vn, rng, medium);
        uint32_t walk = kSssNone;
        if (k.subsurface)
        {
            V guide;
            walk = SssVertex(s, hit.record, hit.attrib, hp, bd, rng, sh.albedo, sh.origin, sh.direction, sh.specular,
                             sh.roughness, vn, guide, bounce_index + 1u > bounces);
        }
        const V scaled_emission = Mul(sh.emission, BrickScale(s, hit.record));
        bounce_index += 1;
        float ew = 1.0f;
        if (segment_light != 0u && (sh.emission.x > 0.0f || sh.emission.y > 0.0f || sh.emission.z > 0.0f))
        {
            ew = LampHitWeight(s, L, segment_light, hit, bo, bd, segment_normal);
        }
        AddD(radiance, MulV(scaled_emission, Mul(throughput, ew)));
        if (walk == kSssDead)
        {
            break;
        }
        if (walk == kSssExit)
        {
            medium = 0u;
        }
        throughput = MulV(throughput, Mk(sh.albedo.x / 1.0f, sh.albedo.y / 1.0f, sh.albedo.z / 1.0f));
        segment_light = 0u;
        sees_sun = true;
        if (claim)
        {
            // The claim of a transmissive primary NEE (caelix_path_trace.hlsl:895-907).
            if (!Opaque(block, s.encoding))
            {
                sees_sun = !k.sun;
                segment_light = 1u;
            }
            else
            {
                claim = false;
            }
        }
        if (!sh.specular && bounce_index <= bounces)
        {
            segment_light = 1u | mis_bit;
            sees_sun = !k.sun;
            if (k.nee_q >= 1.0f || RandomFloat01(nee_rng) < k.nee_q)
            {
                NeeTerm term;
                AddD(radiance, MulV(NextEvent(s, L, k, sh.origin, vn, light_rng, nee_rng, c, term),
                                    Mul(throughput, 1.0f / k.nee_q)));
            }
        }
        const float cp = std::fmax(throughput.x, std::fmax(throughput.y, throughput.z));
        if (cp <= 0.0f)
        {
            break;
        }
        const float u = RandomFloat01(rng);
        if (roulette)
        {
            if (cp < u)
            {
                break;
            }
            throughput = Mul(throughput, 1.0f / cp);
        }
        bo = sh.origin;
        bd = sh.direction;
        segment_normal = vn;
        segment_rough = sh.specular ? sh.roughness : -1.0f;
    }
}

// LtOpenFaces (caelix_light_tree_build.hlsl:92-111): bit f set when voxel v of `record` has a light
// material code and its face f is open (the neighbour, through the hash table in another brick, is not
// opaque); `block` = the voxel's value.
uint32_t OpenLightFaces(const CxScene* s, uint32_t record, uint32_t v, uint32_t& block)
{
    block = s->payload[uint64_t(record) * 512 + v];
    const uint32_t code = MaterialCode(block, s->encoding);
    if (((s->light_codes[code >> 5] >> (code & 31u)) & 1u) == 0u)
    {
        return 0u;
    }
    const int p0[3] = {int(v & 7u), int((v >> 3) & 7u), int(v >> 6)};
    uint32_t open = 0;
    for (uint32_t f = 0; f < 6; ++f)
    {
        int p[3] = {p0[0], p0[1], p0[2]};
        const int axis = int(f >> 1);
        const int sign = (f & 1u) ? -1 : 1;
        p[axis] += sign;
        uint32_t other = record;
        if (p[axis] < 0 || p[axis] > 7)
        {
            int32_t k[3] = {s->keys[3 * uint64_t(record)], s->keys[3 * uint64_t(record) + 1], s->keys[3 * uint64_t(record) + 2]};
            k[axis] += sign;
            p[axis] = (p[axis] + 8) & 7;
            uint32_t probes;
            other = HashGet(*s, k[0], k[1], k[2], probes);
        }
        const uint32_t nb = other == kNoneU ? 0u : s->payload[uint64_t(other) * 512 + p[0] + 8 * p[1] + 64 * p[2]];
        if (!Opaque(nb, s->encoding))
        {
            open |= 1u << f;
        }
    }
    return open;
}

uint32_t LightFaceCount(uint32_t word1)
{
    return (word1 >> kLtRecordFaceCountShift) & kLtRecordFaceCountMask;
}

// LtRecordUnits (caelix_light_tree_group.hlsl:209-212).
uint32_t RecordUnitsOf(uint32_t entries)
{
    return (48u + 4u * entries + 15u) / 16u;
}

// LtBuildBrick's recomputed record (caelix_light_tree_build.hlsl:161-282) at `rec` with room for
// `units` units: the open faces of the brick's light voxels in voxel order per family, at most the
// finalize's face count kept (families in order), face weights (luminance of the emission) summed per
// run of 8 voxels and then over the 64 runs in order, the tight box of the kept voxels (the brick
// corner when none), zero padding. Returns the entries written.
uint32_t RecomputeRecord(const CxScene* s, uint32_t record, uint32_t pos, uint32_t shift, uint32_t units, uint32_t* rec)
{
    uint32_t opens[512], blocks[512];
    uint32_t found[6] = {};
    for (uint32_t v = 0; v < 512; ++v)
    {
        opens[v] = OpenLightFaces(s, record, v, blocks[v]);
        for (uint32_t f = 0; f < 6; ++f)
        {
            found[f] += (opens[v] >> f) & 1u;
        }
    }
    uint32_t left = LightFaceCount(s->info[2 * uint64_t(record) + 1]);
    uint32_t family_count[6], family_base[6], cursor[6] = {};
    uint32_t written = 0;
    for (uint32_t f = 0; f < 6; ++f)
    {
        family_count[f] = std::min(found[f], left);
        family_base[f] = written;
        written += family_count[f];
        left -= family_count[f];
    }
    float face[6] = {};
    uint32_t lo[3] = {8, 8, 8}, hi[3] = {0, 0, 0};
    for (uint32_t run = 0; run < 64; ++run)
    {
        float sums[6] = {};
        for (uint32_t i = 0; i < 8; ++i)
        {
            const uint32_t v = run * 8 + i;
            if (opens[v] == 0)
            {
                continue;
            }
            const float* e = s->materials + uint64_t(MaterialCode(blocks[v], s->encoding)) * 8 + 4;
            const float weight = e[0] * 0.2126f + e[1] * 0.7152f + e[2] * 0.0722f;
            bool kept = false;
            for (uint32_t f = 0; f < 6; ++f)
            {
                if ((opens[v] >> f) & 1u)
                {
                    if (cursor[f] < family_count[f])
                    {
                        rec[12 + family_base[f] + cursor[f]] = v | (blocks[v] << 16);
                        sums[f] += weight;
                        kept = true;
                    }
                    ++cursor[f];
                }
            }
            if (kept)
            {
                const uint32_t c[3] = {v & 7u, (v >> 3) & 7u, v >> 6};
                for (int a = 0; a < 3; ++a)
                {
                    lo[a] = std::min(lo[a], c[a]);
                    hi[a] = std::max(hi[a], c[a] + 1);
                }
            }
        }
        for (uint32_t f = 0; f < 6; ++f)
        {
            face[f] += sums[f];
        }
    }
    const uint32_t mask = (1u << shift) - 1u;
    const uint32_t o[3] = {(pos & mask) * 8, ((pos >> shift) & mask) * 8, (pos >> (2 * shift)) * 8};
    for (int a = 0; a < 3; ++a)
    {
        lo[a] = o[a] + (written == 0 ? 0u : lo[a]);
        hi[a] = o[a] + (written == 0 ? 0u : hi[a]);
    }
    std::memcpy(rec, face, 24);
    rec[6] = lo[0] | (lo[1] << 10) | (lo[2] << 20);
    rec[7] = hi[0] | (hi[1] << 10) | (hi[2] << 20);
    rec[8] = pos;
    rec[9] = written;
    rec[10] = family_count[0] | (family_count[1] << 10) | (family_count[2] << 20);
    rec[11] = family_count[3] | (family_count[4] << 10) | (family_count[5] << 20);
    for (uint32_t e = written; e < (units * 16 - 48) / 4; ++e)
    {
        rec[12 + e] = 0u;
    }
    return written;
}

// LtWorldSummary (caelix_light_gather_emit.hlsl:11-34).
LightSummary WorldSummaryOf(const float face[6], const float lo[3], const float hi[3], const float rows[3][4])
{
    LightSummary w;
    float c[3], h[3];
    for (int k = 0; k < 3; ++k)
    {
        c[k] = 0.5f * (lo[k] + hi[k]);
        h[k] = 0.5f * (hi[k] - lo[k]);
    }
    for (int r = 0; r < 3; ++r)
    {
        const float wc = ((rows[r][0] * c[0] + rows[r][1] * c[1]) + rows[r][2] * c[2]) + rows[r][3];
        const float wh = (std::fabs(rows[r][0]) * h[0] + std::fabs(rows[r][1]) * h[1]) + std::fabs(rows[r][2]) * h[2];
        w.lo[r] = wc - wh;
        w.hi[r] = wc + wh;
    }
    for (int f = 0; f < 6; ++f)
    {
        w.face[f] = 0.0f;
    }
    for (int f = 0; f < 6; ++f)
    {
        const int axis = f >> 1;
        const float sign = (f & 1) != 0 ? -1.0f : 1.0f;
        for (int b = 0; b < 3; ++b)
        {
            const float d = sign * rows[b][axis];
            w.face[2 * b] += face[f] * std::fmax(0.0f, d);
            w.face[2 * b + 1] += face[f] * std::fmax(0.0f, -d);
        }
    }
    return w;
}

}  // namespace

extern "C" {

// The finalize's light state (caelix_finalize.hlsl:22-29, 326-340), after oracle_scene_finalize (which
// leaves word 1's light bits as it found them): per brick the open faces of its light voxels into word 1
// bits 18..29; when the old count or the new one is not 0 the stale bit (bit 30) is set and the brick's
// group row gets kLtGpuLightChanged; a set stale bit is kept.
void oracle_scene_light_state_bricks(CxScene* s, const uint32_t* records, uint32_t count)
{
    for (uint32_t i = 0; i < count; ++i)
    {
        const uint32_t b = records != nullptr ? records[i] : i;
        if (s->group_of_brick[b] == kNoneU)
        {
            continue;  // a free record
        }
        uint32_t faces = 0;
        for (uint32_t v = 0; v < 512; ++v)
        {
            uint32_t block;
            faces += uint32_t(__builtin_popcount(OpenLightFaces(s, b, v, block)));
        }
        const uint32_t old_word1 = s->info[2 * uint64_t(b) + 1];
        uint32_t word1 = (old_word1 & kRecordBoundsMask) | (faces << kLtRecordFaceCountShift) | (old_word1 & kLtRecordStaleBit);
        if (LightFaceCount(old_word1) != 0 || faces != 0)
        {
            word1 |= kLtRecordStaleBit;
            s->light_group_rows[s->group_of_brick[b]].gpu_flags |= kLtGpuLightChanged;
        }
        s->info[2 * uint64_t(b) + 1] = word1;
    }
}

void oracle_scene_finalize_bricks(CxScene* s, const uint32_t* records, uint32_t count)
{
    uint64_t* occ_all = static_cast<uint64_t*>(s->occupancy);
    const bool packed = s->encoding == kEncodingPacked;
    const uint32_t opaque_mask = packed ? kPackedOpaqueMask : kBlockIdOpaqueMask;
    const uint32_t id_mask = packed ? kPackedIdMask : kBlockIdIdMask;
    const uint32_t face_shift = packed ? kPackedFaceShift : kBlockIdFaceShift;
    auto block_at = [&](uint32_t record, int x, int y, int z) -> uint32_t {
        int p[3] = {x, y, z};
        int32_t k[3] = {s->keys[3 * uint64_t(record)], s->keys[3 * uint64_t(record) + 1], s->keys[3 * uint64_t(record) + 2]};
        bool outside = false;
        for (int a = 0; a < 3; ++a)
        {
            if (p[a] < 0 || p[a] > 7)
            {
                k[a] += p[a] < 0 ? -1 : 1;
                p[a] = (p[a] + 8) & 7;
                outside = true;
            }
        }
        uint32_t other = record;
        if (outside)
        {
            uint32_t probes;
            other = HashGet(*s, k[0], k[1], k[2], probes);
            if (other == kNoneU)
            {
                return 0u;
            }
        }
        return s->payload[uint64_t(other) * 512 + p[0] + 8 * p[1] + 64 * p[2]];
    };
    constexpr int kDir[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (uint32_t i = 0; i < count; ++i)
    {
        const uint32_t b = records != nullptr ? records[i] : i;
        uint64_t occ[8] = {};
        uint32_t lo[3] = {7, 7, 7}, hi[3] = {0, 0, 0};
        bool any = false;
        for (uint32_t v = 0; v < 512; ++v)
        {
            const int x = int(v & 7u), y = int((v >> 3) & 7u), z = int(v >> 6);
            const uint32_t raw = s->payload[uint64_t(b) * 512 + v];
            const uint32_t own = (raw & opaque_mask) != 0u ? raw : (raw & id_mask);
            uint32_t opaque_count = 0, faces = 0;
            for (int d = 0; d < 6; ++d)
            {
                const uint32_t n = block_at(b, x + kDir[d][0], y + kDir[d][1], z + kDir[d][2]);
                if ((n & opaque_mask) != 0u)
                {
                    ++opaque_count;
                }
                else if ((n & id_mask) != own)
                {
                    faces |= 1u << d;
                }
            }
            const bool opaque = (own & opaque_mask) != 0u;
            const uint32_t result = opaque ? own : (own | (faces << face_shift));
            s->payload[uint64_t(b) * 512 + v] = uint16_t(result);
            if (!(opaque_count < 6 && result != 0u && (opaque || faces != 0u)))
            {
                continue;
            }
            occ[(x >> 2) | ((y >> 2) << 1) | ((z >> 2) << 2)] |= 1ull << ((x & 3u) | ((y & 3u) << 2) | ((z & 3u) << 4));
            any = true;
            const uint32_t c[3] = {uint32_t(x), uint32_t(y), uint32_t(z)};
            for (int a = 0; a < 3; ++a)
            {
                lo[a] = std::min(lo[a], c[a]);
                hi[a] = std::max(hi[a], c[a]);
            }
        }
        uint32_t coarse = 0;
        for (int c = 0; c < 8; ++c)
        {
            occ_all[uint64_t(b) * 8 + c] = occ[c];
            coarse |= occ[c] != 0 ? (1u << c) : 0u;
        }
        const int32_t* key = s->keys + uint64_t(b) * 3;
        const uint32_t mask = (1u << s->group_shift) - 1u;
        const uint32_t pos = (uint32_t(key[0]) & mask) | ((uint32_t(key[1]) & mask) << s->group_shift) |
                             ((uint32_t(key[2]) & mask) << (2 * s->group_shift));
        s->info[2 * uint64_t(b)] = (pos & kRecordPosMask) | (coarse << kRecordCoarseShift);
        // the light bits of word 1 stay for the light state (0 without it)
        const uint32_t light_bits = s->light_group_rows != nullptr ? (s->info[2 * uint64_t(b) + 1] & ~kRecordBoundsMask) : 0u;
        s->info[2 * uint64_t(b) + 1] =
            (any ? (lo[0] | (lo[1] << 3) | (lo[2] << 6) | (hi[0] << 9) | (hi[1] << 12) | (hi[2] << 15)) : 0u) | light_bits;
        const int32_t g[3] = {key[0] - s->grid_min[0], key[1] - s->grid_min[1], key[2] - s->grid_min[2]};
        if (any && uint32_t(g[0]) < s->grid_dim && uint32_t(g[1]) < s->grid_dim && uint32_t(g[2]) < s->grid_dim)
        {
            const uint64_t i = (uint64_t(g[2]) * s->grid_dim + g[1]) * s->grid_dim + g[0];
            s->brick_bits[i >> 5] |= 1u << (i & 31u);
            const uint32_t sd = s->grid_dim >> 3;
            const uint64_t j = (uint64_t(g[2] >> 3) * sd + (g[1] >> 3)) * sd + (g[0] >> 3);
            s->super_bits[j >> 5] |= 1u << (j & 31u);
        }
    }
    if (s->light_group_rows != nullptr)
    {
        oracle_scene_light_state_bricks(s, records, count);
    }
}

// The whole scene: the bitmaps cleared, every record finalized.
void oracle_scene_finalize(CxScene* s)
{
    const uint64_t cells = uint64_t(s->grid_dim) * s->grid_dim * s->grid_dim;
    std::memset(s->brick_bits, 0, ((cells + 31) / 32) * 4);
    std::memset(s->super_bits, 0, ((cells / 512 + 31) / 32) * 4);
    oracle_scene_finalize_bricks(s, nullptr, s->bricks);
}

// The group pass (shaders/light/records/caelix_light_tree_{size,scan,build,commit}.hlsl,
// caelix_light_tree_group.hlsl:129-243), sequentially over the group indices, on host tables named by a
// descriptor of host addresses. The reference scans a group's dense ids; here the bricks of group g are
// those with group_of_brick = g, and a brick's dense id is its record. Per group: the staged flags merged,
// the mode (released, rebuild: kLtCpuRebuild or kLtGpuLightChanged, else copy); a rebuilt group's light
// bricks by Morton position (a bitmap; above 4096 the lowest positions kept), each planned as a copy of
// its old record (not stale, an old block that is not fresh, an entry at its position) or a recompute;
// sizes; the exclusive scan and the decision; on kLtPassOk the blocks (copied whole, or header, index
// entries and records), rows, stale bits, statistics and the current buffer.
void oracle_light_pass(CxScene* s, const CxLightPass* pass)
{
    const LightTreeDescriptor* d = pass->descriptor;
    LightControl* control = reinterpret_cast<LightControl*>(d->control);
    GroupLightRow* rows = reinterpret_cast<GroupLightRow*>(d->group_rows);
    const GroupInfo* infos = reinterpret_cast<const GroupInfo*>(d->group_infos);
    const uint32_t* motion = reinterpret_cast<const uint32_t*>(d->entity_motion);
    const uint32_t shift = d->group_shift;
    const uint32_t G = pass->group_count;
    const uint32_t current = control->current & 1u;
    const uint32_t* source = reinterpret_cast<const uint32_t*>(d->heap[current]);
    uint32_t* target = reinterpret_cast<uint32_t*>(d->heap[1 - current]);
    std::vector<std::vector<uint32_t>> members(G);
    for (uint32_t b = 0; b < s->bricks; ++b)
    {
        if (s->group_of_brick[b] < G)  // kLtNone: a free record
        {
            members[s->group_of_brick[b]].push_back(b);
        }
    }
    struct Plan
    {
        uint32_t record, pos, units, old_record;  // old_record kNoneU: recompute
    };
    std::vector<uint32_t> mode(G), sizes(G);
    std::vector<std::vector<Plan>> plans(G);
    std::vector<uint32_t> dropped(G, 0);
    uint32_t pass_flags = control->pass_flags;
    // size kernel (caelix_light_tree_size.hlsl)
    for (uint32_t g = 0; g < G; ++g)
    {
        GroupLightRow& row = rows[g];
        uint32_t cpu = row.cpu_flags;
        if (pass->staged_flags[g] != 0)
        {
            cpu = (cpu & kLtCpuFresh) | pass->staged_flags[g];
            row.cpu_flags = cpu;
            pass->staged_flags[g] = 0;
        }
        mode[g] = (cpu & kLtCpuReleased) != 0 ? 2u
                  : ((cpu & kLtCpuRebuild) != 0 || (row.gpu_flags & kLtGpuLightChanged) != 0) ? 1u
                                                                                               : 0u;
        uint32_t units = mode[g] == 0 ? row.block_units : 0u;
        if (mode[g] == 1)
        {
            // LtScanGroup: the light bricks by Morton position, the lowest kLtMaxLightBricksPerGroup kept
            std::vector<std::pair<uint32_t, uint32_t>> by_morton;  // (Morton, record)
            for (uint32_t b : members[g])
            {
                if (LightFaceCount(s->info[2 * uint64_t(b) + 1]) != 0)
                {
                    const int32_t* k = s->keys + 3 * uint64_t(b);
                    const uint32_t m = (1u << shift) - 1u;
                    by_morton.emplace_back(Morton3(uint32_t(k[0]) & m, uint32_t(k[1]) & m, uint32_t(k[2]) & m), b);
                }
            }
            std::sort(by_morton.begin(), by_morton.end());
            if (by_morton.size() > kLtMaxLightBricksPerGroup)
            {
                dropped[g] = uint32_t(by_morton.size()) - kLtMaxLightBricksPerGroup;
                by_morton.resize(kLtMaxLightBricksPerGroup);
            }
            const bool old_valid = row.block_units != 0 && (cpu & kLtCpuFresh) == 0;
            const uint32_t* old = source + uint64_t(row.block_offset) * 4;
            uint32_t records = 0;
            for (const auto& [m, b] : by_morton)
            {
                const uint32_t word1 = s->info[2 * uint64_t(b) + 1];
                const int32_t* k = s->keys + 3 * uint64_t(b);
                const uint32_t mk = (1u << shift) - 1u;
                Plan p{b, (uint32_t(k[0]) & mk) | ((uint32_t(k[1]) & mk) << shift) | ((uint32_t(k[2]) & mk) << (2 * shift)),
                       RecordUnitsOf(LightFaceCount(word1)), kNoneU};
                if ((word1 & kLtRecordStaleBit) == 0 && old_valid)
                {
                    const uint32_t e = FindIndexEntry(old, old[0], m, shift);
                    if (e != kNoneU)
                    {
                        p.old_record = IndexEntryOf(old, e).record_offset;
                        p.units = RecordUnitsOf(old[uint64_t(p.old_record) * 4 + 9]);
                    }
                }
                records += p.units;
                plans[g].push_back(p);
            }
            units = plans[g].empty() ? 0u : 1u + uint32_t(plans[g].size()) + records;
        }
        sizes[g] = units;
        if (mode[g] != 0)
        {
            pass_flags |= kLtPassFlagChanged | (1u << (kLtPassFlagTreeShift + (motion[infos[g].entity] & 1u)));
        }
    }
    // scan kernel (caelix_light_tree_scan.hlsl)
    uint64_t carry = 0;
    std::vector<uint32_t> offsets(G);
    for (uint32_t g = 0; g < G; ++g)
    {
        offsets[g] = uint32_t(std::min(carry, uint64_t(0xFFFFFFFFu)));
        carry += sizes[g];
    }
    const uint32_t changed = pass_flags & kLtPassFlagChanged;
    const uint32_t result = changed == 0 ? kLtPassNoChange : (carry <= pass->heap_units ? kLtPassOk : kLtPassFailed);
    control->pass_result = result;
    if (changed != 0)
    {
        control->needed_units = uint32_t(std::min(carry, uint64_t(0xFFFFFFFFu)));
        control->records_recomputed = 0;
        control->records_copied = 0;
    }
    control->trees_changed = (pass_flags >> kLtPassFlagTreeShift) & 3u;
    control->pass_flags = 0;
    if (result == kLtPassFailed)
    {
        control->passes += 1;
        control->failed_passes += 1;
        return;
    }
    if (result != kLtPassOk)
    {
        return;
    }
    // build kernel (caelix_light_tree_build.hlsl)
    uint32_t recomputed = 0, copied = 0;
    for (uint32_t g = 0; g < G; ++g)
    {
        GroupLightRow& row = rows[g];
        if (mode[g] == 0)
        {
            std::memcpy(target + uint64_t(offsets[g]) * 4, source + uint64_t(row.block_offset) * 4,
                        uint64_t(row.block_units) * 16);
            row.block_offset = row.block_units != 0 ? offsets[g] : 0u;
            continue;
        }
        if (mode[g] == 2)
        {
            row = GroupLightRow{0, 0, ((row.generation >> 16) + 1u) << 16, 0, 0, 0, 0, 0};
            continue;
        }
        const uint32_t kept = uint32_t(plans[g].size());
        uint32_t* block = target + uint64_t(offsets[g]) * 4;
        const uint32_t* old = source + uint64_t(row.block_offset) * 4;
        uint32_t at = 1 + kept;  // record units from the block start
        uint32_t faces = 0;
        if (kept != 0)
        {
            block[0] = kept;
            block[1] = block[2] = block[3] = 0;
        }
        for (uint32_t r = 0; r < kept; ++r)
        {
            const Plan& p = plans[g][r];
            uint32_t* entry = block + 4 + uint64_t(r) * 4;
            entry[0] = p.pos;
            entry[1] = p.record;
            entry[2] = at;
            entry[3] = kNoneU;
            uint32_t* rec = block + uint64_t(at) * 4;
            if (p.old_record != kNoneU)
            {
                std::memcpy(rec, old + uint64_t(p.old_record) * 4, uint64_t(p.units) * 16);
                faces += rec[9];
                ++copied;
            }
            else
            {
                faces += RecomputeRecord(s, p.record, p.pos, shift, p.units, rec);
                ++recomputed;
            }
            at += p.units;
        }
        for (uint32_t b : members[g])
        {
            s->info[2 * uint64_t(b) + 1] &= ~kLtRecordStaleBit;
        }
        const uint32_t units = kept == 0 ? 0u : at;
        const uint32_t generation = (row.generation & ~0xFFFFu) | ((row.generation + 1u) & 0xFFFFu);
        row = GroupLightRow{units != 0 ? offsets[g] : 0u, units, generation, 0, 0,
                            kept | (std::min(dropped[g], 0xFFFFu) << 16), faces, 0};
    }
    // commit kernel (caelix_light_tree_commit.hlsl)
    uint64_t bricks = 0, faces = 0, dropped_total = 0;
    uint32_t groups = 0;
    for (uint32_t g = 0; g < G; ++g)
    {
        bricks += rows[g].bricks & 0xFFFFu;
        faces += rows[g].light_faces;
        dropped_total += rows[g].bricks >> 16;
        groups += (rows[g].bricks & 0xFFFFu) != 0 ? 1u : 0u;
    }
    control->current = 1u - current;
    control->light_groups = groups;
    control->bricks_dropped = uint32_t(std::min(dropped_total, uint64_t(0xFFFFFFFFu)));
    control->light_bricks = bricks;
    control->light_faces = faces;
    control->records_recomputed = recomputed;
    control->records_copied = copied;
    control->passes += 1;
}

// The gather of tree `tree` (caelix_light_gather_count.hlsl, caelix_light_bvh_scan.hlsl,
// caelix_light_gather_emit.hlsl:36-86), gated by state->building; entity_leaf_base (or null) numbers
// the bvh_leaf words per entity.
void oracle_light_gather(const CxLightPass* pass, uint32_t tree, LightBvhLeaf* leaves, uint32_t leaf_capacity,
                         LightTreeState* state, const uint32_t* entity_leaf_base)
{
    if (state->building == 0)
    {
        return;
    }
    const LightTreeDescriptor* d = pass->descriptor;
    const LightControl* control = reinterpret_cast<const LightControl*>(d->control);
    const GroupLightRow* rows = reinterpret_cast<const GroupLightRow*>(d->group_rows);
    const GroupInfo* infos = reinterpret_cast<const GroupInfo*>(d->group_infos);
    const EntityInfo* entities = reinterpret_cast<const EntityInfo*>(d->entity_infos);
    const uint32_t* motion = reinterpret_cast<const uint32_t*>(d->entity_motion);
    uint32_t* heap = reinterpret_cast<uint32_t*>(d->heap[control->current & 1u]);
    const uint32_t G = pass->group_count;
    std::vector<uint32_t> offsets(G + 1, 0);
    for (uint32_t g = 0; g < G; ++g)
    {
        uint32_t count = 0;
        if (rows[g].block_units != 0 && (rows[g].cpu_flags & kLtCpuReleased) == 0 &&
            (motion[infos[g].entity] & 1u) == tree)
        {
            count = heap[uint64_t(rows[g].block_offset) * 4];
        }
        offsets[g + 1] = offsets[g] + count;
    }
    state->leaves_found = offsets[G];
    for (uint32_t g = 0; g < G; ++g)
    {
        uint32_t* block = heap + uint64_t(rows[g].block_offset) * 4;
        const GroupInfo& info = infos[g];
        const int32_t edge = int32_t(8u << d->group_shift);
        const float origin[3] = {float(info.coord[0] * edge), float(info.coord[1] * edge), float(info.coord[2] * edge)};
        const uint32_t base = entity_leaf_base != nullptr ? entity_leaf_base[info.entity] : 0u;
        for (uint32_t k = 0; k < offsets[g + 1] - offsets[g]; ++k)
        {
            const uint32_t leaf = offsets[g] + k;
            uint32_t* entry = block + 4 + uint64_t(k) * 4;
            if (leaf >= leaf_capacity)
            {
                entry[3] = kNoneU;
                continue;
            }
            const uint32_t* rec = block + uint64_t(entry[2]) * 4;
            float face[6], lo[3], hi[3];
            std::memcpy(face, rec, 24);
            for (int c = 0; c < 3; ++c)
            {
                lo[c] = origin[c] + float((rec[6] >> (10 * c)) & 0x3FFu);
                hi[c] = origin[c] + float((rec[7] >> (10 * c)) & 0x3FFu);
            }
            LightBvhLeaf& out = leaves[leaf];
            out.summary = WorldSummaryOf(face, lo, hi, entities[info.entity].cur_rows);
            out.group = g;
            out.entry = k;
            out.pad = 0;
            out.parent = kNoneU;
            entry[3] = (tree << 31) | (leaf - base);
        }
    }
}

void oracle_trace_rays(const CxScene* s, const float* rays, uint32_t count, uint32_t* hits)
{
    for (uint32_t i = 0; i < count; ++i)
    {
        const float* r = rays + uint64_t(i) * 7;
        const WorldHit h = TraceWorld(*s, Mk(r[0], r[1], r[2]), Mk(r[3], r[4], r[5]), r[6]);
        uint32_t* o = hits + uint64_t(i) * 9;
        std::memcpy(&o[0], &h.t, 4);
        o[1] = h.attrib;
        o[2] = h.record;
        o[3] = uint32_t(h.brick[0]);
        o[4] = uint32_t(h.brick[1]);
        o[5] = uint32_t(h.brick[2]);
        o[6] = h.cells;
        o[7] = h.probes;
        o[8] = h.steps;
    }
}

void oracle_light_sample(const CxLightTreeView* view, const float* receivers, const uint32_t* seeds, uint32_t count,
                         float* samples, uint32_t* ok)
{
    const Lights L = OpenLights(*view);
    for (uint32_t i = 0; i < count; ++i)
    {
        const float* r = receivers + uint64_t(i) * 6;
        uint32_t rng = seeds[i];
        LightPoint p;
        const bool drawn = LightTreeSample(L, Mk(r[0], r[1], r[2]), Mk(r[3], r[4], r[5]), rng, p);
        float* o = samples + uint64_t(i) * 12;
        o[0] = p.position.x;
        o[1] = p.position.y;
        o[2] = p.position.z;
        o[3] = p.normal.x;
        o[4] = p.normal.y;
        o[5] = p.normal.z;
        o[6] = p.pdf;
        std::memcpy(&o[7], &p.block, 4);
        std::memcpy(&o[8], &p.face, 4);
        std::memcpy(&o[9], &p.group, 4);
        std::memcpy(&o[10], &p.entry, 4);
        std::memcpy(&o[11], &p.voxel_in_group, 4);
        ok[i] = drawn ? 1u : 0u;
    }
}

void oracle_light_point_pdf(const CxLightTreeView* view, const uint32_t* queries, uint32_t count, float* pdf)
{
    const Lights L = OpenLights(*view);
    for (uint32_t i = 0; i < count; ++i)
    {
        const uint32_t* q = queries + uint64_t(i) * 12;
        float f[6];
        std::memcpy(f, q, 24);
        const uint32_t v[3] = {q[7], q[8], q[9]};
        bool found;
        pdf[i] = LightPointPdf(L, Mk(f[0], f[1], f[2]), Mk(f[3], f[4], f[5]), q[6], v, q[10], q[11], found);
    }
}

void oracle_sky_cube_sample(const float* texels, uint32_t size, uint32_t mips, const float* queries, uint32_t count,
                            float* out)
{
    Sky k;
    k.cube = texels;
    k.cube_size = size;
    k.cube_mips = mips;
    for (uint32_t i = 0; i < count; ++i)
    {
        const float* q = queries + uint64_t(i) * 4;
        const V c = CubeSample(k, Mk(q[0], q[1], q[2]), q[3]);
        out[3 * i] = c.x;
        out[3 * i + 1] = c.y;
        out[3 * i + 2] = c.z;
    }
}

void oracle_sky_sh(const float* texels, uint32_t size, uint32_t mips, float* sh)
{
    Sky k;
    k.cube = texels;
    k.cube_size = size;
    k.cube_mips = mips;
    const float max_lod = float(mips - 1);
    ProjectSh(k, std::min(std::max(std::log2(float(size)) - 5.0f, 0.0f), max_lod));
    for (int b = 0; b < 9; ++b)
    {
        for (int ch = 0; ch < 3; ++ch)
        {
            sh[3 * b + ch] = k.sh[b][ch];
        }
    }
}

void oracle_light_sample_brick_box(const CxLightTreeView* view, const float* boxes, const uint32_t* seeds,
                                   uint32_t count, uint32_t* out)
{
    const Lights L = OpenLights(*view);
    for (uint32_t i = 0; i < count; ++i)
    {
        const float* b = boxes + uint64_t(i) * 9;
        const Recv r{Mk(b[0], b[1], b[2]), Mk(b[3], b[4], b[5]), Mk(b[6], b[7], b[8]), false};
        uint32_t rng = seeds[i];
        BrickRef brick;
        float pdf, importance;
        uint32_t leaf;
        const bool ok = SampleBrickBox(L, r, rng, brick, pdf, leaf, importance);
        uint32_t* o = out + uint64_t(i) * 8;
        o[0] = ok ? 1u : 0u;
        o[1] = brick.group;
        o[2] = brick.entry;
        o[3] = brick.generation;
        o[4] = leaf;
        std::memcpy(o + 5, &pdf, 4);
        std::memcpy(o + 6, &importance, 4);
        o[7] = 0u;
    }
}

void oracle_light_brick_pdf(const CxLightTreeView* view, const uint32_t* queries, uint32_t count, uint32_t box, float* pdf)
{
    const Lights L = OpenLights(*view);
    for (uint32_t i = 0; i < count; ++i)
    {
        const uint32_t* q = queries + uint64_t(i) * 12;
        float f[9];
        std::memcpy(f, q, sizeof(f));
        const V lo = Mk(f[0], f[1], f[2]);
        const Recv r{lo, box != 0u ? Mk(f[3], f[4], f[5]) : lo, Mk(f[6], f[7], f[8]), box == 0u};
        pdf[i] = BrickPdfR(L, r, BrickRef{q[9], q[10], q[11]});
    }
}

void oracle_light_box_support(const CxLightTreeView* view, const uint32_t* queries, uint32_t count, float* out)
{
    const Lights L = OpenLights(*view);
    for (uint32_t i = 0; i < count; ++i)
    {
        const uint32_t* q = queries + uint64_t(i) * 16;
        float f[12];
        std::memcpy(f, q, sizeof(f));
        const Recv r{Mk(f[0], f[1], f[2]), Mk(f[4], f[5], f[6]), Mk(f[8], f[9], f[10]), false};
        const uint32_t v[3] = {q[13] & 0x3FFu, (q[13] >> 10) & 0x3FFu, q[13] >> 20};
        BrickRef brick;
        uint32_t leaf;
        float in_brick;
        float* o = out + uint64_t(i) * 4;
        o[0] = o[1] = o[2] = o[3] = 0.0f;
        if (FindLightFace(L, Mk(f[3], f[7], f[11]), q[12], v, q[14], q[15], brick, leaf, in_brick))
        {
            o[0] = 1.0f;
            o[1] = leaf != kNoneU ? 1.0f : 0.0f;
            o[2] = BrickPdfR(L, r, brick);
            o[3] = LeafImportanceBox(L, brick, r);
        }
    }
}

void oracle_mis_weight(const float* queries, uint32_t count, float* weights)
{
    for (uint32_t i = 0; i < count; ++i)
    {
        const float* q = queries + uint64_t(i) * 9;
        const V n = Mk(q[3], q[4], q[5]);
        const V w = Mk(q[6], q[7], q[8]);
        weights[2 * i] = MisWeight(q[0], n, w, q[1], q[2], true);
        weights[2 * i + 1] = MisWeight(q[0], n, w, q[1], q[2], false);
    }
}

// SampleHenyeyGreenstein for `count` draws: queries [count][4] = (w xyz, g), stream i seeded with
// PcgHash(seed + i); out [count][3] the directions.
void oracle_sss_hg(const float* queries, uint32_t count, uint32_t seed, float* out)
{
    for (uint32_t i = 0; i < count; ++i)
    {
        const float* q = queries + uint64_t(i) * 4;
        uint32_t rng = PcgHash(seed + i);
        const V d = SampleHg(Mk(q[0], q[1], q[2]), q[3], rng);
        out[3 * uint64_t(i)] = d.x;
        out[3 * uint64_t(i) + 1] = d.y;
        out[3 * uint64_t(i) + 2] = d.z;
    }
}

// A frame over pixel rows [row_begin, row_end): color (rgb + t, float64 rgb) and radiance. counters:
// rays (primary + bounce), shadow rays, brick cells visited, hash probes, brick DDA steps, the demand
// bytes of every global load the kernels make for those pixels, and that total split by kind of work
// (primary, bounce and shadow traversal, primary-vertex and bounce-vertex shading).
void oracle_trace_frame(const CxScene* s, const CxLightTreeView* view, const CxCamera* cam, const CxTraceParams* prm,
                        uint32_t row_begin, uint32_t row_end, double* color, double* radiance_out, uint32_t* attrib_out,
                        uint64_t* counters)
{
    const Lights lights = OpenLights(*view);
    const Lights* L = &lights;
    uint64_t rays = 0, shadow = 0, cells = 0, probes = 0, steps = 0;
    uint64_t d0 = 0, d1 = 0, d2 = 0, d3 = 0, d4 = 0;
    Sky sky = SkyOf(*prm);
    sky.subsurface = (prm->flags & CX_TRACE_SUBSURFACE) != 0u && s->subsurface != nullptr;
    sky.glass = sky.subsurface || s->materials_ext != nullptr || prm->bounces_transparent > 0u ||
                (prm->flags & CX_TRACE_SHADOW_TRANSMISSION) != 0u;
    if (!sky.on)
    {
        sky.cube = nullptr;
    }
    if (sky.cube != nullptr)
    {
        // the irradiance SH of the cubemap, at the mip whose face is about 32 texels
        const float max_lod = float(sky.cube_mips - 1);
        ProjectSh(sky, std::min(std::max(std::log2(float(sky.cube_size)) - 5.0f, 0.0f), max_lod));
    }
    // The visibility hint: per pixel of the rows, the word read (sample 0) and the primary shadow rays
    // over the samples (sun traced, sun visible, lamp traced, lamp visible).
    const CxLightHint* hint = prm->hint;
    const uint64_t row_pixels = uint64_t(row_end - row_begin) * cam->width;
    std::vector<uint32_t> hint_words(hint != nullptr ? row_pixels : 0);
    std::vector<uint32_t> hint_counts(hint != nullptr ? row_pixels * 4 : 0);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : rays, shadow, cells, probes, steps, d0, d1, d2, d3, d4)
    for (int64_t py = row_begin; py < int64_t(row_end); ++py)
    {
        Counters local;
        for (int k = 0; k < 5; ++k)
        {
            g_demand[k] = 0;
        }
        for (uint32_t px = 0; px < cam->width; ++px)
        {
            DV col, rad;
            float t = 0.0f;
            uint32_t attrib = 0;
            const uint64_t hi = uint64_t(py - row_begin) * cam->width + px;
            for (uint32_t sample = 0; sample < prm->spp; ++sample)
            {
                DV r;
                V albedo, emission;
                HintIo io;
                io.params = hint;
                if (sky.glass)
                {
                    TracePixelGlass(*s, *L, *cam, prm->bounces, prm->frame_seed + sample, prm->flags, sky, px,
                                    uint32_t(py), r, albedo, emission, t, attrib, local, &io);
                }
                else
                {
                    TracePixel(*s, *L, *cam, prm->bounces, prm->frame_seed + sample, prm->flags, sky, px, uint32_t(py),
                               r, albedo, emission, t, attrib, local, nullptr, &io);
                }
                if (hint != nullptr)
                {
                    if (sample == 0)
                    {
                        hint_words[hi] = io.word;
                    }
                    const uint32_t base = io.traced == 1u ? 0u : 2u;
                    if (io.traced != 0u)
                    {
                        hint_counts[4 * hi + base] += 1u;
                        hint_counts[4 * hi + base + 1] += io.visible ? 1u : 0u;
                    }
                }
                col.x += double(emission.x) + double(albedo.x) * r.x;
                col.y += double(emission.y) + double(albedo.y) * r.y;
                col.z += double(emission.z) + double(albedo.z) * r.z;
                rad.x += r.x;
                rad.y += r.y;
                rad.z += r.z;
            }
            const uint64_t i = uint64_t(py - row_begin) * cam->width + px;
            color[4 * i] = col.x;
            color[4 * i + 1] = col.y;
            color[4 * i + 2] = col.z;
            color[4 * i + 3] = double(t);
            radiance_out[4 * i] = rad.x;
            radiance_out[4 * i + 1] = rad.y;
            radiance_out[4 * i + 2] = rad.z;
            radiance_out[4 * i + 3] = 0.0;
            attrib_out[i] = attrib;
        }
        rays += local.rays;
        shadow += local.shadow;
        cells += local.cells;
        probes += local.probes;
        steps += local.steps;
        d0 += g_demand[0];
        d1 += g_demand[1];
        d2 += g_demand[2];
        d3 += g_demand[3];
        d4 += g_demand[4];
    }
    if (hint != nullptr)
    {
        // WriteLightHint per patch whose rows lie in [row_begin, row_end): the value its first (top-left)
        // pixel read moved toward the fraction of its pixels' shadow rays that reached their light.
        const uint32_t shift = hint->patch_shift;
        const uint32_t columns = (cam->width + (1u << shift) - 1u) >> shift;
        for (uint32_t qy = (row_begin + (1u << shift) - 1u) >> shift; (qy << shift) < row_end; ++qy)
        {
            for (uint32_t qx = 0; qx < columns; ++qx)
            {
                uint32_t n[4] = {0, 0, 0, 0};
                for (uint32_t y = qy << shift; y < std::min(row_end, (qy + 1) << shift); ++y)
                {
                    for (uint32_t x = qx << shift; x < std::min(cam->width, (qx + 1) << shift); ++x)
                    {
                        for (int k = 0; k < 4; ++k)
                        {
                            n[k] += hint_counts[4 * (uint64_t(y - row_begin) * cam->width + x) + k];
                        }
                    }
                }
                const uint32_t first = hint_words[uint64_t((qy << shift) - row_begin) * cam->width + (qx << shift)];
                hint->current[uint64_t(qy) * columns + qx] =
                    uint16_t(HintChannel(first & 0xFFu, n[0], n[1]) | (HintChannel((first >> 8) & 0xFFu, n[2], n[3]) << 8));
            }
        }
        if (hint->read != nullptr)
        {
            for (uint64_t i = 0; i < row_pixels; ++i)
            {
                hint->read[uint64_t(row_begin) * cam->width + i] = uint16_t(hint_words[i]);
            }
        }
    }
    Counters c;
    c.rays = rays;
    c.shadow = shadow;
    c.cells = cells;
    c.probes = probes;
    c.steps = steps;
    counters[0] = c.rays;
    counters[1] = c.shadow;
    counters[2] = c.cells;
    counters[3] = c.probes;
    counters[4] = c.steps;
    counters[5] = d0 + d1 + d2 + d3 + d4;
    counters[6] = d0;
    counters[7] = d1;
    counters[8] = d2;
    counters[9] = d3;
    counters[10] = d4;
}

// The adjoint of oracle_trace_frame over pixel rows [row_begin, row_end): for d_color = dL/dI of the
// whole image (rgba per pixel, I = color / spp), adds dL/d(albedo of material id) to
// grad_albedo[id][3] and dL/d(emission scale of brick) to grad_scale[brick], in float64.
void oracle_trace_frame_adjoint(const CxScene* s, const CxLightTreeView* view, const CxCamera* cam,
                                const CxTraceParams* prm, uint32_t row_begin, uint32_t row_end, const float* d_color,
                                double* grad_albedo, double* grad_scale)
{
    const Lights lights = OpenLights(*view);
    const Lights* L = &lights;
    const int threads = omp_get_max_threads();
    std::vector<std::vector<double>> albedo_parts(threads, std::vector<double>(256 * 3, 0.0));
    std::vector<std::vector<double>> scale_parts(threads, std::vector<double>(s->bricks, 0.0));
    const Sky sky = SkyOf(*prm);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t py = row_begin; py < int64_t(row_end); ++py)
    {
        const int t = omp_get_thread_num();
        Counters local;
        for (uint32_t px = 0; px < cam->width; ++px)
        {
            const float* d = d_color + (uint64_t(py) * cam->width + px) * 4;
            Adjoint adj;
            for (int k = 0; k < 3; ++k)
            {
                adj.dcol[k] = double(d[k]) / double(prm->spp);
            }
            adj.grad_albedo = albedo_parts[t].data();
            adj.grad_scale = scale_parts[t].data();
            for (uint32_t sample = 0; sample < prm->spp; ++sample)
            {
                DV r;
                V albedo, emission;
                float th;
                uint32_t attrib;
                TracePixel(*s, *L, *cam, prm->bounces, prm->frame_seed + sample, prm->flags, sky, px, uint32_t(py), r,
                           albedo, emission, th, attrib, local, &adj);
            }
        }
    }
    for (int t = 0; t < threads; ++t)
    {
        for (int i = 0; i < 256 * 3; ++i)
        {
            grad_albedo[i] += albedo_parts[t][i];
        }
        for (uint32_t b = 0; b < s->bricks; ++b)
        {
            grad_scale[b] += scale_parts[t][b];
        }
    }
}

}  // extern "C"
