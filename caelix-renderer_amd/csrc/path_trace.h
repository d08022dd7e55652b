// The path tracer of one pixel for gfx950 (shaders/render/caelix_path_trace.hlsl:276-452, 737-978;
// shaders/render/caelix_shading.hlsl:267-307), the subset the synthetic workloads use: opaque
// Lambertian voxels (metallic 0, smoothness 0), light NEE through the light tree with MIS on both
// sides, bounce NEE probability 1, no fog, no glass; optionally (CX_TRACE_SKY) the procedural sky on
// misses and the sun, whose NEE shares the vertex's one shadow ray with the lamp (one-sample RIS);
// else a black sky and no sun.
#pragma once
#include "light_sample.h"

namespace cx {

constexpr uint32_t kSegmentLampNee = 1u;
constexpr uint32_t kSegmentLampMis = 2u;

// The sky and the sun (the frame constants sky_zenith/horizon/ground, sun_dir, sun_radiance,
// sun_cos_radius, sun_solid_angle of caelix_frame.hlsl).
struct SkyParams
{
    uint32_t on;   // the procedural sky on misses and the sun disk
    uint32_t sun;  // the sun's NEE runs: `on`, kRenderFlagSunNee and a sun radiance above 0 (caelix_path_trace.hlsl:742)
    V3 sun_dir;
    float sun_solid_angle;
    V3 sun_radiance;
    float sun_cos_radius;
    V3 zenith;
    V3 horizon;
    V3 ground;
    // The host's cubemap in place of the procedural sky (null: procedural): mips of 6 faces of
    // (size >> m)^2 rgba texels, mip-major, faces +X -X +Y -Y +Z -Z, rows from the top; and the
    // irradiance SH of it (9 float4) that the frame projects first.
    const float4* cube;
    uint32_t cube_size;
    uint32_t cube_mips;
    const float4* sh;
};

// The light NEE's visibility hint per pixel patch (kLightNeeHint, caelix_path_trace.hlsl:229-236,
// 338-344, 453-497): a word sun | lamp << 8 of 8-bit fractions k / 255 per patch of 1 << shift pixels.
constexpr float kHintFloor = 0.05f;
constexpr float kHintAlpha = 0.2f;
constexpr uint32_t kHintWordOff = 0xFFFFu;    // (1, 1): the choice weights unchanged
constexpr uint32_t kHintWordReset = 0x8080u;  // (128 / 255, 128 / 255): no history
constexpr uint32_t kNeeTracedNone = 0u;
constexpr uint32_t kNeeTracedSun = 1u;
constexpr uint32_t kNeeTracedLamp = 2u;

struct HintParams
{
    const uint16_t* previous;  // [patch rows][columns]; null: no hint
    uint32_t columns;          // patch columns: (width + patch - 1) >> shift
    uint32_t shift;
    uint32_t reset;            // kLightNeeHintReset: every primary vertex reads kHintWordReset
    uint32_t has_prev;         // a previous camera (prev) to reproject the primary surface into
    uint32_t sample;           // the frame's sample: its counts start at sample 0
    CameraView prev;
};

struct TraceParams
{
    uint32_t bounces;
    uint32_t frame_seed;
    uint32_t mis;
    uint32_t roulette;  // 1: Russian roulette (the reference's), 0: paths run all bounces
    SkyParams sky;
    V3 fog;             // the fog's radiance (fog_color.rgb)
    float fog_density;  // fog_color.w; 0: no fog
    uint32_t delta_budget;  // bounce_count_transparent: smooth interfaces the camera ray crosses
    uint32_t transmissive;  // kRenderFlagShadowTransmission: the primary vertex's shadow ray passes media
    uint32_t checkerboard;  // kRenderFlagDeltaCheckerboard: the chain's first interface split by pixel parity
    float nee_q;            // bounce_nee_probability (1: every bounce vertex runs its NEE)
    HintParams hint;        // the primary vertex's visibility hint (previous null: off)
    uint32_t subsurface;    // kRenderFlagSubsurface with a subsurface table: the walk at SSS vertices
};

// Whether the sun stream (nee_rng) is carried from vertex to vertex: the sun's NEE or the bounce NEE
// budget draws from it.
CX_DEV bool KeepsNeeStream(const TraceParams& p)
{
    return p.sky.sun != 0u || p.nee_q < 1.0f;
}

// TraverseSegment (caelix_shading.hlsl:150-175) in the air: a segment of length t that ends on a
// surface keeps exp(-density t) of the throughput and adds the fog's emission with the throughput
// before it.
CX_DEV void FogSegment(const TraceParams& p, float t, V3& throughput, V3& radiance)
{
    const float transmission = ExpCephes((-p.fog_density) * t);
    radiance = Add(radiance, Mul(MulV(throughput, p.fog), 1.0f - transmission));
    throughput = Mul(throughput, transmission);
}

CX_DEV float Saturate(float x)
{
    return fminf(fmaxf(x, 0.0f), 1.0f);
}
// HLSL lerp(a, b, t) = a + t (b - a).
CX_DEV V3 Lerp(V3 a, V3 b, float t)
{
    return Add(a, Mul(Sub(b, a), t));
}
// ProceduralSky (caelix_shading.hlsl:364-371): above the horizon lerp(horizon, zenith, sqrt(y)),
// below it lerp(horizon, ground, sqrt(-y)); no sun disk.
CX_DEV V3 ProceduralSky(const SkyParams& k, V3 d)
{
    if (d.y >= 0.0f)
    {
        return Lerp(k.horizon, k.zenith, Sqrt(Saturate(d.y)));
    }
    return Lerp(k.horizon, k.ground, Sqrt(Saturate(-d.y)));
}
// SunDisk (caelix_shading.hlsl:394-398).
CX_DEV V3 SunDisk(const SkyParams& k, V3 d)
{
    return Dot(d, k.sun_dir) > k.sun_cos_radius ? k.sun_radiance : Make(0.0f, 0.0f, 0.0f);
}
// ---- the cubemap sky -------------------------------------------------------------------------------
// Cube face selection (the Vulkan cube map rules; the inverse of the host's texel directions, the
// reference test's CubeTexelDirection): the axis of largest magnitude (x, then y, then z on ties),
// the face coordinates s, t in [0, 1] from (sc / |ma| + 1) / 2 with +X (-z, -y), -X (z, -y),
// +Y (x, z), -Y (x, -z), +Z (x, -y), -Z (-x, -y).
CX_DEV uint32_t CubeFace(V3 d, float& s, float& t)
{
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
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
    s = (Div(sc, ma) + 1.0f) * 0.5f;
    t = (Div(tc, ma) + 1.0f) * 0.5f;
    return face;
}

// The texels of mip m of the cube.
CX_DEV const float4* CubeMip(const SkyParams& k, uint32_t m)
{
    uint64_t offset = 0;
    for (uint32_t i = 0; i < m; ++i)
    {
        const uint64_t e = max(k.cube_size >> i, 1u);
        offset += 6 * e * e;
    }
    return k.cube + offset;
}

// Bilinear filtering inside one face of mip m (clamped to the face's edge texels).
CX_DEV V3 CubeBilinear(const SkyParams& k, uint32_t m, uint32_t face, float s, float t)
{
    const uint32_t e = max(k.cube_size >> m, 1u);
    const float4* texels = CubeMip(k, m) + uint64_t(face) * e * e;
    const float u = s * float(e) - 0.5f;
    const float v = t * float(e) - 0.5f;
    const float fu = floorf(u);
    const float fv = floorf(v);
    const float a = u - fu;
    const float b = v - fv;
    const int x0 = min(max(int(fu), 0), int(e) - 1);
    const int x1 = min(max(int(fu) + 1, 0), int(e) - 1);
    const int y0 = min(max(int(fv), 0), int(e) - 1);
    const int y1 = min(max(int(fv) + 1, 0), int(e) - 1);
    const float4 t00 = texels[uint64_t(y0) * e + x0];
    const float4 t10 = texels[uint64_t(y0) * e + x1];
    const float4 t01 = texels[uint64_t(y1) * e + x0];
    const float4 t11 = texels[uint64_t(y1) * e + x1];
    const V3 top = Lerp(Make(t00.x, t00.y, t00.z), Make(t10.x, t10.y, t10.z), a);
    const V3 bottom = Lerp(Make(t01.x, t01.y, t01.z), Make(t11.x, t11.y, t11.z), a);
    return Lerp(top, bottom, b);
}

// The cubemap in direction d at mip level `lod` (SampleLevel with a linear mipmapped sampler):
// the lod clamped to [0, mips - 1], bilinear in the two nearest mips, linear between them.
CX_DEV V3 CubeSample(const SkyParams& k, V3 d, float lod)
{
    float s, t;
    const uint32_t face = CubeFace(d, s, t);
    const float l = fminf(fmaxf(lod, 0.0f), float(k.cube_mips - 1));
    const uint32_t m0 = uint32_t(floorf(l));
    const uint32_t m1 = min(m0 + 1u, k.cube_mips - 1u);
    const float f = l - float(m0);
    const V3 c0 = CubeBilinear(k, m0, face, s, t);
    if (f == 0.0f)
    {
        return c0;
    }
    return Lerp(c0, CubeBilinear(k, m1, face, s, t), f);
}

// SkyRadianceNoSunLod (caelix_shading.hlsl:375-386): the cubemap at mip `lod`, else the procedural sky.
CX_DEV V3 SkyNoSun(const SkyParams& k, V3 d, float lod)
{
    return k.cube != nullptr ? CubeSample(k, d, lod) : ProceduralSky(k, d);
}

// CaelixShBasis (caelix_shading.hlsl:346-359): the 9 real SH of bands 0..2 at the unit direction d.
CX_DEV void ShBasis(V3 d, float (&basis)[9])
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

// SkyIrradianceOverPi (caelix_shading.hlsl:406-419): max(sum c_i Y_i(n), 0) / pi.
CX_DEV V3 SkyIrradianceOverPi(const SkyParams& k, V3 n)
{
    float basis[9];
    ShBasis(n, basis);
    V3 sum = Make(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int i = 0; i < 9; ++i)
    {
        const float4 c = k.sh[i];
        sum = Add(sum, Mul(Make(c.x, c.y, c.z), basis[i]));
    }
    return Mul(Make(fmaxf(sum.x, 0.0f), fmaxf(sum.y, 0.0f), fmaxf(sum.z, 0.0f)), 1.0f / kPi);
}

// SkyRadiance (:400-404): what a primary miss shows (the cubemap at mip 0).
CX_DEV V3 SkyRadiance(const SkyParams& k, V3 d)
{
    return Add(SkyNoSun(k, d, 0.0f), SunDisk(k, d));
}
// SkyDiffuseMiss (:428-437): a diffuse bounce ray from a vertex with normal n that misses: the cubemap's
// SH irradiance / pi at n, else the procedural sky in d; the sun disk when the segment sees the sun.
CX_DEV V3 SkyDiffuseMiss(const SkyParams& k, V3 n, V3 d, bool sees_sun)
{
    const V3 sun = sees_sun ? SunDisk(k, d) : Make(0.0f, 0.0f, 0.0f);
    return Add(k.cube != nullptr ? SkyIrradianceOverPi(k, n) : ProceduralSky(k, d), sun);
}
// SkySpecularMiss (:443-447): a specular bounce ray of a lobe with `roughness`: the cubemap at mip
// roughness * (mips - 1); the sun disk when the segment sees the sun.
CX_DEV V3 SkySpecularMiss(const SkyParams& k, V3 d, float roughness, bool sees_sun)
{
    const V3 sun = sees_sun ? SunDisk(k, d) : Make(0.0f, 0.0f, 0.0f);
    return Add(SkyNoSun(k, d, roughness * float(k.cube_mips > 0 ? k.cube_mips - 1 : 0)), sun);
}
// A diffuse bounce ray of the opaque path misses: every segment leaves a vertex that ran the NEE
// (segment_sees_sun = !nee, caelix_path_trace.hlsl:809, 903), so it sees the sun disk exactly when the
// sun's NEE is off.
CX_DEV V3 SkyMiss(const SkyParams& k, V3 n, V3 d)
{
    return SkyDiffuseMiss(k, n, d, k.sun == 0u);
}
// SampleSunDirection (:467-485): uniform in the sun cone, cos theta = 1 - u1 (1 - cos r), phi =
// 2 pi u2, 1 - cos r from the solid angle; the Duff et al. 2017 basis around sun_dir. sin / cos: the
// Cephes polynomials of RandomUnitVector.
CX_DEV V3 SampleSunDirection(const SkyParams& k, uint32_t& rng)
{
    const float one_minus_cos_r = k.sun_solid_angle * (1.0f / kTwoPi);
    const float a = RandomFloat01(rng) * one_minus_cos_r;
    const float phi = RandomFloat01(rng) * kTwoPi;
    const float cos_theta = 1.0f - a;
    const float sin_theta = Sqrt(fmaxf(a * (2.0f - a), 0.0f));
    const V3 w = k.sun_dir;
    const float sign_z = w.z >= 0.0f ? 1.0f : -1.0f;
    const float p = Div(-1.0f, sign_z + w.z);
    const float q = w.x * w.y * p;
    const V3 t = Make(1.0f + sign_z * w.x * w.x * p, sign_z * q, -sign_z * w.x);
    const V3 b = Make(q, sign_z + w.y * w.y * p, -w.y);
    float sp, cp;
    SinCos(phi, sp, cp);
    return SafeNormalize(Add(Add(Mul(t, cp * sin_theta), Mul(b, sp * sin_theta)), Mul(w, cos_theta)));
}
// SunCandidate (caelix_path_trace.hlsl:276-290): two numbers of `rng`; the demodulated radiance of
// the Lambert lobe (0 below the horizon of `normal`).
CX_DEV V3 SunCandidate(const SkyParams& k, V3 normal, uint32_t& rng, V3& direction)
{
    direction = SampleSunDirection(k, rng);
    const float cos_l = Dot(normal, direction);
    if (!(cos_l > 0.0f))
    {
        return Make(0.0f, 0.0f, 0.0f);
    }
    return Mul(k.sun_radiance, k.sun_solid_angle * cos_l * (1.0f / kPi));
}

struct Shade
{
    V3 albedo;
    V3 emission;
    V3 bounce_origin;
    V3 bounce_direction;
};

CX_DEV V3 BlockAlbedo(const SceneView& s, uint32_t block)
{
    const float4 m = s.materials[2 * uint64_t(block)];
    return Make(m.x, m.y, m.z);
}
CX_DEV V3 BlockEmission(const SceneView& s, uint32_t block)
{
    const float4 m = s.materials[2 * uint64_t(block) + 1];
    return Mul(Make(m.x, m.y, m.z), s.emission_scale);
}
CX_DEV float Luminance(V3 c)
{
    return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f;
}
CX_DEV V3 FaceNormal(uint32_t attrib)
{
    const uint32_t f = uint32_t(__ffs(int(attrib & 0x3Fu)) - 1);
    return FaceDirection(f);
}

// ShadeVoxelHit, opaque branch: the lobe number (specular chance 0), then the cosine direction.
CX_DEV Shade ShadeHit(const SceneView& s, uint32_t block, V3 hit_position, V3 normal, uint32_t& rng)
{
    Shade sh;
    sh.emission = BlockEmission(s, block);
    sh.albedo = BlockAlbedo(s, block);
    (void)RandomFloat01(rng);  // the lobe choice against specular chance 0
    V3 diffuse = SafeNormalize(Add(normal, RandomUnitVector(rng)));
    diffuse = Select(Dot(diffuse, diffuse) > 0.0f, diffuse, normal);
    sh.bounce_origin = Add(hit_position, Mul(normal, kPushOff));
    sh.bounce_direction = diffuse;
    return sh;
}


// ---- the full material model (a scene with the extended material table) -----------------------
CX_DEV bool BlockOpaque(uint32_t block, uint32_t encoding)
{
    return (block & (encoding == kEncodingPacked ? kPackedOpaqueMask : kBlockIdOpaqueMask)) != 0u;
}
// CaelixMaterialCode (caelix_material.hlsl:65-70): an opaque value as it is, a transparent one
// without its face bits (its medium id; 0 = air).
CX_DEV uint32_t BlockCode(uint32_t block, uint32_t encoding)
{
    return BlockOpaque(block, encoding) ? block : (block & (encoding == kEncodingPacked ? kPackedIdMask : kBlockIdIdMask));
}
struct Material
{
    V3 albedo;
    float smoothness;
    V3 emission;
    float metallic;
    float ior;
    float extinction;
};
// CaelixGetMaterial (caelix_material.hlsl:84-100) of a table index; without the extended table
// ior 1.5 and extinction 0 (the reference's default entries).
CX_DEV Material GetMaterial(const SceneView& s, uint32_t code)
{
    const float4 a = s.materials[2 * uint64_t(code)];
    const float4 e = s.materials[2 * uint64_t(code) + 1];
    Material m;
    m.albedo = Make(a.x, a.y, a.z);
    m.smoothness = a.w;
    m.emission = Make(e.x, e.y, e.z);
    m.metallic = e.w;
    m.ior = 1.5f;
    m.extinction = 0.0f;
    if (s.materials_ext != nullptr)
    {
        const float4 x = s.materials_ext[code];
        m.ior = x.x;
        m.extinction = x.y;
    }
    return m;
}
// HLSL reflect(i, n) = i - 2 dot(n, i) n and refract(i, n, eta) (0 past the critical angle).
CX_DEV V3 Reflect(V3 i, V3 n)
{
    return Sub(i, Mul(n, 2.0f * Dot(n, i)));
}
CX_DEV V3 Refract(V3 i, V3 n, float eta)
{
    const float c = Dot(n, i);
    const float k = 1.0f - eta * eta * (1.0f - c * c);
    if (k < 0.0f)
    {
        return Make(0.0f, 0.0f, 0.0f);
    }
    return Sub(Mul(i, eta), Mul(n, eta * c + Sqrt(k)));
}
// FresnelReflectAmountOpaque / FresnelReflectAmountTransparent (caelix_shading.hlsl:116-148): Schlick,
// with the transmitted cosine when n1 > n2 and 1 past the critical angle.
CX_DEV float Schlick(float r0, float cos_x)
{
    const float x = 1.0f - cos_x;
    const float xx = x * x;
    return r0 + (1.0f - r0) * xx * xx * x;
}
CX_DEV float FresnelOpaque(float n1, float n2, V3 incident, V3 normal)
{
    float r0 = Div(n1 - n2, n1 + n2);
    r0 *= r0;
    return Schlick(r0, -Dot(normal, incident));
}
CX_DEV float FresnelTransparent(float n1, float n2, V3 incident, V3 normal)
{
    float r0 = Div(n1 - n2, n1 + n2);
    r0 *= r0;
    float cos_x = -Dot(normal, incident);
    if (n1 > n2)
    {
        const float n = Div(n1, n2);
        const float sin_t2 = n * n * (1.0f - cos_x * cos_x);
        if (sin_t2 >= 1.0f)
        {
            return 1.0f;
        }
        cos_x = Sqrt(1.0f - sin_t2);
    }
    return Schlick(r0, cos_x);
}
struct Interface
{
    V3 normal;  // against the incoming ray
    V3 direction;
    bool refracted;
    uint32_t medium;  // after the interface
    float weight;     // 1, or the split's 2 F / 2 (1 - F)
    bool split_applied;
};
// CrossInterface (caelix_shading.hlsl:203-243). split < 0: refraction when it exists and a number of
// `rng` >= the Fresnel reflectance (the number is drawn only when refraction exists); split 0 / 1: the
// checkerboard's given branch (reflect / refract), no number, weight 2 F / 2 (1 - F); total internal
// reflection is never split.
CX_DEV Interface CrossInterface(const SceneView& s, uint32_t source, uint32_t destination, float destination_ior, V3 dir,
                                V3 normal, uint32_t& rng, int split = -1)
{
    Interface it;
    it.normal = SafeNormalize(normal);
    if (Dot(dir, it.normal) > 0.0f)
    {
        it.normal = Make(-it.normal.x, -it.normal.y, -it.normal.z);
    }
    const float source_ior = source != 0u ? GetMaterial(s, source).ior : 1.0f;
    const float target_ior = destination != 0u ? destination_ior : 1.0f;
    const float eta = Div(source_ior, target_ior);
    const V3 reflection = Reflect(dir, it.normal);
    const V3 refraction = Refract(dir, it.normal, eta);
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
struct ShadeFull
{
    V3 albedo;
    V3 emission;
    V3 bounce_origin;
    V3 bounce_direction;
    bool specular;    // the bounce took the specular lobe (always at a transparent block): no NEE
    float roughness;  // the lobe's roughness: 1 - smoothness, 0 at a transparent block
};
// ShadeVoxelHit (caelix_shading.hlsl:267-307). Transparent block: the interface from `medium` into
// the block's medium, albedo 1, origin pushed to the side of the chosen direction, `medium` updated.
// Opaque block: specular chance lerp(metallic, 1, fresnel smoothness) against the lobe number, the
// cosine direction, the specular direction lerp(diffuse, reflect, smoothness), the chosen direction
// lerp(diffuse, specular, 0 or 1). HLSL lerp(a, b, t) is taken as a + t (b - a).
CX_DEV ShadeFull ShadeVoxel(const SceneView& s, uint32_t block, V3 hit_position, V3 dir, V3 normal, uint32_t& rng,
                            uint32_t& medium)
{
    const uint32_t code = BlockCode(block, s.encoding);
    const Material m = GetMaterial(s, code);
    ShadeFull sh;
    sh.emission = Mul(m.emission, s.emission_scale);
    if (!BlockOpaque(block, s.encoding))
    {
        const Interface it = CrossInterface(s, medium, code, m.ior, dir, normal, rng);
        sh.albedo = Make(1.0f, 1.0f, 1.0f);
        sh.bounce_origin = Add(hit_position, Mul(it.normal, it.refracted ? -kPushOff : kPushOff));
        sh.bounce_direction = it.direction;
        sh.specular = true;
        sh.roughness = 0.0f;
        medium = it.medium;
        return sh;
    }
    sh.albedo = m.albedo;
    const float fresnel = FresnelOpaque(1.0f, m.ior, dir, normal);
    const float chance = m.metallic + fresnel * m.smoothness * (1.0f - m.metallic);
    const float do_specular = RandomFloat01(rng) < chance ? 1.0f : 0.0f;
    V3 diffuse = SafeNormalize(Add(normal, RandomUnitVector(rng)));
    diffuse = Select(Dot(diffuse, diffuse) > 0.0f, diffuse, normal);
    const V3 specular = SafeNormalize(Lerp(diffuse, Reflect(dir, normal), m.smoothness));
    sh.bounce_origin = Add(hit_position, Mul(normal, kPushOff));
    sh.bounce_direction = Lerp(diffuse, specular, do_specular);
    sh.specular = do_specular > 0.5f;
    sh.roughness = 1.0f - m.smoothness;
    return sh;
}
// ---- subsurface scattering (shaders/render/caelix_subsurface.hlsl, docs/SUBSURFACE.md) ---------------
constexpr float kSubsurfaceMaxAnisotropy = 0.99f;
constexpr float kSubsurfaceIsotropicG = 1e-3f;
constexpr uint32_t kSubsurfaceNone = 0;  // not an SSS vertex (or the last vertex): the shade stands
constexpr uint32_t kSubsurfaceBsdf = 1;  // an SSS material whose draw chose the surface BSDF
constexpr uint32_t kSubsurfaceExit = 2;  // the walk left the medium: the shade is the exit vertex
constexpr uint32_t kSubsurfaceDead = 3;  // the walk reached its step cap: the path ends at this vertex

// SubsurfaceVoxel (caelix_scene_trace.hlsl, CAELIX_SUBSURFACE): an opaque block whose code's entry has
// model HG, probability > 0 and steps > 0; its walk albedo 1 - one_minus_albedo.
CX_DEV bool SubsurfaceVoxel(const SceneView& s, uint32_t block, V3& albedo)
{
    if (!BlockOpaque(block, s.encoding))
    {
        return false;
    }
    const CxSubsurface& e = s.subsurface[BlockCode(block, s.encoding)];
    if (e.model != CX_SSS_MODEL_HG || !(e.probability > 0.0f) || (e.steps & 0xFFu) == 0u)
    {
        return false;
    }
    albedo = Make(1.0f - e.one_minus_albedo[0], 1.0f - e.one_minus_albedo[1], 1.0f - e.one_minus_albedo[2]);
    return true;
}

// SampleHenyeyGreenstein (:62-85): the HG cosine by inversion (|g| < 1e-3: the uniform sphere), the
// azimuth around w in the Duff et al. basis of OrthonormalBasis (caelix_shading.hlsl:454-461). Two numbers.
CX_DEV V3 SampleHenyeyGreenstein(V3 w, float g, uint32_t& rng)
{
    const float u1 = RandomFloat01(rng);
    const float u2 = RandomFloat01(rng);
    float cos_theta;
    if (fabsf(g) < kSubsurfaceIsotropicG)
    {
        cos_theta = 1.0f - 2.0f * u1;
    }
    else
    {
        const float sq = Div(1.0f - g * g, 1.0f - g + 2.0f * g * u1);
        cos_theta = Div(1.0f + g * g - sq * sq, 2.0f * g);
    }
    cos_theta = fminf(fmaxf(cos_theta, -1.0f), 1.0f);
    const float sin_theta = Sqrt(fmaxf(1.0f - cos_theta * cos_theta, 0.0f));
    const float phi = kTwoPi * u2;
    const float sign_z = w.z >= 0.0f ? 1.0f : -1.0f;
    const float p = Div(-1.0f, sign_z + w.z);
    const float q = w.x * w.y * p;
    const V3 t = Make(1.0f + sign_z * w.x * w.x * p, sign_z * q, -sign_z * w.x);
    const V3 b = Make(q, sign_z + w.y * w.y * p, -w.y);
    float sp, cp;
    SinCos(phi, sp, cp);
    return Add(Add(Mul(t, cp * sin_theta), Mul(b, sp * sin_theta)), Mul(w, cos_theta));
}

// The voxel cursor of the walk (caelix_voxel_cursor.hlsl): a voxel of the scene's grid (world voxel
// coordinates: the one entity sits at the identity frame) and its brick record, kHashNone for a brick
// that is not resident (its voxels are air).
struct VoxelCursor
{
    int32_t v[3];
    uint32_t record;
};

// VoxelCursorFromHit: the voxel a hit's attrib names in the brick `record`.
CX_DEV VoxelCursor CursorFromHit(const SceneView& s, uint32_t record, uint32_t attrib)
{
    const int32_t* k = BrickXyz(s, record);
    const uint32_t cell = (attrib >> 6) & 0x1FFu;
    VoxelCursor c;
    c.v[0] = k[0] * 8 + int32_t(cell & 7u);
    c.v[1] = k[1] * 8 + int32_t((cell >> 3) & 7u);
    c.v[2] = k[2] * 8 + int32_t(cell >> 6);
    c.record = record;
    return c;
}

// VoxelCursorStep: the cursor moved by (dx, dy, dz), each in -1..1; a new brick is found through the
// hash table.
CX_DEV void CursorStep(const SceneView& s, uint32_t scene, VoxelCursor& c, int32_t dx, int32_t dy, int32_t dz)
{
    const int32_t n[3] = {c.v[0] + dx, c.v[1] + dy, c.v[2] + dz};
    if ((n[0] >> 3) != (c.v[0] >> 3) || (n[1] >> 3) != (c.v[1] >> 3) || (n[2] >> 3) != (c.v[2] >> 3))
    {
        c.record = HashGet(s.slots, s.slot_count, s.pow2, s.hash_info[0],
                           SceneBrickFingerprint(s, scene, n[0] >> 3, n[1] >> 3, n[2] >> 3));
    }
    c.v[0] = n[0];
    c.v[1] = n[1];
    c.v[2] = n[2];
}

// VoxelCursorBlock: the cursor voxel's block value (0, air, in a brick that is not resident).
CX_DEV uint32_t CursorBlock(const SceneView& s, const VoxelCursor& c)
{
    if (c.record == kHashNone)
    {
        return 0u;
    }
    return s.payload[uint64_t(c.record) * 512 + uint32_t(c.v[0] & 7) + 8u * uint32_t(c.v[1] & 7) + 64u * uint32_t(c.v[2] & 7)];
}

// SubsurfaceWalk (:95-190): from the SSS face hit (record, attrib, hit point `hit_point`) of a ray with
// direction `ray_dir`, with the entry material's g, step cap and albedo. Entry jitter: the voxel of
// hit_point - 0.5 n + a uniform offset in [-0.5, 0.5] on the face's tangent axes, reached by one cursor
// step of at most one voxel per tangent axis, if it is an SSS voxel (else the hit voxel). Per step: the
// weight takes the current voxel's albedo, the direction a new HG draw, then the face neighbour along
// its dominant axis (ties x, y, z): an SSS voxel is entered, another opaque voxel is a wall (the cursor
// stays), anything else is the exit: a uniform point on that face of the cursor voxel, pushed off along
// the face normal, a cosine direction around it. False at the step cap (the walk is absorbed). Two
// numbers for the jitter, two per step, four at the exit.
CX_DEV bool SubsurfaceWalk(const SceneView& s, uint32_t scene, uint32_t record, uint32_t attrib, V3 hit_point,
                           V3 ray_dir, float g, uint32_t steps, V3 entry_albedo, uint32_t& rng, V3& origin, V3& normal,
                           V3& direction, V3& weight)
{
    origin = Make(0.0f, 0.0f, 0.0f);
    normal = Make(0.0f, 0.0f, 0.0f);
    direction = Make(0.0f, 0.0f, 0.0f);
    weight = Make(0.0f, 0.0f, 0.0f);
    VoxelCursor c = CursorFromHit(s, record, attrib);
    V3 a = entry_albedo;
    {
        const V3 fn = FaceNormal(attrib);
        const float u1 = RandomFloat01(rng) - 0.5f;
        const float u2 = RandomFloat01(rng) - 0.5f;
        const V3 offset = fn.x != 0.0f ? Make(0.0f, u1, u2) : (fn.y != 0.0f ? Make(u1, 0.0f, u2) : Make(u1, u2, 0.0f));
        const V3 target = Add(Sub(hit_point, Mul(fn, 0.5f)), offset);
        const int32_t tv[3] = {int32_t(floorf(target.x)), int32_t(floorf(target.y)), int32_t(floorf(target.z))};
        const float fv[3] = {fn.x, fn.y, fn.z};
        int32_t delta[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            delta[k] = fv[k] == 0.0f ? min(max(tv[k] - c.v[k], -1), 1) : 0;
        }
        if (delta[0] != 0 || delta[1] != 0 || delta[2] != 0)
        {
            VoxelCursor start = c;
            CursorStep(s, scene, start, delta[0], delta[1], delta[2]);
            V3 start_albedo;
            if (SubsurfaceVoxel(s, CursorBlock(s, start), start_albedo))
            {
                c = start;
                a = start_albedo;
            }
        }
    }
    V3 d = ray_dir;
    V3 w = Make(1.0f, 1.0f, 1.0f);
    for (uint32_t i = 0; i < steps; ++i)
    {
        w = MulV(w, a);
        d = SampleHenyeyGreenstein(d, g, rng);
        const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
        int32_t dx = 0, dy = 0, dz = 0;
        if (ax >= ay && ax >= az)
        {
            dx = d.x >= 0.0f ? 1 : -1;
        }
        else if (ay >= az)
        {
            dy = d.y >= 0.0f ? 1 : -1;
        }
        else
        {
            dz = d.z >= 0.0f ? 1 : -1;
        }
        VoxelCursor next = c;
        CursorStep(s, scene, next, dx, dy, dz);
        const uint32_t next_block = CursorBlock(s, next);
        if (BlockOpaque(next_block, s.encoding))
        {
            V3 next_albedo;
            if (SubsurfaceVoxel(s, next_block, next_albedo))
            {
                c = next;
                a = next_albedo;
            }
            continue;
        }
        const V3 n = Make(float(dx), float(dy), float(dz));
        const float u1 = RandomFloat01(rng) - 0.5f;
        const float u2 = RandomFloat01(rng) - 0.5f;
        const V3 offset = dx != 0 ? Make(0.0f, u1, u2) : (dy != 0 ? Make(u1, 0.0f, u2) : Make(u1, u2, 0.0f));
        const V3 exit_point =
            Add(Add(Add(Make(float(c.v[0]), float(c.v[1]), float(c.v[2])), Make(0.5f, 0.5f, 0.5f)), Mul(n, 0.5f)), offset);
        normal = n;
        origin = Add(exit_point, Mul(normal, kPushOff));
        direction = SafeNormalize(Add(normal, RandomUnitVector(rng)));
        direction = Select(Dot(direction, direction) > 0.0f, direction, normal);
        weight = w;
        return true;
    }
    return false;
}

// SubsurfaceVertex (:192-250): the subsurface choice at a vertex (the hit `record`, `attrib`, hit point,
// ray direction) after ShadeVoxel gave `shade`. Not an SSS vertex (last_vertex: no next segment; a
// transparent block; the entry's model, probability or steps): kSubsurfaceNone, nothing drawn.
// Otherwise guide = lerp(albedo, sss_albedo, p) and one number u: u >= p keeps the shade (Bsdf); else
// the walk: its exit replaces the vertex (Exit: albedo = the walk's weight, the exit's bounce origin,
// direction and normal, the diffuse lobe, roughness 1) or the walk died (Dead: albedo 0).
CX_DEV uint32_t SubsurfaceVertex(const SceneView& s, uint32_t scene, uint32_t record, uint32_t attrib, V3 hit_point,
                                 V3 ray_dir, uint32_t& rng, ShadeFull& shade, V3& vertex_normal, V3& guide_albedo,
                                 bool last_vertex)
{
    guide_albedo = shade.albedo;
    const uint32_t block = (attrib >> 16) & 0xFFFFu;
    if (last_vertex || !BlockOpaque(block, s.encoding))
    {
        return kSubsurfaceNone;
    }
    const CxSubsurface& e = s.subsurface[BlockCode(block, s.encoding)];
    const float probability = e.probability;
    if (e.model != CX_SSS_MODEL_HG || !(probability > 0.0f))
    {
        return kSubsurfaceNone;
    }
    const uint32_t steps = e.steps & 0xFFu;
    if (steps == 0u)
    {
        return kSubsurfaceNone;
    }
    const V3 entry_albedo = Make(1.0f - e.one_minus_albedo[0], 1.0f - e.one_minus_albedo[1], 1.0f - e.one_minus_albedo[2]);
    guide_albedo = Lerp(shade.albedo, entry_albedo, probability);
    if (!(RandomFloat01(rng) < probability))
    {
        return kSubsurfaceBsdf;
    }
    const float g = fminf(fmaxf(e.anisotropy, -kSubsurfaceMaxAnisotropy), kSubsurfaceMaxAnisotropy);
    V3 origin, normal, direction, weight;
    if (!SubsurfaceWalk(s, scene, record, attrib, hit_point, ray_dir, g, steps, entry_albedo, rng, origin, normal,
                        direction, weight))
    {
        shade.albedo = Make(0.0f, 0.0f, 0.0f);
        return kSubsurfaceDead;
    }
    shade.albedo = weight;
    shade.bounce_origin = origin;
    shade.bounce_direction = direction;
    shade.specular = false;
    shade.roughness = 1.0f;
    vertex_normal = normal;
    return kSubsurfaceExit;
}

// The face index (FaceDirection) of an axis-aligned unit normal.
CX_DEV uint32_t FaceOfNormal(V3 n)
{
    return n.x != 0.0f ? (n.x > 0.0f ? 0u : 1u) : (n.y != 0.0f ? (n.y > 0.0f ? 2u : 3u) : (n.z > 0.0f ? 4u : 5u));
}

// TraverseSegment (caelix_shading.hlsl:150-175): in the air the fog; inside a medium the throughput
// keeps exp(-(1 - albedo) t extinction) per channel.
CX_DEV void MediumSegment(const SceneView& s, const TraceParams& p, uint32_t medium, float t, V3& throughput,
                          V3& radiance)
{
    if (medium == 0u)
    {
        if (p.fog_density > 0.0f)
        {
            FogSegment(p, t, throughput, radiance);
        }
        return;
    }
    const Material m = GetMaterial(s, medium);
    const V3 transmission = Make(ExpCephes((-(1.0f - m.albedo.x)) * t * m.extinction),
                                 ExpCephes((-(1.0f - m.albedo.y)) * t * m.extinction),
                                 ExpCephes((-(1.0f - m.albedo.z)) * t * m.extinction));
    throughput = MulV(throughput, transmission);
}
// The sky a segment that misses receives (caelix_path_trace.hlsl:841-846): SkyDiffuseMiss at the
// segment's normal after a diffuse vertex, SkySpecularMiss of its lobe's roughness after a specular one
// (roughness < 0 marks a diffuse segment).
CX_DEV V3 SkySegmentMiss(const SkyParams& k, V3 n, V3 d, float roughness, bool sees_sun)
{
    return roughness < 0.0f ? SkyDiffuseMiss(k, n, d, sees_sun) : SkySpecularMiss(k, d, roughness, sees_sun);
}

CX_DEV WorldHit Trace(const SceneView& s, V3 o, V3 d, float t_max, PathCounters& c, uint32_t scene = 0)
{
    const WorldHit h = TraceWorld(s, o, d, t_max, scene);
    c.cells += h.cells;
    c.probes += h.probes;
    c.steps += h.steps;
    return h;
}

// TraceShadowTransmissive (caelix_scene_trace.hlsl:248-266, 401-477) without subsurface: up to 8 hops;
// an interface voxel is entered (the origin pushed through along the hit normal, the remaining length
// keeping the ray's end), the segments inside media keep exp(-(1 - albedo) t extinction) (the air: 1,
// no fog), an opaque voxel or a transmittance below 1e-3 in every channel occludes, a remaining length
// of 1e29 or more is unbounded (a miss inside a medium adds no segment).
CX_DEV V3 ShadowGlassTransmittance(const SceneView& s, uint32_t medium, float t)
{
    if (medium == 0u)
    {
        return Make(1.0f, 1.0f, 1.0f);
    }
    const Material m = GetMaterial(s, medium);
    return Make(ExpCephes((-(1.0f - m.albedo.x)) * t * m.extinction), ExpCephes((-(1.0f - m.albedo.y)) * t * m.extinction),
                ExpCephes((-(1.0f - m.albedo.z)) * t * m.extinction));
}
CX_DEV bool ShadowOccluded(V3 t)
{
    return fmaxf(t.x, fmaxf(t.y, t.z)) < 1e-3f;
}
CX_DEV V3 ShadowTransmissive(const SceneView& s, V3 origin, V3 dir, float t_max, uint32_t medium, PathCounters& c,
                             uint32_t scene)
{
    V3 transmittance = Make(1.0f, 1.0f, 1.0f);
    uint32_t cur = medium;
    float remaining = t_max;
    const V3 none = Make(0.0f, 0.0f, 0.0f);
    for (uint32_t hop = 0; hop < 8u; ++hop)
    {
        const WorldHit h = Trace(s, origin, dir, remaining, c, scene);
        if (h.attrib == 0u)
        {
            if (cur != 0u && remaining < 1e29f)
            {
                transmittance = MulV(transmittance, ShadowGlassTransmittance(s, cur, remaining));
            }
            return Select(ShadowOccluded(transmittance), none, transmittance);
        }
        transmittance = MulV(transmittance, ShadowGlassTransmittance(s, cur, h.t));
        if (ShadowOccluded(transmittance))
        {
            return none;
        }
        remaining -= h.t;
        const uint32_t block = (h.attrib >> 16) & 0xFFFFu;
        if (BlockOpaque(block, s.encoding))
        {
            return none;
        }
        cur = BlockCode(block, s.encoding);
        const V3 offset = Mul(FaceNormal(h.attrib), -kPushOff);
        origin = Add(Add(origin, Mul(dir, h.t)), offset);
        remaining -= Dot(offset, dir);
    }
    return none;
}

// NextEventAtVertex (:356-413): the lamp candidate (LampCandidate, :301-333) and, with the sun's NEE,
// the sun candidate (two numbers of nee_rng) and the choice (a third), kept in proportion to their
// unoccluded luminance; the kept sun candidate is drawn again from its saved state. Returns false when
// no shadow ray is counted (both weights 0); else `direction` and `t_max` give the shadow ray from
// x_r (t_max 0: counted but not traced) and `visible` the radiance it adds when it reaches its light
// unoccluded. `unscaled` is `visible` without the light brick's emission scale and `light_record` that
// brick (for the adjoint, which runs without the sun).
CX_DEV bool NeeCandidate(const SceneView& s, const LightView& L, const SkyParams& k, V3 x_r, V3 n_r,
                         uint32_t& light_rng, uint32_t& nee_rng, V3& direction, float& t_max, V3& visible, V3& unscaled,
                         uint32_t& light_record, uint32_t hint_word = kHintWordOff, uint32_t* traced = nullptr)
{
    // UnpackHint (caelix_path_trace.hlsl:465-468): the visibility hint that multiplies the choice
    // weights, floored at kHintFloor; kHintWordOff gives exactly 1.
    const float hint_sun = float(hint_word & 0xFFu) * (1.0f / 255.0f);
    const float hint_lamp = float((hint_word >> 8) & 0xFFu) * (1.0f / 255.0f);
    if (traced != nullptr)
    {
        *traced = kNeeTracedNone;
    }
    const uint32_t point_seed = RandomU32(light_rng);
    V3 contribution = Make(0.0f, 0.0f, 0.0f);
    direction = Make(0.0f, 0.0f, 0.0f);
    t_max = 0.0f;
    uint32_t rng = point_seed;
    unscaled = Make(0.0f, 0.0f, 0.0f);
    light_record = kHashNone;
    LightPoint p;
    if (LightTreeSample(L, x_r, n_r, rng, p) && p.pdf > 0.0f)
    {
        const V3 d = Sub(p.position, x_r);
        const float d2 = Dot(d, d);
        const float distance = Sqrt(d2);
        if (distance >= kNearField)
        {
            const V3 w = Make(Div(d.x, distance), Div(d.y, distance), Div(d.z, distance));
            const float cos_r = Dot(n_r, w);
            const float cos_l = -Dot(p.normal, w);
            if (cos_r > 0.0f && cos_l > 0.0f)
            {
                float scale = Div((cos_r * (1.0f / kPi)) * Div(cos_l, d2), p.pdf);
                if (L.mis != 0u)
                {
                    scale *= MisWeight(p.pdf, n_r, w, distance, cos_l, true);
                }
                direction = w;
                t_max = distance - 2.0f * kPushOff;
                unscaled = Mul(BlockEmission(s, p.block), scale);
                contribution = Mul(Mul(BlockEmission(s, p.block), BrickScale(s, p.record)), scale);
                light_record = p.record;
            }
        }
    }
    const float w_lamp = Luminance(contribution) * fmaxf(kHintFloor, hint_lamp);
    const uint32_t sun_start = nee_rng;
    float w_sun = 0.0f;
    float u_choice = 0.0f;
    if (k.sun != 0u)
    {
        V3 sun_direction;
        w_sun = Luminance(SunCandidate(k, n_r, nee_rng, sun_direction)) * fmaxf(kHintFloor, hint_sun);
        u_choice = RandomFloat01(nee_rng);
    }
    const float w_sum = w_sun + w_lamp;
    if (!(w_sum > 0.0f))
    {
        return false;
    }
    if (u_choice * w_sum < w_sun)
    {
        uint32_t replay = sun_start;
        const V3 sun = SunCandidate(k, n_r, replay, direction);
        t_max = kRayTMax;
        visible = Mul(sun, Div(w_sum, w_sun));
        unscaled = Make(0.0f, 0.0f, 0.0f);
        light_record = kHashNone;
        if (traced != nullptr)
        {
            *traced = kNeeTracedSun;
        }
        return true;
    }
    visible = Mul(contribution, Div(w_sum, w_lamp));
    if (traced != nullptr)
    {
        *traced = kNeeTracedLamp;
    }
    return true;
}

// LampHitWeight (:425-452).
CX_DEV float LampHitWeight(const SceneView& s, const LightView& L, uint32_t segment_light, const WorldHit& hit,
                           V3 origin, V3 direction, V3 segment_normal)
{
    if (segment_light == 0u || hit.t < kNearField)
    {
        return 1.0f;
    }
    const V3 n = FaceNormal(hit.attrib);
    const V3 hp = Add(origin, Mul(direction, hit.t));
    const V3 inside = Sub(hp, Mul(n, 0.25f));
    const int32_t voxel[3] = {int32_t(floorf(inside.x)), int32_t(floorf(inside.y)), int32_t(floorf(inside.z))};
    const uint32_t face = uint32_t(__ffs(int(hit.attrib & 0x3Fu)) - 1);
    const uint32_t block = (hit.attrib >> 16) & 0xFFFFu;
    bool found;
    const float p_area = LightFacePdf(s, L, origin, segment_normal, hit.record, voxel, face, block, found);
    if (!found)
    {
        return 1.0f;
    }
    if ((segment_light & kSegmentLampMis) == 0u)
    {
        return 0.0f;
    }
    const float cos_l = -Dot(n, direction);
    return MisWeight(p_area, segment_normal, direction, hit.t, cos_l, false);
}

// PreviousPixel (caelix_path_trace.hlsl:154-172): the pixel (current image size) where the previous
// camera saw the world point p; false when p was not in front of it.
CX_DEV bool PreviousPixel(const CameraView& prev, uint32_t width, uint32_t height, V3 p, float& x, float& y)
{
    const V3 rel = Sub(p, Make(prev.position[0], prev.position[1], prev.position[2]));
    const float z = Dot(rel, Make(prev.forward[0], prev.forward[1], prev.forward[2]));
    x = 0.0f;
    y = 0.0f;
    if (!(z > 1e-6f))
    {
        return false;
    }
    const float ndc_x = Div(Dot(rel, Make(prev.right[0], prev.right[1], prev.right[2])),
                            (z * prev.tan_half_fov_y) * prev.aspect);
    const float ndc_y = Div(Dot(rel, Make(prev.up[0], prev.up[1], prev.up[2])), z * prev.tan_half_fov_y);
    x = ((ndc_x + 1.0f) * 0.5f) * float(width);
    y = ((1.0f - ndc_y) * 0.5f) * float(height);
    return fabsf(x) < 1e30f && fabsf(y) < 1e30f;
}

// ReadLightHint (caelix_path_trace.hlsl:479-498) for the in-image pixel (px, py): kHintWordReset with
// the reset; else the previous hint at the patch of the previous pixel of the primary surface point p
// (identity entity frames: the point did not move), or at the pixel's own patch when there is no
// surface, no previous camera, or p was behind it or outside its image.
CX_DEV uint32_t ReadHint(const HintParams& h, uint32_t width, uint32_t height, uint32_t px, uint32_t py, bool surface,
                         V3 p)
{
    if (h.reset != 0u)
    {
        return kHintWordReset;
    }
    uint32_t qx = px >> h.shift;
    uint32_t qy = py >> h.shift;
    if (surface && h.has_prev != 0u)
    {
        float x, y;
        if (PreviousPixel(h.prev, width, height, p, x, y) && x >= 0.0f && y >= 0.0f && x < float(width) &&
            y < float(height))
        {
            qx = uint32_t(x) >> h.shift;
            qy = uint32_t(y) >> h.shift;
        }
    }
    return h.previous[uint64_t(qy) * h.columns + qx];
}

// One channel of WriteLightHint (caelix_path_trace.hlsl:554-580) in the word's 8 bits: the read value
// moved by kHintAlpha toward visible / traced (kept without a traced ray), rounded as PackHint does.
CX_DEV uint32_t HintChannel(uint32_t k, uint32_t traced, uint32_t visible)
{
    if (traced == 0u)
    {
        return k;
    }
    const float h = float(k) * (1.0f / 255.0f);
    const float cur = h + kHintAlpha * (Div(float(visible), float(traced)) - h);
    return uint32_t(fminf(fmaxf(cur, 0.0f), 1.0f) * 255.0f + 0.5f);
}

// CaelixPrimaryDirection (caelix_frame.hlsl:164-171): the pinhole ray through the pixel centre.
CX_DEV V3 PrimaryDirection(const CameraView& cam, uint32_t px, uint32_t py)
{
    const float sx = float(px) + 0.5f;
    const float sy = float(py) + 0.5f;
    const float ndc_x = Div(sx, float(cam.width)) * 2.0f - 1.0f;
    const float ndc_y = 1.0f - Div(sy, float(cam.height)) * 2.0f;
    const V3 right = Make(cam.right[0], cam.right[1], cam.right[2]);
    const V3 up = Make(cam.up[0], cam.up[1], cam.up[2]);
    const V3 forward = Make(cam.forward[0], cam.forward[1], cam.forward[2]);
    return SafeNormalize(Add(Add(Mul(right, ndc_x * cam.tan_half_fov_y * cam.aspect), Mul(up, ndc_y * cam.tan_half_fov_y)),
                             forward));
}

CX_DEV uint32_t FaceIndex(uint32_t attrib)
{
    return uint32_t(__ffs(int(attrib & 0x3Fu)) - 1);
}

}  // namespace cx
