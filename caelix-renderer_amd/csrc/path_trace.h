// The path tracer of one pixel for gfx950 (shaders/render/caelix_path_trace.hlsl:276-452, 737-978;
// shaders/render/caelix_shading.hlsl:267-307), the subset the synthetic workloads use: opaque
// Lambertian voxels (metallic 0, smoothness 0), light NEE through the light tree with MIS on both
// sides, no sun, no sky light (a miss adds nothing), no fog, no glass, bounce NEE probability 1.
#pragma once
#include "light_sample.h"

namespace cx {

constexpr uint32_t kSegmentLampNee = 1u;
constexpr uint32_t kSegmentLampMis = 2u;

struct TraceParams
{
    uint32_t bounces;
    uint32_t frame_seed;
    uint32_t mis;
    uint32_t pad;
};

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
    diffuse = Dot(diffuse, diffuse) > 0.0f ? diffuse : normal;
    sh.bounce_origin = Add(hit_position, Mul(normal, kPushOff));
    sh.bounce_direction = diffuse;
    return sh;
}

struct PathCounters
{
    uint32_t rays;        // primary + bounce rays traced
    uint32_t shadow_rays;
    uint32_t cells;
    uint32_t probes;
    uint32_t steps;
};

CX_DEV WorldHit Trace(const SceneView& s, V3 o, V3 d, float t_max, PathCounters& c)
{
    const WorldHit h = TraceWorld(s, o, d, t_max);
    c.cells += h.cells;
    c.probes += h.probes;
    c.steps += h.steps;
    return h;
}

// NextEventAtVertex (:356-413) without the sun: the lamp candidate (LampCandidate, :301-333) and its
// shadow ray.
CX_DEV V3 NextEvent(const SceneView& s, const LightView& L, V3 x_r, V3 n_r, uint32_t& light_rng, PathCounters& c)
{
    const uint32_t point_seed = RandomU32(light_rng);
    V3 contribution = Make(0.0f, 0.0f, 0.0f);
    V3 direction = Make(0.0f, 0.0f, 0.0f);
    float t_max = 0.0f;
    uint32_t rng = point_seed;
    LightPoint p;
    if (LightTreeSample(s, L, x_r, n_r, rng, p) && p.pdf > 0.0f)
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
                contribution = Mul(BlockEmission(s, p.block), scale);
            }
        }
    }
    const float w_lamp = Luminance(contribution) * fmaxf(0.05f, 1.0f);
    const float w_sum = 0.0f + w_lamp;
    V3 radiance = Make(0.0f, 0.0f, 0.0f);
    if (w_sum > 0.0f)
    {
        c.shadow_rays += 1;
        const WorldHit h = t_max > 0.0f ? Trace(s, x_r, direction, t_max, c) : WorldHit{0.0f, 0u, kHashNone, {0, 0, 0}, 0, 0, 0};
        if (t_max > 0.0f && h.attrib == 0u)
        {
            radiance = Mul(contribution, Div(w_sum, w_lamp));
        }
    }
    return radiance;
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

struct PixelResult
{
    V3 radiance;  // the demodulated stochastic radiance (TracePath's `stochastic`)
    V3 albedo;    // of the primary hit
    V3 emission;  // of the primary hit
    float t;      // primary hit distance (0 = miss)
    uint32_t attrib;
};

// The main kernel's per-pixel work (:1050-) and TracePath for a primary ray.
CX_DEV PixelResult TracePixel(const SceneView& s, const LightView& L, const CameraView& cam, const TraceParams& prm,
                              uint32_t px, uint32_t py, PathCounters& c)
{
    PixelResult out;
    out.radiance = Make(0.0f, 0.0f, 0.0f);
    out.albedo = Make(0.0f, 0.0f, 0.0f);
    out.emission = Make(0.0f, 0.0f, 0.0f);
    out.t = 0.0f;
    out.attrib = 0u;
    // CaelixPrimaryDirection (caelix_frame.hlsl:164-171), pixel centre.
    const float sx = float(px) + 0.5f;
    const float sy = float(py) + 0.5f;
    const float ndc_x = Div(sx, float(cam.width)) * 2.0f - 1.0f;
    const float ndc_y = 1.0f - Div(sy, float(cam.height)) * 2.0f;
    const V3 right = Make(cam.right[0], cam.right[1], cam.right[2]);
    const V3 up = Make(cam.up[0], cam.up[1], cam.up[2]);
    const V3 forward = Make(cam.forward[0], cam.forward[1], cam.forward[2]);
    const V3 dir = SafeNormalize(Add(Add(Mul(right, ndc_x * cam.tan_half_fov_y * cam.aspect), Mul(up, ndc_y * cam.tan_half_fov_y)),
                                     forward));
    const V3 origin = Make(cam.position[0], cam.position[1], cam.position[2]);
    c.rays += 1;
    const WorldHit primary = Trace(s, origin, dir, kRayTMax, c);
    if (primary.attrib == 0u)
    {
        return out;  // the sky: black
    }
    out.t = primary.t;
    out.attrib = primary.attrib;
    uint32_t rng = InitRandom(px, py, prm.frame_seed);
    uint32_t light_rng = InitRandom(px, py, prm.frame_seed ^ 0x7F4A7C15u);
    const uint32_t mis_bit = L.mis != 0u ? kSegmentLampMis : 0u;
    uint32_t segment_light = 0u;
    const uint32_t surface_block = (primary.attrib >> 16) & 0xFFFFu;
    const V3 primary_normal = FaceNormal(primary.attrib);
    const Shade primary_shade = ShadeHit(s, surface_block, Add(origin, Mul(dir, primary.t)), primary_normal, rng);
    out.albedo = primary_shade.albedo;
    out.emission = primary_shade.emission;
    V3 radiance = Make(0.0f, 0.0f, 0.0f);
    V3 throughput = Make(1.0f, 1.0f, 1.0f);  // 1 / max(0.001, k), k = 1
    V3 segment_normal = primary_normal;
    if (prm.bounces >= 1u)
    {
        radiance = Add(radiance, MulV(NextEvent(s, L, primary_shade.bounce_origin, primary_normal, light_rng, c), throughput));
        segment_light = kSegmentLampNee | mis_bit;
    }
    V3 bounce_origin = primary_shade.bounce_origin;
    V3 bounce_direction = primary_shade.bounce_direction;
    uint32_t bounce_index = 1;
    while (bounce_index <= prm.bounces)
    {
        c.rays += 1;
        const WorldHit hit = Trace(s, bounce_origin, bounce_direction, kRayTMax, c);
        if (hit.attrib == 0u)
        {
            break;  // a miss sees the black sky
        }
        const uint32_t hit_block = (hit.attrib >> 16) & 0xFFFFu;
        const V3 vertex_normal = FaceNormal(hit.attrib);
        const Shade shade = ShadeHit(s, hit_block, Add(bounce_origin, Mul(bounce_direction, hit.t)), vertex_normal, rng);
        bounce_index += 1;
        float emission_weight = 1.0f;
        if (segment_light != 0u && (shade.emission.x > 0.0f || shade.emission.y > 0.0f || shade.emission.z > 0.0f))
        {
            emission_weight = LampHitWeight(s, L, segment_light, hit, bounce_origin, bounce_direction, segment_normal);
        }
        radiance = Add(radiance, MulV(shade.emission, Mul(throughput, emission_weight)));
        throughput = MulV(throughput, Make(Div(shade.albedo.x, 1.0f), Div(shade.albedo.y, 1.0f), Div(shade.albedo.z, 1.0f)));
        segment_light = 0u;
        if (bounce_index <= prm.bounces)
        {
            segment_light = kSegmentLampNee | mis_bit;
            radiance = Add(radiance, MulV(NextEvent(s, L, shade.bounce_origin, vertex_normal, light_rng, c),
                                          Mul(throughput, 1.0f)));
        }
        const float cp = fmaxf(throughput.x, fmaxf(throughput.y, throughput.z));
        if (cp <= 0.0f || cp < RandomFloat01(rng))
        {
            break;
        }
        throughput = Mul(throughput, Div(1.0f, cp));
        bounce_origin = shade.bounce_origin;
        bounce_direction = shade.bounce_direction;
        segment_normal = vertex_normal;
    }
    out.radiance = radiance;
    return out;
}

}  // namespace cx
