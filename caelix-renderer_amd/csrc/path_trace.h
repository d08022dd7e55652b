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
    uint32_t roulette;  // 1: Russian roulette (the reference's), 0: paths run all bounces
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


CX_DEV WorldHit Trace(const SceneView& s, V3 o, V3 d, float t_max, PathCounters& c, uint32_t scene = 0)
{
    const WorldHit h = TraceWorld(s, o, d, t_max, scene);
    c.cells += h.cells;
    c.probes += h.probes;
    c.steps += h.steps;
    return h;
}

// NextEventAtVertex (:356-413) without the sun: the lamp candidate (LampCandidate, :301-333). Returns
// false when no shadow ray is counted (the candidate's weight is 0); else `direction` and `t_max`
// give the shadow ray from x_r (t_max 0: counted but not traced) and `visible` the radiance it adds
// when it reaches the light unoccluded. `unscaled` is `visible` without the light brick's emission
// scale and `light_record` that brick (for the adjoint).
CX_DEV bool NeeCandidate(const SceneView& s, const LightView& L, V3 x_r, V3 n_r, uint32_t& light_rng, V3& direction,
                         float& t_max, V3& visible, V3& unscaled, uint32_t& light_record)
{
    const uint32_t point_seed = RandomU32(light_rng);
    V3 contribution = Make(0.0f, 0.0f, 0.0f);
    direction = Make(0.0f, 0.0f, 0.0f);
    t_max = 0.0f;
    uint32_t rng = point_seed;
    unscaled = Make(0.0f, 0.0f, 0.0f);
    light_record = kHashNone;
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
                unscaled = Mul(BlockEmission(s, p.block), scale);
                contribution = Mul(Mul(BlockEmission(s, p.block), BrickScale(s, p.record)), scale);
                light_record = p.record;
            }
        }
    }
    const float w_lamp = Luminance(contribution) * fmaxf(0.05f, 1.0f);
    const float w_sum = 0.0f + w_lamp;
    if (!(w_sum > 0.0f))
    {
        return false;
    }
    visible = Mul(contribution, Div(w_sum, w_lamp));
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
