// Trace kernels for gfx950: a frame of primary rays + light NEE path tracing, batch ray queries, and
// batch light-tree sample / pdf evaluations (the reference's probe-kernel pattern,
// tests/shaders/caelix_light_tree_probe.hlsl, as C entry points).
#include "path_trace.h"
#include "runtime.h"

namespace cx {
namespace {

SceneView SceneOf(const CxScene& c)
{
    SceneView s;
    s.keys = c.keys;
    s.info = reinterpret_cast<const uint2*>(c.info);
    s.occupancy = static_cast<const uint64_t*>(c.occupancy);
    s.payload = c.payload;
    s.slots = c.slots;
    s.slot_count = c.slot_count;
    s.pow2 = (c.slot_count & (c.slot_count - 1)) == 0 ? 1u : 0u;
    s.hash_info = c.hash_info;
    s.brick_bits = c.brick_bits;
    s.super_bits = c.super_bits;
    for (int k = 0; k < 3; ++k)
    {
        s.grid_min[k] = c.grid_min[k];
    }
    s.grid_dim = c.grid_dim;
    s.materials = reinterpret_cast<const float4*>(c.materials);
    s.light_codes = c.light_codes;
    s.emission_scale = c.emission_scale;
    s.encoding = c.encoding;
    return s;
}

LightView LightsOf(const CxLightTreeView& v)
{
    LightView l;
    l.headers = v.headers;
    l.nodes[0] = v.nodes[0];
    l.nodes[1] = v.nodes[1];
    l.leaves[0] = v.leaves[0];
    l.leaves[1] = v.leaves[1];
    l.heap = v.heap;
    l.record_offset = v.record_offset;
    l.ordinal_of_brick = v.ordinal_of_brick;
    l.group_shift = v.group_shift;
    l.mis = v.mis;
    return l;
}

CameraView CameraOf(const CxCamera& c)
{
    CameraView v;
    for (int k = 0; k < 3; ++k)
    {
        v.position[k] = c.position[k];
        v.forward[k] = c.forward[k];
        v.right[k] = c.right[k];
        v.up[k] = c.up[k];
    }
    v.tan_half_fov_y = c.tan_half_fov_y;
    v.aspect = c.aspect;
    v.width = c.width;
    v.height = c.height;
    return v;
}

struct FrameOut
{
    float4* color;
    float4* radiance;
    uint32_t* attrib;
    unsigned long long* counters;
};

CX_DEV void AddCounter(unsigned long long* c, uint32_t v)
{
    const uint64_t sum = WaveSum(uint64_t(v));
    if (Lane() == 0 && sum != 0)
    {
        atomicAdd(c, static_cast<unsigned long long>(sum));
    }
}

// One wave per 8x8 pixel tile. The path state and the world DDA need ~174 VGPRs unconstrained (2
// waves per SIMD); capping at 5 waves per SIMD (96 VGPRs, some spilled to scratch) hides more of
// the dependent hash/record/occupancy load latency: 1.6x the rays per second of the uncapped
// kernel at 1080p, more than the 4- and 8-wave caps.
constexpr int kFrameWavesPerSimd = 5;

__global__ void __launch_bounds__(64, kFrameWavesPerSimd) FrameKernel(SceneView s, LightView l, CameraView cam, TraceParams prm,
                                                  uint32_t spp, FrameOut out)
{
    const uint32_t px = blockIdx.x * 8 + (threadIdx.x & 7u);
    const uint32_t py = blockIdx.y * 8 + (threadIdx.x >> 3);
    const bool inside = px < cam.width && py < cam.height;
    PathCounters c{0, 0, 0, 0, 0};
    if (inside)
    {
        V3 color = Make(0.0f, 0.0f, 0.0f);
        V3 radiance = Make(0.0f, 0.0f, 0.0f);
        float t = 0.0f;
        uint32_t attrib = 0;
        for (uint32_t sample = 0; sample < spp; ++sample)
        {
            TraceParams p = prm;
            p.frame_seed = prm.frame_seed + sample;
            const PixelResult r = TracePixel(s, l, cam, p, px, py, c);
            color = Add(color, Add(r.emission, MulV(r.albedo, r.radiance)));
            radiance = Add(radiance, r.radiance);
            t = r.t;
            attrib = r.attrib;
        }
        const uint64_t pixel = uint64_t(py) * cam.width + px;
        if (out.color != nullptr)
        {
            out.color[pixel] = make_float4(color.x, color.y, color.z, t);
        }
        if (out.radiance != nullptr)
        {
            out.radiance[pixel] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
        }
        if (out.attrib != nullptr)
        {
            out.attrib[pixel] = attrib;
        }
    }
    AddCounter(&out.counters[0], c.rays);
    AddCounter(&out.counters[1], c.shadow_rays);
    AddCounter(&out.counters[2], c.cells);
    AddCounter(&out.counters[3], c.probes);
    AddCounter(&out.counters[4], c.steps);
}

__global__ void __launch_bounds__(256) RaysKernel(SceneView s, const float* rays, uint32_t count, uint32_t* hits)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
    {
        return;
    }
    const float* r = rays + uint64_t(i) * 7;
    const WorldHit h = TraceWorld(s, Make(r[0], r[1], r[2]), Make(r[3], r[4], r[5]), r[6]);
    uint32_t* o = hits + uint64_t(i) * 9;
    o[0] = __float_as_uint(h.t);
    o[1] = h.attrib;
    o[2] = h.record;
    o[3] = uint32_t(h.brick[0]);
    o[4] = uint32_t(h.brick[1]);
    o[5] = uint32_t(h.brick[2]);
    o[6] = h.cells;
    o[7] = h.probes;
    o[8] = h.steps;
}

__global__ void __launch_bounds__(256) SampleKernel(SceneView s, LightView l, const float* receivers,
                                                    const uint32_t* seeds, uint32_t count, float* samples,
                                                    uint32_t* ok)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
    {
        return;
    }
    const float* r = receivers + uint64_t(i) * 6;
    uint32_t rng = seeds[i];
    LightPoint p;
    const bool drawn = LightTreeSample(s, l, Make(r[0], r[1], r[2]), Make(r[3], r[4], r[5]), rng, p);
    float* o = samples + uint64_t(i) * 9;
    o[0] = p.position.x;
    o[1] = p.position.y;
    o[2] = p.position.z;
    o[3] = p.normal.x;
    o[4] = p.normal.y;
    o[5] = p.normal.z;
    o[6] = p.pdf;
    o[7] = __uint_as_float(p.block);
    o[8] = __uint_as_float(p.face);
    ok[i] = drawn ? 1u : 0u;
}

__global__ void __launch_bounds__(256) FacePdfKernel(SceneView s, LightView l, const uint32_t* queries,
                                                     uint32_t count, float* pdf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
    {
        return;
    }
    const uint32_t* q = queries + uint64_t(i) * 12;
    const V3 x = Make(__uint_as_float(q[0]), __uint_as_float(q[1]), __uint_as_float(q[2]));
    const V3 n = Make(__uint_as_float(q[3]), __uint_as_float(q[4]), __uint_as_float(q[5]));
    const int32_t voxel[3] = {int32_t(q[7]), int32_t(q[8]), int32_t(q[9])};
    bool found;
    pdf[i] = LightFacePdf(s, l, x, n, q[6], voxel, q[10], q[11], found);
}

int CheckScene(const CxScene* s, const char* what)
{
    if (s == nullptr || s->keys == nullptr || s->info == nullptr || s->occupancy == nullptr || s->payload == nullptr ||
        s->slots == nullptr || s->hash_info == nullptr || s->brick_bits == nullptr || s->super_bits == nullptr ||
        s->materials == nullptr)
    {
        return ArgError("%s: the scene has a NULL buffer", what);
    }
    if (s->grid_dim == 0 || s->grid_dim % 8 != 0)
    {
        return ArgError("%s: grid_dim %u is not a positive multiple of 8", what, s->grid_dim);
    }
    return CX_OK;
}

int CheckLights(const CxLightTreeView* l, const char* what)
{
    if (l == nullptr || l->headers == nullptr || l->nodes[0] == nullptr || l->leaves[0] == nullptr || l->heap == nullptr ||
        l->record_offset == nullptr || l->ordinal_of_brick == nullptr)
    {
        return ArgError("%s: the light tree view has a NULL buffer", what);
    }
    return CX_OK;
}

}  // namespace
}  // namespace cx

extern "C" {

CX_API int cx_trace_frame(const CxScene* scene, const CxLightTreeView* lights, const CxCamera* camera,
                          const CxTraceParams* params, const CxTraceOutput* out, void* stream)
{
    if (int r = cx::CheckScene(scene, "cx_trace_frame"); r != CX_OK)
    {
        return r;
    }
    if (int r = cx::CheckLights(lights, "cx_trace_frame"); r != CX_OK)
    {
        return r;
    }
    if (scene->light_codes == nullptr || camera == nullptr || params == nullptr || out == nullptr || out->counters == nullptr)
    {
        return cx::ArgError("cx_trace_frame: NULL light codes, camera, params or counters");
    }
    if (camera->width == 0 || camera->height == 0 || params->spp == 0)
    {
        return cx::ArgError("cx_trace_frame: empty image or 0 samples per pixel");
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::TraceParams p{params->bounces, params->frame_seed, lights->mis, 0};
    cx::FrameOut o{reinterpret_cast<float4*>(out->color), reinterpret_cast<float4*>(out->radiance), out->attrib,
                   reinterpret_cast<unsigned long long*>(out->counters)};
    const dim3 grid(cx::DivUp(camera->width, 8), cx::DivUp(camera->height, 8));
    cx::StageScope scope(s, "trace_frame");
    cx::FrameKernel<<<grid, 64, 0, s>>>(cx::SceneOf(*scene), cx::LightsOf(*lights), cx::CameraOf(*camera), p,
                                        params->spp, o);
    return cx::LaunchResult("cx_trace_frame");
}

CX_API int cx_trace_rays(const CxScene* scene, const float* rays, uint32_t count, uint32_t* hits, void* stream)
{
    if (int r = cx::CheckScene(scene, "cx_trace_rays"); r != CX_OK)
    {
        return r;
    }
    if (rays == nullptr || hits == nullptr)
    {
        return cx::ArgError("cx_trace_rays: NULL rays or hits");
    }
    if (count == 0)
    {
        return CX_OK;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::StageScope scope(s, "trace_rays");
    cx::RaysKernel<<<cx::DivUp(count, 256), 256, 0, s>>>(cx::SceneOf(*scene), rays, count, hits);
    return cx::LaunchResult("cx_trace_rays");
}

CX_API int cx_light_sample(const CxScene* scene, const CxLightTreeView* lights, const float* receivers,
                           const uint32_t* seeds, uint32_t count, float* samples, uint32_t* ok, void* stream)
{
    if (int r = cx::CheckLights(lights, "cx_light_sample"); r != CX_OK)
    {
        return r;
    }
    if (scene == nullptr || scene->keys == nullptr || receivers == nullptr || seeds == nullptr || samples == nullptr ||
        ok == nullptr)
    {
        return cx::ArgError("cx_light_sample: a buffer is NULL");
    }
    if (count == 0)
    {
        return CX_OK;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::StageScope scope(s, "light_sample");
    cx::SampleKernel<<<cx::DivUp(count, 256), 256, 0, s>>>(cx::SceneOf(*scene), cx::LightsOf(*lights), receivers,
                                                           seeds, count, samples, ok);
    return cx::LaunchResult("cx_light_sample");
}

CX_API int cx_light_face_pdf(const CxScene* scene, const CxLightTreeView* lights, const uint32_t* queries,
                             uint32_t count, float* pdf, void* stream)
{
    if (int r = cx::CheckLights(lights, "cx_light_face_pdf"); r != CX_OK)
    {
        return r;
    }
    if (scene == nullptr || scene->keys == nullptr || scene->light_codes == nullptr || queries == nullptr ||
        pdf == nullptr)
    {
        return cx::ArgError("cx_light_face_pdf: a buffer is NULL");
    }
    if (count == 0)
    {
        return CX_OK;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::StageScope scope(s, "light_face_pdf");
    cx::FacePdfKernel<<<cx::DivUp(count, 256), 256, 0, s>>>(cx::SceneOf(*scene), cx::LightsOf(*lights), queries,
                                                            count, pdf);
    return cx::LaunchResult("cx_light_face_pdf");
}

}  // extern "C"
