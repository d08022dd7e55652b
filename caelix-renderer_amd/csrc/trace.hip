// Trace kernels for gfx950: a frame of primary rays + light NEE path tracing, batch ray queries, and
// batch light-tree sample / pdf evaluations (the reference's probe-kernel pattern,
// tests/shaders/caelix_light_tree_probe.hlsl, as C entry points).
#include "path_trace.h"
#include "runtime.h"

#include <cmath>
#include <vector>

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
    s.box = c.occupied_box;
    for (int k = 0; k < 3; ++k)
    {
        s.grid_min[k] = c.grid_min[k];
    }
    s.grid_dim = c.grid_dim;
    s.materials = reinterpret_cast<const float4*>(c.materials);
    s.materials_ext = reinterpret_cast<const float4*>(c.materials_ext);
    s.light_codes = c.light_codes;
    s.emission_scale = c.emission_scale;
    s.encoding = c.encoding;
    s.brick_scale = c.brick_emission_scale;
    s.key_words = c.scene_count > 1 ? 4u : 3u;
    s.scene_cells = uint64_t(c.grid_dim) * c.grid_dim * c.grid_dim;
    s.scene_supers = s.scene_cells / 512;
    s.group_of_brick = c.group_of_brick;
    s.subsurface = c.subsurface;
    return s;
}

LightRef LightsOf(const CxLightTreeView& v)
{
    LightRef r;
    r.descriptor = v.descriptor;
    r.mis = v.mis;
    r.select = v.select;
    return r;
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

CX_DEV void FlushCounters(unsigned long long* counters, const PathCounters& c)
{
    AddCounter(&counters[0], c.rays);
    AddCounter(&counters[1], c.shadow_rays);
    AddCounter(&counters[2], c.cells);
    AddCounter(&counters[3], c.probes);
    AddCounter(&counters[4], c.steps);
}

// ------------------------------------------------------------------------------------------------
// The frame as a wavefront: per sample, PrimaryTrace -> PrimaryShade -> Shadow, then per bounce
// Extend -> BounceShade -> Shadow, then Finish. Each path (one per pixel, numbered so that a wave's
// 64 paths are one 8x8 pixel tile) keeps its state in the workspace between launches; the shade
// stages append the paths that go on and their shadow rays to queues (wave-aggregated atomic
// appends), so the trace launches run only live rays and hold only the traversal's registers. Each
// path does exactly the operations of TracePath (caelix_path_trace.hlsl:737-978) in the same order:
// the radiance terms are added in path order (NEE at vertex 0, emission at vertex 1, NEE at vertex
// 1, ...), because every shadow launch completes before the next bounce's shade launch.

constexpr uint32_t kWaveThreads = 256;



struct Workspace
{
    float4* ray_o;      // [paths] next ray origin
    float4* ray_d;      // [paths] next ray direction
    uint4* hit;         // [paths] t bits, attrib, brick record, -
    float4* throughput; // [paths]
    float4* radiance;   // [paths] this sample's radiance
    uint4* state;       // [paths] rng, light rng, segment light | bounce << 8 | segment face << 16, -
    float4* albedo;     // [paths] primary albedo, t
    float4* emission;   // [paths] primary emission, attrib bits
    float4* shadow_o;   // [paths] shadow ray origin, t_max
    float4* shadow_d;   // [paths] shadow ray direction, path bits
    float4* shadow_add; // [paths] radiance the shadow ray adds when unoccluded, light record bits
    float4* adjoint;    // [paths] adjoint: dL/d(radiance term) (rgb)
    float4* shadow_grad; // [paths] adjoint: the shadow ray's radiance without the emission scale, roulette
    uint32_t* queue[2]; // [paths] path queues (ping-pong by bounce parity)
    uint32_t* nee_rng;  // [paths] the sun's random stream, kept whenever something draws from it: the
                        // sun's NEE, or the bounce NEE budget below 1 (KeepsNeeStream)
    uint32_t* hint_word;  // [paths] the visibility hint the primary vertex read (sample 0) | traced << 16
    uint2* hint_counts;   // [paths] primary shadow rays over the samples: sun traced | visible << 16, lamp
    float4* sky_sh;       // [9] the irradiance SH of the frame's cubemap sky
    uint32_t* ctl;      // [0], [1] queue lengths, [2] shadow queue length
    float* partials;    // [kGradRows][768] adjoint: per-workgroup albedo gradient sums
};

// Albedo gradients: 256 materials x rgb, summed per workgroup in LDS and then in workgroup rows of
// the workspace (no global atomics on the few hot materials); one final kernel sums the rows.
constexpr uint32_t kGradEntries = 768;
constexpr uint32_t kGradRowsMin = 4096;

struct GradView
{
    const float4* d_color;  // null: the forward frame
    float* grad_albedo;     // [256][4]
    float* grad_scale;      // [bricks]
    float inv_spp;
};

uint32_t TilesX(uint32_t width)
{
    return (width + 7) / 8;
}
uint32_t PathCount(uint32_t width, uint32_t height)
{
    return TilesX(width) * ((height + 7) / 8) * 64;
}
uint64_t Align256(uint64_t x)
{
    return (x + 255) & ~uint64_t(255);
}
uint32_t GradRows(uint32_t paths)
{
    return std::max(kGradRowsMin, (paths + 255) / 256);
}
uint64_t WorkspaceBytes(uint32_t paths)
{
    return 13 * Align256(uint64_t(paths) * 16) + 4 * Align256(uint64_t(paths) * 4) + Align256(uint64_t(paths) * 8) + 512 +
           Align256(uint64_t(GradRows(paths)) * kGradEntries * 4);
}
Workspace CarveWorkspace(void* base, uint32_t paths)
{
    uint8_t* p = static_cast<uint8_t*>(base);
    auto take = [&](uint64_t bytes) {
        uint8_t* r = p;
        p += Align256(bytes);
        return r;
    };
    Workspace w;
    const uint64_t v = uint64_t(paths) * 16;
    w.ray_o = reinterpret_cast<float4*>(take(v));
    w.ray_d = reinterpret_cast<float4*>(take(v));
    w.hit = reinterpret_cast<uint4*>(take(v));
    w.throughput = reinterpret_cast<float4*>(take(v));
    w.radiance = reinterpret_cast<float4*>(take(v));
    w.state = reinterpret_cast<uint4*>(take(v));
    w.albedo = reinterpret_cast<float4*>(take(v));
    w.emission = reinterpret_cast<float4*>(take(v));
    w.shadow_o = reinterpret_cast<float4*>(take(v));
    w.shadow_d = reinterpret_cast<float4*>(take(v));
    w.shadow_add = reinterpret_cast<float4*>(take(v));
    w.adjoint = reinterpret_cast<float4*>(take(v));
    w.shadow_grad = reinterpret_cast<float4*>(take(v));
    w.queue[0] = reinterpret_cast<uint32_t*>(take(uint64_t(paths) * 4));
    w.queue[1] = reinterpret_cast<uint32_t*>(take(uint64_t(paths) * 4));
    w.nee_rng = reinterpret_cast<uint32_t*>(take(uint64_t(paths) * 4));
    w.hint_word = reinterpret_cast<uint32_t*>(take(uint64_t(paths) * 4));
    w.hint_counts = reinterpret_cast<uint2*>(take(uint64_t(paths) * 8));
    w.sky_sh = reinterpret_cast<float4*>(take(256));
    w.ctl = reinterpret_cast<uint32_t*>(take(256));
    w.partials = reinterpret_cast<float*>(take(uint64_t(GradRows(paths)) * kGradEntries * 4));
    return w;
}

// The scenes of a launch. OneScene: the plain frame. SceneBatch: chunk scenes of one image size whose
// paths are numbered scene-major (paths_per_scene each); their light trees and cameras are arrays
// in device memory and their outputs are images of pixels_per_scene pixels, scene-major.
struct OneScene
{
    LightRef light;
    CameraView camera;
    CX_DEV uint32_t SceneOf(uint32_t) const { return 0; }
    CX_DEV uint32_t Local(uint32_t path) const { return path; }
    CX_DEV LightView Light(uint32_t) const { return OpenLights(light); }
    CX_DEV const CameraView& Camera(uint32_t) const { return camera; }
    CX_DEV uint64_t PixelBase(uint32_t) const { return 0; }
};
struct SceneBatch
{
    const LightRef* lights;
    const CameraView* cameras;
    uint32_t paths_per_scene;
    uint32_t pixels_per_scene;
    CX_DEV uint32_t SceneOf(uint32_t path) const { return path / paths_per_scene; }
    CX_DEV uint32_t Local(uint32_t path) const { return path % paths_per_scene; }
    CX_DEV LightView Light(uint32_t scene) const { return OpenLights(lights[scene]); }
    CX_DEV const CameraView& Camera(uint32_t scene) const { return cameras[scene]; }
    CX_DEV uint64_t PixelBase(uint32_t scene) const { return uint64_t(scene) * pixels_per_scene; }
};

CX_DEV void PathPixel(const CameraView& cam, uint32_t path, uint32_t& px, uint32_t& py)
{
    const uint32_t tiles_x = (cam.width + 7) / 8;
    const uint32_t tile = path >> 6;
    const uint32_t within = path & 63u;
    px = (tile % tiles_x) * 8 + (within & 7u);
    py = (tile / tiles_x) * 8 + (within >> 3);
}

// Appends `value` to queue[*count ...] for every lane with `active`; one atomic per wave.
CX_DEV uint32_t WaveAppend(uint32_t* count, bool active)
{
    const uint64_t mask = Ballot(active);
    if (mask == 0)
    {
        return 0;
    }
    uint32_t base = 0;
    const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    if (Lane() == uint32_t(leader))
    {
        base = AtomicAddRelaxed(count, uint32_t(PopC(mask)));
    }
    base = __shfl(base, leader);
    return base + uint32_t(PopC(mask & LanesBelow()));
}

CX_DEV void EnqueueShadow(const Workspace& w, uint32_t path, bool active, V3 origin, V3 direction, float t_max, V3 add,
                          uint32_t light, V3 unscaled, float roulette, bool adjoint)
{
    const uint32_t slot = WaveAppend(&w.ctl[2], active);
    if (active)
    {
        w.shadow_o[slot] = make_float4(origin.x, origin.y, origin.z, t_max);
        w.shadow_d[slot] = make_float4(direction.x, direction.y, direction.z, __uint_as_float(path));
        w.shadow_add[slot] = make_float4(add.x, add.y, add.z, __uint_as_float(light));
        if (adjoint)
        {
            w.shadow_grad[slot] = make_float4(unscaled.x, unscaled.y, unscaled.z, roulette);
        }
    }
}

// state.z: segment light | bounce << 8 | segment face << 16 | medium << 19 | segment sees the sun << 28 |
// the transmissive NEE's claim << 29 (the last three only in the full material model).
CX_DEV uint32_t PackState(uint32_t segment_light, uint32_t bounce, uint32_t face, uint32_t medium = 0u,
                          bool sees_sun = false, bool claim = false)
{
    return segment_light | (bounce << 8) | (face << 16) | (medium << 19) | (sees_sun ? (1u << 28) : 0u) |
           (claim ? (1u << 29) : 0u);
}

// ---- adjoint helpers ---------------------------------------------------------------------------
// The bounce vertices of a path whose albedos enter its later terms: material ids (block & 0xFF) of
// vertices 1..3 in bytes 0..2, their count in byte 3.
CX_DEV uint32_t VertexCount(uint32_t verts)
{
    return verts >> 24;
}
CX_DEV uint32_t VertexId(uint32_t verts, uint32_t j)
{
    return (verts >> (8 * j)) & 0xFFu;
}
CX_DEV uint32_t AppendVertex(uint32_t verts, uint32_t id)
{
    const uint32_t n = VertexCount(verts);
    return n >= 3 ? verts : ((verts & 0x00FFFFFFu) | (id << (8 * n)) | ((n + 1) << 24));
}
CX_DEV V3 MaterialAlbedo(const SceneView& s, uint32_t id)
{
    return BlockAlbedo(s, kBlockIdOpaqueMask | id);
}

CX_DEV void LdsGradAdd(float* lds, uint32_t id, V3 v)
{
    atomicAdd(&lds[id * 3 + 0], v.x);
    atomicAdd(&lds[id * 3 + 1], v.y);
    atomicAdd(&lds[id * 3 + 2], v.z);
}

// A term T = base * prod_{j < n} albedo(vertex j) (per channel): adds adj * dT/d(albedo of vertex j)
// = adj * base * prod_{i != j} albedo(vertex i) for every j, and returns T's albedo product.
CX_DEV V3 AddAlbedoTerms(const SceneView& s, float* lds, V3 adj, V3 base, uint32_t verts, uint32_t n)
{
    V3 a[3];
    V3 product = Make(1.0f, 1.0f, 1.0f);
#pragma unroll
    for (uint32_t j = 0; j < 3; ++j)
    {
        a[j] = j < n ? MaterialAlbedo(s, VertexId(verts, j)) : Make(1.0f, 1.0f, 1.0f);
        product = MulV(product, a[j]);
    }
#pragma unroll
    for (uint32_t j = 0; j < 3; ++j)
    {
        if (j < n)
        {
            V3 others = base;
#pragma unroll
            for (uint32_t i = 0; i < 3; ++i)
            {
                if (i != j)
                {
                    others = MulV(others, a[i]);
                }
            }
            LdsGradAdd(lds, VertexId(verts, j), MulV(adj, others));
        }
    }
    return product;
}

CX_DEV void ScaleGradAdd(const GradView& g, uint32_t record, V3 adj, V3 d_term)
{
    const float v = adj.x * d_term.x + adj.y * d_term.y + adj.z * d_term.z;
    if (v != 0.0f)
    {
        atomicAdd(&g.grad_scale[record], v);
    }
}

CX_DEV void GradRowsBegin(float* lds)
{
    for (uint32_t i = threadIdx.x; i < kGradEntries; i += blockDim.x)
    {
        lds[i] = 0.0f;
    }
    __syncthreads();
}
CX_DEV void GradRowsEnd(const Workspace& w, float* lds)
{
    __syncthreads();
    float* row = w.partials + uint64_t(blockIdx.x) * kGradEntries;
    for (uint32_t i = threadIdx.x; i < kGradEntries; i += blockDim.x)
    {
        row[i] += lds[i];
    }
}

// ---- persistent traversal ----------------------------------------------------------------------
// Control words: [0], [1] path queue lengths, [2] shadow queue length, [8] shadow cursor, [9] bounce
// cursor, [10] camera-ray cursor. A wave takes a chunk of consecutive items of its queue per atomic
// and hands them to its idle lanes; the lanes advance their rays together (WorldRayAdvance) and an
// idle lane takes the next item, so a wave's lanes stay busy until the queue is drained.
constexpr uint32_t kCtlShadowCursor = 8;
constexpr uint32_t kCtlBounceCursor = 9;
constexpr uint32_t kCtlPrimaryCursor = 10;

__global__ void __launch_bounds__(64) ResetKernel(Workspace w)
{
    w.ctl[threadIdx.x] = 0u;
}

// Camera rays for the persistent traversal: item = path (its pixel by PathPixel); every path gets its
// hit, a miss for a path outside its image (no ray: t_max < 0 ends it before the grid). A wave claims
// kPrimaryChunk paths at a time (one 8x8 pixel tile): the rays a wave holds stay coherent.
constexpr uint32_t kPrimaryChunk = 64;
template <class Views>
struct PrimarySource
{
    Workspace w;
    Views views;
    uint32_t paths;
    CX_DEV uint32_t Count() const { return paths; }
    CX_DEV uint32_t* Cursor() const { return &w.ctl[kCtlPrimaryCursor]; }
    CX_DEV void Prepare() const {}
    CX_DEV bool Load(uint32_t item, V3& o, V3& d, float& t_max, uint32_t& scene, PathCounters& c) const
    {
        scene = views.SceneOf(item);
        const CameraView& cam = views.Camera(scene);
        uint32_t px, py;
        PathPixel(cam, views.Local(item), px, py);
        o = Make(cam.position[0], cam.position[1], cam.position[2]);
        if (px < cam.width && py < cam.height)
        {
            d = PrimaryDirection(cam, px, py);
            t_max = kRayTMax;
            c.rays += 1;
        }
        else
        {
            d = Make(0.0f, 0.0f, 1.0f);
            t_max = -1.0f;
        }
        return true;
    }
    CX_DEV void Finish(uint32_t item, const WorldHit& h, float*) const
    {
        w.hit[item] = make_uint4(__float_as_uint(h.t), h.attrib, h.record, 0u);
    }
};

// Bounce rays: item -> path of queue[in].
template <class Views>
struct BounceSource
{
    Workspace w;
    uint32_t in;
    Views views;
    CX_DEV uint32_t Count() const { return w.ctl[in]; }
    CX_DEV uint32_t* Cursor() const { return &w.ctl[kCtlBounceCursor]; }
    // The queues the following shade launch appends to, and the shadow cursor, start empty.
    CX_DEV void Prepare() const
    {
        w.ctl[1 - in] = 0u;
        w.ctl[2] = 0u;
        w.ctl[kCtlShadowCursor] = 0u;
    }
    CX_DEV bool Load(uint32_t item, V3& o, V3& d, float& t_max, uint32_t& scene, PathCounters& c) const
    {
        const uint32_t path = w.queue[in][item];
        scene = views.SceneOf(path);
        const float4 o4 = w.ray_o[path];
        const float4 d4 = w.ray_d[path];
        o = Make(o4.x, o4.y, o4.z);
        d = Make(d4.x, d4.y, d4.z);
        t_max = kRayTMax;
        c.rays += 1;
        return true;
    }
    CX_DEV void Finish(uint32_t item, const WorldHit& h, float*) const
    {
        w.hit[w.queue[in][item]] = make_uint4(__float_as_uint(h.t), h.attrib, h.record, 0u);
    }
};

// Shadow rays: an unoccluded one adds its radiance to its path (one shadow ray per path per launch).
template <class Views>
struct ShadowSource
{
    Workspace w;
    Views views;
    SceneView sv;
    GradView g;
    uint32_t hint;  // the primary vertex's shadow rays with the visibility hint: count the visible ones
    CX_DEV uint32_t Count() const { return w.ctl[2]; }
    CX_DEV uint32_t* Cursor() const { return &w.ctl[kCtlShadowCursor]; }
    CX_DEV void Prepare() const {}
    CX_DEV bool Load(uint32_t item, V3& o, V3& d, float& t_max, uint32_t& scene, PathCounters&) const
    {
        const float4 o4 = w.shadow_o[item];
        const float4 d4 = w.shadow_d[item];
        scene = views.SceneOf(__float_as_uint(d4.w));
        o = Make(o4.x, o4.y, o4.z);
        d = Make(d4.x, d4.y, d4.z);
        t_max = o4.w;
        return true;
    }
    CX_DEV void Finish(uint32_t item, const WorldHit& h, float* lds) const
    {
        if (h.attrib != 0u)
        {
            return;
        }
        const uint32_t path = __float_as_uint(w.shadow_d[item].w);
        if (hint != 0u)
        {
            // one shadow ray per path in this launch: no other lane touches the path's counts
            const uint32_t traced = (w.hint_word[path] >> 16) & 3u;
            uint2 counts = w.hint_counts[path];
            counts.x += traced == kNeeTracedSun ? (1u << 16) : 0u;
            counts.y += traced == kNeeTracedLamp ? (1u << 16) : 0u;
            w.hint_counts[path] = counts;
        }
        const float4 a = w.shadow_add[item];
        float4 r = w.radiance[path];
        r.x = r.x + a.x;
        r.y = r.y + a.y;
        r.z = r.z + a.z;
        w.radiance[path] = r;
        if (g.d_color != nullptr)
        {
            // T = unscaled * scale(light) * roulette * prod_{j <= k} albedo(vertex j)
            const float4 sg = w.shadow_grad[item];
            const uint32_t light = __float_as_uint(a.w);
            const float4 adj4 = w.adjoint[path];
            const V3 adj = Make(adj4.x, adj4.y, adj4.z);
            const uint32_t verts = w.state[path].w;
            const V3 unscaled = Mul(Make(sg.x, sg.y, sg.z), sg.w);
            const V3 base = Mul(unscaled, BrickScale(sv, light));
            const V3 product = AddAlbedoTerms(sv, lds, adj, base, verts, VertexCount(verts));
            ScaleGradAdd(g, light, adj, MulV(unscaled, product));
        }
    }
};

template <class Source>
__global__ void __launch_bounds__(kWaveThreads) PersistentTraceKernel(SceneView s, Source src,
                                                                      unsigned long long* counters, Workspace w,
                                                                      bool adjoint, uint32_t chunk)
{
    __shared__ float lds_grad[kGradEntries];
    if (adjoint)
    {
        GradRowsBegin(lds_grad);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        src.Prepare();
    }
    PathCounters c{0, 0, 0, 0, 0};
    const uint32_t count = src.Count();
    uint32_t* cursor = src.Cursor();
    WorldRay r;
    r.phase = kRayDone;
    uint32_t item = 0;
    bool busy = false;
    uint32_t chunk_next = 0;  // wave-uniform: next item of the wave's chunk
    uint32_t chunk_end = 0;   // wave-uniform: end of the wave's chunk
    bool drained = false;     // wave-uniform: the queue has no more chunks
    while (true)
    {
        const uint64_t idle = Ballot(!busy);
        if (idle != 0 && !drained)
        {
            if (chunk_next == chunk_end)
            {
                uint32_t base = 0;
                if (Lane() == 0)
                {
                    base = AtomicAddRelaxed(cursor, chunk);
                }
                base = __shfl(base, 0);
                chunk_next = min(base, count);
                chunk_end = min(base + chunk, count);
                drained = chunk_next == chunk_end;
            }
            const uint32_t take = min(uint32_t(PopC(idle)), chunk_end - chunk_next);
            const uint32_t rank = uint32_t(PopC(idle & LanesBelow()));
            if (!busy && rank < take)
            {
                item = chunk_next + rank;
                V3 o, d;
                float t_max;
                uint32_t scene = 0;
                if (src.Load(item, o, d, t_max, scene, c))
                {
                    WorldRayBegin(s, o, d, t_max, scene, r);
                    busy = r.phase != kRayDone;
                    if (!busy)
                    {
                        src.Finish(item, WorldHit{0.0f, 0u, kHashNone, {0, 0, 0}, 0u, 0u, 0u}, lds_grad);
                    }
                }
            }
            chunk_next += take;
        }
        if (Ballot(busy) == 0)
        {
            if (drained)
            {
                break;
            }
            continue;
        }
        WorldHit h;
        if (WorldRayAdvance(s, r, busy, h, c))
        {
            src.Finish(item, h, lds_grad);
            busy = false;
        }
    }
    if (adjoint)
    {
        GradRowsEnd(w, lds_grad);
    }
    FlushCounters(counters, c);
}

// The primary vertex in the full material model: the delta chain of the camera ray (RunDeltaChain,
// caelix_path_trace.hlsl:613-678; traced here, it is short and rare), then ShadeVoxelHit at the
// surface that ended it, the NEE unless that shade took the specular lobe, and the first bounce ray.
// Outputs: albedo = chain transmission x albedo, emission = chain transmission x emission + the
// fog along the chain (caelix_path_trace.hlsl:1270-1284); a chain that leaves the scene shows the
// sky in its last direction times its transmission; depth = the chain's length, attrib = the last
// surface or interface it met.
CX_DEV void PrimaryGlass(const SceneView& s, const TraceParams& prm, const Workspace& w, uint32_t path, uint32_t scene,
                         const LightView& L, V3 origin, V3 dir, uint4 h, bool even_pixel, uint32_t& rng, uint32_t& light_rng,
                         uint32_t& nee_rng, uint32_t mis_bit, PathCounters& c, bool& shadow, bool& go_on,
                         V3& shadow_origin, V3& shadow_dir, float& shadow_t, V3& shadow_add, uint32_t& shadow_light,
                         V3& shadow_unscaled, const CameraView& cam, uint32_t px, uint32_t py, bool hint,
                         uint32_t& hint_word, uint32_t& traced, uint32_t& seen)
{
    WorldHit surface{__uint_as_float(h.x), h.y, h.z, {0, 0, 0}, 0u, 0u, 0u};
    uint32_t geometry = h.y;
    V3 ray_o = origin;
    V3 ray_d = dir;
    V3 chain = Make(1.0f, 1.0f, 1.0f);
    V3 inscatter = Make(0.0f, 0.0f, 0.0f);
    float t_virtual = 0.0f;
    uint32_t medium = 0u;
    uint32_t interfaces = 0u;
    bool surface_hit = true;
    for (;;)
    {
        MediumSegment(s, prm, medium, surface.t, chain, inscatter);
        t_virtual += surface.t;
        geometry = surface.attrib;
        if (interfaces >= prm.delta_budget)
        {
            break;
        }
        const uint32_t block = (surface.attrib >> 16) & 0xFFFFu;
        if (BlockOpaque(block, s.encoding))
        {
            break;
        }
        const uint32_t code = BlockCode(block, s.encoding);
        const Material m = GetMaterial(s, code);
        if (m.smoothness < 0.99f)
        {
            break;
        }
        const V3 hp = Add(ray_o, Mul(ray_d, surface.t));
        const int split = prm.checkerboard != 0u && interfaces == 0u ? (even_pixel ? 1 : 0) : -1;
        const Interface it = CrossInterface(s, medium, code, m.ior, ray_d, FaceNormal(surface.attrib), rng, split);
        if (it.split_applied)
        {
            chain = Mul(chain, it.weight);
        }
        interfaces += 1u;
        medium = it.medium;
        ray_o = Add(hp, Mul(it.normal, it.refracted ? -kPushOff : kPushOff));
        ray_d = it.direction;
        c.rays += 1;
        surface = Trace(s, ray_o, ray_d, kRayTMax, c, scene);
        if (surface.attrib == 0u)
        {
            surface_hit = false;
            break;
        }
    }
    if (hint)
    {
        hint_word = ReadHint(prm.hint, cam.width, cam.height, px, py, surface_hit, Add(ray_o, Mul(ray_d, surface.t)));
    }
    if (!surface_hit)
    {
        const V3 sky = prm.sky.on != 0u ? SkyRadiance(prm.sky, ray_d) : Make(0.0f, 0.0f, 0.0f);
        const V3 deterministic = Add(MulV(chain, sky), inscatter);
        w.albedo[path] = make_float4(0.0f, 0.0f, 0.0f, t_virtual);
        w.emission[path] = make_float4(deterministic.x, deterministic.y, deterministic.z, __uint_as_float(geometry));
        return;
    }
    const uint32_t face = FaceIndex(surface.attrib);
    V3 normal = FaceDirection(face);
    const V3 hit_point = Add(ray_o, Mul(ray_d, surface.t));
    ShadeFull sh = ShadeVoxel(s, (surface.attrib >> 16) & 0xFFFFu, hit_point, ray_d, normal, rng, medium);
    // Subsurface (caelix_path_trace.hlsl:757-781): the walk may replace the vertex; the primary albedo
    // output is then the guide albedo and the branch weight over it the throughput.
    V3 throughput = Make(1.0f, 1.0f, 1.0f);
    uint32_t walk = kSubsurfaceNone;
    if (prm.subsurface != 0u)
    {
        V3 guide;
        walk = SubsurfaceVertex(s, scene, surface.record, surface.attrib, hit_point, ray_d, rng, sh, normal, guide, false);
        if (walk != kSubsurfaceNone)
        {
            throughput = Make(Div(sh.albedo.x, fmaxf(guide.x, 0.001f)), Div(sh.albedo.y, fmaxf(guide.y, 0.001f)),
                              Div(sh.albedo.z, fmaxf(guide.z, 0.001f)));
            sh.albedo = guide;
            if (walk == kSubsurfaceExit)
            {
                medium = 0u;  // the exit vertex lies in the air
            }
        }
    }
    const uint32_t vertex_face = walk == kSubsurfaceExit ? FaceOfNormal(normal) : face;
    const V3 emission = Mul(sh.emission, BrickScale(s, surface.record));
    const V3 albedo = MulV(chain, sh.albedo);
    const V3 deterministic = Add(MulV(chain, emission), inscatter);
    w.albedo[path] = make_float4(albedo.x, albedo.y, albedo.z, t_virtual);
    w.emission[path] = make_float4(deterministic.x, deterministic.y, deterministic.z, __uint_as_float(geometry));
    if (walk == kSubsurfaceDead)
    {
        return;  // no NEE, no segment: the path ends with its deterministic emission
    }
    uint32_t segment_light = 0u;
    bool sees_sun = true;
    bool claim = false;
    if (prm.bounces >= 1u)
    {
        if (!sh.specular)
        {
            V3 visible;
            if (NeeCandidate(s, L, prm.sky, sh.bounce_origin, normal, light_rng, nee_rng, shadow_dir, shadow_t, visible,
                             shadow_unscaled, shadow_light, hint_word, &traced))
            {
                c.shadow_rays += 1;
                if (prm.transmissive != 0u)
                {
                    // The shadow ray through the media, traced here (at most 8 hops): its radiance is
                    // the path's first term.
                    if (shadow_t > 0.0f)
                    {
                        const V3 tr = ShadowTransmissive(s, sh.bounce_origin, shadow_dir, shadow_t, medium, c, scene);
                        const V3 add = MulV(MulV(visible, tr), throughput);
                        w.radiance[path] = make_float4(add.x, add.y, add.z, 0.0f);
                        seen = tr.x > 0.0f || tr.y > 0.0f || tr.z > 0.0f ? 1u : 0u;
                    }
                }
                else
                {
                    shadow = shadow_t > 0.0f;
                    shadow_origin = sh.bounce_origin;
                    shadow_add = MulV(visible, throughput);
                }
            }
            segment_light = kSegmentLampNee | mis_bit;
            sees_sun = prm.sky.sun == 0u;
            claim = prm.transmissive != 0u;
        }
        go_on = true;
    }
    w.throughput[path] = make_float4(throughput.x, throughput.y, throughput.z, 1.0f);
    w.state[path] = make_uint4(rng, light_rng, PackState(segment_light, 1u, vertex_face, medium, sees_sun, claim), 0u);
    if (KeepsNeeStream(prm))
    {
        w.nee_rng[path] = nee_rng;
    }
    w.ray_o[path] = make_float4(sh.bounce_origin.x, sh.bounce_origin.y, sh.bounce_origin.z,
                                sh.specular ? sh.roughness : -1.0f);
    w.ray_d[path] = make_float4(sh.bounce_direction.x, sh.bounce_direction.y, sh.bounce_direction.z, 0.0f);
}

// The primary vertex: outputs of the hit, the first shade, the NEE at vertex 0 and the first
// bounce ray (queue 0). Adjoint: the path's adjoint dL/dI / spp * albedo, and the primary emission's
// emission-scale gradient.
template <class Views, bool Glass>
__global__ void __launch_bounds__(kWaveThreads) PrimaryShadeKernel(SceneView s, Views views, TraceParams prm,
                                                                   Workspace w, uint32_t paths, GradView g,
                                                                   unsigned long long* counters)
{
    const uint32_t path = blockIdx.x * kWaveThreads + threadIdx.x;
    PathCounters c{0, 0, 0, 0, 0};
    bool shadow = false;
    bool go_on = false;
    V3 shadow_origin = Make(0.0f, 0.0f, 0.0f), shadow_dir = Make(0.0f, 0.0f, 0.0f), shadow_add = Make(0.0f, 0.0f, 0.0f);
    V3 shadow_unscaled = Make(0.0f, 0.0f, 0.0f);
    uint32_t shadow_light = 0;
    float shadow_t = 0.0f;
    if (path < paths)
    {
        const uint4 h = w.hit[path];
        w.radiance[path] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const CameraView& pixel_cam = views.Camera(views.SceneOf(path));
        uint32_t hx, hy;
        PathPixel(pixel_cam, views.Local(path), hx, hy);
        // the visibility hint of the primary vertex (every pixel of the image reads one)
        const bool hint = prm.hint.previous != nullptr && hx < pixel_cam.width && hy < pixel_cam.height;
        uint32_t hint_word = kHintWordOff;
        uint32_t traced = kNeeTracedNone;
        uint32_t seen = 0u;  // a transmissive shadow ray that reached its light
        if (h.y == 0u)
        {
            // A primary miss: the sky with the sun disk in the sky's frame (the composite's
            // kRenderFlagSkyOnMiss, caelix_composite.hlsl:130-133), else black.
            V3 sky = Make(0.0f, 0.0f, 0.0f);
            if (prm.sky.on != 0u)
            {
                sky = SkyRadiance(prm.sky, PrimaryDirection(pixel_cam, hx, hy));
            }
            if (hint)
            {
                hint_word = ReadHint(prm.hint, pixel_cam.width, pixel_cam.height, hx, hy, false, Make(0.0f, 0.0f, 0.0f));
            }
            w.albedo[path] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            w.emission[path] = make_float4(sky.x, sky.y, sky.z, 0.0f);
        }
        else
        {
            const uint32_t scene = views.SceneOf(path);
            const LightView L = views.Light(scene);
            const CameraView& cam = views.Camera(scene);
            uint32_t px, py;
            PathPixel(cam, views.Local(path), px, py);
            const float t = __uint_as_float(h.x);
            const V3 dir = PrimaryDirection(cam, px, py);
            const V3 origin = Make(cam.position[0], cam.position[1], cam.position[2]);
            uint32_t rng = InitRandom(px, py, prm.frame_seed);
            uint32_t light_rng = InitRandom(px, py, prm.frame_seed ^ 0x7F4A7C15u);
            uint32_t nee_rng = InitRandom(px, py, prm.frame_seed ^ 0x68E31DA4u);
            const uint32_t mis_bit = L.mis != 0u ? kSegmentLampMis : 0u;
            const uint32_t face = FaceIndex(h.y);
            const V3 normal = FaceDirection(face);
            if constexpr (Glass)
            {
                PrimaryGlass(s, prm, w, path, views.SceneOf(path), L, origin, dir, h, ((px + py) & 1u) == 0u, rng,
                             light_rng, nee_rng, mis_bit, c, shadow, go_on, shadow_origin, shadow_dir, shadow_t,
                             shadow_add, shadow_light, shadow_unscaled, cam, px, py, hint, hint_word, traced, seen);
            }
            else
            {
                const Shade sh = ShadeHit(s, (h.y >> 16) & 0xFFFFu, Add(origin, Mul(dir, t)), normal, rng);
                const V3 emission = Mul(sh.emission, BrickScale(s, h.z));
                if (prm.fog_density > 0.0f)
                {
                    // The camera segment in the air (the delta chain's TraverseSegment): its transmission
                    // multiplies what the surface sends, its fog emission is deterministic
                    // (caelix_path_trace.hlsl:1273-1283).
                    V3 chain = Make(1.0f, 1.0f, 1.0f);
                    V3 inscatter = Make(0.0f, 0.0f, 0.0f);
                    FogSegment(prm, t, chain, inscatter);
                    const V3 albedo = MulV(chain, sh.albedo);
                    const V3 deterministic = Add(MulV(chain, emission), inscatter);
                    w.albedo[path] = make_float4(albedo.x, albedo.y, albedo.z, t);
                    w.emission[path] = make_float4(deterministic.x, deterministic.y, deterministic.z, __uint_as_float(h.y));
                }
                else
                {
                    w.albedo[path] = make_float4(sh.albedo.x, sh.albedo.y, sh.albedo.z, t);
                    w.emission[path] = make_float4(emission.x, emission.y, emission.z, __uint_as_float(h.y));
                }
                if (g.d_color != nullptr)
                {
                    const float4 d = g.d_color[views.PixelBase(scene) + uint64_t(py) * cam.width + px];
                    const V3 dc = Mul(Make(d.x, d.y, d.z), g.inv_spp);
                    const V3 adj = MulV(dc, sh.albedo);
                    w.adjoint[path] = make_float4(adj.x, adj.y, adj.z, 0.0f);
                    ScaleGradAdd(g, h.z, dc, sh.emission);  // color = emission * scale + albedo * radiance
                }
                const V3 throughput = Make(1.0f, 1.0f, 1.0f);
                uint32_t segment_light = 0u;
                if (hint)
                {
                    hint_word = ReadHint(prm.hint, cam.width, cam.height, px, py, true, Add(origin, Mul(dir, t)));
                }
                if (prm.bounces >= 1u)
                {
                    V3 visible;
                    if (NeeCandidate(s, L, prm.sky, sh.bounce_origin, normal, light_rng, nee_rng, shadow_dir, shadow_t, visible,
                                     shadow_unscaled, shadow_light, hint_word, &traced))
                    {
                        c.shadow_rays += 1;
                        shadow = shadow_t > 0.0f;
                        shadow_origin = sh.bounce_origin;
                        shadow_add = MulV(visible, throughput);
                    }
                    segment_light = kSegmentLampNee | mis_bit;
                    go_on = true;
                }
                w.throughput[path] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);  // w: the roulette's normalisation
                w.state[path] = make_uint4(rng, light_rng, PackState(segment_light, 1u, face), 0u);
                if (KeepsNeeStream(prm))
                {
                    w.nee_rng[path] = nee_rng;
                }
                w.ray_o[path] = make_float4(sh.bounce_origin.x, sh.bounce_origin.y, sh.bounce_origin.z, -1.0f);
                w.ray_d[path] = make_float4(sh.bounce_direction.x, sh.bounce_direction.y, sh.bounce_direction.z, 0.0f);
            }
        }
        if (hint)
        {
            uint2 counts = prm.hint.sample == 0u ? make_uint2(0u, 0u) : w.hint_counts[path];
            const uint32_t add = 1u + (seen << 16);
            counts.x += traced == kNeeTracedSun ? add : 0u;
            counts.y += traced == kNeeTracedLamp ? add : 0u;
            w.hint_counts[path] = counts;
            const uint32_t read = prm.hint.sample == 0u ? hint_word : (w.hint_word[path] & 0xFFFFu);
            w.hint_word[path] = read | (traced << 16);
        }
    }
    EnqueueShadow(w, path, shadow, shadow_origin, shadow_dir, shadow_t, shadow_add, shadow_light, shadow_unscaled, 1.0f,
                  g.d_color != nullptr);
    const uint32_t slot = WaveAppend(&w.ctl[0], go_on);
    if (go_on)
    {
        w.queue[0][slot] = path;
    }
    FlushCounters(counters, c);
}

// One iteration of TracePath's bounce loop for the paths in queue[(bounce - 1) & 1]. Adjoint: the
// gradients of the emission term of this vertex.
template <class Views, bool Glass>
__global__ void __launch_bounds__(kWaveThreads) BounceShadeKernel(SceneView s, Views views, TraceParams prm, Workspace w,
                                                                  uint32_t bounce, uint32_t paths, GradView g,
                                                                  unsigned long long* counters)
{
    __shared__ float lds_grad[kGradEntries];
    const bool adjoint = g.d_color != nullptr;
    if (adjoint)
    {
        GradRowsBegin(lds_grad);
    }
    const uint32_t i = blockIdx.x * kWaveThreads + threadIdx.x;
    const uint32_t in = (bounce - 1) & 1u;
    if (i == 0)
    {
        w.ctl[kCtlBounceCursor] = 0u;  // for the next bounce's rays
    }
    PathCounters c{0, 0, 0, 0, 0};
    const uint32_t count = w.ctl[in];
    bool shadow = false;
    bool go_on = false;
    uint32_t path = 0;
    V3 shadow_origin = Make(0.0f, 0.0f, 0.0f), shadow_dir = Make(0.0f, 0.0f, 0.0f), shadow_add = Make(0.0f, 0.0f, 0.0f);
    V3 shadow_unscaled = Make(0.0f, 0.0f, 0.0f);
    uint32_t shadow_light = 0;
    float shadow_t = 0.0f;
    float shadow_roulette = 1.0f;
    if (i < count && i < paths)
    {
        path = w.queue[in][i];
        const uint4 h = w.hit[path];
        if (h.y == 0u && prm.sky.on != 0u)
        {
            // A bounce miss: the sky the diffuse segment receives (SkyDiffuseMiss, caelix_path_trace.hlsl:840-846).
            const float4 d4 = w.ray_d[path];
            const float4 t4 = w.throughput[path];
            const float4 r4 = w.radiance[path];
            // The full material model keeps whether the segment sees the sun (it does after a specular
            // vertex, which ran no NEE); else every segment leaves an NEE vertex.
            const uint32_t z = w.state[path].z;
            const V3 n = FaceDirection((z >> 16) & 7u);
            const V3 sky = Glass ? SkySegmentMiss(prm.sky, n, Make(d4.x, d4.y, d4.z), w.ray_o[path].w, ((z >> 28) & 1u) != 0u)
                                 : SkyMiss(prm.sky, n, Make(d4.x, d4.y, d4.z));
            const V3 add = MulV(sky, Make(t4.x, t4.y, t4.z));
            w.radiance[path] = make_float4(r4.x + add.x, r4.y + add.y, r4.z + add.z, 0.0f);
        }
        if (h.y != 0u)
        {
            const LightView L = views.Light(views.SceneOf(path));
            uint4 st = w.state[path];
            uint32_t rng = st.x;
            uint32_t light_rng = st.y;
            uint32_t nee_rng = KeepsNeeStream(prm) ? w.nee_rng[path] : 0u;
            uint32_t segment_light = st.z & 0xFFu;
            uint32_t bounce_index = (st.z >> 8) & 0xFFu;
            const uint32_t segment_face = (st.z >> 16) & 7u;
            uint32_t medium = (st.z >> 19) & 0x1FFu;
            bool claim = ((st.z >> 29) & 1u) != 0u;
            uint32_t verts = st.w;
            const float4 o4 = w.ray_o[path];
            const float4 d4 = w.ray_d[path];
            const V3 origin = Make(o4.x, o4.y, o4.z);
            const V3 dir = Make(d4.x, d4.y, d4.z);
            const float4 t4 = w.throughput[path];
            V3 throughput = Make(t4.x, t4.y, t4.z);
            float roulette = t4.w;
            const float4 r4 = w.radiance[path];
            V3 radiance = Make(r4.x, r4.y, r4.z);
            WorldHit hit;
            hit.t = __uint_as_float(h.x);
            hit.attrib = h.y;
            hit.record = h.z;
            if constexpr (Glass)
            {
                MediumSegment(s, prm, medium, hit.t, throughput, radiance);
            }
            else if (prm.fog_density > 0.0f)
            {
                FogSegment(prm, hit.t, throughput, radiance);
            }
            const uint32_t face = FaceIndex(h.y);
            const uint32_t block = (h.y >> 16) & 0xFFFFu;
            V3 vertex_normal = FaceDirection(face);
            ShadeFull sh;
            uint32_t walk = kSubsurfaceNone;
            if constexpr (Glass)
            {
                const V3 hit_point = Add(origin, Mul(dir, hit.t));
                sh = ShadeVoxel(s, block, hit_point, dir, vertex_normal, rng, medium);
                if (prm.subsurface != 0u)
                {
                    // caelix_path_trace.hlsl:861-867: a vertex without a next segment skips the walk
                    V3 guide;
                    walk = SubsurfaceVertex(s, views.SceneOf(path), h.z, h.y, hit_point, dir, rng, sh, vertex_normal,
                                            guide, bounce_index + 1u > prm.bounces);
                }
            }
            else
            {
                const Shade lambert = ShadeHit(s, block, Add(origin, Mul(dir, hit.t)), vertex_normal, rng);
                sh = ShadeFull{lambert.albedo, lambert.emission, lambert.bounce_origin, lambert.bounce_direction, false,
                               1.0f};
            }
            const V3 emission = Mul(sh.emission, BrickScale(s, h.z));
            bounce_index += 1;
            float emission_weight = 1.0f;
            if (segment_light != 0u && (sh.emission.x > 0.0f || sh.emission.y > 0.0f || sh.emission.z > 0.0f))
            {
                emission_weight = LampHitWeight(s, L, segment_light, hit, origin, dir, FaceDirection(segment_face));
            }
            radiance = Add(radiance, MulV(emission, Mul(throughput, emission_weight)));
            if (adjoint && (sh.emission.x > 0.0f || sh.emission.y > 0.0f || sh.emission.z > 0.0f))
            {
                // T = emission * scale * weight * roulette * prod_{j < k} albedo(vertex j)
                const float4 adj4 = w.adjoint[path];
                const V3 adj = Make(adj4.x, adj4.y, adj4.z);
                const V3 base = Mul(emission, emission_weight * roulette);
                const V3 product = AddAlbedoTerms(s, lds_grad, adj, base, verts, VertexCount(verts));
                ScaleGradAdd(g, h.z, adj, MulV(Mul(sh.emission, emission_weight * roulette), product));
            }
            // A dead walk ends the path after the hit's emission; an exit vertex lies in the air.
            const bool dead = walk == kSubsurfaceDead;
            if (walk == kSubsurfaceExit)
            {
                medium = 0u;
            }
            throughput = MulV(throughput, Make(Div(sh.albedo.x, 1.0f), Div(sh.albedo.y, 1.0f), Div(sh.albedo.z, 1.0f)));
            verts = AppendVertex(verts, block & 0xFFu);
            segment_light = 0u;
            bool sees_sun = true;
            if (Glass && claim)
            {
                // The claim of a transmissive primary NEE (caelix_path_trace.hlsl:895-907): through a
                // transparent vertex the next segment sees no sun disk and counts a samplable lamp as 0.
                if (!BlockOpaque(block, s.encoding))
                {
                    sees_sun = prm.sky.sun == 0u;
                    segment_light = kSegmentLampNee;
                }
                else
                {
                    claim = false;
                }
            }
            if (!sh.specular && bounce_index <= prm.bounces && !dead)
            {
                segment_light = kSegmentLampNee | (L.mis != 0u ? kSegmentLampMis : 0u);
                sees_sun = prm.sky.sun == 0u;
                // The bounce NEE budget: with probability q, scaled by 1 / q (q = 1 draws no number).
                const bool run = prm.nee_q >= 1.0f || RandomFloat01(nee_rng) < prm.nee_q;
                V3 visible;
                if (run && NeeCandidate(s, L, prm.sky, sh.bounce_origin, vertex_normal, light_rng, nee_rng, shadow_dir,
                                        shadow_t, visible, shadow_unscaled, shadow_light))
                {
                    c.shadow_rays += 1;
                    shadow = shadow_t > 0.0f;
                    shadow_origin = sh.bounce_origin;
                    shadow_add = MulV(visible, Mul(throughput, Div(1.0f, prm.nee_q)));
                    shadow_roulette = roulette;
                }
            }
            const float cp = fmaxf(throughput.x, fmaxf(throughput.y, throughput.z));
            const float u = RandomFloat01(rng);
            const bool survive = !dead && (prm.roulette != 0u ? !(cp <= 0.0f || cp < u) : !(cp <= 0.0f));
            if (survive)
            {
                if (prm.roulette != 0u)
                {
                    const float inv_cp = Div(1.0f, cp);
                    throughput = Mul(throughput, inv_cp);
                    roulette = roulette * inv_cp;
                }
                go_on = bounce_index <= prm.bounces;
                // w: the segment's lobe for the sky it may miss into (roughness; -1: diffuse)
                w.ray_o[path] = make_float4(sh.bounce_origin.x, sh.bounce_origin.y, sh.bounce_origin.z,
                                            sh.specular ? sh.roughness : -1.0f);
                w.ray_d[path] = make_float4(sh.bounce_direction.x, sh.bounce_direction.y, sh.bounce_direction.z, 0.0f);
            }
            w.radiance[path] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
            w.throughput[path] = make_float4(throughput.x, throughput.y, throughput.z, roulette);
            w.state[path] = make_uint4(rng, light_rng,
                                       Glass ? PackState(segment_light, bounce_index,
                                                         walk == kSubsurfaceExit ? FaceOfNormal(vertex_normal) : face,
                                                         medium, sees_sun, claim)
                                             : PackState(segment_light, bounce_index, face),
                                       verts);
            if (KeepsNeeStream(prm))
            {
                w.nee_rng[path] = nee_rng;
            }
        }
    }
    EnqueueShadow(w, path, shadow, shadow_origin, shadow_dir, shadow_t, shadow_add, shadow_light, shadow_unscaled,
                  shadow_roulette, adjoint);
    const uint32_t slot = WaveAppend(&w.ctl[1 - in], go_on);
    if (go_on)
    {
        w.queue[1 - in][slot] = path;
    }
    FlushCounters(counters, c);
    if (adjoint)
    {
        GradRowsEnd(w, lds_grad);
    }
}

// color += emission + albedo * radiance, radiance += radiance, t and attrib of the last sample.
// Adjoint: the primary albedo's gradient dL/dI / spp * radiance.
template <class Views>
__global__ void __launch_bounds__(kWaveThreads) FinishKernel(Views views, Workspace w, uint32_t paths, uint32_t sample,
                                                             FrameOut out, GradView g)
{
    __shared__ float lds_grad[kGradEntries];
    const bool adjoint = g.d_color != nullptr;
    if (adjoint)
    {
        GradRowsBegin(lds_grad);
    }
    const uint32_t path = blockIdx.x * kWaveThreads + threadIdx.x;
    uint32_t px = 0, py = 0;
    const uint32_t scene = path < paths ? views.SceneOf(path) : 0u;
    const CameraView& cam = views.Camera(scene);
    if (path < paths)
    {
        PathPixel(cam, views.Local(path), px, py);
    }
    if (path < paths && px < cam.width && py < cam.height)
    {
        const float4 a = w.albedo[path];
        const float4 e = w.emission[path];
        const float4 r = w.radiance[path];
        const uint64_t pixel = views.PixelBase(scene) + uint64_t(py) * cam.width + px;
        if (adjoint)
        {
            const uint32_t attrib = __float_as_uint(e.w);
            if (attrib != 0u)
            {
                const float4 d = g.d_color[pixel];
                LdsGradAdd(lds_grad, (attrib >> 16) & 0xFFu,
                           MulV(Mul(Make(d.x, d.y, d.z), g.inv_spp), Make(r.x, r.y, r.z)));
            }
        }
        else
        {
            const V3 x = Add(Make(e.x, e.y, e.z), MulV(Make(a.x, a.y, a.z), Make(r.x, r.y, r.z)));
            if (out.color != nullptr)
            {
                float4 c = sample == 0 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : out.color[pixel];
                c = make_float4(c.x + x.x, c.y + x.y, c.z + x.z, a.w);
                out.color[pixel] = c;
            }
            if (out.radiance != nullptr)
            {
                float4 c = sample == 0 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : out.radiance[pixel];
                c = make_float4(c.x + r.x, c.y + r.y, c.z + r.z, 0.0f);
                out.radiance[pixel] = c;
            }
            if (out.attrib != nullptr)
            {
                out.attrib[pixel] = __float_as_uint(e.w);
            }
        }
    }
    if (adjoint)
    {
        GradRowsEnd(w, lds_grad);
    }
}

// WriteLightHint (caelix_path_trace.hlsl:548-580) after the frame's samples, one thread per patch: the
// first pixel of the patch in the image (its top-left one, the lowest lane of its 8x8 tile) gives the
// value read; each channel moves by kHintAlpha toward the fraction of visible primary shadow rays of
// the patch's pixels over the samples, and keeps its value without one. `read` (optional): the hint
// each pixel's primary vertex read.
__global__ void __launch_bounds__(256) HintUpdateKernel(Workspace w, CameraView cam, uint32_t shift, uint32_t columns,
                                                        uint32_t patch_rows, uint16_t* current, uint16_t* read)
{
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q >= columns * patch_rows)
    {
        return;
    }
    const uint32_t tiles_x = (cam.width + 7) / 8;
    const uint32_t x0 = (q % columns) << shift;
    const uint32_t y0 = (q / columns) << shift;
    const uint32_t x1 = min(x0 + (1u << shift), cam.width);
    const uint32_t y1 = min(y0 + (1u << shift), cam.height);
    auto path_of = [&](uint32_t x, uint32_t y) { return ((y >> 3) * tiles_x + (x >> 3)) * 64u + (y & 7u) * 8u + (x & 7u); };
    const uint32_t first = w.hint_word[path_of(x0, y0)] & 0xFFFFu;
    uint32_t sun = 0, sun_visible = 0, lamp = 0, lamp_visible = 0;
    for (uint32_t y = y0; y < y1; ++y)
    {
        for (uint32_t x = x0; x < x1; ++x)
        {
            const uint32_t path = path_of(x, y);
            const uint2 c = w.hint_counts[path];
            sun += c.x & 0xFFFFu;
            sun_visible += c.x >> 16;
            lamp += c.y & 0xFFFFu;
            lamp_visible += c.y >> 16;
            if (read != nullptr)
            {
                read[uint64_t(y) * cam.width + x] = uint16_t(w.hint_word[path] & 0xFFFFu);
            }
        }
    }
    current[q] = uint16_t(HintChannel(first & 0xFFu, sun, sun_visible) | (HintChannel(first >> 8, lamp, lamp_visible) << 8));
}

// The sky SH kernel (shaders/render/caelix_sky_sh.hlsl): projects the cubemap at mip `lod` onto the 9
// SH of bands 0..2 over the 6 x 32 x 32 texel-centre directions (24 per thread) with the solid angles
// (2 / 32)^2 / (1 + s^2 + t^2)^1.5, renormalized by 4 pi / their sum, and convolved with the clamped
// cosine (A_l = pi, 2 pi / 3, pi / 4), so that E(n) = sum c_i Y_i(n). One workgroup of 256 threads; the
// sums of the threads are added by a tree over the workgroup (strides 128 .. 1).
constexpr uint32_t kShThreads = 256;
constexpr uint32_t kShFace = 32;

// CubeTexelDirection (caelix_sky_sh.hlsl:41-58): the direction of face coordinates (s, t) in [-1, 1].
CX_DEV V3 CubeTexelDirection(uint32_t face, float s, float t)
{
    switch (face)
    {
        case 0u:
            return Make(1.0f, -t, -s);
        case 1u:
            return Make(-1.0f, -t, s);
        case 2u:
            return Make(s, 1.0f, t);
        case 3u:
            return Make(s, -1.0f, -t);
        case 4u:
            return Make(s, -t, 1.0f);
        default:
            return Make(-s, -t, -1.0f);
    }
}

CX_DEV float ShWorkgroupSum(float* lds, float v)
{
    const uint32_t t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (uint32_t stride = kShThreads / 2; stride > 0; stride >>= 1)
    {
        if (t < stride)
        {
            lds[t] = lds[t] + lds[t + stride];
        }
        __syncthreads();
    }
    const float total = lds[0];
    __syncthreads();
    return total;
}

__global__ void __launch_bounds__(kShThreads) SkyShKernel(SkyParams k, float lod, float4* sh)
{
    __shared__ float lds[kShThreads];
    const uint32_t t = threadIdx.x;
    float sums[9][3];
#pragma unroll
    for (int i = 0; i < 9; ++i)
    {
        sums[i][0] = sums[i][1] = sums[i][2] = 0.0f;
    }
    float weight_sum = 0.0f;
    const float texel = 2.0f / float(kShFace);
    for (uint32_t i = t; i < 6 * kShFace * kShFace; i += kShThreads)
    {
        const uint32_t face = i / (kShFace * kShFace);
        const uint32_t x = i % kShFace;
        const uint32_t y = (i / kShFace) % kShFace;
        const float s = (float(x) + 0.5f) * texel - 1.0f;
        const float tt = (float(y) + 0.5f) * texel - 1.0f;
        const V3 v = CubeTexelDirection(face, s, tt);
        const V3 direction = Mul(v, Div(1.0f, Sqrt(Dot(v, v))));
        const float r2 = 1.0f + s * s + tt * tt;
        const float weight = Div(texel * texel, r2 * Sqrt(r2));
        const V3 radiance = CubeSample(k, direction, lod);
        float basis[9];
        ShBasis(direction, basis);
#pragma unroll
        for (int b = 0; b < 9; ++b)
        {
            const float w = basis[b] * weight;
            sums[b][0] += radiance.x * w;
            sums[b][1] += radiance.y * w;
            sums[b][2] += radiance.z * w;
        }
        weight_sum += weight;
    }
    const float normalization = Div(4.0f * kPi, ShWorkgroupSum(lds, weight_sum));
    for (int b = 0; b < 9; ++b)
    {
        const float band = b == 0 ? kPi : (b < 4 ? 2.0f * kPi / 3.0f : 0.25f * kPi);
        float c[3];
        for (int ch = 0; ch < 3; ++ch)
        {
            c[ch] = ShWorkgroupSum(lds, sums[b][ch]) * (band * normalization);
        }
        if (t == 0)
        {
            sh[b] = make_float4(c[0], c[1], c[2], 0.0f);
        }
    }
}

// Zeroes the albedo gradient rows of the workspace.
__global__ void __launch_bounds__(256) ClearRowsKernel(float* rows, uint64_t count)
{
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < count)
    {
        rows[i] = 0.0f;
    }
}

// grad_albedo[id][c] += sum over the rows (one thread per entry, a fixed order: deterministic).
__global__ void __launch_bounds__(256) ReduceRowsKernel(const float* rows, uint32_t row_count, float* grad_albedo)
{
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= kGradEntries)
    {
        return;
    }
    float sum = 0.0f;
    for (uint32_t r = 0; r < row_count; ++r)
    {
        sum += rows[uint64_t(r) * kGradEntries + e];
    }
    grad_albedo[(e / 3) * 4 + e % 3] += sum;
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

__global__ void __launch_bounds__(256) SampleKernel(LightRef lights, const float* receivers,
                                                    const uint32_t* seeds, uint32_t count, float* samples,
                                                    uint32_t* ok)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
    {
        return;
    }
    const LightView l = OpenLights(lights);
    const float* r = receivers + uint64_t(i) * 6;
    uint32_t rng = seeds[i];
    LightPoint p;
    const bool drawn = LightTreeSample(l, Make(r[0], r[1], r[2]), Make(r[3], r[4], r[5]), rng, p);
    float* o = samples + uint64_t(i) * 12;
    o[0] = p.position.x;
    o[1] = p.position.y;
    o[2] = p.position.z;
    o[3] = p.normal.x;
    o[4] = p.normal.y;
    o[5] = p.normal.z;
    o[6] = p.pdf;
    o[7] = __uint_as_float(p.block);
    o[8] = __uint_as_float(p.face);
    o[9] = __uint_as_float(p.group);
    o[10] = __uint_as_float(p.entry);
    o[11] = __uint_as_float(p.voxel_in_group);
    ok[i] = drawn ? 1u : 0u;
}

__global__ void __launch_bounds__(256) PointPdfKernel(LightRef lights, const uint32_t* queries, uint32_t count,
                                                      float* pdf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
    {
        return;
    }
    const LightView l = OpenLights(lights);
    const uint32_t* q = queries + uint64_t(i) * 12;
    const V3 x = Make(__uint_as_float(q[0]), __uint_as_float(q[1]), __uint_as_float(q[2]));
    const V3 n = Make(__uint_as_float(q[3]), __uint_as_float(q[4]), __uint_as_float(q[5]));
    const uint32_t voxel[3] = {q[7], q[8], q[9]};
    bool found;
    pdf[i] = LightPointPdf(l, x, n, q[6], voxel, q[10], q[11], found);
}

// LightTreeSampleBrickBox per box receiver: boxes[i] = lo(3), hi(3), n(3); out[i] = {ok, group, entry,
// generation, leaf, pdf_brick bits, importance bits, 0}.
__global__ void __launch_bounds__(256) BrickBoxSampleKernel(LightRef lights, const float* boxes, const uint32_t* seeds,
                                                            uint32_t count, uint32_t* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
    {
        return;
    }
    const LightView l = OpenLights(lights);
    const float* b = boxes + uint64_t(i) * 9;
    uint32_t rng = seeds[i];
    BrickRef brick;
    float pdf, importance;
    uint32_t leaf;
    const bool ok = SampleBrickBox(l, Make(b[0], b[1], b[2]), Make(b[3], b[4], b[5]), Make(b[6], b[7], b[8]), rng, brick,
                                   pdf, leaf, importance);
    uint4* o = reinterpret_cast<uint4*>(out + uint64_t(i) * 8);
    o[0] = make_uint4(ok ? 1u : 0u, brick.group, brick.entry, brick.generation);
    o[1] = make_uint4(leaf, __float_as_uint(pdf), __float_as_uint(importance), 0u);
}

// LightTreeBrickPdf (point) or LightTreeBrickPdfBox (box) per query: lo(3), hi(3), n(3) as floats; group,
// entry, generation. A point receiver is lo (hi is not read).
template <bool kPoint>
__global__ void __launch_bounds__(256) BrickPdfKernel(LightRef lights, const uint32_t* queries, uint32_t count, float* pdf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
    {
        return;
    }
    const LightView l = OpenLights(lights);
    const uint32_t* q = queries + uint64_t(i) * 12;
    const float* f = reinterpret_cast<const float*>(q);
    const V3 lo = Make(f[0], f[1], f[2]);
    const Receiver r{lo, kPoint ? lo : Make(f[3], f[4], f[5]), Make(f[6], f[7], f[8])};
    pdf[i] = BrickPdfR<kPoint>(l, r, BrickRef{q[9], q[10], q[11]});
}

// The reference probe's support query (tests/shaders/caelix_light_nee_probe.hlsl mode 0): lo(3), x.x,
// hi(3), x.y, n(3), x.z as floats; group, voxel in group (x | y << 10 | z << 20), face, block. The face
// found from the point x in front of it (LightTreeFindLightFace), then its box pdf
// (LightTreeBrickPdfBox) and target (LightTreeLeafImportanceBox): out[i] = (found, has leaf, pdf, target).
__global__ void __launch_bounds__(256) BoxSupportKernel(LightRef lights, const uint32_t* queries, uint32_t count,
                                                        float* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
    {
        return;
    }
    const LightView l = OpenLights(lights);
    const uint32_t* q = queries + uint64_t(i) * 16;
    const float* f = reinterpret_cast<const float*>(q);
    const V3 lo = Make(f[0], f[1], f[2]);
    const V3 hi = Make(f[4], f[5], f[6]);
    const V3 n = Make(f[8], f[9], f[10]);
    const V3 front = Make(f[3], f[7], f[11]);
    const uint32_t v[3] = {q[13] & kLtVoxelMask, (q[13] >> kLtVoxelBits) & kLtVoxelMask, q[13] >> (2 * kLtVoxelBits)};
    BrickRef brick;
    uint32_t leaf;
    float in_brick;
    float4 result = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (FindLightFace(l, front, q[12], v, q[14], q[15], brick, leaf, in_brick))
    {
        result.x = 1.0f;
        result.y = leaf != kLtNone ? 1.0f : 0.0f;
        result.z = BrickPdfR<false>(l, Receiver{lo, hi, n}, brick);
        result.w = LeafImportanceBox(l, brick, lo, hi, n);
    }
    reinterpret_cast<float4*>(out)[i] = result;
}

// The cubemap sky read as the frame reads it, per query (dir(3), lod): out = rgb.
__global__ void __launch_bounds__(256) SkyCubeSampleKernel(SkyParams k, const float* queries, uint32_t count,
                                                           float* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
    {
        return;
    }
    const float* q = queries + uint64_t(i) * 4;
    const V3 c = CubeSample(k, Make(q[0], q[1], q[2]), q[3]);
    out[3 * uint64_t(i)] = c.x;
    out[3 * uint64_t(i) + 1] = c.y;
    out[3 * uint64_t(i) + 2] = c.z;
}

// LightNeeMisWeight of both sides per query: q = {p_area, d, cos_l, n_r (3), w (3)} -> {w_nee, w_bsdf}.
__global__ void __launch_bounds__(256) MisWeightKernel(const float* queries, uint32_t count, float* weights)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
    {
        return;
    }
    const float* q = queries + uint64_t(i) * 9;
    const V3 n = Make(q[3], q[4], q[5]);
    const V3 w = Make(q[6], q[7], q[8]);
    weights[2 * uint64_t(i)] = MisWeight(q[0], n, w, q[1], q[2], true);
    weights[2 * uint64_t(i) + 1] = MisWeight(q[0], n, w, q[1], q[2], false);
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
    if (l == nullptr || l->descriptor == nullptr)
    {
        return ArgError("%s: no light tree descriptor", what);
    }
    return CX_OK;
}

// Tunables of the trace (cx_trace_set_tuning): the ray-queue depth a wave claims per refill, and the
// persistent traversal's workgroups per CU (0: as many as fit).
CxTraceTuning g_tuning{64u, 0u};

uint32_t DeviceCUs()
{
    static int cus = 0;
    if (cus == 0)
    {
        int device = 0;
        if (hipGetDevice(&device) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        {
            cus = 1;
        }
    }
    return uint32_t(cus);
}

uint32_t TunedBlocks(uint32_t fit)
{
    return g_tuning.workgroups_per_cu == 0 ? fit : std::min(fit, g_tuning.workgroups_per_cu * DeviceCUs());
}

// Workgroups of a persistent kernel: as many as fit on the device at once.
template <class Kernel>
uint32_t PersistentBlocks(Kernel kernel)
{
    static uint32_t blocks = 0;
    if (blocks == 0)
    {
        int device = 0;
        int cus = 0;
        int per_cu = 0;
        if (hipGetDevice(&device) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kWaveThreads, 0) != hipSuccess)
        {
            return 0;
        }
        blocks = uint32_t(std::max(1, per_cu) * std::max(1, cus));
    }
    return blocks;
}

// The sky constants of `params`: all zero without CX_TRACE_SKY; the sun's NEE runs unless
// CX_TRACE_NO_SUN_NEE and when a component of its radiance is above 0 (caelix_path_trace.hlsl:742).
SkyParams SkyOf(const CxTraceParams* params)
{
    SkyParams k{};
    if ((params->flags & CX_TRACE_SKY) == 0)
    {
        return k;
    }
    auto v = [](const float* x) { return V3{x[0], x[1], x[2]}; };
    k.on = 1u;
    const bool radiant = params->sun_radiance[0] > 0.0f || params->sun_radiance[1] > 0.0f || params->sun_radiance[2] > 0.0f;
    k.sun = (params->flags & CX_TRACE_NO_SUN_NEE) == 0 && radiant ? 1u : 0u;
    k.sun_dir = v(params->sun_dir);
    k.sun_solid_angle = params->sun_solid_angle;
    k.sun_radiance = v(params->sun_radiance);
    k.sun_cos_radius = params->sun_cos_radius;
    k.zenith = v(params->sky_zenith);
    k.horizon = v(params->sky_horizon);
    k.ground = v(params->sky_ground);
    if (params->sky_cubemap != nullptr)
    {
        k.cube = reinterpret_cast<const float4*>(params->sky_cubemap->texels);
        k.cube_size = params->sky_cubemap->size;
        k.cube_mips = params->sky_cubemap->mip_levels;
    }
    return k;
}

// The launches of a frame (g.d_color null) or of its adjoint replay (set), in the same order, for the
// scenes of `views` (paths = all of their paths).
template <class Views>
void RecordFrame(const SceneView& sv, const Views& views, uint32_t paths, const CxTraceParams* params, uint32_t mis,
                 const GradView& g, const Workspace& w, const FrameOut& o, hipStream_t s,
                 const CameraView* frame_camera = nullptr)
{
    const bool adjoint = g.d_color != nullptr;
    // The full material model: the scene's extended material table (or a delta chain) asks for it.
    // Subsurface scattering: the flag with a subsurface table (the walk lives in the full model's kernels).
    const bool subsurface = (params->flags & CX_TRACE_SUBSURFACE) != 0 && sv.subsurface != nullptr;
    const bool glass = subsurface || sv.materials_ext != nullptr || params->bounces_transparent > 0u ||
                       (params->flags & CX_TRACE_SHADOW_TRANSMISSION) != 0;
    const uint32_t blocks = DivUp(paths, kWaveThreads);
    const uint32_t rows = GradRows(paths);
    const uint32_t bounce_blocks = std::min(TunedBlocks(PersistentBlocks(PersistentTraceKernel<BounceSource<Views>>)), rows);
    const uint32_t shadow_blocks = std::min(TunedBlocks(PersistentBlocks(PersistentTraceKernel<ShadowSource<Views>>)), rows);
    const uint32_t primary_blocks =
        std::min(TunedBlocks(PersistentBlocks(PersistentTraceKernel<PrimarySource<Views>>)), blocks);
    const uint32_t chunk = g_tuning.queue_chunk;
    if (adjoint)
    {
        const uint64_t n = uint64_t(rows) * kGradEntries;
        ClearRowsKernel<<<uint32_t(DivUp(n, uint64_t(256))), 256, 0, s>>>(w.partials, n);
    }
    const CxSkyCubemap* cube = (params->flags & CX_TRACE_SKY) != 0 ? params->sky_cubemap : nullptr;
    if (cube != nullptr)
    {
        // the irradiance SH of the cubemap, at the mip whose face is about 32 texels
        SkyParams k = SkyOf(params);
        const float max_lod = float(cube->mip_levels - 1);
        const float sh_lod = std::min(std::max(std::log2(float(cube->size)) - 5.0f, 0.0f), max_lod);
        StageScope scope(s, "trace_sky_sh");
        SkyShKernel<<<1, kShThreads, 0, s>>>(k, sh_lod, w.sky_sh);
    }
    const ShadowSource<Views> shadow_source{w, views, sv, g, 0u};
    // the visibility hint: one scene's frame (frame_camera) only
    const CxLightHint* hint = frame_camera != nullptr ? params->hint : nullptr;
    const ShadowSource<Views> primary_shadow_source{w, views, sv, g, hint != nullptr ? 1u : 0u};
    for (uint32_t sample = 0; sample < params->spp; ++sample)
    {
        TraceParams p{params->bounces, params->frame_seed + sample, mis,
                      (params->flags & CX_TRACE_NO_RUSSIAN_ROULETTE) != 0 ? 0u : 1u, SkyOf(params),
                      V3{params->fog_color[0], params->fog_color[1], params->fog_color[2]}, params->fog_density,
                      params->bounces_transparent, (params->flags & CX_TRACE_SHADOW_TRANSMISSION) != 0 ? 1u : 0u,
                      (params->flags & CX_TRACE_DELTA_CHECKERBOARD) != 0 ? 1u : 0u,
                      params->bounce_nee_probability > 0.0f ? params->bounce_nee_probability : 1.0f, HintParams{},
                      subsurface ? 1u : 0u};
        if (cube != nullptr)
        {
            p.sky.sh = w.sky_sh;
        }
        if (hint != nullptr)
        {
            const CameraView& cam = *frame_camera;
            p.hint.previous = hint->previous;
            p.hint.shift = hint->patch_shift;
            p.hint.columns = (cam.width + (1u << hint->patch_shift) - 1u) >> hint->patch_shift;
            p.hint.reset = hint->reset != 0u ? 1u : 0u;
            p.hint.has_prev = hint->previous_camera != nullptr ? 1u : 0u;
            p.hint.sample = sample;
            if (hint->previous_camera != nullptr)
            {
                p.hint.prev = CameraOf(*hint->previous_camera);
            }
        }
        ResetKernel<<<1, 64, 0, s>>>(w);
        {
            StageScope scope(s, adjoint ? "adjoint_trace" : "trace_primary");
            PersistentTraceKernel<<<primary_blocks, kWaveThreads, 0, s>>>(sv, PrimarySource<Views>{w, views, paths},
                                                                          o.counters, w, false, kPrimaryChunk);
        }
        {
            StageScope scope(s, adjoint ? "adjoint_shade" : "trace_shade_primary");
            if (glass)
            {
                PrimaryShadeKernel<Views, true><<<blocks, kWaveThreads, 0, s>>>(sv, views, p, w, paths, g, o.counters);
            }
            else
            {
                PrimaryShadeKernel<Views, false><<<blocks, kWaveThreads, 0, s>>>(sv, views, p, w, paths, g, o.counters);
            }
        }
        {
            StageScope scope(s, adjoint ? "adjoint_trace" : "trace_shadow");
            PersistentTraceKernel<<<shadow_blocks, kWaveThreads, 0, s>>>(sv, primary_shadow_source, o.counters, w,
                                                                         adjoint, chunk);
        }
        for (uint32_t bounce = 1; bounce <= params->bounces; ++bounce)
        {
            {
                StageScope scope(s, adjoint ? "adjoint_trace" : "trace_extend");
                PersistentTraceKernel<<<bounce_blocks, kWaveThreads, 0, s>>>(
                    sv, BounceSource<Views>{w, (bounce - 1) & 1u, views}, o.counters, w, false, chunk);
            }
            {
                StageScope scope(s, adjoint ? "adjoint_shade" : "trace_shade_bounce");
                if (glass)
                {
                    BounceShadeKernel<Views, true><<<blocks, kWaveThreads, 0, s>>>(sv, views, p, w, bounce, paths, g,
                                                                                   o.counters);
                }
                else
                {
                    BounceShadeKernel<Views, false><<<blocks, kWaveThreads, 0, s>>>(sv, views, p, w, bounce, paths, g,
                                                                                    o.counters);
                }
            }
            {
                StageScope scope(s, adjoint ? "adjoint_trace" : "trace_shadow");
                PersistentTraceKernel<<<shadow_blocks, kWaveThreads, 0, s>>>(sv, shadow_source, o.counters, w,
                                                                             adjoint, chunk);
            }
        }
        {
            StageScope scope(s, adjoint ? "adjoint_shade" : "trace_finish");
            FinishKernel<<<blocks, kWaveThreads, 0, s>>>(views, w, paths, sample, o, g);
        }
    }
    if (adjoint)
    {
        ReduceRowsKernel<<<DivUp(kGradEntries, 256u), 256, 0, s>>>(w.partials, rows, g.grad_albedo);
    }
    if (hint != nullptr)
    {
        const CameraView& cam = *frame_camera;
        const uint32_t patch = 1u << hint->patch_shift;
        const uint32_t columns = (cam.width + patch - 1u) >> hint->patch_shift;
        const uint32_t patch_rows = (cam.height + patch - 1u) >> hint->patch_shift;
        StageScope scope(s, "trace_hint");
        HintUpdateKernel<<<DivUp(columns * patch_rows, 256u), 256, 0, s>>>(w, cam, hint->patch_shift, columns, patch_rows,
                                                                          hint->current, hint->read);
    }
}

int CheckParams(const CxTraceParams* params, const CxTraceOutput* out, uint32_t width, uint32_t height, bool adjoint,
                const char* what)
{
    if (params == nullptr || out == nullptr || out->counters == nullptr)
    {
        return ArgError("%s: NULL params, output or counters", what);
    }
    if (width == 0 || height == 0 || params->spp == 0)
    {
        return ArgError("%s: empty image or 0 samples per pixel", what);
    }
    if (params->bounces > 255)
    {
        return ArgError("%s: more than 255 bounces", what);
    }
    if (adjoint && params->bounces > 4)
    {
        return ArgError("%s: at most 4 bounces (3 albedo vertices per path)", what);
    }
    if ((params->flags & ~(CX_TRACE_NO_RUSSIAN_ROULETTE | CX_TRACE_SKY | CX_TRACE_NO_SUN_NEE | CX_TRACE_SHADOW_TRANSMISSION |
                           CX_TRACE_DELTA_CHECKERBOARD | CX_TRACE_SUBSURFACE)) != 0)
    {
        return ArgError("%s: unknown flags", what);
    }
    if (!(params->fog_density >= 0.0f) || !(params->fog_density < 3.0e38f) || !(params->fog_color[0] >= 0.0f) ||
        !(params->fog_color[1] >= 0.0f) || !(params->fog_color[2] >= 0.0f) || !(params->fog_color[0] < 3.0e38f) ||
        !(params->fog_color[1] < 3.0e38f) || !(params->fog_color[2] < 3.0e38f))
    {
        return ArgError("%s: fog density or colour negative or not finite", what);
    }
    if (params->bounces_transparent > 15)
    {
        return ArgError("%s: bounces_transparent above 15", what);
    }
    if (!(params->bounce_nee_probability >= 0.0f) || !(params->bounce_nee_probability <= 1.0f))
    {
        return ArgError("%s: bounce_nee_probability outside [0, 1]", what);
    }
    if (adjoint && params->bounce_nee_probability > 0.0f && params->bounce_nee_probability < 1.0f)
    {
        Log(3, "%s: the adjoint runs every bounce NEE (bounce_nee_probability below 1)", what);
        return CX_ERROR_UNSUPPORTED;
    }
    if (adjoint && ((params->flags & (CX_TRACE_SKY | CX_TRACE_SHADOW_TRANSMISSION)) != 0 || params->fog_density > 0.0f ||
                    params->bounces_transparent > 0))
    {
        Log(3, "%s: the adjoint has no sky, sun, fog or delta chain", what);
        return CX_ERROR_UNSUPPORTED;
    }
    if (params->hint != nullptr)
    {
        const CxLightHint* h = params->hint;
        if (adjoint)
        {
            Log(3, "%s: the adjoint has no visibility hint", what);
            return CX_ERROR_UNSUPPORTED;
        }
        if (h->previous == nullptr || h->current == nullptr || h->patch_shift > 2)
        {
            return ArgError("%s: visibility hint without its images, or patch_shift above 2", what);
        }
        if (params->spp > 0xFFFFu)
        {
            return ArgError("%s: the visibility hint counts at most 65535 samples per pixel", what);
        }
        if (h->previous_camera != nullptr &&
            (h->previous_camera->width != width || h->previous_camera->height != height))
        {
            return ArgError("%s: the previous camera's image size differs", what);
        }
    }
    if (params->sky_cubemap != nullptr)
    {
        const CxSkyCubemap* c = params->sky_cubemap;
        uint32_t levels = 1;
        while ((c->size >> levels) != 0u)
        {
            ++levels;
        }
        if ((params->flags & CX_TRACE_SKY) == 0 || c->texels == nullptr || c->size == 0 || c->mip_levels == 0 ||
            c->mip_levels > levels)
        {
            return ArgError("%s: sky cubemap without CX_TRACE_SKY or texels, or a size or mip count out of range", what);
        }
    }
    if ((params->flags & CX_TRACE_SKY) != 0)
    {
        const float* d = params->sun_dir;
        const float n2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (!(std::fabs(n2 - 1.0f) < 1e-3f) || !(params->sun_solid_angle >= 0.0f) ||
            !(params->sun_solid_angle <= 4.0f * 3.14159265f) || !(params->sun_cos_radius >= -1.0f) ||
            !(params->sun_cos_radius <= 1.0f))
        {
            return ArgError("%s: sun_dir not a unit vector, or sun solid angle / cos radius out of range", what);
        }
        for (int c = 0; c < 3; ++c)
        {
            if (!(params->sun_radiance[c] >= 0.0f) || !(params->sun_radiance[c] < 3.0e38f))
            {
                return ArgError("%s: sun radiance negative or not finite", what);
            }
        }
    }
    return CX_OK;
}

// The frame (grads null) or its adjoint replay (grads set) of one scene.
int RunFrame(const CxScene* scene, const CxLightTreeView* lights, const CxCamera* camera, const CxTraceParams* params,
             const CxTraceGradients* grads, const CxTraceOutput* out, void* stream, const char* what)
{
    if (int r = CheckScene(scene, what); r != CX_OK)
    {
        return r;
    }
    if (int r = CheckLights(lights, what); r != CX_OK)
    {
        return r;
    }
    if (scene->light_codes == nullptr || camera == nullptr)
    {
        return ArgError("%s: NULL light codes or camera", what);
    }
    if (scene->scene_count > 1)
    {
        return ArgError("%s: a batch of scenes goes through cx_trace_batch", what);
    }
    if (int r = CheckParams(params, out, camera->width, camera->height, grads != nullptr, what); r != CX_OK)
    {
        return r;
    }
    if (grads != nullptr && scene->materials_ext != nullptr)
    {
        Log(3, "%s: the adjoint shades opaque Lambertian voxels only (materials_ext set)", what);
        return CX_ERROR_UNSUPPORTED;
    }
    if (grads != nullptr && (params->flags & CX_TRACE_SUBSURFACE) != 0 && scene->subsurface != nullptr)
    {
        Log(3, "%s: the adjoint has no subsurface walk (CX_TRACE_SUBSURFACE)", what);
        return CX_ERROR_UNSUPPORTED;
    }
    if (uint64_t(camera->width) * camera->height > (uint64_t(1) << 28))
    {
        return ArgError("%s: image larger than 2^28 pixels", what);
    }
    const uint32_t paths = PathCount(camera->width, camera->height);
    if (out->workspace == nullptr || out->workspace_bytes < WorkspaceBytes(paths))
    {
        return ArgError("%s: workspace missing or smaller than cx_trace_workspace_bytes", what);
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const Workspace w = CarveWorkspace(out->workspace, paths);
    FrameOut o{reinterpret_cast<float4*>(out->color), reinterpret_cast<float4*>(out->radiance), out->attrib,
               reinterpret_cast<unsigned long long*>(out->counters)};
    GradView g{nullptr, nullptr, nullptr, 1.0f / float(params->spp)};
    if (grads != nullptr)
    {
        g.d_color = reinterpret_cast<const float4*>(grads->d_color);
        g.grad_albedo = grads->grad_albedo;
        g.grad_scale = grads->grad_scale;
    }
    if (PersistentBlocks(PersistentTraceKernel<BounceSource<OneScene>>) == 0 ||
        PersistentBlocks(PersistentTraceKernel<ShadowSource<OneScene>>) == 0)
    {
        return HipResult(hipErrorInvalidDeviceFunction, "cx_trace_frame: occupancy query");
    }
    const CameraView frame_camera = CameraOf(*camera);
    RecordFrame(SceneOf(*scene), OneScene{LightsOf(*lights), frame_camera}, paths, params, lights->mis, g, w, o, s,
                &frame_camera);
    return LaunchResult(what);
}

// Per-scene device views of a batch: the light tree views, then the cameras.
uint64_t BatchViewBytes(uint32_t scenes)
{
    return Align256(uint64_t(scenes) * sizeof(LightRef)) + uint64_t(scenes) * sizeof(CameraView);
}

}  // namespace
}  // namespace cx

extern "C" {

CX_API uint64_t cx_trace_workspace_bytes(uint32_t width, uint32_t height)
{
    return cx::WorkspaceBytes(cx::PathCount(width, height));
}

CX_API int cx_trace_frame(const CxScene* scene, const CxLightTreeView* lights, const CxCamera* camera,
                          const CxTraceParams* params, const CxTraceOutput* out, void* stream)
{
    return cx::RunFrame(scene, lights, camera, params, nullptr, out, stream, "cx_trace_frame");
}

CX_API int cx_trace_frame_adjoint(const CxScene* scene, const CxLightTreeView* lights, const CxCamera* camera,
                                  const CxTraceParams* params, const CxTraceGradients* grads,
                                  const CxTraceOutput* out, void* stream)
{
    if (grads == nullptr || grads->d_color == nullptr || grads->grad_albedo == nullptr || grads->grad_scale == nullptr)
    {
        return cx::ArgError("cx_trace_frame_adjoint: NULL d_color or gradient buffer");
    }
    return cx::RunFrame(scene, lights, camera, params, grads, out, stream, "cx_trace_frame_adjoint");
}

CX_API int cx_trace_set_tuning(const CxTraceTuning* tuning)
{
    if (tuning == nullptr || tuning->queue_chunk == 0 || tuning->queue_chunk > 4096 || tuning->workgroups_per_cu > 64)
    {
        return cx::ArgError("cx_trace_set_tuning: queue_chunk in 1..4096, workgroups_per_cu in 0..64");
    }
    cx::g_tuning = *tuning;
    return CX_OK;
}

CX_API void cx_trace_get_tuning(CxTraceTuning* tuning)
{
    if (tuning != nullptr)
    {
        *tuning = cx::g_tuning;
    }
}

CX_API uint64_t cx_trace_batch_bytes(uint32_t scenes)
{
    return cx::BatchViewBytes(scenes);
}

CX_API uint64_t cx_trace_batch_workspace_bytes(uint32_t width, uint32_t height, uint32_t scenes)
{
    return cx::WorkspaceBytes(uint32_t(std::min<uint64_t>(uint64_t(cx::PathCount(width, height)) * scenes, 0xFFFFFFFFull)));
}

CX_API int cx_trace_batch_prepare(const CxLightTreeView* lights, const CxCamera* cameras, uint32_t scenes,
                                  void* device_views, void* stream)
{
    if (lights == nullptr || cameras == nullptr || device_views == nullptr || scenes == 0)
    {
        return cx::ArgError("cx_trace_batch_prepare: no scenes or a NULL array");
    }
    std::vector<uint8_t> host(cx::BatchViewBytes(scenes), 0);
    cx::LightRef* lv = reinterpret_cast<cx::LightRef*>(host.data());
    cx::CameraView* cv = reinterpret_cast<cx::CameraView*>(host.data() + cx::Align256(uint64_t(scenes) * sizeof(cx::LightRef)));
    for (uint32_t i = 0; i < scenes; ++i)
    {
        if (int r = cx::CheckLights(&lights[i], "cx_trace_batch_prepare"); r != CX_OK)
        {
            return r;
        }
        if (cameras[i].width != cameras[0].width || cameras[i].height != cameras[0].height)
        {
            return cx::ArgError("cx_trace_batch_prepare: the cameras of a batch share one image size");
        }
        lv[i] = cx::LightsOf(lights[i]);
        cv[i] = cx::CameraOf(cameras[i]);
    }
    return cx::HipResult(hipMemcpyAsync(device_views, host.data(), host.size(), hipMemcpyHostToDevice,
                                        static_cast<hipStream_t>(stream)),
                         "cx_trace_batch_prepare");
}

CX_API int cx_trace_batch(const CxScene* scene, const void* device_views, uint32_t scenes, uint32_t width,
                          uint32_t height, const CxTraceParams* params, const CxTraceOutput* out, void* stream)
{
    const char* what = "cx_trace_batch";
    if (int r = cx::CheckScene(scene, what); r != CX_OK)
    {
        return r;
    }
    if (scene->light_codes == nullptr || device_views == nullptr || scenes == 0 ||
        scenes != (scene->scene_count > 1 ? scene->scene_count : 1u))
    {
        return cx::ArgError("%s: NULL light codes or views, or scenes != scene->scene_count", what);
    }
    if (int r = cx::CheckParams(params, out, width, height, false, what); r != CX_OK)
    {
        return r;
    }
    if (params->hint != nullptr)
    {
        cx::Log(3, "%s: the batch has no visibility hint", what);
        return CX_ERROR_UNSUPPORTED;
    }
    const uint64_t per_scene = cx::PathCount(width, height);
    if (per_scene * scenes > (uint64_t(1) << 30))
    {
        return cx::ArgError("%s: more than 2^30 paths", what);
    }
    const uint32_t paths = uint32_t(per_scene * scenes);
    if (out->workspace == nullptr || out->workspace_bytes < cx::WorkspaceBytes(paths))
    {
        return cx::ArgError("%s: workspace missing or smaller than cx_trace_batch_workspace_bytes", what);
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const cx::Workspace w = cx::CarveWorkspace(out->workspace, paths);
    cx::FrameOut o{reinterpret_cast<float4*>(out->color), reinterpret_cast<float4*>(out->radiance), out->attrib,
                   reinterpret_cast<unsigned long long*>(out->counters)};
    const uint8_t* base = static_cast<const uint8_t*>(device_views);
    const cx::SceneBatch views{reinterpret_cast<const cx::LightRef*>(base),
                               reinterpret_cast<const cx::CameraView*>(base + cx::Align256(uint64_t(scenes) * sizeof(cx::LightRef))),
                               uint32_t(per_scene), width * height};
    if (cx::PersistentBlocks(cx::PersistentTraceKernel<cx::BounceSource<cx::SceneBatch>>) == 0 ||
        cx::PersistentBlocks(cx::PersistentTraceKernel<cx::ShadowSource<cx::SceneBatch>>) == 0)
    {
        return cx::HipResult(hipErrorInvalidDeviceFunction, "cx_trace_batch: occupancy query");
    }
    const cx::GradView g{nullptr, nullptr, nullptr, 1.0f / float(params->spp)};
    // every scene's light view carries its mis flag; the launch takes the first one's
    cx::RecordFrame(cx::SceneOf(*scene), views, paths, params, 1u, g, w, o, s);
    return cx::LaunchResult(what);
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

CX_API int cx_light_sample(const CxLightTreeView* lights, const float* receivers, const uint32_t* seeds,
                           uint32_t count, float* samples, uint32_t* ok, void* stream)
{
    if (int r = cx::CheckLights(lights, "cx_light_sample"); r != CX_OK)
    {
        return r;
    }
    if (receivers == nullptr || seeds == nullptr || samples == nullptr || ok == nullptr)
    {
        return cx::ArgError("cx_light_sample: a buffer is NULL");
    }
    if (count == 0)
    {
        return CX_OK;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::StageScope scope(s, "light_sample");
    cx::SampleKernel<<<cx::DivUp(count, 256), 256, 0, s>>>(cx::LightsOf(*lights), receivers, seeds, count, samples,
                                                           ok);
    return cx::LaunchResult("cx_light_sample");
}

CX_API int cx_light_point_pdf(const CxLightTreeView* lights, const uint32_t* queries, uint32_t count, float* pdf,
                              void* stream)
{
    if (int r = cx::CheckLights(lights, "cx_light_point_pdf"); r != CX_OK)
    {
        return r;
    }
    if (queries == nullptr || pdf == nullptr)
    {
        return cx::ArgError("cx_light_point_pdf: a buffer is NULL");
    }
    if (count == 0)
    {
        return CX_OK;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::StageScope scope(s, "light_point_pdf");
    cx::PointPdfKernel<<<cx::DivUp(count, 256), 256, 0, s>>>(cx::LightsOf(*lights), queries, count, pdf);
    return cx::LaunchResult("cx_light_point_pdf");
}

CX_API int cx_light_sample_brick_box(const CxLightTreeView* lights, const float* boxes, const uint32_t* seeds,
                                     uint32_t count, uint32_t* out, void* stream)
{
    if (int r = cx::CheckLights(lights, "cx_light_sample_brick_box"); r != CX_OK)
    {
        return r;
    }
    if (count == 0)
    {
        return CX_OK;
    }
    if (boxes == nullptr || seeds == nullptr || out == nullptr)
    {
        return cx::ArgError("cx_light_sample_brick_box: a buffer is NULL");
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::BrickBoxSampleKernel<<<cx::DivUp(count, 256), 256, 0, s>>>(cx::LightsOf(*lights), boxes, seeds, count, out);
    return cx::LaunchResult("cx_light_sample_brick_box");
}

CX_API int cx_light_brick_pdf(const CxLightTreeView* lights, const uint32_t* queries, uint32_t count, uint32_t box,
                              float* pdf, void* stream)
{
    if (int r = cx::CheckLights(lights, "cx_light_brick_pdf"); r != CX_OK)
    {
        return r;
    }
    if (count == 0)
    {
        return CX_OK;
    }
    if (queries == nullptr || pdf == nullptr)
    {
        return cx::ArgError("cx_light_brick_pdf: a buffer is NULL");
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (box != 0)
    {
        cx::BrickPdfKernel<false><<<cx::DivUp(count, 256), 256, 0, s>>>(cx::LightsOf(*lights), queries, count, pdf);
    }
    else
    {
        cx::BrickPdfKernel<true><<<cx::DivUp(count, 256), 256, 0, s>>>(cx::LightsOf(*lights), queries, count, pdf);
    }
    return cx::LaunchResult("cx_light_brick_pdf");
}

CX_API int cx_light_box_support(const CxLightTreeView* lights, const uint32_t* queries, uint32_t count, float* out,
                                void* stream)
{
    if (int r = cx::CheckLights(lights, "cx_light_box_support"); r != CX_OK)
    {
        return r;
    }
    if (count == 0)
    {
        return CX_OK;
    }
    if (queries == nullptr || out == nullptr)
    {
        return cx::ArgError("cx_light_box_support: a buffer is NULL");
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::BoxSupportKernel<<<cx::DivUp(count, 256), 256, 0, s>>>(cx::LightsOf(*lights), queries, count, out);
    return cx::LaunchResult("cx_light_box_support");
}

CX_API int cx_sky_cube_sample(const CxSkyCubemap* cube, const float* queries, uint32_t count, float* out,
                              void* stream)
{
    if (cube == nullptr || cube->texels == nullptr || cube->size == 0 || cube->mip_levels == 0 ||
        (count > 0 && (queries == nullptr || out == nullptr)))
    {
        return cx::ArgError("cx_sky_cube_sample: NULL cubemap, texels or buffer, or an empty cube");
    }
    if (count == 0)
    {
        return CX_OK;
    }
    cx::SkyParams k{};
    k.cube = reinterpret_cast<const float4*>(cube->texels);
    k.cube_size = cube->size;
    k.cube_mips = cube->mip_levels;
    cx::SkyCubeSampleKernel<<<cx::DivUp(count, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(k, queries, count,
                                                                                                   out);
    return cx::LaunchResult("cx_sky_cube_sample");
}

CX_API int cx_sky_sh(const CxSkyCubemap* cube, float* sh, void* stream)
{
    if (cube == nullptr || cube->texels == nullptr || cube->size == 0 || cube->mip_levels == 0 || sh == nullptr)
    {
        return cx::ArgError("cx_sky_sh: NULL cubemap, texels or output, or an empty cube");
    }
    cx::SkyParams k{};
    k.cube = reinterpret_cast<const float4*>(cube->texels);
    k.cube_size = cube->size;
    k.cube_mips = cube->mip_levels;
    const float max_lod = float(cube->mip_levels - 1);
    const float lod = std::min(std::max(std::log2(float(cube->size)) - 5.0f, 0.0f), max_lod);
    cx::SkyShKernel<<<1, cx::kShThreads, 0, static_cast<hipStream_t>(stream)>>>(k, lod, reinterpret_cast<float4*>(sh));
    return cx::LaunchResult("cx_sky_sh");
}

CX_API int cx_light_mis_weight(const float* queries, uint32_t count, float* weights, void* stream)
{
    if (count == 0)
    {
        return CX_OK;
    }
    if (queries == nullptr || weights == nullptr)
    {
        return cx::ArgError("cx_light_mis_weight: a buffer is NULL");
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::MisWeightKernel<<<cx::DivUp(count, 256), 256, 0, s>>>(queries, count, weights);
    return cx::LaunchResult("cx_light_mis_weight");
}

}  // extern "C"
