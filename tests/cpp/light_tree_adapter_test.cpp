// The light tree adapter (cx_light_tree_*) driven from C++ through the C ABI only, as a host engine would
// drive it (the reference's tests/light_tree_test.cpp rig: upload, flush, recycle, stats). Self-checking;
// exit code 0 = every check passed. Checks:
//   begin        caelix_light_bvh_begin.hlsl:9-17 over every combination of reason, pass recorded,
//                pass result and trees_changed, for both trees
//   build-up     a 3 x 3 x 3 block of random lamp bricks across group borders, a 4096-byte heap: the
//                passes fail until the heap has grown (light_tree_test.cpp:1700-1722), then light bricks,
//                light faces and both trees' leaves equal a host count of the light faces
//   edit         one brick replaced, it and its face neighbours finalized: the records of the light
//                bricks among them are recomputed, the other light bricks of their groups copied
//   sampling     draws over the tree: the draw's pdf equals the point pdf of the drawn face
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <vector>

#include "caelix_amd/caelix_amd.h"

namespace {

int g_failures = 0;

#define CHECK(cond)                                                            \
    do                                                                         \
    {                                                                          \
        if (!(cond))                                                           \
        {                                                                      \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++g_failures;                                                      \
        }                                                                      \
    } while (0)

#define OK(call) CHECK((call) == CX_OK)

constexpr uint16_t kStone = 0x8000 | 1;
constexpr uint16_t kLamp = 0x8000 | 5;
constexpr uint32_t kShift = 4;

uint32_t g_lcg = 12345u;
uint32_t Next()
{
    g_lcg = g_lcg * 1664525u + 1013904223u;
    return g_lcg >> 8;
}

std::vector<uint16_t> RandomBrick()
{
    std::vector<uint16_t> b(512);
    for (uint16_t& v : b)
    {
        const uint32_t r = Next() % 100;
        v = r < 50 ? 0 : (r < 94 ? kStone : kLamp);
    }
    return b;
}

using Key = std::tuple<int, int, int>;

struct World
{
    std::vector<Key> keys;
    std::vector<std::vector<uint16_t>> payload;
    std::map<Key, uint32_t> record;

    uint16_t At(uint32_t r, int x, int y, int z) const
    {
        int k[3] = {std::get<0>(keys[r]), std::get<1>(keys[r]), std::get<2>(keys[r])};
        int p[3] = {x, y, z};
        for (int a = 0; a < 3; ++a)
        {
            if (p[a] < 0 || p[a] > 7)
            {
                k[a] += p[a] < 0 ? -1 : 1;
                p[a] = (p[a] + 8) & 7;
            }
        }
        const auto it = record.find(Key{k[0], k[1], k[2]});
        return it == record.end() ? 0 : payload[it->second][p[0] + 8 * p[1] + 64 * p[2]];
    }
    // Open faces of the lamp voxels of record r (a neighbour that is not opaque: bit 15 clear).
    uint32_t LightFaces(uint32_t r) const
    {
        static const int d[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        uint32_t faces = 0;
        for (int v = 0; v < 512; ++v)
        {
            if (payload[r][v] != kLamp)
            {
                continue;
            }
            const int x = v & 7, y = (v >> 3) & 7, z = v >> 6;
            for (const auto& f : d)
            {
                faces += (At(r, x + f[0], y + f[1], z + f[2]) & 0x8000) == 0 ? 1u : 0u;
            }
        }
        return faces;
    }
};

template <typename T>
T* DeviceCopy(const std::vector<T>& host)
{
    T* p = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(host.size(), 1) * sizeof(T)) != hipSuccess)
    {
        return nullptr;
    }
    if (!host.empty())
    {
        (void)hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice);
    }
    return p;
}

void* DeviceZero(size_t bytes)
{
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 4)) != hipSuccess)
    {
        return nullptr;
    }
    (void)hipMemset(p, 0, std::max<size_t>(bytes, 4));
    return p;
}

// caelix_light_bvh_begin over every combination.
void TestBegin()
{
    LightControl* control = static_cast<LightControl*>(DeviceZero(sizeof(LightControl)));
    for (uint32_t tree = 0; tree < 2; ++tree)
    {
        for (uint32_t reason = 0; reason < 2; ++reason)
        {
            for (uint32_t recorded = 0; recorded < 2; ++recorded)
            {
                for (uint32_t result : {kLtPassFailed, kLtPassOk, kLtPassNoChange})
                {
                    for (uint32_t changed = 0; changed < 4; ++changed)
                    {
                        LightControl host{};
                        host.pass_result = result;
                        host.trees_changed = changed;
                        host.states[tree].building = 7;  // overwritten
                        (void)hipMemcpy(control, &host, sizeof(host), hipMemcpyHostToDevice);
                        OK(cx_light_bvh_begin(control, &control->states[tree], tree, reason, recorded, nullptr));
                        (void)hipMemcpy(&host, control, sizeof(host), hipMemcpyDeviceToHost);
                        const bool want = reason != 0 ||
                                          (recorded != 0 && result == kLtPassOk && ((changed >> tree) & 1u) != 0);
                        CHECK(host.states[tree].building == (want ? 1u : 0u));
                    }
                }
            }
        }
    }
    (void)hipFree(control);
}

struct Scene
{
    World world;
    uint32_t n = 0;
    std::vector<uint32_t> group_of_brick;
    std::vector<GroupInfo> groups;
    CxScene scene{};
    void* slots = nullptr;
    uint32_t slot_count = 0;
};

Scene MakeScene()
{
    Scene s;
    for (int x = 14; x <= 16; ++x)
    {
        for (int y = 14; y <= 16; ++y)
        {
            for (int z = 14; z <= 16; ++z)
            {
                s.world.record[Key{x, y, z}] = uint32_t(s.world.keys.size());
                s.world.keys.push_back(Key{x, y, z});
                s.world.payload.push_back(RandomBrick());
            }
        }
    }
    s.n = uint32_t(s.world.keys.size());
    std::map<Key, uint32_t> group_index;
    for (const Key& k : s.world.keys)
    {
        group_index[Key{std::get<0>(k) >> kShift, std::get<1>(k) >> kShift, std::get<2>(k) >> kShift}] = 0;
    }
    for (auto& [g, index] : group_index)
    {
        index = uint32_t(s.groups.size());
        GroupInfo info{};
        info.coord[0] = std::get<0>(g);
        info.coord[1] = std::get<1>(g);
        info.coord[2] = std::get<2>(g);
        s.groups.push_back(info);
    }
    std::vector<int32_t> keys;
    for (const Key& k : s.world.keys)
    {
        keys.insert(keys.end(), {std::get<0>(k), std::get<1>(k), std::get<2>(k)});
        s.group_of_brick.push_back(
            group_index[Key{std::get<0>(k) >> kShift, std::get<1>(k) >> kShift, std::get<2>(k) >> kShift}]);
    }
    std::vector<uint16_t> payload;
    for (const auto& b : s.world.payload)
    {
        payload.insert(payload.end(), b.begin(), b.end());
    }
    std::vector<float> materials(65536 * 8, 0.0f);
    materials[kStone * 8 + 0] = materials[kStone * 8 + 1] = materials[kStone * 8 + 2] = 0.5f;
    materials[kLamp * 8 + 4] = materials[kLamp * 8 + 5] = materials[kLamp * 8 + 6] = 60.0f;
    s.slot_count = s.n * 2 + 1;
    s.slots = DeviceZero(uint64_t(s.slot_count) * 16);
    uint32_t* hash_info = static_cast<uint32_t*>(DeviceZero(16));
    void* hash_scratch = DeviceZero(cx_hash_build_scratch_bytes(s.n));
    int32_t* d_keys = DeviceCopy(keys);
    OK(cx_hash_build_bricks(d_keys, s.n, static_cast<HashKeySlot*>(s.slots), s.slot_count, hash_scratch, hash_info,
                            nullptr));
    const uint32_t grid = 32;
    CxScene& c = s.scene;
    c.keys = d_keys;
    c.info = static_cast<uint32_t*>(DeviceZero(uint64_t(s.n) * 8));
    c.occupancy = DeviceZero(uint64_t(s.n) * 64);
    c.payload = DeviceCopy(payload);
    c.slots = static_cast<const HashKeySlot*>(s.slots);
    c.slot_count = s.slot_count;
    c.bricks = s.n;
    c.hash_info = hash_info;
    c.brick_bits = static_cast<uint32_t*>(DeviceZero(grid * grid * grid / 8));
    c.super_bits = static_cast<uint32_t*>(DeviceZero(64));
    c.grid_dim = grid;
    c.group_shift = kShift;
    c.encoding = kEncodingBlockId;
    c.materials = DeviceCopy(materials);
    c.emission_scale = 1.0f;
    c.group_of_brick = DeviceCopy(s.group_of_brick);
    return s;
}

struct Flusher
{
    CxLightTree* tree;
    Scene* s;
    std::vector<float> materials;
    CxLightSceneTables tables{};
    uint64_t frame = 0;

    // One flush: prepare, stage, transfers, finalize (all records or the listed ones), record, end, recycle.
    void Flush(const std::vector<uint32_t>* list)
    {
        ++frame;
        uint32_t requeue = 0;
        uint32_t staged = 0;
        OK(cx_light_tree_prepare(tree, materials.data(), &tables, &requeue));
        const bool finalize = frame == 1 || list != nullptr;
        OK(cx_light_tree_stage(tree, finalize ? 1u : 0u, &staged));
        OK(cx_light_tree_record_transfers(tree, nullptr));
        s->scene.light_codes = cx_light_tree_light_codes(tree);
        s->scene.light_group_rows = cx_light_tree_group_rows(tree);
        if (frame == 1)
        {
            OK(cx_scene_finalize(&s->scene, nullptr));
        }
        else if (list != nullptr)
        {
            uint32_t* d_list = DeviceCopy(*list);
            OK(cx_scene_finalize_bricks(&s->scene, d_list, uint32_t(list->size()), nullptr));
            (void)hipDeviceSynchronize();
            (void)hipFree(d_list);
        }
        OK(cx_light_tree_record(tree, &s->scene, frame, nullptr));
        OK(cx_light_tree_end_flush(tree, frame));
        CHECK(hipDeviceSynchronize() == hipSuccess);
        OK(cx_light_tree_recycle(tree, frame));
    }
    int Settle(const std::vector<uint32_t>* list)
    {
        for (int i = 0; i < 16; ++i)
        {
            Flush(i == 0 ? list : nullptr);
            if (cx_light_tree_pending(tree) == 0)
            {
                return i + 1;
            }
        }
        CHECK(false);
        return 16;
    }
};

}  // namespace

int main()
{
    TestBegin();
    Scene s = MakeScene();
    std::vector<float> materials(65536 * 8, 0.0f);
    materials[kLamp * 8 + 4] = materials[kLamp * 8 + 5] = materials[kLamp * 8 + 6] = 60.0f;
    CxLightOptions options{1, 1.0f, 4096};
    CxLightTree* tree = nullptr;
    OK(cx_light_tree_create(kShift, kEncodingBlockId, &options, &tree));
    if (tree == nullptr)
    {
        std::fprintf(stderr, "no light tree\n");
        return 1;
    }
    std::vector<EntityInfo> entities(1);
    std::memset(entities.data(), 0, sizeof(EntityInfo));
    for (int r = 0; r < 3; ++r)
    {
        entities[0].cur_rows[r][r] = 1.0f;
        entities[0].prev_rows[r][r] = 1.0f;
    }
    GroupInfo* d_groups = DeviceCopy(s.groups);
    EntityInfo* d_entities = DeviceCopy(entities);
    OK(cx_light_tree_on_entity_added(tree, 0, CX_ENTITY_STATIC));
    for (uint32_t g = 0; g < s.groups.size(); ++g)
    {
        OK(cx_light_tree_on_group_created(tree, g));
    }
    Flusher f{tree, &s, materials};
    f.tables = CxLightSceneTables{d_groups, d_entities, uint32_t(s.groups.size()), 1, s.n, 0};
    f.Settle(nullptr);
    CxLightStats st{};
    cx_light_tree_get_stats(tree, &st);
    uint64_t bricks = 0, faces = 0;
    std::vector<uint32_t> count(s.n);
    for (uint32_t r = 0; r < s.n; ++r)
    {
        count[r] = s.world.LightFaces(r);
        bricks += count[r] > 0 ? 1 : 0;
        faces += count[r];
    }
    std::printf("build-up: %llu light bricks, %llu faces, %u passes (%u failed), heap %llu bytes\n",
                (unsigned long long)st.light_bricks, (unsigned long long)st.light_faces, st.passes, st.failed_passes,
                (unsigned long long)st.heap_bytes_capacity);
    CHECK(st.enabled == 1);
    CHECK(bricks > 10 && st.light_bricks == bricks && st.light_faces == faces);
    CHECK(st.tree_leaves[0] == bricks && st.tree_leaves[1] == 0 && st.tree_leaves_dropped[0] == 0);
    CHECK(st.failed_passes >= 1 && st.passes > st.failed_passes);
    CHECK(st.heap_bytes_capacity > 4096 && st.heap_bytes_used <= st.heap_bytes_capacity);
    CHECK(st.records_recomputed == bricks && st.records_copied == 0);
    CHECK(st.light_groups >= 2);

    // edit: the centre brick replaced; it and its six neighbours finalized again
    const uint32_t centre = s.world.record[Key{15, 15, 15}];
    s.world.payload[centre] = RandomBrick();
    (void)hipMemcpy(const_cast<uint16_t*>(s.scene.payload) + uint64_t(centre) * 512, s.world.payload[centre].data(), 1024,
                    hipMemcpyHostToDevice);
    std::vector<uint32_t> list = {centre};
    static const int d[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (const auto& o : d)
    {
        list.push_back(s.world.record[Key{15 + o[0], 15 + o[1], 15 + o[2]}]);
    }
    std::vector<uint32_t> after(s.n);
    std::set<uint32_t> flagged;
    for (uint32_t r = 0; r < s.n; ++r)
    {
        after[r] = s.world.LightFaces(r);
    }
    uint32_t recomputed = 0, copied = 0;
    for (uint32_t r : list)
    {
        if (count[r] != 0 || after[r] != 0)
        {
            flagged.insert(s.group_of_brick[r]);
        }
        recomputed += after[r] != 0 ? 1u : 0u;
    }
    for (uint32_t r = 0; r < s.n; ++r)
    {
        const bool listed = std::find(list.begin(), list.end(), r) != list.end();
        copied += !listed && after[r] != 0 && flagged.count(s.group_of_brick[r]) != 0 ? 1u : 0u;
    }
    f.Settle(&list);
    cx_light_tree_get_stats(tree, &st);
    std::printf("edit: %u recomputed (%u expected), %u copied (%u expected)\n", st.records_recomputed, recomputed,
                st.records_copied, copied);
    CHECK(st.records_recomputed == recomputed && st.records_copied == copied && copied > 0);
    uint64_t faces_after = 0, bricks_after = 0;
    for (uint32_t r = 0; r < s.n; ++r)
    {
        faces_after += after[r];
        bricks_after += after[r] != 0 ? 1 : 0;
    }
    CHECK(st.light_bricks == bricks_after && st.light_faces == faces_after && st.tree_leaves[0] == bricks_after);

    // sampling: the draw's pdf is the point pdf of the drawn face
    const uint32_t draws = 256;
    std::vector<float> receivers(draws * 6);
    std::vector<uint32_t> seeds(draws);
    for (uint32_t i = 0; i < draws; ++i)
    {
        receivers[6 * i + 0] = 100.0f + float(Next() % 200);
        receivers[6 * i + 1] = 100.0f + float(Next() % 200);
        receivers[6 * i + 2] = 300.0f;
        receivers[6 * i + 5] = -1.0f;
        seeds[i] = Next();
    }
    float* d_receivers = DeviceCopy(receivers);
    uint32_t* d_seeds = DeviceCopy(seeds);
    float* d_samples = static_cast<float*>(DeviceZero(draws * 48));
    uint32_t* d_ok = static_cast<uint32_t*>(DeviceZero(draws * 4));
    CxLightTreeView view{cx_light_tree_descriptor(tree), 1, 0};
    OK(cx_light_sample(&view, d_receivers, d_seeds, draws, d_samples, d_ok, nullptr));
    std::vector<float> samples(draws * 12);
    std::vector<uint32_t> ok(draws);
    (void)hipMemcpy(samples.data(), d_samples, draws * 48, hipMemcpyDeviceToHost);
    (void)hipMemcpy(ok.data(), d_ok, draws * 4, hipMemcpyDeviceToHost);
    std::vector<uint32_t> queries;
    std::vector<float> pdfs;
    for (uint32_t i = 0; i < draws; ++i)
    {
        if (ok[i] == 0)
        {
            continue;
        }
        uint32_t w[12];
        std::memcpy(w, &samples[12 * i], 48);
        uint32_t q[12];
        std::memcpy(q, &receivers[6 * i], 24);
        q[6] = w[9];
        q[7] = w[11] & 0x3FF;
        q[8] = (w[11] >> 10) & 0x3FF;
        q[9] = w[11] >> 20;
        q[10] = w[8];
        q[11] = w[7];
        queries.insert(queries.end(), q, q + 12);
        pdfs.push_back(samples[12 * i + 6]);
    }
    CHECK(pdfs.size() > draws / 2);
    uint32_t* d_queries = DeviceCopy(queries);
    float* d_pdf = static_cast<float*>(DeviceZero(pdfs.size() * 4));
    OK(cx_light_point_pdf(&view, d_queries, uint32_t(pdfs.size()), d_pdf, nullptr));
    std::vector<float> point(pdfs.size());
    (void)hipMemcpy(point.data(), d_pdf, point.size() * 4, hipMemcpyDeviceToHost);
    for (size_t i = 0; i < point.size(); ++i)
    {
        CHECK(pdfs[i] > 0.0f && std::fabs(double(point[i]) / double(pdfs[i]) - 1.0) < 1e-4);
    }
    cx_light_tree_destroy(tree);
    std::printf("%s (%d failures)\n", g_failures == 0 ? "PASSED" : "FAILED", g_failures);
    return g_failures == 0 ? 0 : 1;
}
