// Scene kernels for gfx950: the record producer and the light pass of the trace workloads.
//
// finalize     per brick (one 64-lane workgroup, lane = one x row): the rules of
//              shaders/scene/caelix_finalize.hlsl:205-247 per voxel over its six neighbours (in this
//              brick, or in the face-neighbour brick found through the hash table, air when absent):
//              a transparent or air voxel gets the face bits of the neighbours that are not opaque and
//              of another medium, written back into the payload; a voxel is occupied when it is alive
//              (fewer than six opaque neighbours), not 0 and an interface (opaque, or a face bit).
//              Then micro occupancy words, coarse byte, tight bounds and the resident bits of the
//              brick grid and its 8^3-brick super cells for the world traversal.
// light pass   per brick: the open faces of its light voxels (a light material code; the neighbour,
//              in this brick or found through the hash table, not opaque), the light face count into
//              record word 1 (caelix_finalize.hlsl:206-243), then the brick light records in the
//              reference's format (BrickLightRecordHeader + face entries, families in order, entries
//              in voxel order; shaders/light/records/caelix_light_tree_build.hlsl:93-283) and one
//              LightBvhLeaf per light brick in world space (caelix_light_gather_emit.hlsl:36-86 for an
//              identity frame). Face sums: the lane sums over 8 consecutive voxels, added lane by lane.
#include "hash_device.h"
#include "light_block.h"
#include "runtime.h"

#include <algorithm>

namespace cx {
namespace {

struct SceneArgs
{
    const int32_t* keys;
    uint2* info;
    uint64_t* occupancy;
    const uint16_t* payload;
    const HashKeySlot* slots;
    uint32_t slot_count;
    uint32_t pow2;
    const uint32_t* hash_info;
    uint32_t* brick_bits;
    uint32_t* super_bits;
    int32_t grid_min[3];
    uint32_t grid_dim;
    uint32_t bricks;
    uint32_t group_shift;
    uint32_t encoding;
    const float4* materials;
    const uint32_t* light_codes;
    uint32_t key_words;  // 3: (x, y, z); 4: (scene, x, y, z) for a batch of chunk scenes
    int32_t* box;        // per scene: super-cell box of the bricks with a voxel, or null
    uint16_t* payload_out;  // = payload: finalize writes the face bits back
    GroupLightRow* light_rows;        // null: no light state in word 1
    const uint32_t* group_of_brick;
    const uint32_t* list;             // the records to finalize (block x = list index), or null: every record
};

// The (x, y, z) brick coordinates of a record and its scene (0 for one scene).
CX_DEV const int32_t* BrickXyz(const SceneArgs& s, uint32_t record)
{
    return s.keys + uint64_t(record) * s.key_words + (s.key_words - 3);
}
CX_DEV uint32_t BrickScene(const SceneArgs& s, uint32_t record)
{
    return s.key_words == 4 ? uint32_t(s.keys[uint64_t(record) * 4]) : 0u;
}
CX_DEV uint2 KeyFingerprint(uint32_t key_words, uint32_t scene, int32_t x, int32_t y, int32_t z)
{
    if (key_words == 4)
    {
        const uint32_t w[4] = {scene, uint32_t(x), uint32_t(y), uint32_t(z)};
        return Fingerprint(w, 4);
    }
    return BrickFingerprint(x, y, z);
}

CX_DEV bool Opaque(uint32_t block, uint32_t encoding)
{
    return (block & (encoding == kEncodingPacked ? kPackedOpaqueMask : kBlockIdOpaqueMask)) != 0;
}
CX_DEV uint32_t IdMask(uint32_t encoding)
{
    return encoding == kEncodingPacked ? kPackedIdMask : kBlockIdIdMask;
}
// BlockMaterialCode (caelix_record.hlsl:209): an opaque value as it is, a transparent one without its
// face bits.
CX_DEV uint32_t MaterialCode(uint32_t block, uint32_t encoding)
{
    return Opaque(block, encoding) ? block : (block & IdMask(encoding));
}
CX_DEV uint32_t RowBlock(uint4 row, uint32_t x)
{
    const uint32_t w = x < 4 ? (x < 2 ? row.x : row.y) : (x < 6 ? row.z : row.w);
    return (x & 1u) != 0 ? (w >> 16) : (w & 0xFFFFu);
}

// Empty boxes (lo = INT32_MAX, hi = INT32_MIN) for the finalize kernel's atomic min / max.
__global__ void BoxInitKernel(int32_t* box, uint32_t scenes)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < scenes * 6; i += gridDim.x * blockDim.x)
    {
        box[i] = i % 6 < 3 ? INT32_MAX : INT32_MIN;
    }
}

__global__ void FillKernel(uint32_t* p, uint64_t words)
{
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += uint64_t(gridDim.x) * blockDim.x)
    {
        p[i] = 0u;
    }
}

__global__ void __launch_bounds__(64) FinalizeKernel(SceneArgs s)
{
    const uint32_t brick = s.list != nullptr ? s.list[blockIdx.x] : blockIdx.x;
    const uint32_t lane = threadIdx.x;
    __shared__ unsigned long long lds_occ[8];
    __shared__ uint4 lds_rows[64];
    __shared__ uint32_t lds_nbr[6];
    if (lane < 8)
    {
        lds_occ[lane] = 0ull;
    }
    // The own rows (lane = y + 8 z, 8 blocks of 16 bits each) and the six face-neighbour records.
    const uint4 own = reinterpret_cast<const uint4*>(s.payload + uint64_t(brick) * 512)[lane];
    lds_rows[lane] = own;
    if (lane < 6)
    {
        const int32_t* key = BrickXyz(s, brick);
        int32_t nk[3] = {key[0], key[1], key[2]};
        nk[lane >> 1] += (lane & 1u) != 0 ? -1 : 1;
        lds_nbr[lane] = HashGet(s.slots, s.slot_count, s.pow2, s.hash_info[0],
                                KeyFingerprint(s.key_words, BrickScene(s, brick), nk[0], nk[1], nk[2]));
    }
    __syncthreads();
    const uint32_t y = lane & 7u, z = lane >> 3;
    auto neighbour_row = [&](uint32_t f, uint32_t ny, uint32_t nz) {
        const uint32_t record = lds_nbr[f];
        return record == kHashNone ? make_uint4(0u, 0u, 0u, 0u)
                                   : reinterpret_cast<const uint4*>(s.payload + uint64_t(record) * 512)[ny + 8 * nz];
    };
    const uint4 row_yp = y < 7 ? lds_rows[lane + 1] : neighbour_row(2, 0, z);
    const uint4 row_yn = y > 0 ? lds_rows[lane - 1] : neighbour_row(3, 7, z);
    const uint4 row_zp = z < 7 ? lds_rows[lane + 8] : neighbour_row(4, y, 0);
    const uint4 row_zn = z > 0 ? lds_rows[lane - 8] : neighbour_row(5, y, 7);
    const uint32_t block_xp = lds_nbr[0] == kHashNone ? 0u : s.payload[uint64_t(lds_nbr[0]) * 512 + 8 * y + 64 * z];
    const uint32_t block_xn = lds_nbr[1] == kHashNone ? 0u : s.payload[uint64_t(lds_nbr[1]) * 512 + 7 + 8 * y + 64 * z];
    const uint32_t opaque_mask = s.encoding == kEncodingPacked ? kPackedOpaqueMask : kBlockIdOpaqueMask;
    const uint32_t id_mask = IdMask(s.encoding);
    const uint32_t face_shift = s.encoding == kEncodingPacked ? kPackedFaceShift : kBlockIdFaceShift;
    uint32_t lo[3] = {7, 7, 7};
    uint32_t hi[3] = {0, 0, 0};
    bool any = false;
    for (uint32_t x = 0; x < 8; ++x)
    {
        const uint32_t raw = RowBlock(own, x);
        const uint32_t v = (raw & opaque_mask) != 0u ? raw : (raw & id_mask);
        const uint32_t n[6] = {x < 7 ? RowBlock(own, x + 1) : block_xp, x > 0 ? RowBlock(own, x - 1) : block_xn,
                               RowBlock(row_yp, x), RowBlock(row_yn, x), RowBlock(row_zp, x), RowBlock(row_zn, x)};
        uint32_t opaque_count = 0;
        uint32_t faces = 0;
#pragma unroll
        for (uint32_t d = 0; d < 6; ++d)
        {
            const bool neighbour_opaque = (n[d] & opaque_mask) != 0u;
            opaque_count += neighbour_opaque ? 1u : 0u;
            faces |= (!neighbour_opaque && (n[d] & id_mask) != v) ? (1u << d) : 0u;
        }
        const bool opaque = (v & opaque_mask) != 0u;
        const uint32_t result = opaque ? v : (v | (faces << face_shift));
        if (result != raw)
        {
            s.payload_out[uint64_t(brick) * 512 + x + 8 * y + 64 * z] = uint16_t(result);
        }
        if (opaque_count < 6 && result != 0u && (opaque || faces != 0u))
        {
            const uint32_t coarse = (x >> 2) | ((y >> 2) << 1) | ((z >> 2) << 2);
            const uint32_t micro = (x & 3u) | ((y & 3u) << 2) | ((z & 3u) << 4);
            atomicOr(&lds_occ[coarse], 1ull << micro);
            any = true;
            lo[0] = min(lo[0], x);
            lo[1] = min(lo[1], y);
            lo[2] = min(lo[2], z);
            hi[0] = max(hi[0], x);
            hi[1] = max(hi[1], y);
            hi[2] = max(hi[2], z);
        }
    }
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = min(lo[k], uint32_t(__shfl_xor(int(lo[k]), offset, 64)));
            hi[k] = max(hi[k], uint32_t(__shfl_xor(int(hi[k]), offset, 64)));
        }
    }
    const bool any_voxel = Ballot(any) != 0ull;
    __syncthreads();
    if (lane < 8)
    {
        s.occupancy[uint64_t(brick) * 8 + lane] = lds_occ[lane];
    }
    if (lane == 0)
    {
        uint32_t coarse = 0;
        for (int c = 0; c < 8; ++c)
        {
            coarse |= lds_occ[c] != 0ull ? (1u << c) : 0u;
        }
        const int32_t* key = BrickXyz(s, brick);
        const uint64_t scene = BrickScene(s, brick);
        const uint32_t mask = (1u << s.group_shift) - 1u;
        const uint32_t pos = (uint32_t(key[0]) & mask) | ((uint32_t(key[1]) & mask) << s.group_shift) |
                             ((uint32_t(key[2]) & mask) << (2 * s.group_shift));
        const uint32_t bounds = any_voxel ? (lo[0] | (lo[1] << 3) | (lo[2] << 6) | (hi[0] << 9) | (hi[1] << 12) | (hi[2] << 15))
                                          : 0u;
        // With the light state on, the light bits of word 1 are kept for LightFacesKernel (it compares the
        // old face count with the new one); without it they are 0.
        const uint32_t light_bits = s.light_rows != nullptr ? (s.info[brick].y & ~kRecordBoundsMask) : 0u;
        s.info[brick] = make_uint2((pos & kRecordPosMask) | (coarse << kRecordCoarseShift), bounds | light_bits);
        const int32_t gx = key[0] - s.grid_min[0];
        const int32_t gy = key[1] - s.grid_min[1];
        const int32_t gz = key[2] - s.grid_min[2];
        if (any_voxel && uint32_t(gx) < s.grid_dim && uint32_t(gy) < s.grid_dim && uint32_t(gz) < s.grid_dim)
        {
            const uint64_t cells = uint64_t(s.grid_dim) * s.grid_dim * s.grid_dim;
            const uint64_t index = scene * cells + (uint64_t(gz) * s.grid_dim + uint32_t(gy)) * s.grid_dim + uint32_t(gx);
            atomicOr(&s.brick_bits[index >> 5], 1u << (index & 31u));
            const uint32_t sd = s.grid_dim >> 3;
            const uint64_t super_index = scene * (cells / 512) + (uint64_t(gz >> 3) * sd + uint32_t(gy >> 3)) * sd +
                                         uint32_t(gx >> 3);
            // Every brick of a super cell, and every brick of the scene for the box, would hit the same
            // words: the atomics are issued only when they change something. The words only gain bits
            // and the box only widens during the kernel, and the launch starts from cleared caches, so a
            // stale read can cost an extra atomic but never skip a needed one.
            const uint32_t super_bit = 1u << (super_index & 31u);
            if ((s.super_bits[super_index >> 5] & super_bit) == 0u)
            {
                atomicOr(&s.super_bits[super_index >> 5], super_bit);
            }
            if (s.box != nullptr)
            {
                int32_t* box = s.box + scene * 6;
                const int32_t c[3] = {gx >> 3, gy >> 3, gz >> 3};
                for (int k = 0; k < 3; ++k)
                {
                    if (c[k] < box[k])
                    {
                        atomicMin(&box[k], c[k]);
                    }
                    if (c[k] > box[3 + k])
                    {
                        atomicMax(&box[3 + k], c[k]);
                    }
                }
            }
        }
    }
}

CX_DEV bool LightCode(const uint32_t* light_codes, uint32_t block)
{
    return ((light_codes[block >> 5] >> (block & 31u)) & 1u) != 0u;
}

// The block of the voxel (x, y, z) of brick `record` moved by one voxel along face f's direction:
// inside the brick, or in the neighbour brick found through the hash table (air when absent).
CX_DEV uint32_t NeighbourBlock(const SceneArgs& s, uint32_t record, int x, int y, int z, uint32_t f)
{
    const int axis = int(f >> 1);
    const int sign = (f & 1u) != 0 ? -1 : 1;
    int p[3] = {x, y, z};
    p[axis] += sign;
    if (p[axis] >= 0 && p[axis] < 8)
    {
        return s.payload[uint64_t(record) * 512 + p[0] + 8 * p[1] + 64 * p[2]];
    }
    const int32_t* key = BrickXyz(s, record);
    int32_t nk[3] = {key[0], key[1], key[2]};
    nk[axis] += sign;
    p[axis] = (p[axis] + 8) & 7;
    const uint32_t other = HashGet(s.slots, s.slot_count, s.pow2, s.hash_info[0],
                                   KeyFingerprint(s.key_words, BrickScene(s, record), nk[0], nk[1], nk[2]));
    if (other == kHashNone)
    {
        return 0u;
    }
    return s.payload[uint64_t(other) * 512 + p[0] + 8 * p[1] + 64 * p[2]];
}

// Open light faces of voxel v (LtOpenFaces, caelix_light_tree_build.hlsl:92-111): bit f set when v's
// material code is a light code and its face f is open.
CX_DEV uint32_t OpenLightFaces(const SceneArgs& s, uint32_t record, uint32_t v, uint32_t& block)
{
    block = s.payload[uint64_t(record) * 512 + v];
    if (!LightCode(s.light_codes, MaterialCode(block, s.encoding)))
    {
        return 0u;
    }
    const int x = int(v & 7u), y = int((v >> 3) & 7u), z = int(v >> 6);
    uint32_t open = 0;
    for (uint32_t f = 0; f < 6; ++f)
    {
        if (!Opaque(NeighbourBlock(s, record, x, y, z, f), s.encoding))
        {
            open |= 1u << f;
        }
    }
    return open;
}

// The tables of the light tree, as its descriptor names them (device addresses).
struct LightTables
{
    LightControl* control;
    uint32_t* heap[2];
    GroupLightRow* rows;
    const GroupInfo* group_infos;
    const EntityInfo* entity_infos;
    const uint32_t* entity_motion;
    uint32_t group_shift;
    uint32_t group_count;
};

CX_DEV LightTables TablesOf(const LightTreeDescriptor* d)
{
    LightTables t;
    t.control = reinterpret_cast<LightControl*>(d->control);
    t.heap[0] = reinterpret_cast<uint32_t*>(d->heap[0]);
    t.heap[1] = reinterpret_cast<uint32_t*>(d->heap[1]);
    t.rows = reinterpret_cast<GroupLightRow*>(d->group_rows);
    t.group_infos = reinterpret_cast<const GroupInfo*>(d->group_infos);
    t.entity_infos = reinterpret_cast<const EntityInfo*>(d->entity_infos);
    t.entity_motion = reinterpret_cast<const uint32_t*>(d->entity_motion);
    t.group_shift = d->group_shift;
    t.group_count = d->group_count;
    return t;
}

CX_DEV uint32_t LightFaceCount(uint32_t word1)
{
    return (word1 >> kLtRecordFaceCountShift) & kLtRecordFaceCountMask;
}

// The light state of the finalize (caelix_finalize.hlsl:22-29, 326-340), one wave per brick: the open
// faces of its light voxels counted into record word 1; when the count found in word 1 before or the new
// one is not 0, the stale bit is set (and a set one kept) and the brick's group row gets
// kLtGpuLightChanged.
__global__ void __launch_bounds__(64) LightFacesKernel(SceneArgs s, GroupLightRow* rows, const uint32_t* group_of_brick)
{
    const uint32_t record = s.list != nullptr ? s.list[blockIdx.x] : blockIdx.x;
    const uint32_t lane = threadIdx.x;
    if (group_of_brick[record] == kLtNone)
    {
        return;  // a free record
    }
    uint32_t faces = 0;
    for (uint32_t i = 0; i < 8; ++i)
    {
        uint32_t block;
        faces += __popc(OpenLightFaces(s, record, lane * 8 + i, block));
    }
    faces = WaveSum(faces);
    if (lane == 0)
    {
        uint2 w = s.info[record];
        const uint32_t old_word1 = w.y;  // the light bits as the finalize found them
        uint32_t word1 = (w.y & kRecordBoundsMask) | (faces << kLtRecordFaceCountShift) | (old_word1 & kLtRecordStaleBit);
        if (LightFaceCount(old_word1) != 0 || faces != 0)
        {
            word1 |= kLtRecordStaleBit;
            atomicOr(&rows[group_of_brick[record]].gpu_flags, kLtGpuLightChanged);
        }
        w.y = word1;
        s.info[record] = w;
    }
}

// Group modes of a pass (caelix_light_tree_size.hlsl:16-18, caelix_light_tree_build.hlsl:22-24).
constexpr uint32_t kModeCopy = 0;
constexpr uint32_t kModeRebuild = 1;
constexpr uint32_t kModeReleased = 2;
constexpr uint32_t kNoKey = 0xFFFFFFFFu;

// The light pass's scratch (per brick, per sorted position, per group) and its bookkeeping words.
struct PassArgs
{
    const LightTreeDescriptor* descriptor;
    uint32_t* staged;           // [groups] staged kLtCpu* flags
    const uint32_t* group_of_brick;
    uint32_t* keys;             // [bricks] sort keys (group << 3 shift | Morton of the brick in its group)
    uint32_t* values;           // [bricks] brick records, sorted with the keys
    uint32_t* units;            // [bricks + 1] by sorted position: record units of the plan, then offsets
    uint32_t* recompute;        // [bricks + 1] by sorted position: 1 = recomputed record, then offsets
    uint32_t* plans;            // [bricks] by sorted position: the old record (units in the old block) or kLtNone
    uint32_t* mode;             // [groups] kMode*
    uint32_t* first;            // [groups] sorted position of the group's first light brick
    uint32_t* count;            // [groups] the group's light bricks (kept and dropped)
    uint32_t* sizes;            // [groups] block units, then block offsets in the target buffer
    uint32_t* faces;            // [groups] face entries written into the group's records
    uint32_t* totals;           // after the scan: [0] recomputed records, [1] record units
    uint64_t heap_units;
    uint32_t bricks;
    uint32_t groups;
};

// The group-local brick position (x | y << shift | z << 2 shift) and its Morton code.
CX_DEV uint32_t LocalPos(const SceneArgs& s, uint32_t record, uint32_t& morton)
{
    const int32_t* key = BrickXyz(s, record);
    const uint32_t mask = (1u << s.group_shift) - 1u;
    const uint32_t bx = uint32_t(key[0]) & mask, by = uint32_t(key[1]) & mask, bz = uint32_t(key[2]) & mask;
    morton = Morton3(bx, by, bz);
    return bx | (by << s.group_shift) | (bz << (2 * s.group_shift));
}

// Per group index (caelix_light_tree_size.hlsl:37-56, 95-100): the staged flags merged into the row
// (cpu_flags = (cpu_flags & fresh) | staged when staged != 0, the staged word cleared), the group's mode
// (released; rebuild: kLtCpuRebuild or the finalize's kLtGpuLightChanged; else its block is copied), and
// for a released or rebuilt group the pass flags: changed, and the tree of its entity's motion.
__global__ void __launch_bounds__(256) PassFlagsKernel(PassArgs p)
{
    const LightTables t = TablesOf(p.descriptor);
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= p.groups)
    {
        return;
    }
    p.first[g] = kLtNone;
    p.count[g] = 0;
    p.faces[g] = 0;
    GroupLightRow& row = t.rows[g];
    uint32_t cpu_flags = row.cpu_flags;
    const uint32_t staged = p.staged[g];
    if (staged != 0)
    {
        cpu_flags = (cpu_flags & kLtCpuFresh) | staged;
        row.cpu_flags = cpu_flags;
        p.staged[g] = 0;
    }
    const uint32_t mode = (cpu_flags & kLtCpuReleased) != 0 ? kModeReleased
                          : ((cpu_flags & kLtCpuRebuild) != 0 || (row.gpu_flags & kLtGpuLightChanged) != 0)
                              ? kModeRebuild
                              : kModeCopy;
    p.mode[g] = mode;
    if (mode != kModeCopy)
    {
        const uint32_t motion = t.entity_motion[t.group_infos[g].entity] & 1u;
        atomicOr(&t.control->pass_flags, kLtPassFlagChanged | (1u << (kLtPassFlagTreeShift + motion)));
    }
}

// Sort keys of the light bricks (record word 1 counts open light faces) of the groups the pass rebuilds:
// group << 3 shift | the Morton code of the brick in its group, the index entry order; every other brick
// sorts last.
__global__ void __launch_bounds__(256) LightKeyKernel(SceneArgs s, PassArgs p)
{
    const uint32_t record = blockIdx.x * blockDim.x + threadIdx.x;
    if (record >= p.bricks)
    {
        return;
    }
    const uint32_t g = p.group_of_brick[record];
    uint32_t key = kNoKey;
    if (LightFaceCount(s.info[record].y) != 0 && g < p.groups && p.mode[g] == kModeRebuild)
    {
        uint32_t morton;
        LocalPos(s, record, morton);
        key = (g << (3 * s.group_shift)) | morton;
    }
    p.keys[record] = key;
    p.values[record] = record;
}

// Per sorted light brick: its group's first position and light brick count.
__global__ void __launch_bounds__(256) GroupSpanKernel(SceneArgs s, PassArgs p)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.bricks || p.keys[i] == kNoKey)
    {
        return;
    }
    const uint32_t g = p.keys[i] >> (3 * s.group_shift);
    atomicMin(&p.first[g], i);
    atomicAdd(&p.count[g], 1u);
}

CX_DEV uint32_t Kept(uint32_t count)
{
    return min(count, kLtMaxLightBricksPerGroup);
}

// The group's block before the pass in the current heap buffer, or null when it has none or the group
// index starts a new life (LtOpenOldBlock, caelix_light_tree_group.hlsl:189-206).
CX_DEV const uint32_t* OldBlock(const LightTables& t, const GroupLightRow& row)
{
    if (row.block_units == 0 || (row.cpu_flags & kLtCpuFresh) != 0)
    {
        return nullptr;
    }
    return t.heap[t.control->current & 1u] + uint64_t(row.block_offset) * 4;
}

// Per sorted position, the plan of a kept light brick (a group keeps its first
// kLtMaxLightBricksPerGroup light bricks in Morton order; LtPlanBrick, caelix_light_tree_group.hlsl:
// 214-243): its record is copied from the old block when the brick is not stale and the old block has an
// entry at its position, else recomputed; the units of either record. The scan turns units and the
// recompute flags into offsets and totals.
__global__ void __launch_bounds__(256) BrickPlanKernel(SceneArgs s, PassArgs p)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > p.bricks)
    {
        return;
    }
    uint32_t units = 0, recompute = 0, plan = kLtNone;
    if (i < p.bricks && p.keys[i] != kNoKey)
    {
        const uint32_t shift = s.group_shift;
        const uint32_t g = p.keys[i] >> (3 * shift);
        if (i - p.first[g] < kLtMaxLightBricksPerGroup)
        {
            const LightTables t = TablesOf(p.descriptor);
            const uint32_t word1 = s.info[p.values[i]].y;
            units = RecordUnits(LightFaceCount(word1));
            recompute = 1;
            const GroupLightRow row = t.rows[g];
            const uint32_t* old = OldBlock(t, row);
            if ((word1 & kLtRecordStaleBit) == 0 && old != nullptr)
            {
                const uint32_t entry = FindIndexEntry(old, old[0], p.keys[i] & ((1u << (3 * shift)) - 1u), shift);
                if (entry != kLtNone)
                {
                    plan = IndexEntry(old, entry)->record_offset;
                    units = RecordUnits(old[uint64_t(plan) * 4 + 9]);  // its entry_count
                    recompute = 0;
                }
            }
        }
    }
    p.units[i] = units;
    p.recompute[i] = recompute;
    if (i < p.bricks)
    {
        p.plans[i] = plan;
    }
}

// The exclusive prefix at sorted position j (j may be the end: the total).
CX_DEV uint32_t PrefixAt(const uint32_t* prefix, uint32_t j, uint32_t n, uint32_t total)
{
    return j < n ? prefix[j] : total;
}

// Per group: the size of its block after the pass (caelix_light_tree_size.hlsl:58-94): a copied group
// keeps its block_units; a released one 0; a rebuilt one header, index entries and records of its kept
// light bricks (LtBlockUnits, 0 without light bricks).
__global__ void __launch_bounds__(256) GroupSizeKernel(PassArgs p)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= p.groups)
    {
        return;
    }
    const LightTables t = TablesOf(p.descriptor);
    const uint32_t mode = p.mode[g];
    uint32_t size = mode == kModeCopy ? t.rows[g].block_units : 0u;
    const uint32_t count = p.count[g];
    if (mode == kModeRebuild && count > 0)
    {
        const uint32_t first = p.first[g];
        const uint32_t records =
            PrefixAt(p.units, first + count, p.bricks, p.totals[1]) - PrefixAt(p.units, first, p.bricks, p.totals[1]);
        size = 1u + Kept(count) + records;
    }
    p.sizes[g] = size;
}

// One workgroup (caelix_light_tree_scan.hlsl): the block sizes become their exclusive prefix sum (the
// offsets in the target heap buffer, 64-bit carry), and the pass is decided: no released or rebuilt
// group -> kLtPassNoChange; the total fits heap_units -> kLtPassOk; else kLtPassFailed. A pass with
// changes records needed_units and resets the record counters.
__global__ void __launch_bounds__(1024) GroupScanKernel(PassArgs p)
{
    __shared__ uint32_t lds[32];
    __shared__ uint32_t lds_flags;
    const LightTables t = TablesOf(p.descriptor);
    if (threadIdx.x == 0)
    {
        lds_flags = t.control->pass_flags;
    }
    uint64_t carry = 0;
    for (uint32_t base = 0; base < p.groups; base += 1024)
    {
        const uint32_t g = base + threadIdx.x;
        const uint32_t size = g < p.groups ? p.sizes[g] : 0u;
        uint32_t total;
        const uint32_t before = BlockExclusiveSum<1024>(size, lds, total);
        if (g < p.groups)
        {
            const uint64_t offset = carry + before;
            p.sizes[g] = uint32_t(min(offset, uint64_t(0xFFFFFFFFu)));
        }
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        const uint32_t changed = lds_flags & kLtPassFlagChanged;
        const bool fits = carry <= p.heap_units;
        const uint32_t result = changed == 0 ? kLtPassNoChange : (fits ? kLtPassOk : kLtPassFailed);
        t.control->pass_result = result;
        if (changed != 0)
        {
            t.control->needed_units = uint32_t(min(carry, uint64_t(0xFFFFFFFFu)));
            t.control->records_recomputed = 0u;
            t.control->records_copied = 0u;
        }
        t.control->trees_changed = (lds_flags >> kLtPassFlagTreeShift) & ((1u << kLtTrees) - 1u);
        t.control->pass_flags = 0u;
    }
}

// One workgroup per copied group, after a kLtPassOk decision: its block copied as it is to its new
// offset in the target buffer (caelix_light_tree_build.hlsl:363-378).
__global__ void __launch_bounds__(256) BlockCopyKernel(PassArgs p)
{
    const uint32_t g = blockIdx.x;
    const LightTables t = TablesOf(p.descriptor);
    if (p.mode[g] != kModeCopy || t.control->pass_result != kLtPassOk)
    {
        return;
    }
    const GroupLightRow& row = t.rows[g];
    const uint32_t current = t.control->current & 1u;
    const uint4* from = reinterpret_cast<const uint4*>(t.heap[current]) + row.block_offset;
    uint4* to = reinterpret_cast<uint4*>(t.heap[1u - current]) + p.sizes[g];
    for (uint32_t u = threadIdx.x; u < row.block_units; u += blockDim.x)
    {
        to[u] = from[u];
    }
}

// One wave per sorted position: a kept light brick of a rebuilt group writes its index entry (bvh_leaf
// kLtNone until the gather) and its brick light record into the group's block in the target heap buffer
// (LtBuildBrick, caelix_light_tree_build.hlsl:114-282); the group's first brick writes the block header.
// A copied record is copied whole; a recomputed one finds the open faces of the brick's light voxels
// (lane l owns voxels [8 l, 8 l + 8), so each family's entries come out in voxel order), keeps at most
// the finalize's face count of entries (families in order), and sums the face weights (luminance of the
// emission) and the tight box of the kept entries. Only after a kLtPassOk decision.
__global__ void __launch_bounds__(64) BlockBuildKernel(SceneArgs s, PassArgs p)
{
    const uint32_t i = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const uint32_t key = p.keys[i];
    if (key == kNoKey)
    {
        return;
    }
    const LightTables t = TablesOf(p.descriptor);
    if (t.control->pass_result != kLtPassOk)
    {
        return;
    }
    const uint32_t g = key >> (3 * s.group_shift);
    const uint32_t first = p.first[g];
    const uint32_t k = i - first;
    if (k >= kLtMaxLightBricksPerGroup)
    {
        return;
    }
    const uint32_t record = p.values[i];
    const uint32_t kept = Kept(p.count[g]);
    const uint32_t current = t.control->current & 1u;
    uint32_t* block = t.heap[1u - current] + uint64_t(p.sizes[g]) * 4;
    const uint32_t record_units = 1u + kept + (p.units[i] - p.units[first]);
    const uint32_t plan_units = PrefixAt(p.units, i + 1, p.bricks, p.totals[1]) - p.units[i];
    uint32_t* rec = block + uint64_t(record_units) * 4;
    uint32_t morton;
    const uint32_t pos = LocalPos(s, record, morton);
    if (lane == 0)
    {
        *reinterpret_cast<uint4*>(block + 4 + uint64_t(k) * 4) = make_uint4(pos, record, record_units, kLtNone);
        if (k == 0)
        {
            *reinterpret_cast<uint4*>(block) = make_uint4(kept, 0u, 0u, 0u);
        }
    }
    const uint32_t plan = p.plans[i];
    if (plan != kLtNone)
    {
        const uint32_t* from = t.heap[current] + uint64_t(t.rows[g].block_offset + plan) * 4;
        for (uint32_t u = lane; u < plan_units; u += 64)
        {
            reinterpret_cast<uint4*>(rec)[u] = reinterpret_cast<const uint4*>(from)[u];
        }
        if (lane == 0)
        {
            atomicAdd(&p.faces[g], from[9]);
        }
        return;
    }
    __shared__ float lds_sums[6][64];
    uint32_t counts[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t v = 0; v < 8; ++v)
    {
        uint32_t b;
        const uint32_t open = OpenLightFaces(s, record, lane * 8 + v, b);
        for (uint32_t f = 0; f < 6; ++f)
        {
            counts[f] += (open >> f) & 1u;
        }
    }
    uint32_t family_base[6];
    uint32_t family_count[6];
    uint32_t cursor[6];
    uint32_t left = LightFaceCount(s.info[record].y);
    uint32_t written = 0;
    for (uint32_t f = 0; f < 6; ++f)
    {
        const uint32_t inclusive = WaveInclusiveSum(counts[f]);
        const uint32_t found = __shfl(inclusive, 63, 64);
        cursor[f] = inclusive - counts[f];
        family_count[f] = min(found, left);
        family_base[f] = written;
        written += family_count[f];
        left -= family_count[f];
    }
    float sums[6] = {0, 0, 0, 0, 0, 0};
    uint32_t lo[3] = {8, 8, 8};
    uint32_t hi[3] = {0, 0, 0};
    for (uint32_t v8 = 0; v8 < 8; ++v8)
    {
        const uint32_t vi = lane * 8 + v8;
        uint32_t b;
        const uint32_t open = OpenLightFaces(s, record, vi, b);
        if (open == 0)
        {
            continue;
        }
        const float4 e = s.materials[2 * uint64_t(MaterialCode(b, s.encoding)) + 1];
        const float weight = e.x * 0.2126f + e.y * 0.7152f + e.z * 0.0722f;
        bool kept_voxel = false;
        for (uint32_t f = 0; f < 6; ++f)
        {
            if (((open >> f) & 1u) != 0)
            {
                if (cursor[f] < family_count[f])
                {
                    rec[12 + family_base[f] + cursor[f]] = vi | (b << kLtFaceEntryBlockShift);
                    sums[f] += weight;
                    kept_voxel = true;
                }
                ++cursor[f];
            }
        }
        if (kept_voxel)
        {
            const uint32_t x = vi & 7u, y = (vi >> 3) & 7u, z = vi >> 6;
            lo[0] = min(lo[0], x);
            lo[1] = min(lo[1], y);
            lo[2] = min(lo[2], z);
            hi[0] = max(hi[0], x + 1);
            hi[1] = max(hi[1], y + 1);
            hi[2] = max(hi[2], z + 1);
        }
    }
    for (uint32_t f = 0; f < 6; ++f)
    {
        lds_sums[f][lane] = sums[f];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
    {
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
            lo[c] = min(lo[c], uint32_t(__shfl_xor(int(lo[c]), o, 64)));
            hi[c] = max(hi[c], uint32_t(__shfl_xor(int(hi[c]), o, 64)));
        }
    }
    __syncthreads();
    // zero padding up to the end of the record's units (records are copied whole)
    const uint32_t capacity = (plan_units * kLtUnitBytes - 48u) / 4u;
    for (uint32_t e = written + lane; e < capacity; e += 64)
    {
        rec[12 + e] = 0u;
    }
    if (lane != 0)
    {
        return;
    }
    float face[6];
    for (uint32_t f = 0; f < 6; ++f)
    {
        float sum = 0.0f;
        for (uint32_t l = 0; l < 64; ++l)
        {
            sum += lds_sums[f][l];
        }
        face[f] = sum;
    }
    const uint32_t mask = (1u << s.group_shift) - 1u;
    const uint32_t o[3] = {(pos & mask) * 8, ((pos >> s.group_shift) & mask) * 8, (pos >> (2 * s.group_shift)) * 8};
    const bool empty = written == 0;
    for (int c = 0; c < 3; ++c)
    {
        lo[c] = o[c] + (empty ? 0u : lo[c]);
        hi[c] = o[c] + (empty ? 0u : hi[c]);
    }
    for (uint32_t f = 0; f < 6; ++f)
    {
        rec[f] = __float_as_uint(face[f]);
    }
    rec[6] = lo[0] | (lo[1] << kLtVoxelBits) | (lo[2] << (2 * kLtVoxelBits));
    rec[7] = hi[0] | (hi[1] << kLtVoxelBits) | (hi[2] << (2 * kLtVoxelBits));
    rec[8] = pos;
    rec[9] = written;
    rec[10] = family_count[0] | (family_count[1] << kLtFamilyCountBits) | (family_count[2] << (2 * kLtFamilyCountBits));
    rec[11] = family_count[3] | (family_count[4] << kLtFamilyCountBits) | (family_count[5] << (2 * kLtFamilyCountBits));
    atomicAdd(&p.faces[g], written);
}

// Per brick of a rebuilt group, after a kLtPassOk decision: the stale bit of record word 1 cleared
// (caelix_light_tree_build.hlsl:398-412: every resident brick of the group).
__global__ void __launch_bounds__(256) StaleClearKernel(SceneArgs s, PassArgs p)
{
    const uint32_t record = blockIdx.x * blockDim.x + threadIdx.x;
    if (record >= p.bricks)
    {
        return;
    }
    const LightTables t = TablesOf(p.descriptor);
    const uint32_t g = p.group_of_brick[record];
    if (t.control->pass_result != kLtPassOk || g >= p.groups || p.mode[g] != kModeRebuild)
    {
        return;
    }
    const uint32_t word1 = s.info[record].y;
    if ((word1 & kLtRecordStaleBit) != 0)
    {
        s.info[record].y = word1 & ~kLtRecordStaleBit;
    }
}

CX_DEV uint32_t NextLife(uint32_t generation)
{
    return ((generation >> kLtGenerationLifeShift) + 1u) << kLtGenerationLifeShift;
}
CX_DEV uint32_t NextVersion(uint32_t generation)
{
    return (generation & ~kLtGenerationVersionMask) | ((generation + 1u) & kLtGenerationVersionMask);
}

// Per group, after a kLtPassOk decision, its row (caelix_light_tree_build.hlsl:372-388, 414-431): a
// copied group gets its new block offset; a released group's row is cleared into the next life of its
// index (LtNextLife); a rebuilt group gets its offset, units, kept and dropped light bricks, face entries
// and the next version of its life (LtNextVersion), flags cleared.
__global__ void __launch_bounds__(256) RowCommitKernel(PassArgs p)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const LightTables t = TablesOf(p.descriptor);
    if (g >= p.groups || t.control->pass_result != kLtPassOk)
    {
        return;
    }
    GroupLightRow& row = t.rows[g];
    const uint32_t mode = p.mode[g];
    if (mode == kModeCopy)
    {
        row.block_offset = row.block_units != 0 ? p.sizes[g] : 0u;
        return;
    }
    GroupLightRow out;
    out.gpu_flags = 0;
    out.cpu_flags = 0;
    out.pad = 0;
    if (mode == kModeReleased)
    {
        out.block_offset = 0;
        out.block_units = 0;
        out.generation = NextLife(row.generation);
        out.bricks = 0;
        out.light_faces = 0;
        row = out;
        return;
    }
    const uint32_t count = p.count[g];
    const uint32_t kept = Kept(count);
    uint32_t units = 0;
    if (count > 0)
    {
        const uint32_t first = p.first[g];
        units = 1u + kept + (PrefixAt(p.units, first + count, p.bricks, p.totals[1]) -
                             PrefixAt(p.units, first, p.bricks, p.totals[1]));
    }
    out.block_offset = units != 0 ? p.sizes[g] : 0u;
    out.block_units = units;
    out.generation = NextVersion(row.generation);
    out.bricks = kept | (min(count - kept, kLtRowBricksMask) << kLtRowDroppedShift);
    out.light_faces = p.faces[g];
    row = out;
}

// One workgroup closes the pass (caelix_light_tree_commit.hlsl): kLtPassOk -> the target buffer becomes
// current, the totals over the rows are recounted, the records recomputed and copied by the rebuilt
// groups counted, passes + 1; kLtPassFailed -> passes + 1, failed_passes + 1; kLtPassNoChange ->
// nothing.
__global__ void __launch_bounds__(256) PassCommitKernel(PassArgs p)
{
    __shared__ unsigned long long lds_sums[4][256];
    __shared__ uint32_t lds_groups[256];
    const LightTables t = TablesOf(p.descriptor);
    const uint32_t result = t.control->pass_result;
    if (result == kLtPassNoChange)
    {
        return;
    }
    const uint32_t tid = threadIdx.x;
    if (result == kLtPassFailed)
    {
        if (tid == 0)
        {
            t.control->passes += 1u;
            t.control->failed_passes += 1u;
        }
        return;
    }
    unsigned long long bricks = 0, faces = 0, dropped = 0, rebuilt = 0;
    uint32_t groups = 0;
    for (uint32_t g = tid; g < p.groups; g += 256)
    {
        const GroupLightRow& row = t.rows[g];
        bricks += row.bricks & kLtRowBricksMask;
        faces += row.light_faces;
        dropped += row.bricks >> kLtRowDroppedShift;
        groups += (row.bricks & kLtRowBricksMask) != 0 ? 1u : 0u;
        rebuilt += p.mode[g] == kModeRebuild ? Kept(p.count[g]) : 0u;
    }
    lds_sums[0][tid] = bricks;
    lds_sums[1][tid] = faces;
    lds_sums[2][tid] = dropped;
    lds_sums[3][tid] = rebuilt;
    lds_groups[tid] = groups;
    __syncthreads();
    for (uint32_t stride = 128; stride > 0; stride >>= 1)
    {
        if (tid < stride)
        {
            for (int k = 0; k < 4; ++k)
            {
                lds_sums[k][tid] += lds_sums[k][tid + stride];
            }
            lds_groups[tid] += lds_groups[tid + stride];
        }
        __syncthreads();
    }
    if (tid == 0)
    {
        const uint32_t recomputed = p.totals[0];
        t.control->current = 1u - (t.control->current & 1u);
        t.control->light_groups = lds_groups[0];
        t.control->bricks_dropped = uint32_t(min(lds_sums[2][0], 0xFFFFFFFFull));
        t.control->light_bricks = lds_sums[0][0];
        t.control->light_faces = lds_sums[1][0];
        t.control->records_recomputed = recomputed;
        t.control->records_copied = uint32_t(lds_sums[3][0]) - recomputed;
        t.control->passes += 1u;
    }
}

// ---- gather (caelix_light_gather_count.hlsl, caelix_light_bvh_scan.hlsl, caelix_light_gather_emit.hlsl)

struct GatherArgs
{
    const LightTreeDescriptor* descriptor;
    uint32_t* offsets;  // [groups] leaf counts, then exclusive offsets
    LightBvhLeaf* leaves;
    LightTreeState* state;
    const uint32_t* entity_leaf_base;  // null, or per entity the leaf index its bvh_leaf words count from
    uint32_t leaf_capacity;
    uint32_t tree;
    uint32_t groups;
};

CX_DEV const uint32_t* CurrentHeap(const LightTables& t)
{
    return t.heap[t.control->current & 1u];
}

// Per group (caelix_light_gather_count.hlsl:9-35): its leaves for this tree: the block's record count
// when the group has a block, is not released and its entity has the tree's motion, else 0.
__global__ void __launch_bounds__(256) GatherCountKernel(GatherArgs a)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.groups || a.state->building == 0)
    {
        return;
    }
    const LightTables t = TablesOf(a.descriptor);
    const GroupLightRow& row = t.rows[g];
    uint32_t count = 0;
    if (row.block_units != 0 && (row.cpu_flags & kLtCpuReleased) == 0 &&
        (t.entity_motion[t.group_infos[g].entity] & 1u) == a.tree)
    {
        count = CurrentHeap(t)[uint64_t(row.block_offset) * 4];
    }
    a.offsets[g] = count;
}

// One workgroup: the exclusive scan of the counts (leaf offsets), the total = leaves_found.
__global__ void __launch_bounds__(1024) GatherScanKernel(GatherArgs a)
{
    __shared__ uint32_t lds[32];
    if (a.state->building == 0)
    {
        return;
    }
    uint32_t carry = 0;
    for (uint32_t base = 0; base < a.groups; base += 1024)
    {
        const uint32_t g = base + threadIdx.x;
        const uint32_t count = g < a.groups ? a.offsets[g] : 0u;
        uint32_t total;
        const uint32_t before = BlockExclusiveSum<1024>(count, lds, total);
        if (g < a.groups)
        {
            a.offsets[g] = carry + before;
        }
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        a.state->leaves_found = carry;
    }
}

// LtWorldSummary (caelix_light_gather_emit.hlsl:11-34): the AABB of the rotated box (R c + t, |R| h)
// and the projected face sums W[+a] = sum_f F[f] max(0, (R d_f)_a), W[-a] = sum_f F[f] max(0, -(R d_f)_a).
CX_DEV LightSummary WorldSummary(const float face[6], const float lo[3], const float hi[3], const float rows[3][4])
{
    LightSummary w;
    float centre[3], half_size[3];
    for (int k = 0; k < 3; ++k)
    {
        centre[k] = 0.5f * (lo[k] + hi[k]);
        half_size[k] = 0.5f * (hi[k] - lo[k]);
    }
    for (int r = 0; r < 3; ++r)
    {
        const float wc = ((rows[r][0] * centre[0] + rows[r][1] * centre[1]) + rows[r][2] * centre[2]) + rows[r][3];
        const float wh = (fabsf(rows[r][0]) * half_size[0] + fabsf(rows[r][1]) * half_size[1]) +
                         fabsf(rows[r][2]) * half_size[2];
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
            w.face[2 * b] += face[f] * fmaxf(0.0f, d);
            w.face[2 * b + 1] += face[f] * fmaxf(0.0f, -d);
        }
    }
    return w;
}

// One wave per group, lanes over its index entries (caelix_light_gather_emit.hlsl:36-86): leaf = the
// group's offset + entry, its summary the record's box (+ the group origin) and face sums moved to world
// space; the index entry's bvh_leaf = tree << 31 | leaf (counted from the entity's leaf base when the
// gather numbers leaves per entity), or kLtNone past the leaf capacity.
__global__ void __launch_bounds__(64) GatherEmitKernel(GatherArgs a)
{
    const uint32_t g = blockIdx.x;
    if (g >= a.groups || a.state->building == 0)
    {
        return;
    }
    const LightTables t = TablesOf(a.descriptor);
    const uint32_t found = a.state->leaves_found;
    const uint32_t first = a.offsets[g];
    const uint32_t next = g + 1 < a.groups ? a.offsets[g + 1] : found;
    if (next == first)
    {
        return;
    }
    const GroupLightRow& row = t.rows[g];
    uint32_t* block = const_cast<uint32_t*>(CurrentHeap(t)) + uint64_t(row.block_offset) * 4;
    const GroupInfo& info = t.group_infos[g];
    const int32_t edge = int32_t(8u << t.group_shift);
    const float origin[3] = {float(info.coord[0] * edge), float(info.coord[1] * edge), float(info.coord[2] * edge)};
    float rows[3][4];
    for (int r = 0; r < 3; ++r)
    {
        for (int c = 0; c < 4; ++c)
        {
            rows[r][c] = t.entity_infos[info.entity].cur_rows[r][c];
        }
    }
    const uint32_t base = a.entity_leaf_base != nullptr ? a.entity_leaf_base[info.entity] : 0u;
    for (uint32_t k = threadIdx.x; k < next - first; k += 64)
    {
        const uint32_t leaf = first + k;
        uint32_t* index_entry = block + 4 + uint64_t(k) * 4;
        if (leaf >= a.leaf_capacity)
        {
            index_entry[3] = kLtNone;
            continue;
        }
        const uint32_t* rec = block + uint64_t(index_entry[2]) * 4;
        float face[6], lo[3], hi[3];
        for (int f = 0; f < 6; ++f)
        {
            face[f] = __uint_as_float(rec[f]);
        }
        for (int c = 0; c < 3; ++c)
        {
            lo[c] = origin[c] + float((rec[6] >> (kLtVoxelBits * c)) & kLtVoxelMask);
            hi[c] = origin[c] + float((rec[7] >> (kLtVoxelBits * c)) & kLtVoxelMask);
        }
        LightBvhLeaf out;
        out.summary = WorldSummary(face, lo, hi, rows);
        out.group = g;
        out.entry = k;
        out.pad = 0;
        out.parent = kLtNone;
        a.leaves[leaf] = out;
        index_entry[3] = (a.tree << kLtBvhLeafTreeShift) | (leaf - base);
    }
}

// Exclusive scan of two u32 arrays at once (units, flags), in place, with the totals: reduce per
// 1024-element tile, scan of the tile sums in one workgroup, add.
constexpr uint32_t kScanTileElems = 1024;

__global__ void __launch_bounds__(256) TileSumKernel(const uint32_t* a, const uint32_t* b, uint32_t n, uint2* tile_sums)
{
    __shared__ uint32_t lds[8];
    const uint32_t base = blockIdx.x * kScanTileElems;
    uint32_t sa = 0, sb = 0;
    for (uint32_t k = threadIdx.x; k < kScanTileElems; k += 256)
    {
        if (base + k < n)
        {
            sa += a[base + k];
            sb += b[base + k];
        }
    }
    sa = WaveSum(sa);
    sb = WaveSum(sb);
    const uint32_t wave = threadIdx.x / 64;
    if (Lane() == 0)
    {
        lds[wave] = sa;
        lds[4 + wave] = sb;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        tile_sums[blockIdx.x] = make_uint2(lds[0] + lds[1] + lds[2] + lds[3], lds[4] + lds[5] + lds[6] + lds[7]);
    }
}

__global__ void __launch_bounds__(1024) TileScanKernel(uint2* tile_sums, uint32_t tiles, uint32_t* totals,
                                                       LightTreeState* state)
{
    __shared__ uint32_t lds[32];
    uint32_t carry_a = 0, carry_b = 0;
    for (uint32_t base = 0; base < tiles; base += 1024)
    {
        const uint32_t i = base + threadIdx.x;
        const uint2 v = i < tiles ? tile_sums[i] : make_uint2(0, 0);
        uint32_t ta, tb;
        const uint32_t ea = BlockExclusiveSum<1024>(v.x, lds, ta);
        const uint32_t eb = BlockExclusiveSum<1024>(v.y, lds + 16, tb);
        if (i < tiles)
        {
            tile_sums[i] = make_uint2(carry_a + ea, carry_b + eb);
        }
        carry_a += ta;
        carry_b += tb;
    }
    if (threadIdx.x == 0)
    {
        totals[1] = carry_a;  // heap units
        totals[0] = carry_b;  // light bricks
        if (state != nullptr)
        {
            state->leaves_found = carry_b;  // as the gather's scan writes it
        }
    }
}

__global__ void __launch_bounds__(256) TileApplyKernel(uint32_t* a, uint32_t* b, uint32_t n, const uint2* tile_sums)
{
    __shared__ uint32_t lds[8];
    const uint32_t base = blockIdx.x * kScanTileElems;
    const uint32_t t = threadIdx.x;
    uint32_t va[4], vb[4];
    uint32_t sa = 0, sb = 0;
    for (int k = 0; k < 4; ++k)
    {
        const uint32_t i = base + t * 4 + k;
        va[k] = i < n ? a[i] : 0u;
        vb[k] = i < n ? b[i] : 0u;
        sa += va[k];
        sb += vb[k];
    }
    uint32_t ta, tb;
    uint32_t ea = BlockExclusiveSum<256>(sa, lds, ta);
    uint32_t eb = BlockExclusiveSum<256>(sb, lds + 4, tb);
    const uint2 carry = tile_sums[blockIdx.x];
    ea += carry.x;
    eb += carry.y;
    for (int k = 0; k < 4; ++k)
    {
        const uint32_t i = base + t * 4 + k;
        if (i < n)
        {
            a[i] = ea;
            b[i] = eb;
        }
        ea += va[k];
        eb += vb[k];
    }
}

SceneArgs ArgsOf(const CxScene& c)
{
    SceneArgs s;
    s.keys = c.keys;
    s.info = reinterpret_cast<uint2*>(c.info);
    s.occupancy = reinterpret_cast<uint64_t*>(c.occupancy);
    s.payload = c.payload;
    s.payload_out = c.payload;
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
    s.bricks = c.bricks;
    s.group_shift = c.group_shift;
    s.encoding = c.encoding;
    s.materials = reinterpret_cast<const float4*>(c.materials);
    s.light_codes = c.light_codes;
    s.key_words = c.scene_count > 1 ? 4u : 3u;
    s.box = c.occupied_box;
    s.light_rows = c.light_group_rows;
    s.group_of_brick = c.group_of_brick;
    s.list = nullptr;
    return s;
}

}  // namespace
}  // namespace cx

extern "C" {

CX_API int cx_scene_finalize_bricks(const CxScene* scene, const uint32_t* records, uint32_t count, void* stream)
{
    if (scene == nullptr || scene->keys == nullptr || scene->info == nullptr || scene->occupancy == nullptr ||
        scene->payload == nullptr || scene->brick_bits == nullptr || scene->super_bits == nullptr ||
        scene->slots == nullptr || scene->hash_info == nullptr || (records == nullptr && count > 0))
    {
        return cx::ArgError("cx_scene_finalize_bricks: a buffer is NULL");
    }
    if ((reinterpret_cast<uintptr_t>(scene->payload) & 15u) != 0 || scene->grid_dim == 0 || scene->grid_dim % 8 != 0 ||
        scene->group_shift == 0 || scene->group_shift > 7)
    {
        return cx::ArgError("cx_scene_finalize_bricks: payload alignment, grid_dim or group_shift invalid");
    }
    if (scene->light_group_rows != nullptr &&
        (scene->light_codes == nullptr || scene->group_of_brick == nullptr || scene->materials == nullptr))
    {
        return cx::ArgError("cx_scene_finalize_bricks: the light state needs light_codes, materials and group_of_brick");
    }
    if (count == 0)
    {
        return CX_OK;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::SceneArgs a = cx::ArgsOf(*scene);
    a.list = records;
    cx::StageScope scope(s, "scene_finalize");
    cx::FinalizeKernel<<<count, 64, 0, s>>>(a);
    if (scene->light_group_rows != nullptr)
    {
        cx::LightFacesKernel<<<count, 64, 0, s>>>(a, scene->light_group_rows, scene->group_of_brick);
    }
    return cx::LaunchResult("cx_scene_finalize_bricks");
}

CX_API int cx_scene_finalize(const CxScene* scene, void* stream)
{
    if (scene == nullptr || scene->keys == nullptr || scene->info == nullptr || scene->occupancy == nullptr ||
        scene->payload == nullptr || scene->brick_bits == nullptr || scene->super_bits == nullptr ||
        scene->slots == nullptr || scene->hash_info == nullptr)
    {
        return cx::ArgError("cx_scene_finalize: a buffer is NULL");
    }
    if ((reinterpret_cast<uintptr_t>(scene->payload) & 15u) != 0)
    {
        return cx::ArgError("cx_scene_finalize: payload not 16-byte aligned");
    }
    if (scene->grid_dim == 0 || scene->grid_dim % 8 != 0 || scene->group_shift == 0 || scene->group_shift > 7)
    {
        return cx::ArgError("cx_scene_finalize: grid_dim must be a positive multiple of 8, group_shift in 1..7");
    }
    if (scene->light_group_rows != nullptr &&
        (scene->light_codes == nullptr || scene->group_of_brick == nullptr || scene->materials == nullptr))
    {
        return cx::ArgError("cx_scene_finalize: the light state needs light_codes, materials and group_of_brick");
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const uint64_t scenes = scene->scene_count > 1 ? scene->scene_count : 1;
    const uint64_t cells = uint64_t(scene->grid_dim) * scene->grid_dim * scene->grid_dim * scenes;
    const uint64_t supers = cells / 512;
    cx::StageScope scope(s, "scene_finalize");
    cx::FillKernel<<<2048, 256, 0, s>>>(scene->brick_bits, (cells + 31) / 32);
    cx::FillKernel<<<256, 256, 0, s>>>(scene->super_bits, (supers + 31) / 32);
    if (scene->occupied_box != nullptr)
    {
        cx::BoxInitKernel<<<uint32_t(cx::DivUp(scenes * 6, uint64_t(256))), 256, 0, s>>>(scene->occupied_box,
                                                                                        uint32_t(scenes));
    }
    if (scene->bricks > 0)
    {
        cx::FinalizeKernel<<<scene->bricks, 64, 0, s>>>(cx::ArgsOf(*scene));
        if (scene->light_group_rows != nullptr)
        {
            cx::LightFacesKernel<<<scene->bricks, 64, 0, s>>>(cx::ArgsOf(*scene), scene->light_group_rows,
                                                              scene->group_of_brick);
        }
    }
    return cx::LaunchResult("cx_scene_finalize");
}

}  // extern "C"

namespace cx {
namespace {

// The light pass's scratch: bookkeeping words, per group (mode, first, count, sizes, faces: first, so
// that their place does not depend on the brick count and the gather finds them), per brick (sort keys
// and values twice), per sorted position (units and recompute flags with one more for the totals, plans),
// the scan's tile sums and the sort's state.
struct PassScratch
{
    uint32_t* words;  // [0] brick count (the sort reads it), [4..8) totals
    uint32_t* mode;
    uint32_t* first;
    uint32_t* count;
    uint32_t* sizes;
    uint32_t* faces;
    uint32_t* keys0;
    uint32_t* values0;
    uint32_t* keys1;
    uint32_t* values1;
    uint32_t* units;
    uint32_t* recompute;
    uint32_t* plans;
    uint2* tile_sums;
    void* sort_state;
    uint64_t bytes;
};

uint64_t Align256(uint64_t v)
{
    return (v + 255) & ~uint64_t(255);
}

PassScratch PassScratchOf(void* base, uint32_t bricks, uint32_t groups)
{
    const uint64_t n = std::max(bricks, 1u);
    const uint64_t g = std::max(groups, 1u);
    const uint64_t tiles = (n + kScanTileElems - 1) / kScanTileElems;
    uint8_t* p = static_cast<uint8_t*>(base);
    uint64_t at = 0;
    auto take = [&](uint64_t bytes) {
        uint8_t* q = p + at;
        at = Align256(at + bytes);
        return q;
    };
    PassScratch s;
    s.words = reinterpret_cast<uint32_t*>(take(64));
    s.mode = reinterpret_cast<uint32_t*>(take(4 * g));
    s.first = reinterpret_cast<uint32_t*>(take(4 * g));
    s.count = reinterpret_cast<uint32_t*>(take(4 * g));
    s.sizes = reinterpret_cast<uint32_t*>(take(4 * g));
    s.faces = reinterpret_cast<uint32_t*>(take(4 * g));
    s.keys0 = reinterpret_cast<uint32_t*>(take(4 * n));
    s.values0 = reinterpret_cast<uint32_t*>(take(4 * n));
    s.keys1 = reinterpret_cast<uint32_t*>(take(4 * n));
    s.values1 = reinterpret_cast<uint32_t*>(take(4 * n));
    s.units = reinterpret_cast<uint32_t*>(take(4 * (n + 1)));
    s.recompute = reinterpret_cast<uint32_t*>(take(4 * (n + 1)));
    s.plans = reinterpret_cast<uint32_t*>(take(4 * n));
    s.tile_sums = reinterpret_cast<uint2*>(take(8 * (tiles + 1)));
    s.sort_state = take(cx_radix_sort_state_bytes(uint32_t(n)));
    s.bytes = at;
    return s;
}

__global__ void SetWordKernel(uint32_t* word, uint32_t value)
{
    *word = value;
}

int Gather(const CxLightPass* pass, uint32_t tree, LightBvhLeaf* leaves, uint32_t leaf_capacity,
           LightTreeState* state, const uint32_t* entity_leaf_base, void* stream, const char* what)
{
    if (pass == nullptr || pass->descriptor == nullptr || pass->scratch == nullptr || leaves == nullptr ||
        state == nullptr || tree >= kLtTrees || pass->group_count == 0)
    {
        return ArgError("%s: a buffer is NULL, tree >= 2 or no groups", what);
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // The gather's offsets reuse the pass's per-group `first` array (the pass is done with it).
    const PassScratch sc = PassScratchOf(pass->scratch, 1, pass->group_count);
    GatherArgs a;
    a.descriptor = pass->descriptor;
    a.offsets = sc.first;
    a.leaves = leaves;
    a.state = state;
    a.entity_leaf_base = entity_leaf_base;
    a.leaf_capacity = leaf_capacity;
    a.tree = tree;
    a.groups = pass->group_count;
    StageScope scope(s, "light_gather");
    GatherCountKernel<<<DivUp(a.groups, 256), 256, 0, s>>>(a);
    GatherScanKernel<<<1, 1024, 0, s>>>(a);
    GatherEmitKernel<<<a.groups, 64, 0, s>>>(a);
    return LaunchResult(what);
}

}  // namespace
}  // namespace cx

extern "C" {

CX_API uint64_t cx_light_pass_scratch_bytes(uint32_t bricks, uint32_t groups)
{
    return cx::PassScratchOf(nullptr, bricks, groups).bytes;
}

CX_API int cx_light_pass(const CxScene* scene, const CxLightPass* pass, void* stream)
{
    if (scene == nullptr || pass == nullptr || scene->slots == nullptr || scene->hash_info == nullptr ||
        scene->materials == nullptr || scene->light_codes == nullptr || scene->group_of_brick == nullptr ||
        scene->info == nullptr || pass->descriptor == nullptr || pass->staged_flags == nullptr ||
        pass->scratch == nullptr)
    {
        return cx::ArgError("cx_light_pass: a buffer is NULL");
    }
    if (pass->group_count == 0 || (uint64_t(pass->group_count) << (3 * scene->group_shift)) > 0xFFFFFFFFull)
    {
        return cx::ArgError("cx_light_pass: %u groups: not 1 .. 2^32 / 8^%u", pass->group_count, scene->group_shift);
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const cx::SceneArgs a = cx::ArgsOf(*scene);
    const uint32_t n = scene->bricks;
    const cx::PassScratch sc = cx::PassScratchOf(pass->scratch, n, pass->group_count);
    cx::PassArgs p;
    p.descriptor = pass->descriptor;
    p.staged = pass->staged_flags;
    p.group_of_brick = scene->group_of_brick;
    p.keys = sc.keys0;
    p.values = sc.values0;
    p.units = sc.units;
    p.recompute = sc.recompute;
    p.plans = sc.plans;
    p.mode = sc.mode;
    p.first = sc.first;
    p.count = sc.count;
    p.sizes = sc.sizes;
    p.faces = sc.faces;
    p.totals = sc.words + 4;
    p.heap_units = pass->heap_units;
    p.bricks = n;
    p.groups = pass->group_count;
    const uint32_t g = pass->group_count;
    const uint32_t tiles = cx::DivUp(std::max(n, 1u), cx::kScanTileElems);
    cx::StageScope scope(s, "light_pass");
    cx::FillKernel<<<1, 64, 0, s>>>(sc.words, 16);
    cx::PassFlagsKernel<<<cx::DivUp(g, 256), 256, 0, s>>>(p);
    if (n > 0)
    {
        cx::LightKeyKernel<<<cx::DivUp(n, 256), 256, 0, s>>>(a, p);
        cx::SetWordKernel<<<1, 1, 0, s>>>(sc.words, n);
        if (int r = cx_radix_sort_pairs(sc.keys0, sc.values0, sc.keys1, sc.values1, sc.words, nullptr, n,
                                        sc.sort_state, stream);
            r != CX_OK)
        {
            return r;
        }
        cx::GroupSpanKernel<<<cx::DivUp(n, 256), 256, 0, s>>>(a, p);
        cx::BrickPlanKernel<<<cx::DivUp(n + 1, 256), 256, 0, s>>>(a, p);
        cx::TileSumKernel<<<tiles, 256, 0, s>>>(sc.units, sc.recompute, n, sc.tile_sums);
        cx::TileScanKernel<<<1, 1024, 0, s>>>(sc.tile_sums, tiles, p.totals, nullptr);
        cx::TileApplyKernel<<<tiles, 256, 0, s>>>(sc.units, sc.recompute, n, sc.tile_sums);
    }
    cx::GroupSizeKernel<<<cx::DivUp(g, 256), 256, 0, s>>>(p);
    cx::GroupScanKernel<<<1, 1024, 0, s>>>(p);
    cx::BlockCopyKernel<<<g, 256, 0, s>>>(p);
    if (n > 0)
    {
        cx::BlockBuildKernel<<<n, 64, 0, s>>>(a, p);
        cx::StaleClearKernel<<<cx::DivUp(n, 256), 256, 0, s>>>(a, p);
    }
    cx::RowCommitKernel<<<cx::DivUp(g, 256), 256, 0, s>>>(p);
    cx::PassCommitKernel<<<1, 256, 0, s>>>(p);
    return cx::LaunchResult("cx_light_pass");
}

CX_API int cx_light_gather(const CxLightPass* pass, uint32_t tree, LightBvhLeaf* leaves, uint32_t leaf_capacity,
                           LightTreeState* state, void* stream)
{
    return cx::Gather(pass, tree, leaves, leaf_capacity, state, nullptr, stream, "cx_light_gather");
}

CX_API int cx_light_gather_entities(const CxLightPass* pass, uint32_t tree, LightBvhLeaf* leaves,
                                    uint32_t leaf_capacity, LightTreeState* state, const uint32_t* entity_leaf_base,
                                    void* stream)
{
    if (entity_leaf_base == nullptr)
    {
        return cx::ArgError("cx_light_gather_entities: entity_leaf_base is NULL");
    }
    return cx::Gather(pass, tree, leaves, leaf_capacity, state, entity_leaf_base, stream, "cx_light_gather_entities");
}

}  // extern "C"
