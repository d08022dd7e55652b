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
#include "runtime.h"

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
    const uint32_t brick = blockIdx.x;
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
        s.info[brick] = make_uint2((pos & kRecordPosMask) | (coarse << kRecordCoarseShift), bounds);
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

struct LightArgs
{
    uint32_t* heap;
    uint32_t heap_units;
    uint32_t* units;          // per brick: record units (count pass), then exclusive offsets
    uint32_t* flags;          // per brick: 1 for a light brick (count pass), then exclusive ordinals
    uint32_t* record_offset;  // per ordinal
    uint32_t* ordinal_of_brick;
    LightBvhLeaf* leaves;
    uint32_t leaf_capacity;
    uint32_t* totals;  // [0] light bricks, [1] heap units, [2] light faces
};

__global__ void __launch_bounds__(64) LightCountKernel(SceneArgs s, LightArgs l)
{
    const uint32_t record = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    uint32_t faces = 0;
    for (uint32_t i = 0; i < 8; ++i)
    {
        uint32_t block;
        faces += __popc(OpenLightFaces(s, record, lane * 8 + i, block));
    }
    faces = WaveSum(faces);
    if (lane == 0)
    {
        const uint32_t count = min(faces, kLtRecordFaceCountMask);
        uint2 w = s.info[record];
        w.y = (w.y & kRecordBoundsMask) | (count << kLtRecordFaceCountShift);
        s.info[record] = w;
        l.units[record] = count > 0 ? (kLtUnitBytes * 3 + count * 4 + kLtUnitBytes - 1) / kLtUnitBytes : 0u;
        l.flags[record] = count > 0 ? 1u : 0u;
    }
}

__global__ void __launch_bounds__(64) LightRecordKernel(SceneArgs s, LightArgs l)
{
    const uint32_t record = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const uint32_t count = (s.info[record].y >> kLtRecordFaceCountShift) & kLtRecordFaceCountMask;
    if (count == 0)
    {
        if (lane == 0)
        {
            l.ordinal_of_brick[record] = kLtNone;
        }
        return;
    }
    __shared__ float lds_sums[6][64];
    const uint32_t ordinal = l.flags[record];
    const uint32_t offset = l.units[record];
    const uint32_t units = (kLtUnitBytes * 3 + count * 4 + kLtUnitBytes - 1) / kLtUnitBytes;
    if (uint64_t(offset) + units > l.heap_units)
    {
        if (lane == 0)
        {
            l.ordinal_of_brick[record] = kLtNone;
            atomicMax(&l.totals[3], 1u);  // the heap is too small: the caller grows it and runs again
        }
        return;
    }
    uint32_t* rec = l.heap + uint64_t(offset) * 4;
    uint32_t counts[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t i = 0; i < 8; ++i)
    {
        uint32_t block;
        const uint32_t open = OpenLightFaces(s, record, lane * 8 + i, block);
        for (uint32_t f = 0; f < 6; ++f)
        {
            counts[f] += (open >> f) & 1u;
        }
    }
    uint32_t family_base[6];
    uint32_t family_count[6];
    uint32_t cursor[6];
    uint32_t written = 0;
    for (uint32_t f = 0; f < 6; ++f)
    {
        const uint32_t inclusive = WaveInclusiveSum(counts[f]);
        family_count[f] = __shfl(inclusive, 63, 64);
        cursor[f] = inclusive - counts[f];
        family_base[f] = written;
        written += family_count[f];
    }
    float sums[6] = {0, 0, 0, 0, 0, 0};
    uint32_t lo[3] = {8, 8, 8};
    uint32_t hi[3] = {0, 0, 0};
    for (uint32_t i = 0; i < 8; ++i)
    {
        const uint32_t vi = lane * 8 + i;
        uint32_t block;
        const uint32_t open = OpenLightFaces(s, record, vi, block);
        if (open == 0)
        {
            continue;
        }
        const float4 e = s.materials[2 * uint64_t(MaterialCode(block, s.encoding)) + 1];
        const float weight = e.x * 0.2126f + e.y * 0.7152f + e.z * 0.0722f;
        for (uint32_t f = 0; f < 6; ++f)
        {
            if (((open >> f) & 1u) != 0)
            {
                rec[12 + family_base[f] + cursor[f]] = vi | (block << kLtFaceEntryBlockShift);
                ++cursor[f];
                sums[f] += weight;
            }
        }
        const uint32_t x = vi & 7u, y = (vi >> 3) & 7u, z = vi >> 6;
        lo[0] = min(lo[0], x);
        lo[1] = min(lo[1], y);
        lo[2] = min(lo[2], z);
        hi[0] = max(hi[0], x + 1);
        hi[1] = max(hi[1], y + 1);
        hi[2] = max(hi[2], z + 1);
    }
    for (uint32_t f = 0; f < 6; ++f)
    {
        lds_sums[f][lane] = sums[f];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = min(lo[k], uint32_t(__shfl_xor(int(lo[k]), o, 64)));
            hi[k] = max(hi[k], uint32_t(__shfl_xor(int(hi[k]), o, 64)));
        }
    }
    __syncthreads();
    if (lane == 0)
    {
        float face[6];
        for (uint32_t f = 0; f < 6; ++f)
        {
            float sum = 0.0f;
            for (uint32_t k = 0; k < 64; ++k)
            {
                sum += lds_sums[f][k];
            }
            face[f] = sum;
        }
        const int32_t* key = BrickXyz(s, record);
        const int32_t g = int32_t(s.group_shift);
        const uint32_t mask = (1u << s.group_shift) - 1u;
        const uint32_t bx = uint32_t(key[0]) & mask, by = uint32_t(key[1]) & mask, bz = uint32_t(key[2]) & mask;
        const uint32_t box_lo = (bx * 8 + lo[0]) | ((by * 8 + lo[1]) << kLtVoxelBits) | ((bz * 8 + lo[2]) << (2 * kLtVoxelBits));
        const uint32_t box_hi = (bx * 8 + hi[0]) | ((by * 8 + hi[1]) << kLtVoxelBits) | ((bz * 8 + hi[2]) << (2 * kLtVoxelBits));
        for (uint32_t f = 0; f < 6; ++f)
        {
            rec[f] = __float_as_uint(face[f]);
        }
        rec[6] = box_lo;
        rec[7] = box_hi;
        rec[8] = bx | (by << s.group_shift) | (bz << (2 * s.group_shift));
        rec[9] = written;
        rec[10] = family_count[0] | (family_count[1] << kLtFamilyCountBits) | (family_count[2] << (2 * kLtFamilyCountBits));
        rec[11] = family_count[3] | (family_count[4] << kLtFamilyCountBits) | (family_count[5] << (2 * kLtFamilyCountBits));
        const uint32_t end = ((12 + written + 3) / 4) * 4;
        for (uint32_t w = 12 + written; w < end; ++w)
        {
            rec[w] = 0u;
        }
        l.record_offset[ordinal] = offset;
        l.ordinal_of_brick[record] = ordinal;
        atomicAdd(&l.totals[2], written);
        if (ordinal < l.leaf_capacity)
        {
            // The leaf: the record box moved by the group origin, the face sums (identity frame).
            const float ox = float((key[0] >> g) << g) * 8.0f;
            const float oy = float((key[1] >> g) << g) * 8.0f;
            const float oz = float((key[2] >> g) << g) * 8.0f;
            LightBvhLeaf leaf;
            for (int f = 0; f < 6; ++f)
            {
                leaf.summary.face[f] = face[f];
            }
            leaf.summary.lo[0] = ox + float(bx * 8 + lo[0]);
            leaf.summary.lo[1] = oy + float(by * 8 + lo[1]);
            leaf.summary.lo[2] = oz + float(bz * 8 + lo[2]);
            leaf.summary.hi[0] = ox + float(bx * 8 + hi[0]);
            leaf.summary.hi[1] = oy + float(by * 8 + hi[1]);
            leaf.summary.hi[2] = oz + float(bz * 8 + hi[2]);
            leaf.group = record;
            leaf.entry = ordinal;
            leaf.pad = 0;
            leaf.parent = kLtNone;
            l.leaves[ordinal] = leaf;
        }
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
    return s;
}

}  // namespace
}  // namespace cx

extern "C" {

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
    }
    return cx::LaunchResult("cx_scene_finalize");
}

CX_API uint64_t cx_scene_light_scratch_bytes(uint32_t bricks)
{
    const uint64_t tiles = (uint64_t(bricks) + cx::kScanTileElems - 1) / cx::kScanTileElems;
    return 2 * 4 * uint64_t(bricks + 1) + 8 * (tiles + 1) + 256;
}

CX_API int cx_scene_light_records(const CxScene* scene, const CxLightScene* light, void* stream)
{
    if (scene == nullptr || light == nullptr || scene->slots == nullptr || scene->hash_info == nullptr ||
        scene->materials == nullptr || scene->light_codes == nullptr || light->heap == nullptr ||
        light->record_offset == nullptr || light->ordinal_of_brick == nullptr || light->leaves == nullptr ||
        light->totals == nullptr || light->scratch == nullptr)
    {
        return cx::ArgError("cx_scene_light_records: a buffer is NULL");
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const cx::SceneArgs a = cx::ArgsOf(*scene);
    const uint32_t n = scene->bricks;
    uint32_t* scratch = static_cast<uint32_t*>(light->scratch);
    cx::LightArgs l;
    l.heap = light->heap;
    l.heap_units = light->heap_units;
    l.units = scratch;
    l.flags = scratch + (n + 1);
    l.record_offset = light->record_offset;
    l.ordinal_of_brick = light->ordinal_of_brick;
    l.leaves = light->leaves;
    l.leaf_capacity = light->leaf_capacity;
    l.totals = light->totals;
    uint2* tile_sums = reinterpret_cast<uint2*>(scratch + 2 * (uint64_t(n) + 1));
    const uint32_t tiles = cx::DivUp(std::max(n, 1u), cx::kScanTileElems);
    cx::StageScope scope(s, "scene_light_records");
    cx::FillKernel<<<1, 64, 0, s>>>(light->totals, 4);
    if (n == 0)
    {
        return cx::LaunchResult("cx_scene_light_records");
    }
    cx::LightCountKernel<<<n, 64, 0, s>>>(a, l);
    cx::TileSumKernel<<<tiles, 256, 0, s>>>(l.units, l.flags, n, tile_sums);
    cx::TileScanKernel<<<1, 1024, 0, s>>>(tile_sums, tiles, light->totals, light->state);
    cx::TileApplyKernel<<<tiles, 256, 0, s>>>(l.units, l.flags, n, tile_sums);
    cx::LightRecordKernel<<<n, 64, 0, s>>>(a, l);
    return cx::LaunchResult("cx_scene_light_records");
}

}  // extern "C"
