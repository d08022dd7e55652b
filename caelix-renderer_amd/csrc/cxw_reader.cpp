// Reader for Caelix .cxw world files (format version 8 and older): the host-side producer of the
// bricks the scene build consumes (host-rs/caelix-cxw/src/lib.rs:1-656 is the reference reader).
//
// The file is memory-mapped. cx_cxw_open parses the header, the entity table and every sector
// index and checks every region (lib.rs:305-452); cx_cxw_decode inflates one sector's raw-deflate
// payload and returns its non-air bricks (lib.rs:454-656); cx_cxw_decode_all decodes every sector on
// a pool of threads into per-sector buffers that cx_cxw_copy_decoded concatenates in sector order
// (the reference streams them to a sink in completion order; the order here is fixed so the copy is
// deterministic). Brick values are returned as stored.
//
// Layout (little-endian):
//   header (64 B)       magic u32 "CXLS" or legacy "VXLS", version u16 (<= 8), flags u16 (== 1,
//                       deflate), entity table offset u64
//   entity table        [v >= 7: world record u8 flags, f32 x3 position, f32 x4 rotation]
//                       entity count u32, then per entity: guid 16 B, position f32 x3, rotation f32
//                       x4, require-update flags u16, [v >= 3: flags u8 (v < 6: legacy body state)],
//                       [v >= 4: velocity f32 x6], [v >= 5: protected u8], sector index offset u64,
//                       sector count u32
//   sector index        per sector: coord i32 x3, region offset u64, region size u32
//   region              preview u32 x 4096, payload size u32, raw-deflate payload; region size ==
//                       16388 + payload size
//   payload (inflated)  magic u32 "VXS2", capacity, count, free count (u32), brick index table
//                       i16 x 4096 (-1 = none), free list i16 x free, require-update flags
//                       2 + 8192 B, slot count u32 (<= 16), per slot: id u8, stride u16, data
//                       capacity * 512 * stride B; slot 0 = Block (stride 2, u16 values)
#include "caelix_amd/caelix_amd.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

namespace {

constexpr uint32_t kFileMagic = 0x534C5843u;        // "CXLS"
constexpr uint32_t kLegacyFileMagic = 0x534C5856u;  // "VXLS"
constexpr uint16_t kCurrentVersion = 8;
constexpr uint16_t kFlagDeflate = 1;
constexpr uint32_t kPayloadMagic = 0x32535856u;  // "VXS2"
constexpr uint8_t kEntityHasBody = 1;
constexpr uint8_t kEntityStatic = 2;
constexpr uint8_t kWorldHasPlayerPose = 1;
constexpr int32_t kBricksPerAxis = 16;
constexpr uint32_t kBricksInSector = 4096;
constexpr uint32_t kBlocksInBrick = 512;
constexpr uint64_t kHeaderBytes = 64;
constexpr uint64_t kPreviewBytes = uint64_t(kBricksInSector) * 4;
constexpr uint64_t kRegionOverhead = kPreviewBytes + 4;
constexpr uint64_t kSectorEntryBytes = 24;
constexpr uint32_t kMaxSlots = 16;
constexpr uint8_t kSlotBlock = 0;
constexpr uint16_t kBlockStride = 2;
constexpr size_t kMaxPayloadBytes = size_t(1) << 30;
constexpr uint16_t kVersionEntityFlags = 3;
constexpr uint16_t kVersionVelocity = 4;
constexpr uint16_t kVersionProtected = 5;
constexpr uint16_t kVersionFlagsV6 = 6;
constexpr uint16_t kVersionWorldRecord = 7;

// An error of the reference's Error enum (lib.rs:69-109), with its fields.
struct Failure
{
    CxCxwError e{};
};

Failure Fail(int32_t kind, const char* fmt, ...)
{
    Failure f;
    f.e.kind = kind;
    va_list args;
    va_start(args, fmt);
    vsnprintf(f.e.message, sizeof(f.e.message), fmt, args);
    va_end(args);
    return f;
}

Failure Truncated(const char* what, uint64_t offset)
{
    Failure f = Fail(CX_CXW_TRUNCATED, "truncated data reading %s at offset %llu", what,
                     static_cast<unsigned long long>(offset));
    snprintf(f.e.what, sizeof(f.e.what), "%s", what);
    f.e.offset = offset;
    return f;
}

Failure BadPayload(const int32_t coord[3], const char* reason)
{
    Failure f = Fail(CX_CXW_BAD_PAYLOAD, "sector [%d, %d, %d]: invalid payload: %s", coord[0], coord[1], coord[2],
                     reason);
    std::memcpy(f.e.coord, coord, sizeof(f.e.coord));
    snprintf(f.e.what, sizeof(f.e.what), "%s", reason);
    return f;
}

// Bounds-checked little-endian reader (lib.rs:267-303).
struct Cursor
{
    const uint8_t* data;
    size_t size;
    size_t pos;

    const uint8_t* Take(size_t n, const char* what)
    {
        if (n > size || pos > size - n)
        {
            throw Truncated(what, pos);
        }
        const uint8_t* p = data + pos;
        pos += n;
        return p;
    }
    template <class T>
    T Get(const char* what)
    {
        T v;
        std::memcpy(&v, Take(sizeof(T), what), sizeof(T));
        return v;
    }
    template <size_t N>
    void Floats(float (&out)[N], const char* what)
    {
        std::memcpy(out, Take(N * 4, what), N * 4);
    }
};

uint8_t FlagsFromLegacyBodyState(uint8_t body_state)  // lib.rs:305-312
{
    switch (body_state)
    {
    case 1:
        return kEntityHasBody | kEntityStatic;
    case 2:
        return kEntityHasBody;
    default:
        return kEntityStatic;
    }
}

struct Region
{
    const uint8_t* preview;
    const uint8_t* payload;
    uint32_t payload_size;
};

// Checks one region and returns its preview and compressed payload (lib.rs:415-452).
Region CheckRegion(const uint8_t* data, uint64_t len, const int32_t coord[3], uint64_t offset, uint32_t size)
{
    if (offset > len || len - offset < kRegionOverhead)
    {
        throw Truncated("sector region", offset);
    }
    uint32_t payload_size;
    std::memcpy(&payload_size, data + offset + kPreviewBytes, 4);
    const uint64_t expected = kRegionOverhead + payload_size;
    if (expected != size)
    {
        Failure f = Fail(CX_CXW_BAD_REGION_SIZE, "sector [%d, %d, %d]: region size is %u, expected %llu (16388 + payload size)",
                         coord[0], coord[1], coord[2], size, static_cast<unsigned long long>(expected));
        std::memcpy(f.e.coord, coord, sizeof(f.e.coord));
        f.e.expected = expected;
        f.e.actual = size;
        throw f;
    }
    const uint64_t payload_start = offset + kRegionOverhead;
    if (payload_start + payload_size > len)
    {
        throw Truncated("sector payload", payload_start);
    }
    return Region{data + offset, data + payload_start, payload_size};
}

struct Decoded
{
    std::vector<int32_t> keys;  // 3 per brick
    std::vector<uint16_t> blocks;
    std::vector<uint32_t> summaries;
    uint64_t voxels = 0;
    uint32_t entity = 0;
};

// Inflates a raw deflate stream (no zlib header), growing the buffer (lib.rs:471-513).
void Inflate(z_stream& z, const uint8_t* input, uint32_t size, std::vector<uint8_t>& out, const int32_t coord[3])
{
    auto fail = [&](const char* message) {
        return Fail(CX_CXW_INFLATE, "inflate failed: sector [%d, %d, %d]: %s", coord[0], coord[1], coord[2], message);
    };
    if (inflateReset(&z) != Z_OK)
    {
        throw fail("inflateReset");
    }
    size_t guess = std::min<size_t>(std::max<size_t>(size_t(size) * 4, 64 * 1024), kMaxPayloadBytes);
    out.resize(guess);
    z.next_in = const_cast<Bytef*>(input);
    z.avail_in = size;
    size_t produced = 0;
    while (true)
    {
        if (produced == out.size())
        {
            if (out.size() >= kMaxPayloadBytes)
            {
                throw fail("payload inflates to more than 1 GiB");
            }
            out.resize(std::min(out.size() * 2, kMaxPayloadBytes));
        }
        z.next_out = out.data() + produced;
        z.avail_out = static_cast<uInt>(std::min<size_t>(out.size() - produced, 1u << 30));
        const uInt before_in = z.avail_in;
        const uInt before_out = z.avail_out;
        const int status = inflate(&z, Z_NO_FLUSH);
        produced += before_out - z.avail_out;
        if (status == Z_STREAM_END)
        {
            out.resize(produced);
            return;
        }
        if (status != Z_OK && status != Z_BUF_ERROR)
        {
            throw fail(z.msg != nullptr ? z.msg : "corrupt deflate stream");
        }
        const bool stalled = z.avail_in == before_in && z.avail_out == before_out;
        if (stalled && produced < out.size())
        {
            throw fail("deflate stream ends before its final block");
        }
    }
}

// The non-air bricks of one inflated payload (lib.rs:515-656).
void DecodePayload(const std::vector<uint8_t>& raw, const uint8_t* preview, const int32_t coord[3], Decoded& out)
{
    Cursor c{raw.data(), raw.size(), 0};
    const uint32_t magic = c.Get<uint32_t>("payload header");
    if (magic != kPayloadMagic)
    {
        Failure f = Fail(CX_CXW_BAD_SECTOR_MAGIC, "bad sector payload magic 0x%08X (expected VXS2)", magic);
        f.e.actual = magic;
        throw f;
    }
    const uint32_t capacity = c.Get<uint32_t>("payload header");
    const uint32_t count = c.Get<uint32_t>("payload header");
    const uint32_t free_count = c.Get<uint32_t>("payload header");
    if (capacity > kBricksInSector || count > capacity || free_count > capacity || count + free_count != capacity)
    {
        Failure f = Fail(CX_CXW_BAD_COUNTS, "invalid brick map counters: capacity %u, count %u, free %u", capacity,
                         count, free_count);
        f.e.expected = capacity;
        f.e.actual = count;
        f.e.offset = free_count;
        throw f;
    }
    const uint8_t* indices = c.Take(kBricksInSector * 2, "payload brick indices");
    c.Take(size_t(free_count) * 2, "payload free list");
    c.Take(2 + kBricksInSector * 2, "payload require-update flags");
    const uint32_t slot_count = c.Get<uint32_t>("payload slot record count");
    if (slot_count > kMaxSlots)
    {
        throw BadPayload(coord, "more than 16 slot records");
    }
    uint32_t seen = 0;
    const uint8_t* block_data = nullptr;
    for (uint32_t s = 0; s < slot_count; ++s)
    {
        const uint8_t slot_id = c.Get<uint8_t>("payload slot record header");
        const uint16_t stride = c.Get<uint16_t>("payload slot record header");
        if (slot_id >= kMaxSlots || (seen & (1u << slot_id)) != 0)
        {
            throw BadPayload(coord, "invalid or duplicate slot id");
        }
        seen |= 1u << slot_id;
        if (stride == 0)
        {
            throw BadPayload(coord, "slot stride is 0");
        }
        const uint8_t* data = c.Take(size_t(capacity) * kBlocksInBrick * stride, "payload slot data");
        if (slot_id == kSlotBlock)
        {
            if (stride != kBlockStride)
            {
                throw BadPayload(coord, "Block slot stride is not 2");
            }
            block_data = data;
        }
    }
    // Trailing bytes are tolerated, as by the reference.
    const int32_t base[3] = {int32_t(uint32_t(coord[0]) * uint32_t(kBricksPerAxis)),
                             int32_t(uint32_t(coord[1]) * uint32_t(kBricksPerAxis)),
                             int32_t(uint32_t(coord[2]) * uint32_t(kBricksPerAxis))};
    out.keys.reserve(size_t(count) * 3);
    out.blocks.reserve(size_t(count) * kBlocksInBrick);
    out.summaries.reserve(count);
    for (uint32_t bi = 0; bi < kBricksInSector; ++bi)
    {
        int16_t id;
        std::memcpy(&id, indices + bi * 2, 2);
        if (id == -1)
        {
            continue;
        }
        if (id < -1 || uint32_t(id) >= capacity)
        {
            throw BadPayload(coord, "brick index table entry outside -1..capacity-1");
        }
        if (block_data == nullptr)
        {
            throw BadPayload(coord, "allocated bricks but no Block slot");
        }
        const uint8_t* src = block_data + size_t(id) * kBlocksInBrick * kBlockStride;
        const size_t start = out.blocks.size();
        out.blocks.resize(start + kBlocksInBrick);
        std::memcpy(out.blocks.data() + start, src, kBlocksInBrick * 2);
        uint64_t non_zero = 0;
        for (uint32_t v = 0; v < kBlocksInBrick; ++v)
        {
            non_zero += out.blocks[start + v] != 0 ? 1u : 0u;
        }
        if (non_zero == 0)
        {
            out.blocks.resize(start);
            continue;
        }
        out.voxels += non_zero;
        out.keys.push_back(int32_t(uint32_t(base[0]) + (bi & 15u)));
        out.keys.push_back(int32_t(uint32_t(base[1]) + ((bi >> 4) & 15u)));
        out.keys.push_back(int32_t(uint32_t(base[2]) + (bi >> 8)));
        uint32_t word;
        std::memcpy(&word, preview + size_t(bi) * 4, 4);
        out.summaries.push_back(word);
    }
}

struct Inflater
{
    z_stream z{};
    bool ready = false;
    Inflater() { ready = inflateInit2(&z, -MAX_WBITS) == Z_OK; }
    ~Inflater()
    {
        if (ready)
        {
            inflateEnd(&z);
        }
    }
};

void Report(CxCxwError* error, const Failure& f)
{
    if (error != nullptr)
    {
        *error = f.e;
    }
}

}  // namespace

struct CxCxw
{
    int fd = -1;
    const uint8_t* data = nullptr;
    uint64_t size = 0;
    CxCxwWorld world{};
    std::vector<CxCxwEntity> entities;
    std::vector<CxCxwSector> sectors;
    std::vector<Decoded> decoded;  // per sector, after cx_cxw_decode_all

    ~CxCxw()
    {
        if (data != nullptr && size > 0)
        {
            munmap(const_cast<uint8_t*>(data), size);
        }
        if (fd >= 0)
        {
            close(fd);
        }
    }

    // lib.rs:314-413
    void ParseIndex()
    {
        Cursor header{data, size, 0};
        const uint32_t magic = header.Get<uint32_t>("header");
        if (magic != kFileMagic && magic != kLegacyFileMagic)
        {
            Failure f = Fail(CX_CXW_BAD_MAGIC, "bad file magic 0x%08X (expected CXLS or VXLS)", magic);
            f.e.actual = magic;
            throw f;
        }
        const uint16_t version = header.Get<uint16_t>("header");
        if (version > kCurrentVersion)
        {
            Failure f = Fail(CX_CXW_UNSUPPORTED_VERSION, "format version %u is not supported (newest supported: %u)",
                             version, kCurrentVersion);
            f.e.actual = version;
            throw f;
        }
        const uint16_t flags = header.Get<uint16_t>("header");
        if (flags != kFlagDeflate)
        {
            Failure f = Fail(CX_CXW_BAD_FLAGS, "unsupported save flags 0x%04X (must be 1 = Deflate)", flags);
            f.e.actual = flags;
            throw f;
        }
        const uint64_t table = header.Get<uint64_t>("header");
        if (table < kHeaderBytes || table > size)
        {
            throw Truncated("entity table", table);
        }
        Cursor c{data, size, size_t(table)};
        world.version = version;
        if (version >= kVersionWorldRecord)
        {
            const uint8_t world_flags = c.Get<uint8_t>("world record");
            float position[3];
            float rotation[4];
            c.Floats(position, "world record");
            c.Floats(rotation, "world record");
            if ((world_flags & kWorldHasPlayerPose) != 0)
            {
                world.has_player_pose = 1;
                std::memcpy(world.pose_position, position, sizeof(position));
                std::memcpy(world.pose_rotation, rotation, sizeof(rotation));
            }
        }
        const uint32_t count = c.Get<uint32_t>("entity count");
        std::vector<uint64_t> index_offsets;
        for (uint32_t i = 0; i < count; ++i)
        {
            const char* what = "entity record";
            CxCxwEntity e{};
            std::memcpy(e.guid, c.Take(16, what), 16);
            c.Floats(e.position, what);
            c.Floats(e.rotation, what);
            e.require_update_flags = c.Get<uint16_t>(what);
            if (version >= kVersionFlagsV6)
            {
                e.flags = c.Get<uint8_t>(what) & (kEntityHasBody | kEntityStatic);
            }
            else if (version >= kVersionEntityFlags)
            {
                e.flags = FlagsFromLegacyBodyState(c.Get<uint8_t>(what));
            }
            else
            {
                e.flags = kEntityStatic;
            }
            if (version >= kVersionVelocity)
            {
                c.Floats(e.velocity, what);
            }
            e.is_protected = version >= kVersionProtected && c.Get<uint8_t>(what) != 0 ? 1 : 0;
            index_offsets.push_back(c.Get<uint64_t>(what));
            e.sector_count = c.Get<uint32_t>(what);
            entities.push_back(e);
        }
        std::set<std::tuple<int32_t, int32_t, int32_t>> seen;
        for (uint32_t entity = 0; entity < count; ++entity)
        {
            const uint64_t start = index_offsets[entity];
            const uint64_t bytes = uint64_t(entities[entity].sector_count) * kSectorEntryBytes;
            if (start < kHeaderBytes || start > size || bytes > size - start)
            {
                throw Truncated("sector index", start);
            }
            Cursor s{data, size, size_t(start)};
            seen.clear();
            for (uint32_t k = 0; k < entities[entity].sector_count; ++k)
            {
                CxCxwSector ref{};
                ref.entity = entity;
                std::memcpy(ref.coord, s.Take(12, "sector index"), 12);
                ref.region_offset = s.Get<uint64_t>("sector index");
                ref.region_size = s.Get<uint32_t>("sector index");
                CheckRegion(data, size, ref.coord, ref.region_offset, ref.region_size);
                if (!seen.insert({ref.coord[0], ref.coord[1], ref.coord[2]}).second)
                {
                    Failure f = Fail(CX_CXW_DUPLICATE_SECTOR, "sector [%d, %d, %d] appears twice in one entity's sector index",
                                     ref.coord[0], ref.coord[1], ref.coord[2]);
                    std::memcpy(f.e.coord, ref.coord, sizeof(f.e.coord));
                    throw f;
                }
                sectors.push_back(ref);
            }
        }
        world.entity_count = count;
        world.sector_count = sectors.size();
    }

    void Decode(uint64_t index, Inflater& inflater, std::vector<uint8_t>& buffer, Decoded& out) const
    {
        const CxCxwSector& ref = sectors[index];
        const Region r = CheckRegion(data, size, ref.coord, ref.region_offset, ref.region_size);
        if (!inflater.ready)
        {
            throw Fail(CX_CXW_INFLATE, "inflate failed: inflateInit2");
        }
        Inflate(inflater.z, r.payload, r.payload_size, buffer, ref.coord);
        out = Decoded{};
        out.entity = ref.entity;
        DecodePayload(buffer, r.preview, ref.coord, out);
    }
};

extern "C" {

CX_API int cx_cxw_open(const char* path, CxCxw** out, CxCxwError* error)
{
    if (path == nullptr || out == nullptr)
    {
        Report(error, Fail(CX_CXW_IO, "cx_cxw_open: NULL path or handle"));
        return CX_ERROR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    CxCxw* f = new (std::nothrow) CxCxw();
    if (f == nullptr)
    {
        return CX_ERROR_OUT_OF_MEMORY;
    }
    try
    {
        f->fd = open(path, O_RDONLY | O_CLOEXEC);
        struct stat st;
        if (f->fd < 0 || fstat(f->fd, &st) != 0)
        {
            throw Fail(CX_CXW_IO, "I/O error: %s: %s", path, strerror(errno));
        }
        f->size = uint64_t(st.st_size);
        if (f->size < kHeaderBytes)
        {
            throw Truncated("header", 0);
        }
        void* map = mmap(nullptr, f->size, PROT_READ, MAP_PRIVATE, f->fd, 0);
        if (map == MAP_FAILED)
        {
            throw Fail(CX_CXW_IO, "I/O error: mmap %s: %s", path, strerror(errno));
        }
        f->data = static_cast<const uint8_t*>(map);
        f->ParseIndex();
    }
    catch (const Failure& failure)
    {
        Report(error, failure);
        delete f;
        return CX_ERROR_INVALID_ARGUMENT;
    }
    catch (const std::bad_alloc&)
    {
        delete f;
        return CX_ERROR_OUT_OF_MEMORY;
    }
    if (error != nullptr)
    {
        *error = CxCxwError{};
    }
    *out = f;
    return CX_OK;
}

CX_API void cx_cxw_close(CxCxw* file)
{
    delete file;
}

CX_API int cx_cxw_world(const CxCxw* file, CxCxwWorld* out)
{
    if (file == nullptr || out == nullptr)
    {
        return CX_ERROR_INVALID_ARGUMENT;
    }
    *out = file->world;
    return CX_OK;
}

CX_API int cx_cxw_entity(const CxCxw* file, uint32_t index, CxCxwEntity* out)
{
    if (file == nullptr || out == nullptr || index >= file->entities.size())
    {
        return CX_ERROR_INVALID_ARGUMENT;
    }
    *out = file->entities[index];
    return CX_OK;
}

CX_API int cx_cxw_sector(const CxCxw* file, uint64_t index, CxCxwSector* out)
{
    if (file == nullptr || out == nullptr || index >= file->sectors.size())
    {
        return CX_ERROR_INVALID_ARGUMENT;
    }
    *out = file->sectors[index];
    return CX_OK;
}

CX_API int cx_cxw_decode(const CxCxw* file, uint64_t sector, int32_t* keys, uint16_t* blocks, uint32_t* summaries,
                         uint32_t* brick_count, uint64_t* voxels, CxCxwError* error)
{
    if (file == nullptr || sector >= file->sectors.size() || keys == nullptr || blocks == nullptr ||
        summaries == nullptr || brick_count == nullptr || voxels == nullptr)
    {
        return CX_ERROR_INVALID_ARGUMENT;
    }
    try
    {
        Inflater inflater;
        std::vector<uint8_t> buffer;
        Decoded d;
        file->Decode(sector, inflater, buffer, d);
        const uint32_t n = uint32_t(d.summaries.size());
        std::memcpy(keys, d.keys.data(), d.keys.size() * 4);
        std::memcpy(blocks, d.blocks.data(), d.blocks.size() * 2);
        std::memcpy(summaries, d.summaries.data(), size_t(n) * 4);
        *brick_count = n;
        *voxels = d.voxels;
    }
    catch (const Failure& failure)
    {
        Report(error, failure);
        return CX_ERROR_INVALID_ARGUMENT;
    }
    catch (const std::bad_alloc&)
    {
        return CX_ERROR_OUT_OF_MEMORY;
    }
    if (error != nullptr)
    {
        *error = CxCxwError{};
    }
    return CX_OK;
}

CX_API int cx_cxw_decode_all(CxCxw* file, uint32_t threads, uint64_t* brick_count, uint64_t* voxels,
                             CxCxwError* error)
{
    if (file == nullptr || brick_count == nullptr || voxels == nullptr)
    {
        return CX_ERROR_INVALID_ARGUMENT;
    }
    const uint64_t count = file->sectors.size();
    file->decoded.clear();
    try
    {
        file->decoded.resize(count);
    }
    catch (const std::bad_alloc&)
    {
        return CX_ERROR_OUT_OF_MEMORY;
    }
    if (threads == 0)
    {
        threads = std::max(1u, std::thread::hardware_concurrency());
    }
    threads = uint32_t(std::min<uint64_t>(threads, std::max<uint64_t>(count, 1)));
    std::atomic<uint64_t> next{0};
    std::atomic<bool> failed{false};
    std::vector<Failure> failures(threads);
    std::vector<int> bad(threads, 0);
    auto worker = [&](uint32_t w) {
        Inflater inflater;
        std::vector<uint8_t> buffer;
        while (!failed.load(std::memory_order_relaxed))
        {
            const uint64_t i = next.fetch_add(1, std::memory_order_relaxed);
            if (i >= count)
            {
                return;
            }
            try
            {
                file->Decode(i, inflater, buffer, file->decoded[i]);
            }
            catch (const Failure& failure)
            {
                failures[w] = failure;
                bad[w] = CX_ERROR_INVALID_ARGUMENT;
                failed.store(true);
                return;
            }
            catch (const std::bad_alloc&)
            {
                failures[w] = Fail(CX_CXW_IO, "out of memory");
                bad[w] = CX_ERROR_OUT_OF_MEMORY;
                failed.store(true);
                return;
            }
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t w = 1; w < threads; ++w)
    {
        pool.emplace_back(worker, w);
    }
    worker(0);
    for (std::thread& t : pool)
    {
        t.join();
    }
    for (uint32_t w = 0; w < threads; ++w)
    {
        if (bad[w] != 0)
        {
            Report(error, failures[w]);
            file->decoded.clear();
            return bad[w];
        }
    }
    uint64_t bricks = 0;
    uint64_t non_zero = 0;
    for (const Decoded& d : file->decoded)
    {
        bricks += d.summaries.size();
        non_zero += d.voxels;
    }
    *brick_count = bricks;
    *voxels = non_zero;
    if (error != nullptr)
    {
        *error = CxCxwError{};
    }
    return CX_OK;
}

CX_API int cx_cxw_copy_decoded(const CxCxw* file, int32_t* keys, uint16_t* blocks, uint32_t* summaries,
                               uint32_t* entity_of_brick, uint64_t capacity_bricks)
{
    if (file == nullptr || keys == nullptr || blocks == nullptr)
    {
        return CX_ERROR_INVALID_ARGUMENT;
    }
    uint64_t total = 0;
    for (const Decoded& d : file->decoded)
    {
        total += d.summaries.size();
    }
    if (total > capacity_bricks || (file->decoded.empty() && !file->sectors.empty()))
    {
        return CX_ERROR_INVALID_ARGUMENT;
    }
    uint64_t at = 0;
    for (const Decoded& d : file->decoded)
    {
        const uint64_t n = d.summaries.size();
        std::memcpy(keys + at * 3, d.keys.data(), n * 12);
        std::memcpy(blocks + at * kBlocksInBrick, d.blocks.data(), n * kBlocksInBrick * 2);
        if (summaries != nullptr)
        {
            std::memcpy(summaries + at, d.summaries.data(), n * 4);
        }
        if (entity_of_brick != nullptr)
        {
            std::fill(entity_of_brick + at, entity_of_brick + at + n, d.entity);
        }
        at += n;
    }
    return CX_OK;
}

}  // extern "C"
