// Hash table kernels for gfx950.
//
// The reference's table (lib/src/rhi/hash_table.cpp:79-130, shaders/hash/*.hlsl,
// shaders/common/caelix_hash_table.hlsl): batch lookups and lookups-with-insert (one thread per key,
// CAS claims, wave-aggregated result counters), and the per-frame maintenance (init, begin, key fill,
// maintain = evict by age + rebuild + active list, finish).
//
// The bulk brick table: a deterministic build of a whole key set at any load (ordered linear probing).
// Keys are sorted by home slot (stable in input order) with the onesweep sort; a key's slot is then
// p_i = max(home_i, p_(i-1) + 1) = i + max_(j <= i)(home_j - j), one max-scan; the few keys past the
// table end wrap to the first free slots from slot 0 in order. That is the layout a sequential insert
// of the keys in (home, input index) order gives, which is what the oracle does. The probe limit of
// the table is the longest displacement + 1.
#include "hash_device.h"
#include "radix_sort.h"
#include "runtime.h"

namespace cx {
namespace {

constexpr uint32_t kHashBlock = 256;

HashTableView ViewOf(const CxHashTable& t)
{
    HashTableView v;
    v.keys = static_cast<HashKeySlot*>(t.keys);
    v.stack = static_cast<uint32_t*>(t.stack);
    v.counters = static_cast<HashCounters*>(t.counters);
    v.payloads = static_cast<uint8_t*>(t.payloads);
    v.slot_count = t.slot_count;
    v.payload_stride = t.payload_stride;
    v.probe_limit = t.probe_limit;
    v.pow2 = (t.slot_count & (t.slot_count - 1)) == 0 ? 1u : 0u;
    return v;
}

__global__ void FingerprintKernel(const uint32_t* words, uint32_t word_count, uint32_t count, uint32_t* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
    {
        return;
    }
    uint32_t w[kHashMaxKeyWords];
    for (uint32_t k = 0; k < word_count; ++k)
    {
        w[k] = words[uint64_t(i) * word_count + k];
    }
    const uint2 fp = Fingerprint(w, word_count);
    out[2 * uint64_t(i)] = fp.x;
    out[2 * uint64_t(i) + 1] = fp.y;
}

// HashCountResult (caelix_hash_table.hlsl:243-256): one atomic per counter and wave.
CX_DEV void CountResult(const HashTableView& t, bool active, uint32_t result)
{
    const uint32_t inserted = PopC(Ballot(active && result == kHashInserted));
    const uint32_t pending = PopC(Ballot(active && result == kHashPending));
    const uint32_t failed_slot = PopC(Ballot(active && result == kHashFullWindow));
    const uint32_t failed_payload = PopC(Ballot(active && result == kHashNoPayload));
    if (Lane() == 0)
    {
        if (inserted != 0)
        {
            AtomicAddRelaxed(&t.counters->inserted, inserted);
        }
        if (pending != 0)
        {
            AtomicAddRelaxed(&t.counters->pending, pending);
        }
        if (failed_slot != 0)
        {
            AtomicAddRelaxed(&t.counters->failed_slot, failed_slot);
        }
        if (failed_payload != 0)
        {
            AtomicAddRelaxed(&t.counters->failed_payload, failed_payload);
        }
    }
}

__global__ void __launch_bounds__(kHashBlock) LookupKernel(HashTableView t, const uint32_t* words, uint32_t word_count,
                                                           uint32_t count, uint32_t frame, uint32_t insert,
                                                           uint32_t* results, uint32_t* payloads)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < count;
    uint32_t result = kHashMissing;
    if (active)
    {
        uint32_t w[kHashMaxKeyWords] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t k = 0; k < word_count; ++k)
        {
            w[k] = words[uint64_t(i) * word_count + k];
        }
        const uint2 key = Fingerprint(w, word_count);
        uint32_t payload;
        result = HashFind(t, key, w, word_count, frame, insert != 0, payload);
        results[i] = result;
        payloads[i] = payload;
    }
    if (insert != 0)
    {
        CountResult(t, active, result);
    }
}

// caelix_hash_init.hlsl: every payload free, the free stack 0 .. capacity - 1.
__global__ void HashInitKernel(HashTableView t, uint32_t capacity)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= capacity)
    {
        return;
    }
    t.stack[kHashStackEntries / 4 + i] = i;
    *reinterpret_cast<uint32_t*>(t.payloads + uint64_t(i) * t.payload_stride + kHashPayloadState) = kHashStateFree;
    if (i == 0)
    {
        t.stack[kHashStackTop / 4] = capacity;
    }
}

// caelix_hash_begin.hlsl: clamp the stack top, clear the counters and the dispatch.
__global__ void HashBeginKernel(HashTableView t, uint4* dispatch)
{
    const int32_t top = int32_t(t.stack[kHashStackTop / 4]);
    if (top < 0)
    {
        t.stack[kHashStackTop / 4] = 0;
    }
    uint32_t* c = reinterpret_cast<uint32_t*>(t.counters);
    for (uint32_t w = 0; w < sizeof(HashCounters) / 4; ++w)
    {
        c[w] = 0;
    }
    *dispatch = make_uint4(0, 1, 1, 0);
}

__global__ void FillKeysKernel(uint4* keys, uint32_t slot_count)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < slot_count; i += gridDim.x * blockDim.x)
    {
        keys[i] = make_uint4(0, 0, 0, 0);
    }
}

// caelix_hash_maintain.hlsl: evict by age, rebuild the keys of the live payloads, the active list.
__global__ void __launch_bounds__(kHashBlock) HashMaintainKernel(HashTableView t, uint32_t capacity,
                                                                 uint32_t key_words, uint32_t frame,
                                                                 uint32_t max_age, uint32_t* active_list,
                                                                 uint4* dispatch)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool freed = false;
    bool evicted = false;
    bool dropped = false;
    bool live = false;
    bool active = false;
    if (i < capacity)
    {
        uint32_t* header = reinterpret_cast<uint32_t*>(t.payloads + uint64_t(i) * t.payload_stride);
        const uint32_t state = header[kHashPayloadState / 4];
        if (state != kHashStateFree)
        {
            const uint32_t touched = header[kHashPayloadTouched / 4];
            if (frame - touched > max_age)
            {
                evicted = true;
            }
            else
            {
                uint32_t w[kHashMaxKeyWords] = {0, 0, 0, 0, 0, 0, 0, 0};
                for (uint32_t k = 0; k < key_words; ++k)
                {
                    w[k] = header[kHashPayloadKey / 4 + k];
                }
                const uint2 key = Fingerprint(w, key_words);
                bool inserted = false;
                for (uint32_t probe = 0; probe < t.probe_limit && !inserted; ++probe)
                {
                    HashKeySlot* slot = t.keys + HashSlotIndex(t, key.x, probe);
                    uint32_t expected = 0u;
                    if (__hip_atomic_compare_exchange_strong(&slot->key_lo, &expected, key.x, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    {
                        StoreRelaxed(&slot->payload, i);
                        StoreRelaxed(&slot->key_hi, key.y);
                        inserted = true;
                    }
                }
                if (inserted)
                {
                    live = true;
                    active = touched == frame - 1u;
                }
                else
                {
                    dropped = true;
                }
            }
            freed = evicted || dropped;
            if (freed)
            {
                header[kHashPayloadState / 4] = kHashStateFree;
            }
        }
    }
    const uint64_t freed_mask = Ballot(freed);
    uint32_t push_base = 0;
    if (Lane() == 0 && freed_mask != 0)
    {
        push_base = AtomicAddRelaxed(t.stack + kHashStackTop / 4, PopC(freed_mask));
    }
    push_base = __shfl(push_base, 0, kWave);
    if (freed)
    {
        t.stack[kHashStackEntries / 4 + push_base + PopC(freed_mask & LanesBelow())] = i;
    }
    const uint64_t active_mask = Ballot(active);
    uint32_t active_base = 0;
    if (Lane() == 0 && active_mask != 0)
    {
        active_base = AtomicAddRelaxed(&reinterpret_cast<uint32_t*>(dispatch)[3], PopC(active_mask));
    }
    active_base = __shfl(active_base, 0, kWave);
    if (active)
    {
        active_list[active_base + PopC(active_mask & LanesBelow())] = i;
    }
    const uint32_t live_count = PopC(Ballot(live));
    const uint32_t evicted_count = PopC(Ballot(evicted));
    const uint32_t dropped_count = PopC(Ballot(dropped));
    if (Lane() == 0)
    {
        if (live_count != 0)
        {
            AtomicAddRelaxed(&t.counters->live, live_count);
        }
        if (evicted_count != 0)
        {
            AtomicAddRelaxed(&t.counters->evicted, evicted_count);
        }
        if (dropped_count != 0)
        {
            AtomicAddRelaxed(&t.counters->rebuild_dropped, dropped_count);
        }
    }
}

// caelix_hash_finish.hlsl: the user pass's dispatch from the active count.
__global__ void HashFinishKernel(HashTableView t, uint4* dispatch, uint32_t items_per_group)
{
    const uint32_t count = reinterpret_cast<uint32_t*>(dispatch)[3];
    const uint32_t per_group = max(items_per_group, 1u);
    const uint32_t groups = (count + per_group - 1u) / per_group;
    const uint32_t row = 32768u;
    const uint32_t x = min(groups, row);
    const uint32_t y = (groups + row - 1u) / row;
    *dispatch = make_uint4(x, max(y, 1u), 1u, count);
    t.counters->active = count;
}

// ------------------------------------------------------------------------------------------------
// Bulk brick table.

struct BulkScratch
{
    uint32_t* homes[2];
    uint32_t* index[2];
    uint32_t* key_lo;
    uint32_t* key_hi;
    int32_t* position;
    uint64_t* scan_status;
    uint32_t* ctl;  // [0] scan tile counter, [1] first wrapped sorted index, [2] max displacement
    SortState sort;
};

constexpr uint32_t kBulkScanTile = 4096;

uint64_t BulkScratchBytes(uint32_t count, BulkScratch* out, uint8_t* base)
{
    const uint64_t n = count;
    const uint32_t tiles = DivUp(count, kBulkScanTile);
    uint64_t offset = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t at = offset;
        offset += (bytes + 255) & ~uint64_t(255);
        return base + at;
    };
    uint8_t* h0 = take(4 * n);
    uint8_t* h1 = take(4 * n);
    uint8_t* i0 = take(4 * n);
    uint8_t* i1 = take(4 * n);
    uint8_t* lo = take(4 * n);
    uint8_t* hi = take(4 * n);
    uint8_t* pos = take(4 * n);
    uint8_t* status = take(8 * uint64_t(tiles));
    uint8_t* ctl = take(256);
    uint8_t* sort = take(SortStateBytes(count));
    if (out != nullptr)
    {
        out->homes[0] = reinterpret_cast<uint32_t*>(h0);
        out->homes[1] = reinterpret_cast<uint32_t*>(h1);
        out->index[0] = reinterpret_cast<uint32_t*>(i0);
        out->index[1] = reinterpret_cast<uint32_t*>(i1);
        out->key_lo = reinterpret_cast<uint32_t*>(lo);
        out->key_hi = reinterpret_cast<uint32_t*>(hi);
        out->position = reinterpret_cast<int32_t*>(pos);
        out->scan_status = reinterpret_cast<uint64_t*>(status);
        out->ctl = reinterpret_cast<uint32_t*>(ctl);
        out->sort = SortStateAt(sort, count);
    }
    return offset;
}

__global__ void BulkPrepareKernel(const uint32_t* words, uint32_t word_count, uint32_t count, uint32_t slot_count,
                                  uint32_t pow2, BulkScratch s, uint32_t tiles)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < tiles)
    {
        s.scan_status[i] = 0;
    }
    if (i == 0)
    {
        s.ctl[0] = 0;
        s.ctl[1] = count;
        s.ctl[2] = 0;
    }
    if (i >= count)
    {
        return;
    }
    const uint2 key = Fingerprint(words + uint64_t(i) * word_count, word_count);
    s.homes[0][i] = pow2 != 0 ? (key.x & (slot_count - 1)) : (key.x % slot_count);
    s.index[0][i] = i;
    s.key_lo[i] = key.x;
    s.key_hi[i] = key.y;
}

constexpr uint64_t kMaxAggregate = 1ull << 32;
constexpr uint64_t kMaxPrefix = 2ull << 32;
constexpr uint64_t kMaxFlagMask = 3ull << 32;

// position[i] = i + max over j <= i of (home_j - j), over the sorted homes (decoupled look-back with
// max as the operator). Then the keys that stay inside the table are written to their slots.
__global__ void __launch_bounds__(256) BulkPlaceKernel(uint32_t count, uint32_t slot_count, BulkScratch s,
                                                       HashKeySlot* slots)
{
    __shared__ uint32_t lds_tile;
    __shared__ int32_t lds_wave[4];
    __shared__ int32_t lds_prefix;
    const uint32_t t = threadIdx.x;
    if (t == 0)
    {
        lds_tile = AtomicAddRelaxed(&s.ctl[0], 1u);
    }
    __syncthreads();
    const uint32_t tile = lds_tile;
    const uint32_t base = tile * kBulkScanTile + t * 16;
    int32_t v[16];
    int32_t running = INT32_MIN;
#pragma unroll
    for (int k = 0; k < 16; ++k)
    {
        const uint32_t i = base + k;
        const int32_t x = i < count ? int32_t(s.homes[0][i]) - int32_t(i) : INT32_MIN;
        running = max(running, x);
        v[k] = running;  // inclusive within the thread
    }
    // Exclusive max over the threads of the block (wave inclusive scan, then the waves).
    int32_t incl = running;
    const uint32_t lane = Lane();
#pragma unroll
    for (int offset = 1; offset < kWave; offset <<= 1)
    {
        const int32_t up = __shfl_up(incl, offset, kWave);
        if (lane >= uint32_t(offset))
        {
            incl = max(incl, up);
        }
    }
    int32_t excl = __shfl_up(incl, 1, kWave);
    if (lane == 0)
    {
        excl = INT32_MIN;
    }
    const uint32_t wave = t / kWave;
    if (lane == kWave - 1)
    {
        lds_wave[wave] = incl;
    }
    __syncthreads();
    int32_t before_waves = INT32_MIN;
    int32_t tile_max = INT32_MIN;
    for (uint32_t w = 0; w < 4; ++w)
    {
        if (w < wave)
        {
            before_waves = max(before_waves, lds_wave[w]);
        }
        tile_max = max(tile_max, lds_wave[w]);
    }
    // Look-back over the earlier tiles.
    if (t == 0)
    {
        int32_t prefix = INT32_MIN;
        if (tile == 0)
        {
            StoreRelaxed(&s.scan_status[0], kMaxPrefix | uint32_t(tile_max));
        }
        else
        {
            StoreRelaxed(&s.scan_status[tile], kMaxAggregate | uint32_t(tile_max));
            int64_t p = int64_t(tile) - 1;
            while (true)
            {
                const uint64_t word = LoadRelaxed(&s.scan_status[p]);
                const uint64_t flag = word & kMaxFlagMask;
                if (flag == 0)
                {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                prefix = max(prefix, int32_t(uint32_t(word)));
                if (flag == kMaxPrefix)
                {
                    break;
                }
                --p;
            }
            StoreRelaxed(&s.scan_status[tile], kMaxPrefix | uint32_t(max(prefix, tile_max)));
        }
        lds_prefix = prefix;
    }
    __syncthreads();
    const int32_t carry = max(lds_prefix, max(before_waves, excl));
    uint32_t local_max_disp = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k)
    {
        const uint32_t i = base + k;
        if (i < count)
        {
            const int64_t position = int64_t(i) + int64_t(max(carry, v[k]));
            s.position[i] = int32_t(min(position, int64_t(INT32_MAX)));
            if (position < int64_t(slot_count))
            {
                const uint32_t key_index = s.index[0][i];
                slots[position] = HashKeySlot{s.key_lo[key_index], s.key_hi[key_index], key_index, 0u};
                local_max_disp = max(local_max_disp, uint32_t(position) - s.homes[0][i]);
            }
            else
            {
                atomicMin(&s.ctl[1], i);  // positions increase: the smallest wrapped index is the first
            }
        }
    }
    if (local_max_disp != 0)
    {
        atomicMax(&s.ctl[2], local_max_disp);
    }
}

// The wrapped keys (sorted indices ctl[1] .. count - 1) take the first free slots from slot 0.
__global__ void __launch_bounds__(1024) BulkWrapKernel(uint32_t count, uint32_t slot_count, BulkScratch s,
                                                       HashKeySlot* slots, uint32_t* info)
{
    __shared__ uint32_t lds_assigned;
    __shared__ uint32_t lds_wave_free[16];
    const uint32_t first = s.ctl[1];
    const uint32_t wrapped = count > first ? count - first : 0u;
    if (threadIdx.x == 0)
    {
        lds_assigned = 0;
    }
    __syncthreads();
    uint32_t max_disp = s.ctl[2];
    for (uint32_t chunk = 0; chunk < slot_count && lds_assigned < wrapped; chunk += 1024)
    {
        const uint32_t slot = chunk + threadIdx.x;
        const bool empty = slot < slot_count && slots[slot].key_lo == 0u;
        const uint64_t mask = Ballot(empty);
        const uint32_t wave = threadIdx.x / kWave;
        if (Lane() == 0)
        {
            lds_wave_free[wave] = PopC(mask);
        }
        __syncthreads();
        uint32_t before = lds_assigned;
        for (uint32_t w = 0; w < wave; ++w)
        {
            before += lds_wave_free[w];
        }
        const uint32_t rank = before + PopC(mask & LanesBelow());
        if (empty && rank < wrapped)
        {
            const uint32_t sorted = first + rank;
            const uint32_t key_index = s.index[0][sorted];
            slots[slot] = HashKeySlot{s.key_lo[key_index], s.key_hi[key_index], key_index, 0u};
            max_disp = max(max_disp, slot + slot_count - s.homes[0][sorted]);
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            uint32_t total = 0;
            for (uint32_t w = 0; w < 16; ++w)
            {
                total += lds_wave_free[w];
            }
            lds_assigned += total;
        }
        __syncthreads();
    }
    // max displacement over the block
    for (int offset = kWave / 2; offset > 0; offset >>= 1)
    {
        max_disp = max(max_disp, __shfl_xor(max_disp, offset, kWave));
    }
    if (Lane() == 0)
    {
        lds_wave_free[threadIdx.x / kWave] = max_disp;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t m = 0;
        for (uint32_t w = 0; w < 16; ++w)
        {
            m = max(m, lds_wave_free[w]);
        }
        info[0] = count > 0 ? m + 1 : 1u;  // probe limit
        info[1] = m;                        // longest displacement
        info[2] = wrapped;
    }
}

__global__ void __launch_bounds__(256) GetBricksKernel(const HashKeySlot* slots, uint32_t slot_count,
                                                       const uint32_t* info, const int32_t* keys_xyz, uint32_t count,
                                                       uint32_t* payloads, uint32_t* probes)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
    {
        return;
    }
    const uint2 key = BrickFingerprint(keys_xyz[3 * uint64_t(i)], keys_xyz[3 * uint64_t(i) + 1],
                                       keys_xyz[3 * uint64_t(i) + 2]);
    uint32_t p = 0;
    const uint32_t pow2 = (slot_count & (slot_count - 1)) == 0 ? 1u : 0u;
    payloads[i] = HashGet(slots, slot_count, pow2, info[0], key, &p);
    if (probes != nullptr)
    {
        probes[i] = p;
    }
}

}  // namespace
}  // namespace cx

extern "C" {

CX_API int cx_hash_fingerprint(const uint32_t* words, uint32_t word_count, uint32_t count, uint32_t* out,
                               void* stream)
{
    if (words == nullptr || out == nullptr || word_count == 0 || word_count > kHashMaxKeyWords)
    {
        return cx::ArgError("cx_hash_fingerprint: NULL buffer or word count outside 1..8");
    }
    if (count == 0)
    {
        return CX_OK;
    }
    cx::FingerprintKernel<<<cx::DivUp(count, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(words, word_count,
                                                                                               count, out);
    return cx::LaunchResult("cx_hash_fingerprint");
}

static int CheckTable(const CxHashTable* table, const char* what)
{
    if (table == nullptr || table->keys == nullptr || table->slot_count == 0 || table->probe_limit == 0)
    {
        return cx::ArgError("%s: NULL table, no key slots or probe limit 0", what);
    }
    return CX_OK;
}

CX_API int cx_hash_lookup(const CxHashTable* table, const uint32_t* words, uint32_t word_count, uint32_t count,
                          uint32_t frame, uint32_t insert, uint32_t* results, uint32_t* payloads, void* stream)
{
    if (int r = CheckTable(table, "cx_hash_lookup"); r != CX_OK)
    {
        return r;
    }
    if (words == nullptr || results == nullptr || payloads == nullptr || word_count == 0 ||
        word_count > kHashMaxKeyWords || (insert != 0 && (table->stack == nullptr || table->counters == nullptr ||
                                                          table->payloads == nullptr)))
    {
        return cx::ArgError("cx_hash_lookup: NULL buffer or word count outside 1..8");
    }
    if (count == 0)
    {
        return CX_OK;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::StageScope scope(s, "hash_lookup");
    cx::LookupKernel<<<cx::DivUp(count, cx::kHashBlock), cx::kHashBlock, 0, s>>>(
        cx::ViewOf(*table), words, word_count, count, frame, insert, results, payloads);
    return cx::LaunchResult("cx_hash_lookup");
}

CX_API int cx_hash_maintain(const CxHashTable* table, uint32_t payload_capacity, uint32_t key_words, uint32_t frame,
                            uint32_t max_age, uint32_t items_per_group, uint32_t initialise, uint32_t* active,
                            uint32_t* dispatch, void* stream)
{
    if (int r = CheckTable(table, "cx_hash_maintain"); r != CX_OK)
    {
        return r;
    }
    if (table->stack == nullptr || table->counters == nullptr || table->payloads == nullptr || active == nullptr ||
        dispatch == nullptr || payload_capacity == 0 || key_words == 0 || key_words > kHashMaxKeyWords ||
        table->payload_stride < kHashPayloadKey + 4 * key_words || table->payload_stride % 4 != 0)
    {
        return cx::ArgError("cx_hash_maintain: NULL buffer, or a payload stride or key word count out of range");
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const cx::HashTableView v = cx::ViewOf(*table);
    uint4* d = reinterpret_cast<uint4*>(dispatch);
    if (initialise != 0)
    {
        cx::HashInitKernel<<<cx::DivUp(payload_capacity, 256), 256, 0, s>>>(v, payload_capacity);
    }
    cx::HashBeginKernel<<<1, 1, 0, s>>>(v, d);
    cx::FillKeysKernel<<<std::min(cx::DivUp(table->slot_count, 256), 2048u), 256, 0, s>>>(
        reinterpret_cast<uint4*>(table->keys), table->slot_count);
    cx::HashMaintainKernel<<<cx::DivUp(payload_capacity, cx::kHashBlock), cx::kHashBlock, 0, s>>>(
        v, payload_capacity, key_words, frame, max_age, active, d);
    cx::HashFinishKernel<<<1, 1, 0, s>>>(v, d, items_per_group);
    return cx::LaunchResult("cx_hash_maintain");
}

CX_API uint64_t cx_hash_build_scratch_bytes(uint32_t count)
{
    return cx::BulkScratchBytes(count, nullptr, nullptr);
}

CX_API int cx_hash_build_bricks(const int32_t* keys_xyz, uint32_t count, HashKeySlot* slots, uint32_t slot_count,
                                void* scratch, uint32_t* info, void* stream)
{
    return cx_hash_build_keys(reinterpret_cast<const uint32_t*>(keys_xyz), 3, count, slots, slot_count, scratch, info,
                              stream);
}

CX_API int cx_hash_build_keys(const uint32_t* words, uint32_t word_count, uint32_t count, HashKeySlot* slots,
                              uint32_t slot_count, void* scratch, uint32_t* info, void* stream)
{
    if (words == nullptr || slots == nullptr || scratch == nullptr || info == nullptr)
    {
        return cx::ArgError("cx_hash_build_keys: a buffer is NULL");
    }
    if (word_count == 0 || word_count > 8)
    {
        return cx::ArgError("cx_hash_build_keys: %u key words (1..8)", word_count);
    }
    if (slot_count == 0 || slot_count > (1u << 31) || count >= slot_count)
    {
        return cx::ArgError("cx_hash_build_keys: %u keys do not fit %u slots (need count < slots <= 2^31)", count,
                            slot_count);
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::BulkScratch b;
    cx::BulkScratchBytes(count, &b, static_cast<uint8_t*>(scratch));
    const uint32_t pow2 = (slot_count & (slot_count - 1)) == 0 ? 1u : 0u;
    const uint32_t n = std::max(count, 1u);
    const uint32_t tiles = cx::DivUp(n, cx::kBulkScanTile);
    cx::StageScope scope(s, "hash_build");
    cx::FillKeysKernel<<<std::min(cx::DivUp(slot_count, 256), 4096u), 256, 0, s>>>(reinterpret_cast<uint4*>(slots),
                                                                                  slot_count);
    cx::BulkPrepareKernel<<<cx::DivUp(std::max(n, tiles), 256), 256, 0, s>>>(words, word_count, count, slot_count, pow2,
                                                                            b, tiles);
    if (count > 0)
    {
        // count lives in ctl[1] (set by the prepare kernel) for the sort's device-side count.
        cx::LaunchSortInit(b.sort, s);
        cx::LaunchSortHistogram(b.homes[0], &b.ctl[1], nullptr, b.sort, s);
        for (uint32_t pass = 0; pass < cx::kSortPasses; ++pass)
        {
            const uint32_t in = pass & 1u;
            cx::LaunchSortPass(b.homes[in], b.index[in], b.homes[1 - in], b.index[1 - in], &b.ctl[1], nullptr,
                               b.sort, pass, s);
        }
        cx::BulkPlaceKernel<<<tiles, 256, 0, s>>>(count, slot_count, b, slots);
    }
    cx::BulkWrapKernel<<<1, 1024, 0, s>>>(count, slot_count, b, slots, info);
    return cx::LaunchResult("cx_hash_build_keys");
}

CX_API int cx_hash_get_bricks(const HashKeySlot* slots, uint32_t slot_count, const uint32_t* info,
                              const int32_t* keys_xyz, uint32_t count, uint32_t* payloads, uint32_t* probes,
                              void* stream)
{
    if (slots == nullptr || info == nullptr || keys_xyz == nullptr || payloads == nullptr || slot_count == 0)
    {
        return cx::ArgError("cx_hash_get_bricks: a buffer is NULL or no slots");
    }
    if (count == 0)
    {
        return CX_OK;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::StageScope scope(s, "hash_get");
    cx::GetBricksKernel<<<cx::DivUp(count, 256), 256, 0, s>>>(slots, slot_count, info, keys_xyz, count, payloads,
                                                             probes);
    return cx::LaunchResult("cx_hash_get_bricks");
}

}  // extern "C"
