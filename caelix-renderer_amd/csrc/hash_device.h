// Device API of the 64-bit-key hash table (the reference's shaders/common/caelix_hash_table.hlsl:
// 82-275) for gfx950 kernels: fingerprint, lookup, lookup-or-insert with a CAS claim, rebuild
// insert. The slot format, payload header words, probe window and result codes are the reference's
// (include/caelix_amd/layout.h). One generalisation: the home slot is key_lo mod slot_count, which is
// the reference's key_lo & mask when slot_count is a power of two, and lets a bulk-built table run at
// any load (the 0.7-load brick table); probe_limit is the window (kHashProbes for the reference form,
// the longest displacement + 1 for a bulk-built table).
#pragma once
#include "device_common.h"

namespace cx {

struct HashTableView
{
    HashKeySlot* keys;
    uint32_t* stack;  // int top, then uint entries[capacity]
    HashCounters* counters;
    uint8_t* payloads;
    uint32_t slot_count;
    uint32_t payload_stride;
    uint32_t probe_limit;
    uint32_t pow2;  // 1: slot_count is a power of two (home = key_lo & (slot_count - 1))
};

CX_DEV uint32_t HashSlotIndex(const HashTableView& t, uint32_t key_lo, uint32_t probe)
{
    if (t.pow2 != 0)
    {
        return (key_lo + probe) & (t.slot_count - 1);
    }
    const uint32_t home = key_lo % t.slot_count;
    const uint32_t s = home + probe;
    return s >= t.slot_count ? s - t.slot_count : s;
}

// HashSplitMix64 (caelix_hash_table.hlsl:103-109).
CX_DEV uint64_t SplitMix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// HashFingerprint (caelix_hash_table.hlsl:113-128): one splitmix64 round per word pair; key_lo >= 1
// and key_hi != 0 are forced.
CX_DEV uint2 Fingerprint(const uint32_t* words, uint32_t count)
{
    uint64_t h = 0;
    for (uint32_t i = 0; i < count; i += 2)
    {
        const uint64_t high = i + 1 < count ? words[i + 1] : 0u;
        h = SplitMix64(h ^ ((high << 32) | words[i]));
    }
    const uint32_t lo = uint32_t(h);
    const uint32_t hi = uint32_t(h >> 32);
    return make_uint2(lo > 1u ? lo : 1u, hi != 0u ? hi : 1u);
}

// The brick key fingerprint: the three signed brick coordinates as key words.
CX_DEV uint2 BrickFingerprint(int32_t x, int32_t y, int32_t z)
{
    const uint32_t words[3] = {uint32_t(x), uint32_t(y), uint32_t(z)};
    return Fingerprint(words, 3);
}

CX_DEV uint32_t HashLoad(const uint32_t* p)
{
    return LoadRelaxed(p);
}

CX_DEV void HashTouch(const HashTableView& t, uint32_t payload, uint32_t frame)
{
    uint32_t* touched = reinterpret_cast<uint32_t*>(t.payloads + uint64_t(payload) * t.payload_stride + kHashPayloadTouched);
    if (*touched != frame)
    {
        *touched = frame;
    }
}

// HashClaim (caelix_hash_table.hlsl:143-167): pops a payload for a slot this thread claimed, writes
// the payload header, the slot's payload word, then publishes key_hi (release).
CX_DEV uint32_t HashClaim(const HashTableView& t, HashKeySlot* slot, uint2 key, const uint32_t* words,
                          uint32_t count, uint32_t frame, uint32_t& payload)
{
    const int32_t index = int32_t(__hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(t.stack) + kHashStackTop / 4,
                                                         0xFFFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) - 1;
    payload = index >= 0 ? t.stack[kHashStackEntries / 4 + index] : kHashNone;
    if (payload != kHashNone)
    {
        uint32_t* header = reinterpret_cast<uint32_t*>(t.payloads + uint64_t(payload) * t.payload_stride);
        StoreRelaxed(&header[kHashPayloadState / 4], kHashStateUninitialised);
        header[kHashPayloadTouched / 4] = frame;
        for (uint32_t w = 0; w < count; ++w)
        {
            header[kHashPayloadKey / 4 + w] = words[w];
        }
    }
    StoreRelaxed(&slot->payload, payload);
    StoreRelease(&slot->key_hi, key.y);
    return payload != kHashNone ? kHashInserted : kHashNoPayload;
}

// HashFind (caelix_hash_table.hlsl:171-214).
CX_DEV uint32_t HashFind(const HashTableView& t, uint2 key, const uint32_t* words, uint32_t count, uint32_t frame,
                         bool insert, uint32_t& payload)
{
    payload = kHashNone;
    for (uint32_t probe = 0; probe < t.probe_limit; ++probe)
    {
        HashKeySlot* slot = t.keys + HashSlotIndex(t, key.x, probe);
        uint32_t lo = LoadRelaxed(&slot->key_lo);
        if (lo == 0u)
        {
            if (!insert)
            {
                return kHashMissing;
            }
            uint32_t expected = 0u;
            __hip_atomic_compare_exchange_strong(&slot->key_lo, &expected, key.x, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            lo = expected;
            if (lo == 0u)
            {
                return HashClaim(t, slot, key, words, count, frame, payload);
            }
        }
        if (lo == key.x)
        {
            const uint32_t hi = LoadAcquire(&slot->key_hi);
            if (hi == 0u)
            {
                return kHashPending;
            }
            if (hi == key.y)
            {
                payload = LoadRelaxed(&slot->payload);
                if (payload == kHashNone)
                {
                    return kHashEmptyEntry;
                }
                HashTouch(t, payload, frame);
                return kHashHit;
            }
        }
    }
    return insert ? kHashFullWindow : kHashMissing;
}

// A read-only lookup of a bulk-built table (no touch, no pending states): the payload or kHashNone.
// Probes until the key, an empty slot or the probe limit; `probes_out` = the slots that sequential
// probing examines. The slots are read four at a time (independent loads of consecutive slots,
// usually one or two 64-byte lines), so the typical lookup at load 0.7 (about 2.2 probes) costs one
// memory round trip; the home slot is computed once (one integer division for a table whose size
// is not a power of two).
CX_DEV uint32_t HashGet(const HashKeySlot* keys, uint32_t slot_count, uint32_t pow2, uint32_t probe_limit, uint2 key,
                        uint32_t* probes_out = nullptr)
{
    const uint32_t home = pow2 != 0 ? (key.x & (slot_count - 1)) : (key.x % slot_count);
    for (uint32_t base = 0; base < probe_limit; base += 4)
    {
        uint4 s[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j)
        {
            // home + base + j < 2 * slot_count + 3: two subtractions wrap it when slot_count >= 4.
            uint32_t index = home + base + j;
            if (pow2 != 0)
            {
                index &= slot_count - 1;
            }
            else if (slot_count >= 4)
            {
                index = index >= slot_count ? index - slot_count : index;
                index = index >= slot_count ? index - slot_count : index;
            }
            else
            {
                index %= slot_count;
            }
            s[j] = *reinterpret_cast<const uint4*>(keys + index);
        }
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j)
        {
            const uint32_t probe = base + j;
            if (probe >= probe_limit)
            {
                break;
            }
            if (s[j].x == 0u)
            {
                if (probes_out != nullptr)
                {
                    *probes_out = probe + 1;
                }
                return kHashNone;
            }
            if (s[j].x == key.x && s[j].y == key.y)
            {
                if (probes_out != nullptr)
                {
                    *probes_out = probe + 1;
                }
                return s[j].z;
            }
        }
    }
    if (probes_out != nullptr)
    {
        *probes_out = probe_limit;
    }
    return kHashNone;
}

}  // namespace cx
