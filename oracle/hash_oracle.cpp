// ORACLE — test infrastructure only. CPU restatement of the reference's hash table
// (shaders/common/caelix_hash_table.hlsl:103-128 fingerprint; :171-214 probing) and of the bulk
// brick table's layout as a sequential insert.
//
// Pinning:
//   fingerprint   the reference test's own CPU model (tests/hash_table_test.cpp:48-67) is the same
//                 SplitMix64 fold; tests also check a published SplitMix64 output (Vigna's reference
//                 sequence from seed 0) and the forced key_lo >= 1 / key_hi != 0 rules.
//   bulk build    tests check it against a Python dict-based linear-probing insert and the invariants
//                 (every key found within the probe limit, no key in an empty-terminated run).
#include "oracle_common.h"

#include <algorithm>
#include <numeric>
#include <vector>

namespace {

uint64_t SplitMix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

void FingerprintOf(const uint32_t* words, uint32_t count, uint32_t* lo_out, uint32_t* hi_out)
{
    uint64_t h = 0;
    for (uint32_t i = 0; i < count; i += 2)
    {
        const uint64_t high = i + 1 < count ? words[i + 1] : 0u;
        h = SplitMix64(h ^ ((high << 32) | words[i]));
    }
    const uint32_t lo = static_cast<uint32_t>(h);
    const uint32_t hi = static_cast<uint32_t>(h >> 32);
    *lo_out = std::max(lo, 1u);
    *hi_out = hi != 0 ? hi : 1u;
}

uint32_t Home(uint32_t key_lo, uint32_t slot_count)
{
    const bool pow2 = (slot_count & (slot_count - 1)) == 0;
    return pow2 ? (key_lo & (slot_count - 1)) : (key_lo % slot_count);
}

}  // namespace

extern "C" {

uint64_t oracle_splitmix64(uint64_t x)
{
    return SplitMix64(x);
}

void oracle_hash_fingerprint(const uint32_t* words, uint32_t word_count, uint32_t count, uint32_t* out)
{
    for (uint32_t i = 0; i < count; ++i)
    {
        FingerprintOf(words + uint64_t(i) * word_count, word_count, &out[2 * uint64_t(i)], &out[2 * uint64_t(i) + 1]);
    }
}

// Sequential insert of the brick keys in (home slot, key index) order with linear probing (wrapping
// at the table end): slot = the first empty slot from home. Returns the longest displacement.
uint32_t oracle_hash_build_bricks(const int32_t* keys_xyz, uint32_t count, HashKeySlot* slots, uint32_t slot_count)
{
    std::vector<uint32_t> lo(count), hi(count), home(count);
    for (uint32_t i = 0; i < count; ++i)
    {
        const uint32_t w[3] = {uint32_t(keys_xyz[3 * uint64_t(i)]), uint32_t(keys_xyz[3 * uint64_t(i) + 1]),
                               uint32_t(keys_xyz[3 * uint64_t(i) + 2])};
        FingerprintOf(w, 3, &lo[i], &hi[i]);
        home[i] = Home(lo[i], slot_count);
    }
    std::vector<uint32_t> order(count);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return home[a] < home[b]; });
    for (uint32_t s = 0; s < slot_count; ++s)
    {
        slots[s] = HashKeySlot{0, 0, 0, 0};
    }
    uint32_t longest = 0;
    for (uint32_t k : order)
    {
        uint32_t s = home[k];
        uint32_t d = 0;
        while (slots[s].key_lo != 0)
        {
            s = s + 1 == slot_count ? 0 : s + 1;
            ++d;
        }
        slots[s] = HashKeySlot{lo[k], hi[k], k, 0};
        longest = std::max(longest, d);
    }
    return longest;
}

// A lookup of a brick key in a bulk-built table: probes from home until the key, an empty slot or the
// probe limit; returns the payload or kHashNone and the probes read.
uint32_t oracle_hash_get_brick(const HashKeySlot* slots, uint32_t slot_count, uint32_t probe_limit, int32_t x,
                               int32_t y, int32_t z, uint32_t* probes)
{
    const uint32_t w[3] = {uint32_t(x), uint32_t(y), uint32_t(z)};
    uint32_t lo, hi;
    FingerprintOf(w, 3, &lo, &hi);
    uint32_t s = Home(lo, slot_count);
    for (uint32_t p = 0; p < probe_limit; ++p)
    {
        if (slots[s].key_lo == 0)
        {
            *probes = p + 1;
            return kHashNone;
        }
        if (slots[s].key_lo == lo && slots[s].key_hi == hi)
        {
            *probes = p + 1;
            return slots[s].payload;
        }
        s = s + 1 == slot_count ? 0 : s + 1;
    }
    *probes = probe_limit;
    return kHashNone;
}

}  // extern "C"
