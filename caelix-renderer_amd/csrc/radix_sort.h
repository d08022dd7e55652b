// Onesweep LSD radix sort of (uint32 key, uint32 value) pairs for gfx950: 8-bit digits, 4 passes.
// One upfront kernel builds the four digit histograms (folded into the light BVH's Morton kernel),
// then each pass ranks an 8192-element tile with wave64 ballots, finds the tile's global digit
// offsets by decoupled look-back over earlier tiles, stages the tile in LDS in digit order and
// writes it out coalesced. Stable: equal digits keep their input order.
#pragma once
#include "device_common.h"

namespace cx {

constexpr int kSortThreads = 512;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // 8192 elements
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kSortRadix = 256;
constexpr int kSortPasses = 4;
constexpr int kSortSlices = 8;  // histogram copies, one per XCD (blockIdx % 8)
constexpr int kSortLookBack = 4;  // predecessor tiles read per look-back round trip
constexpr int kSortWavesPerSimd = 4;  // two workgroups per CU: at most 128 VGPRs per lane

// Sort state bytes: control words, 8 x 4 x 256 histogram slices, 4 x tiles x 256 tile status.
constexpr uint64_t kSortCtlBytes = 256;
constexpr uint64_t kSortHistBytes = uint64_t{kSortSlices} * kSortPasses * kSortRadix * 4;
__host__ __device__ inline uint32_t SortTiles(uint32_t capacity)
{
    return (capacity + kSortTile - 1) / kSortTile;
}
__host__ __device__ inline uint64_t SortStateBytes(uint32_t capacity)
{
    return kSortCtlBytes + kSortHistBytes + uint64_t{kSortPasses} * SortTiles(capacity) * kSortRadix * 4;
}

// Tile status words of the look-back: flag bits 30..31, count bits 0..29.
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kFlagMask = 3u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1u;

struct SortState
{
    uint32_t* ctl;     // [0..3] tile counters of the 4 passes
    uint32_t* hist;    // [slice][pass][digit]
    uint32_t* status;  // [pass][tile][digit]: flag bits 30..31, count bits 0..29
    uint32_t tiles;
};

__host__ __device__ inline SortState SortStateAt(void* base, uint32_t capacity)
{
    uint8_t* b = static_cast<uint8_t*>(base);
    SortState s;
    s.ctl = reinterpret_cast<uint32_t*>(b);
    s.hist = reinterpret_cast<uint32_t*>(b + kSortCtlBytes);
    s.status = reinterpret_cast<uint32_t*>(b + kSortCtlBytes + kSortHistBytes);
    s.tiles = SortTiles(capacity);
    return s;
}

// Device part used by kernels that produce sort keys (the Morton kernel): accumulates the four
// 8-bit digit histograms of the calling block's keys into LDS (`lds_hist`, 4 x 256, zeroed by the
// caller), then SortFlushHistogram adds them to the block's slice.
CX_DEV void SortCountKey(uint32_t* lds_hist, uint32_t key)
{
#pragma unroll
    for (int p = 0; p < kSortPasses; ++p)
    {
        atomicAdd(&lds_hist[p * kSortRadix + ((key >> (8 * p)) & 0xFFu)], 1u);
    }
}

// After a barrier: thread t < 256 adds digit t of the four LDS histograms to slice blockIdx % 8.
CX_DEV void SortFlushHistogram(const uint32_t* lds_hist, uint32_t* hist)
{
    const uint32_t t = threadIdx.x;
    if (t < kSortRadix)
    {
        uint32_t* slice = hist + (blockIdx.x % kSortSlices) * kSortPasses * kSortRadix;
#pragma unroll
        for (int p = 0; p < kSortPasses; ++p)
        {
            const uint32_t v = lds_hist[p * kSortRadix + t];
            if (v != 0)
            {
                AtomicAddRelaxed(&slice[p * kSortRadix + t], v);
            }
        }
    }
}

// Zeroes the tile status rows of tile `tile` for the four passes (called by the kernel that owns
// that tile before the passes run).
CX_DEV void SortClearTileStatus(const SortState& s, uint32_t tile)
{
    const uint32_t t = threadIdx.x;
    if (t < kSortRadix && tile < s.tiles)
    {
#pragma unroll
        for (int p = 0; p < kSortPasses; ++p)
        {
            s.status[(uint64_t(p) * s.tiles + tile) * kSortRadix + t] = 0;
        }
    }
}

// One onesweep pass over 8192-element tiles. Element order inside a tile: wave w owns elements
// [w * 1024, (w + 1) * 1024); item k of lane l is element w * 1024 + k * 64 + l. Ranking walks the
// items in order, so (wave, item, lane) is the input order and equal digits stay stable.
// The body of a pass for one workgroup of kSortThreads threads (kernels of one sort, or of one
// sort per scene of a batch, call it).
CX_DEV void SortPassTile(const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ values_in,
                         uint32_t* __restrict__ keys_out, uint32_t* __restrict__ values_out, const uint32_t* count_ptr,
                         const uint32_t* gate, const SortState& s, uint32_t pass)
{
    if (gate != nullptr && *gate == 0)
    {
        return;
    }
    __shared__ uint32_t lds_keys[kSortTile];
    __shared__ uint32_t lds_values[kSortTile];
    __shared__ uint32_t lds_wave_count[kSortWaves][kSortRadix];  // per wave and digit: counts, then offsets
    __shared__ uint32_t lds_digit_start[kSortRadix];             // tile-local start of each digit
    __shared__ uint32_t lds_destination[kSortRadix];             // global index of each digit's first element
    __shared__ uint32_t lds_scan[kSortWaves];
    __shared__ uint32_t lds_tile;

    const uint32_t t = threadIdx.x;
    const uint32_t lane = Lane();
    const uint32_t wave = t / kWave;
    const uint32_t shift = 8 * pass;
    if (t == 0)
    {
        lds_tile = AtomicAddRelaxed(&s.ctl[pass], 1u);
    }
    for (uint32_t i = t; i < kSortWaves * kSortRadix; i += kSortThreads)
    {
        (&lds_wave_count[0][0])[i] = 0;
    }
    __syncthreads();
    const uint32_t tile = lds_tile;
    const uint32_t n = *count_ptr;
    const uint32_t tile_start = tile * kSortTile;
    if (tile_start >= n)
    {
        return;  // uniform; no later tile has elements either
    }

    uint32_t key[kSortItems];
    uint32_t value[kSortItems];
    uint32_t offset[kSortItems];
    const uint32_t wave_start = tile_start + wave * (kSortItems * kWave);
#pragma unroll
    for (int k = 0; k < kSortItems; ++k)
    {
        const uint32_t e = wave_start + k * kWave + lane;
        const bool valid = e < n;
        key[k] = valid ? keys_in[e] : 0u;
        value[k] = valid ? values_in[e] : 0u;
    }
    // Rank: peers with the same digit in this item round, from 8 ballots; the lowest peer updates
    // the wave's running count of the digit.
#pragma unroll
    for (int k = 0; k < kSortItems; ++k)
    {
        const bool valid = wave_start + k * kWave + lane < n;
        const uint32_t digit = (key[k] >> shift) & 0xFFu;
        uint64_t peers = Ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b)
        {
            const bool bit = ((digit >> b) & 1u) != 0;
            const uint64_t votes = Ballot(bit);
            peers &= bit ? votes : ~votes;
        }
        const uint32_t rank = PopC(peers & LanesBelow());
        const uint32_t leader = valid ? static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(peers)) - 1) : 0u;
        uint32_t before = 0;
        if (valid && lane == leader)
        {
            before = lds_wave_count[wave][digit];
            lds_wave_count[wave][digit] = before + PopC(peers);
        }
        before = __shfl(before, static_cast<int>(leader), kWave);
        offset[k] = before + rank;
    }
    __syncthreads();

    // Thread t < 256 owns digit t: wave offsets (in place: only the owner reads its column), tile
    // count, look-back.
    const bool owner = t < uint32_t(kSortRadix);
    uint32_t count = 0;
    if (owner)
    {
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w)
        {
            const uint32_t c = lds_wave_count[w][t];
            lds_wave_count[w][t] = count;
            count += c;
        }
    }
    uint32_t* status = s.status + (uint64_t(pass) * s.tiles) * kSortRadix;
    if (owner)
    {
        StoreRelaxed(&status[uint64_t(tile) * kSortRadix + t], (tile == 0 ? kFlagPrefix : kFlagAggregate) | count);
    }
    // Digit bases of this pass: the exclusive scan over digits of the summed histogram slices.
    uint32_t total_d = 0;
    if (owner)
    {
#pragma unroll
        for (int sl = 0; sl < kSortSlices; ++sl)
        {
            total_d += s.hist[(sl * kSortPasses + pass) * kSortRadix + t];
        }
    }
    uint32_t unused;
    const uint32_t digit_base = BlockExclusiveSum<kSortThreads>(total_d, lds_scan, unused);
    const uint32_t digit_start = BlockExclusiveSum<kSortThreads>(owner ? count : 0u, lds_scan, unused);
    // Look-back, kSortLookBack predecessors per round trip: aggregates are summed up to the first
    // prefix; a tile not yet published ends the round there.
    uint32_t prefix = 0;
    if (owner && tile > 0)
    {
        int64_t p = int64_t(tile) - 1;
        while (true)
        {
            uint32_t word[kSortLookBack];
#pragma unroll
            for (int j = 0; j < kSortLookBack; ++j)
            {
                word[j] = p - j >= 0 ? LoadRelaxed(&status[uint64_t(p - j) * kSortRadix + t]) : kFlagPrefix;
            }
            int used = kSortLookBack;
            bool finished = false;
#pragma unroll
            for (int j = kSortLookBack - 1; j >= 0; --j)
            {
                const uint32_t flag = word[j] & kFlagMask;
                if (flag == 0)
                {
                    used = j;
                    finished = false;
                }
                else if (flag == kFlagPrefix)
                {
                    used = j;
                    finished = true;
                }
            }
#pragma unroll
            for (int j = 0; j < kSortLookBack; ++j)
            {
                if (j < used || (finished && j == used))
                {
                    prefix += word[j] & kCountMask;
                }
            }
            if (finished)
            {
                break;
            }
            p -= used;
            if (used < kSortLookBack)
            {
                __builtin_amdgcn_s_sleep(1);
            }
        }
        StoreRelaxed(&status[uint64_t(tile) * kSortRadix + t], kFlagPrefix | (prefix + count));
    }
    if (owner)
    {
        lds_digit_start[t] = digit_start;
        lds_destination[t] = digit_base + prefix;
    }
    __syncthreads();

    // Stage the tile in LDS in digit order, then write it out: consecutive LDS slots of one digit
    // go to consecutive global slots.
#pragma unroll
    for (int k = 0; k < kSortItems; ++k)
    {
        if (wave_start + k * kWave + lane < n)
        {
            const uint32_t digit = (key[k] >> shift) & 0xFFu;
            const uint32_t slot = lds_digit_start[digit] + lds_wave_count[wave][digit] + offset[k];
            lds_keys[slot] = key[k];
            lds_values[slot] = value[k];
        }
    }
    __syncthreads();
    const uint32_t tile_count = min(uint32_t(kSortTile), n - tile_start);
#pragma unroll 4
    for (int k = 0; k < kSortItems; ++k)
    {
        const uint32_t j = k * kSortThreads + t;
        if (j < tile_count)
        {
            const uint32_t kv = lds_keys[j];
            const uint32_t digit = (kv >> shift) & 0xFFu;
            const uint32_t destination = lds_destination[digit] + (j - lds_digit_start[digit]);
            keys_out[destination] = kv;
            values_out[destination] = lds_values[j];
        }
    }
}


// Host launchers (radix_sort.hip).
// Zeroes the control words and histograms (1 workgroup).
void LaunchSortInit(const SortState& s, hipStream_t stream);
// Digit histograms of keys[0 .. *count) and clear of the tile status (one workgroup per tile).
void LaunchSortHistogram(const uint32_t* keys, const uint32_t* count, const uint32_t* gate, const SortState& s,
                         hipStream_t stream);
// One pass: digit = (key >> 8 * pass) & 255.
void LaunchSortPass(const uint32_t* keys_in, const uint32_t* values_in, uint32_t* keys_out, uint32_t* values_out,
                    const uint32_t* count, const uint32_t* gate, const SortState& s, uint32_t pass,
                    hipStream_t stream);

}  // namespace cx
