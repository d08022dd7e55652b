// Onesweep LSD radix sort for gfx950, in two key layouts:
//   SortPairs32  (uint32 key, uint32 value) pairs, 8-bit digits, 4 passes, 8192-element tiles;
//   SortKeys64   uint64 keys alone, 9-bit digits over bits 0..62, 7 passes, 4096-element tiles.
// One upfront kernel builds the digit histograms of every pass (folded into the light BVH's Morton
// kernel), then each pass ranks a tile with wave64 ballots, finds the tile's global digit offsets by
// decoupled look-back over earlier tiles, stages the tile in LDS in digit order and writes it out
// coalesced. Stable: equal digits keep their input order.
#pragma once
#include "device_common.h"

namespace cx {

struct SortPairs32
{
    using Key = uint32_t;
    static constexpr int kDigitBits = 8;
    static constexpr int kRadix = 256;
    static constexpr int kPasses = 4;
    static constexpr int kThreads = 512;
    static constexpr int kItems = 16;
    static constexpr bool kValues = true;
};

struct SortKeys64
{
    using Key = uint64_t;
    static constexpr int kDigitBits = 9;
    static constexpr int kRadix = 512;
    static constexpr int kPasses = 7;
    static constexpr int kThreads = 512;
    static constexpr int kItems = 8;
    static constexpr bool kValues = false;
};

// The light BVH build's sorts of 30-bit Morton keys with leaf values: kBits-bit digits, as many
// passes as cover bits 0..29, tiles of 512 x kItemsPerLane elements. The build picks one at run time
// (cx_light_bvh_set_tuning); every layout gives the same stable order.
template <int kBits, int kItemsPerLane>
struct SortMorton
{
    using Key = uint32_t;
    static constexpr int kDigitBits = kBits;
    static constexpr int kRadix = 1 << kBits;
    static constexpr int kPasses = (30 + kBits - 1) / kBits;
    static constexpr int kThreads = 512;
    static constexpr int kItems = kItemsPerLane;
    static constexpr bool kValues = true;
};

// The (uint32 key, uint32 value) layout under its historical names.
constexpr int kSortThreads = SortPairs32::kThreads;
constexpr int kSortItems = SortPairs32::kItems;
constexpr int kSortTile = kSortThreads * kSortItems;  // 8192 elements
constexpr int kSortRadix = SortPairs32::kRadix;
constexpr int kSortPasses = SortPairs32::kPasses;
constexpr int kSortSlices = 8;        // histogram copies, one per XCD (blockIdx % 8)
constexpr int kSortLookBack = 4;      // predecessor tiles read per look-back round trip
constexpr int kSortWavesPerSimd = 4;  // two workgroups per CU: at most 128 VGPRs per lane

// Sort state: control words, kSortSlices x passes x radix histogram slices, passes x tiles x radix
// tile status words.
constexpr uint64_t kSortCtlBytes = 256;

template <class L>
__host__ __device__ inline uint32_t SortTilesOf(uint32_t capacity)
{
    constexpr uint32_t tile = L::kThreads * L::kItems;
    return (capacity + tile - 1) / tile;
}
template <class L>
__host__ __device__ inline uint64_t SortHistBytesOf()
{
    return uint64_t{kSortSlices} * L::kPasses * L::kRadix * 4;
}
template <class L>
__host__ __device__ inline uint64_t SortStateBytesOf(uint32_t capacity)
{
    return kSortCtlBytes + SortHistBytesOf<L>() + uint64_t{L::kPasses} * SortTilesOf<L>(capacity) * L::kRadix * 4;
}

constexpr uint64_t kSortHistBytes = uint64_t{kSortSlices} * kSortPasses * kSortRadix * 4;
__host__ __device__ inline uint32_t SortTiles(uint32_t capacity)
{
    return SortTilesOf<SortPairs32>(capacity);
}
__host__ __device__ inline uint64_t SortStateBytes(uint32_t capacity)
{
    return SortStateBytesOf<SortPairs32>(capacity);
}

// Tile status words of the look-back: flag bits 30..31, count bits 0..29.
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kFlagMask = 3u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1u;

struct SortState
{
    uint32_t* ctl;     // [0 .. passes) tile counters of the passes
    uint32_t* hist;    // [slice][pass][digit]
    uint32_t* status;  // [pass][tile][digit]: flag bits 30..31, count bits 0..29
    uint32_t tiles;
};

template <class L>
__host__ __device__ inline SortState SortStateOf(void* base, uint32_t capacity)
{
    uint8_t* b = static_cast<uint8_t*>(base);
    SortState s;
    s.ctl = reinterpret_cast<uint32_t*>(b);
    s.hist = reinterpret_cast<uint32_t*>(b + kSortCtlBytes);
    s.status = reinterpret_cast<uint32_t*>(b + kSortCtlBytes + SortHistBytesOf<L>());
    s.tiles = SortTilesOf<L>(capacity);
    return s;
}
__host__ __device__ inline SortState SortStateAt(void* base, uint32_t capacity)
{
    return SortStateOf<SortPairs32>(base, capacity);
}

template <class L>
CX_DEV uint32_t SortDigit(typename L::Key key, uint32_t pass)
{
    return uint32_t(key >> (L::kDigitBits * pass)) & uint32_t(L::kRadix - 1);
}

// Device part used by kernels that produce sort keys (the Morton kernel): accumulates the digit
// histograms of every pass of the calling block's keys into LDS (`lds_hist`, passes x radix, zeroed
// by the caller), then SortFlushHistogram adds them to the block's slice.
template <class L>
CX_DEV void SortCountKeyOf(uint32_t* lds_hist, typename L::Key key)
{
#pragma unroll
    for (int p = 0; p < L::kPasses; ++p)
    {
        atomicAdd(&lds_hist[p * L::kRadix + SortDigit<L>(key, p)], 1u);
    }
}
CX_DEV void SortCountKey(uint32_t* lds_hist, uint32_t key)
{
    SortCountKeyOf<SortPairs32>(lds_hist, key);
}

// After a barrier: the block's threads add the LDS histograms to slice blockIdx % 8.
template <class L>
CX_DEV void SortFlushHistogramOf(const uint32_t* lds_hist, uint32_t* hist)
{
    uint32_t* slice = hist + (blockIdx.x % kSortSlices) * L::kPasses * L::kRadix;
    for (uint32_t i = threadIdx.x; i < uint32_t(L::kPasses * L::kRadix); i += blockDim.x)
    {
        const uint32_t v = lds_hist[i];
        if (v != 0)
        {
            AtomicAddRelaxed(&slice[i], v);
        }
    }
}
CX_DEV void SortFlushHistogram(const uint32_t* lds_hist, uint32_t* hist)
{
    SortFlushHistogramOf<SortPairs32>(lds_hist, hist);
}

// Zeroes the tile status rows of tile `tile` for every pass (called by the kernel that owns that
// tile before the passes run).
template <class L>
CX_DEV void SortClearTileStatusOf(const SortState& s, uint32_t tile)
{
    if (tile < s.tiles)
    {
        for (uint32_t i = threadIdx.x; i < uint32_t(L::kPasses * L::kRadix); i += blockDim.x)
        {
            const uint32_t p = i / L::kRadix;
            s.status[(uint64_t(p) * s.tiles + tile) * L::kRadix + i % L::kRadix] = 0;
        }
    }
}
CX_DEV void SortClearTileStatus(const SortState& s, uint32_t tile)
{
    SortClearTileStatusOf<SortPairs32>(s, tile);
}

// Where a pass writes: pairs to (keys, values), 64-bit keys to `keys`, or, when `keys` is null (the
// last pass of the light-tree build's 64-bit sort), split into the high word (the Morton code) at
// `high` and the low word (the leaf) at `low`.
template <class L>
struct SortSink
{
    typename L::Key* keys;
    uint32_t* values;
    uint32_t* high;
    uint32_t* low;
    CX_DEV void operator()(uint32_t destination, typename L::Key key, uint32_t value) const
    {
        if constexpr (L::kValues)
        {
            keys[destination] = key;
            values[destination] = value;
        }
        else if (keys != nullptr)
        {
            keys[destination] = key;
        }
        else
        {
            high[destination] = uint32_t(uint64_t(key) >> 32);
            low[destination] = uint32_t(key);
        }
    }
};

// One onesweep pass. Element order inside a tile: wave w owns elements [w * 64k, (w + 1) * 64k) for
// k = items per lane; item j of lane l is element w * 64k + j * 64 + l. Ranking walks the items in
// order, so (wave, item, lane) is the input order and equal digits stay stable. The body of a pass
// for one workgroup of L::kThreads threads (kernels of one sort, or of one sort per scene of a batch,
// call it). `values_in` is ignored for key-only layouts.
template <class L>
CX_DEV void SortPassTileOf(const typename L::Key* __restrict__ keys_in, const uint32_t* __restrict__ values_in,
                           const SortSink<L>& out, const uint32_t* count_ptr, const uint32_t* gate,
                           const SortState& s, uint32_t pass)
{
    using Key = typename L::Key;
    constexpr int kThreads = L::kThreads;
    constexpr int kItems = L::kItems;
    constexpr int kTile = kThreads * kItems;
    constexpr int kWaves = kThreads / kWave;
    constexpr int kRadix = L::kRadix;
    static_assert(kRadix <= kThreads, "every digit has an owner thread");
    if (gate != nullptr && *gate == 0)
    {
        return;
    }
    __shared__ Key lds_keys[kTile];
    __shared__ uint32_t lds_values[L::kValues ? kTile : 1];
    __shared__ uint32_t lds_wave_count[kWaves][kRadix];  // per wave and digit: counts, then offsets
    __shared__ uint32_t lds_digit_start[kRadix];         // tile-local start of each digit
    __shared__ uint32_t lds_destination[kRadix];         // global index of each digit's first element
    __shared__ uint32_t lds_scan[kWaves];
    __shared__ uint32_t lds_tile;

    const uint32_t t = threadIdx.x;
    const uint32_t lane = Lane();
    const uint32_t wave = t / kWave;
    if (t == 0)
    {
        lds_tile = AtomicAddRelaxed(&s.ctl[pass], 1u);
    }
    for (uint32_t i = t; i < uint32_t(kWaves * kRadix); i += kThreads)
    {
        (&lds_wave_count[0][0])[i] = 0;
    }
    __syncthreads();
    const uint32_t tile = lds_tile;
    const uint32_t n = *count_ptr;
    const uint32_t tile_start = tile * kTile;
    if (tile_start >= n)
    {
        return;  // uniform; no later tile has elements either
    }

    Key key[kItems];
    uint32_t value[kItems];
    uint32_t offset[kItems];
    const uint32_t wave_start = tile_start + wave * (kItems * kWave);
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        const uint32_t e = wave_start + k * kWave + lane;
        const bool valid = e < n;
        key[k] = valid ? keys_in[e] : Key(0);
        if constexpr (L::kValues)
        {
            value[k] = valid ? values_in[e] : 0u;
        }
        else
        {
            value[k] = 0u;
        }
    }
    // Rank: peers with the same digit in this item round, from one ballot per digit bit; the
    // lowest peer updates the wave's running count of the digit.
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        const bool valid = wave_start + k * kWave + lane < n;
        const uint32_t digit = SortDigit<L>(key[k], pass);
        uint64_t peers = Ballot(valid);
#pragma unroll
        for (int b = 0; b < L::kDigitBits; ++b)
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

    // Thread t < radix owns digit t: wave offsets (in place: only the owner reads its column), tile
    // count, look-back.
    const bool owner = t < uint32_t(kRadix);
    uint32_t count = 0;
    if (owner)
    {
#pragma unroll
        for (int w = 0; w < kWaves; ++w)
        {
            const uint32_t c = lds_wave_count[w][t];
            lds_wave_count[w][t] = count;
            count += c;
        }
    }
    uint32_t* status = s.status + (uint64_t(pass) * s.tiles) * kRadix;
    if (owner)
    {
        StoreRelaxed(&status[uint64_t(tile) * kRadix + t], (tile == 0 ? kFlagPrefix : kFlagAggregate) | count);
    }
    // Digit bases of this pass: the exclusive scan over digits of the summed histogram slices.
    uint32_t total_d = 0;
    if (owner)
    {
#pragma unroll
        for (int sl = 0; sl < kSortSlices; ++sl)
        {
            total_d += s.hist[(sl * L::kPasses + pass) * kRadix + t];
        }
    }
    uint32_t unused;
    const uint32_t digit_base = BlockExclusiveSum<kThreads>(total_d, lds_scan, unused);
    const uint32_t digit_start = BlockExclusiveSum<kThreads>(owner ? count : 0u, lds_scan, unused);
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
                word[j] = p - j >= 0 ? LoadRelaxed(&status[uint64_t(p - j) * kRadix + t]) : kFlagPrefix;
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
        StoreRelaxed(&status[uint64_t(tile) * kRadix + t], kFlagPrefix | (prefix + count));
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
    for (int k = 0; k < kItems; ++k)
    {
        if (wave_start + k * kWave + lane < n)
        {
            const uint32_t digit = SortDigit<L>(key[k], pass);
            const uint32_t slot = lds_digit_start[digit] + lds_wave_count[wave][digit] + offset[k];
            lds_keys[slot] = key[k];
            if constexpr (L::kValues)
            {
                lds_values[slot] = value[k];
            }
        }
    }
    __syncthreads();
    const uint32_t tile_count = min(uint32_t(kTile), n - tile_start);
#pragma unroll 4
    for (int k = 0; k < kItems; ++k)
    {
        const uint32_t j = k * kThreads + t;
        if (j < tile_count)
        {
            const Key kv = lds_keys[j];
            const uint32_t digit = SortDigit<L>(kv, pass);
            const uint32_t destination = lds_destination[digit] + (j - lds_digit_start[digit]);
            out(destination, kv, L::kValues ? lds_values[j] : 0u);
        }
    }
}

// The (uint32 key, uint32 value) pass.
CX_DEV void SortPassTile(const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ values_in,
                         uint32_t* __restrict__ keys_out, uint32_t* __restrict__ values_out, const uint32_t* count_ptr,
                         const uint32_t* gate, const SortState& s, uint32_t pass)
{
    const SortSink<SortPairs32> out{keys_out, values_out, nullptr, nullptr};
    SortPassTileOf<SortPairs32>(keys_in, values_in, out, count_ptr, gate, s, pass);
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
