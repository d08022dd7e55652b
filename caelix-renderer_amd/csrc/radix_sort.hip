// Onesweep radix sort kernels (gfx950): histogram, and one ranking + decoupled look-back + LDS-staged
// scatter kernel per 8-bit digit. Replaces the reference's 8 x (histogram, single-workgroup scan,
// scatter) sequence of 4-bit passes (shaders/light/bvh/caelix_light_bvh_sort_{histogram,scatter}.hlsl,
// caelix_light_bvh_scan.hlsl); the result (a stable sort by key) is the same.
#include "radix_sort.h"
#include "runtime.h"

namespace cx {
namespace {

constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kFlagMask = 3u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1u;

__global__ void __launch_bounds__(256) SortInitKernel(SortState s)
{
    for (uint32_t i = threadIdx.x; i < kSortCtlBytes / 4; i += blockDim.x)
    {
        s.ctl[i] = 0;
    }
    for (uint32_t i = threadIdx.x; i < kSortHistBytes / 4; i += blockDim.x)
    {
        s.hist[i] = 0;
    }
}

__global__ void __launch_bounds__(kSortThreads) SortHistogramKernel(const uint32_t* __restrict__ keys,
                                                                    const uint32_t* count_ptr,
                                                                    const uint32_t* gate, SortState s)
{
    if (gate != nullptr && *gate == 0)
    {
        return;
    }
    __shared__ uint32_t lds_hist[kSortPasses * kSortRadix];
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < kSortPasses * kSortRadix; i += kSortThreads)
    {
        lds_hist[i] = 0;
    }
    SortClearTileStatus(s, blockIdx.x);
    __syncthreads();
    const uint32_t n = *count_ptr;
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll 4
    for (int k = 0; k < kSortItems; ++k)
    {
        const uint32_t e = base + k * kSortThreads + t;
        if (e < n)
        {
            SortCountKey(lds_hist, keys[e]);
        }
    }
    __syncthreads();
    SortFlushHistogram(lds_hist, s.hist);
}

// One onesweep pass over 4096-element tiles. Element order inside a tile: wave w owns elements
// [w * 1024, (w + 1) * 1024); item k of lane l is element w * 1024 + k * 64 + l. Ranking walks the
// items in order, so (wave, item, lane) is the input order and equal digits stay stable.
__global__ void __launch_bounds__(kSortThreads) SortPassKernel(const uint32_t* __restrict__ keys_in,
                                                               const uint32_t* __restrict__ values_in,
                                                               uint32_t* __restrict__ keys_out,
                                                               uint32_t* __restrict__ values_out,
                                                               const uint32_t* count_ptr, const uint32_t* gate,
                                                               SortState s, uint32_t pass)
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

    // Thread t owns digit t: wave offsets, tile count, look-back.
    const uint32_t c0 = lds_wave_count[0][t];
    const uint32_t c1 = lds_wave_count[1][t];
    const uint32_t c2 = lds_wave_count[2][t];
    const uint32_t c3 = lds_wave_count[3][t];
    const uint32_t count = c0 + c1 + c2 + c3;
    uint32_t* status = s.status + (uint64_t(pass) * s.tiles) * kSortRadix;
    StoreRelaxed(&status[uint64_t(tile) * kSortRadix + t], (tile == 0 ? kFlagPrefix : kFlagAggregate) | count);
    // Digit bases of this pass: the exclusive scan over digits of the summed histogram slices.
    uint32_t total_d = 0;
#pragma unroll
    for (int sl = 0; sl < kSortSlices; ++sl)
    {
        total_d += s.hist[(sl * kSortPasses + pass) * kSortRadix + t];
    }
    uint32_t unused;
    const uint32_t digit_base = BlockExclusiveSum<kSortThreads>(total_d, lds_scan, unused);
    const uint32_t digit_start = BlockExclusiveSum<kSortThreads>(count, lds_scan, unused);
    uint32_t prefix = 0;
    if (tile > 0)
    {
        int64_t p = int64_t(tile) - 1;
        while (true)
        {
            const uint32_t word = LoadRelaxed(&status[uint64_t(p) * kSortRadix + t]);
            const uint32_t flag = word & kFlagMask;
            if (flag == 0)
            {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            prefix += word & kCountMask;
            if (flag == kFlagPrefix)
            {
                break;
            }
            --p;
        }
        StoreRelaxed(&status[uint64_t(tile) * kSortRadix + t], kFlagPrefix | (prefix + count));
    }
    lds_wave_count[0][t] = 0;
    lds_wave_count[1][t] = c0;
    lds_wave_count[2][t] = c0 + c1;
    lds_wave_count[3][t] = c0 + c1 + c2;
    lds_digit_start[t] = digit_start;
    lds_destination[t] = digit_base + prefix;
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

}  // namespace

void LaunchSortInit(const SortState& s, hipStream_t stream)
{
    StageScope scope(stream, "sort_init");
    SortInitKernel<<<1, 256, 0, stream>>>(s);
}

void LaunchSortHistogram(const uint32_t* keys, const uint32_t* count, const uint32_t* gate, const SortState& s,
                         hipStream_t stream)
{
    StageScope scope(stream, "sort_histogram");
    SortHistogramKernel<<<s.tiles, kSortThreads, 0, stream>>>(keys, count, gate, s);
}

void LaunchSortPass(const uint32_t* keys_in, const uint32_t* values_in, uint32_t* keys_out, uint32_t* values_out,
                    const uint32_t* count, const uint32_t* gate, const SortState& s, uint32_t pass,
                    hipStream_t stream)
{
    StageScope scope(stream, "sort_pass");
    SortPassKernel<<<s.tiles, kSortThreads, 0, stream>>>(keys_in, values_in, keys_out, values_out, count, gate, s,
                                                         pass);
}

}  // namespace cx

extern "C" {

CX_API uint64_t cx_radix_sort_state_bytes(uint32_t capacity)
{
    return cx::SortStateBytes(capacity);
}

CX_API int cx_radix_sort_pairs(uint32_t* keys0, uint32_t* values0, uint32_t* keys1, uint32_t* values1,
                               const uint32_t* count_address, const uint32_t* gate_address, uint32_t capacity,
                               void* sort_state, void* stream)
{
    if (keys0 == nullptr || values0 == nullptr || keys1 == nullptr || values1 == nullptr ||
        count_address == nullptr || sort_state == nullptr)
    {
        return cx::ArgError("cx_radix_sort_pairs: a buffer is NULL");
    }
    if (capacity == 0)
    {
        return cx::ArgError("cx_radix_sort_pairs: capacity 0");
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const cx::SortState state = cx::SortStateAt(sort_state, capacity);
    cx::LaunchSortInit(state, s);
    cx::LaunchSortHistogram(keys0, count_address, gate_address, state, s);
    uint32_t* keys[2] = {keys0, keys1};
    uint32_t* values[2] = {values0, values1};
    for (uint32_t pass = 0; pass < cx::kSortPasses; ++pass)
    {
        const uint32_t in = pass & 1u;
        cx::LaunchSortPass(keys[in], values[in], keys[1 - in], values[1 - in], count_address, gate_address, state,
                           pass, s);
    }
    return cx::LaunchResult("cx_radix_sort_pairs");
}

}  // extern "C"
