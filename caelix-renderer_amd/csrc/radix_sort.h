// Onesweep LSD radix sort of (uint32 key, uint32 value) pairs for gfx950: 8-bit digits, 4 passes.
// One upfront kernel builds the four digit histograms (folded into the light BVH's Morton kernel),
// then each pass ranks a 4096-element tile with wave64 ballots, finds the tile's global digit
// offsets by decoupled look-back over earlier tiles, stages the tile in LDS in digit order and
// writes it out coalesced. Stable: equal digits keep their input order.
#pragma once
#include "device_common.h"

namespace cx {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096 elements
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kSortRadix = 256;
constexpr int kSortPasses = 4;
constexpr int kSortSlices = 8;  // histogram copies, one per XCD (blockIdx % 8)

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
