// Onesweep radix sort kernels (gfx950): histogram, and one ranking + decoupled look-back + LDS-staged
// scatter kernel per digit (8-bit digits of (uint32 key, value) pairs; 9-bit digits of key-only
// uint64 keys). Replaces the reference's 8 x (histogram, single-workgroup scan,
// scatter) sequence of 4-bit passes (shaders/light/bvh/caelix_light_bvh_sort_{histogram,scatter}.hlsl,
// caelix_light_bvh_scan.hlsl); the result (a stable sort by key) is the same.
#include "radix_sort.h"
#include "runtime.h"

namespace cx {
namespace {


__global__ void __launch_bounds__(256) SortInitKernel(SortState s, uint32_t hist_words)
{
    for (uint32_t i = threadIdx.x; i < kSortCtlBytes / 4; i += blockDim.x)
    {
        s.ctl[i] = 0;
    }
    for (uint32_t i = threadIdx.x; i < hist_words; i += blockDim.x)
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

// The 64-bit key-only layout: histograms of its 7 digits, and one pass.
__global__ void __launch_bounds__(SortKeys64::kThreads) SortHistogram64Kernel(const uint64_t* __restrict__ keys,
                                                                             const uint32_t* count_ptr,
                                                                             const uint32_t* gate, SortState s)
{
    if (gate != nullptr && *gate == 0)
    {
        return;
    }
    constexpr uint32_t kWords = SortKeys64::kPasses * SortKeys64::kRadix;
    constexpr uint32_t kTile = SortKeys64::kThreads * SortKeys64::kItems;
    __shared__ uint32_t lds_hist[kWords];
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < kWords; i += blockDim.x)
    {
        lds_hist[i] = 0;
    }
    SortClearTileStatusOf<SortKeys64>(s, blockIdx.x);
    __syncthreads();
    const uint32_t n = *count_ptr;
    const uint32_t base = blockIdx.x * kTile;
#pragma unroll 4
    for (int k = 0; k < SortKeys64::kItems; ++k)
    {
        const uint32_t e = base + k * SortKeys64::kThreads + t;
        if (e < n)
        {
            SortCountKeyOf<SortKeys64>(lds_hist, keys[e]);
        }
    }
    __syncthreads();
    SortFlushHistogramOf<SortKeys64>(lds_hist, s.hist);
}

__global__ void __launch_bounds__(SortKeys64::kThreads, kSortWavesPerSimd) SortPass64Kernel(
    const uint64_t* __restrict__ keys_in, uint64_t* __restrict__ keys_out, const uint32_t* count_ptr,
    const uint32_t* gate, SortState s, uint32_t pass)
{
    const SortSink<SortKeys64> out{keys_out, nullptr, nullptr, nullptr};
    SortPassTileOf<SortKeys64>(keys_in, nullptr, out, count_ptr, gate, s, pass);
}

__global__ void __launch_bounds__(kSortThreads, kSortWavesPerSimd) SortPassKernel(const uint32_t* __restrict__ keys_in,
                                                               const uint32_t* __restrict__ values_in,
                                                               uint32_t* __restrict__ keys_out,
                                                               uint32_t* __restrict__ values_out,
                                                               const uint32_t* count_ptr, const uint32_t* gate,
                                                               SortState s, uint32_t pass)
{
    SortPassTile(keys_in, values_in, keys_out, values_out, count_ptr, gate, s, pass);
}

}  // namespace

// The 64-bit key-only sort (7 passes: the result is in keys1).
void LaunchSortKeys64(uint64_t* keys0, uint64_t* keys1, const uint32_t* count, const uint32_t* gate,
                      const SortState& s, hipStream_t stream)
{
    {
        StageScope scope(stream, "sort_init");
        SortInitKernel<<<1, 256, 0, stream>>>(s, uint32_t(SortHistBytesOf<SortKeys64>() / 4));
    }
    {
        StageScope scope(stream, "sort64_histogram");
        SortHistogram64Kernel<<<s.tiles, SortKeys64::kThreads, 0, stream>>>(keys0, count, gate, s);
    }
    uint64_t* keys[2] = {keys0, keys1};
    for (uint32_t pass = 0; pass < uint32_t(SortKeys64::kPasses); ++pass)
    {
        StageScope scope(stream, "sort64_pass");
        const uint32_t in = pass & 1u;
        SortPass64Kernel<<<s.tiles, SortKeys64::kThreads, 0, stream>>>(keys[in], keys[1 - in], count, gate, s, pass);
    }
}

void LaunchSortInit(const SortState& s, hipStream_t stream)
{
    StageScope scope(stream, "sort_init");
    SortInitKernel<<<1, 256, 0, stream>>>(s, uint32_t(kSortHistBytes / 4));
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

CX_API uint64_t cx_radix_sort_keys64_state_bytes(uint32_t capacity)
{
    return cx::SortStateBytesOf<cx::SortKeys64>(capacity);
}

CX_API int cx_radix_sort_keys64(uint64_t* keys0, uint64_t* keys1, const uint32_t* count_address,
                                const uint32_t* gate_address, uint32_t capacity, void* sort_state, void* stream)
{
    if (keys0 == nullptr || keys1 == nullptr || count_address == nullptr || sort_state == nullptr)
    {
        return cx::ArgError("cx_radix_sort_keys64: a buffer is NULL");
    }
    if (capacity == 0)
    {
        return cx::ArgError("cx_radix_sort_keys64: capacity 0");
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cx::LaunchSortKeys64(keys0, keys1, count_address, gate_address, cx::SortStateOf<cx::SortKeys64>(sort_state, capacity),
                         s);
    return cx::LaunchResult("cx_radix_sort_keys64");
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
