// Device helpers shared by the gfx950 kernels: wave64 primitives, agent-scope atomics, block scans.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "caelix_amd/layout.h"

#define CX_DEV __device__ __forceinline__

namespace cx {

constexpr int kWave = 64;

CX_DEV uint32_t Lane() { return __lane_id(); }

CX_DEV uint64_t Ballot(bool p) { return __ballot(p); }

CX_DEV uint64_t LanesBelow() { return __lanemask_lt(); }

CX_DEV uint32_t PopC(uint64_t m) { return static_cast<uint32_t>(__popcll(m)); }

// ---- agent-scope atomics (device-wide coherence across the 8 XCD L2s) ---------------------------
CX_DEV uint32_t AtomicAddAcqRel(uint32_t* p, uint32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
CX_DEV uint32_t AtomicAddRelaxed(uint32_t* p, uint32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
CX_DEV int32_t AtomicAddRelaxed(int32_t* p, int32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
CX_DEV uint32_t LoadRelaxed(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
CX_DEV uint64_t LoadRelaxed(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
CX_DEV void StoreRelaxed(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
CX_DEV void StoreRelaxed(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
CX_DEV uint32_t LoadAcquire(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
CX_DEV void StoreRelease(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- wave scans (DPP through __shfl_up; wave64) ---------------------------------------------------
template <typename T>
CX_DEV T WaveInclusiveSum(T v)
{
    const uint32_t lane = Lane();
#pragma unroll
    for (int offset = 1; offset < kWave; offset <<= 1)
    {
        const T up = __shfl_up(v, offset, kWave);
        if (lane >= static_cast<uint32_t>(offset))
        {
            v += up;
        }
    }
    return v;
}

template <typename T>
CX_DEV T WaveSum(T v)
{
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset >>= 1)
    {
        v += __shfl_xor(v, offset, kWave);
    }
    return v;
}

// Exclusive sum over a block of `kThreads` threads (one value per thread); returns the exclusive
// prefix, writes the block total. `lds` holds at least kThreads / 64 elements. Ends with a barrier.
template <int kThreads, typename T>
CX_DEV T BlockExclusiveSum(T v, T* lds, T& total)
{
    constexpr int kWaves = kThreads / kWave;
    const uint32_t lane = Lane();
    const uint32_t wave = threadIdx.x / kWave;
    const T inclusive = WaveInclusiveSum(v);
    if (lane == kWave - 1)
    {
        lds[wave] = inclusive;
    }
    __syncthreads();
    T before = 0;
    T sum = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
    {
        const T x = lds[w];
        before += (static_cast<uint32_t>(w) < wave) ? x : T(0);
        sum += x;
    }
    total = sum;
    __syncthreads();
    return before + inclusive - v;
}

// 30-bit Morton code of a cell with 10-bit coordinates (LtPart1By2 / LtMorton,
// caelix_light_tree_layout.hlsl:248-261).
CX_DEV uint32_t Part1By2(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

CX_DEV uint32_t Morton3(uint32_t x, uint32_t y, uint32_t z)
{
    return Part1By2(x) | (Part1By2(y) << 1) | (Part1By2(z) << 2);
}

}  // namespace cx
