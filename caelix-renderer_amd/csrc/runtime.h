// Host runtime of the library: logging, argument errors, HIP error translation and the optional
// per-kernel stage timer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>

#include "caelix_amd/caelix_amd.h"

namespace cx {

void Log(int level, const char* format, ...) __attribute__((format(printf, 2, 3)));

// Logs an argument error and returns CX_ERROR_INVALID_ARGUMENT.
int ArgError(const char* format, ...) __attribute__((format(printf, 1, 2)));

// CX_OK, or logs the HIP error of `what` and returns CX_ERROR_HIP (CX_ERROR_OUT_OF_MEMORY for an
// allocation failure).
int HipResult(hipError_t error, const char* what);

// The result of the launches since the last check (hipGetLastError).
int LaunchResult(const char* what);

// Stage timing: Begin/End bracket one kernel launch on `stream` when timing is enabled.
bool ProfileEnabled();
void ProfileBegin(hipStream_t stream, const char* stage);
void ProfileEnd(hipStream_t stream);

// RAII bracket of one stage (`timed` false: never timed, for callers that must not create events).
struct StageScope
{
    hipStream_t stream;
    bool on;
    StageScope(hipStream_t s, const char* stage, bool timed = true) : stream(s), on(timed && ProfileEnabled())
    {
        if (on)
        {
            ProfileBegin(stream, stage);
        }
    }
    ~StageScope()
    {
        if (on)
        {
            ProfileEnd(stream);
        }
    }
};

__host__ __device__ inline uint32_t DivUp(uint64_t a, uint64_t b)
{
    return static_cast<uint32_t>((a + b - 1) / b);
}

}  // namespace cx
