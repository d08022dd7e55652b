// Host runtime: log callback, error translation, stage timer (HIP events on the launch stream).
#include "runtime.h"

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace cx {
namespace {

std::mutex g_log_mutex;
CxLogCallback g_log_callback = nullptr;
void* g_log_user = nullptr;

struct PendingStage
{
    const char* name;
    hipEvent_t start;
    hipEvent_t end;
};

struct StageTotal
{
    const char* name;
    double ms;
    uint32_t launches;
};

std::mutex g_profile_mutex;
bool g_profile_enabled = false;
std::vector<PendingStage> g_pending;
std::vector<StageTotal> g_totals;
// The stage opened by ProfileBegin on this thread, closed by ProfileEnd.
thread_local PendingStage t_open{nullptr, nullptr, nullptr};

void Resolve()
{
    for (PendingStage& p : g_pending)
    {
        float ms = 0.0f;
        if (hipEventSynchronize(p.end) == hipSuccess && hipEventElapsedTime(&ms, p.start, p.end) == hipSuccess)
        {
            bool found = false;
            for (StageTotal& t : g_totals)
            {
                if (std::strcmp(t.name, p.name) == 0)
                {
                    t.ms += ms;
                    ++t.launches;
                    found = true;
                    break;
                }
            }
            if (!found)
            {
                g_totals.push_back(StageTotal{p.name, double(ms), 1});
            }
        }
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.end);
    }
    g_pending.clear();
}

}  // namespace

void Log(int level, const char* format, ...)
{
    char message[1024];
    va_list args;
    va_start(args, format);
    std::vsnprintf(message, sizeof(message), format, args);
    va_end(args);
    std::lock_guard<std::mutex> lock(g_log_mutex);
    if (g_log_callback != nullptr)
    {
        g_log_callback(level, message, g_log_user);
    }
}

int ArgError(const char* format, ...)
{
    char message[1024];
    va_list args;
    va_start(args, format);
    std::vsnprintf(message, sizeof(message), format, args);
    va_end(args);
    Log(3, "%s", message);
    return CX_ERROR_INVALID_ARGUMENT;
}

int HipResult(hipError_t error, const char* what)
{
    if (error == hipSuccess)
    {
        return CX_OK;
    }
    Log(3, "%s: %s", what, hipGetErrorString(error));
    return error == hipErrorOutOfMemory ? CX_ERROR_OUT_OF_MEMORY : CX_ERROR_HIP;
}

int LaunchResult(const char* what)
{
    return HipResult(hipGetLastError(), what);
}

bool ProfileEnabled()
{
    return g_profile_enabled;
}

void ProfileBegin(hipStream_t stream, const char* stage)
{
    t_open.name = stage;
    t_open.start = nullptr;
    t_open.end = nullptr;
    if (hipEventCreate(&t_open.start) != hipSuccess || hipEventCreate(&t_open.end) != hipSuccess)
    {
        t_open.name = nullptr;
        return;
    }
    (void)hipEventRecord(t_open.start, stream);
}

void ProfileEnd(hipStream_t stream)
{
    if (t_open.name == nullptr)
    {
        return;
    }
    (void)hipEventRecord(t_open.end, stream);
    std::lock_guard<std::mutex> lock(g_profile_mutex);
    g_pending.push_back(t_open);
    t_open.name = nullptr;
}

}  // namespace cx

extern "C" {

CX_API void cx_set_log_callback(CxLogCallback callback, void* user_data)
{
    std::lock_guard<std::mutex> lock(cx::g_log_mutex);
    cx::g_log_callback = callback;
    cx::g_log_user = user_data;
}

CX_API const char* cx_version_string(void)
{
    return "caelix-mi355x 0.1.0 (gfx950)";
}

CX_API void cx_profile_enable(int enabled)
{
    std::lock_guard<std::mutex> lock(cx::g_profile_mutex);
    cx::g_profile_enabled = enabled != 0;
}

CX_API void cx_profile_reset(void)
{
    std::lock_guard<std::mutex> lock(cx::g_profile_mutex);
    cx::Resolve();
    cx::g_totals.clear();
}

CX_API int cx_profile_read(uint32_t max_stages, const char** names, double* total_ms, uint32_t* launches,
                           uint32_t* stage_count)
{
    std::lock_guard<std::mutex> lock(cx::g_profile_mutex);
    cx::Resolve();
    const uint32_t count = static_cast<uint32_t>(cx::g_totals.size());
    for (uint32_t i = 0; i < count && i < max_stages; ++i)
    {
        if (names != nullptr)
        {
            names[i] = cx::g_totals[i].name;
        }
        if (total_ms != nullptr)
        {
            total_ms[i] = cx::g_totals[i].ms;
        }
        if (launches != nullptr)
        {
            launches[i] = cx::g_totals[i].launches;
        }
    }
    if (stage_count != nullptr)
    {
        *stage_count = count;
    }
    return CX_OK;
}

}  // extern "C"
