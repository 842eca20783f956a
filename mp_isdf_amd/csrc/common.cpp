#include "common.hpp"

namespace mpisdf {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& message) { g_last_error = message; }
const char* last_error() { return g_last_error.c_str(); }

void ensure_device()
{
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        throw NoDevice("no HIP device is available: mp_isdf_amd runs its selection path on an "
                       "MI355X (gfx950) GPU and has no CPU fallback");
    }
}

HipStream::HipStream() { MPISDF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }

HipStream::~HipStream()
{
    if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
}

HipEvent::HipEvent() { MPISDF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }

HipEvent::~HipEvent()
{
    if (e) (void)hipEventDestroy(e);
}

void HipEvent::record(hipStream_t s) { MPISDF_HIP(hipEventRecord(e, s)); }

void HipEvent::wait(hipStream_t s) const { MPISDF_HIP(hipStreamWaitEvent(s, e, 0)); }

void LaunchTimer::begin(hipStream_t s)
{
    hipEvent_t e0, e1;
    MPISDF_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        throw std::runtime_error("hipEventCreate failed");
    }
    events.emplace_back(e0, e1);
    MPISDF_HIP(hipEventRecord(e0, s));
}

void LaunchTimer::end(hipStream_t s) { MPISDF_HIP(hipEventRecord(events.back().second, s)); }

double LaunchTimer::seconds() const
{
    double total = 0;
    for (const auto& ev : events) {
        float ms = 0.f;
        MPISDF_HIP(hipEventElapsedTime(&ms, ev.first, ev.second));
        total += 1e-3 * ms;
    }
    return total;
}

LaunchTimer::~LaunchTimer()
{
    for (auto& ev : events) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
}

}  // namespace mpisdf
