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

}  // namespace mpisdf
