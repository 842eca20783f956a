// Shared host/device definitions for the mp_isdf_amd HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace mpisdf {

// ---------------------------------------------------------------------------
// Errors.  Every C-ABI entry point converts exceptions into a status code and
// a thread-local message, which the pybind layer maps back onto the same
// Python exception classes the reference raises (ValueError, OverflowError,
// RuntimeError).
// ---------------------------------------------------------------------------
enum Status : int {
    kOk = 0,
    kInvalid = 1,    // std::invalid_argument  -> ValueError
    kOverflow = 2,   // std::overflow_error    -> OverflowError
    kRuntime = 3,    // std::runtime_error     -> RuntimeError
    kLogic = 4,      // std::logic_error       -> RuntimeError
    kNoDevice = 5,   // no usable HIP device   -> RuntimeError
};

struct NoDevice : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void set_last_error(const std::string& message);
const char* last_error();

#define MPISDF_HIP(expr)                                                              \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess)                                                         \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(_e) + \
                                     " at " __FILE__ ":" + std::to_string(__LINE__) + \
                                     " (" #expr ")");                                 \
    } while (0)

void ensure_device();  // throws NoDevice when no HIP device is present

// Output element types of the AO and sketch kernels.
// kI8Digits / kI8Digits4: a sketch held as three / four signed base-128 digit
// planes per column in place of its FP32 entries (digitize_columns), with a
// per-column power-of-two scale; kI8Digits254: four signed base-254 planes in
// place of FP64 entries (digitize_columns_f64).
enum DType : int { kF64 = 0, kF32 = 1, kBF16 = 2, kFP8 = 3, kI8Digits = 4, kI8Digits4 = 5, kI8Digits254 = 6 };
__host__ __device__ inline bool is_digit_type(DType t) { return t == kI8Digits || t == kI8Digits4 || t == kI8Digits254; }
__host__ __device__ inline int digit_planes_of(DType t) { return t == kI8Digits254 || t == kI8Digits4 ? 4 : 3; }
__host__ __device__ inline int digit_base_of(DType t) { return t == kI8Digits254 ? 254 : 128; }
// Bytes per pitch entry of a digit-plane column: the entries it replaced.
__host__ __device__ inline int digit_column_bytes(DType t) { return t == kI8Digits254 ? 8 : 4; }

// Derivative modes of AO evaluation.
enum AoMode : int { kValues = 0, kGradient = 1, kValuesGradient = 2 };

// One contracted shell as the device kernels see it.  ``tab`` indexes the
// per-angular-momentum transform table (spherical or cartesian), ``rcut2``
// is the squared distance beyond which every value and derivative of the
// shell is below the culling threshold of the evaluation precision.
struct alignas(16) DevShell {
    double x, y, z;
    float rcut2_f64, rcut2_f32;
    int l, nprim, nctr, ncomp;
    int exp_off, coeff_off, ao_off, tab;
};

constexpr int kAngMax = 6;
constexpr int kMaxCart = (kAngMax + 1) * (kAngMax + 2) / 2;

// RAII device buffer.
template <class T>
struct DeviceBuffer {
    T* ptr = nullptr;
    size_t count = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) { resize(n); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), count(o.count) { o.ptr = nullptr; o.count = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; count = o.count; o.ptr = nullptr; o.count = 0; }
        return *this;
    }
    ~DeviceBuffer() { release(); }
    void resize(size_t n) {
        if (n == count) return;
        release();
        if (n) MPISDF_HIP(hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T)));
        count = n;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    void upload(const T* host, size_t n, hipStream_t s) {
        resize(n);
        if (n) MPISDF_HIP(hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    std::vector<T> download(hipStream_t s) const {
        std::vector<T> h(count);
        if (count) {
            MPISDF_HIP(hipMemcpyAsync(h.data(), ptr, count * sizeof(T), hipMemcpyDeviceToHost, s));
            MPISDF_HIP(hipStreamSynchronize(s));
        }
        return h;
    }
};

// A stream of the default priority, destroyed with its owner (after the work
// queued on it has finished).
struct HipStream {
    hipStream_t s = nullptr;
    HipStream();
    ~HipStream();
    HipStream(const HipStream&) = delete;
    HipStream& operator=(const HipStream&) = delete;
};
// An event without timing, for ordering work across streams.
struct HipEvent {
    hipEvent_t e = nullptr;
    HipEvent();
    ~HipEvent();
    HipEvent(const HipEvent&) = delete;
    HipEvent& operator=(const HipEvent&) = delete;
    void record(hipStream_t s);
    void wait(hipStream_t s) const;  // s waits for the last record
};

// Durations of a sequence of launches on one stream, bracketed by HIP events
// (begin/end pairs); read after the stream has been synchronised.
struct LaunchTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    void begin(hipStream_t s);
    void end(hipStream_t s);
    double seconds() const;  // sum of the recorded durations
    size_t launches() const { return events.size(); }
    LaunchTimer() = default;
    LaunchTimer(const LaunchTimer&) = delete;
    LaunchTimer& operator=(const LaunchTimer&) = delete;
    ~LaunchTimer();
};

}  // namespace mpisdf
