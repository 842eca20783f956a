// C ABI of the mp_isdf_amd library (declarations in include/mp_isdf_amd.h).
#include "../../include/mp_isdf_amd.h"

#include "axis.hpp"
#include "context.hpp"
#include "coulomb.hpp"
#include "cvt.hpp"
#include "eri.hpp"
#include "rbrp.hpp"
#include "kernels.hpp"
#include "qrcp.hpp"
#include "util.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>
#include <new>

struct mpisdf_context {
    std::unique_ptr<mpisdf::Context> impl;
};

struct mpisdf_eri_cholesky {
    mpisdf::EriCholesky impl;
};

namespace {

template <class F>
int guard(F&& f)
{
    try {
        f();
        return mpisdf::kOk;
    } catch (const mpisdf::NoDevice& e) {
        mpisdf::set_last_error(e.what());
        return mpisdf::kNoDevice;
    } catch (const std::invalid_argument& e) {
        mpisdf::set_last_error(e.what());
        return mpisdf::kInvalid;
    } catch (const std::overflow_error& e) {
        mpisdf::set_last_error(e.what());
        return mpisdf::kOverflow;
    } catch (const std::bad_alloc&) {
        mpisdf::set_last_error("host allocation failed");
        return mpisdf::kRuntime;
    } catch (const std::logic_error& e) {
        mpisdf::set_last_error(e.what());
        return mpisdf::kLogic;
    } catch (const std::exception& e) {
        mpisdf::set_last_error(e.what());
        return mpisdf::kRuntime;
    }
}

mpisdf::Context& unwrap(const mpisdf_context* ctx)
{
    if (!ctx || !ctx->impl) throw std::invalid_argument("null mp_isdf context");
    return *ctx->impl;
}

// Points per device batch for host-buffer evaluation (bounds device memory).
constexpr int64_t kHostBatch = 1 << 18;

}  // namespace

// Device scratch of the allocation-free entry point (mpisdf_kr_sketch_device_noalloc):
// 2^22 doubles (32 MiB) per GPU, held by each GPU the library runs on until
// the process ends.  Host code never uses the symbol itself as an address:
// the device address comes from hipGetSymbolAddress, once per device.
constexpr int64_t kNoAllocScratchDoubles = int64_t(1) << 22;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kScratchDevices] = {};

// The current device's copy of ``scratch`` (looked up on the first call on
// each device, with that device current).
double* scratch_on_current_device()
{
    int dev = -1;
    MPISDF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices) throw std::runtime_error("device ordinal outside the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

const char* mpisdf_last_error(void) { return mpisdf::last_error(); }

int mpisdf_device_count(int* count)
{
    return guard([&] {
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) { (void)hipGetLastError(); c = 0; }
        *count = c;
    });
}

int mpisdf_context_create(const int32_t* atm, int64_t natm, const int32_t* bas, int64_t nbas,
                          const double* env, int64_t nenv, int cart, mpisdf_context** out)
{
    return guard([&] {
        if (!out) throw std::invalid_argument("null output handle");
        auto handle = std::make_unique<mpisdf_context>();
        handle->impl = mpisdf::make_context(atm, natm, bas, nbas, env, nenv, cart != 0);
        *out = handle.release();
    });
}

int mpisdf_context_destroy(mpisdf_context* ctx)
{
    return guard([&] { delete ctx; });
}

int mpisdf_context_release_workspace(mpisdf_context* ctx)
{
    return guard([&] {
        if (!ctx || !ctx->impl) throw std::invalid_argument("null mp_isdf context");
        ctx->impl->release_sketch_store();
    });
}

int mpisdf_context_nao(const mpisdf_context* ctx, int64_t* nao)
{
    return guard([&] { *nao = unwrap(ctx).nao; });
}

int mpisdf_overlap(const mpisdf_context* ctx, double* s)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (!s) throw std::invalid_argument("null overlap output");
        mpisdf::overlap_matrix(c, s);
    });
}

int mpisdf_eval_ao(const mpisdf_context* ctx, const double* coords, int64_t n, int gradient,
                   double* out)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (n < 0) throw std::invalid_argument("expected coords[ngrid,3] in Bohr");
        const int64_t nao = c.nao;
        if (n == 0 || nao == 0) return;
        for (int64_t i = 0; i < 3 * n; ++i)
            if (!std::isfinite(coords[i])) throw std::invalid_argument("grid coordinates must be finite");
        const int comps = gradient ? 3 : 1;
        const int64_t batch = std::min(n, kHostBatch);
        mpisdf::DeviceBuffer<double> d_xyz(3 * batch), d_out(comps * batch * nao);
        hipStream_t s = nullptr;
        for (int64_t start = 0; start < n; start += batch) {
            const int64_t count = std::min(batch, n - start);
            MPISDF_HIP(hipMemcpyAsync(d_xyz.ptr, coords + 3 * start, 3 * count * sizeof(double),
                                      hipMemcpyHostToDevice, s));
            mpisdf::launch_ao_eval(c, d_xyz.ptr, count, gradient ? mpisdf::kGradient : mpisdf::kValues,
                                   mpisdf::kF64, d_out.ptr, nao, count * nao, s);
            for (int comp = 0; comp < comps; ++comp)
                MPISDF_HIP(hipMemcpyAsync(out + comp * n * nao + start * nao, d_out.ptr + comp * count * nao,
                                          count * nao * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        MPISDF_HIP(hipStreamSynchronize(s));
    });
}

int mpisdf_ao_eval_device(const mpisdf_context* ctx, const double* xyz, int64_t n, int mode,
                          int dtype, void* out, int64_t ld, int64_t cstride, void* stream)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (n < 0) throw std::invalid_argument("negative point count");
        if (mode < 0 || mode > 2) throw std::invalid_argument("mode must be 0, 1 or 2");
        mpisdf::launch_ao_eval(c, xyz, n, static_cast<mpisdf::AoMode>(mode),
                               static_cast<mpisdf::DType>(dtype), out, ld, cstride,
                               static_cast<hipStream_t>(stream));
    });
}

int mpisdf_evaluate_axis(const mpisdf_context* ctx, const double* points, int64_t n,
                         const int32_t* axes, int32_t naxes, const double* mo_coeff, int64_t nocc,
                         double* out, int64_t* width)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        const auto spec = mpisdf::axis_spec(c, axes, naxes, mo_coeff, nocc);
        if (width) *width = spec.width;
        if (!out || n <= 0) return;
        for (int64_t i = 0; i < 3 * n; ++i)
            if (!std::isfinite(points[i])) throw std::invalid_argument("grid coordinates must be finite");
        const int64_t batch = std::min(n, kHostBatch);
        mpisdf::DeviceBuffer<double> d_xyz(3 * batch), d_feat(batch * spec.width);
        mpisdf::AxisWorkspace ws;
        hipStream_t s = nullptr;
        for (int64_t start = 0; start < n; start += batch) {
            const int64_t count = std::min(batch, n - start);
            MPISDF_HIP(hipMemcpyAsync(d_xyz.ptr, points + 3 * start, 3 * count * sizeof(double),
                                      hipMemcpyHostToDevice, s));
            mpisdf::evaluate_axis_device(c, spec, d_xyz.ptr, count, nullptr, mo_coeff, nocc, d_feat.ptr,
                                         ws, s);
            MPISDF_HIP(hipMemcpyAsync(out + start * spec.width, d_feat.ptr,
                                      count * spec.width * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        MPISDF_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"

#include "fit.hpp"
#include "grid.hpp"
#include "lebedev.hpp"
#include "sketch.hpp"
#include "sketch_mfma.hpp"
#include "select.hpp"

namespace {

int select_impl(const mpisdf_context* ctx, const double* points, const double* weights, int64_t n,
                const int32_t* axes, int32_t naxes, const double* mo_coeff, int64_t nocc, const double* axis_weights,
                int64_t naxis_weights, double tolerance, int64_t max_rank, int32_t algorithm, int64_t batch_size,
                int32_t debug, int64_t sketch_size, int64_t oversampling, uint64_t seed, int32_t precision,
                const double* point_importance, double ao_screening_threshold, int64_t* out, int64_t capacity,
                int64_t* count, double* diagnostics)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (n < 0 || !count) throw std::invalid_argument("expected points[ngrid,3] and weights[ngrid]");
        if (naxes <= 0 || !axes) throw std::invalid_argument("axis list must not be empty");
        if (algorithm < 0 || algorithm > 5) throw std::invalid_argument("unknown anchor selection algorithm");
        if (algorithm == 3) throw std::invalid_argument("CVT selection is not part of this build");
        if (precision < 0 || precision > 5) throw std::invalid_argument("precision must be 0..5");
        if (batch_size < 0 || sketch_size < 0 || oversampling < 0 || max_rank < 0)
            throw std::invalid_argument("sizes must be non-negative");
        mpisdf::SelectArgs a;
        a.points = points;
        a.weights = weights;
        a.n = n;
        a.axes.assign(axes, axes + naxes);
        a.mo_coeff = mo_coeff;
        a.nocc = nocc;
        if (axis_weights && naxis_weights > 0) a.axis_weights.assign(axis_weights, axis_weights + naxis_weights);
        a.tolerance = tolerance;
        a.max_rank = max_rank;
        a.algorithm = algorithm;
        a.batch_size = batch_size;
        a.debug = (debug & 1) != 0;
        a.audit = (debug & 2) != 0;
        a.sketch_size = sketch_size;
        a.oversampling = oversampling;
        a.seed = seed;
        a.precision = precision;
        a.importance = point_importance;
        a.ao_screening_threshold = ao_screening_threshold;
        const auto sel = mpisdf::select_anchors(c, a, diagnostics);
        if ((int64_t)sel.size() > capacity) throw std::invalid_argument("output capacity is too small");
        std::copy(sel.begin(), sel.end(), out);
        *count = (int64_t)sel.size();
    });
}

}  // namespace

extern "C" int mpisdf_select(const mpisdf_context* ctx, const double* points, const double* weights, int64_t n,
                             const int32_t* axes, int32_t naxes, const double* mo_coeff, int64_t nocc,
                             const double* axis_weights, int64_t naxis_weights, double tolerance,
                             int64_t max_rank, int32_t algorithm, int64_t batch_size, int32_t debug,
                             int64_t sketch_size, int64_t oversampling, uint64_t seed, int32_t precision,
                             int64_t* out, int64_t capacity, int64_t* count, double* diagnostics)
{
    return select_impl(ctx, points, weights, n, axes, naxes, mo_coeff, nocc, axis_weights, naxis_weights, tolerance,
                       max_rank, algorithm, batch_size, debug, sketch_size, oversampling, seed, precision, nullptr,
                       1e-24, out, capacity, count, diagnostics);
}

extern "C" int mpisdf_select_importance(const mpisdf_context* ctx, const double* points, const double* weights,
                                        int64_t n, const int32_t* axes, int32_t naxes, const double* mo_coeff,
                                        int64_t nocc, const double* axis_weights, int64_t naxis_weights,
                                        double tolerance, int64_t max_rank, int32_t algorithm, int64_t batch_size,
                                        int32_t debug, int64_t sketch_size, int64_t oversampling, uint64_t seed,
                                        int32_t precision, const double* point_importance,
                                        double ao_screening_threshold, int64_t* out, int64_t capacity,
                                        int64_t* count, double* diagnostics)
{
    return select_impl(ctx, points, weights, n, axes, naxes, mo_coeff, nocc, axis_weights, naxis_weights, tolerance,
                       max_rank, algorithm, batch_size, debug, sketch_size, oversampling, seed, precision,
                       point_importance, ao_screening_threshold, out, capacity, count, diagnostics);
}

extern "C" int mpisdf_becke_grid(const int32_t* charges, const double* coords, int64_t natm, int32_t level,
                                 int32_t device, double* xyz, double* weights, int64_t capacity, int64_t* count)
{
    return guard([&] {
        if (!charges || !coords || !count) throw std::invalid_argument("null grid argument");
        std::vector<int> z(charges, charges + natm);
        const int64_t n = mpisdf::becke_grid_size(z.data(), natm, level);
        *count = n;
        if (!xyz) return;
        if (!weights) throw std::invalid_argument("null grid argument");
        if (capacity < n) throw std::invalid_argument("grid output capacity is too small");
        auto grid = mpisdf::becke_grid(z.data(), coords, natm, level, device != 0);
        std::copy(grid.first.begin(), grid.first.end(), xyz);
        std::copy(grid.second.begin(), grid.second.end(), weights);
    });
}

extern "C" int mpisdf_lebedev(int32_t npoints, double* xyz, double* weights)
{
    return guard([&] {
        const mpisdf::LebedevRule* rule = mpisdf::lebedev_rule(npoints);
        if (!rule) throw std::invalid_argument("no tabulated Lebedev rule has " + std::to_string(npoints) + " points");
        if (!xyz || !weights) throw std::invalid_argument("null Lebedev output");
        std::vector<double> x, w;
        mpisdf::lebedev_points(*rule, x, w);
        std::copy(x.begin(), x.end(), xyz);
        std::copy(w.begin(), w.end(), weights);
    });
}

extern "C" int mpisdf_omega(int64_t d, int64_t ell, uint64_t seed, double* u, double* v)
{
    return guard([&] {
        if (d < 0 || ell < 0 || !u || !v) throw std::invalid_argument("invalid omega arguments");
        std::vector<double> hu, hv;
        mpisdf::make_omega(d, ell, seed, hu, hv);
        std::copy(hu.begin(), hu.end(), u);
        std::copy(hv.begin(), hv.end(), v);
    });
}

extern "C" int mpisdf_kr_sketch_f64_device(const double* phi, int64_t count, int64_t d, const double* u,
                                           const double* v, int64_t ell, const double* sw, double* y, int64_t ldy,
                                           unsigned long long* bad, void* stream)
{
    return guard([&] {
        if (count < 0 || d <= 0 || ell <= 0 || ldy < ell) throw std::invalid_argument("invalid sketch shape");
        mpisdf::SketchScratch scratch;
        mpisdf::DeviceBuffer<unsigned long long> own;
        if (!bad) {
            own.resize(1);
            MPISDF_HIP(hipMemsetAsync(own.ptr, 0, sizeof(unsigned long long), static_cast<hipStream_t>(stream)));
            bad = own.ptr;
        }
        mpisdf::sketch_f64(phi, count, d, u, v, ell, sw, y, ldy, scratch, bad, static_cast<hipStream_t>(stream));
        MPISDF_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    });
}

extern "C" int mpisdf_sketch_prepare_omega(const double* u, const double* v, int64_t d, int64_t ell, int64_t ldo,
                                           int dtype, void* ut, void* vt, void* stream)
{
    return guard([&] {
        if (ldo < d || ldo % 16) throw std::invalid_argument("ldo must be >= d and a multiple of 16");
        mpisdf::prepare_omega_lowp(u, v, d, ell, ldo, dtype, ut, vt, static_cast<hipStream_t>(stream));
    });
}

extern "C" int mpisdf_sketch_split_bf16x3(const double* phi, int64_t ld, int64_t count, int64_t d, void* out,
                                          int64_t ldo, void* stream)
{
    return guard([&] {
        if (ldo < d || ldo % 16 || ld < d) throw std::invalid_argument("invalid split row strides");
        mpisdf::split_rows_bf16x3(phi, ld, count, d, ldo, out, static_cast<hipStream_t>(stream));
    });
}

extern "C" int mpisdf_sketch_scales_bf16(const double* sw, int64_t count, int64_t ell, float* scale, void* stream)
{
    return guard([&] { mpisdf::bf16_point_scales(sw, count, ell, scale, static_cast<hipStream_t>(stream)); });
}

extern "C" int mpisdf_sketch_quantize_fp8(const float* phi, int64_t ld, int64_t count, int64_t d, const double* sw,
                                          int64_t ell, uint8_t* out, int64_t ldo, float* scale, void* stream)
{
    return guard([&] {
        if (ldo < d || ld < d) throw std::invalid_argument("invalid fp8 row strides");
        mpisdf::quantize_fp8_rows(phi, ld, count, d, sw, ell, out, ldo, scale, static_cast<hipStream_t>(stream));
    });
}

extern "C" int mpisdf_kr_sketch_device(int dtype, const void* ut, const void* vt, int64_t ell, int64_t d,
                                       const void* phi, int64_t ldphi, int64_t count, const float* scale, float* y,
                                       int64_t ldy, double* ynorm2, unsigned long long* bad, void* stream)
{
    return guard([&] {
        if (ldy < ell || ldphi < d || !ynorm2 || !bad) throw std::invalid_argument("invalid sketch arguments");
        mpisdf::DeviceBuffer<double> part;
        mpisdf::kr_sketch_lowp(static_cast<mpisdf::DType>(dtype), ut, vt, ell, d, phi, ldphi, count, scale, y, ldy,
                               ynorm2, bad, part, static_cast<hipStream_t>(stream));
        MPISDF_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    });
}

extern "C" int mpisdf_kr_sketch_device_noalloc(int dtype, const void* ut, const void* vt, int64_t ell, int64_t d,
                                               const void* phi, int64_t ldphi, int64_t count, const float* scale,
                                               float* y, int64_t ldy, double* ynorm2, unsigned long long* bad,
                                               void* stream)
{
    return guard([&] {
        if (ldy < ell || ldphi < d || !ynorm2 || !bad) throw std::invalid_argument("invalid sketch arguments");
        const int64_t need = mpisdf::kr_sketch_lowp_partials(ell, count);
        if (need > kNoAllocScratchDoubles)
            throw std::invalid_argument("sketch too large for the allocation-free entry point: count * ceil(ell / 256) = " +
                                        std::to_string(need) + " exceeds its scratch of " +
                                        std::to_string(kNoAllocScratchDoubles) + " doubles");
        double* part = scratch_on_current_device();
        mpisdf::kr_sketch_lowp_into(static_cast<mpisdf::DType>(dtype), ut, vt, ell, d, phi, ldphi, count, scale, y,
                                    ldy, ynorm2, bad, part, static_cast<hipStream_t>(stream));
    });
}

extern "C" int mpisdf_qrcp_fold_device(int64_t b, int64_t n, int64_t m, const double* Qb, const void* Y, int y_f32,
                                       int64_t ldy, double* proj2, double* energy, void* stream)
{
    return guard([&] {
        if (b < 0 || n < 0 || m <= 0 || ldy < m || !proj2 || y_f32 < 0 || y_f32 > 5)
            throw std::invalid_argument("invalid fold arguments");
        if (y_f32 == 5) {
            // Four base-254 digit planes of a copy of FP64 Y (pitch P = m
            // rounded up to 64), then the INT8 fold.
            mpisdf::FoldI8Scratch scratch;
            auto st = static_cast<hipStream_t>(stream);
            const int64_t P = (m + mpisdf::kDigitSlice - 1) / mpisdf::kDigitSlice * mpisdf::kDigitSlice;
            mpisdf::DeviceBuffer<double> yn(std::max<int64_t>(n, 1)), sc(std::max<int64_t>(n, 1));
            mpisdf::DeviceBuffer<double> yd((size_t)std::max<int64_t>(n * P, 1));
            mpisdf::launch_column_norms2_f64(static_cast<const double*>(Y), m, n, ldy, yn.ptr, st);
            MPISDF_HIP(hipMemcpy2DAsync(yd.ptr, P * sizeof(double), Y, ldy * sizeof(double), m * sizeof(double), n,
                                        hipMemcpyDeviceToDevice, st));
            mpisdf::digitize_columns_f64(yd.ptr, m, n, P, yn.ptr, sc.ptr, st);
            mpisdf::fold_projections_i8(b, n, m, Qb, m, reinterpret_cast<const int8_t*>(yd.ptr), P, 4, 254, 8, sc.ptr,
                                        proj2, energy, scratch, st);
            MPISDF_HIP(hipStreamSynchronize(st));
        } else if (y_f32 == 3 || y_f32 == 4) {
            // Digits (3 or 4 planes) of a copy of Y (pitch P = m rounded up to
            // 64), then the INT8 fold.
            mpisdf::FoldI8Scratch scratch;
            auto st = static_cast<hipStream_t>(stream);
            const int64_t P = (m + mpisdf::kDigitSlice - 1) / mpisdf::kDigitSlice * mpisdf::kDigitSlice;
            mpisdf::DeviceBuffer<double> yn(std::max<int64_t>(n, 1)), sc(std::max<int64_t>(n, 1));
            mpisdf::DeviceBuffer<float> yd((size_t)std::max<int64_t>(n * P, 1));
            mpisdf::launch_column_norms2_f32(static_cast<const float*>(Y), m, n, ldy, yn.ptr, st);
            MPISDF_HIP(hipMemcpy2DAsync(yd.ptr, P * sizeof(float), Y, ldy * sizeof(float), m * sizeof(float), n,
                                        hipMemcpyDeviceToDevice, st));
            const int planes = y_f32 == 4 ? 4 : 3;
            mpisdf::digitize_columns(yd.ptr, m, n, P, planes, yn.ptr, sc.ptr, st);
            mpisdf::fold_projections_i8(b, n, m, Qb, m, reinterpret_cast<const int8_t*>(yd.ptr), P, planes, 128, 4,
                                        sc.ptr, proj2, energy, scratch, st);
            MPISDF_HIP(hipStreamSynchronize(st));
        } else if (y_f32 == 2) {
            mpisdf::FoldScratch scratch;
            mpisdf::fold_projections_bf16x3(b, n, m, Qb, m, static_cast<const float*>(Y), ldy, proj2, energy, scratch,
                                            static_cast<hipStream_t>(stream));
            MPISDF_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
        } else {
            mpisdf::DeviceBuffer<double> part;
            mpisdf::fold_projections(b, n, m, Qb, m, Y, y_f32 != 0, ldy, proj2, energy, part,
                                     static_cast<hipStream_t>(stream));
            MPISDF_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
        }
    });
}

extern "C" int mpisdf_digitize_columns_device(float* y, int64_t m, int64_t n, int64_t P, int32_t planes,
                                              double* ynorm2, double* scale, void* stream)
{
    return guard([&] {
        if (m <= 0 || n < 0 || !y || !ynorm2 || !scale) throw std::invalid_argument("invalid digitise arguments");
        mpisdf::digitize_columns(y, m, n, P, planes, ynorm2, scale, static_cast<hipStream_t>(stream));
        MPISDF_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    });
}

extern "C" int mpisdf_digitize_columns_f64_device(double* y, int64_t m, int64_t n, int64_t P, double* ynorm2,
                                                  double* scale, void* stream)
{
    return guard([&] {
        if (m <= 0 || n < 0 || !y || !ynorm2 || !scale) throw std::invalid_argument("invalid digitise arguments");
        mpisdf::digitize_columns_f64(y, m, n, P, ynorm2, scale, static_cast<hipStream_t>(stream));
        MPISDF_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    });
}

extern "C" int mpisdf_qrcp_device(const void* y, int ytype, int64_t m, int64_t n, int64_t ldy, const double* ynorm2,
                                  const double* yscale, int64_t kmax, double stop_rel, int64_t* pivots,
                                  double* rdiag, int64_t* count, void* stream)
{
    return guard([&] {
        if (!y || !pivots || !count || m <= 0 || n <= 0 || kmax <= 0 || !(stop_rel >= 0))
            throw std::invalid_argument("invalid QRCP arguments");
        if (ytype != mpisdf::kF64 && ytype != mpisdf::kF32 && !mpisdf::is_digit_type((mpisdf::DType)ytype))
            throw std::invalid_argument("QRCP input type must be 0 (f64), 1 (f32), 4 or 5 (3 or 4 base-128 planes) or 6 (4 base-254 planes)");
        mpisdf::QrcpOptions opt;
        opt.kmax = kmax;
        opt.stop_rel = stop_rel;
        opt.candidates = std::min<int64_t>(n, 2048);
        opt.ynorm2 = ynorm2;
        if (mpisdf::is_digit_type((mpisdf::DType)ytype)) {
            opt.yscale = yscale;
            opt.margin = mpisdf::kFoldI8Margin;
        }
        const auto r = mpisdf::qrcp_device(y, static_cast<mpisdf::DType>(ytype), m, n, ldy, opt,
                                           static_cast<hipStream_t>(stream));
        const int64_t k = std::min<int64_t>(kmax, (int64_t)r.pivots.size());
        std::copy(r.pivots.begin(), r.pivots.begin() + k, pivots);
        if (rdiag) std::copy(r.rdiag.begin(), r.rdiag.begin() + k, rdiag);
        *count = k;
    });
}

extern "C" int mpisdf_qrcp_device_ordered(const void* y, int ytype, int64_t m, int64_t n, int64_t ldy,
                                          const double* ynorm2, const double* yscale, const int64_t* tie_rank,
                                          int32_t dormant, int64_t kmax, double stop_rel, int64_t* pivots,
                                          double* rdiag, int64_t* count, int64_t* folded_columns, void* stream)
{
    return guard([&] {
        if (!y || !pivots || !count || m <= 0 || n <= 0 || kmax <= 0 || !(stop_rel >= 0))
            throw std::invalid_argument("invalid QRCP arguments");
        if (ytype != mpisdf::kF64 && ytype != mpisdf::kF32 && !mpisdf::is_digit_type((mpisdf::DType)ytype))
            throw std::invalid_argument("QRCP input type must be 0 (f64), 1 (f32), 4 or 5 (3 or 4 base-128 planes) or 6 (4 base-254 planes)");
        mpisdf::QrcpOptions opt;
        opt.kmax = kmax;
        opt.stop_rel = stop_rel;
        opt.candidates = std::min<int64_t>(n, 1024);
        opt.ynorm2 = ynorm2;
        opt.tie_rank = tie_rank;
        opt.dormant = dormant != 0;
        if (mpisdf::is_digit_type((mpisdf::DType)ytype)) {
            opt.yscale = yscale;
            opt.margin = mpisdf::kFoldI8Margin;
        }
        const auto r = mpisdf::qrcp_device(y, static_cast<mpisdf::DType>(ytype), m, n, ldy, opt,
                                           static_cast<hipStream_t>(stream));
        const int64_t k = std::min<int64_t>(kmax, (int64_t)r.pivots.size());
        std::copy(r.pivots.begin(), r.pivots.begin() + k, pivots);
        if (rdiag) std::copy(r.rdiag.begin(), r.rdiag.begin() + k, rdiag);
        *count = k;
        if (folded_columns) *folded_columns = r.folded_columns;
    });
}

namespace {

void run_fit(mpisdf::Context& c, const double* anchor_xyz_dev, const double* anchor_w, int64_t k,
             const double* xyz_dev, const double* w_dev, int64_t n, int64_t batch, double* theta, double* tsum,
             double* results, hipStream_t s, mpisdf::FitProducts products = mpisdf::FitProducts::F64,
             bool extended = false)
{
    if (k <= 0) throw std::invalid_argument("the fit needs at least one anchor");
    for (int64_t i = 0; i < k; ++i)
        if (!std::isfinite(anchor_w[i]) || anchor_w[i] == 0)
            throw std::invalid_argument("anchor weights must be finite and nonzero");
    if (batch <= 0) throw std::invalid_argument("fit_batch_size must be positive");
    const auto t0 = std::chrono::steady_clock::now();
    mpisdf::FitFactor f;
    mpisdf::fit_factor(c, anchor_xyz_dev, k, f, s);
    const auto t1 = std::chrono::steady_clock::now();
    double err[2] = {0, 0};
    double gemm_stats[5] = {0, 0, 0, 0, 0};
    if (n > 0)
        mpisdf::fit_stream(c, f, anchor_w, xyz_dev, w_dev, n, batch, theta, err, tsum, s, gemm_stats, products);
    MPISDF_HIP(hipStreamSynchronize(s));
    const auto t2 = std::chrono::steady_clock::now();
    if (results) {
        results[0] = std::sqrt(err[0]);
        results[1] = err[1] > 0 ? std::sqrt(err[0] / err[1]) : (err[0] == 0 ? 0.0 : INFINITY);
        results[2] = f.condition_estimate;
        results[3] = f.gram_diag_max;
        results[4] = std::chrono::duration<double>(t1 - t0).count();
        results[5] = std::chrono::duration<double>(t2 - t1).count();
        results[6] = gemm_stats[0];
        results[7] = gemm_stats[1];
        results[8] = (double)f.retained_rank;
        results[9] = f.triangular ? 0.0 : 1.0;
        if (extended) {
            results[10] = gemm_stats[3];
            results[11] = gemm_stats[4];
        }
    }
}

}  // namespace

extern "C" int mpisdf_isdf_fit(const mpisdf_context* ctx, const double* anchor_xyz, const double* anchor_w, int64_t k,
                               const double* xyz, const double* w, int64_t n, int64_t batch, double* theta,
                               double* tsum, double* results)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (n < 0 || k < 0) throw std::invalid_argument("negative sizes");
        for (int64_t i = 0; i < 3 * n; ++i)
            if (!std::isfinite(xyz[i])) throw std::invalid_argument("grid coordinates must be finite");
        mpisdf::DeviceBuffer<double> dxyz, dw, dax;
        dxyz.upload(xyz, 3 * n, nullptr);
        dw.upload(w, n, nullptr);
        dax.upload(anchor_xyz, 3 * k, nullptr);
        run_fit(c, dax.ptr, anchor_w, k, dxyz.ptr, dw.ptr, n, batch, theta, tsum, results, nullptr);
    });
}

extern "C" int mpisdf_isdf_fit_device(const mpisdf_context* ctx, const double* anchor_xyz_dev, const double* anchor_w,
                                      int64_t k, const double* xyz_dev, const double* w_dev, int64_t n, int64_t batch,
                                      double* tsum, double* results, void* stream)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        run_fit(c, anchor_xyz_dev, anchor_w, k, xyz_dev, w_dev, n, batch, nullptr, tsum, results,
                static_cast<hipStream_t>(stream));
    });
}

extern "C" int mpisdf_isdf_fit_device_form(const mpisdf_context* ctx, const double* anchor_xyz_dev,
                                           const double* anchor_w, int64_t k, const double* xyz_dev,
                                           const double* w_dev, int64_t n, int64_t batch, int32_t form, double* tsum,
                                           double* results, void* stream)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (form != 0 && form != 1) throw std::invalid_argument("fit form must be 0 (FP64) or 1 (INT8 digits)");
        if (!anchor_xyz_dev || !anchor_w || (n > 0 && (!xyz_dev || !w_dev)))
            throw std::invalid_argument("invalid fit arguments");
        run_fit(c, anchor_xyz_dev, anchor_w, k, xyz_dev, w_dev, n, batch, nullptr, tsum, results,
                static_cast<hipStream_t>(stream),
                form == 1 ? mpisdf::FitProducts::Int8Digits : mpisdf::FitProducts::F64, true);
    });
}

extern "C" int mpisdf_pair_space_error(const mpisdf_context* ctx, const double* anchor_xyz, const double* anchor_w,
                                       int64_t k, const double* xyz, const double* w, int64_t n, const double* theta,
                                       int64_t batch, double* results)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (k <= 0 || n < 0 || batch <= 0 || !anchor_xyz || !anchor_w || !results || (n > 0 && (!xyz || !w || !theta)))
            throw std::invalid_argument("invalid pair-space error arguments");
        mpisdf::DeviceBuffer<double> dax, dxyz, dw;
        dax.upload(anchor_xyz, 3 * k, nullptr);
        double err[2] = {0, 0};
        if (n > 0) {
            dxyz.upload(xyz, 3 * n, nullptr);
            dw.upload(w, n, nullptr);
            mpisdf::theta_pair_error(c, dax.ptr, anchor_w, k, dxyz.ptr, dw.ptr, n, theta, batch, err, nullptr);
        }
        results[0] = std::sqrt(err[0]);
        results[1] = err[1] > 0 ? std::sqrt(err[0] / err[1]) : (err[0] == 0 ? 0.0 : INFINITY);
    });
}

extern "C" int mpisdf_quadrature_overlap(const mpisdf_context* ctx, const double* xyz, const double* w, int64_t n,
                                         int64_t batch, double* out)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (n < 0 || batch <= 0 || !out || (n > 0 && (!xyz || !w)))
            throw std::invalid_argument("invalid quadrature overlap arguments");
        mpisdf::DeviceBuffer<double> dxyz, dw, ds((size_t)(c.nao * c.nao));
        if (n > 0) {
            dxyz.upload(xyz, 3 * n, nullptr);
            dw.upload(w, n, nullptr);
        }
        mpisdf::quadrature_overlap(c, dxyz.ptr, dw.ptr, n, batch, ds.ptr, nullptr);
        MPISDF_HIP(hipMemcpy(out, ds.ptr, (size_t)(c.nao * c.nao) * sizeof(double), hipMemcpyDeviceToHost));
    });
}

extern "C" int mpisdf_ao_adjoint_device(const mpisdf_context* ctx, const double* xyz, int64_t n, const double* abar,
                                        int64_t ldab, double* gout, void* stream)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (n < 0 || !gout || (n > 0 && (!xyz || !abar))) throw std::invalid_argument("invalid AO adjoint arguments");
        mpisdf::DeviceBuffer<double> part;
        mpisdf::launch_ao_adjoint(c, xyz, n, abar, ldab, gout, part, static_cast<hipStream_t>(stream));
        MPISDF_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    });
}

namespace {
// grad_atoms[3 A + c] = -sum_{mu on A} gao[c nao + mu]: moving a nucleus moves its AOs.
void atoms_from_ao_gradient(const mpisdf::Context& c, const std::vector<double>& gao, double* grad_atoms)
{
    const int64_t natm = (int64_t)c.atm.size() / 6;
    std::fill(grad_atoms, grad_atoms + 3 * natm, 0.0);
    for (size_t sidx = 0; sidx < c.shells.size(); ++sidx) {
        const int atom = c.bas[sidx * 8 + 0];
        const auto& sh = c.shells[sidx];
        for (int mu = sh.ao_off; mu < sh.ao_off + sh.nctr * sh.ncomp; ++mu)
            for (int comp = 0; comp < 3; ++comp) grad_atoms[3 * atom + comp] -= gao[comp * c.nao + mu];
    }
}
}  // namespace

extern "C" int mpisdf_isdf_fit_gradient(const mpisdf_context* ctx, const double* anchor_xyz, const double* anchor_w,
                                        int64_t k, const double* xyz, const double* w, int64_t n, int64_t batch,
                                        double* loss, double* grad_ao, double* grad_atoms)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (n < 0 || k <= 0 || batch <= 0) throw std::invalid_argument("invalid fit gradient sizes");
        for (int64_t i = 0; i < k; ++i)
            if (!std::isfinite(anchor_w[i]) || anchor_w[i] == 0)
                throw std::invalid_argument("anchor weights must be finite and nonzero");
        mpisdf::DeviceBuffer<double> dxyz, dw, dax;
        dxyz.upload(xyz, 3 * n, nullptr);
        dw.upload(w, n, nullptr);
        dax.upload(anchor_xyz, 3 * k, nullptr);
        mpisdf::FitFactor f;
        mpisdf::fit_factor(c, dax.ptr, k, f, nullptr);
        std::vector<double> gao(3 * c.nao, 0.0);
        double l = 0;
        mpisdf::fit_loss_gradient(c, f, dax.ptr, dxyz.ptr, dw.ptr, n, batch, &l, gao.data(), nullptr);
        if (loss) *loss = l;
        if (grad_ao) std::copy(gao.begin(), gao.end(), grad_ao);
        if (grad_atoms) atoms_from_ao_gradient(c, gao, grad_atoms);
    });
}

extern "C" int mpisdf_coulomb_energy_gradient(const mpisdf_context* ctx, const double* anchor_xyz, int64_t k,
                                              const double* xyz, const double* w, int64_t n, const double* dm,
                                              int64_t batch, double* energy, double* grad_ao, double* grad_atoms)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (n < 0 || k <= 0 || batch <= 0 || !anchor_xyz || !dm || (n > 0 && (!xyz || !w)))
            throw std::invalid_argument("invalid Coulomb-energy arguments");
        for (int64_t g = 0; g < n; ++g)
            if (!std::isfinite(w[g]) || w[g] == 0)
                throw std::invalid_argument("grid weights must be finite and nonzero (the metric's self term)");
        const int64_t nao = c.nao;
        for (int64_t i = 0; i < nao; ++i)
            for (int64_t j = 0; j < nao; ++j) {
                const double a = dm[i * nao + j], b = dm[j * nao + i];
                if (!std::isfinite(a)) throw std::invalid_argument("dm must be a finite AO density matrix");
                if (std::abs(a - b) > 1e-12 + 1e-12 * std::abs(b)) throw std::invalid_argument("dm must be symmetric");
            }
        mpisdf::DeviceBuffer<double> dxyz, dw, dax, ddm;
        dxyz.upload(xyz, 3 * n, nullptr);
        dw.upload(w, n, nullptr);
        dax.upload(anchor_xyz, 3 * k, nullptr);
        ddm.upload(dm, nao * nao, nullptr);
        mpisdf::FitFactor f;
        mpisdf::fit_factor(c, dax.ptr, k, f, nullptr);
        const bool want = grad_ao || grad_atoms;
        std::vector<double> gao(want ? 3 * nao : 0, 0.0);
        double e = 0;
        mpisdf::coulomb_energy_gradient(c, f, dax.ptr, dxyz.ptr, dw.ptr, n, ddm.ptr, batch, &e,
                                        want ? gao.data() : nullptr, nullptr);
        MPISDF_HIP(hipStreamSynchronize(nullptr));
        if (energy) *energy = e;
        if (grad_ao) std::copy(gao.begin(), gao.end(), grad_ao);
        if (grad_atoms) atoms_from_ao_gradient(c, gao, grad_atoms);
    });
}

extern "C" int mpisdf_ao_adjoint_host(const mpisdf_context* ctx, const double* xyz, int64_t n, const double* abar,
                                      double* gout)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (n < 0 || !gout) throw std::invalid_argument("invalid AO adjoint arguments");
        mpisdf::DeviceBuffer<double> dxyz, dab, dg(3 * c.nao);
        MPISDF_HIP(hipMemsetAsync(dg.ptr, 0, 3 * c.nao * sizeof(double), nullptr));
        if (n > 0) {
            dxyz.upload(xyz, 3 * n, nullptr);
            dab.upload(abar, n * c.nao, nullptr);
            mpisdf::DeviceBuffer<double> part;
            mpisdf::launch_ao_adjoint(c, dxyz.ptr, n, dab.ptr, c.nao, dg.ptr, part, nullptr);
            MPISDF_HIP(hipStreamSynchronize(nullptr));
        }
        MPISDF_HIP(hipMemcpy(gout, dg.ptr, 3 * c.nao * sizeof(double), hipMemcpyDeviceToHost));
    });
}

namespace {

// Copies rows x n host values into a device buffer with rows at a stride of
// round-up(n, 4) doubles (the TN product's vector load path).
int64_t upload_rows(const double* host, int64_t rows, int64_t n, mpisdf::DeviceBuffer<double>& dev)
{
    const int64_t ld = std::max<int64_t>((n + 3) / 4 * 4, 4);
    dev.resize((size_t)(std::max<int64_t>(rows, 1) * ld));
    if (rows > 0 && n > 0)
        MPISDF_HIP(hipMemcpy2D(dev.ptr, ld * sizeof(double), host, n * sizeof(double), n * sizeof(double),
                               (size_t)rows, hipMemcpyHostToDevice));
    return ld;
}

void require_finite(const double* p, int64_t count, const char* msg)
{
    for (int64_t i = 0; i < count; ++i)
        if (!std::isfinite(p[i])) throw std::invalid_argument(msg);
}

}  // namespace

extern "C" int mpisdf_grid_coulomb_metric(const double* left, int64_t kl, const double* right, int64_t kr,
                                          const double* xyz, const double* w, int64_t n, int symmetrize,
                                          double* metric)
{
    return guard([&] {
        mpisdf::ensure_device();
        if (kl < 0 || kr < 0 || n < 0) throw std::invalid_argument("negative metric sizes");
        if (symmetrize && kl != kr) throw std::invalid_argument("a symmetrised metric needs as many left as right rows");
        if (kl == 0 || kr == 0) return;
        if (!metric || (n > 0 && (!left || !right || !xyz || !w)))
            throw std::invalid_argument("null metric argument");
        require_finite(xyz, 3 * n, "grid coordinates must be finite");
        for (int64_t i = 0; i < n; ++i)
            if (!std::isfinite(w[i]) || w[i] == 0.0)
                throw std::invalid_argument("grid weights must be finite and nonzero (the self-cell term is "
                                            "|w|^(-1/3))");
        require_finite(left, kl * n, "interpolation vectors must be finite");
        if (right != left) require_finite(right, kr * n, "interpolation vectors must be finite");
        mpisdf::DeviceBuffer<double> dl, dr, dxyz, dw, dm((size_t)(kl * kr));
        const int64_t ldl = upload_rows(left, kl, n, dl);
        int64_t ldr = ldl;
        const double* rp = dl.ptr;
        if (right != left) {
            ldr = upload_rows(right, kr, n, dr);
            rp = dr.ptr;
        }
        dxyz.upload(xyz, 3 * n, nullptr);
        dw.upload(w, n, nullptr);
        mpisdf::grid_coulomb_metric(dl.ptr, kl, ldl, rp, kr, ldr, dxyz.ptr, dw.ptr, n, symmetrize != 0, dm.ptr, 0,
                                    nullptr);
        MPISDF_HIP(hipMemcpy(metric, dm.ptr, kl * kr * sizeof(double), hipMemcpyDeviceToHost));
    });
}

extern "C" int mpisdf_grid_coulomb_metric_device(const double* left, int64_t kl, int64_t ldl, const double* right,
                                                 int64_t kr, int64_t ldr, const double* xyz, const double* w,
                                                 int64_t n, int symmetrize, double* metric, int64_t block_rows,
                                                 void* stream)
{
    return guard([&] {
        if (kl > 0 && kr > 0 && (!metric || (n > 0 && (!left || !right || !xyz || !w))))
            throw std::invalid_argument("null metric argument");
        mpisdf::grid_coulomb_metric(left, kl, ldl, right, kr, ldr, xyz, w, n, symmetrize != 0, metric, block_rows,
                                    static_cast<hipStream_t>(stream));
        MPISDF_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    });
}

extern "C" int mpisdf_isdf_jk(const double* x, int64_t k, int64_t nao, const double* z, const double* dm,
                              double* j, double* kmat)
{
    return guard([&] {
        mpisdf::ensure_device();
        if (k < 0 || nao < 0) throw std::invalid_argument("negative J/K sizes");
        if (!j || !kmat || (nao > 0 && !dm) || (k > 0 && (!x || !z))) throw std::invalid_argument("null J/K argument");
        require_finite(dm, nao * nao, "dm must be a finite AO density matrix of shape (nao,nao)");
        // numpy.allclose(dm, dm.T, rtol=1e-12, atol=1e-12) (python/isdf.py:143)
        for (int64_t a = 0; a < nao; ++a)
            for (int64_t b = 0; b < nao; ++b)
                if (std::fabs(dm[a * nao + b] - dm[b * nao + a]) > 1e-12 + 1e-12 * std::fabs(dm[b * nao + a]))
                    throw std::invalid_argument("dm must be symmetric");
        require_finite(x, k * nao, "anchor AO values must be finite");
        require_finite(z, k * k, "the Coulomb metric must be finite");
        if (nao == 0) return;
        mpisdf::DeviceBuffer<double> dx, dz, ddm, dj((size_t)(nao * nao)), dk((size_t)(nao * nao));
        dx.upload(x, k * nao, nullptr);
        dz.upload(z, k * k, nullptr);
        ddm.upload(dm, nao * nao, nullptr);
        mpisdf::isdf_jk(dx.ptr, k, nao, dz.ptr, ddm.ptr, dj.ptr, dk.ptr, nullptr);
        MPISDF_HIP(hipMemcpy(j, dj.ptr, nao * nao * sizeof(double), hipMemcpyDeviceToHost));
        MPISDF_HIP(hipMemcpy(kmat, dk.ptr, nao * nao * sizeof(double), hipMemcpyDeviceToHost));
    });
}

extern "C" int mpisdf_isdf_jk_device(const double* x, int64_t k, int64_t nao, const double* z, const double* dm,
                                     double* j, double* kmat, void* stream)
{
    return guard([&] {
        if (nao > 0 && (!j || !kmat || !dm || (k > 0 && (!x || !z)))) throw std::invalid_argument("null J/K argument");
        mpisdf::isdf_jk(x, k, nao, z, dm, j, kmat, static_cast<hipStream_t>(stream));
    });
}

extern "C" int mpisdf_select_cvt(const double* points, const double* weights, const double* cvt_weights, int64_t n,
                                 int64_t rank, uint64_t seed, int64_t max_iter, double tol, int64_t* out,
                                 int64_t capacity, int64_t* count)
{
    return guard([&] {
        if (n < 0 || !count || (n > 0 && (!points || !weights || !cvt_weights)))
            throw std::invalid_argument("CVT points, quadrature weights and CVT weights must have equal lengths");
        const auto sel = mpisdf::select_cvt(points, weights, cvt_weights, n, rank, seed, max_iter, tol, nullptr);
        if ((int64_t)sel.size() > capacity) throw std::invalid_argument("output capacity is too small");
        if (!sel.empty() && !out) throw std::invalid_argument("CVT output buffer is NULL");
        std::copy(sel.begin(), sel.end(), out);
        *count = (int64_t)sel.size();
    });
}

namespace {

int rbrp_impl(const mpisdf_context* ctx, const double* points, const double* ao_in, const double* weights,
              int64_t n, int64_t nao_in, int64_t batch_size, int64_t block_size, double tolerance, int64_t max_rank,
              uint64_t seed, double tau_block, double rank_rtol, int64_t refresh_every, double svd_rtol,
              int compute_id, int64_t* indices, int64_t capacity, int64_t* count, double* interp_vecs,
              double* transform, double* history, int64_t* nblocks, double* scalars, int32_t* status)
{
    return guard([&] {
        if (n < 0 || !count || !indices || !nblocks || !scalars || !status || (n > 0 && !weights))
            throw std::invalid_argument("invalid RBRP arguments");
        if (batch_size < 0 || block_size < 0 || max_rank < 0 || refresh_every < 0)
            throw std::invalid_argument("invalid RBRP dimensions, batch size or tolerances");
        std::vector<double> ao;
        int64_t nao = nao_in;
        const double* aop = ao_in;
        if (ctx) {
            auto& c = unwrap(ctx);
            nao = c.nao;
            if (n > 0 && !points) throw std::invalid_argument("expected points[grid,3] and weights[grid]");
            for (int64_t g = 0; g < n; ++g) {
                if (!std::isfinite(weights[g]) || weights[g] == 0)
                    throw std::invalid_argument("device RBRP requires finite nonzero weights");
                for (int d = 0; d < 3; ++d)
                    if (!std::isfinite(points[3 * g + d]))
                        throw std::invalid_argument("device RBRP requires finite coordinates");
            }
            ao.resize((size_t)(n * nao));
            if (n > 0) {
                mpisdf::DeviceBuffer<double> dxyz, dao((size_t)(n * nao));
                dxyz.upload(points, 3 * n, nullptr);
                mpisdf::launch_ao_eval(c, dxyz.ptr, n, mpisdf::kValues, mpisdf::kF64, dao.ptr, nao, n * nao, nullptr);
                MPISDF_HIP(hipMemcpy(ao.data(), dao.ptr, ao.size() * sizeof(double), hipMemcpyDeviceToHost));
            }
            aop = ao.data();
        } else if (n > 0 && !ao_in) {
            throw std::invalid_argument("expected ao[grid,nao] and weights[grid]");
        }
        mpisdf::RbrpOptions o;
        o.batch_size = batch_size;
        o.block_size = block_size;
        o.tolerance = tolerance;
        o.max_rank = max_rank;
        o.seed = seed;
        o.tau_block = tau_block;
        o.rank_rtol = rank_rtol;
        o.refresh_every = refresh_every;
        o.svd_rtol = svd_rtol;
        o.compute_id = compute_id != 0;
        const auto r = mpisdf::rbrp_ao(aop, nullptr, weights, n, nao, o, nullptr);
        const int64_t k = (int64_t)r.indices.size();
        if (k > capacity) throw std::invalid_argument("output capacity is too small");
        std::copy(r.indices.begin(), r.indices.end(), indices);
        *count = k;
        *nblocks = (int64_t)r.history.size() / 4;
        if (history) std::copy(r.history.begin(), r.history.end(), history);
        if (compute_id && interp_vecs) std::copy(r.interp_vecs.begin(), r.interp_vecs.end(), interp_vecs);
        if (compute_id && transform) std::copy(r.transform.begin(), r.transform.end(), transform);
        scalars[0] = r.relative_residual;
        scalars[1] = r.condition;
        scalars[2] = (double)r.retained_rank;
        static const char* names[] = {"rank_cap", "tolerance", "no_candidates", "numerical_rank", "roundoff_limited"};
        *status = 0;
        for (int32_t i = 0; i < 5; ++i)
            if (r.status == names[i]) *status = i;
    });
}

}  // namespace

extern "C" int mpisdf_rbrp_ao(const double* ao, const double* weights, int64_t n, int64_t nao, int64_t batch_size,
                              int64_t block_size, double tolerance, int64_t max_rank, uint64_t seed,
                              double tau_block, double rank_rtol, int64_t refresh_every, double svd_rtol,
                              int compute_id, int64_t* indices, int64_t capacity, int64_t* count,
                              double* interp_vecs, double* transform, double* history, int64_t* nblocks,
                              double* scalars, int32_t* status)
{
    return rbrp_impl(nullptr, nullptr, ao, weights, n, nao, batch_size, block_size, tolerance, max_rank, seed,
                     tau_block, rank_rtol, refresh_every, svd_rtol, compute_id, indices, capacity, count, interp_vecs,
                     transform, history, nblocks, scalars, status);
}

extern "C" int mpisdf_rbrp_points(const mpisdf_context* ctx, const double* points, const double* weights, int64_t n,
                                  int64_t batch_size, int64_t block_size, double tolerance, int64_t max_rank,
                                  uint64_t seed, double tau_block, double rank_rtol, int64_t refresh_every,
                                  double svd_rtol, int compute_id, int64_t* indices, int64_t capacity,
                                  int64_t* count, double* interp_vecs, double* transform, double* history,
                                  int64_t* nblocks, double* scalars, int32_t* status)
{
    if (!ctx) {
        mpisdf::set_last_error("null mp_isdf context");
        return mpisdf::kInvalid;
    }
    return rbrp_impl(ctx, points, nullptr, weights, n, 0, batch_size, block_size, tolerance, max_rank, seed,
                     tau_block, rank_rtol, refresh_every, svd_rtol, compute_id, indices, capacity, count, interp_vecs,
                     transform, history, nblocks, scalars, status);
}

/* ---- integral-direct Cholesky Coulomb metric ---------------------------- */

namespace {

mpisdf::CholeskyAlgorithm cholesky_algorithm(int32_t algorithm)
{
    if (algorithm != 0 && algorithm != 1) throw std::invalid_argument("unknown Cholesky algorithm");
    return algorithm == 1 ? mpisdf::CholeskyAlgorithm::UnblockedUniform : mpisdf::CholeskyAlgorithm::UnblockedGreedy;
}

}  // namespace

extern "C" int mpisdf_cholesky_eri(const mpisdf_context* ctx, double tolerance, int64_t max_rank, int32_t algorithm,
                                   uint64_t seed, int32_t debug, mpisdf_eri_cholesky** out)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (!out) throw std::invalid_argument("null Cholesky handle pointer");
        *out = nullptr;
        auto h = std::make_unique<mpisdf_eri_cholesky>();
        mpisdf::eri_cholesky(c, tolerance, max_rank, cholesky_algorithm(algorithm), seed, debug != 0, h->impl,
                             nullptr);
        MPISDF_HIP(hipStreamSynchronize(nullptr));
        *out = h.release();
    });
}

extern "C" int mpisdf_eri_cholesky_size(const mpisdf_eri_cholesky* h, int64_t* rank, int64_t* npair)
{
    return guard([&] {
        if (!h || !rank || !npair) throw std::invalid_argument("null Cholesky handle or output");
        *rank = h->impl.rank;
        *npair = h->impl.npair;
    });
}

extern "C" int mpisdf_eri_cholesky_copy(const mpisdf_eri_cholesky* h, double* factors, int64_t* pivots,
                                        double* residual)
{
    return guard([&] {
        if (!h) throw std::invalid_argument("null Cholesky handle");
        const auto& f = h->impl;
        if (factors && f.rank > 0)
            MPISDF_HIP(hipMemcpy(factors, f.factors.ptr, (size_t)(f.rank * f.npair) * sizeof(double),
                                 hipMemcpyDeviceToHost));
        if (pivots) std::copy(f.pivots.begin(), f.pivots.end(), pivots);
        if (residual)
            MPISDF_HIP(hipMemcpy(residual, f.residual.ptr, (size_t)f.npair * sizeof(double), hipMemcpyDeviceToHost));
    });
}

extern "C" int mpisdf_eri_cholesky_destroy(mpisdf_eri_cholesky* h)
{
    return guard([&] { delete h; });
}

extern "C" int mpisdf_coulomb_metric(const mpisdf_context* ctx, const double* transform, int64_t nvec,
                                     int32_t algorithm, double tolerance, int64_t max_rank, int64_t batch_size,
                                     int32_t debug, uint64_t seed, double* out)
{
    return guard([&] {
        auto& c = unwrap(ctx);
        if (nvec < 0 || (nvec > 0 && (!transform || !out))) throw std::invalid_argument("invalid Coulomb metric arguments");
        mpisdf::eri_coulomb_metric(c, transform, nvec, cholesky_algorithm(algorithm), tolerance, max_rank,
                                   batch_size, debug != 0, seed, out, nullptr);
    });
}
