// Anchor selection on the GPU.
//
// Mirrors the reference's AnchorSelection::select (cpp/src/anchor_select.cpp:79-173):
// validation, global multi-axis Frobenius scales, then one of
//   * QRColumnPivoting          cpp/src/anchor_select_qrcp.cpp:95-225
//   * OnlineQRColumnPivoting    cpp/src/anchor_select_online_qr.cpp:127-375
//   * RandomizedQuadratureKhatriRao  cpp/src/anchor_select_randomized_id.cpp:92-374
// with the same validation messages, rank/tolerance rules and returned
// index order (original grid indices, pivot order).  All grid work runs on the
// device; only pivots, norms and small vectors return to the host.
#include "select.hpp"
#include "audit.hpp"

#include "kernels.hpp"
#include "online.hpp"
#include "qrcp.hpp"
#include "sketch.hpp"
#include "sketch_mfma.hpp"
#include "util.hpp"

#include <algorithm>
#include <atomic>
#include <exception>
#include <chrono>
#include <cmath>
#include <iostream>
#include <limits>
#include <memory>
#include <random>
#include <thread>
#include <utility>

namespace mpisdf {
namespace {

using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

constexpr double kEps = std::numeric_limits<double>::epsilon();

struct Kept {
    std::vector<double> xyz, w, importance;
    std::vector<int64_t> original;
};

Kept filter_points(const SelectArgs& a, const double* importance = nullptr)
{
    Kept k;
    k.xyz.reserve(3 * a.n);
    for (int64_t i = 0; i < a.n; ++i) {
        if (!std::isfinite(a.weights[i])) throw std::invalid_argument("grid weights must be finite");
        if (a.weights[i] == 0 || (importance && importance[i] == 0)) continue;
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(a.points[3 * i + c])) throw std::invalid_argument("grid coordinates must be finite");
        k.xyz.insert(k.xyz.end(), a.points + 3 * i, a.points + 3 * i + 3);
        k.w.push_back(a.weights[i]);
        k.original.push_back(i);
        if (importance) k.importance.push_back(importance[i]);
    }
    return k;
}

// q[g] = sum_a f[g * d + a]^2 for the rows of a feature batch: one wave per
// row (coalesced reads), its lanes' sums added by a butterfly.
constexpr int kRowNormRows = 4;  // rows per 256-thread workgroup
__global__ __launch_bounds__(256) void row_norm2_kernel(const double* __restrict__ f, int64_t rows, int64_t d,
                                                        double* __restrict__ q)
{
    const int64_t g = (int64_t)blockIdx.x * kRowNormRows + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (g >= rows) return;
    double sum = 0;
    for (int64_t a = lane; a < d; a += 64) sum += f[g * d + a] * f[g * d + a];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if (lane == 0) q[g] = sum;
}

// Fixed-rank randomized selections on grids of this many kept points or more
// order their columns in levels of expected sketch norm and leave the far
// smaller ones unfolded in the QRCP (below that the extra pass does not pay).
constexpr int64_t kDormantColumns = int64_t(1) << 18;

// Device batch for feature evaluation (independent of the reference's
// host-memory batch_size, which does not change results).
int64_t device_batch(int64_t width, int64_t nao, int64_t extra_per_point)
{
    const int64_t per_point = 8 * (width + 4 * nao + extra_per_point);
    return std::max<int64_t>(1024, std::min<int64_t>(1 << 16, (int64_t(2) << 30) / std::max<int64_t>(per_point, 1)));
}

// Global per-axis scales (cpp/src/anchor_select.cpp:103-147).
std::vector<double> axis_scales(Context& ctx, const SelectArgs& a, const AxisSpec& spec, hipStream_t s)
{
    std::vector<double> scales;
    if (a.axes.size() <= 1) return scales;
    if (a.n > 0)
        for (int64_t g = 0; g < a.n; ++g)
            if (!std::isfinite(a.weights[g])) throw std::invalid_argument("grid weights must be finite");
    Kept kept = filter_points(a);
    const int64_t n = (int64_t)kept.original.size();
    if (n == 0) return scales;
    const int64_t batch = std::min(n, device_batch(spec.width, ctx.nao, 0));
    DeviceBuffer<double> d_xyz(3 * batch), feat(batch * spec.width), sums(spec.width), part;
    MPISDF_HIP(hipMemsetAsync(sums.ptr, 0, spec.width * sizeof(double), s));
    AxisWorkspace ws;
    for (int64_t start = 0; start < n; start += batch) {
        const int64_t count = std::min(batch, n - start);
        MPISDF_HIP(hipMemcpyAsync(d_xyz.ptr, kept.xyz.data() + 3 * start, 3 * count * sizeof(double),
                                  hipMemcpyHostToDevice, s));
        evaluate_axis_device(ctx, spec, d_xyz.ptr, count, nullptr, a.mo_coeff, a.nocc, feat.ptr, ws, s);
        launch_colsumsq(feat.ptr, count, spec.width, sums.ptr, part, s);
    }
    const auto h = sums.download(s);
    int64_t off = 0;
    for (size_t ax = 0; ax < spec.kinds.size(); ++ax) {
        double total = 0;
        for (int64_t j = 0; j < spec.widths[ax]; ++j) total += h[off + j];
        off += spec.widths[ax];
        const double norm = std::sqrt(total);
        if (!(norm > 0) || !std::isfinite(norm))
            throw std::invalid_argument("axis Frobenius norm must be finite and positive");
        const double scale = (a.axis_weights.empty() ? 1.0 : a.axis_weights[ax]) / norm;
        if (!(scale > 0) || !std::isfinite(scale))
            throw std::invalid_argument("normalized axis scale must be finite and positive");
        scales.push_back(scale);
    }
    return scales;
}

// ---------------------------------------------------------------------------
// qrcp_dense (cpp/src/anchor_select_qrcp.cpp:95-225)
// ---------------------------------------------------------------------------
std::vector<int64_t> select_dense(Context& ctx, const SelectArgs& a, const AxisSpec& spec,
                                  const std::vector<double>& scales, double* diag, hipStream_t s)
{
    if (!std::isfinite(a.tolerance) || a.tolerance < 0)
        throw std::invalid_argument("invalid weights, axis list, or tolerance");
    Kept kept = filter_points(a);
    const int64_t ng = (int64_t)kept.original.size();
    if (ng == 0) return {};
    const int64_t w = spec.width;
    if (w > (1 << 15) || ng > std::numeric_limits<int64_t>::max() / (w * w))
        throw std::overflow_error("pair-product matrix is too large");
    const int64_t m = w * w;
    DeviceBuffer<double> d_xyz, feat(ng * w), sw(ng), A((size_t)(m * ng));
    d_xyz.upload(kept.xyz.data(), kept.xyz.size(), s);
    std::vector<double> hsw(ng);
    for (int64_t g = 0; g < ng; ++g) hsw[g] = std::sqrt(std::abs(kept.w[g]));
    sw.upload(hsw.data(), ng, s);
    AxisWorkspace ws;
    const int64_t batch = std::min(ng, device_batch(w, ctx.nao, 0));
    for (int64_t start = 0; start < ng; start += batch) {
        const int64_t count = std::min(batch, ng - start);
        evaluate_axis_device(ctx, spec, d_xyz.ptr + 3 * start, count, scales.empty() ? nullptr : scales.data(),
                             a.mo_coeff, a.nocc, feat.ptr + start * w, ws, s);
    }
    DeviceBuffer<unsigned long long> bad(1);
    MPISDF_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(unsigned long long), s));
    launch_pair_product(feat.ptr, ng, w, sw.ptr, A.ptr, bad.ptr, s);
    if (bad.download(s)[0]) throw std::overflow_error("non-finite weighted pair product");
    const int64_t cap = std::min({m, ng, a.max_rank ? a.max_rank : std::min(m, ng)});
    QrcpOptions opt;
    opt.kmax = cap;
    opt.stop_rel = a.tolerance;
    opt.candidates = std::min<int64_t>(ng, 2048);
    opt.debug = a.debug;
    const auto t0 = Clock::now();
    const QrcpResult r = qrcp_device(A.ptr, kF64, m, ng, m, opt, s);
    std::vector<int64_t> out;
    for (int64_t p : r.pivots) {
        if (p < 0 || p >= ng) throw std::runtime_error("QRCP returned an invalid grid pivot");
        out.push_back(kept.original[p]);
    }
    if (diag) {
        diag[kDiagRows] = (double)m;
        diag[kDiagColumns] = (double)ng;
        diag[kDiagCap] = (double)cap;
        diag[kDiagRank] = (double)out.size();
        diag[kDiagBlocks] = (double)r.blocks;
        diag[kDiagFolds] = (double)r.folds;
        diag[kDiagFoldSeconds] = r.fold_seconds;
        diag[kDiagFoldFlops] = r.fold_flops;
        diag[kDiagFoldLaunches] = (double)r.folds;
        diag[kDiagGsSeconds] = r.gs_seconds;
        diag[kDiagGsFlops] = r.gs_flops;
        diag[kDiagSecondsQrcp] = seconds_since(t0);
    }
    return out;
}

// Dense audit of the selected pivots (cpp/src/anchor_select_randomized_id.cpp:19-88):
// the weighted pair matrix of the retained grid (column g scaled by col[g];
// the full d^2 rows have the packed matrix's inner products), its first
// rank(selected) dense QRCP pivots, and the relative projection residual of
// both pivot sets, printed as the reference prints them.  Returns
// (selected, dense).
std::pair<double, double> audit_projection(Context& ctx, const SelectArgs& a, const AxisSpec& spec,
                                           const std::vector<double>& scales, const double* d_xyz, int64_t n,
                                           const std::vector<double>& col, const std::vector<int64_t>& selected,
                                           const char* objective, hipStream_t s)
{
    const int64_t d = spec.width, m = d * d;
    DeviceBuffer<double> feat((size_t)(n * d)), sw, A((size_t)(m * n));
    AxisWorkspace ws;
    const int64_t batch = std::min(n, device_batch(d, ctx.nao, 0));
    for (int64_t start = 0; start < n; start += batch) {
        const int64_t count = std::min(batch, n - start);
        evaluate_axis_device(ctx, spec, d_xyz + 3 * start, count, scales.empty() ? nullptr : scales.data(),
                             a.mo_coeff, a.nocc, feat.ptr + start * d, ws, s);
    }
    sw.upload(col.data(), n, s);
    DeviceBuffer<unsigned long long> bad(1);
    MPISDF_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(unsigned long long), s));
    launch_pair_product(feat.ptr, n, d, sw.ptr, A.ptr, bad.ptr, s);
    if (bad.download(s)[0]) throw std::overflow_error("non-finite weighted pair product in the audit");
    const int64_t k = (int64_t)selected.size();
    std::vector<int64_t> dense;
    if (k > 0) {
        QrcpOptions opt;
        opt.kmax = std::min({k, m, n});
        opt.candidates = std::min<int64_t>(n, 2048);
        dense = qrcp_device(A.ptr, kF64, m, n, m, opt, s).pivots;
    }
    const double actual = projection_relative_error(A.ptr, m, n, selected, s);
    const double ref = projection_relative_error(A.ptr, m, n, dense, s);
    std::cerr << "[randomized id audit] objective=" << objective << " rank=" << k
              << " actual_projection_relative_error=" << actual << " dense_qrcp_projection_relative_error=" << ref
              << " (equal rank; not ID interpolation error)\n";
    return {actual, ref};
}

// ---------------------------------------------------------------------------
// rand_sketch with the quadrature and density policies
// (cpp/src/anchor_select_randomized_id.cpp:92-374)
// ---------------------------------------------------------------------------
std::vector<int64_t> select_randomized(Context& ctx, const SelectArgs& a, const AxisSpec& spec,
                                       const std::vector<double>& scales, double* diag, hipStream_t s)
{
    const auto started = Clock::now();
    if (!std::isfinite(a.tolerance) || a.tolerance < 0 || a.batch_size == 0 || a.oversampling == 0 ||
        (!a.max_rank && a.sketch_size <= a.oversampling))
        throw std::invalid_argument(
            "invalid randomized ID grid, tolerance, batch, sketch or rank/oversampling");
    if (a.max_rank > std::numeric_limits<int64_t>::max() - a.oversampling)
        throw std::overflow_error("randomized ID rank + oversampling overflows size_t");
    const int64_t requested = a.max_rank ? a.max_rank + a.oversampling : a.sketch_size;
    const int64_t d = spec.width;
    // Density policies (randomized_id.cpp:102-115, :139, :231-260, :290-293):
    // importance replaces |w| as the column weight, points with zero
    // importance or an AO squared norm at or below the screening threshold
    // are not candidates, and the normalized policy divides each pair
    // column by that squared norm.
    const bool density = a.algorithm == kDensity || a.algorithm == kDensityNormalized;
    const bool normalized = a.algorithm == kDensityNormalized;
    if (density) {
        if (!a.importance) throw std::invalid_argument("point_importance must have shape (ngrid,)");
        for (int64_t i = 0; i < a.n; ++i)
            if (!std::isfinite(a.importance[i]) || a.importance[i] < 0)
                throw std::invalid_argument("point_importance must be finite and nonnegative");
        if (!std::isfinite(a.ao_screening_threshold) || a.ao_screening_threshold < 0)
            throw std::invalid_argument("AO screening threshold must be finite and nonnegative");
        if (a.axes.size() != 1 || a.axes[0] != 0)
            throw std::invalid_argument("density Khatri-Rao sketches with non-AO axis features are NYI");
    }
    Kept kept = filter_points(a, density ? a.importance : nullptr);
    int64_t n = (int64_t)kept.original.size();
    if (a.audit && d * (d + 1) / 2 > 2000000 / std::max<int64_t>(1, n))
        throw std::invalid_argument("randomized ID dense audit is limited to 2 million pair-matrix elements");
    if (n == 0) {
        if (density && a.max_rank) throw std::invalid_argument("requested anchor_rank exceeds eligible candidates");
        return {};
    }
    const int64_t physical_rank = d * (d + 1) / 2;
    const int64_t physical_limit = physical_rank + a.oversampling;
    const int64_t backend_limit = std::numeric_limits<int64_t>::max() / n;
    const int64_t ell = requested <= backend_limit ? requested
                                                   : std::min({requested, physical_limit, backend_limit});
    if (ell <= a.oversampling || (a.max_rank && std::min(a.max_rank, physical_rank) > ell - a.oversampling))
        throw std::invalid_argument(
            "randomized ID sketch/rank exceeds the ILP64 QRCP backend limit for this grid; "
            "reduce the sketch/rank or use a tiled QRCP backend");
    // Omega is drawn on host threads (it depends on d, ell and the seed only)
    // while the GPU screens and orders the points.
    std::vector<double> hu, hv;
    struct OmegaThread {
        std::thread t;
        std::exception_ptr err;
        void join()
        {
            if (t.joinable()) t.join();
            if (err) std::rethrow_exception(std::exchange(err, nullptr));
        }
        ~OmegaThread()
        {
            if (t.joinable()) t.join();
        }
    } omega;
    omega.t = std::thread([&hu, &hv, &omega, d, ell, seed = a.seed] {
        try {
            make_omega(d, ell, seed, hu, hv);
        } catch (...) {
            omega.err = std::current_exception();
        }
    });
    std::vector<double> ao_norm2;
    if (density) {
        // Screening pass: only O(n) squared norms come back to the host.
        DeviceBuffer<double> d_all, q(n);
        d_all.upload(kept.xyz.data(), kept.xyz.size(), s);
        const int64_t batch = std::min(n, device_batch(d, ctx.nao, 0));
        DeviceBuffer<double> feat((size_t)batch * d);
        AxisWorkspace ws0;
        for (int64_t start = 0; start < n; start += batch) {
            const int64_t count = std::min(batch, n - start);
            evaluate_axis_device(ctx, spec, d_all.ptr + 3 * start, count, scales.empty() ? nullptr : scales.data(),
                                 a.mo_coeff, a.nocc, feat.ptr, ws0, s);
            hipLaunchKernelGGL(row_norm2_kernel, dim3((unsigned)((count + kRowNormRows - 1) / kRowNormRows)), dim3(256), 0, s, feat.ptr,
                               count, d, q.ptr + start);
            MPISDF_HIP(hipGetLastError());
        }
        const auto hq = q.download(s);
        Kept screened;
        for (int64_t g = 0; g < n; ++g) {
            if (!std::isfinite(hq[g])) throw std::overflow_error("non-finite AO squared norm");
            if (hq[g] <= a.ao_screening_threshold) continue;
            screened.xyz.insert(screened.xyz.end(), kept.xyz.begin() + 3 * g, kept.xyz.begin() + 3 * g + 3);
            screened.w.push_back(kept.w[g]);
            screened.importance.push_back(kept.importance[g]);
            screened.original.push_back(kept.original[g]);
            ao_norm2.push_back(hq[g]);
        }
        kept = std::move(screened);
        n = (int64_t)kept.original.size();
    }
    if (a.max_rank > std::min(n, physical_rank))
        throw std::invalid_argument("requested anchor_rank exceeds eligible candidates or physical pair rank");
    if (n == 0) return {};
    if (a.precision == kPrecF32)
        throw std::invalid_argument("precision 'fp32' is not provided; use fp64, bf16x3, bf16 or fp8");
    const bool lowp = a.precision != kPrecF64 && a.precision != kPrecF64I8;
    // The FP64 sketch on four base-254 digit planes: pitch a multiple of 64,
    // ell within the FP64 digitiser's register budget.
    const bool digits5 = a.precision == kPrecF64I8;
    if (digits5 && ell > kDigitRowsMaxF64)
        throw std::invalid_argument("precision 'fp64i8' supports sketches of at most 16384 rows");
    // MFMA operand form of the reduced-precision sketch (sketch_mfma.hpp):
    // BF16, two-term FP8 or split BF16 (hi.hi + lo.hi + hi.lo).
    const int form = a.precision == kPrecFP8 ? kSketchFP8 : a.precision == kPrecBF16x3 ? kSketchBF16x3 : kSketchBF16;
    const DType in_type = form == kSketchFP8 ? kFP8 : kBF16;
    // The split-BF16 sketch is accurate to ~1e-7 of each column, so its digit
    // form keeps four planes (three would round it to ~3e-5).
    const int planes = form == kSketchBF16x3 ? 4 : 3;

    // Columns in levels of their expected sketch norm E||y_g||^2 = c_g^2
    // ||f_g||^4 (c_g the column weight below, f_g the point's features), in
    // factors of four down from the largest, grid order kept within a level:
    // the QRCP then leaves the trailing levels unfolded while they stay far
    // below its pivots (QrcpOptions::dormant), and ties between equal norms
    // still go to the lower grid position (QrcpOptions::tie_rank).  Fixed
    // rank only, on grids large enough for the pass to pay.
    DeviceBuffer<int64_t> tie_rank;
    const bool dormant = a.max_rank > 0 && n >= kDormantColumns && !a.audit;
    if (dormant) {
        DeviceBuffer<double> d_all, q(n);
        d_all.upload(kept.xyz.data(), kept.xyz.size(), s);
        const int64_t batch = std::min(n, device_batch(d, ctx.nao, 0));
        DeviceBuffer<double> feat((size_t)batch * d);
        AxisWorkspace ws0;
        for (int64_t start = 0; start < n; start += batch) {
            const int64_t count = std::min(batch, n - start);
            evaluate_axis_device(ctx, spec, d_all.ptr + 3 * start, count, scales.empty() ? nullptr : scales.data(),
                                 a.mo_coeff, a.nocc, feat.ptr, ws0, s);
            hipLaunchKernelGGL(row_norm2_kernel, dim3((unsigned)((count + kRowNormRows - 1) / kRowNormRows)), dim3(256), 0, s, feat.ptr,
                               count, d, q.ptr + start);
            MPISDF_HIP(hipGetLastError());
        }
        const auto hq = q.download(s);
        constexpr int kLevels = 48;
        std::vector<double> est(n);
        std::vector<int> level(n);
        double emax = 0;
        for (int64_t g = 0; g < n; ++g) {
            double c2 = density ? kept.importance[g] : std::abs(kept.w[g]);
            if (normalized) c2 /= ao_norm2[g] * ao_norm2[g];
            est[g] = c2 * hq[g] * hq[g];
            if (std::isfinite(est[g])) emax = std::max(emax, est[g]);
        }
        std::vector<int64_t> count(kLevels + 1, 0);
        const int top = emax > 0 ? std::ilogb(emax) : 0;
        for (int64_t g = 0; g < n; ++g) {
            const double e = est[g];
            int lv = kLevels - 1;  // zero or non-finite estimates last
            if (e > 0 && std::isfinite(e)) lv = std::min(kLevels - 1, std::max(0, (top - std::ilogb(e)) / 2));
            level[g] = lv;
            ++count[lv + 1];
        }
        for (int l = 0; l < kLevels; ++l) count[l + 1] += count[l];
        std::vector<int64_t> order(n);  // new position -> position in grid order
        for (int64_t g = 0; g < n; ++g) order[count[level[g]]++] = g;
        Kept sorted;
        sorted.xyz.resize(kept.xyz.size());
        sorted.w.resize(n);
        sorted.importance.resize(kept.importance.size());
        sorted.original.resize(n);
        std::vector<double> sorted_norm2(ao_norm2.size());
        for (int64_t j = 0; j < n; ++j) {
            const int64_t g = order[j];
            for (int c = 0; c < 3; ++c) sorted.xyz[3 * j + c] = kept.xyz[3 * g + c];
            sorted.w[j] = kept.w[g];
            if (!kept.importance.empty()) sorted.importance[j] = kept.importance[g];
            sorted.original[j] = kept.original[g];
            if (!ao_norm2.empty()) sorted_norm2[j] = ao_norm2[g];
        }
        kept = std::move(sorted);
        ao_norm2 = std::move(sorted_norm2);
        tie_rank.upload(order.data(), n, s);
    }

    omega.join();
    DeviceBuffer<double> u, v;
    u.upload(hu.data(), hu.size(), s);
    v.upload(hv.data(), hv.size(), s);

    DeviceBuffer<double> d_xyz, sw(n);
    d_xyz.upload(kept.xyz.data(), kept.xyz.size(), s);
    std::vector<double> hsw(n);
    for (int64_t g = 0; g < n; ++g) {
        // (phi/sqrt(q) . u)(phi/sqrt(q) . v) = (phi . u)(phi . v) / q
        hsw[g] = density ? std::sqrt(kept.importance[g]) : std::sqrt(std::abs(kept.w[g]));
        if (normalized) hsw[g] /= ao_norm2[g];
    }
    sw.upload(hsw.data(), n, s);
    DeviceBuffer<unsigned long long> bad(1);
    MPISDF_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(unsigned long long), s));
    // Y lives in the context's sketch storage (held between selections).
    std::unique_lock<std::mutex> store_lock(ctx.sketch_mutex);
    struct { double* ptr = nullptr; } y64;
    struct { float* ptr = nullptr; } y32;
    // The reduced-precision forms fold on INT8 digit planes of the FP32 sketch
    // (three or four digits per entry in place of each column), which needs a
    // pitch that is a multiple of 64 and ell within the digit kernels' limits.
    const bool use_digits = lowp && ell <= kDigitRowsMax;
    const int64_t ldy32 = use_digits ? (ell + kDigitSlice - 1) / kDigitSlice * kDigitSlice
                                     : (ell + 3) / 4 * 4;  // 256- or 16-byte aligned FP32 columns
    const int64_t ldy64 = digits5 ? (ell + kDigitSlice - 1) / kDigitSlice * kDigitSlice
                                  : (ell + 3) / 4 * 4;  // 32-byte aligned FP64 columns
    DeviceBuffer<double> ynorm2;
    AxisWorkspace ws;
    double t_sketch_kernel = 0, sketch_launches = 0;
    ynorm2.resize(n);
    MPISDF_HIP(hipMemsetAsync(ynorm2.ptr, 0, n * sizeof(double), s));
    if (digits5) {
        // The FP64 sketch from digit products: the rows of U and V (stacked,
        // 64 of U then the same 64 of V) and each point's features as four
        // base-254 digits under a power of two (within ~1.2e-10 of the row's
        // or point's largest entry); the ten digit-pair products p + r <= 3 run
        // exactly in INT32 on INT8 MFMA and are combined in FP64, and the
        // kernel's epilogue forms y = sqrt(h / ell) (u . z)(v . z), ||y||^2 and
        // the non-finite count as the FP64 sketch does.
        y64.ptr = static_cast<double*>(ctx.reserve_sketch((size_t)ldy64 * n * sizeof(double)));
        const int64_t dpad = (d + kDigitSlice - 1) / kDigitSlice * kDigitSlice;
        FoldI8Rows uv;
        split_khatri_rao_rows_i8(u.ptr, v.ptr, d, ell, uv, s);
        // Every 128-row group re-reads the batch's feature digits: batches of
        // 16,384 points keep them (~130 MB at 1536 features) in the
        // last-level cache across the groups.
        const int64_t batch = std::min({n, device_batch(d, ctx.nao, 0), int64_t(16384)});
        // Two batch buffers: the features of batch i + 1 and their digits are
        // prepared on a second stream while batch i's digit products run.
        DeviceBuffer<double> feat[2], featp[2], fscale[2], fnorm[2], cscale(n);
        for (int b = 0; b < 2; ++b) {
            feat[b].resize((size_t)batch * d);
            fscale[b].resize(batch);
            fnorm[b].resize(batch);
            if (dpad != d) {
                featp[b].resize((size_t)batch * dpad);
                MPISDF_HIP(hipMemsetAsync(featp[b].ptr, 0, featp[b].count * sizeof(double), s));  // pad columns stay zero
            }
        }
        {
            std::vector<double> hc(n);
            const double inv = 1.0 / std::sqrt((double)ell);
            for (int64_t g = 0; g < n; ++g) hc[g] = hsw[g] * inv;
            cscale.upload(hc.data(), n, s);
            MPISDF_HIP(hipStreamSynchronize(s));
        }
        HipStream prep;
        HipEvent ready[2], consumed[2];
        const int64_t nbatch = (n + batch - 1) / batch;
        auto digits_of = [&](int b) { return dpad != d ? featp[b].ptr : feat[b].ptr; };
        auto prepare = [&](int64_t i) {
            const int b = (int)(i & 1);
            const int64_t start = i * batch, count = std::min(batch, n - start);
            if (i >= 2) consumed[b].wait(prep.s);  // batch i - 2's products have read this buffer
            evaluate_axis_device(ctx, spec, d_xyz.ptr + 3 * start, count, scales.empty() ? nullptr : scales.data(),
                                 a.mo_coeff, a.nocc, feat[b].ptr, ws, prep.s);
            if (dpad != d)
                MPISDF_HIP(hipMemcpy2DAsync(featp[b].ptr, dpad * sizeof(double), feat[b].ptr, d * sizeof(double),
                                            d * sizeof(double), count, hipMemcpyDeviceToDevice, prep.s));
            digitize_columns_f64_max(digits_of(b), d, count, dpad, fnorm[b].ptr, fscale[b].ptr, prep.s, kSketchDigits254);
            ready[b].record(prep.s);
        };
        LaunchTimer timer;
        prepare(0);
        for (int64_t i = 0; i < nbatch; ++i) {
            if (i + 1 < nbatch) prepare(i + 1);
            const int b = (int)(i & 1);
            const int64_t start = i * batch, count = std::min(batch, n - start);
            ready[b].wait(s);
            timer.begin(s);
            fold_rows_khatri_rao_i8(uv, reinterpret_cast<const int8_t*>(digits_of(b)), dpad, fscale[b].ptr, count,
                                    cscale.ptr + start, y64.ptr + start * ldy64, ldy64, ynorm2.ptr + start, bad.ptr,
                                    s);
            timer.end(s);
            consumed[b].record(s);
        }
        MPISDF_HIP(hipStreamSynchronize(s));
        t_sketch_kernel = timer.seconds();
        sketch_launches = (double)timer.launches();
    } else if (!lowp) {
        y64.ptr = static_cast<double*>(ctx.reserve_sketch((size_t)ldy64 * n * sizeof(double)));
        const int64_t batch = std::min(n, device_batch(d, ctx.nao, 0));
        DeviceBuffer<double> feat((size_t)batch * d);
        SketchScratch scratch;
        for (int64_t start = 0; start < n; start += batch) {
            const int64_t count = std::min(batch, n - start);
            evaluate_axis_device(ctx, spec, d_xyz.ptr + 3 * start, count, scales.empty() ? nullptr : scales.data(),
                                 a.mo_coeff, a.nocc, feat.ptr, ws, s);
            sketch_f64(feat.ptr, count, d, u.ptr, v.ptr, ell, sw.ptr + start, y64.ptr + start * ldy64, ldy64, scratch,
                       bad.ptr, s, ynorm2.ptr + start);
        }
        MPISDF_HIP(hipStreamSynchronize(s));
        t_sketch_kernel = scratch.timer.seconds();
        sketch_launches = (double)scratch.timer.launches();
    } else {
        // Reduced precision: operand segments padded to a multiple of 16
        // features, sketch_segments(form) of them along K.
        const int64_t dp = (d + 15) / 16 * 16;
        const int64_t dk = sketch_segments(form) * dp;
        const size_t esz = in_type == kFP8 ? 1 : 2;
        DeviceBuffer<uint8_t> ut((size_t)ell * dk * esz), vt((size_t)ell * dk * esz);
        prepare_omega_lowp(u.ptr, v.ptr, d, ell, dp, form, ut.ptr, vt.ptr, s);
        y32.ptr = static_cast<float*>(ctx.reserve_sketch((size_t)ldy32 * n * sizeof(float)));
        const bool ao_only = spec.kinds.size() == 1 && spec.kinds[0] == 0;
        LaunchTimer timer;
        const int64_t batch = std::min<int64_t>(n, 1 << 16);
        DeviceBuffer<uint8_t> phi((size_t)batch * dk * esz);
        MPISDF_HIP(hipMemsetAsync(phi.ptr, 0, phi.count, s));  // padding columns stay zero
        DeviceBuffer<float> phi32, scale(batch);
        DeviceBuffer<double> ynorm_part;
        DeviceBuffer<double> feat;
        if (!ao_only || form == kSketchBF16x3) feat.resize((size_t)batch * d);
        if (in_type == kFP8) phi32.resize((size_t)batch * d);
        for (int64_t start = 0; start < n; start += batch) {
            const int64_t count = std::min(batch, n - start);
            const double* xyz = d_xyz.ptr + 3 * start;
            if (form == kSketchBF16x3) {
                if (ao_only) launch_ao_eval(ctx, xyz, count, kValues, kF64, feat.ptr, d, count * d, s);
                else
                    evaluate_axis_device(ctx, spec, xyz, count, scales.empty() ? nullptr : scales.data(), a.mo_coeff,
                                         a.nocc, feat.ptr, ws, s);
                split_rows_bf16x3(feat.ptr, d, count, d, dp, phi.ptr, s);
                bf16_point_scales(sw.ptr + start, count, ell, scale.ptr, s);
            } else if (in_type == kBF16) {
                if (ao_only) launch_ao_eval(ctx, xyz, count, kValues, kBF16, phi.ptr, dp, count * dp, s);
                else {
                    evaluate_axis_device(ctx, spec, xyz, count, scales.empty() ? nullptr : scales.data(), a.mo_coeff,
                                         a.nocc, feat.ptr, ws, s);
                    cast_rows(feat.ptr, count, d, dp, kBF16, phi.ptr, s);
                }
                bf16_point_scales(sw.ptr + start, count, ell, scale.ptr, s);
            } else {
                if (ao_only) launch_ao_eval(ctx, xyz, count, kValues, kF32, phi32.ptr, d, count * d, s);
                else {
                    evaluate_axis_device(ctx, spec, xyz, count, scales.empty() ? nullptr : scales.data(), a.mo_coeff,
                                         a.nocc, feat.ptr, ws, s);
                    cast_rows(feat.ptr, count, d, d, kF32, phi32.ptr, s);
                }
                quantize_fp8_rows(phi32.ptr, d, count, d, sw.ptr + start, ell, phi.ptr, dp, scale.ptr, s);
            }
            timer.begin(s);
            kr_sketch_lowp(in_type, ut.ptr, vt.ptr, ell, dk, phi.ptr, dk, count, scale.ptr, y32.ptr + start * ldy32,
                           ldy32, ynorm2.ptr + start, bad.ptr, ynorm_part, s);
            timer.end(s);
        }
        MPISDF_HIP(hipStreamSynchronize(s));
        t_sketch_kernel = timer.seconds();
        sketch_launches = (double)timer.launches();
    }
    if (bad.download(s)[0]) throw std::overflow_error("non-finite randomized ID sketch");
    DeviceBuffer<double> yscale;
    if (use_digits) {
        yscale.resize(n);
        digitize_columns(y32.ptr, ell, n, ldy32, planes, ynorm2.ptr, yscale.ptr, s);
    } else if (digits5) {
        yscale.resize(n);
        digitize_columns_f64(y64.ptr, ell, n, ldy64, ynorm2.ptr, yscale.ptr, s);
    }
    const double t_sketch = seconds_since(started);

    const int64_t steps = std::min(ell, n);
    const int64_t structural = std::min({ell - a.oversampling, steps, physical_rank});
    QrcpOptions opt;
    opt.kmax = a.max_rank ? std::min(a.max_rank, structural) : structural;
    // Numerical-rank floor eps * max(ell, n) * max|R_ii| (|R_00| is the max).
    opt.stop_rel = kEps * (double)std::max(ell, n);
    // Candidates per block: fewer make each block's Gram-Schmidt passes and
    // per-pivot updates cheaper but end blocks sooner; 1024 is the measured
    // optimum on (H2O)64 (QRCP 7.7 s; 8.4 s with 2048, 8.8 s with 512).
    opt.candidates = std::min<int64_t>(n, 1024);
    opt.want_energy = a.max_rank == 0;
    opt.want_tail = a.max_rank == 0;
    opt.debug = a.debug;
    opt.dormant = dormant;
    opt.tie_rank = dormant ? tie_rank.ptr : nullptr;
    const auto tq = Clock::now();
    opt.ynorm2 = ynorm2.ptr;
    // Certifying greedy pivots late in the factorisation needs folded norms
    // accurate to ~1e-9 of ||y_j||^2.  The FP64 form folds with FP64 MFMA; the
    // reduced-precision forms fold the digit planes of the FP32 sketch with
    // INT8 MFMA (exact digit products, ~1e-11 of ||y_j||^2 per fold, covered
    // by the margin) and factorise y_hat, which equals the FP32 sketch to
    // 2^-29 of each column's scale.
    if (use_digits || digits5) {
        opt.yscale = yscale.ptr;
        opt.margin = kFoldI8Margin;
    }
    const QrcpResult r = use_digits ? qrcp_device(y32.ptr, planes == 4 ? kI8Digits4 : kI8Digits, ell, n, ldy32, opt, s)
                         : digits5  ? qrcp_device(y64.ptr, kI8Digits254, ell, n, ldy64, opt, s)
                         : lowp     ? qrcp_device(y32.ptr, kF32, ell, n, ldy32, opt, s)
                                    : qrcp_device(y64.ptr, kF64, ell, n, ldy64, opt, s);
    const double t_qrcp = seconds_since(tq);
    const int64_t K = (int64_t)r.pivots.size();
    // numerical_rank: first i with |R_ii| <= floor; the engine stops there.
    const int64_t numerical_rank = K < opt.kmax ? K : steps;
    const int64_t cap = std::min({a.max_rank ? a.max_rank : ell - a.oversampling, ell - a.oversampling, steps,
                                  physical_rank, numerical_rank});
    if (a.max_rank && numerical_rank < a.max_rank)
        throw std::invalid_argument("requested anchor_rank exceeds the numerical rank of the randomized sketch");
    int64_t rank = a.max_rank ? a.max_rank : cap;
    double norm = std::sqrt(r.frob2), rel_tail = 0;
    if (!a.max_rank) {
        // tails[i]^2 = tail2 + sum_{i' >= i} energy[i'] (row norms of R).
        std::vector<double> tails2(K + 1, 0.0);
        tails2[K] = r.tail2;
        for (int64_t i = K; i-- > 0;) tails2[i] = tails2[i + 1] + r.energy[i];
        norm = std::sqrt(tails2[0]);
        if (!std::isfinite(norm)) throw std::overflow_error("non-finite randomized ID QR residual");
        for (int64_t i = 0; i <= std::min(cap, K); ++i)
            if (std::sqrt(tails2[i]) <= a.tolerance * norm) { rank = i; break; }
        rank = std::min(rank, K);
        rel_tail = norm > 0 ? std::sqrt(tails2[rank]) / norm : 0.0;
    }
    std::vector<int64_t> out;
    for (int64_t i = 0; i < rank; ++i) {
        if (i >= K || r.pivots[i] < 0 || r.pivots[i] >= n)
            throw std::runtime_error("randomized ID QRCP returned invalid pivot");
        out.push_back(kept.original[r.pivots[i]]);
    }
    if (a.audit) {
        // Positions of the selected anchors within the retained grid.
        const std::vector<int64_t> local(r.pivots.begin(), r.pivots.begin() + rank);
        std::vector<double> quad(n);
        for (int64_t g = 0; g < n; ++g) quad[g] = std::sqrt(std::abs(kept.w[g]));
        const auto q = audit_projection(ctx, a, spec, scales, d_xyz.ptr, n, quad, local, "physical_quadrature", s);
        if (diag) {
            diag[kDiagAuditActual] = q.first;
            diag[kDiagAuditDense] = q.second;
        }
        if (density) {
            const auto e = audit_projection(ctx, a, spec, scales, d_xyz.ptr, n, hsw, local,
                                            normalized ? "density_normalized" : "density_raw", s);
            if (diag) {
                diag[kDiagAuditDensityActual] = e.first;
                diag[kDiagAuditDensityDense] = e.second;
            }
        }
    }
    if (a.debug)
        std::cerr << "[rand_sketch:quadrature] rank=" << rank << " sketch_size=" << ell
                  << " oversampling=" << a.oversampling << " columns=" << n
                  << " sketch_relative_frobenius=" << rel_tail << " numerical_rank=" << numerical_rank
                  << " rank_cap=" << cap << " seed=" << a.seed << " space=device blocks=" << r.blocks
                  << " folds=" << r.folds << " elapsed=" << seconds_since(started) << " s\n";
    if (diag) {
        diag[kDiagRows] = (double)ell;
        diag[kDiagColumns] = (double)n;
        diag[kDiagNumericalRank] = (double)numerical_rank;
        diag[kDiagCap] = (double)cap;
        diag[kDiagRank] = (double)rank;
        diag[kDiagNorm] = norm;
        diag[kDiagRelTail] = rel_tail;
        diag[kDiagBlocks] = (double)r.blocks;
        diag[kDiagFolds] = (double)r.folds;
        diag[kDiagSecondsSketch] = t_sketch;
        diag[kDiagSecondsQrcp] = t_qrcp;
        diag[kDiagStalls] = (double)r.stalls;
        diag[kDiagRefreshes] = (double)r.refreshes;
        diag[kDiagFoldSeconds] = r.fold_seconds;
        diag[kDiagFoldFlops] = r.fold_flops;
        diag[kDiagFoldLaunches] = (double)r.folds;
        diag[kDiagGsSeconds] = r.gs_seconds;
        diag[kDiagGsFlops] = r.gs_flops;
        diag[kDiagSketchSeconds] = t_sketch_kernel;
        // MFMA work of the sketch: 4 ell d n per product term (three for the
        // split-BF16 form, two for the two-term FP8 form).
        diag[kDiagSketchFlops] = 4.0 * (double)ell * (double)d * (double)n * (lowp ? sketch_segments(form) : 1);
        // fp64i8: the INT8 operations, ten digit pairs x 2 (2 ell64) dpad per point
        if (digits5)
            diag[kDiagSketchOps] = 10.0 * 2.0 * (2.0 * (double)ldy64) * (double)((d + 63) / 64 * 64) * (double)n;
        diag[kDiagSketchLaunches] = sketch_launches;
        diag[kDiagFoldKind] = digits5 ? 4.0 : use_digits ? (planes == 4 ? 3.0 : 2.0) : (lowp ? 1.0 : 0.0);
    }
    return out;
}

}  // namespace

namespace {

// Raw mt19937_64 outputs, replayed in order while a producer thread is still
// recording them (a read waits until its draw is published).  Past the
// recorded range it returns 3/4 of the range, a draw the polar method accepts,
// so a call that runs off the end terminates (its value is discarded).
struct ReplayUrng {
    using result_type = std::mt19937_64::result_type;
    const result_type* raw;
    size_t size, pos;
    const std::atomic<size_t>* ready;
    static constexpr result_type min() { return std::mt19937_64::min(); }
    static constexpr result_type max() { return std::mt19937_64::max(); }
    result_type operator()()
    {
        const size_t i = pos++;
        if (i >= size) return result_type(3) << 62;
        while (ready->load(std::memory_order_acquire) <= i) std::this_thread::yield();
        return raw[i];
    }
};

// out[0 .. count) = the first count values of std::normal_distribution<double>(0, 1)
// over std::mt19937_64(seed), exactly as one sequential loop would draw them.
//
// libstdc++ draws normals by the polar method: every attempt consumes two
// uniforms (one raw 64-bit draw each for mt19937_64 and double), and each
// accepted attempt yields two values, the second held for the next call.
// Attempts are independent, so the calling thread records the raw stream
// (the only sequential part) while workers run the expensive part (log,
// sqrt) on chunks of attempts: a fresh distribution replaying from an attempt
// boundary yields exactly the values of the accepted attempts that start
// inside its chunk.
void fill_normal(uint64_t seed, int64_t count, double* out, int workers)
{
    std::mt19937_64 gen(seed);
    int64_t produced = 0;
    while (produced < count) {
        const int64_t want = (count - produced + 1) / 2;  // accepted attempts still needed
        const int64_t attempts = want + want / 3 + 4096;  // acceptance is pi/4 ~ 0.785
        const size_t nraw = (size_t)(2 * attempts);
        std::unique_ptr<std::mt19937_64::result_type[]> raw(new std::mt19937_64::result_type[nraw]);
        std::atomic<size_t> ready{0};
        constexpr int64_t chunk = 1 << 16;
        const int64_t nchunks = (attempts + chunk - 1) / chunk;
        std::vector<std::vector<double>> parts((size_t)nchunks);
        std::atomic<int64_t> next{0};
        auto work = [&]() {
            for (int64_t c; (c = next.fetch_add(1)) < nchunks;) {
                const int64_t a0 = c * chunk, a1 = std::min(attempts, a0 + chunk);
                ReplayUrng g{raw.get(), nraw, (size_t)(2 * a0), &ready};
                std::normal_distribution<double> normal(0.0, 1.0);
                auto& o = parts[(size_t)c];
                o.reserve((size_t)((a1 - a0) * 8 / 5));
                for (;;) {
                    const double first = normal(g);
                    const int64_t accepted = (int64_t)(g.pos / 2) - 1;  // the attempt that produced it
                    if (accepted >= a1) break;
                    o.push_back(first);
                    o.push_back(normal(g));  // held value: consumes no draw
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 0; t < std::max(1, workers); ++t) pool.emplace_back(work);
        constexpr size_t publish = 1 << 14;
        for (size_t i = 0; i < nraw; i += publish) {
            const size_t e = std::min(nraw, i + publish);
            for (size_t j = i; j < e; ++j) raw[j] = gen();
            ready.store(e, std::memory_order_release);
        }
        for (auto& t : pool) t.join();
        // The chunks' values in chunk order, copied by the workers in parallel.
        std::vector<int64_t> at((size_t)nchunks + 1, produced);
        for (int64_t c = 0; c < nchunks; ++c)
            at[(size_t)c + 1] = std::min<int64_t>(count, at[(size_t)c] + (int64_t)parts[(size_t)c].size());
        std::atomic<int64_t> next_copy{0};
        auto copy = [&]() {
            for (int64_t c; (c = next_copy.fetch_add(1)) < nchunks;)
                std::copy_n(parts[(size_t)c].begin(), at[(size_t)c + 1] - at[(size_t)c], out + at[(size_t)c]);
        };
        pool.clear();
        for (int t = 1; t < std::max(1, workers); ++t) pool.emplace_back(copy);
        copy();
        for (auto& t : pool) t.join();
        produced = at[(size_t)nchunks];
    }
}

}  // namespace

void make_omega(int64_t d, int64_t ell, uint64_t seed, std::vector<double>& u, std::vector<double>& v)
{
    // Two independent streams, each recorded by one thread and expanded by
    // half of the remaining host threads (at most 16 in all).
    const int threads = (int)std::max(4u, std::min(16u, std::thread::hardware_concurrency()));
    const int workers = (threads - 2) / 2;
    std::exception_ptr verr;
    std::thread tv([&] {
        try {
            v.resize((size_t)(d * ell));
            fill_normal(seed ^ 0x9e3779b97f4a7c15ULL, d * ell, v.data(), workers);
        } catch (...) {
            verr = std::current_exception();
        }
    });
    try {
        u.resize((size_t)(d * ell));
        fill_normal(seed, d * ell, u.data(), workers);
    } catch (...) {
        tv.join();
        throw;
    }
    tv.join();
    if (verr) std::rethrow_exception(verr);
}

std::vector<int64_t> select_anchors(Context& ctx, const SelectArgs& a, double* diag)
{
    const auto started = Clock::now();
    if (diag) {
        std::fill(diag, diag + kDiagCount, 0.0);
        std::fill(diag + kDiagAuditActual, diag + kDiagAuditDensityDense + 1, -1.0);
    }
    if (a.n < 0) throw std::invalid_argument("points and weights must have equal lengths");
    if (!a.axis_weights.empty() && a.axis_weights.size() != a.axes.size())
        throw std::invalid_argument("axis weights must match axes");
    for (double w : a.axis_weights)
        if (!std::isfinite(w) || w <= 0) throw std::invalid_argument("axis weights must be finite and positive");
    const AxisSpec spec = axis_spec(ctx, a.axes.data(), (int)a.axes.size(), a.mo_coeff, a.nocc);
    hipStream_t s = nullptr;
    std::vector<int64_t> out;
    switch (a.algorithm) {
    case kDense: {
        const auto scales = axis_scales(ctx, a, spec, s);
        out = select_dense(ctx, a, spec, scales, diag, s);
        break;
    }
    case kOnline: {
        const auto scales = axis_scales(ctx, a, spec, s);
        out = select_online(ctx, a, spec, scales, diag, s);
        break;
    }
    case kRandomized:
    case kDensity:
    case kDensityNormalized: {
        const auto scales = axis_scales(ctx, a, spec, s);
        out = select_randomized(ctx, a, spec, scales, diag, s);
        break;
    }
    default:
        throw std::invalid_argument("unknown anchor selection algorithm");
    }
    if (diag) diag[kDiagSecondsTotal] = seconds_since(started);
    return out;
}

}  // namespace mpisdf
