// Dense audit of sketch pivots: the orthogonal-projection residual of the
// weighted pair matrix on the span of a pivot set
// (cpp/src/anchor_select_randomized_id.cpp:19-88).  The audit is limited to
// small problems (2 million packed pair-matrix elements), so the Gram-Schmidt
// steps are matrix-vector products on the device, one pivot at a time.
#include <cmath>
#include <limits>

#include "audit.hpp"
#include "kernels.hpp"
#include "util.hpp"

namespace mpisdf {
namespace {

constexpr int kThreads = 256;

// Scales v to unit norm, or to zero when its norm is at or below
// drop * ||orig|| (the column adds no direction).  One workgroup; the sums are
// reduced in a fixed order.
__global__ __launch_bounds__(kThreads) void normalize_or_drop_kernel(double* __restrict__ v,
                                                                     const double* __restrict__ orig, int64_t m,
                                                                     double drop)
{
    __shared__ double sv[kThreads], so[kThreads];
    double a = 0, b = 0;
    for (int64_t i = threadIdx.x; i < m; i += kThreads) {
        a += v[i] * v[i];
        b += orig[i] * orig[i];
    }
    sv[threadIdx.x] = a;
    so[threadIdx.x] = b;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sv[threadIdx.x] += sv[threadIdx.x + w];
            so[threadIdx.x] += so[threadIdx.x + w];
        }
        __syncthreads();
    }
    const double norm = sqrt(sv[0]);
    const double f = norm <= drop * sqrt(so[0]) || !(norm > 0) ? 0.0 : 1.0 / norm;
    for (int64_t i = threadIdx.x; i < m; i += kThreads) v[i] *= f;
}

double total_sumsq(const double* a, int64_t m, int64_t n, DeviceBuffer<double>& out, DeviceBuffer<double>& part,
                   hipStream_t s)
{
    MPISDF_HIP(hipMemsetAsync(out.ptr, 0, m * sizeof(double), s));
    launch_colsumsq(a, n, m, out.ptr, part, s);  // out[i] = sum_g a[g m + i]^2
    double t = 0;
    for (double x : out.download(s)) t += x;
    return t;
}

}  // namespace

double projection_relative_error(const double* a, int64_t m, int64_t n, const std::vector<int64_t>& pivots,
                                 hipStream_t s)
{
    if (m <= 0 || n <= 0) return 0.0;
    const int64_t k = (int64_t)pivots.size();
    for (int64_t p : pivots)
        if (p < 0 || p >= n) throw std::invalid_argument("audit pivot outside the grid");
    DeviceBuffer<double> sums(m), part;
    const double total = total_sumsq(a, m, n, sums, part, s);
    if (!(total > 0)) return 0.0;
    if (k == 0) return 1.0;
    const double drop = 32 * std::numeric_limits<double>::epsilon();
    DeviceBuffer<double> q((size_t)(m * k)), c(k), t((size_t)(k * n)), r((size_t)(m * n));
    for (int64_t j = 0; j < k; ++j) {
        double* v = q.ptr + j * m;
        const double* col = a + pivots[j] * m;
        MPISDF_HIP(hipMemcpyAsync(v, col, m * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (j > 0)
            for (int pass = 0; pass < 2; ++pass) {
                // c = Q_j^T v, then v -= Q_j c (dropped columns of Q are zero).
                gemm_f64(j, 1, m, 1.0, MatView{q.ptr, m, 1}, MatView{v, 1, 0}, 0.0, MutView{c.ptr, 1, 1}, s);
                gemm_f64(m, 1, j, -1.0, MatView{q.ptr, 1, m}, MatView{c.ptr, 1, 1}, 1.0, MutView{v, 1, 1}, s);
            }
        hipLaunchKernelGGL(normalize_or_drop_kernel, dim3(1), dim3(kThreads), 0, s, v, col, m, drop);
        MPISDF_HIP(hipGetLastError());
    }
    MPISDF_HIP(hipMemcpyAsync(r.ptr, a, (size_t)(m * n) * sizeof(double), hipMemcpyDeviceToDevice, s));
    for (int pass = 0; pass < 2; ++pass) {
        // T = Q^T R, then R -= Q T.
        gemm_f64(k, n, m, 1.0, MatView{q.ptr, m, 1}, MatView{r.ptr, 1, m}, 0.0, MutView{t.ptr, 1, k}, s);
        gemm_f64(m, n, k, -1.0, MatView{q.ptr, 1, m}, MatView{t.ptr, 1, k}, 1.0, MutView{r.ptr, 1, m}, s);
    }
    const double residual = total_sumsq(r.ptr, m, n, sums, part, s);
    return std::sqrt(residual / total);
}

}  // namespace mpisdf
