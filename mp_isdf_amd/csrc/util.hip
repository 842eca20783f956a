// Small elementwise / reduction kernels of the selection path.
#include "util.hpp"

#include <algorithm>

namespace mpisdf {
namespace {

int grid_for(int64_t work, int64_t cap = 65536)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((work + 255) / 256, cap));
}

// a[g * m + p] = sw[g] * f[g, p / w] * f[g, p % w]   (anchor_select_qrcp.cpp:178-184)
__global__ void pair_product_kernel(const double* __restrict__ f, int64_t n, int64_t w, const double* __restrict__ sw,
                                    double* __restrict__ a, unsigned long long* __restrict__ bad)
{
    const int64_t m = w * w, total = n * m;
    unsigned long long nonfinite = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = i / m, p = i - g * m;
        const double v = sw[g] * f[g * w + p / w] * f[g * w + p % w];
        a[i] = v;
        nonfinite += !isfinite(v);
    }
    if (nonfinite) atomicAdd(bad, nonfinite);
}

// out[j] += sum_g f[g, j]^2
__global__ void colsumsq_kernel(const double* __restrict__ f, int64_t n, int64_t w, double* __restrict__ out)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= w) return;
    double s = 0;
    for (int64_t g = blockIdx.y; g < n; g += gridDim.y) { const double v = f[g * w + j]; s += v * v; }
    atomicAdd(&out[j], s);
}

}  // namespace

void launch_pair_product(const double* f, int64_t n, int64_t w, const double* sw, double* a,
                         unsigned long long* bad, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(pair_product_kernel, dim3(grid_for(n * w * w)), dim3(256), 0, s, f, n, w, sw, a, bad);
    MPISDF_HIP(hipGetLastError());
}

void launch_colsumsq(const double* f, int64_t n, int64_t w, double* out, hipStream_t s)
{
    if (n <= 0 || w <= 0) return;
    const unsigned gy = (unsigned)std::min<int64_t>(n, 256);
    hipLaunchKernelGGL(colsumsq_kernel, dim3((unsigned)((w + 255) / 256), gy), dim3(256), 0, s, f, n, w, out);
    MPISDF_HIP(hipGetLastError());
}

namespace {
__global__ void column_norms2_f32_kernel(const float* __restrict__ y, int64_t m, int64_t n, int64_t ldy,
                                         double* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (j >= n) return;
    double s = 0;
    for (int64_t i = lane; i < m; i += 64) {
        const double v = (double)y[j * ldy + i];
        s += v * v;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) out[j] = s;
}
}  // namespace

void launch_column_norms2_f32(const float* y, int64_t m, int64_t n, int64_t ldy, double* out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(column_norms2_f32_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, y, m, n, ldy, out);
    MPISDF_HIP(hipGetLastError());
}

}  // namespace mpisdf
