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

// part[y * w + j] = sum over the rows g = y, y + gridDim.y, ... of f[g, j]^2
__global__ void colsumsq_kernel(const double* __restrict__ f, int64_t n, int64_t w, double* __restrict__ part)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= w) return;
    double s = 0;
    for (int64_t g = blockIdx.y; g < n; g += gridDim.y) { const double v = f[g * w + j]; s += v * v; }
    part[(int64_t)blockIdx.y * w + j] = s;
}

__global__ void sum_partials_kernel(const double* __restrict__ part, int64_t groups, int64_t len,
                                    double* __restrict__ out, int accumulate)
{
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < len; j += (int64_t)gridDim.x * blockDim.x) {
        double s = accumulate ? out[j] : 0.0;
        for (int64_t r = 0; r < groups; ++r) s += part[r * len + j];
        out[j] = s;
    }
}

// out[c] += sum_i x[c * ld + i]: one workgroup per vector, strided per-thread
// sums, then a butterfly over each wave and the waves in index order.
__global__ __launch_bounds__(256) void sum_fixed_kernel(const double* __restrict__ x, int64_t n, int64_t ld,
                                                        double* __restrict__ out)
{
    __shared__ double part[4];
    const double* v = x + (int64_t)blockIdx.x * ld;
    double s = 0;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += v[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] += ((part[0] + part[1]) + (part[2] + part[3]));
}

}  // namespace

void launch_sum_partials(const double* part, int64_t groups, int64_t len, double* out, bool accumulate,
                         hipStream_t s)
{
    if (len <= 0) return;
    hipLaunchKernelGGL(sum_partials_kernel, dim3(grid_for(len)), dim3(256), 0, s, part, groups, len, out,
                       accumulate ? 1 : 0);
    MPISDF_HIP(hipGetLastError());
}

void launch_sum_fixed(const double* x, int64_t n, int64_t ld, int64_t nvec, double* out, hipStream_t s)
{
    if (nvec <= 0) return;
    if (nvec > 0x7fffffff) throw std::overflow_error("too many vectors to sum in one launch");
    hipLaunchKernelGGL(sum_fixed_kernel, dim3((unsigned)nvec), dim3(256), 0, s, x, n, ld, out);
    MPISDF_HIP(hipGetLastError());
}

void launch_pair_product(const double* f, int64_t n, int64_t w, const double* sw, double* a,
                         unsigned long long* bad, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(pair_product_kernel, dim3(grid_for(n * w * w)), dim3(256), 0, s, f, n, w, sw, a, bad);
    MPISDF_HIP(hipGetLastError());
}

void launch_colsumsq(const double* f, int64_t n, int64_t w, double* out, DeviceBuffer<double>& part, hipStream_t s)
{
    if (n <= 0 || w <= 0) return;
    const unsigned gy = (unsigned)std::min<int64_t>(n, 256);
    if (part.count < (size_t)gy * w) part.resize((size_t)gy * w);
    hipLaunchKernelGGL(colsumsq_kernel, dim3((unsigned)((w + 255) / 256), gy), dim3(256), 0, s, f, n, w, part.ptr);
    MPISDF_HIP(hipGetLastError());
    launch_sum_partials(part.ptr, gy, w, out, true, s);
}

namespace {
template <class T>
__global__ void column_norms2_kernel(const T* __restrict__ y, int64_t m, int64_t n, int64_t ldy,
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
    hipLaunchKernelGGL(column_norms2_kernel<float>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, y, m, n, ldy, out);
    MPISDF_HIP(hipGetLastError());
}

void launch_column_norms2_f64(const double* y, int64_t m, int64_t n, int64_t ldy, double* out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(column_norms2_kernel<double>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, y, m, n, ldy,
                       out);
    MPISDF_HIP(hipGetLastError());
}

}  // namespace mpisdf
