// Device kernels of the online (multipass) QRCP.
//
// Pair columns a_g = sqrt|w_g| vec(f_g f_g^T) (m = d^2, pair = mu*d + nu) are
// never stored for the whole grid.  Two kinds of work exist:
//   * explicit columns (candidate checks, refreshes, direct rescans): one
//     workgroup builds a_g in scratch and applies the stored Householder
//     reflectors in order, then takes the tail norm from a given row;
//   * the per-pivot full pass r_g = q^T a_g and sum_i |q_i a_gi|, evaluated as
//     quadratic forms sqrt|w_g| f_g^T M f_g and sqrt|w_g| |f_g|^T |M| |f_g|
//     with M = reshape(q, d, d) (a GEMM F M followed by row dots).
#include "online_kernels.hpp"

#include "kernels.hpp"

#include <algorithm>

namespace mpisdf {
namespace {

constexpr int kThreads = 256;

__device__ double block_sum(double v, double* part)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = v;
    __syncthreads();
    double s = 0;
    for (int i = 0; i < kThreads / 64; ++i) s += part[i];
    return s;
}

// For each listed column: a = sw * vec(f f^T); apply reflectors 0..rank-1
// (H[rank-1] ... H[0] a, v_j[i] = 0 for i < j); norm[c] = ||a[first:]||.
// When ``out`` is given the transformed column is stored there (m values per
// listed column).
__global__ __launch_bounds__(kThreads) void transform_norm_kernel(
    const double* __restrict__ feat, const double* __restrict__ sw, int64_t d,
    const int64_t* __restrict__ ids, int64_t count, const double* __restrict__ refl,
    const double* __restrict__ tau, int64_t rank, int64_t first, double* __restrict__ scratch,
    double* __restrict__ norms, double* __restrict__ out)
{
    __shared__ double part[kThreads / 64];
    const int64_t m = d * d;
    double* a = scratch + (int64_t)blockIdx.x * m;
    for (int64_t c = blockIdx.x; c < count; c += gridDim.x) {
        const int64_t g = ids[c];
        const double* f = feat + g * d;
        const double s = sw[g];
        for (int64_t i = threadIdx.x; i < m; i += kThreads) a[i] = s * f[i / d] * f[i % d];
        __syncthreads();
        for (int64_t j = 0; j < rank; ++j) {
            const double* v = refl + j * m;
            double dot = 0;
            for (int64_t i = j + threadIdx.x; i < m; i += kThreads) dot += v[i] * a[i];
            dot = block_sum(dot, part) * tau[j];
            for (int64_t i = j + threadIdx.x; i < m; i += kThreads) a[i] -= v[i] * dot;
            __syncthreads();
        }
        double ss = 0;
        for (int64_t i = first + threadIdx.x; i < m; i += kThreads) ss += a[i] * a[i];
        ss = block_sum(ss, part);
        if (threadIdx.x == 0) norms[c] = sqrt(ss);
        if (out)
            for (int64_t i = threadIdx.x; i < m; i += kThreads) out[c * m + i] = a[i];
        __syncthreads();
    }
}

// q_k = H[0] ... H[k] e_k, one workgroup (k + 1 reflector applications).
__global__ __launch_bounds__(kThreads) void q_kernel(const double* __restrict__ refl, const double* __restrict__ tau,
                                                     int64_t m, int64_t k, double* __restrict__ q)
{
    __shared__ double part[kThreads / 64];
    for (int64_t i = threadIdx.x; i < m; i += kThreads) q[i] = i == k ? 1.0 : 0.0;
    __syncthreads();
    for (int64_t j = k + 1; j-- > 0;) {
        const double* v = refl + j * m;
        double dot = 0;
        for (int64_t i = j + threadIdx.x; i < m; i += kThreads) dot += v[i] * q[i];
        const double factor = tau[j] * block_sum(dot, part);
        for (int64_t i = j + threadIdx.x; i < m; i += kThreads) q[i] -= factor * v[i];
        __syncthreads();
    }
}

__global__ void abs_kernel(const double* __restrict__ x, int64_t count, double* __restrict__ y)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = fabs(x[i]);
}

// row[g] = sw_g * sum_nu T[g, nu] f[g, nu];  absrow[g] likewise with |T|, |f| inputs.
__global__ void rowdot_kernel(const double* __restrict__ t, const double* __restrict__ f, const double* __restrict__ sw,
                              int64_t n, int64_t d, int absolute, double* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (g >= n) return;
    double s = 0;
    for (int64_t j = lane; j < d; j += 64) {
        const double fv = absolute ? fabs(f[g * d + j]) : f[g * d + j];
        s += t[g * d + j] * fv;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) out[g] = sw[g] * s;
}

int grid_for(int64_t work, int64_t cap = 65536)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((work + 255) / 256, cap));
}

}  // namespace

void online_transform_norms(const double* feat, const double* sw, int64_t d, const int64_t* ids, int64_t count,
                            const double* refl, const double* tau, int64_t rank, int64_t first, double* scratch,
                            int64_t scratch_columns, double* norms, double* out, hipStream_t s)
{
    if (count <= 0) return;
    const int64_t blocks = std::min(count, scratch_columns);
    hipLaunchKernelGGL(transform_norm_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, feat, sw, d, ids, count,
                       refl, tau, rank, first, scratch, norms, out);
    MPISDF_HIP(hipGetLastError());
}

void online_make_q(const double* refl, const double* tau, int64_t m, int64_t k, double* q, hipStream_t s)
{
    hipLaunchKernelGGL(q_kernel, dim3(1), dim3(kThreads), 0, s, refl, tau, m, k, q);
    MPISDF_HIP(hipGetLastError());
}

void online_abs(const double* x, int64_t count, double* y, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(abs_kernel, dim3(grid_for(count)), dim3(256), 0, s, x, count, y);
    MPISDF_HIP(hipGetLastError());
}

void online_row_pass(const double* feat, const double* absfeat, const double* sw, int64_t n, int64_t d,
                     const double* q, double* absq, double* t, double* row, double* absrow, hipStream_t s)
{
    if (n <= 0) return;
    // T = F M with M[mu][nu] = q[mu * d + nu];  row[g] = sw_g sum_nu T[g][nu] f[g][nu]
    gemm_f64(n, d, d, 1.0, MatView{feat, d, 1}, MatView{q, d, 1}, 0.0, MutView{t, d, 1}, s);
    hipLaunchKernelGGL(rowdot_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, t, feat, sw, n, d, 0, row);
    MPISDF_HIP(hipGetLastError());
    // absrow[g] = sw_g |f_g|^T |M| |f_g|
    online_abs(q, d * d, absq, s);
    gemm_f64(n, d, d, 1.0, MatView{absfeat, d, 1}, MatView{absq, d, 1}, 0.0, MutView{t, d, 1}, s);
    hipLaunchKernelGGL(rowdot_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, t, absfeat, sw, n, d, 0,
                       absrow);
    MPISDF_HIP(hipGetLastError());
}

}  // namespace mpisdf
