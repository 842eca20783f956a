// Grid Coulomb metric of ISDF interpolation vectors and the J/K contractions
// that consume it (python/isdf.py:48-67 grid_coulomb_metric, :76-92
// grid_cross_metric, :130-151 evaluate_isdf_jk).
//
// The metric is M = L diag(w) K diag(w) R^T over n grid points, K the
// point-to-point Coulomb kernel with a spherical-cell self term.  Its cost is
// 2 n^2 kr + 2 n kl kr flops: an N-body product.  Rows of the weighted kernel
// K'_ij = w_i K_ij w_j are generated in blocks of R rows (R x n FP64, sized
// so that the product below runs ~1024 workgroups) by one VALU kernel, then
// two FP64 MFMA TN products consume the block:  V = K'_blk R^T (rows x kr),
// M += L_blk V.  For the symmetric metric of one vector set only the R x R
// blocks on and above the diagonal are generated and multiplied (off-diagonal
// ones twice weighted), which halves the dominant product.  The kernel block is
// written once and read by the product once per 128 columns of R, so for
// kr >= 128 the block costs 8 (1 + kr/128) bytes against 2 kr flops per
// element: FP64 MFMA bound above kr ~ 40 at 8 TB/s.
#include "coulomb.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>

namespace mpisdf {
namespace {

constexpr double kPi = 3.14159265358979323846;

// Kernel rows one thread writes for its column j (its point's coordinates and
// weight stay in registers; row i's are workgroup-uniform scalar loads).
constexpr int kRowsPerThread = 16;

// K'[r * ldk + c] = w_i w_j K_ij for grid points i = i0 + r (r < rows) and
// j = j0 + c (c < cols).
__global__ __launch_bounds__(256) void coulomb_rows_kernel(const double* __restrict__ xyz,
                                                           const double* __restrict__ w, int64_t i0, int64_t rows,
                                                           int64_t j0, int64_t cols, int64_t ldk, double csphere,
                                                           double* __restrict__ kb)
{
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= cols) return;
    const int64_t j = j0 + c;
    const double xj = xyz[3 * j], yj = xyz[3 * j + 1], zj = xyz[3 * j + 2], wj = w[j];
    for (int64_t r0 = (int64_t)blockIdx.y * kRowsPerThread; r0 < rows; r0 += (int64_t)gridDim.y * kRowsPerThread) {
        const int64_t rend = r0 + kRowsPerThread < rows ? r0 + kRowsPerThread : rows;
        for (int64_t r = r0; r < rend; ++r) {
            const int64_t i = i0 + r;
            double v;
            if (i == j) {
                v = csphere * pow(fabs(wj), -1.0 / 3.0) * wj * wj;
            } else {
                const double dx = xyz[3 * i] - xj, dy = xyz[3 * i + 1] - yj, dz = xyz[3 * i + 2] - zj;
                const double d2 = dx * dx + dy * dy + dz * dz;
                v = d2 > 0.0 ? w[i] * wj / sqrt(d2) : 0.0;
            }
            kb[r * ldk + c] = v;
        }
    }
}

// 1/sqrt(x) for finite x > 0: the hardware estimate v_rsq_f64 refined by two
// Newton-Raphson steps y <- y + y (1 - x y^2) / 2, each of which squares the
// relative error, to within a few ulp of the rounded quotient.
__device__ inline double inv_sqrt(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// v[i] = w_i sum_j K_ij w_j u_j, the kernel of grid_coulomb_metric applied to
// one weighted vector.  Each thread owns two target points; source points
// stream through LDS 256 at a time, and every target sums its sources in
// index order, so the result is the same on every run.
constexpr int kPotTargets = 2;
__global__ __launch_bounds__(256) void coulomb_potential_kernel(const double* __restrict__ xyz,
                                                                const double* __restrict__ w,
                                                                const double* __restrict__ u, int64_t n,
                                                                double csphere, double* __restrict__ v)
{
    __shared__ double sx[256], sy[256], sz[256], sq[256];
    double xi[kPotTargets], yi[kPotTargets], zi[kPotTargets], acc[kPotTargets];
    int64_t idx[kPotTargets];
#pragma unroll
    for (int t = 0; t < kPotTargets; ++t) {
        idx[t] = ((int64_t)blockIdx.x * kPotTargets + t) * 256 + threadIdx.x;
        const bool ok = idx[t] < n;
        xi[t] = ok ? xyz[3 * idx[t]] : 0.0;
        yi[t] = ok ? xyz[3 * idx[t] + 1] : 0.0;
        zi[t] = ok ? xyz[3 * idx[t] + 2] : 0.0;
        acc[t] = 0.0;
    }
    for (int64_t j0 = 0; j0 < n; j0 += 256) {
        const int64_t j = j0 + threadIdx.x;
        __syncthreads();
        if (j < n) {
            sx[threadIdx.x] = xyz[3 * j];
            sy[threadIdx.x] = xyz[3 * j + 1];
            sz[threadIdx.x] = xyz[3 * j + 2];
            sq[threadIdx.x] = w[j] * u[j];
        }
        __syncthreads();
        const int cnt = (int)(n - j0 < 256 ? n - j0 : 256);
        for (int c = 0; c < cnt; ++c) {
            const double px = sx[c], py = sy[c], pz = sz[c], q = sq[c];
#pragma unroll
            for (int t = 0; t < kPotTargets; ++t) {
                const double dx = xi[t] - px, dy = yi[t] - py, dz = zi[t] - pz;
                const double d2 = dx * dx + dy * dy + dz * dz;
                // the point itself and coincident distinct points add nothing here
                acc[t] += d2 > 0.0 ? q * inv_sqrt(d2) : 0.0;
            }
        }
    }
#pragma unroll
    for (int t = 0; t < kPotTargets; ++t) {
        const int64_t i = idx[t];
        if (i < n) {
            const double wi = w[i];
            v[i] = wi * (acc[t] + csphere * pow(fabs(wi), -1.0 / 3.0) * wi * u[i]);
        }
    }
}

void launch_rows(const double* xyz, const double* w, int64_t i0, int64_t rows, int64_t j0, int64_t cols, int64_t ldk,
                 double csphere, double* kb, hipStream_t s)
{
    const unsigned gx = (unsigned)((cols + 255) / 256);
    const unsigned gy = (unsigned)std::min<int64_t>((rows + kRowsPerThread - 1) / kRowsPerThread, 65535);
    hipLaunchKernelGGL(coulomb_rows_kernel, dim3(gx, gy), dim3(256), 0, s, xyz, w, i0, rows, j0, cols, ldk, csphere,
                       kb);
    MPISDF_HIP(hipGetLastError());
}

// In place a <- (a + a^T) / 2 for a row-major m x m matrix.
__global__ void symmetrize_kernel(double* __restrict__ a, int64_t m)
{
    const int64_t r = (int64_t)blockIdx.y * 16 + threadIdx.y, c = (int64_t)blockIdx.x * 16 + threadIdx.x;
    if (r >= m || c >= m || c <= r) return;
    const double s = 0.5 * (a[r * m + c] + a[c * m + r]);
    a[r * m + c] = s;
    a[c * m + r] = s;
}

// out[p * cols + c] = u[p] * x[p * cols + c]  (rows x cols)
__global__ void scale_rows_kernel(const double* __restrict__ x, const double* __restrict__ u, int64_t rows,
                                  int64_t cols, double* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= rows * cols) return;
    out[t] = u[t / cols] * x[t];
}

__global__ void hadamard_kernel(const double* __restrict__ a, const double* __restrict__ b, int64_t count,
                                double* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < count) out[t] = a[t] * b[t];
}

__global__ void fill_kernel(double* __restrict__ a, int64_t count, double value)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < count) a[t] = value;
}

void symmetrize(double* a, int64_t m, hipStream_t s)
{
    if (m <= 1) return;
    const unsigned g = (unsigned)((m + 15) / 16);
    hipLaunchKernelGGL(symmetrize_kernel, dim3(g, g), dim3(16, 16), 0, s, a, m);
    MPISDF_HIP(hipGetLastError());
}

unsigned blocks_of(int64_t count) { return (unsigned)((count + 255) / 256); }

}  // namespace

int64_t coulomb_rows_per_block(int64_t n, int64_t kr)
{
    const int64_t ldk = (std::max<int64_t>(n, 1) + 3) / 4 * 4;
    // The V = K'_blk R^T product runs one 256-thread workgroup per 128 x 128
    // tile, two per CU (73.7 KB of LDS each): 1024 tiles fill the 256 CUs in
    // exactly two rounds.
    const int64_t colgroups = std::max<int64_t>(1, (kr + 127) / 128);
    int64_t rows = std::max<int64_t>(1, 1024 / colgroups) * 128;
    // The block takes at most a third of the free device memory and 32 GiB.
    size_t free_bytes = 0, total_bytes = 0;
    MPISDF_HIP(hipMemGetInfo(&free_bytes, &total_bytes));
    const int64_t budget = std::min<int64_t>(int64_t{32} << 30, (int64_t)(free_bytes / 3));
    rows = std::min(rows, std::max<int64_t>(128, budget / (8 * ldk) / 128 * 128));
    return std::max<int64_t>(128, std::min(rows, (n + 127) / 128 * 128));
}

void grid_coulomb_potential(const double* xyz, const double* w, const double* u, int64_t n, double* v, hipStream_t s)
{
    if (n <= 0) return;
    const double csphere = (32.0 / 15.0) * kPi * kPi * std::pow(3.0 / (4.0 * kPi), 5.0 / 3.0);
    const int64_t per_block = 256 * kPotTargets;
    hipLaunchKernelGGL(coulomb_potential_kernel, dim3((unsigned)((n + per_block - 1) / per_block)), dim3(256), 0, s,
                       xyz, w, u, n, csphere, v);
    MPISDF_HIP(hipGetLastError());
}

void grid_coulomb_metric(const double* left, int64_t kl, int64_t ldl, const double* right, int64_t kr, int64_t ldr,
                         const double* xyz, const double* w, int64_t n, bool symmetrize_out, double* metric,
                         int64_t block_rows, hipStream_t s)
{
    if (kl < 0 || kr < 0 || n < 0) throw std::invalid_argument("negative metric sizes");
    if (symmetrize_out && kl != kr) throw std::invalid_argument("a symmetrised metric needs kl == kr");
    if (kl == 0 || kr == 0) return;
    if (n == 0) {
        hipLaunchKernelGGL(fill_kernel, dim3(blocks_of(kl * kr)), dim3(256), 0, s, metric, kl * kr, 0.0);
        MPISDF_HIP(hipGetLastError());
        return;
    }
    if (ldl < n || ldr < n) throw std::invalid_argument("metric operand leading dimensions are below n");
    const double csphere = (32.0 / 15.0) * kPi * kPi * std::pow(3.0 / (4.0 * kPi), 5.0 / 3.0);
    const int64_t ldk = (n + 3) / 4 * 4;
    if (block_rows < 0 || block_rows % 128 != 0)
        throw std::invalid_argument("block_rows must be 0 (automatic) or a positive multiple of 128");
    const int64_t rblk = block_rows ? block_rows : coulomb_rows_per_block(n, kr);
    if (symmetrize_out && left == right && ldl == ldr) {
        // One vector set: K' is symmetric, so only the block pairs J >= I are
        // generated.  V_I = K'_II T_I^T + 2 sum_{J>I} K'_IJ T_J^T and
        // M += T_I V_I give M + M^T of the full sum after symmetrisation.
        DeviceBuffer<double> kb((size_t)(rblk * rblk)), v((size_t)(rblk * kr));
        for (int64_t i0 = 0; i0 < n; i0 += rblk) {
            const int64_t rows = std::min(rblk, n - i0);
            for (int64_t j0 = i0; j0 < n; j0 += rblk) {
                const int64_t cols = std::min(rblk, n - j0);
                launch_rows(xyz, w, i0, rows, j0, cols, rblk, csphere, kb.ptr, s);
                // V[r + b * rblk] (+)= alpha sum_c K'[r][c] right[b][j0 + c]
                gemm_tn_f64(rows, kr, cols, j0 == i0 ? 1.0 : 2.0, kb.ptr, rblk, right + j0, ldr,
                            j0 == i0 ? 0.0 : 1.0, v.ptr, rblk, s);
            }
            gemm_tn_f64(kr, kl, rows, 1.0, v.ptr, rblk, left + i0, ldl, i0 == 0 ? 0.0 : 1.0, metric, kr, s);
        }
    } else {
        DeviceBuffer<double> kb((size_t)(rblk * ldk)), v((size_t)(rblk * kr));
        for (int64_t i0 = 0; i0 < n; i0 += rblk) {
            const int64_t rows = std::min(rblk, n - i0);
            launch_rows(xyz, w, i0, rows, 0, n, ldk, csphere, kb.ptr, s);
            // V[r + b * rblk] = sum_j K'[r][j] right[b][j]
            gemm_tn_f64(rows, kr, n, 1.0, kb.ptr, ldk, right, ldr, 0.0, v.ptr, rblk, s);
            // metric[a * kr + b] (+)= sum_r V[r + b * rblk] left[a][i0 + r]
            gemm_tn_f64(kr, kl, rows, 1.0, v.ptr, rblk, left + i0, ldl, i0 == 0 ? 0.0 : 1.0, metric, kr, s);
        }
    }
    if (symmetrize_out) symmetrize(metric, kl, s);
    MPISDF_HIP(hipStreamSynchronize(s));  // before the block buffers are freed
}

void isdf_jk(const double* x, int64_t k, int64_t nao, const double* z, const double* dm, double* j, double* kmat,
             hipStream_t s)
{
    if (k < 0 || nao < 0) throw std::invalid_argument("negative J/K sizes");
    if (nao == 0) return;
    if (k == 0) {
        hipLaunchKernelGGL(fill_kernel, dim3(blocks_of(nao * nao)), dim3(256), 0, s, j, nao * nao, 0.0);
        hipLaunchKernelGGL(fill_kernel, dim3(blocks_of(nao * nao)), dim3(256), 0, s, kmat, nao * nao, 0.0);
        MPISDF_HIP(hipGetLastError());
        return;
    }
    DeviceBuffer<double> t((size_t)(k * nao)), p((size_t)(k * k)), u((size_t)k), zp((size_t)(k * k));
    // T = x dm (k x nao), P = T x^T (k x k)
    gemm_f64(k, nao, nao, 1.0, MatView{x, nao, 1}, MatView{dm, nao, 1}, 0.0, MutView{t.ptr, nao, 1}, s);
    gemm_f64(k, k, nao, 1.0, MatView{t.ptr, nao, 1}, MatView{x, 1, nao}, 0.0, MutView{p.ptr, k, 1}, s);
    // u = z diag(P); J = x^T diag(u) x  (T is reused for diag(u) x)
    gemm_f64(k, 1, k, 1.0, MatView{z, k, 1}, MatView{p.ptr, k + 1, 1}, 0.0, MutView{u.ptr, 1, 1}, s);
    hipLaunchKernelGGL(scale_rows_kernel, dim3(blocks_of(k * nao)), dim3(256), 0, s, x, u.ptr, k, nao, t.ptr);
    MPISDF_HIP(hipGetLastError());
    gemm_f64(nao, nao, k, 1.0, MatView{x, 1, nao}, MatView{t.ptr, nao, 1}, 0.0, MutView{j, nao, 1}, s);
    // K = x^T (z o P) x  (T is reused for (z o P) x)
    hipLaunchKernelGGL(hadamard_kernel, dim3(blocks_of(k * k)), dim3(256), 0, s, z, p.ptr, k * k, zp.ptr);
    MPISDF_HIP(hipGetLastError());
    gemm_f64(k, nao, k, 1.0, MatView{zp.ptr, k, 1}, MatView{x, nao, 1}, 0.0, MutView{t.ptr, nao, 1}, s);
    gemm_f64(nao, nao, k, 1.0, MatView{x, 1, nao}, MatView{t.ptr, nao, 1}, 0.0, MutView{kmat, nao, 1}, s);
    symmetrize(j, nao, s);
    symmetrize(kmat, nao, s);
    MPISDF_HIP(hipStreamSynchronize(s));  // before the work buffers are freed
}

}  // namespace mpisdf
