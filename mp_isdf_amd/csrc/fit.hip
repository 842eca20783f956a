// ISDF fit on the selected anchors (FP64, GPU).
//
// The reference refits with an SVD of the k x nao^2 anchor-pair matrix P
// (python/isdf.py:238-263: theta_g = V S^-1 U^T (phi_g (x) phi_g) / w_anchor)
// and measures the grid-weighted pair-space error with one nao^2 outer
// product per grid point (isdf.py:176-191).  Both touch nao^2-length
// vectors.  Here the same least-squares solution is reached through the
// k x k Gram of the pair columns,
//     P P^T = (Phi_a Phi_a^T) o (Phi_a Phi_a^T) = G2,   P a_g = (Phi_a phi_g)^2 = s_g,
// factorised G2 = L L^T in FP64:
//     theta_g = L^-T L^-1 s_g / w_a,
//     ||a_g - P^T theta_phys||^2 = ||phi_g||^4 - ||L^-1 s_g||^2.
// Gram, RHS and triangular products are FP64 MFMA GEMMs; the diagonal
// blocks of the Cholesky factor are factorised and inverted in LDS.
#include "fit.hpp"
#include "coulomb.hpp"

#include "kernels.hpp"
#include "util.hpp"

#include <algorithm>
#include <cmath>

namespace mpisdf {
namespace {

constexpr int NB = 64;  // diagonal block of the blocked Cholesky / inverse
constexpr double kDependent = 1e-14;
// Pseudo-inverse fallback (dependent anchor pairs): eigenvalues of G2 at or
// below this fraction of the largest are dropped (singular values of the pair
// matrix below ~1e-7 of the largest: the resolution of the normal equations
// in FP64), and the host eigensolver takes at most this many anchors.
constexpr double kEigenCut = 1e-14;
constexpr int64_t kPinvMaxAnchors = 2048;

// Eigen-decomposition of the symmetric n x n matrix a (row-major, destroyed)
// by cyclic Jacobi rotations: on return a holds the eigenvalues on its
// diagonal and v (row-major) the eigenvectors in its columns.  Rotations
// zero a_pq with t = sgn(theta) / (|theta| + sqrt(theta^2 + 1)),
// theta = (a_qq - a_pp) / (2 a_pq), c = 1 / sqrt(t^2 + 1), s = t c, applied as
// A <- P^T A P with P_pp = P_qq = c, P_pq = s, P_qp = -s (the classical
// Jacobi method, e.g. Golub and Van Loan, Matrix Computations, 8.5).
void jacobi_eigen(std::vector<double>& a, std::vector<double>& v, int64_t n)
{
    v.assign((size_t)(n * n), 0.0);
    for (int64_t i = 0; i < n; ++i) v[i * n + i] = 1.0;
    auto A = [&](int64_t i, int64_t j) -> double& { return a[i * n + j]; };
    double total = 0;
    for (double x : a) total += x * x;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0;
        for (int64_t i = 0; i < n; ++i)
            for (int64_t j = i + 1; j < n; ++j) off += A(i, j) * A(i, j);
        if (!(off > 1e-32 * total)) return;
        for (int64_t p = 0; p < n; ++p)
            for (int64_t q = p + 1; q < n; ++q) {
                const double apq = A(p, q);
                if (apq == 0.0) continue;
                const double theta = (A(q, q) - A(p, p)) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int64_t r = 0; r < n; ++r) {  // columns p, q
                    const double arp = A(r, p), arq = A(r, q);
                    A(r, p) = c * arp - sn * arq;
                    A(r, q) = sn * arp + c * arq;
                }
                for (int64_t r = 0; r < n; ++r) {  // rows p, q
                    const double apr = A(p, r), aqr = A(q, r);
                    A(p, r) = c * apr - sn * aqr;
                    A(q, r) = sn * apr + c * aqr;
                }
                A(p, q) = A(q, p) = 0.0;
                for (int64_t r = 0; r < n; ++r) {
                    const double vrp = v[r * n + p], vrq = v[r * n + q];
                    v[r * n + p] = c * vrp - sn * vrq;
                    v[r * n + q] = sn * vrp + c * vrq;
                }
            }
    }
    throw std::runtime_error("Jacobi eigensolver of the fit Gram did not converge");
}

int grid_for(int64_t work, int64_t cap = 65536)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((work + 255) / 256, cap));
}

__global__ void diagonal_kernel(const double* __restrict__ a, int64_t lda, int64_t k, double* __restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = a[i * lda + i];
}

// The diagonal of a row-major k x k matrix, gathered on the device and read
// back with one contiguous copy.  (A strided device-to-host 2-D copy into
// pageable memory has returned before its data landed when the runtime is
// traced, so no strided copy ends in host memory here.)
std::vector<double> diagonal(const double* a, int64_t k, hipStream_t s)
{
    DeviceBuffer<double> dg((size_t)k);
    hipLaunchKernelGGL(diagonal_kernel, dim3(grid_for(k)), dim3(256), 0, s, a, k, k, dg.ptr);
    MPISDF_HIP(hipGetLastError());
    return dg.download(s);
}

// Factor the jb x jb diagonal block in place (lower Cholesky) and write its
// inverse (lower, row-major jb x jb) to ``inv``.  A non-positive pivot sets
// *info to the 1-based global row.
__global__ __launch_bounds__(256) void potrf_block_kernel(double* __restrict__ a, int64_t lda, int jb, int64_t row0,
                                                          const double* __restrict__ gdiag, double* __restrict__ inv,
                                                          int* __restrict__ info)
{
    __shared__ double s[NB][NB + 1];
    __shared__ double x[NB][NB + 1];
    __shared__ int bad;
    const int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    for (int i = tid; i < jb * jb; i += 256) s[i / jb][i % jb] = a[(i / jb) * lda + (i % jb)];
    __syncthreads();
    for (int c = 0; c < jb; ++c) {
        if (tid == 0) {
            // A Schur pivot below 1e-14 of the column's own G2 diagonal means the
            // pair column is dependent to ~1e-7 relative: beyond what the FP64
            // normal equations resolve.
            const double d = s[c][c];
            if (!(d > kDependent * gdiag[row0 + c]) && !bad) { bad = 1; *info = (int)(row0 + c + 1); }
            s[c][c] = d > 0 ? sqrt(d) : 1.0;
        }
        __syncthreads();
        for (int r = c + 1 + tid; r < jb; r += 256) s[r][c] /= s[c][c];
        __syncthreads();
        const int rem = jb - c - 1;
        for (int i = tid; i < rem * rem; i += 256) {
            const int r = c + 1 + i / rem, q = c + 1 + i % rem;
            if (q <= r) s[r][q] -= s[r][c] * s[q][c];
        }
        __syncthreads();
    }
    // inverse of the lower factor, one column per thread
    if (tid < jb) {
        const int c = tid;
        for (int r = 0; r < jb; ++r) {
            if (r < c) { x[r][c] = 0.0; continue; }
            double v = r == c ? 1.0 : 0.0;
            for (int p = c; p < r; ++p) v -= s[r][p] * x[p][c];
            x[r][c] = v / s[r][r];
        }
    }
    __syncthreads();
    for (int i = tid; i < jb * jb; i += 256) {
        const int r = i / jb, q = i % jb;
        a[r * lda + q] = q <= r ? s[r][q] : 0.0;
        inv[r * jb + q] = x[r][q];
    }
}

// Per-point pair errors from Z (k x nb, column-major ld k), one wave per point:
//   qn = (sum_a phi_ga^2)^2 ; out[g] = |w_g| max(qn - ||z_g||^2, 0), out[ld + g] = |w_g| qn
// (summed over the batch in a fixed order afterwards).
__global__ void pair_error_kernel(const double* __restrict__ z, int64_t k, const double* __restrict__ phi, int64_t d,
                                  const double* __restrict__ w, int64_t nb, double* __restrict__ out, int64_t ld)
{
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    double e = 0, t = 0;
    if (g < nb) {
        double zz = 0, pp = 0;
        for (int64_t i = lane; i < k; i += 64) { const double v = z[i + g * k]; zz += v * v; }
        for (int64_t a = lane; a < d; a += 64) { const double v = phi[g * d + a]; pp += v * v; }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) { zz += __shfl_xor(zz, off); pp += __shfl_xor(pp, off); }
        const double qn = pp * pp;
        const double aw = fabs(w[g]);
        e = aw * fmax(qn - zz, 0.0);
        t = aw * qn;
    }
    if (lane == 0 && g < nb) { out[g] = e; out[ld + g] = t; }
}

// The same per-point terms from the row-group partials of ||z_g||^2
// (zpart[r * nb + g], added in row-group order): qn = ||phi_g||^4,
// out[g] = |w_g| max(qn - zz_g, 0), out[ld + g] = |w_g| qn; one wave per point.
__global__ void pair_error_norms_kernel(const double* __restrict__ zpart, int64_t groups,
                                        const double* __restrict__ phi, int64_t d, const double* __restrict__ w,
                                        int64_t nb, double* __restrict__ out, int64_t ld)
{
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (g >= nb) return;
    double pp = 0;
    for (int64_t a = lane; a < d; a += 64) { const double v = phi[g * d + a]; pp += v * v; }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) pp += __shfl_xor(pp, off);
    if (lane == 0) {
        double zz = 0;
        for (int64_t r = 0; r < groups; ++r) zz += zpart[r * nb + g];
        const double qn = pp * pp, aw = fabs(w[g]);
        out[g] = aw * fmax(qn - zz, 0.0);
        out[ld + g] = aw * qn;
    }
}

// pp[g] = ||phi_g||^2 (rows of phi, leading dimension ld), one wave per point.
__global__ void row_norms2_kernel(const double* __restrict__ phi, int64_t d, int64_t ld, int64_t nb,
                                  double* __restrict__ pp)
{
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (g >= nb) return;
    double acc = 0;
    for (int64_t a = lane; a < d; a += 64) { const double v = phi[g * ld + a]; acc += v * v; }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) pp[g] = acc;
}

// Per-point pair-error terms from ||z_g||^2 and ||phi_g||^2:
// out[g] = |w_g| max(pp^2 - zz, 0), out[ld + g] = |w_g| pp^2.
__global__ void pair_error_from_norms_kernel(const double* __restrict__ zz, const double* __restrict__ pp,
                                             const double* __restrict__ w, int64_t nb, double* __restrict__ out,
                                             int64_t ld)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nb) return;
    const double qn = pp[g] * pp[g], aw = fabs(w[g]);
    out[g] = aw * fmax(qn - zz[g], 0.0);
    out[ld + g] = aw * qn;
}

// part[y * k + r] = sum over the columns g of column block y (128 columns) of
// w[g] z[r + g * k]: this batch's share of sum_g w_g z_g, added in block order.
__global__ void zw_rows_kernel(const double* __restrict__ z, int64_t k, int64_t ld, int64_t nb,
                               const double* __restrict__ w, double* __restrict__ part)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= k) return;
    const int64_t g0 = (int64_t)blockIdx.y * 128, g1 = min(nb, g0 + 128);
    double s = 0;
    for (int64_t g = g0; g < g1; ++g) s += w[g] * z[r + g * ld];
    part[(int64_t)blockIdx.y * k + r] = s;
}

// out[c * rows + r] = in[r * cols + c] / (div ? div[c] : 1)
__global__ void transpose_kernel(const double* __restrict__ in, int64_t rows, int64_t cols,
                                 const double* __restrict__ div, double* __restrict__ out)
{
    __shared__ double tile[32][33];
    const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int64_t rr = r0 + r, cc = c0 + tx;
        tile[r][tx] = (rr < rows && cc < cols) ? in[rr * cols + cc] : 0.0;
    }
    __syncthreads();
    for (int c = ty; c < 32; c += 8) {
        const int64_t cc = c0 + c, rr = r0 + tx;
        if (rr < rows && cc < cols) out[cc * rows + rr] = div ? tile[tx][c] / div[cc] : tile[tx][c];
    }
}

// out[i] = a[i] * (b ? b[i] : 1) * (w ? |w[i / rows]| : 1)   (k x nb column-major)
__global__ void hadamard_weight_kernel(const double* a, const double* __restrict__ b,
                                       const double* __restrict__ w, int64_t rows, int64_t total, double* out)
{
    // a and out may alias (in-place weighting)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = a[i] * (b ? b[i] : 1.0) * (w ? fabs(w[i / rows]) : 1.0);
}

// Adjoint rows of the grid AOs: abar[g][mu] = 4 |w_g| (||phi_g||^2 phi_g[mu] - v[g][mu]),
// one wave per point.
__global__ void grid_adjoint_kernel(const double* __restrict__ phi, const double* __restrict__ v, int64_t d,
                                    const double* __restrict__ w, int64_t nb, double* __restrict__ abar)
{
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (g >= nb) return;
    const double* row = phi + g * d;
    double pp = 0;
    for (int64_t a = lane; a < d; a += 64) pp += row[a] * row[a];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) pp += __shfl_xor(pp, off);
    const double f = 4.0 * fabs(w[g]);
    for (int64_t a = lane; a < d; a += 64) abar[g * d + a] = f * (pp * row[a] - v[g * d + a]);
}

// m2 = h o g elementwise (k x k), then used as m2 Phi_a
__global__ void mul_kernel(double* __restrict__ h, const double* __restrict__ g, int64_t total)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        h[i] *= g[i];
}

// abar_a = -4 (m1 - m2)
__global__ void anchor_adjoint_kernel(const double* __restrict__ m1, const double* __restrict__ m2, int64_t total,
                                      double* __restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = -4.0 * (m1[i] - m2[i]);
}

// out[c * ldo + r] = in[r * ldi + c] for r < rows, c < cols.
__global__ void transpose_into_kernel(const double* __restrict__ in, int64_t rows, int64_t cols, int64_t ldi,
                                      double* __restrict__ out, int64_t ldo)
{
    __shared__ double tile[32][33];
    const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int64_t rr = r0 + r, cc = c0 + tx;
        tile[r][tx] = (rr < rows && cc < cols) ? in[rr * ldi + cc] : 0.0;
    }
    __syncthreads();
    for (int c = ty; c < 32; c += 8) {
        const int64_t cc = c0 + c, rr = r0 + tx;
        if (rr < rows && cc < cols) out[cc * ldo + rr] = tile[tx][c];
    }
}

__global__ void set_identity_kernel(double* __restrict__ x, int64_t k)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) x[i * k + i] = 1.0;
}

void transpose(const double* in, int64_t rows, int64_t cols, const double* div, double* out, hipStream_t s)
{
    const dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
    hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, s, in, rows, cols, div, out);
    MPISDF_HIP(hipGetLastError());
}

// Pair error of given coefficients, one wave per point g of the batch:
// c = column g of C (k x nb), s and t = columns of S = (Phi_a phi_g)^2 and
// T = G2 C;  out[g] = |w_g| (||phi_g||^4 - 2 c.s + c.t), out[ld + g] = |w_g| ||phi_g||^4.
__global__ void theta_pair_error_kernel(const double* __restrict__ c, const double* __restrict__ sm,
                                        const double* __restrict__ t, int64_t k, const double* __restrict__ phi,
                                        int64_t d, const double* __restrict__ w, int64_t nb, double* __restrict__ out,
                                        int64_t ld)
{
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (g >= nb) return;
    double cs = 0, ct = 0, pp = 0;
    for (int64_t i = lane; i < k; i += 64) {
        const double ci = c[i + g * k];
        cs += ci * sm[i + g * k];
        ct += ci * t[i + g * k];
    }
    for (int64_t a = lane; a < d; a += 64) { const double v = phi[g * d + a]; pp += v * v; }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        cs += __shfl_xor(cs, off);
        ct += __shfl_xor(ct, off);
        pp += __shfl_xor(pp, off);
    }
    if (lane == 0) {
        const double qn = pp * pp, aw = fabs(w[g]);
        out[g] = aw * fmax(qn - 2.0 * cs + ct, 0.0);
        out[ld + g] = aw * qn;
    }
}

// out[g * d + a] = w[g] in[g * d + a]
__global__ void scale_rows_kernel(const double* __restrict__ in, int64_t count, int64_t d,
                                  const double* __restrict__ w, double* __restrict__ out)
{
    const int64_t total = count * d;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = w[i / d] * in[i];
}

// Pseudo-inverse factor of G2 (see fit_factor): eigen-decomposition of G2 on
// the host, W = diag(lambda)^-1/2 V^T over the eigenvalues above kEigenCut of
// the largest (largest first), as a k x k row-major linv padded with zero
// rows and its transpose linvt.
void fit_pseudo_inverse(FitFactor& f, int64_t k, int64_t d, hipStream_t s)
{
    DeviceBuffer<double> g((size_t)(k * k));
    gemm_tn_f64_fused(TnFused::Square, k, k, d, 1.0, f.phi_a.ptr, d, f.phi_a.ptr, d, g.ptr, k, nullptr, nullptr,
                      nullptr, s);
    std::vector<double> a = g.download(s), v;
    jacobi_eigen(a, v, k);
    std::vector<int64_t> order((size_t)k);
    for (int64_t i = 0; i < k; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return a[x * k + x] > a[y * k + y]; });
    const double lmax = a[order[0] * k + order[0]];
    if (!(lmax > 0)) throw std::runtime_error("ISDF fit Gram has zero numerical rank");
    std::vector<double> w((size_t)(k * k), 0.0), wt((size_t)(k * k), 0.0);
    int64_t r = 0;
    double lmin = lmax;
    for (int64_t t = 0; t < k; ++t) {
        const int64_t j = order[t];
        const double lam = a[j * k + j];
        if (!(lam > kEigenCut * lmax)) break;
        const double sc = 1.0 / std::sqrt(lam);
        for (int64_t i = 0; i < k; ++i) {
            w[t * k + i] = v[i * k + j] * sc;
            wt[i * k + t] = v[i * k + j] * sc;
        }
        lmin = lam;
        ++r;
    }
    f.linv.upload(w.data(), w.size(), s);
    f.linvt.upload(wt.data(), wt.size(), s);
    MPISDF_HIP(hipStreamSynchronize(s));
    f.triangular = false;
    f.retained_rank = r;
    f.condition_estimate = lmax / lmin;
}

}  // namespace

void fit_factor(Context& ctx, const double* anchor_xyz, int64_t k, FitFactor& f, hipStream_t s)
{
    const int64_t d = ctx.nao;
    f.k = k;
    f.d = d;
    f.phi_a.resize((size_t)(k * d));
    launch_ao_eval(ctx, anchor_xyz, k, kValues, kF64, f.phi_a.ptr, d, k * d, s);
    // G2 = (Phi_a Phi_a^T) o (Phi_a Phi_a^T), row-major k x k
    f.linv.resize((size_t)(k * k));
    DeviceBuffer<double> gdiag;
    DeviceBuffer<double> a((size_t)(k * k));
    gemm_tn_f64_fused(TnFused::Square, k, k, d, 1.0, f.phi_a.ptr, d, f.phi_a.ptr, d, a.ptr, k, nullptr, nullptr,
                      nullptr, s);  // symmetric: layout-free
    MPISDF_HIP(hipGetLastError());
    // Diagonal of G2 gives the pair column norms; keep the max for diagnostics.
    {
        const std::vector<double> diag = diagonal(a.ptr, k, s);
        f.gram_diag_max = *std::max_element(diag.begin(), diag.end());
        gdiag.upload(diag.data(), k, s);
        MPISDF_HIP(hipStreamSynchronize(s));
    }
    // Blocked right-looking Cholesky, lower, in place.
    DeviceBuffer<int> info(1);
    MPISDF_HIP(hipMemsetAsync(info.ptr, 0, sizeof(int), s));
    const int64_t nblocks = (k + NB - 1) / NB;
    DeviceBuffer<double> invd((size_t)(nblocks * NB * NB)), panel;
    for (int64_t j0 = 0; j0 < k; j0 += NB) {
        const int jb = (int)std::min<int64_t>(NB, k - j0);
        double* inv_j = invd.ptr + (j0 / NB) * NB * NB;
        hipLaunchKernelGGL(potrf_block_kernel, dim3(1), dim3(256), 0, s, a.ptr + j0 * k + j0, k, jb, j0, gdiag.ptr,
                           inv_j, info.ptr);
        MPISDF_HIP(hipGetLastError());
        const int64_t rem = k - j0 - jb;
        if (rem <= 0) continue;
        // panel <- A[j0+jb:, j0:j0+jb] inv(L_jj)^T
        if (panel.count < (size_t)(rem * jb)) panel.resize((size_t)(rem * jb));
        gemm_f64(rem, jb, jb, 1.0, MatView{a.ptr + (j0 + jb) * k + j0, k, 1}, MatView{inv_j, 1, jb}, 0.0,
                 MutView{panel.ptr, jb, 1}, s);
        MPISDF_HIP(hipMemcpy2DAsync(a.ptr + (j0 + jb) * k + j0, k * sizeof(double), panel.ptr, jb * sizeof(double),
                                    jb * sizeof(double), rem, hipMemcpyDeviceToDevice, s));
        // trailing A[j0+jb:, j0+jb:] -= panel panel^T (symmetric, so the TN
        // product's column-major store is the row-major block)
        gemm_tn_f64(rem, rem, jb, -1.0, panel.ptr, jb, panel.ptr, jb, 1.0, a.ptr + (j0 + jb) * k + (j0 + jb), k, s);
    }
    int hinfo = 0;
    MPISDF_HIP(hipMemcpyAsync(&hinfo, info.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    MPISDF_HIP(hipStreamSynchronize(s));
    f.triangular = true;
    f.retained_rank = k;
    if (hinfo) {
        // Dependent anchor pairs (a Schur pivot below 1e-14 of its diagonal,
        // e.g. more anchors than independent pairs).  The reference's SVD
        // refit keeps the directions above its cutoff (python/isdf.py:238-263);
        // here G2 = V diag(lambda) V^T, and the rows of W = diag(lambda)^-1/2 V^T
        // over the kept eigenvalues replace L^-1: ||W s||^2 = s^T G2^+ s and
        // W^T W s = G2^+ s, the minimum-norm least-squares coefficients.
        if (k > kPinvMaxAnchors)
            throw std::runtime_error("ISDF fit Gram (Phi_a Phi_a^T)^2 is not positive definite at row " +
                                     std::to_string(hinfo) + " (pivot below 1e-14 of its diagonal): the anchor "
                                     "pair columns are numerically dependent, and the pseudo-inverse fallback "
                                     "takes at most " + std::to_string(kPinvMaxAnchors) + " anchors");
        fit_pseudo_inverse(f, k, d, s);
        return;
    }
    // L^-1 right-looking, in place on X = I (row-major, lower triangular):
    // for each block J of NB rows (w = j0 + jb),
    //   X[J, :w] <- inv(L_JJ) X[J, :w],
    //   X[r, :w] -= L[r, J] X[J, :w] for every row r below J (one TN product, K = jb),
    // which is forward substitution of L X = I by blocks.  The transpose of each
    // finished block row goes to linvt: the TN product's K-contiguous operand
    // here, and L^-T for the fit pass.
    f.linvt.resize((size_t)(k * k));
    MPISDF_HIP(hipMemsetAsync(f.linv.ptr, 0, (size_t)(k * k) * sizeof(double), s));
    MPISDF_HIP(hipMemsetAsync(f.linvt.ptr, 0, (size_t)(k * k) * sizeof(double), s));
    hipLaunchKernelGGL(set_identity_kernel, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, s, f.linv.ptr, k);
    MPISDF_HIP(hipGetLastError());
    DeviceBuffer<double> xs((size_t)(NB * k));
    for (int64_t j0 = 0; j0 < k; j0 += NB) {
        const int jb = (int)std::min<int64_t>(NB, k - j0);
        const int64_t w = j0 + jb;
        double* xj = f.linv.ptr + j0 * k;
        gemm_f64(jb, w, jb, 1.0, MatView{invd.ptr + (j0 / NB) * NB * NB, jb, 1}, MatView{xj, k, 1}, 0.0,
                 MutView{xs.ptr, w, 1}, s);
        MPISDF_HIP(hipMemcpy2DAsync(xj, k * sizeof(double), xs.ptr, w * sizeof(double), w * sizeof(double), jb,
                                    hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(transpose_into_kernel, dim3((unsigned)((w + 31) / 32), (unsigned)((jb + 31) / 32)),
                           dim3(256), 0, s, xs.ptr, (int64_t)jb, w, w, f.linvt.ptr + j0, k);
        MPISDF_HIP(hipGetLastError());
        const int64_t rem = k - w;
        if (rem > 0)
            gemm_tn_f64(w, rem, jb, -1.0, f.linvt.ptr + j0, k, a.ptr + w * k + j0, k, 1.0, f.linv.ptr + w * k, k, s);
    }
    const std::vector<double> ldiag = diagonal(a.ptr, k, s);
    double lo = ldiag[0], hi = ldiag[0];
    for (double v : ldiag) { lo = std::min(lo, v); hi = std::max(hi, v); }
    // cond(G2) estimate from the factor's diagonal (a lower bound of the true condition)
    f.condition_estimate = (hi / lo) * (hi / lo);
}

void fit_stream(Context& ctx, const FitFactor& f, const double* anchor_w, const double* xyz, const double* w, int64_t n,
                int64_t batch, double* theta_host, double* err_out, double* tsum_out, hipStream_t s,
                double* gemm_stats, FitProducts products)
{
    const int64_t k = f.k, d = f.d;
    // The pseudo-inverse's rows carry the inverse square roots of eigenvalues
    // down to 1e-14 of the largest, which would amplify the digits' rounding
    // (~5e-13) to ~1e-6 of z: that (small, k <= 2048) case keeps FP64 products.
    const bool digits = products == FitProducts::Int8Digits && f.triangular;
    if (digits && theta_host)
        throw std::invalid_argument("the INT8-digit grid pass gives the pair error and theta . w, not theta");
    if (digits && k > kDigitRowsMaxF64)
        throw std::invalid_argument("the INT8-digit grid pass takes at most 16384 anchors");
    LaunchTimer timer, ztimer;
    double gemm_flops = 0;
    int64_t zlaunches = 0;
    batch = std::max<int64_t>(1, std::min(batch, n));
    // digit form: S is held with a column pitch of P (a multiple of 64) and
    // digitised in place; the Z product folds it against the digits of L^-1
    const int64_t P = digits ? (k + kDigitSlice - 1) / kDigitSlice * kDigitSlice : k;
    // digit form: phi rows with a pitch of dpad (a multiple of 64), digitised
    // in place after their norms are taken
    const int64_t dpad = digits ? (d + kDigitSlice - 1) / kDigitSlice * kDigitSlice : d;
    DeviceBuffer<double> phi((size_t)(batch * dpad)), sm((size_t)(P * batch)), z, acc(2), wa;
    if (digits && dpad != d) MPISDF_HIP(hipMemsetAsync(phi.ptr, 0, phi.count * sizeof(double), s));
    if (!digits) z.resize((size_t)(k * batch));
    const int64_t zgroups = digits ? 1 : (k + 127) / 128;  // row groups of the Z product
    DeviceBuffer<double> zz((size_t)(zgroups * batch)), ept((size_t)(2 * batch));
    DeviceBuffer<double> th, thr, tsum, tpart, zw, yscale, ynorm, pscale, pp;
    FoldI8Rows lrows, arows;
    if (digits) {
        yscale.resize((size_t)batch);
        ynorm.resize((size_t)batch);
        pscale.resize((size_t)batch);
        pp.resize((size_t)batch);
        split_rows_i8(f.linv.ptr, k, k, k, lrows, s);
        split_rows_i8(f.phi_a.ptr, d, k, d, arows, s);
    }
    // theta . w alone needs no theta: sum_g w_g theta_phys[:, g] =
    // L^-T (sum_g w_g z_g), one k x k triangular product after the grid pass
    // (digit form: L^-T L^-1 (sum_g w_g s_g), two).
    const bool theta_w_only = tsum_out && !theta_host;
    std::vector<double> stage;
    wa.upload(anchor_w, k, s);
    if (theta_host) {
        th.resize((size_t)(k * batch));
        thr.resize((size_t)(k * batch));
    }
    if (tsum_out) {
        tsum.resize(k);
        tpart.resize((size_t)(tn_column_blocks(batch) * k));
        MPISDF_HIP(hipMemsetAsync(tsum.ptr, 0, k * sizeof(double), s));
    }
    if (theta_w_only) {
        zw.resize(k);
        MPISDF_HIP(hipMemsetAsync(zw.ptr, 0, k * sizeof(double), s));
    }
    MPISDF_HIP(hipMemsetAsync(acc.ptr, 0, 2 * sizeof(double), s));
    for (int64_t start = 0; start < n; start += batch) {
        const int64_t nb = std::min(batch, n - start);
        launch_ao_eval(ctx, xyz + 3 * start, nb, kValues, kF64, phi.ptr, dpad, nb * dpad, s);
        if (digits) {
            // ||phi_g||^2, then phi_g and the anchor rows as five base-254
            // digits; S = (Phi_a Phi_b^T)^2 (k x nb, pitch P) from the 15
            // digit-pair products on INT8 MFMA
            hipLaunchKernelGGL(row_norms2_kernel, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, s, phi.ptr, d, dpad,
                               nb, pp.ptr);
            digitize_columns_f64_max(phi.ptr, d, nb, dpad, ynorm.ptr, pscale.ptr, s);
            timer.begin(s);
            fold_rows_square_i8(arows, reinterpret_cast<const int8_t*>(phi.ptr), dpad, pscale.ptr, nb, sm.ptr, P, s);
            timer.end(s);
            gemm_flops += 2.0 * (double)k * (double)d * (double)nb;
            if (theta_w_only) {  // sum_g w_g s_g, before S becomes digits
                const dim3 grid((unsigned)((k + 255) / 256), (unsigned)((nb + 127) / 128));
                hipLaunchKernelGGL(zw_rows_kernel, grid, dim3(256), 0, s, sm.ptr, k, P, nb, w + start, tpart.ptr);
                MPISDF_HIP(hipGetLastError());
                launch_sum_partials(tpart.ptr, grid.y, k, zw.ptr, true, s);
            }
            // five base-254 planes per entry under each column's power of two
            // (~4.8e-13 of its largest entry), then ||L^-1 s_hat_g||^2 on INT8
            digitize_columns_f64_max(sm.ptr, k, nb, P, ynorm.ptr, yscale.ptr, s);
            MPISDF_HIP(hipMemsetAsync(zz.ptr, 0, (size_t)nb * sizeof(double), s));
            timer.begin(s);
            ztimer.begin(s);
            fold_rows_i8(lrows, f.triangular, reinterpret_cast<const int8_t*>(sm.ptr), P, yscale.ptr, nb, zz.ptr, s);
            ztimer.end(s);
            timer.end(s);
            zlaunches += (k + 127) / 128;
            gemm_flops += (double)k * (double)(k + 1) * (double)nb;
            if (err_out) {
                hipLaunchKernelGGL(pair_error_from_norms_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s,
                                   zz.ptr, pp.ptr, w + start, nb, ept.ptr, batch);
                launch_sum_fixed(ept.ptr, nb, batch, 2, acc.ptr, s);
            }
            MPISDF_HIP(hipGetLastError());
            continue;
        }
        // S = (Phi_a Phi_b^T)^2, k x nb column-major (squared in the epilogue)
        timer.begin(s);
        gemm_tn_f64_fused(TnFused::Square, k, nb, d, 1.0, f.phi_a.ptr, d, phi.ptr, d, sm.ptr, k, nullptr, nullptr,
                          nullptr, s);
        timer.end(s);
        gemm_flops += 2.0 * (double)k * (double)d * (double)nb;
        // Z = L^-1 S (lower triangular A), with ||z_g||^2 summed in the epilogue
        timer.begin(s);
        ztimer.begin(s);
        gemm_tn_f64_fused(TnFused::ColumnSumSquares, k, nb, k, 1.0, f.linv.ptr, k, sm.ptr, k, z.ptr, k, zz.ptr,
                          nullptr, nullptr, s, f.triangular ? 1 : 0);
        ztimer.end(s);
        timer.end(s);
        ++zlaunches;
        gemm_flops += (double)k * (double)(k + 1) * (double)nb;
        if (err_out) {
            hipLaunchKernelGGL(pair_error_norms_kernel, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, s, zz.ptr, zgroups,
                               phi.ptr, d, w + start, nb, ept.ptr, batch);
            launch_sum_fixed(ept.ptr, nb, batch, 2, acc.ptr, s);
        }
        if (theta_w_only) {
            const dim3 grid((unsigned)((k + 255) / 256), (unsigned)((nb + 127) / 128));
            hipLaunchKernelGGL(zw_rows_kernel, grid, dim3(256), 0, s, z.ptr, k, k, nb, w + start, tpart.ptr);
            MPISDF_HIP(hipGetLastError());
            launch_sum_partials(tpart.ptr, grid.y, k, zw.ptr, true, s);
        }
        if (theta_host) {
            // theta_phys = L^-T Z (upper triangular A), column-major k x nb;
            // theta = theta_phys / w_anchor (with t_phys += theta_phys w_grid
            // summed in the epilogue; divided by w_anchor at the end)
            timer.begin(s);
            if (tsum_out)
                gemm_tn_f64_fused(TnFused::RowDot, k, nb, k, 1.0, f.linvt.ptr, k, z.ptr, k, th.ptr, k, nullptr,
                                  tpart.ptr, w + start, s, f.triangular ? 2 : 0);
            else
                gemm_tn_f64(k, nb, k, 1.0, f.linvt.ptr, k, z.ptr, k, 0.0, th.ptr, k, s, f.triangular ? 2 : 0);
            timer.end(s);
            if (tsum_out) launch_sum_partials(tpart.ptr, tn_column_blocks(nb), k, tsum.ptr, true, s);
            gemm_flops += (double)k * (double)(k + 1) * (double)nb;
            {  // rows of theta[k][n] on the host, through a contiguous staging copy
                transpose(th.ptr, nb, k, wa.ptr, thr.ptr, s);
                if (nb == n) {
                    MPISDF_HIP(hipMemcpyAsync(theta_host, thr.ptr, (size_t)(k * n) * sizeof(double),
                                              hipMemcpyDeviceToHost, s));
                } else {
                    stage.resize((size_t)(k * nb));
                    MPISDF_HIP(hipMemcpyAsync(stage.data(), thr.ptr, (size_t)(k * nb) * sizeof(double),
                                              hipMemcpyDeviceToHost, s));
                    MPISDF_HIP(hipStreamSynchronize(s));
                    for (int64_t t = 0; t < k; ++t)
                        std::copy(stage.data() + t * nb, stage.data() + (t + 1) * nb, theta_host + t * n + start);
                }
            }
        }
        MPISDF_HIP(hipGetLastError());
    }
    if (err_out) {
        const auto h = acc.download(s);
        err_out[0] = h[0];
        err_out[1] = h[1];
    }
    if (theta_w_only && digits) {  // zw = sum_g w_g s_g: first L^-1 zw (linv: row-major)
        DeviceBuffer<double> tmp((size_t)k);
        gemm_f64(k, 1, k, 1.0, MatView{f.linv.ptr, k, 1}, MatView{zw.ptr, 1, 1}, 0.0, MutView{tmp.ptr, 1, 1}, s);
        MPISDF_HIP(hipMemcpyAsync(zw.ptr, tmp.ptr, (size_t)k * sizeof(double), hipMemcpyDeviceToDevice, s));
        MPISDF_HIP(hipStreamSynchronize(s));
    }
    if (theta_w_only)  // t_phys = L^-T zw (linvt: row-major upper triangular)
        gemm_f64(k, 1, k, 1.0, MatView{f.linvt.ptr, k, 1}, MatView{zw.ptr, 1, 1}, 0.0, MutView{tsum.ptr, 1, 1}, s);
    if (tsum_out) {
        MPISDF_HIP(hipMemcpyAsync(tsum_out, tsum.ptr, k * sizeof(double), hipMemcpyDeviceToHost, s));
        MPISDF_HIP(hipStreamSynchronize(s));
        for (int64_t t = 0; t < k; ++t) tsum_out[t] /= anchor_w[t];
    }
    if (gemm_stats) {
        MPISDF_HIP(hipStreamSynchronize(s));
        gemm_stats[0] = timer.seconds();
        gemm_stats[1] = gemm_flops;
        gemm_stats[2] = (double)timer.launches();
        gemm_stats[3] = ztimer.seconds();
        gemm_stats[4] = (double)zlaunches;
    }
}

void theta_pair_error(Context& ctx, const double* anchor_xyz, const double* anchor_w, int64_t k, const double* xyz,
                      const double* w, int64_t n, const double* theta_host, int64_t batch, double* err_out,
                      hipStream_t s)
{
    const int64_t d = ctx.nao;
    batch = std::max<int64_t>(1, std::min(batch, std::max<int64_t>(n, 1)));
    DeviceBuffer<double> phi_a((size_t)(k * d)), g2((size_t)(k * k)), phi((size_t)(batch * d)),
        sm((size_t)(k * batch)), cm((size_t)(k * batch)), tm((size_t)(k * batch)), ept((size_t)(2 * batch)), acc(2);
    launch_ao_eval(ctx, anchor_xyz, k, kValues, kF64, phi_a.ptr, d, k * d, s);
    gemm_tn_f64_fused(TnFused::Square, k, k, d, 1.0, phi_a.ptr, d, phi_a.ptr, d, g2.ptr, k, nullptr, nullptr, nullptr,
                      s);
    MPISDF_HIP(hipMemsetAsync(acc.ptr, 0, 2 * sizeof(double), s));
    std::vector<double> stage((size_t)(k * batch));
    for (int64_t start = 0; start < n; start += batch) {
        const int64_t nb = std::min(batch, n - start);
        // c[t + g k] = w_anchor[t] theta[t][start + g] (the physical
        // coefficients); the previous batch's copy out of ``stage`` and its
        // reads of cm are finished first.
        MPISDF_HIP(hipStreamSynchronize(s));
        for (int64_t t = 0; t < k; ++t) {
            const double* row = theta_host + t * n + start;
            for (int64_t g = 0; g < nb; ++g) stage[(size_t)(t + g * k)] = anchor_w[t] * row[g];
        }
        MPISDF_HIP(hipMemcpyAsync(cm.ptr, stage.data(), (size_t)(k * nb) * sizeof(double), hipMemcpyHostToDevice, s));
        launch_ao_eval(ctx, xyz + 3 * start, nb, kValues, kF64, phi.ptr, d, nb * d, s);
        gemm_tn_f64_fused(TnFused::Square, k, nb, d, 1.0, phi_a.ptr, d, phi.ptr, d, sm.ptr, k, nullptr, nullptr,
                          nullptr, s);
        gemm_tn_f64(k, nb, k, 1.0, g2.ptr, k, cm.ptr, k, 0.0, tm.ptr, k, s);  // G2 symmetric
        hipLaunchKernelGGL(theta_pair_error_kernel, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, s, cm.ptr, sm.ptr,
                           tm.ptr, k, phi.ptr, d, w + start, nb, ept.ptr, batch);
        MPISDF_HIP(hipGetLastError());
        launch_sum_fixed(ept.ptr, nb, batch, 2, acc.ptr, s);
    }
    const auto h = acc.download(s);
    err_out[0] = h[0];
    err_out[1] = h[1];
}

void quadrature_overlap(Context& ctx, const double* xyz, const double* w, int64_t n, int64_t batch, double* out,
                        hipStream_t s)
{
    const int64_t d = ctx.nao;
    batch = std::max<int64_t>(1, std::min(batch, std::max<int64_t>(n, 1)));
    DeviceBuffer<double> phi((size_t)(batch * d)), wphi((size_t)(batch * d));
    MPISDF_HIP(hipMemsetAsync(out, 0, (size_t)(d * d) * sizeof(double), s));
    for (int64_t start = 0; start < n; start += batch) {
        const int64_t nb = std::min(batch, n - start);
        launch_ao_eval(ctx, xyz + 3 * start, nb, kValues, kF64, phi.ptr, d, nb * d, s);
        hipLaunchKernelGGL(scale_rows_kernel, dim3((unsigned)grid_for(nb * d)), dim3(256), 0, s, phi.ptr, nb, d,
                           w + start, wphi.ptr);
        MPISDF_HIP(hipGetLastError());
        // out += Phi_b^T diag(w) Phi_b, batches in order
        gemm_f64(d, d, nb, 1.0, MatView{phi.ptr, 1, d}, MatView{wphi.ptr, d, 1}, 1.0, MutView{out, d, 1}, s);
    }
}

// ---------------------------------------------------------------------------
// Adjoint of the fit with respect to the AO values (no reference counterpart).
//
// L = sum_g |w_g| || phi_g phi_g^T - sum_k c_gk phi_k phi_k^T ||_F^2 at the
// least-squares coefficients c_g = theta_phys[:, g].  Because c is optimal,
// dL/dc = 0 and the total derivative of L is its partial derivative at fixed c
// (envelope theorem).  With t_g = Phi_a phi_g, U = C o T (k x n), G = Phi_a Phi_a^T:
//   dL/dphi_g = 4 |w_g| ( ||phi_g||^2 phi_g - Phi_a^T (c_g o t_g) )
//   dL/dPhi_a = -4 ( (U W) Phi_grid - ((C W C^T) o G) Phi_a )
// Both are k x batch or k x k GEMMs; no nao^2 pair object is formed.  The
// AO adjoint kernel then contracts them with AO gradients recomputed on the
// fly: gao[c][mu] = sum_points dL/dphi_mu(r) d phi_mu(r)/d r_c.
// ---------------------------------------------------------------------------
void fit_loss_gradient(Context& ctx, const FitFactor& f, const double* anchor_xyz, const double* xyz,
                       const double* w, int64_t n, int64_t batch, double* loss_out, double* gao_out, hipStream_t s)
{
    const int64_t k = f.k, d = f.d, nao = ctx.nao;
    batch = std::max<int64_t>(1, std::min(batch, std::max<int64_t>(n, 1)));
    auto blocks = [](int64_t work) { return dim3((unsigned)grid_for(work)); };
    DeviceBuffer<double> phi((size_t)(batch * d)), tm((size_t)(k * batch)), sm((size_t)(k * batch)),
        z((size_t)(k * batch)), cm((size_t)(k * batch)), um((size_t)(k * batch)), vt((size_t)(batch * d)),
        abar((size_t)(std::max(batch, k) * d)), acc(2), gao((size_t)(3 * nao)), m1((size_t)(k * d)),
        cwc((size_t)(k * k)), gm((size_t)(k * k)), m2((size_t)(k * d)), phiat((size_t)(d * k)),
        ept((size_t)(2 * batch)), adj_part;
    MPISDF_HIP(hipMemsetAsync(acc.ptr, 0, 2 * sizeof(double), s));
    MPISDF_HIP(hipMemsetAsync(gao.ptr, 0, 3 * nao * sizeof(double), s));
    MPISDF_HIP(hipMemsetAsync(m1.ptr, 0, (size_t)(k * d) * sizeof(double), s));
    MPISDF_HIP(hipMemsetAsync(cwc.ptr, 0, (size_t)(k * k) * sizeof(double), s));
    transpose(f.phi_a.ptr, k, d, nullptr, phiat.ptr, s);  // Phi_a^T, row-major d x k
    for (int64_t start = 0; start < n; start += batch) {
        const int64_t nb = std::min(batch, n - start);
        const double* wb = w + start;
        launch_ao_eval(ctx, xyz + 3 * start, nb, kValues, kF64, phi.ptr, d, nb * d, s);
        // T = Phi_a Phi_b^T, S = T o T, Z = L^-1 S, C = L^-T Z  (k x nb column-major)
        gemm_tn_f64(k, nb, d, 1.0, f.phi_a.ptr, d, phi.ptr, d, 0.0, tm.ptr, k, s);
        hipLaunchKernelGGL(hadamard_weight_kernel, blocks(k * nb), dim3(256), 0, s, tm.ptr, tm.ptr,
                           (const double*)nullptr, k, k * nb, sm.ptr);
        gemm_tn_f64(k, nb, k, 1.0, f.linv.ptr, k, sm.ptr, k, 0.0, z.ptr, k, s, f.triangular ? 1 : 0);
        hipLaunchKernelGGL(pair_error_kernel, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, s, z.ptr, k, phi.ptr, d,
                           wb, nb, ept.ptr, batch);
        launch_sum_fixed(ept.ptr, nb, batch, 2, acc.ptr, s);
        gemm_tn_f64(k, nb, k, 1.0, f.linvt.ptr, k, z.ptr, k, 0.0, cm.ptr, k, s, f.triangular ? 2 : 0);
        // Grid adjoint: V = Phi_a^T (C o T) stored [g][mu]; abar_g = 4|w|(||phi||^2 phi - V)
        hipLaunchKernelGGL(hadamard_weight_kernel, blocks(k * nb), dim3(256), 0, s, cm.ptr, tm.ptr,
                           (const double*)nullptr, k, k * nb, um.ptr);
        gemm_tn_f64(d, nb, k, 1.0, phiat.ptr, k, um.ptr, k, 0.0, vt.ptr, d, s);
        hipLaunchKernelGGL(grid_adjoint_kernel, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, s, phi.ptr, vt.ptr, d,
                           wb, nb, abar.ptr);
        launch_ao_adjoint(ctx, xyz + 3 * start, nb, abar.ptr, d, gao.ptr, adj_part, s);
        // Anchor terms: M1 += (U W) Phi_b, CWC^T += (C W) C^T
        hipLaunchKernelGGL(hadamard_weight_kernel, blocks(k * nb), dim3(256), 0, s, um.ptr,
                           (const double*)nullptr, wb, k, k * nb, um.ptr);
        gemm_f64(k, d, nb, 1.0, MatView{um.ptr, 1, k}, MatView{phi.ptr, d, 1}, 1.0, MutView{m1.ptr, d, 1}, s);
        hipLaunchKernelGGL(hadamard_weight_kernel, blocks(k * nb), dim3(256), 0, s, cm.ptr,
                           (const double*)nullptr, wb, k, k * nb, tm.ptr);
        gemm_f64(k, k, nb, 1.0, MatView{tm.ptr, 1, k}, MatView{cm.ptr, k, 1}, 1.0, MutView{cwc.ptr, k, 1}, s);
        MPISDF_HIP(hipGetLastError());
    }
    // dL/dPhi_a = -4 (M1 - ((C W C^T) o G) Phi_a)
    gemm_tn_f64(k, k, d, 1.0, f.phi_a.ptr, d, f.phi_a.ptr, d, 0.0, gm.ptr, k, s);  // G, symmetric
    hipLaunchKernelGGL(mul_kernel, blocks(k * k), dim3(256), 0, s, cwc.ptr, gm.ptr, k * k);
    gemm_f64(k, d, k, 1.0, MatView{cwc.ptr, k, 1}, MatView{f.phi_a.ptr, d, 1}, 0.0, MutView{m2.ptr, d, 1}, s);
    hipLaunchKernelGGL(anchor_adjoint_kernel, blocks(k * d), dim3(256), 0, s, m1.ptr, m2.ptr, k * d, abar.ptr);
    MPISDF_HIP(hipGetLastError());
    launch_ao_adjoint(ctx, anchor_xyz, k, abar.ptr, d, gao.ptr, adj_part, s);
    const auto h = acc.download(s);
    if (loss_out) *loss_out = h[0];
    if (gao_out) MPISDF_HIP(hipMemcpyAsync(gao_out, gao.ptr, 3 * nao * sizeof(double), hipMemcpyDeviceToHost, s));
    MPISDF_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// Coulomb energy of the fitted density and its nuclear adjoint
// ---------------------------------------------------------------------------
namespace {

// out[i] = alpha a[i] (rowf ? rowf[i % rows] : 1) (colf ? colf[i / rows] : 1)
// for a column-major rows x (total / rows) matrix a.
__global__ void scale_rc_kernel(const double* __restrict__ a, int64_t rows, int64_t total,
                                const double* __restrict__ rowf, const double* __restrict__ colf, double alpha,
                                double* __restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = alpha * a[i] * (rowf ? rowf[i % rows] : 1.0) * (colf ? colf[i / rows] : 1.0);
}

// out[r] = sum_c a[r * d + c] b[r * d + c], one wavefront per row.
__global__ void row_dot_kernel(const double* __restrict__ a, const double* __restrict__ b, int64_t rows, int64_t d,
                               double* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (r >= rows) return;
    double s = 0;
    for (int64_t c = lane; c < d; c += 64) s += a[r * d + c] * b[r * d + c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) out[r] = s;
}

// h[i k + j] = g[i k + j] (x_i y_j + y_i x_j)
__global__ void sym_outer_hadamard_kernel(const double* __restrict__ g, const double* __restrict__ x,
                                          const double* __restrict__ y, int64_t k, double* __restrict__ h)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < k * k; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / k, j = e % k;
        h[e] = g[e] * (x[i] * y[j] + y[i] * x[j]);
    }
}

// abar[p][mu] = 2 (lam_p m1[p][mu] + beta_p y[p][mu] - m2[p][mu])
__global__ void anchor_energy_adjoint_kernel(const double* __restrict__ m1, const double* __restrict__ lam,
                                             const double* __restrict__ y, const double* __restrict__ beta,
                                             const double* __restrict__ m2, int64_t k, int64_t d,
                                             double* __restrict__ out)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < k * d; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = e / d;
        out[e] = 2.0 * (lam[p] * m1[e] + beta[p] * y[e] - m2[e]);
    }
}

// x <- G2^-1 x through the explicit factor inverse: L^-T (L^-1 x).
void solve_g2(const FitFactor& f, double* x, double* tmp, hipStream_t s)
{
    const int64_t k = f.k;
    gemm_tn_f64(k, 1, k, 1.0, f.linv.ptr, k, x, k, 0.0, tmp, k, s, f.triangular ? 1 : 0);
    gemm_tn_f64(k, 1, k, 1.0, f.linvt.ptr, k, tmp, k, 0.0, x, k, s, f.triangular ? 2 : 0);
}

}  // namespace

void coulomb_energy_gradient(Context& ctx, const FitFactor& f, const double* anchor_xyz, const double* xyz,
                             const double* w, int64_t n, const double* dm, int64_t batch, double* energy_out,
                             double* gao_out, hipStream_t s)
{
    const int64_t k = f.k, d = f.d, nao = ctx.nao;
    if (gao_out && !f.triangular)
        throw std::runtime_error("the Coulomb-energy gradient needs linearly independent anchor pairs; the fit fell "
                                 "back to the pseudo-inverse of the pair Gram");
    batch = std::max<int64_t>(1, std::min(batch, std::max<int64_t>(n, 1)));
    auto blocks = [](int64_t work) { return dim3((unsigned)grid_for(work)); };
    DeviceBuffer<double> y((size_t)(k * d)), rho(k), lam(k), tmp(k), u(std::max<int64_t>(n, 1)),
        v(std::max<int64_t>(n, 1)), uv(std::max<int64_t>(n, 1)), acc(1), phi((size_t)(batch * d)),
        tm((size_t)(k * batch)), sm((size_t)(k * batch));
    // rho_p = phi_p^T D phi_p (the density at the anchors), lam = G2^-1 rho.
    gemm_f64(k, d, d, 1.0, MatView{f.phi_a.ptr, d, 1}, MatView{dm, d, 1}, 0.0, MutView{y.ptr, d, 1}, s);
    hipLaunchKernelGGL(row_dot_kernel, dim3((unsigned)((k + 3) / 4)), dim3(256), 0, s, y.ptr, f.phi_a.ptr, k, d,
                       rho.ptr);
    MPISDF_HIP(hipMemcpyAsync(lam.ptr, rho.ptr, k * sizeof(double), hipMemcpyDeviceToDevice, s));
    solve_g2(f, lam.ptr, tmp.ptr, s);
    // u_g = s_g^T lam = (C^T rho)_g, the fitted density at grid point g.
    for (int64_t start = 0; start < n; start += batch) {
        const int64_t nb = std::min(batch, n - start);
        launch_ao_eval(ctx, xyz + 3 * start, nb, kValues, kF64, phi.ptr, d, nb * d, s);
        gemm_tn_f64_fused(TnFused::Square, k, nb, d, 1.0, f.phi_a.ptr, d, phi.ptr, d, sm.ptr, k, nullptr, nullptr,
                          nullptr, s);
        gemm_tn_f64(nb, 1, k, 1.0, sm.ptr, k, lam.ptr, k, 0.0, u.ptr + start, nb, s);
    }
    // v = W K W u; E_J = u^T v / 2.
    grid_coulomb_potential(xyz, w, u.ptr, n, v.ptr, s);
    hipLaunchKernelGGL(hadamard_weight_kernel, blocks(n), dim3(256), 0, s, u.ptr, v.ptr, (const double*)nullptr,
                       n, n, uv.ptr);
    MPISDF_HIP(hipMemsetAsync(acc.ptr, 0, sizeof(double), s));
    if (n > 0) launch_sum_fixed(uv.ptr, n, n, 1, acc.ptr, s);
    MPISDF_HIP(hipGetLastError());
    const double e = 0.5 * acc.download(s)[0];
    if (energy_out) *energy_out = e;
    if (!gao_out) return;

    // Adjoint.  dE = v^T du with u_g = s_g^T lam, lam = G2^-1 rho:
    //   dE = sum_g v_g lam^T ds_g + beta^T (d rho - dG2 lam),  beta = G2^-1 S v.
    DeviceBuffer<double> lt((size_t)(k * batch)), vt((size_t)(batch * d)), abar((size_t)(std::max(batch, k) * d)),
        m1((size_t)(k * d)), m2((size_t)(k * d)), sigma(k), gm((size_t)(k * k)), gao((size_t)(3 * nao)), spart,
        adj_part, phiat((size_t)(d * k));
    MPISDF_HIP(hipMemsetAsync(m1.ptr, 0, (size_t)(k * d) * sizeof(double), s));
    MPISDF_HIP(hipMemsetAsync(sigma.ptr, 0, k * sizeof(double), s));
    MPISDF_HIP(hipMemsetAsync(gao.ptr, 0, 3 * nao * sizeof(double), s));
    spart.resize((size_t)(k * ((batch + 127) / 128)));
    transpose(f.phi_a.ptr, k, d, nullptr, phiat.ptr, s);  // Phi_a^T, row-major d x k
    for (int64_t start = 0; start < n; start += batch) {
        const int64_t nb = std::min(batch, n - start);
        const double* vb = v.ptr + start;
        launch_ao_eval(ctx, xyz + 3 * start, nb, kValues, kF64, phi.ptr, d, nb * d, s);
        gemm_tn_f64(k, nb, d, 1.0, f.phi_a.ptr, d, phi.ptr, d, 0.0, tm.ptr, k, s);  // T = Phi_a Phi_b^T
        hipLaunchKernelGGL(hadamard_weight_kernel, blocks(k * nb), dim3(256), 0, s, tm.ptr, tm.ptr,
                           (const double*)nullptr, k, k * nb, sm.ptr);
        // sigma += S v_b, per 128-point block then in block order.
        const dim3 grid((unsigned)((k + 255) / 256), (unsigned)((nb + 127) / 128));
        hipLaunchKernelGGL(zw_rows_kernel, grid, dim3(256), 0, s, sm.ptr, k, k, nb, vb, spart.ptr);
        launch_sum_partials(spart.ptr, grid.y, k, sigma.ptr, true, s);
        // Grid AOs: abar_g = 2 v_g Phi_a^T (lam o t_g).
        hipLaunchKernelGGL(scale_rc_kernel, blocks(k * nb), dim3(256), 0, s, tm.ptr, k, k * nb, lam.ptr,
                           (const double*)nullptr, 1.0, lt.ptr);
        gemm_tn_f64(d, nb, k, 1.0, phiat.ptr, k, lt.ptr, k, 0.0, vt.ptr, d, s);
        hipLaunchKernelGGL(scale_rc_kernel, blocks(nb * d), dim3(256), 0, s, vt.ptr, d, nb * d,
                           (const double*)nullptr, vb, 2.0, abar.ptr);
        launch_ao_adjoint(ctx, xyz + 3 * start, nb, abar.ptr, d, gao.ptr, adj_part, s);
        // Anchor AOs (s_g part): M1 += (T o v) Phi_b, i.e. sum_g v_g t_gk phi_g.
        hipLaunchKernelGGL(scale_rc_kernel, blocks(k * nb), dim3(256), 0, s, tm.ptr, k, k * nb,
                           (const double*)nullptr, vb, 1.0, lt.ptr);
        gemm_f64(k, d, nb, 1.0, MatView{lt.ptr, 1, k}, MatView{phi.ptr, d, 1}, 1.0, MutView{m1.ptr, d, 1}, s);
        MPISDF_HIP(hipGetLastError());
    }
    // beta = G2^-1 sigma; anchor adjoint 2 (lam_p M1_p + beta_p (D phi_p) - [(G o (beta lam^T + lam beta^T)) Phi_a]_p).
    solve_g2(f, sigma.ptr, tmp.ptr, s);
    gemm_tn_f64(k, k, d, 1.0, f.phi_a.ptr, d, f.phi_a.ptr, d, 0.0, gm.ptr, k, s);  // G, symmetric
    hipLaunchKernelGGL(sym_outer_hadamard_kernel, blocks(k * k), dim3(256), 0, s, gm.ptr, sigma.ptr, lam.ptr, k,
                       gm.ptr);
    gemm_f64(k, d, k, 1.0, MatView{gm.ptr, k, 1}, MatView{f.phi_a.ptr, d, 1}, 0.0, MutView{m2.ptr, d, 1}, s);
    hipLaunchKernelGGL(anchor_energy_adjoint_kernel, blocks(k * d), dim3(256), 0, s, m1.ptr, lam.ptr, y.ptr,
                       sigma.ptr, m2.ptr, k, d, abar.ptr);
    MPISDF_HIP(hipGetLastError());
    launch_ao_adjoint(ctx, anchor_xyz, k, abar.ptr, d, gao.ptr, adj_part, s);
    MPISDF_HIP(hipMemcpyAsync(gao_out, gao.ptr, 3 * nao * sizeof(double), hipMemcpyDeviceToHost, s));
    MPISDF_HIP(hipStreamSynchronize(s));
}

}  // namespace mpisdf
