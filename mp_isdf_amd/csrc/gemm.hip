// Generic strided GEMMs with FP64 accumulation used by assembly, QRCP and fit.
//
// C[M,N] = alpha * A[M,K] B[K,N] + beta * C with arbitrary row/column
// strides.  64x64 output tiles per 256-thread workgroup, K staged through LDS
// in slices of 16; each thread accumulates a 4x4 register block.  B may be
// FP32 (sketch matrices) and is widened on load.
#include "kernels.hpp"

#include <algorithm>

namespace mpisdf {
namespace {

constexpr int BM = 64, BN = 64, BK = 16;

template <class TB>
__global__ __launch_bounds__(256) void gemm_kernel(int64_t M, int64_t N, int64_t K, double alpha,
                                                   const double* __restrict__ Ap, int64_t ars, int64_t acs,
                                                   const TB* __restrict__ Bp, int64_t brs, int64_t bcs,
                                                   double beta, double* __restrict__ Cp, int64_t crs,
                                                   int64_t ccs)
{
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN + 1];
    const int tid = threadIdx.x;
    const int64_t m0 = (int64_t)blockIdx.y * BM, n0 = (int64_t)blockIdx.x * BN;
    const int tm = (tid / 16) * 4, tn = (tid % 16) * 4;
    double acc[4][4] = {};
    for (int64_t k0 = 0; k0 < K; k0 += BK) {
        for (int i = tid; i < BM * BK; i += 256) {
            // Consecutive threads walk the operand's unit-stride direction.
            const int mm = acs == 1 ? i / BK : i % BM, kk = acs == 1 ? i % BK : i / BM;
            const int64_t gm = m0 + mm, gk = k0 + kk;
            As[kk][mm] = (gm < M && gk < K) ? Ap[gm * ars + gk * acs] : 0.0;
        }
        for (int i = tid; i < BN * BK; i += 256) {
            const int nn = brs == 1 ? i / BK : i % BN, kk = brs == 1 ? i % BK : i / BN;
            const int64_t gn = n0 + nn, gk = k0 + kk;
            Bs[kk][nn] = (gn < N && gk < K) ? (double)Bp[gk * brs + gn * bcs] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; ++kk) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) { a[i] = As[kk][tm + i]; b[i] = Bs[kk][tn + i]; }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t gm = m0 + tm + i, gn = n0 + tn + j;
            if (gm < M && gn < N) {
                double* c = Cp + gm * crs + gn * ccs;
                *c = alpha * acc[i][j] + (beta == 0.0 ? 0.0 : beta * *c);
            }
        }
}

template <class TB>
void launch(int64_t M, int64_t N, int64_t K, double alpha, MatView A, const TB* B, int64_t brs,
            int64_t bcs, double beta, MutView C, hipStream_t stream)
{
    if (M <= 0 || N <= 0) return;
    const int64_t gx = (N + BN - 1) / BN, gy = (M + BM - 1) / BM;
    if (gx > 0x7fffffff || gy > 65535) throw std::overflow_error("GEMM grid too large");
    hipLaunchKernelGGL(gemm_kernel<TB>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, stream, M, N,
                       std::max<int64_t>(K, 0), alpha, A.p, A.rs, A.cs, B, brs, bcs, beta, C.p, C.rs,
                       C.cs);
    MPISDF_HIP(hipGetLastError());
}

}  // namespace

void gemm_f64(int64_t M, int64_t N, int64_t K, double alpha, MatView A, MatView B, double beta,
              MutView C, hipStream_t stream)
{
    launch<double>(M, N, K, alpha, A, B.p, B.rs, B.cs, beta, C, stream);
}

void gemm_f64_f32b(int64_t M, int64_t N, int64_t K, double alpha, MatView A, const float* B,
                   int64_t brs, int64_t bcs, double beta, MutView C, hipStream_t stream)
{
    launch<float>(M, N, K, alpha, A, B, brs, bcs, beta, C, stream);
}

}  // namespace mpisdf
