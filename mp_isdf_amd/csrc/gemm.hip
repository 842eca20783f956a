// Strided FP64 GEMMs on MFMA (v_mfma_f64_16x16x4_f64) for gfx950.
//
// C[M,N] = alpha * A[M,K] B[K,N] + beta * C with arbitrary row/column strides
// (element (i, j) of a view at p[i * rs + j * cs]).  B may be FP32 (sketch
// matrices) and is widened exactly on load.  128 x 128 output tiles per
// 256-thread workgroup, four waves of 64 x 64 (4 x 4 MFMA tiles), K staged
// through LDS in slices of 16.  Products whose operands are both contiguous
// along K go through the TN kernel (gemm_tn.hip) instead.
//
// Operand lane layout of the 16x16x4 f64 MFMA: A[l % 16][l / 16],
// B[l / 16][l % 16]; result element i of lane l is D[l / 16 + 4 i][l % 16].
#include "kernels.hpp"

#include <algorithm>

namespace mpisdf {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 16;

enum Epilogue { kStore = 0 };

template <class TB, int EPI>
__global__ __launch_bounds__(256) void gemm_mfma_kernel(int64_t M, int64_t N, int64_t K, double alpha,
                                                        const double* __restrict__ Ap, int64_t ars, int64_t acs,
                                                        const TB* __restrict__ Bp, int64_t brs, int64_t bcs,
                                                        double beta, double* __restrict__ Cp, int64_t crs,
                                                        int64_t ccs, double* __restrict__ colacc,
                                                        double* __restrict__ rowacc)
{
    __shared__ double As[BK][BM + 2];
    __shared__ double Bs[BK][BN + 2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t m0 = (int64_t)blockIdx.y * BM, n0 = (int64_t)blockIdx.x * BN;
    d4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    double ra[8], rb[8];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int i = tid + 256 * r;
            // Consecutive threads walk each operand's unit-stride direction.
            const int mm = acs == 1 ? i / BK : i % BM, kk = acs == 1 ? i % BK : i / BM;
            const int64_t gm = m0 + mm, gk = k0 + kk;
            ra[r] = (gm < M && gk < K) ? Ap[gm * ars + gk * acs] : 0.0;
            const int nn = brs == 1 ? i / BK : i % BN, kb = brs == 1 ? i % BK : i / BN;
            const int64_t gn = n0 + nn, gkb = k0 + kb;
            rb[r] = (gn < N && gkb < K) ? (double)Bp[gkb * brs + gn * bcs] : 0.0;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int i = tid + 256 * r;
            const int mm = acs == 1 ? i / BK : i % BM, kk = acs == 1 ? i % BK : i / BM;
            As[kk][mm] = ra[r];
            const int nn = brs == 1 ? i / BK : i % BN, kb = brs == 1 ? i % BK : i / BN;
            Bs[kb][nn] = rb[r];
        }
    };

    load(0);
    for (int64_t k0 = 0; k0 < K; k0 += BK) {
        __syncthreads();
        stash();
        __syncthreads();
        if (k0 + BK < K) load(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double a[4], b[4];
#pragma unroll
            for (int mi = 0; mi < 4; ++mi) a[mi] = As[kk + (lane >> 4)][wm * 64 + mi * 16 + (lane & 15)];
#pragma unroll
            for (int ni = 0; ni < 4; ++ni) b[ni] = Bs[kk + (lane >> 4)][wn * 64 + ni * 16 + (lane & 15)];
#pragma unroll
            for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                for (int ni = 0; ni < 4; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mi], b[ni], acc[mi][ni], 0, 0, 0);
        }
    }

    if (EPI == kStore) {
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
            for (int ni = 0; ni < 4; ++ni)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int64_t gm = m0 + wm * 64 + mi * 16 + (lane >> 4) + 4 * i;
                    const int64_t gn = n0 + wn * 64 + ni * 16 + (lane & 15);
                    if (gm < M && gn < N) {
                        double* c = Cp + gm * crs + gn * ccs;
                        *c = alpha * acc[mi][ni][i] + (beta == 0.0 ? 0.0 : beta * *c);
                    }
                }
    }
}

template <class TB, int EPI>
void launch(int64_t M, int64_t N, int64_t K, double alpha, MatView A, const TB* B, int64_t brs, int64_t bcs,
            double beta, MutView C, double* colacc, double* rowacc, hipStream_t stream)
{
    if (M <= 0 || N <= 0) return;
    const int64_t gx = (N + BN - 1) / BN, gy = (M + BM - 1) / BM;
    if (gx > 0x7fffffff || gy > 65535) throw std::overflow_error("GEMM grid too large");
    hipLaunchKernelGGL((gemm_mfma_kernel<TB, EPI>), dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, stream, M, N,
                       std::max<int64_t>(K, 0), alpha, A.p, A.rs, A.cs, B, brs, bcs, beta, C.p, C.rs, C.cs, colacc,
                       rowacc);
    MPISDF_HIP(hipGetLastError());
}

}  // namespace

void gemm_f64(int64_t M, int64_t N, int64_t K, double alpha, MatView A, MatView B, double beta, MutView C,
              hipStream_t stream)
{
    launch<double, kStore>(M, N, K, alpha, A, B.p, B.rs, B.cs, beta, C, nullptr, nullptr, stream);
}

void gemm_f64_f32b(int64_t M, int64_t N, int64_t K, double alpha, MatView A, const float* B, int64_t brs,
                   int64_t bcs, double beta, MutView C, hipStream_t stream)
{
    launch<float, kStore>(M, N, K, alpha, A, B, brs, bcs, beta, C, nullptr, nullptr, stream);
}

}  // namespace mpisdf
