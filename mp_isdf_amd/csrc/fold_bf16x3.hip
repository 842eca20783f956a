// QRCP fold on FP32 sketches with split-BF16 MFMA ("BF16x3").
//
// W = Qb^T Y with Qb (m x b, FP64) and Y (m x n, FP32), consumed as
//     proj2[j] += sum_t W[t,j]^2  (and optionally energy[t] += sum_j W[t,j]^2).
// Each operand is split into bf16 hi + lo parts (x = hi + lo + O(2^-16 x)) and
// W accumulates hi*hi + hi*lo + lo*hi on v_mfma_f32_32x32x16_bf16.  The FP32
// MFMA accumulators are flushed into FP64 registers every 64 K values, so the
// error of each W entry stays at the 1e-6 level of ||q|| ||y_j||; the QRCP
// carries that bound as its certification margin.
//
// The fold streams Y once per block of pivots; with b ~ 100 rows it is HBM
// bound (4 bytes of Y per 6 b flops), so the schedule maps all M tiles of a
// column tile onto one XCD (shared L2) and reads Y with 16-byte loads.
#include "kernels.hpp"

#include <algorithm>

namespace mpisdf {
namespace {

typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

constexpr int BM = 128, BN = 128, BK = 32, LDK = BK + 8;
constexpr int kFlushSlices = 2;  // FP32 accumulation over 64 K values

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo)
{
    hi = static_cast<__bf16>(x);
    lo = static_cast<__bf16>(x - static_cast<float>(hi));
}

// Qh/Ql[t * ldq + k] from Qb[k + t * ldqb] (FP64, column t), zero padded.
__global__ void split_q_kernel(const double* __restrict__ qb, int64_t ldqb, int64_t m, int64_t b, int64_t ldq,
                               int64_t bpad,
                               __bf16* __restrict__ qh, __bf16* __restrict__ ql)
{
    const int64_t total = bpad * ldq;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = i / ldq, k = i - t * ldq;
        const float x = (t < b && k < m) ? (float)qb[k + t * ldqb] : 0.f;
        __bf16 h, l;
        split_bf16(x, h, l);
        // hi/lo of the FP64 value: the lo part absorbs the FP32 rounding too.
        const double full = (t < b && k < m) ? qb[k + t * ldqb] : 0.0;
        l = static_cast<__bf16>((float)(full - (double)static_cast<float>(h)));
        qh[i] = h;
        ql[i] = l;
    }
}

template <bool ENERGY>
__global__ __launch_bounds__(256) void fold_bf16x3_kernel(const __bf16* __restrict__ qh, const __bf16* __restrict__ ql,
                                                          int64_t ldq, int64_t b, const float* __restrict__ y,
                                                          int64_t ldy, int64_t m, int64_t n, int64_t ntiles,
                                                          int64_t mtiles, double* __restrict__ proj2,
                                                          double* __restrict__ energy,
                                                          const double* __restrict__ ynorm2,
                                                          double* __restrict__ err2, double eps_w)
{
    __shared__ __align__(16) __bf16 sQh[BM * LDK];
    __shared__ __align__(16) __bf16 sQl[BM * LDK];
    __shared__ __align__(16) __bf16 sYh[BN * LDK];
    __shared__ __align__(16) __bf16 sYl[BN * LDK];
    __shared__ double colpart[2][BN];
    // XCD-aware tile order: workgroups id, id + 8, id + 16, ... share an XCD;
    // give them the same column tile so its Y panel is fetched into one L2.
    const int64_t id = blockIdx.x, xcd = id % 8, seq = id / 8;
    const int64_t ntile = (seq / mtiles) * 8 + xcd, mtile = seq % mtiles;
    if (ntile >= ntiles) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t m0 = mtile * BM, n0 = ntile * BN;

    f16v acc[2][2];
    double dacc[2][2][16];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) { acc[i][j][e] = 0.f; dacc[i][j][e] = 0.0; }

    // Per slice: Q tile 128 rows x 32 K (two bf16 planes) and Y tile 128
    // columns x 32 K floats; each thread moves 2 + 2 Q chunks and 4 Y chunks.
    uint4 rqh[2], rql[2];
    float4 ry[4];
    const bool vec = (ldy & 3) == 0;
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int c = tid + 256 * r;
            const int row = c >> 2, kc = (c & 3) * 8;
            const int64_t off = (m0 + row) * ldq + k0 + kc;  // padded: always in bounds
            rqh[r] = *reinterpret_cast<const uint4*>(qh + off);
            rql[r] = *reinterpret_cast<const uint4*>(ql + off);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = tid + 256 * r;
            const int col = c >> 3, kc = (c & 7) * 4;
            const int64_t j = n0 + col, k = k0 + kc;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (j < n) {
                const float* src = y + j * ldy + k;
                if (vec && k + 4 <= m) {
                    v = *reinterpret_cast<const float4*>(src);
                } else {
                    v.x = k < m ? src[0] : 0.f;
                    v.y = k + 1 < m ? src[1] : 0.f;
                    v.z = k + 2 < m ? src[2] : 0.f;
                    v.w = k + 3 < m ? src[3] : 0.f;
                }
            }
            ry[r] = v;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int c = tid + 256 * r;
            const int row = c >> 2, kc = (c & 3) * 8;
            *reinterpret_cast<uint4*>(sQh + row * LDK + kc) = rqh[r];
            *reinterpret_cast<uint4*>(sQl + row * LDK + kc) = rql[r];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = tid + 256 * r;
            const int col = c >> 3, kc = (c & 7) * 4;
            __bf16 h[4], l[4];
            split_bf16(ry[r].x, h[0], l[0]);
            split_bf16(ry[r].y, h[1], l[1]);
            split_bf16(ry[r].z, h[2], l[2]);
            split_bf16(ry[r].w, h[3], l[3]);
            *reinterpret_cast<uint2*>(sYh + col * LDK + kc) = *reinterpret_cast<uint2*>(h);
            *reinterpret_cast<uint2*>(sYl + col * LDK + kc) = *reinterpret_cast<uint2*>(l);
        }
    };

    const int64_t mk = (m + BK - 1) / BK * BK;
    load(0);
    int slice = 0;
    for (int64_t k0 = 0; k0 < mk; k0 += BK, ++slice) {
        __syncthreads();
        stash();
        __syncthreads();
        if (k0 + BK < mk) load(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            bf8v ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi) {
                const int row = wm * 64 + mi * 32 + (lane & 31);
                ah[mi] = *reinterpret_cast<const bf8v*>(sQh + row * LDK + kk + 8 * (lane >> 5));
                al[mi] = *reinterpret_cast<const bf8v*>(sQl + row * LDK + kk + 8 * (lane >> 5));
            }
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) {
                const int col = wn * 64 + ni * 32 + (lane & 31);
                bh[ni] = *reinterpret_cast<const bf8v*>(sYh + col * LDK + kk + 8 * (lane >> 5));
                bl[ni] = *reinterpret_cast<const bf8v*>(sYl + col * LDK + kk + 8 * (lane >> 5));
            }
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < 2; ++ni) {
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[mi], bh[ni], acc[mi][ni], 0, 0, 0);
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mi], bl[ni], acc[mi][ni], 0, 0, 0);
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mi], bh[ni], acc[mi][ni], 0, 0, 0);
                }
        }
        if ((slice + 1) % kFlushSlices == 0 || k0 + BK >= mk) {
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < 2; ++ni)
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        dacc[mi][ni][e] += (double)acc[mi][ni][e];
                        acc[mi][ni][e] = 0.f;
                    }
        }
    }

    // Epilogue.  Lane l, element e of a 32x32 tile: row 8(e/4) + 4(l/32) + e%4, column l%32.
#pragma unroll
    for (int ni = 0; ni < 2; ++ni) {
        double s = 0;
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int64_t row = m0 + wm * 64 + mi * 32 + 8 * (e / 4) + 4 * (lane >> 5) + (e % 4);
                const double w = row < b ? dacc[mi][ni][e] : 0.0;
                s += w * w;
            }
        s += __shfl_xor(s, 32);
        if (lane < 32) colpart[wm][wn * 64 + ni * 32 + lane] = s;
    }
    __syncthreads();
    if (tid < BN) {
        const int64_t j = n0 + tid;
        if (j < n) {
            const double sj = colpart[0][tid] + colpart[1][tid];
            atomicAdd(&proj2[j], sj);
            // Error of this fold's contribution: |w~ - w| <= eps_w ||y_j|| per
            // entry with independent signs -> 2 eps_w ||y_j|| sqrt(s_j) + b (eps_w ||y_j||)^2.
            if (err2 && mtiles == 1) {
                const double yn = ynorm2[j];
                err2[j] += 2.0 * eps_w * sqrt(yn * sj) + (double)b * eps_w * eps_w * yn;
            } else if (err2) {
                const double yn = ynorm2[j];
                atomicAdd(&err2[j], 2.0 * eps_w * sqrt(yn * sj) + (double)BM * eps_w * eps_w * yn);
            }
        }
    }
    if (ENERGY) {
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                double s = 0;
#pragma unroll
                for (int ni = 0; ni < 2; ++ni) {
                    const int64_t j = n0 + wn * 64 + ni * 32 + (lane & 31);
                    const double w = j < n ? dacc[mi][ni][e] : 0.0;
                    s += w * w;
                }
#pragma unroll
                for (int off = 1; off < 32; off <<= 1) s += __shfl_xor(s, off);
                const int64_t row = m0 + wm * 64 + mi * 32 + 8 * (e / 4) + 4 * (lane >> 5) + (e % 4);
                if ((lane & 31) == 0 && row < b) atomicAdd(&energy[row], s);
            }
    }
}

}  // namespace

void fold_projections_bf16x3(int64_t b, int64_t n, int64_t m, const double* Qb, int64_t ldqb, const float* Y,
                             int64_t ldy,
                             double* proj2, double* energy, FoldScratch& scratch, hipStream_t stream,
                             const double* ynorm2, double* err2)
{
    if (b <= 0 || n <= 0) return;
    const int64_t bpad = (b + BM - 1) / BM * BM;
    const int64_t ldq = (m + BK - 1) / BK * BK;
    const size_t need = (size_t)(bpad * ldq);
    if (scratch.hi.count < need) { scratch.hi.resize(need); scratch.lo.resize(need); }
    const int64_t total = bpad * ldq;
    hipLaunchKernelGGL(split_q_kernel, dim3((unsigned)std::min<int64_t>((total + 255) / 256, 65536)), dim3(256), 0,
                       stream, Qb, ldqb, m, b, ldq, bpad, reinterpret_cast<__bf16*>(scratch.hi.ptr),
                       reinterpret_cast<__bf16*>(scratch.lo.ptr));
    MPISDF_HIP(hipGetLastError());
    const int64_t ntiles = (n + BN - 1) / BN, mtiles = bpad / BM;
    const int64_t groups = (ntiles + 7) / 8;
    const int64_t blocks = groups * 8 * mtiles;
    if (blocks > 0x7fffffff) throw std::overflow_error("fold grid too large");
    if (energy)
        hipLaunchKernelGGL(fold_bf16x3_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, stream,
                           reinterpret_cast<const __bf16*>(scratch.hi.ptr), reinterpret_cast<const __bf16*>(scratch.lo.ptr),
                           ldq, b, Y, ldy, m, n, ntiles, mtiles, proj2, energy, ynorm2, err2, kFoldEpsW);
    else
        hipLaunchKernelGGL(fold_bf16x3_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, stream,
                           reinterpret_cast<const __bf16*>(scratch.hi.ptr), reinterpret_cast<const __bf16*>(scratch.lo.ptr),
                           ldq, b, Y, ldy, m, n, ntiles, mtiles, proj2, nullptr, ynorm2, err2, kFoldEpsW);
    MPISDF_HIP(hipGetLastError());
}

}  // namespace mpisdf
