// Khatri-Rao pair sketch on MFMA (gfx950), reduced-precision forms.
//
//   Y[g * ldy + s] = scale_g * (sum_a U[s,a] phi[g,a]) * (sum_a V[s,a] phi[g,a])
//
// (the reference's id_factored_sketch, cpp/src/anchor_select_randomized_id.cpp:299-309,
// with scale_g = sqrt(|w_g| / ell) times the inverse input scales).  Both
// projections share the phi tile, so one workgroup computes a 128 (s) x 128
// (g) block of U.phi and V.phi with two accumulator sets and multiplies them
// in the epilogue; Y is written once in FP32 and the FP64 column norms
// ||y_g||^2 needed by the QRCP are accumulated on the way out.
//
// Operands are K-contiguous ("NT"): Ut[s * d + a], Vt[s * d + a] and
// phi[g * ldphi + a], staged through LDS in K slices.
//   * BF16: v_mfma_f32_32x32x16_bf16, 4 waves as 2 x 2, each 64 x 64.
//   * FP8 e4m3 (OCP): v_mfma_scale_f32_16x16x128_f8f6f4 with unit block
//     scales; phi rows carry a per-point power-of-two scale chosen on
//     quantisation, Ut/Vt a global 2^5.
#include "sketch_mfma.hpp"

#include <algorithm>
#include <cmath>

namespace mpisdf {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef int i8v __attribute__((ext_vector_type(8)));

constexpr int BM = 128, BN = 128;

// ---------------------------------------------------------------------------
// Epilogue helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ void store_column_group(float* __restrict__ y, int64_t ldy, int64_t ell, int64_t s,
                                                   int64_t g, const float v[4])
{
    float* dst = y + g * ldy + s;
    if (s + 3 < ell && ((ldy & 3) == 0)) {
        *reinterpret_cast<float4*>(dst) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
        for (int i = 0; i < 4; ++i)
            if (s + i < ell) dst[i] = v[i];
    }
}

// ---------------------------------------------------------------------------
// BF16 kernel
// ---------------------------------------------------------------------------
constexpr int BK16 = 32;
constexpr int LDK16 = BK16 + 8;  // padded LDS row (bf16 elements)

__global__ __launch_bounds__(256) void kr_sketch_bf16_kernel(
    const __bf16* __restrict__ ut, const __bf16* __restrict__ vt, int64_t ell, int64_t d,
    const __bf16* __restrict__ phi, int64_t ldphi, int64_t count, const float* __restrict__ scale,
    float* __restrict__ y, int64_t ldy, double* __restrict__ ynorm2, unsigned long long* __restrict__ bad)
{
    __shared__ __align__(16) __bf16 sU[BM * LDK16];
    __shared__ __align__(16) __bf16 sV[BM * LDK16];
    __shared__ __align__(16) __bf16 sP[BN * LDK16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t s0 = (int64_t)blockIdx.y * BM, g0 = (int64_t)blockIdx.x * BN;

    f16v accU[2][2], accV[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) { accU[i][j][e] = 0.f; accV[i][j][e] = 0.f; }

    // Each thread moves 2 chunks of 8 bf16 per operand per K slice.
    uint4 ru[2], rv[2], rp[2];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int c = tid + 256 * r;
            const int row = c >> 2, kc = (c & 3) * 8;
            const int64_t k = k0 + kc;
            const bool kin = k + 8 <= d;
            const int64_t s = s0 + row, g = g0 + row;
            uint4 zero = make_uint4(0, 0, 0, 0);
            if (kin) {
                ru[r] = s < ell ? *reinterpret_cast<const uint4*>(ut + s * d + k) : zero;
                rv[r] = s < ell ? *reinterpret_cast<const uint4*>(vt + s * d + k) : zero;
                rp[r] = g < count ? *reinterpret_cast<const uint4*>(phi + g * ldphi + k) : zero;
            } else {
                // K tail: element-wise with zero fill.
                __bf16 tu[8], tv[8], tp[8];
                for (int e = 0; e < 8; ++e) {
                    const bool ok = k + e < d;
                    tu[e] = (ok && s < ell) ? ut[s * d + k + e] : (__bf16)0.f;
                    tv[e] = (ok && s < ell) ? vt[s * d + k + e] : (__bf16)0.f;
                    tp[e] = (ok && g < count) ? phi[g * ldphi + k + e] : (__bf16)0.f;
                }
                ru[r] = *reinterpret_cast<uint4*>(tu);
                rv[r] = *reinterpret_cast<uint4*>(tv);
                rp[r] = *reinterpret_cast<uint4*>(tp);
            }
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int c = tid + 256 * r;
            const int row = c >> 2, kc = (c & 3) * 8;
            *reinterpret_cast<uint4*>(sU + row * LDK16 + kc) = ru[r];
            *reinterpret_cast<uint4*>(sV + row * LDK16 + kc) = rv[r];
            *reinterpret_cast<uint4*>(sP + row * LDK16 + kc) = rp[r];
        }
    };

    load(0);
    for (int64_t k0 = 0; k0 < d; k0 += BK16) {
        __syncthreads();
        stash();
        __syncthreads();
        if (k0 + BK16 < d) load(k0 + BK16);  // prefetch next slice into registers
#pragma unroll
        for (int kk = 0; kk < BK16; kk += 16) {
            bf8v a_u[2], a_v[2], b_p[2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi) {
                const int row = wm * 64 + mi * 32 + (lane & 31);
                a_u[mi] = *reinterpret_cast<const bf8v*>(sU + row * LDK16 + kk + 8 * (lane >> 5));
                a_v[mi] = *reinterpret_cast<const bf8v*>(sV + row * LDK16 + kk + 8 * (lane >> 5));
            }
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) {
                const int row = wn * 64 + ni * 32 + (lane & 31);
                b_p[ni] = *reinterpret_cast<const bf8v*>(sP + row * LDK16 + kk + 8 * (lane >> 5));
            }
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < 2; ++ni) {
                    accU[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_u[mi], b_p[ni], accU[mi][ni], 0, 0, 0);
                    accV[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_v[mi], b_p[ni], accV[mi][ni], 0, 0, 0);
                }
        }
    }

    // Epilogue: D[8(i/4) + 4(lane/32) + i%4][lane%32] per 32x32 tile.
    unsigned long long nonfinite = 0;
#pragma unroll
    for (int ni = 0; ni < 2; ++ni) {
        const int64_t g = g0 + wn * 64 + ni * 32 + (lane & 31);
        const bool gin = g < count;
        const float sc = gin ? scale[g] : 0.f;
        double colsq = 0;
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) {
#pragma unroll
            for (int grp = 0; grp < 4; ++grp) {
                const int64_t s = s0 + wm * 64 + mi * 32 + 8 * grp + 4 * (lane >> 5);
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    v[e] = sc * accU[mi][ni][4 * grp + e] * accV[mi][ni][4 * grp + e];
                    if (s + e < ell && gin) {
                        colsq += (double)v[e] * (double)v[e];
                        nonfinite += !isfinite(v[e]);
                    }
                }
                if (gin) store_column_group(y, ldy, ell, s, g, v);
            }
        }
        colsq += __shfl_xor(colsq, 32);
        if (gin && lane < 32) atomicAdd(&ynorm2[g], colsq);
    }
    if (nonfinite) atomicAdd(bad, nonfinite);
}

// ---------------------------------------------------------------------------
// FP8 (e4m3) kernel: 16x16x128 MFMA, per wave 64 x 64 = 4 x 4 tiles.
// ---------------------------------------------------------------------------
constexpr int BK8 = 128;
constexpr int LDK8 = BK8 + 16;  // padded LDS row (bytes)

__global__ __launch_bounds__(256) void kr_sketch_fp8_kernel(
    const uint8_t* __restrict__ ut, const uint8_t* __restrict__ vt, int64_t ell, int64_t d,
    const uint8_t* __restrict__ phi, int64_t ldphi, int64_t count, const float* __restrict__ scale,
    float* __restrict__ y, int64_t ldy, double* __restrict__ ynorm2, unsigned long long* __restrict__ bad)
{
    __shared__ __align__(16) uint8_t sU[BM * LDK8];
    __shared__ __align__(16) uint8_t sV[BM * LDK8];
    __shared__ __align__(16) uint8_t sP[BN * LDK8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t s0 = (int64_t)blockIdx.y * BM, g0 = (int64_t)blockIdx.x * BN;

    f4 accU[4][4], accV[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) { accU[i][j][e] = 0.f; accV[i][j][e] = 0.f; }

    // Tile = 128 rows x 128 bytes per operand = 1024 chunks of 16 B; 4 per thread.
    uint4 ru[4], rv[4], rp[4];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = tid + 256 * r;
            const int row = c >> 3, kc = (c & 7) * 16;
            const int64_t k = k0 + kc;
            const int64_t s = s0 + row, g = g0 + row;
            const uint4 zero = make_uint4(0, 0, 0, 0);
            if (k + 16 <= d) {
                ru[r] = s < ell ? *reinterpret_cast<const uint4*>(ut + s * d + k) : zero;
                rv[r] = s < ell ? *reinterpret_cast<const uint4*>(vt + s * d + k) : zero;
                rp[r] = g < count ? *reinterpret_cast<const uint4*>(phi + g * ldphi + k) : zero;
            } else {
                uint8_t tu[16], tv[16], tp[16];
                for (int e = 0; e < 16; ++e) {
                    const bool ok = k + e < d;
                    tu[e] = (ok && s < ell) ? ut[s * d + k + e] : 0;
                    tv[e] = (ok && s < ell) ? vt[s * d + k + e] : 0;
                    tp[e] = (ok && g < count) ? phi[g * ldphi + k + e] : 0;
                }
                ru[r] = *reinterpret_cast<uint4*>(tu);
                rv[r] = *reinterpret_cast<uint4*>(tv);
                rp[r] = *reinterpret_cast<uint4*>(tp);
            }
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = tid + 256 * r;
            const int row = c >> 3, kc = (c & 7) * 16;
            *reinterpret_cast<uint4*>(sU + row * LDK8 + kc) = ru[r];
            *reinterpret_cast<uint4*>(sV + row * LDK8 + kc) = rv[r];
            *reinterpret_cast<uint4*>(sP + row * LDK8 + kc) = rp[r];
        }
    };
    // Lane's 32 K bytes: 32 * (lane / 16) .. +31, read as two 16-byte halves.
    auto frag = [&](const uint8_t* base, int row) {
        const uint8_t* p = base + row * LDK8 + 32 * (lane >> 4);
        const uint4 lo = *reinterpret_cast<const uint4*>(p);
        const uint4 hi = *reinterpret_cast<const uint4*>(p + 16);
        i8v v;
        v[0] = (int)lo.x; v[1] = (int)lo.y; v[2] = (int)lo.z; v[3] = (int)lo.w;
        v[4] = (int)hi.x; v[5] = (int)hi.y; v[6] = (int)hi.z; v[7] = (int)hi.w;
        return v;
    };

    load(0);
    for (int64_t k0 = 0; k0 < d; k0 += BK8) {
        __syncthreads();
        stash();
        __syncthreads();
        if (k0 + BK8 < d) load(k0 + BK8);
        i8v b_p[4];
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) b_p[ni] = frag(sP, wn * 64 + ni * 16 + (lane & 15));
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) {
            const i8v a_u = frag(sU, wm * 64 + mi * 16 + (lane & 15));
            const i8v a_v = frag(sV, wm * 64 + mi * 16 + (lane & 15));
#pragma unroll
            for (int ni = 0; ni < 4; ++ni) {
                accU[mi][ni] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a_u, b_p[ni], accU[mi][ni], 0, 0, 0, 127,
                                                                                 0, 127);
                accV[mi][ni] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a_v, b_p[ni], accV[mi][ni], 0, 0, 0, 127,
                                                                                 0, 127);
            }
        }
    }

    // Epilogue: D[4(lane/16) + e][lane%16] per 16x16 tile.
    unsigned long long nonfinite = 0;
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
        const int64_t g = g0 + wn * 64 + ni * 16 + (lane & 15);
        const bool gin = g < count;
        const float sc = gin ? scale[g] : 0.f;
        double colsq = 0;
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) {
            const int64_t s = s0 + wm * 64 + mi * 16 + 4 * (lane >> 4);
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v[e] = sc * accU[mi][ni][e] * accV[mi][ni][e];
                if (s + e < ell && gin) {
                    colsq += (double)v[e] * (double)v[e];
                    nonfinite += !isfinite(v[e]);
                }
            }
            if (gin) store_column_group(y, ldy, ell, s, g, v);
        }
        colsq += __shfl_xor(colsq, 16);
        colsq += __shfl_xor(colsq, 32);
        if (gin && lane < 16) atomicAdd(&ynorm2[g], colsq);
    }
    if (nonfinite) atomicAdd(bad, nonfinite);
}

// ---------------------------------------------------------------------------
// Operand preparation
// ---------------------------------------------------------------------------
// out[s * ldo + a] = cast(scale * in[a * ell + s]) for a < d, zero for d <= a < ldo
template <class T>
__global__ void transpose_cast_kernel(const double* __restrict__ in, int64_t d, int64_t ell, int64_t ldo, float scale,
                                      T* __restrict__ out)
{
    __shared__ float tile[32][33];
    const int64_t a0 = (int64_t)blockIdx.y * 32, s0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 8 rows per pass
    for (int r = ty; r < 32; r += 8) {
        const int64_t a = a0 + r, s = s0 + tx;
        tile[r][tx] = (a < d && s < ell) ? (float)in[a * ell + s] * scale : 0.f;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t s = s0 + r, a = a0 + tx;
        if (s < ell && a < ldo) {
            if constexpr (sizeof(T) == 1) {
                const int packed = __builtin_amdgcn_cvt_pk_fp8_f32(tile[tx][r], 0.f, 0, false);
                out[s * ldo + a] = (uint8_t)(packed & 0xff);
            } else {
                out[s * ldo + a] = static_cast<T>(tile[tx][r]);
            }
        }
    }
}

// Per point g: e_g = largest power of two with max|phi_g| * 2^e <= 256; fp8
// row = phi_g * 2^e; scale_g = sqrt(|w_g| / ell) * 2^(-2 e_g) * post.
__global__ void quantize_fp8_rows_kernel(const float* __restrict__ phi, int64_t ld, int64_t count, int64_t d,
                                         const double* __restrict__ sw, double inv_sqrt_ell, double post,
                                         uint8_t* __restrict__ out, int64_t ldo, float* __restrict__ scale)
{
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (g >= count) return;
    const float* row = phi + g * ld;
    float mx = 0.f;
    for (int64_t a = lane; a < d; a += 64) mx = fmaxf(mx, fabsf(row[a]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    int e = 0;
    if (mx > 0.f) {
        int ex;
        frexpf(mx, &ex);      // mx = m * 2^ex, m in [0.5, 1)
        e = 8 - ex;           // mx * 2^e in [128, 256)
    }
    const float mul = ldexpf(1.f, e);
    uint8_t* dst = out + g * ldo;
    for (int64_t a = lane; a < d; a += 64) {
        const int packed = __builtin_amdgcn_cvt_pk_fp8_f32(row[a] * mul, 0.f, 0, false);
        dst[a] = (uint8_t)(packed & 0xff);
    }
    if (lane == 0) scale[g] = (float)(sw[g] * inv_sqrt_ell * post * ldexp(1.0, -2 * e));
}

__global__ void bf16_scale_kernel(const double* __restrict__ sw, int64_t count, double inv_sqrt_ell,
                                  float* __restrict__ scale)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < count) scale[g] = (float)(sw[g] * inv_sqrt_ell);
}

// out[g * ldo + a] = cast(in[g * d + a]), zero padding for d <= a < ldo
template <class T>
__global__ void cast_rows_kernel(const double* __restrict__ in, int64_t count, int64_t d, int64_t ldo,
                                 T* __restrict__ out)
{
    const int64_t total = count * ldo;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = i / ldo, a = i - g * ldo;
        out[i] = static_cast<T>(a < d ? (float)in[g * d + a] : 0.f);
    }
}

int grid_for(int64_t work, int64_t cap = 65536)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((work + 255) / 256, cap));
}

}  // namespace

void prepare_omega_lowp(const double* u, const double* v, int64_t d, int64_t ell, int64_t ldo, DType dtype,
                        void* ut, void* vt, hipStream_t s)
{
    const dim3 grid((unsigned)((ell + 31) / 32), (unsigned)((ldo + 31) / 32));
    if (dtype == kBF16) {
        hipLaunchKernelGGL(transpose_cast_kernel<__bf16>, grid, dim3(256), 0, s, u, d, ell, ldo, 1.0f,
                           static_cast<__bf16*>(ut));
        hipLaunchKernelGGL(transpose_cast_kernel<__bf16>, grid, dim3(256), 0, s, v, d, ell, ldo, 1.0f,
                           static_cast<__bf16*>(vt));
    } else if (dtype == kFP8) {
        hipLaunchKernelGGL(transpose_cast_kernel<uint8_t>, grid, dim3(256), 0, s, u, d, ell, ldo, kOmegaFp8Scale,
                           static_cast<uint8_t*>(ut));
        hipLaunchKernelGGL(transpose_cast_kernel<uint8_t>, grid, dim3(256), 0, s, v, d, ell, ldo, kOmegaFp8Scale,
                           static_cast<uint8_t*>(vt));
    } else {
        throw std::invalid_argument("reduced-precision omega must be bf16 or fp8");
    }
    MPISDF_HIP(hipGetLastError());
}

void bf16_point_scales(const double* sw, int64_t count, int64_t ell, float* scale, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(bf16_scale_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, sw, count,
                       1.0 / std::sqrt((double)ell), scale);
    MPISDF_HIP(hipGetLastError());
}

void quantize_fp8_rows(const float* phi, int64_t ld, int64_t count, int64_t d, const double* sw, int64_t ell,
                       uint8_t* out, int64_t ldo, float* scale, hipStream_t s)
{
    if (count <= 0) return;
    const double post = 1.0 / ((double)kOmegaFp8Scale * (double)kOmegaFp8Scale);
    hipLaunchKernelGGL(quantize_fp8_rows_kernel, dim3((unsigned)((count + 3) / 4)), dim3(256), 0, s, phi, ld, count,
                       d, sw, 1.0 / std::sqrt((double)ell), post, out, ldo, scale);
    MPISDF_HIP(hipGetLastError());
}

void cast_rows(const double* in, int64_t count, int64_t d, int64_t ldo, DType dtype, void* out, hipStream_t s)
{
    if (count <= 0) return;
    if (dtype == kBF16)
        hipLaunchKernelGGL(cast_rows_kernel<__bf16>, dim3(grid_for(count * ldo)), dim3(256), 0, s, in, count, d, ldo,
                           static_cast<__bf16*>(out));
    else if (dtype == kF32)
        hipLaunchKernelGGL(cast_rows_kernel<float>, dim3(grid_for(count * ldo)), dim3(256), 0, s, in, count, d, ldo,
                           static_cast<float*>(out));
    else
        throw std::invalid_argument("cast_rows supports bf16 and f32");
    MPISDF_HIP(hipGetLastError());
}

void kr_sketch_lowp(DType dtype, const void* ut, const void* vt, int64_t ell, int64_t d, const void* phi, int64_t ldphi,
                    int64_t count, const float* scale, float* y, int64_t ldy, double* ynorm2, unsigned long long* bad,
                    hipStream_t s)
{
    if (count <= 0) return;
    const int64_t gx = (count + BN - 1) / BN, gy = (ell + BM - 1) / BM;
    if (gx > 0x7fffffff || gy > 65535) throw std::overflow_error("sketch grid too large");
    if (dtype == kBF16) {
        if (ldphi % 8 || d % 8) throw std::invalid_argument("bf16 sketch needs d and ldphi multiples of 8");
        hipLaunchKernelGGL(kr_sketch_bf16_kernel, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s,
                           static_cast<const __bf16*>(ut), static_cast<const __bf16*>(vt), ell, d,
                           static_cast<const __bf16*>(phi), ldphi, count, scale, y, ldy, ynorm2, bad);
    } else if (dtype == kFP8) {
        if (ldphi % 16 || d % 16) throw std::invalid_argument("fp8 sketch needs d and ldphi multiples of 16");
        hipLaunchKernelGGL(kr_sketch_fp8_kernel, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s,
                           static_cast<const uint8_t*>(ut), static_cast<const uint8_t*>(vt), ell, d,
                           static_cast<const uint8_t*>(phi), ldphi, count, scale, y, ldy, ynorm2, bad);
    } else {
        throw std::invalid_argument("reduced-precision sketch must be bf16 or fp8");
    }
    MPISDF_HIP(hipGetLastError());
}

}  // namespace mpisdf
