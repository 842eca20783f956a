// Khatri-Rao pair sketch on MFMA (gfx950), reduced-precision forms.
//
//   Y[g * ldy + s] = scale_g * (sum_a U[s,a] phi[g,a]) * (sum_a V[s,a] phi[g,a])
//
// (the reference's id_factored_sketch, cpp/src/anchor_select_randomized_id.cpp:299-309,
// with scale_g = sqrt(|w_g| / ell) times the inverse input scales).  Both
// projections share the phi tile, so one workgroup computes U.phi and V.phi
// with two accumulator sets and multiplies them in the epilogue; Y is written
// once in FP32 and the FP64 column norms ||y_g||^2 needed by the QRCP are
// accumulated on the way out.
//
// Operands are K-contiguous: Ut[s * d + a], Vt[s * d + a] and
// phi[g * ldphi + a].  BF16 and FP8 e4m3 (OCP) inputs, FP32 accumulation;
// FP8 phi rows carry a per-point power-of-two scale chosen on quantisation,
// Ut/Vt a global 2^5.  The split forms lay their terms side by side along K
// (sketch_mfma.hpp), so one K sum forms every product term.
#include "sketch_mfma.hpp"
#include "util.hpp"

#include <algorithm>
#include <cmath>

namespace mpisdf {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef int i8v __attribute__((ext_vector_type(8)));


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

// Sketch kernel (BF16 and FP8)
//
// Workgroup: 256 Omega rows (s) x 128 grid points (g); wave w owns rows
// 64 w .. 64 w + 63 for both U and V and all 128 points, i.e. 2 x 4 MFMA
// tiles of 32 x 32 per operand (256 FP32 accumulators per lane).  K advances
// in slices of 128 bytes (64 BF16 / 128 FP8 values).
//
// Both MFMA operands are read K-contiguously.  Lane l feeds row / column
// l % 32; the K order inside a slice is permuted so that each lane consumes
// the 64 consecutive bytes at offset 64 (l / 32) of the slice (the reduction
// is indifferent to the order, as long as A and B agree).  The U / V
// fragments therefore go straight from memory into registers, one slice
// ahead; only the phi slice, shared by the four waves, is staged in
// double-buffered LDS with a 144-byte row pitch (conflict-free 16-byte reads).
//   BF16: v_mfma_f32_32x32x16_bf16, 4 steps of 16 bytes per lane per slice.
//   FP8:  v_mfma_scale_f32_32x32x64_f8f6f4 (e4m3, unit block scales),
//         2 steps of 32 bytes per lane per slice.
// Result element i of lane l: D[8 (i / 4) + 4 (l / 32) + i % 4][l % 32].
// ---------------------------------------------------------------------------
constexpr int SK_S = 256;        // Omega rows per workgroup
constexpr int SK_G = 128;        // grid points per workgroup
constexpr int SK_SLICE = 128;    // K bytes per slice
constexpr int SK_PITCH = 144;    // LDS row pitch of the phi slice (bytes)
// Word 3 of a raw buffer descriptor on gfx9-class GPUs (32-bit data format).
constexpr int kBufferWord3 = 0x00020000;

template <int FMT>  // 0: BF16, 1: FP8 e4m3
__global__ __launch_bounds__(256) void kr_sketch_kernel(
    const uint8_t* __restrict__ ut, const uint8_t* __restrict__ vt, int64_t ell, int64_t dbytes,
    const uint8_t* __restrict__ phi, int64_t ldphi_bytes, int64_t count, const float* __restrict__ scale,
    float* __restrict__ y, int64_t ldy, double* __restrict__ ynorm2, unsigned long long* __restrict__ bad)
{
    __shared__ __align__(16) uint8_t sP[2][SK_G * SK_PITCH];
    __shared__ double colred[4][SK_G];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, l32 = lane & 31;
    const int64_t s0 = (int64_t)blockIdx.y * SK_S + wave * 64, g0 = (int64_t)blockIdx.x * SK_G;

    f16v accU[2][4], accV[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int gt = 0; gt < 4; ++gt)
#pragma unroll
            for (int e = 0; e < 16; ++e) { accU[t][gt][e] = 0.f; accV[t][gt][e] = 0.f; }

    // Raw buffer loads: an offset past the descriptor's range returns zeros,
    // which pads rows past ell / count and K past d without branches.  The
    // host guarantees every offset below fits in 32 bits.
    const uint32_t db = (uint32_t)dbytes;
    const __amdgpu_buffer_rsrc_t ursrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(ut), (short)0, (int)(ell * dbytes), kBufferWord3);
    const __amdgpu_buffer_rsrc_t vrsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(vt), (short)0, (int)(ell * dbytes), kBufferWord3);
    const int64_t prows = std::min<int64_t>(SK_G, count - g0);
    const __amdgpu_buffer_rsrc_t prsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint8_t*>(phi + g0 * ldphi_bytes), (short)0, (int)(prows * ldphi_bytes), kBufferWord3);
    uint32_t uoff[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) uoff[t] = (uint32_t)(s0 + 32 * t + l32) * db;  // >= range for rows past ell
    // phi chunks staged by this thread: chunk c = tid + 256 r -> row c / 8, byte 16 (c % 8).
    const int prow = tid >> 3, pkc = (tid & 7) * 16;

    // ua/va[t][c]: 16-byte chunk c of this lane's 64 bytes of the current
    // slice.  A chunk is reloaded for the next slice right after the step that
    // consumed it, so one register set covers both slices.
    uint4 ua[2][4], va[2][4];
    uint4 rp[4];
#define MPISDF_LOAD_UV(k0, t, c)                                                              \
    do {                                                                                     \
        const uint32_t k_ = (uint32_t)(k0) + 64u * half + 16u * (c);                         \
        const uint32_t off_ = k_ < db ? uoff[t] + k_ : 0x80000000u;                          \
        ua[t][c] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(ursrc, off_, 0, 0)); \
        va[t][c] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(vrsrc, off_, 0, 0)); \
    } while (0)
#define MPISDF_LOAD_P(k0)                                                                     \
    do {                                                                                     \
        _Pragma("unroll") for (int r_ = 0; r_ < 4; ++r_) {                                   \
            const uint32_t k_ = (uint32_t)(k0) + (uint32_t)pkc;                              \
            const uint32_t off_ = k_ < db ? (uint32_t)(prow + 32 * r_) * (uint32_t)ldphi_bytes + k_ : 0x80000000u; \
            rp[r_] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(prsrc, off_, 0, 0)); \
        }                                                                                    \
    } while (0)
#define MPISDF_STASH_P(buf)                                                                   \
    do {                                                                                     \
        _Pragma("unroll") for (int r_ = 0; r_ < 4; ++r_)                                     \
            *reinterpret_cast<uint4*>(&sP[buf][(prow + 32 * r_) * SK_PITCH + pkc]) = rp[r_];  \
    } while (0)

    const int64_t nsl = (dbytes + SK_SLICE - 1) / SK_SLICE;
    if (nsl > 0) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int c = 0; c < 4; ++c) MPISDF_LOAD_UV(0, t, c);
        MPISDF_LOAD_P(0);
        MPISDF_STASH_P(0);
    }
    __syncthreads();
    for (int64_t sl = 0; sl < nsl; ++sl) {
        const int buf = (int)(sl & 1);
        // Prefetches run unconditionally: past the last slice their offsets
        // fall outside the descriptors and return zeros (no branch, so the
        // wait counts stay exact).
        const int64_t knext = (sl + 1) * SK_SLICE;
        MPISDF_LOAD_P(knext);
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the MFMAs
        const uint8_t* P = sP[buf];
        if (FMT == 0) {
            // phi operands of K-step st + 1 are read from LDS while step st's
            // MFMAs run (one wave per SIMD: nothing else hides that latency)
            bf8v b[2][4];
#pragma unroll
            for (int gt = 0; gt < 4; ++gt)
                b[0][gt] = *reinterpret_cast<const bf8v*>(P + (gt * 32 + l32) * SK_PITCH + 64 * half);
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                if (st < 3) {
#pragma unroll
                    for (int gt = 0; gt < 4; ++gt)
                        b[(st + 1) & 1][gt] = *reinterpret_cast<const bf8v*>(P + (gt * 32 + l32) * SK_PITCH +
                                                                            64 * half + 16 * (st + 1));
                }
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const bf8v au = __builtin_bit_cast(bf8v, ua[t][st]);
                    const bf8v av = __builtin_bit_cast(bf8v, va[t][st]);
#pragma unroll
                    for (int gt = 0; gt < 4; ++gt) {
                        accU[t][gt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(au, b[st & 1][gt], accU[t][gt], 0, 0, 0);
                        accV[t][gt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b[st & 1][gt], accV[t][gt], 0, 0, 0);
                    }
                    MPISDF_LOAD_UV(knext, t, st);
                }
                // next step's reads first, then MFMAs with this step's loads between them
                if (st < 3) __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
#pragma unroll
                for (int i_ = 0; i_ < 4; ++i_) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
                    __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
#pragma unroll
            for (int st = 0; st < 2; ++st) {
                i8v b[4];
#pragma unroll
                for (int gt = 0; gt < 4; ++gt) {
                    const uint8_t* p = P + (gt * 32 + l32) * SK_PITCH + 64 * half + 32 * st;
                    const uint4 lo = *reinterpret_cast<const uint4*>(p);
                    const uint4 hi = *reinterpret_cast<const uint4*>(p + 16);
                    b[gt] = i8v{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w, (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
                }
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const uint4 ul = ua[t][2 * st], uh = ua[t][2 * st + 1];
                    const uint4 vl = va[t][2 * st], vh = va[t][2 * st + 1];
                    const i8v au = i8v{(int)ul.x, (int)ul.y, (int)ul.z, (int)ul.w, (int)uh.x, (int)uh.y, (int)uh.z, (int)uh.w};
                    const i8v av = i8v{(int)vl.x, (int)vl.y, (int)vl.z, (int)vl.w, (int)vh.x, (int)vh.y, (int)vh.z, (int)vh.w};
#pragma unroll
                    for (int gt = 0; gt < 4; ++gt) {
                        accU[t][gt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(au, b[gt], accU[t][gt], 0, 0, 0,
                                                                                      127, 0, 127);
                        accV[t][gt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, b[gt], accV[t][gt], 0, 0, 0,
                                                                                      127, 0, 127);
                    }
                    MPISDF_LOAD_UV(knext, t, 2 * st);
                    MPISDF_LOAD_UV(knext, t, 2 * st + 1);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
        MPISDF_STASH_P(buf ^ 1);
        __syncthreads();
    }
#undef MPISDF_LOAD_UV
#undef MPISDF_LOAD_P
#undef MPISDF_STASH_P

    // Epilogue: y = scale_g * (U phi) * (V phi), FP32 stores of four
    // consecutive s per lane; column sums of y^2 reduced over the workgroup.
    unsigned long long nonfinite = 0;
#pragma unroll
    for (int gt = 0; gt < 4; ++gt) {
        const int64_t g = g0 + gt * 32 + l32;
        const bool gin = g < count;
        const float sc = gin ? scale[g] : 0.f;
        double colsq = 0;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
#pragma unroll
            for (int grp = 0; grp < 4; ++grp) {
                const int64_t srow = s0 + 32 * t + 8 * grp + 4 * half;
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    v[e] = sc * accU[t][gt][4 * grp + e] * accV[t][gt][4 * grp + e];
                    if (srow + e < ell && gin) {
                        colsq += (double)v[e] * (double)v[e];
                        nonfinite += !isfinite(v[e]);
                    }
                }
                if (gin && srow < ell) store_column_group(y, ldy, ell, srow, g, v);
            }
        }
        colsq += __shfl_xor(colsq, 32);
        if (half == 0) colred[wave][gt * 32 + l32] = colsq;
    }
    __syncthreads();
    if (tid < SK_G) {
        // This workgroup's Omega rows' part of ||y_g||^2; the row blocks are
        // added in order by the host-side launcher.
        const int64_t g = g0 + tid;
        if (g < count)
            ynorm2[(int64_t)blockIdx.y * count + g] = colred[0][tid] + colred[1][tid] + colred[2][tid] + colred[3][tid];
    }
    if (nonfinite) atomicAdd(bad, nonfinite);
}

// ---------------------------------------------------------------------------
// Operand preparation
// ---------------------------------------------------------------------------
// out[s * ldo + a] = cast(scale * in[a * ell + s]) for a < d, zero for d <= a < ldo
template <class T>
__global__ void transpose_cast_kernel(const double* __restrict__ in, int64_t d, int64_t ell, int64_t ldo,
                                      int64_t pitch, float scale, T* __restrict__ out)
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
                out[s * pitch + a] = (uint8_t)(packed & 0xff);
            } else {
                out[s * pitch + a] = static_cast<T>(tile[tx][r]);
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
    // Two terms on the one scale: hi = e4m3(x), lo = e4m3(x - hi) (x - hi is
    // exact in FP32), at offsets 0 and ldo of the 2 ldo row.
    uint8_t* dst = out + g * 2 * ldo;
    for (int64_t a = lane; a < d; a += 64) {
        const float x = row[a] * mul;
        const int hi = __builtin_amdgcn_cvt_pk_fp8_f32(x, 0.f, 0, false);
        const float r = x - __builtin_amdgcn_cvt_f32_fp8(hi, 0);
        const int lo = __builtin_amdgcn_cvt_pk_fp8_f32(r, 0.f, 0, false);
        dst[a] = (uint8_t)(hi & 0xff);
        dst[ldo + a] = (uint8_t)(lo & 0xff);
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

// BF16 splits x = hi + lo of FP64 values: hi = bf16(x), lo = bf16(x - hi).
__device__ __forceinline__ void split_bf16(double x, __bf16& hi, __bf16& lo)
{
    hi = static_cast<__bf16>((float)x);
    lo = static_cast<__bf16>((float)(x - (double)(float)hi));
}

// Split-BF16 phi rows: out[g * 3 ldo + {0, ldo, 2 ldo} + a] = {hi, lo, hi} of
// in[g * ld + a] (zero for d <= a < ldo).
__global__ void split_rows_bf16x3_kernel(const double* __restrict__ in, int64_t ld, int64_t count, int64_t d,
                                         int64_t ldo, __bf16* __restrict__ out)
{
    const int64_t total = count * ldo;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = i / ldo, a = i - g * ldo;
        __bf16 hi, lo;
        split_bf16(a < d ? in[g * ld + a] : 0.0, hi, lo);
        __bf16* row = out + g * 3 * ldo;
        row[a] = hi;
        row[ldo + a] = lo;
        row[2 * ldo + a] = hi;
    }
}

// Split-BF16 Omega rows: out[s * 3 ldo + {0, ldo, 2 ldo} + a] = {hi, hi, lo}
// of in[a * ell + s] (zero for d <= a < ldo).
__global__ void split_omega_bf16x3_kernel(const double* __restrict__ in, int64_t d, int64_t ell, int64_t ldo,
                                          __bf16* __restrict__ out)
{
    __shared__ double tile[32][33];
    const int64_t a0 = (int64_t)blockIdx.y * 32, s0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int64_t a = a0 + r, s = s0 + tx;
        tile[r][tx] = (a < d && s < ell) ? in[a * ell + s] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t s = s0 + r, a = a0 + tx;
        if (s < ell && a < ldo) {
            __bf16 hi, lo;
            split_bf16(tile[tx][r], hi, lo);
            __bf16* row = out + s * 3 * ldo;
            row[a] = hi;
            row[ldo + a] = hi;
            row[2 * ldo + a] = lo;
        }
    }
}

int grid_for(int64_t work, int64_t cap = 65536)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((work + 255) / 256, cap));
}

}  // namespace

void prepare_omega_lowp(const double* u, const double* v, int64_t d, int64_t ell, int64_t ldo, int form,
                        void* ut, void* vt, hipStream_t s)
{
    if (ldo < d || ldo % 16) throw std::invalid_argument("omega segment width must be >= d and a multiple of 16");
    const dim3 grid((unsigned)((ell + 31) / 32), (unsigned)((ldo + 31) / 32));
    if (form == kSketchBF16) {
        hipLaunchKernelGGL(transpose_cast_kernel<__bf16>, grid, dim3(256), 0, s, u, d, ell, ldo, ldo, 1.0f,
                           static_cast<__bf16*>(ut));
        hipLaunchKernelGGL(transpose_cast_kernel<__bf16>, grid, dim3(256), 0, s, v, d, ell, ldo, ldo, 1.0f,
                           static_cast<__bf16*>(vt));
    } else if (form == kSketchFP8) {
        // [u | u]: the same e4m3 Omega row meets phi's hi and lo terms.
        for (int seg = 0; seg < 2; ++seg) {
            hipLaunchKernelGGL(transpose_cast_kernel<uint8_t>, grid, dim3(256), 0, s, u, d, ell, ldo, 2 * ldo,
                               kOmegaFp8Scale, static_cast<uint8_t*>(ut) + seg * ldo);
            hipLaunchKernelGGL(transpose_cast_kernel<uint8_t>, grid, dim3(256), 0, s, v, d, ell, ldo, 2 * ldo,
                               kOmegaFp8Scale, static_cast<uint8_t*>(vt) + seg * ldo);
        }
    } else if (form == kSketchBF16x3) {
        hipLaunchKernelGGL(split_omega_bf16x3_kernel, grid, dim3(256), 0, s, u, d, ell, ldo, static_cast<__bf16*>(ut));
        hipLaunchKernelGGL(split_omega_bf16x3_kernel, grid, dim3(256), 0, s, v, d, ell, ldo, static_cast<__bf16*>(vt));
    } else {
        throw std::invalid_argument("reduced-precision sketch form must be bf16, fp8 or bf16x3");
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

void split_rows_bf16x3(const double* in, int64_t ld, int64_t count, int64_t d, int64_t ldo, void* out, hipStream_t s)
{
    if (count <= 0) return;
    if (ldo < d || ld < d) throw std::invalid_argument("invalid split row strides");
    hipLaunchKernelGGL(split_rows_bf16x3_kernel, dim3(grid_for(count * ldo)), dim3(256), 0, s, in, ld, count, d, ldo,
                       static_cast<__bf16*>(out));
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

int64_t kr_sketch_lowp_partials(int64_t ell, int64_t count)
{
    return count > 0 ? (ell + SK_S - 1) / SK_S * count : 0;
}

void kr_sketch_lowp(DType dtype, const void* ut, const void* vt, int64_t ell, int64_t d, const void* phi, int64_t ldphi,
                    int64_t count, const float* scale, float* y, int64_t ldy, double* ynorm2, unsigned long long* bad,
                    DeviceBuffer<double>& part, hipStream_t s)
{
    const int64_t need = kr_sketch_lowp_partials(ell, count);
    if (part.count < (size_t)need) part.resize((size_t)need);
    kr_sketch_lowp_into(dtype, ut, vt, ell, d, phi, ldphi, count, scale, y, ldy, ynorm2, bad, part.ptr, s);
}

void kr_sketch_lowp_into(DType dtype, const void* ut, const void* vt, int64_t ell, int64_t d, const void* phi,
                         int64_t ldphi, int64_t count, const float* scale, float* y, int64_t ldy, double* ynorm2,
                         unsigned long long* bad, double* part, hipStream_t s)
{
    if (count <= 0) return;
    const int64_t gx = (count + SK_G - 1) / SK_G, gy = (ell + SK_S - 1) / SK_S;
    if (gx > 0x7fffffff || gy > 65535) throw std::overflow_error("sketch grid too large");
    // Buffer offsets inside the kernel are 32-bit: Omega rows (padded to the
    // 256-row tile) and one 128-point phi tile must stay below 2^31 bytes.
    const int64_t dbytes = dtype == kBF16 ? 2 * d : d, pbytes = dtype == kBF16 ? 2 * ldphi : ldphi;
    if ((ell + SK_S) * dbytes >= (int64_t(1) << 31) || SK_G * pbytes >= (int64_t(1) << 31))
        throw std::overflow_error("sketch operands exceed the 32-bit buffer range");
    const auto* u8 = static_cast<const uint8_t*>(ut);
    const auto* v8 = static_cast<const uint8_t*>(vt);
    const auto* p8 = static_cast<const uint8_t*>(phi);
    if (dtype == kBF16) {
        if (ldphi % 8 || d % 8) throw std::invalid_argument("bf16 sketch needs d and ldphi multiples of 8");
        hipLaunchKernelGGL(kr_sketch_kernel<0>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, u8, v8, ell, 2 * d,
                           p8, 2 * ldphi, count, scale, y, ldy, part, bad);
    } else if (dtype == kFP8) {
        if (ldphi % 16 || d % 16) throw std::invalid_argument("fp8 sketch needs d and ldphi multiples of 16");
        hipLaunchKernelGGL(kr_sketch_kernel<1>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, u8, v8, ell, d, p8,
                           ldphi, count, scale, y, ldy, part, bad);
    } else {
        throw std::invalid_argument("reduced-precision sketch must be bf16 or fp8");
    }
    MPISDF_HIP(hipGetLastError());
    launch_sum_partials(part, gy, count, ynorm2, true, s);
}

}  // namespace mpisdf
