// QRCP fold on INT8 MFMA with digit splitting, for sketches stored as
// signed base-128 digit planes.
//
//     W = Qb^T Y,  proj2[j] += sum_t W[t,j]^2  (energy[t] += sum_j W[t,j]^2)
//
// Y is held as four digit planes per column (digitize_columns, in place over
// the FP32 sketch, same footprint): with the power of two s_j >= 2 ||y_j||,
//     y_hat[i] = s_j sum_{r<4} d_r[i] 128^-(r+1),  |d_r| <= 64,
// which equals the FP32 entry exactly when it lies within 2^-4 of s_j and
// differs by at most 2^-29 s_j otherwise (a perturbation of ~1e-7 relative,
// FP32-level).  The QRCP then factorises y_hat: candidate columns are
// rebuilt from the same digits, and ||y_hat_j||^2 replaces ||y_j||^2.
//
// Products use the splitting method of Ozaki, Ogita, Oishi and Rump (2012) in
// the INT8 form of Ootomo, Ozaki and Yokota (2024): row t of Qb^T is written
// as five digits with its own power of two u_t >= 2 max|q_t|; the 17 digit
// pairs with p + r <= 5 run on v_mfma_i32_16x16x64_i8, pairs of equal p + r
// sharing an exact INT32 accumulator (|sum| < 4 * 64^2 * m < 2^31 for
// m < 2^16), and the six accumulators are combined in FP64:
//     W = u_t s_j sum_g 128^-(g+2) acc_g.
// The dropped pairs and the truncation of q leave ~1e-11 of ||y_j||^2 in the
// folded sums (measured), covered by the QRCP's relative margin.
//
// Lane layout of v_mfma_i32_16x16x64_i8: A row l % 16 and B column l % 16,
// each supplying 16 consecutive K bytes at offset 16 (l / 16); result element
// e of lane l is D[4 (l / 16) + e][l % 16].  Workgroup: 4 waves x 16 columns
// x 128 rows (8 row tiles; grid.y walks further row groups).  The Q digit
// slice (5 planes x 128 rows x 64 K) is double-buffered in LDS; Y digit
// fragments go straight from HBM to registers, one slice ahead.
#include "kernels.hpp"

#include <algorithm>

namespace mpisdf {
namespace {

typedef int i4v __attribute__((ext_vector_type(4)));

constexpr int ND = 5;            // digits of q
constexpr int NY = 4;            // digits of y (planes per column)
constexpr int NG = 6;            // accumulator groups, p + r = 0..5
constexpr int FBM = 128;         // rows per workgroup
constexpr int FBN = 64;          // columns per workgroup
constexpr int FBK = 64;          // K per slice (one MFMA)
constexpr int FPITCH = FBK + 16; // LDS row pitch (bytes): 80 -> conflict-free 16-byte reads

// Q digit planes: qd[(p * bpad + t) * ldq + k] = digit p of q_t[k] / u_t, qs[t] = u_t.
__global__ void split_q_kernel(const double* __restrict__ qb, int64_t ldqb, int64_t m, int64_t b, int64_t ldq,
                               int64_t bpad, int8_t* __restrict__ qd, double* __restrict__ qs)
{
    __shared__ double red[256];
    const int64_t t = blockIdx.x;
    double mx = 0;
    if (t < b)
        for (int64_t k = threadIdx.x; k < m; k += blockDim.x) mx = fmax(mx, fabs(qb[k + t * ldqb]));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    mx = red[0];
    int e = 0;
    if (mx > 0) frexp(mx, &e);  // mx < 2^e
    const double inv = mx > 0 ? ldexp(1.0, -(e + 1)) : 0.0;
    if (threadIdx.x == 0) qs[t] = mx > 0 ? ldexp(1.0, e + 1) : 0.0;
    for (int64_t k = threadIdx.x; k < ldq; k += blockDim.x) {
        double x = (t < b && k < m) ? qb[k + t * ldqb] * inv : 0.0;  // |x| < 1/2, exact scaling
#pragma unroll
        for (int p = 0; p < ND; ++p) {
            x *= 128.0;
            const double dgt = rint(x);
            qd[((int64_t)p * bpad + t) * ldq + k] = (int8_t)dgt;
            x -= dgt;  // exact: |x - dgt| <= 1/2
        }
    }
}

// In place: column j of y (FP32, m entries, column pitch 4 P bytes with P a
// multiple of 16 >= m) becomes four digit planes of P bytes (plane r at byte
// offset r P, zero past m); scale[j] = s_j and ynorm2[j] = ||y_hat_j||^2.
// One workgroup per column; the column is first read into LDS.
__global__ __launch_bounds__(256) void digitize_columns_kernel(float* __restrict__ y, int64_t m, int64_t P,
                                                               double* __restrict__ ynorm2,
                                                               double* __restrict__ scale)
{
    extern __shared__ float col[];
    __shared__ double part[4];
    const int64_t j = blockIdx.x;
    float* src = y + j * P;  // 4 P bytes = P floats
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x) col[i] = src[i];
    const double nrm = sqrt(ynorm2[j]);  // of the FP32 entries, >= max |y_ij|
    int e = 0;
    if (nrm > 0) frexp(nrm, &e);
    const double s = nrm > 0 ? ldexp(1.0, e + 1) : 0.0;
    const float inv = nrm > 0 ? ldexpf(1.0f, -(e + 1)) : 0.f;
    __syncthreads();
    int8_t* planes = reinterpret_cast<int8_t*>(src);
    double acc = 0;
    for (int64_t i = threadIdx.x; i < P; i += blockDim.x) {
        float x = i < m ? col[i] * inv : 0.f;  // exact: power-of-two scaling, |x| < 1/2
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            x *= 128.0f;
            const float dgt = rintf(x);
            x -= dgt;  // exact (Sterbenz)
            planes[r * P + i] = (int8_t)dgt;
        }
    }
    __syncthreads();
    // ||y_hat||^2 from the stored digits, in FP64
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x) {
        double v = 0;
#pragma unroll
        for (int r = NY - 1; r >= 0; --r) v = (v + (double)planes[r * P + i]) * (1.0 / 128.0);
        v *= s;
        acc += v * v;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        ynorm2[j] = part[0] + part[1] + part[2] + part[3];
        scale[j] = s;
    }
}

template <bool ENERGY>
__global__ __launch_bounds__(256) void fold_i8_kernel(const int8_t* __restrict__ qd, const double* __restrict__ qs,
                                                      int64_t ldq, int64_t bpad, int64_t b,
                                                      const int8_t* __restrict__ yd, int64_t P,
                                                      const double* __restrict__ yscale, int64_t m, int64_t n,
                                                      double* __restrict__ proj2, double* __restrict__ energy)
{
    __shared__ __align__(16) int8_t sQ[2][ND][FBM * FPITCH];
    __shared__ double red[4][FBM];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int quarter = lane >> 4, l16 = lane & 15;
    const int64_t m0 = (int64_t)blockIdx.y * FBM;
    const int64_t col = (int64_t)blockIdx.x * FBN + wave * 16 + l16;  // this lane's Y column
    const int R = (int)std::min<int64_t>(8, (b - m0 + 15) / 16);

    i4v acc[NG][8];
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int t = 0; t < 8; ++t) acc[g][t] = i4v{0, 0, 0, 0};

    const bool cok = col < n;
    const double sj = cok ? yscale[col] : 0.0;
    const int8_t* ycol = yd + (cok ? col : 0) * 4 * P;
    // This lane's 16 K bytes of plane r at slice k0 (zero past the plane).
    i4v bd[NY], bn[NY];
#define MPISDF_LOAD_YD(k0, dst)                                                                    \
    do {                                                                                          \
        const int64_t k_ = (k0) + 16 * quarter;                                                   \
        const bool in_ = cok && k_ < P;                                                           \
        _Pragma("unroll") for (int r_ = 0; r_ < NY; ++r_) {                                       \
            const i4v v_ = *reinterpret_cast<const i4v*>(ycol + r_ * P + (in_ ? k_ : 0));          \
            dst[r_] = in_ ? v_ : i4v{0, 0, 0, 0};                                                 \
        }                                                                                         \
    } while (0)
    // Q digits: ND planes x 128 rows x 64 bytes per slice = 2560 chunks of 16
    // bytes, ten per thread.
    constexpr int QCHUNKS = ND * FBM * (FBK / 16) / 256;
    uint4 rq[QCHUNKS];
#define MPISDF_LOAD_Q(k0)                                                                          \
    do {                                                                                          \
        _Pragma("unroll") for (int r_ = 0; r_ < QCHUNKS; ++r_) {                                  \
            const int c_ = tid + 256 * r_;                                                        \
            const int p_ = c_ / (FBM * 4), rem_ = c_ - p_ * FBM * 4, row_ = rem_ >> 2, kc_ = (rem_ & 3) * 16; \
            rq[r_] = *reinterpret_cast<const uint4*>(qd + ((int64_t)p_ * bpad + m0 + row_) * ldq + (k0) + kc_); \
        }                                                                                         \
    } while (0)
#define MPISDF_STASH_Q(buf)                                                                        \
    do {                                                                                          \
        _Pragma("unroll") for (int r_ = 0; r_ < QCHUNKS; ++r_) {                                  \
            const int c_ = tid + 256 * r_;                                                        \
            const int p_ = c_ / (FBM * 4), rem_ = c_ - p_ * FBM * 4, row_ = rem_ >> 2, kc_ = (rem_ & 3) * 16; \
            *reinterpret_cast<uint4*>(&sQ[buf][p_][row_ * FPITCH + kc_]) = rq[r_];                \
        }                                                                                         \
    } while (0)

    const int64_t nsl = (m + FBK - 1) / FBK;
    MPISDF_LOAD_Q(0);
    MPISDF_LOAD_YD(0, bd);
    MPISDF_STASH_Q(0);
    __syncthreads();
    for (int64_t sl = 0; sl < nsl; ++sl) {
        const int buf = (int)(sl & 1);
        const bool more = sl + 1 < nsl;
        if (more) {
            MPISDF_LOAD_Q((sl + 1) * FBK);
            MPISDF_LOAD_YD((sl + 1) * FBK, bn);
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            if (t >= R) continue;
#pragma unroll
            for (int p = 0; p < ND; ++p) {
                const i4v a = *reinterpret_cast<const i4v*>(&sQ[buf][p][(t * 16 + l16) * FPITCH + 16 * quarter]);
#pragma unroll
                for (int r = 0; r + p < NG && r < NY; ++r)
                    acc[p + r][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, bd[r], acc[p + r][t], 0, 0, 0);
            }
        }
        if (more) {
            MPISDF_STASH_Q(buf ^ 1);
#pragma unroll
            for (int r = 0; r < NY; ++r) bd[r] = bn[r];
        }
        __syncthreads();
    }
#undef MPISDF_LOAD_YD
#undef MPISDF_LOAD_Q
#undef MPISDF_STASH_Q

    // W[t][j] = u_t s_j sum_g 128^-(g+2) acc_g; element e of lane l is row 4 quarter + e.
    double colsq = 0;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            double v = 0;
#pragma unroll
            for (int g = NG - 1; g >= 0; --g) v = v * (1.0 / 128.0) + (double)acc[g][t][e];
            const int64_t row = m0 + t * 16 + 4 * quarter + e;
            const double wv = (t < R && row < b) ? v * (1.0 / 16384.0) * qs[row] * sj : 0.0;
            colsq += wv * wv;
            if (ENERGY) {
                double s = wv * wv;
#pragma unroll
                for (int off = 1; off < 16; off <<= 1) s += __shfl_xor(s, off);
                if (l16 == 0) red[wave][t * 16 + 4 * quarter + e] = s;
            }
        }
    }
    colsq += __shfl_xor(colsq, 16);
    colsq += __shfl_xor(colsq, 32);
    if (quarter == 0 && cok) {
        if (gridDim.y == 1) proj2[col] += colsq;
        else atomicAdd(&proj2[col], colsq);
    }
    if (ENERGY) {
        __syncthreads();
        if (tid < FBM) {
            const int64_t row = m0 + tid;
            if (row < b) atomicAdd(&energy[row], red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid]);
        }
    }
}

}  // namespace

void digitize_columns(float* y, int64_t m, int64_t n, int64_t P, double* ynorm2, double* scale, hipStream_t stream)
{
    if (n <= 0) return;
    if (P % 16 || P < m) throw std::invalid_argument("digit planes need a pitch that is a multiple of 16 and >= m");
    const size_t lds = (size_t)m * sizeof(float);
    if (lds > 144 * 1024) throw std::invalid_argument("digitize_columns holds one column in LDS (m <= 36864)");
    if (lds > 64 * 1024)
        MPISDF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(digitize_columns_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (n > 0x7fffffff) throw std::overflow_error("too many columns to digitise in one launch");
    hipLaunchKernelGGL(digitize_columns_kernel, dim3((unsigned)n), dim3(256), lds, stream, y, m, P, ynorm2, scale);
    MPISDF_HIP(hipGetLastError());
}

void fold_projections_i8(int64_t b, int64_t n, int64_t m, const double* Qb, int64_t ldqb, const int8_t* Yd,
                         int64_t P, const double* yscale, double* proj2, double* energy, FoldI8Scratch& scratch,
                         hipStream_t stream)
{
    if (b <= 0 || n <= 0) return;
    if (m >= (int64_t(1) << 16)) throw std::invalid_argument("the INT8 fold needs fewer than 65536 sketch rows");
    const int64_t bpad = (b + FBM - 1) / FBM * FBM;
    const int64_t ldq = (m + FBK - 1) / FBK * FBK;
    const size_t need = (size_t)(ND * bpad * ldq);
    if (scratch.digits.count < need) scratch.digits.resize(need);
    if (scratch.scales.count < (size_t)bpad) scratch.scales.resize(bpad);
    hipLaunchKernelGGL(split_q_kernel, dim3((unsigned)bpad), dim3(256), 0, stream, Qb, ldqb, m, b, ldq, bpad,
                       scratch.digits.ptr, scratch.scales.ptr);
    MPISDF_HIP(hipGetLastError());
    const int64_t gx = (n + FBN - 1) / FBN, gy = bpad / FBM;
    if (gx > 0x7fffffff || gy > 65535) throw std::overflow_error("fold grid too large");
    if (energy)
        hipLaunchKernelGGL(fold_i8_kernel<true>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, stream,
                           scratch.digits.ptr, scratch.scales.ptr, ldq, bpad, b, Yd, P, yscale, m, n, proj2, energy);
    else
        hipLaunchKernelGGL(fold_i8_kernel<false>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, stream,
                           scratch.digits.ptr, scratch.scales.ptr, ldq, bpad, b, Yd, P, yscale, m, n, proj2,
                           static_cast<double*>(nullptr));
    MPISDF_HIP(hipGetLastError());
}

}  // namespace mpisdf
