// TN GEMM on FP64 MFMA with both operands K-contiguous:
//     W[t, j] = sum_k A[k + t * lda] * B[k + j * ldb]      (t < M, j < N, k < K)
// A is FP64 (the QRCP's orthonormal columns), B is FP32 or FP64 (sketch
// columns, candidate residuals).  Used for the QRCP fold (W consumed by a
// sum-of-squares epilogue, never stored) and for the projections H = Q^T R
// of the candidate block.
//
// v_mfma_f64_16x16x4_f64 takes A[l % 16][k(l)] and B[k(l)][l % 16] from lane
// l, where the four K values of one step are spread over the lane quarters.
// The reduction over K does not care which K index a quarter supplies, so
// quarter g feeds K = 2 g + s (s = 0, 1) and 8 + 2 g + s - 2 (s = 2, 3) of a
// 16-wide K chunk: every lane then consumes two pairs of consecutive K values
// of one row / column, read as two 16-byte LDS vectors (A) and two 8- or
// 16-byte global vectors (B).  Result element i of lane l is
// D[l / 16 + 4 i][l % 16].
//
// Workgroup: 4 waves, 128 columns (32 per wave = 2 column tiles) x up to 128
// rows (8 row tiles, grid.y walks further row groups).  The A slice of 128 rows
// x 32 K is staged in double-buffered LDS and shared by the four waves; B goes
// straight from HBM to registers, one slice ahead.
#include "kernels.hpp"
#include "util.hpp"

#include <algorithm>

namespace mpisdf {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int TBM = 128;       // rows per workgroup (row group)
constexpr int TBN = 128;       // columns per workgroup
constexpr int TBK = 32;        // K per slice
// LDS row pitch (doubles): 288 bytes.  Each lane reads 16 bytes at row * 288 +
// 16 q (+ 64): on gfx950 a pitch of 32 mod 64 bytes keeps these ds_read_b128
// free of bank conflicts (measured with SQ_LDS_BANK_CONFLICT; 272-byte rows
// and a 32-byte lane stride conflict two ways).
constexpr int LDA_S = TBK + 4;

enum TnEpilogue { kTnStore = 0, kTnFold = 1, kTnFoldEnergy = 2, kTnKhatriRao = 3,
                  kTnStoreSquare = 4, kTnStoreColSq = 5, kTnStoreRowDot = 6, kTnFoldPart = 7 };

template <class TB>
struct BVec2;
template <>
struct BVec2<float> {
    typedef float2 type;
};
template <>
struct BVec2<double> {
    typedef double2 type;
};

// out[0..1] = p[ka], p[ka + 1]; out[2..3] = p[kb], p[kb + 1] (zero past K or
// for an absent column).
template <class TB, bool VEC>
__device__ __forceinline__ void load2x2(const TB* __restrict__ p, int64_t ka, int64_t kb, int64_t K, bool col_ok,
                                        TB out[4])
{
    const int64_t ks[2] = {ka, kb};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int64_t k = ks[h];
        if (VEC && col_ok && k + 2 <= K) {
            const typename BVec2<TB>::type v = *reinterpret_cast<const typename BVec2<TB>::type*>(p + k);
            out[2 * h] = v.x;
            out[2 * h + 1] = v.y;
        } else {
            out[2 * h] = (col_ok && k < K) ? p[k] : TB(0);
            out[2 * h + 1] = (col_ok && k + 1 < K) ? p[k + 1] : TB(0);
        }
    }
}

template <class TB>
struct BVec;
template <>
struct BVec<float> {
    typedef float4 type;
};
template <>
struct BVec<double> {
    typedef double4 type;
};

template <class TB, bool VEC, class TO>
__device__ __forceinline__ void load4(const TB* __restrict__ p, int64_t k, int64_t K, bool col_ok, TO out[4])
{
    if (VEC) {
        if (col_ok && k + 4 <= K) {
            const typename BVec<TB>::type v = *reinterpret_cast<const typename BVec<TB>::type*>(p + k);
            out[0] = v.x;
            out[1] = v.y;
            out[2] = v.z;
            out[3] = v.w;
            return;
        }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) out[s] = (col_ok && k + s < K) ? p[k + s] : TB(0);
}

template <class TB, int EPI, bool VEC>
__global__ __launch_bounds__(256, 2) void tn_gemm_kernel(int64_t M, int64_t N, int64_t K,
                                                          const double* __restrict__ A, int64_t lda,
                                                          const TB* __restrict__ B, int64_t ldb, double alpha,
                                                          double beta, double* __restrict__ C, int64_t ldc,
                                                          double* __restrict__ colacc, double* __restrict__ rowacc,
                                                          const double* __restrict__ A2,
                                                          const double* __restrict__ cscale,
                                                          unsigned long long* __restrict__ bad, int tri,
                                                          int64_t kchunk, int64_t czstride)
{
    __shared__ __align__(16) double sA[2][TBM * LDA_S];
    __shared__ double red[4][TBM];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int quarter = lane >> 4, l16 = lane & 15;
    const int64_t n0 = (int64_t)blockIdx.x * TBN + wave * 32;
    // Khatri-Rao mode: the 128 A rows are 64 rows of A (U^T) stacked on the
    // same 64 rows of A2 (V^T); row tiles t and t + 4 then hold U phi and V phi
    // for identical (s, g) positions in every lane.
    constexpr bool KR = EPI == kTnKhatriRao;
    const int64_t m0 = (int64_t)blockIdx.y * (KR ? TBM / 2 : TBM);
    const int R = KR ? 8 : (int)std::min<int64_t>(8, (M - m0 + 15) / 16);  // row tiles in this group (uniform)

    d4 acc[8][2];
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[t][c] = d4{0.0, 0.0, 0.0, 0.0};

    // A staging: 128 rows x 32 K = 1024 four-wide chunks, four per thread.
    double qa[4][4];
    auto load_a = [&](int64_t k0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = tid + 256 * r;
            const int row = c >> 3, kc = (c & 7) * 4;
            const int64_t gm = m0 + (KR ? (row & 63) : row);
            const double* base = (KR && row >= 64) ? A2 : A;
            load4<double, VEC, double>(base + gm * lda, k0 + kc, K, gm < M, qa[r]);
        }
    };
    auto stash_a = [&](int buf) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = tid + 256 * r;
            const int row = c >> 3, kc = (c & 7) * 4;
            double* dst = &sA[buf][row * LDA_S + kc];
            *reinterpret_cast<double2*>(dst) = make_double2(qa[r][0], qa[r][1]);
            *reinterpret_cast<double2*>(dst + 2) = make_double2(qa[r][2], qa[r][3]);
        }
    };
    // B: per column tile and 16-K half, the four K of this lane's quarter q:
    // 2q, 2q + 1, 8 + 2q, 9 + 2q (the order the A reads from LDS supply).
    TB yb[2][2][4], yn[2][2][4];  // raw B values, widened at use
    const bool ok0 = n0 + l16 < N, ok1 = n0 + 16 + l16 < N;
    const TB* b0 = B + (ok0 ? (n0 + l16) : 0) * ldb;
    const TB* b1 = B + (ok1 ? (n0 + 16 + l16) : 0) * ldb;
    auto load_b = [&](int64_t k0, TB dst[2][2][4]) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int64_t k = k0 + 16 * h + 2 * quarter;
            load2x2<TB, VEC>(b0, k, k + 8, K, ok0, dst[0][h]);
            load2x2<TB, VEC>(b1, k, k + 8, K, ok1, dst[1][h]);
        }
    };

    // Triangular A (tri = 1 lower, 2 upper): this row group's nonzero columns
    // are k < m0 + 128 (lower) or k >= m0 (upper); the rest of K is skipped.
    // Split K (kchunk > 0, a multiple of TBK): workgroup z takes K chunk z and
    // stores its partial product at C + z czstride.
    const int64_t kz0 = kchunk > 0 ? (int64_t)blockIdx.z * kchunk : 0;
    const int64_t kz1 = kchunk > 0 ? kz0 + kchunk : K;
    if (kchunk > 0) C += (int64_t)blockIdx.z * czstride;
    const int64_t k_lo = std::max<int64_t>(tri == 2 ? m0 : 0, kz0);
    const int64_t k_hi = std::min<int64_t>(tri == 1 ? std::min<int64_t>(K, m0 + TBM) : K, kz1);
    const int64_t nslices = k_hi > k_lo ? (k_hi - k_lo + TBK - 1) / TBK : 0;
    if (nslices > 0) {
        load_a(k_lo);
        load_b(k_lo, yb);
        stash_a(0);
    }
    __syncthreads();
    for (int64_t sl = 0; sl < nslices; ++sl) {
        const int buf = (int)(sl & 1);
        const bool more = sl + 1 < nslices;
        if (more) {
            load_a(k_lo + (sl + 1) * TBK);
            load_b(k_lo + (sl + 1) * TBK, yn);
        }
        const double* As = sA[buf];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                if (t < R) {
                    const double* ap = As + (t * 16 + l16) * LDA_S + 16 * h + 2 * quarter;
                    const double2 a01 = *reinterpret_cast<const double2*>(ap);
                    const double2 a23 = *reinterpret_cast<const double2*>(ap + 8);
                    const double av[4] = {a01.x, a01.y, a23.x, a23.y};
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        acc[t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], (double)yb[0][h][s], acc[t][0], 0, 0, 0);
                        acc[t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], (double)yb[1][h][s], acc[t][1], 0, 0, 0);
                    }
                }
            }
        }
        if (more) {
            stash_a(buf ^ 1);
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int s = 0; s < 4; ++s) yb[c][h][s] = yn[c][h][s];
        }
        __syncthreads();
    }

    if (KR) {
        // y[g * ldc + s] = cscale[g] (U phi)[s, g] (V phi)[s, g]; colacc[r * N + g] = sum of y^2
        // over the rows s of row group r (one writer per entry; added in row-group order later).
        unsigned long long nonfinite = 0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int64_t gn = n0 + 16 * c + l16;
            const bool gin = gn < N;
            const double sc = gin ? cscale[gn] : 0.0;
            double s2 = 0;
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int64_t gm = m0 + t * 16 + quarter + 4 * i;
                    const double v = sc * acc[t][c][i] * acc[t + 4][c][i];
                    if (gin && gm < M) {
                        C[gm + gn * ldc] = v;
                        s2 += v * v;
                        nonfinite += !isfinite(v);
                    }
                }
            s2 += __shfl_xor(s2, 16);
            s2 += __shfl_xor(s2, 32);
            if (quarter == 0 && gin && colacc) colacc[(int64_t)blockIdx.y * N + gn] = s2;
        }
        if (nonfinite) atomicAdd(bad, nonfinite);
        return;
    }
    if (EPI == kTnStore || EPI == kTnStoreSquare || EPI == kTnStoreColSq || EPI == kTnStoreRowDot) {
        // Store C = alpha acc + beta C (kTnStoreSquare: C = acc^2).
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            if (t >= R) continue;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int64_t gn = n0 + 16 * c + l16;
                if (gn >= N) continue;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int64_t gm = m0 + t * 16 + quarter + 4 * i;
                    if (gm < M) {
                        double* p = C + gm + gn * ldc;
                        if (EPI == kTnStoreSquare) *p = acc[t][c][i] * acc[t][c][i];
                        else *p = alpha * acc[t][c][i] + (beta == 0.0 ? 0.0 : beta * *p);
                    }
                }
            }
        }
        if (EPI == kTnStoreColSq) {
            // colacc[g * N + n] = sum over row group g of (alpha acc)^2 (rows >=
            // M hold zeros); a partial per row group keeps the total independent
            // of the order the workgroups finish in.
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                double s = 0;
#pragma unroll
                for (int t = 0; t < 8; ++t)
#pragma unroll
                    for (int i = 0; i < 4; ++i) s += acc[t][c][i] * acc[t][c][i];
                s += __shfl_xor(s, 16);
                s += __shfl_xor(s, 32);
                const int64_t gn = n0 + 16 * c + l16;
                if (quarter == 0 && gn < N) colacc[(int64_t)blockIdx.y * N + gn] = alpha * alpha * s;
            }
        }
        if (EPI == kTnStoreRowDot) {
            // rowacc[x * M + m] = sum over the columns of column block x of
            // alpha acc * cscale[n] (columns >= N hold zeros and weigh zero).
            double wc[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int64_t gn = n0 + 16 * c + l16;
                wc[c] = gn < N ? alpha * cscale[gn] : 0.0;
            }
#pragma unroll
            for (int t = 0; t < 8; ++t)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    double s = acc[t][0][i] * wc[0] + acc[t][1][i] * wc[1];
#pragma unroll
                    for (int off = 1; off < 16; off <<= 1) s += __shfl_xor(s, off);
                    if (l16 == 0) red[wave][t * 16 + quarter + 4 * i] = s;
                }
            __syncthreads();
            if (tid < TBM) {
                const int64_t gm = m0 + tid;
                const double s = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
                if (gm < M && tid < R * 16) rowacc[(int64_t)blockIdx.x * M + gm] = s;
            }
        }
        return;
    }
    // Column sums of W^2 over the (single) row group: each column lives in one
    // wave, and launches of successive row groups add in order.
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        double s = 0;
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) s += acc[t][c][i] * acc[t][c][i];  // rows >= M hold zeros
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        const int64_t gn = n0 + 16 * c + l16;
        if (EPI == kTnFoldPart) {
            // one partial per row group (launches of many groups), added in order later
            if (quarter == 0 && gn < N) colacc[(int64_t)blockIdx.y * N + gn] = s;
        } else if (quarter == 0 && gn < N) {
            colacc[gn] += s;
        }
    }
    if (EPI == kTnFoldEnergy) {
        // Row sums of W^2 over the workgroup's columns: lanes of one quarter
        // share rows; reduce over l16 and the two column tiles, then the waves.
        // rowacc[x * M + m] is column block x's partial.
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                double s = acc[t][0][i] * acc[t][0][i] + acc[t][1][i] * acc[t][1][i];
#pragma unroll
                for (int off = 1; off < 16; off <<= 1) s += __shfl_xor(s, off);
                if (l16 == 0) red[wave][t * 16 + quarter + 4 * i] = s;
            }
        __syncthreads();
        if (tid < TBM) {
            const int64_t gm = m0 + tid;
            const double s = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
            if (gm < M && tid < R * 16) rowacc[(int64_t)blockIdx.x * M + gm] = s;
        }
    }
}

// C[i + j ldc] = (accumulate ? C : 0) + sum_z part[z zstride + i + j ldc] for
// i < M, j < N, z ascending (the rows between M and ldc are left alone).
__global__ void sum_split_partials_kernel(const double* __restrict__ part, int64_t splits, int64_t zstride,
                                          int64_t M, int64_t N, double* __restrict__ C, int64_t ldc, int accumulate)
{
    const int64_t total = M * N;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = idx / M, off = idx - j * M + j * ldc;
        double v = accumulate ? C[off] : 0.0;
        for (int64_t z = 0; z < splits; ++z) v += part[z * zstride + off];
        C[off] = v;
    }
}

template <class TB, int EPI>
void launch_tn(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const TB* B, int64_t ldb, double alpha,
               double beta, double* C, int64_t ldc, double* colacc, double* rowacc, hipStream_t stream,
               const double* A2 = nullptr, const double* cscale = nullptr, unsigned long long* bad = nullptr,
               int tri = 0, int64_t splits = 1, int64_t kchunk = 0, int64_t czstride = 0)
{
    if (M <= 0 || N <= 0) return;
    const int64_t rows = EPI == kTnKhatriRao ? TBM / 2 : TBM;
    const int64_t gx = (N + TBN - 1) / TBN, gy = (M + rows - 1) / rows;
    if (gx > 0x7fffffff || gy > 65535) throw std::overflow_error("TN GEMM grid too large");
    if ((EPI == kTnFold || EPI == kTnFoldEnergy) && gy != 1)
        throw std::logic_error("a fold launch covers one row group of at most 128 rows");
    const bool vec = lda % 4 == 0 && ldb % 4 == 0 && (reinterpret_cast<uintptr_t>(A) % 32) == 0 &&
                     (reinterpret_cast<uintptr_t>(A2) % 32) == 0 &&
                     (reinterpret_cast<uintptr_t>(B) % (4 * sizeof(TB))) == 0;
    const dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)splits);
    if (vec)
        hipLaunchKernelGGL((tn_gemm_kernel<TB, EPI, true>), grid, dim3(256), 0, stream, M, N, std::max<int64_t>(K, 0),
                           A, lda, B, ldb, alpha, beta, C, ldc, colacc, rowacc, A2, cscale, bad, tri, kchunk, czstride);
    else
        hipLaunchKernelGGL((tn_gemm_kernel<TB, EPI, false>), grid, dim3(256), 0, stream, M, N, std::max<int64_t>(K, 0),
                           A, lda, B, ldb, alpha, beta, C, ldc, colacc, rowacc, A2, cscale, bad, tri, kchunk, czstride);
    MPISDF_HIP(hipGetLastError());
}

}  // namespace

void gemm_tn_f64(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda, const double* B,
                 int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t stream, int triangular)
{
    launch_tn<double, kTnStore>(M, N, K, A, lda, B, ldb, alpha, beta, C, ldc, nullptr, nullptr, stream, nullptr,
                                nullptr, nullptr, triangular);
}

void gemm_tn_f64_fused(TnFused mode, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                       const double* B, int64_t ldb, double* C, int64_t ldc, double* colacc, double* rowacc,
                       const double* colw, hipStream_t stream, int triangular)
{
    switch (mode) {
    case TnFused::Square:
        launch_tn<double, kTnStoreSquare>(M, N, K, A, lda, B, ldb, 1.0, 0.0, C, ldc, nullptr, nullptr, stream, nullptr,
                                          nullptr, nullptr, triangular);
        break;
    case TnFused::ColumnSumSquares:
        if (!colacc) throw std::invalid_argument("column sums need an accumulator");
        launch_tn<double, kTnStoreColSq>(M, N, K, A, lda, B, ldb, alpha, 0.0, C, ldc, colacc, nullptr, stream, nullptr,
                                         nullptr, nullptr, triangular);
        break;
    case TnFused::RowDot:
        if (!rowacc || !colw) throw std::invalid_argument("row sums need an accumulator and column weights");
        launch_tn<double, kTnStoreRowDot>(M, N, K, A, lda, B, ldb, alpha, 0.0, C, ldc, nullptr, rowacc, stream, nullptr,
                                          colw, nullptr, triangular);
        break;
    }
}

int64_t tn_column_blocks(int64_t N) { return (N + TBN - 1) / TBN; }

void gemm_tn_f64_split(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda, const double* B,
                       int64_t ldb, double beta, double* C, int64_t ldc, DeviceBuffer<double>& part,
                       hipStream_t stream)
{
    if (M <= 0 || N <= 0) return;
    if (beta != 0.0 && beta != 1.0) throw std::invalid_argument("split-K products take beta 0 or 1");
    // Workgroups of the plain product (two fit on a CU); split K until the
    // grid covers about two rounds of the chip, in chunks of whole slices.
    int dev = 0, cus = 0;
    MPISDF_HIP(hipGetDevice(&dev));
    MPISDF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t tiles = ((N + TBN - 1) / TBN) * ((M + TBM - 1) / TBM);
    const int64_t slots = 2 * std::max(cus, 1);
    int64_t splits = std::min<int64_t>(16, std::max<int64_t>(1, 2 * slots / std::max<int64_t>(tiles, 1)));
    const int64_t kchunk = ((K + splits - 1) / splits + TBK - 1) / TBK * TBK;
    splits = kchunk > 0 ? (K + kchunk - 1) / kchunk : 1;
    if (splits <= 1) {
        gemm_tn_f64(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, stream);
        return;
    }
    // partial z: alpha A^T B over K chunk z (M x N, leading dimension ldc);
    // C = beta C + the partials in chunk order
    const int64_t zstride = ldc * N;
    if (part.count < (size_t)(splits * zstride)) part.resize((size_t)(splits * zstride));
    launch_tn<double, kTnStore>(M, N, K, A, lda, B, ldb, alpha, 0.0, part.ptr, ldc, nullptr, nullptr, stream, nullptr,
                                nullptr, nullptr, 0, splits, kchunk, zstride);
    const int64_t total = M * N;
    hipLaunchKernelGGL(sum_split_partials_kernel, dim3((unsigned)std::min<int64_t>((total + 255) / 256, 65536)),
                       dim3(256), 0, stream, part.ptr, splits, zstride, M, N, C, ldc, beta == 1.0 ? 1 : 0);
    MPISDF_HIP(hipGetLastError());
}

void fold_projections(int64_t b, int64_t n, int64_t m, const double* Qb, int64_t ldq, const void* Y, bool y_f32,
                      int64_t ldy, double* proj2, double* energy, DeviceBuffer<double>& part, hipStream_t stream)
{
    // W (b x n) = Qb^T Y; proj2[j] += sum_t W[t,j]^2; energy[t] += sum_j W[t,j]^2.
    // One launch per group of 128 rows, so every column's sum has one writer
    // per launch and the groups add in order; the energies are written as
    // per-column-block partials and added in block order.
    if (b <= 0 || n <= 0) return;
    if (!energy && b > TBM) {
        // Many row groups (a fold of every pivot so far): all groups in one
        // launch, each writing its column partial, added in group order; the
        // columns in chunks that bound the partials.
        const int64_t groups = (b + TBM - 1) / TBM;
        if (groups > 65535) throw std::overflow_error("too many pivot groups for one fold launch");
        const int64_t chunk = std::max<int64_t>(TBN, (int64_t(1) << 25) / groups / TBN * TBN);
        if (part.count < (size_t)(groups * std::min(chunk, n))) part.resize((size_t)(groups * std::min(chunk, n)));
        for (int64_t c0 = 0; c0 < n; c0 += chunk) {
            const int64_t nc = std::min(chunk, n - c0);
            if (y_f32)
                launch_tn<float, kTnFoldPart>(b, nc, m, Qb, ldq, static_cast<const float*>(Y) + c0 * ldy, ldy, 1.0, 0.0,
                                              nullptr, 0, part.ptr, nullptr, stream);
            else
                launch_tn<double, kTnFoldPart>(b, nc, m, Qb, ldq, static_cast<const double*>(Y) + c0 * ldy, ldy, 1.0,
                                               0.0, nullptr, 0, part.ptr, nullptr, stream);
            launch_sum_partials(part.ptr, groups, nc, proj2 + c0, true, stream);
        }
        return;
    }
    const int64_t gx = tn_column_blocks(n);
    if (energy && part.count < (size_t)(gx * TBM)) part.resize((size_t)(gx * TBM));
    for (int64_t r0 = 0; r0 < b; r0 += TBM) {
        const int64_t rows = std::min<int64_t>(TBM, b - r0);
        const double* q = Qb + r0 * ldq;
        if (y_f32) {
            const float* y = static_cast<const float*>(Y);
            if (energy) launch_tn<float, kTnFoldEnergy>(rows, n, m, q, ldq, y, ldy, 1.0, 0.0, nullptr, 0, proj2, part.ptr, stream);
            else launch_tn<float, kTnFold>(rows, n, m, q, ldq, y, ldy, 1.0, 0.0, nullptr, 0, proj2, nullptr, stream);
        } else {
            const double* y = static_cast<const double*>(Y);
            if (energy) launch_tn<double, kTnFoldEnergy>(rows, n, m, q, ldq, y, ldy, 1.0, 0.0, nullptr, 0, proj2, part.ptr, stream);
            else launch_tn<double, kTnFold>(rows, n, m, q, ldq, y, ldy, 1.0, 0.0, nullptr, 0, proj2, nullptr, stream);
        }
        if (energy) launch_sum_partials(part.ptr, gx, rows, energy + r0, true, stream);
    }
}

void kr_sketch_f64_tn(const double* ut, const double* vt, int64_t ell, int64_t d, const double* phi, int64_t ldphi,
                      int64_t count, const double* scale, double* y, int64_t ldy, double* ynorm2,
                      DeviceBuffer<double>& part, unsigned long long* bad, hipStream_t stream)
{
    // ||y_g||^2 per 64-row group of the launch, then added in group order.
    const int64_t groups = (ell + TBM / 2 - 1) / (TBM / 2);
    if (ynorm2 && part.count < (size_t)(groups * count)) part.resize((size_t)(groups * count));
    launch_tn<double, kTnKhatriRao>(ell, count, d, ut, d, phi, ldphi, 1.0, 0.0, y, ldy, ynorm2 ? part.ptr : nullptr,
                                    nullptr, stream, vt, scale, bad);
    if (ynorm2) launch_sum_partials(part.ptr, groups, count, ynorm2, true, stream);
}

}  // namespace mpisdf
