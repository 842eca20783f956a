// FP64-accurate products on INT8 MFMA with digit splitting.  The QRCP fold
// on sketches stored as signed digit planes,
//
//     W = Qb^T Y,  proj2[j] += sum_t W[t,j]^2  (energy[t] += sum_j W[t,j]^2)
//
// and, through the same kernel with other epilogues (below), the Khatri-Rao
// sketch of FP64 point features and the fit's S = (Phi_a Phi^T)^2 and
// Z = L^-1 S products, all in base 254.
//
// Y is held as NY = 3 or 4 digit planes per column (digitize_columns, in place
// over the FP32 sketch, same footprint): with the power of two s_j >= 2 ||y_j||,
//     y_hat[i] = s_j sum_{r<NY} d_r[i] 128^-(r+1),  |d_r| <= 64,
// which rounds each entry to within 2^-(7 NY + 1) s_j: with three planes 3e-5
// to 6e-5 of ||y_j|| in RMS over 13,200 rows (s_j < 4 ||y_j||; well below the
// rounding noise of a BF16 or FP8 sketch), with four ~2e-7 (below that of a
// split-BF16 sketch).  An FP64 sketch is held as four base-254 planes
// (digitize_columns_f64, |d_r| <= 127, within s_j 254^-4 / 2 ~ 1.2e-10 s_j:
// the FP64 sketch's pivots), in place over its FP64 entries.
// The QRCP then factorises y_hat: candidate columns are rebuilt from the same
// digits, and ||y_hat_j||^2 replaces ||y_j||^2.
//
// Products use the splitting method of Ozaki, Ogita, Oishi and Rump (2012) in
// the INT8 form of Ootomo, Ozaki and Yokota (2024): row t of Qb^T is written
// as five digits with its own power of two u_t >= 2 max|q_t|; the digit
// pairs with p + r <= 5 (14 with three Y planes, 17 with four) run on
// v_mfma_i32_32x32x32_i8, pairs of equal p + r sharing an exact INT32
// accumulator (at most four pairs: |sum| < 4 * 64^2 * m < 2^31 for m < 2^16),
// and the six accumulators are combined in FP64:
//     W = u_t s_j sum_g 128^-(g+2) acc_g.
// In base 254 the pairs p + r <= 4 run in five groups: 13 for the QRCP fold
// (q in four digits, four Y planes: at most four pairs per group,
// |sum| < 4 * 127^2 * m < 2^31 for m < 2^15), 15 with five digits and five
// planes (at most five: m <= 26624).  The Khatri-Rao sketch takes four digits
// on both sides and the ten pairs p + r <= 3 in four groups (the NG template
// argument).
// The dropped pairs (p + r >= 6 in base 128, >= 5 in base 254) and the
// truncation of q leave ~1e-11 of ||y_j||^2 in the folded sums (measured),
// covered by the QRCP's relative margin.
//
// v_mfma_i32_32x32x32_i8 runs at about 1.7x the rate of the 16x16x64 form on
// gfx950.  Q and Y digit slices are staged through double-buffered LDS, with
// loads issued two slices ahead; the lane layout and workgroup shapes are
// described at the kernel.
#include "kernels.hpp"
#include "util.hpp"

#include <algorithm>

namespace mpisdf {
namespace {

typedef int i4v __attribute__((ext_vector_type(4)));

// Digit schemes.  Base 128 (|d| <= 64): q in five digits, Y in three or four
// planes, the pairs p + r <= 5 in six accumulator groups.  Base 254
// (|d| <= 127): q in five digits (four in the QRCP fold's column sums, below),
// Y in four planes (an FP64 sketch or the sketch's point features) or five
// (fit features, S), the pairs p + r <= 4 (13, 14 or 15) in five groups; the
// dropped pairs are 254^-7 of u_t s_j and smaller.  A kernel can take fewer
// groups (NG): the sketch runs p + r <= 3.
template <int BASE>
struct DigitScheme;
template <>
struct DigitScheme<128> {
    static constexpr int ND = 5, NG = 6;
};
template <>
struct DigitScheme<254> {
    static constexpr int ND = 5, NG = 5;
};
// The QRCP fold of an FP64 sketch (four base-254 Y planes) writes each q_t in
// four digits: its truncation, within u_t 254^-4 / 2 per entry, perturbs
// W[t,j] by ~u_t 254^-4 ||y_j|| / (2 sqrt 3) (independent roundings), and the
// folded sums by a zero-mean error whose deviation over a whole
// factorisation is below 2e-11 ||y_j||^2 (sum_t W[t,j]^2 <= ||y_j||^2), far
// inside the QRCP's margin; the fifth digit's pair (4, 0) would cost a
// fourteenth MFMA per K-step.  A row energy sum_j W[t,j]^2 reuses one rounded
// q_t for every column, so its error does not average out (~1e-10 relative):
// folds that return row energies, and the sketch and fit products, keep five
// digits.
constexpr int kFoldQDigits254 = 4;
// Digit pairs (p, r) with p + r < NG, p < ND and r < NY: the MFMAs of one
// K-step.
template <int ND, int NY, int NG>
struct DigitPairs {
    static constexpr int value = [] {
        int c = 0;
        for (int p = 0; p < ND; ++p)
            for (int r = 0; r < NY; ++r) c += p + r < NG;
        return c;
    }();
};
// Instruction classes of __builtin_amdgcn_sched_group_barrier.
constexpr int kSchedMfma = 0x008, kSchedVmemRead = 0x020, kSchedDsRead = 0x100, kSchedDsWrite = 0x200;
constexpr int FBM = 128;         // rows per workgroup
constexpr int FBN = 64;          // columns per workgroup (two tiles of 32)
constexpr int FBK = 64;          // K bytes per slice (two MFMA K-steps)
// LDS layout of the digit slices: lanes read 16-byte chunk 2 h + s of row
// lane % 32 (h = lane / 32).  Two layouts are free of bank conflicts on gfx950
// (measured with SQ_LDS_BANK_CONFLICT; unswizzled 64-, 96- and 128-byte rows
// conflict): rows padded to 80 bytes, and unpadded 64-byte rows whose chunk
// index is XORed with (row >> 2) & 3.  The 8-wave workgroups use the padded
// rows (the two measured within 2% of each other there); the 4-wave ones the
// swizzled rows, whose smaller footprint fits two workgroups per CU.
template <int NR>
struct SliceLayout {
    static constexpr bool kSwizzle = NR < 4;
    static constexpr int kPitch = kSwizzle ? FBK : FBK + 16;
    // Byte offset of 16-byte chunk c of row r.
    static __device__ __forceinline__ int at(int r, int c)
    {
        return r * kPitch + 16 * (kSwizzle ? c ^ ((r >> 2) & 3) : c);
    }
};
// Word 3 of a raw buffer descriptor on gfx9-class GPUs (32-bit data format).
constexpr int kBufferWord3 = 0x00020000;

// Q digits, slice-major: digit p of q_t[k] / u_t at
// qd[((k / 64 * ND + p) * bpad + t) * 64 + k % 64] (one slice of every plane
// and row is 64 ND bpad contiguous bytes), qs[t] = u_t.
template <int BASE, int ND = DigitScheme<BASE>::ND>
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
            x *= (double)BASE;
            const double dgt = rint(x);
            qd[(((k / FBK) * ND + p) * bpad + t) * FBK + k % FBK] = (int8_t)dgt;
            x -= dgt;  // exact: |x - dgt| <= 1/2
        }
    }
}

// In place: column j of y (FP32, m entries, column pitch 4 P bytes with P a
// multiple of 64 >= m) becomes NY digits per entry at digit_offset (zero
// past m and in the column's last (4 - NY) P bytes); scale[j] = s_j and
// ynorm2[j] = ||y_hat_j||^2.  One workgroup of kDigitizeThreads per column;
// each thread holds its entries in registers between the read of the whole
// column and the in-place writes (one barrier apart).
constexpr int kDigitizeThreads = 1024;
constexpr int kDigitizeEntries = (int)(kDigitRowsMax / kDigitizeThreads);  // entries per thread, P <= 36864
template <int NY>
__global__ __launch_bounds__(kDigitizeThreads) void digitize_columns_kernel(float* __restrict__ y, int64_t m,
                                                                            int64_t P, double* __restrict__ ynorm2,
                                                                            double* __restrict__ scale)
{
    __shared__ double part[kDigitizeThreads / 64];
    const int64_t j = blockIdx.x;
    float* src = y + j * P;  // 4 P bytes = P floats
    float x[kDigitizeEntries];
#pragma unroll
    for (int e = 0; e < kDigitizeEntries; ++e) {
        const int64_t i = threadIdx.x + (int64_t)kDigitizeThreads * e;
        x[e] = i < m ? src[i] : 0.f;
    }
    const double nrm = sqrt(ynorm2[j]);  // of the FP32 entries, >= max |y_ij|
    int ex = 0;
    if (nrm > 0) frexp(nrm, &ex);
    const double sc = nrm > 0 ? ldexp(1.0, ex + 1) : 0.0;
    const float inv = nrm > 0 ? ldexpf(1.0f, -(ex + 1)) : 0.f;
    __syncthreads();  // every entry of the column is read before its bytes are rewritten
    int8_t* planes = reinterpret_cast<int8_t*>(src);
    double acc = 0;
#pragma unroll
    for (int e = 0; e < kDigitizeEntries; ++e) {
        const int64_t i = threadIdx.x + (int64_t)kDigitizeThreads * e;
        if (i < P) {
            float v = x[e] * inv;  // exact: power-of-two scaling, |v| < 1/2
            double yh = 0, w = 1.0;
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                v *= 128.0f;
                const float dgt = rintf(v);
                v -= dgt;  // exact (Sterbenz)
                planes[digit_offset(i, r, NY)] = (int8_t)dgt;
                w *= 1.0 / 128.0;
                yh += (double)dgt * w;
            }
#pragma unroll
            for (int r = NY; r < 4; ++r) planes[(int64_t)r * P + i] = 0;  // the column's unused last bytes
            yh *= sc;
            acc += yh * yh;
        }
    }
    // ||y_hat||^2 in FP64 from the digits, summed in a fixed order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
#pragma unroll
        for (int w = 0; w < kDigitizeThreads / 64; ++w) t += part[w];
        ynorm2[j] = t;
        scale[j] = sc;
    }
}

// Four base-254 planes (|d| <= 127) in place of FP64 column j (pitch 8 P
// bytes): y_hat[i] = s_j sum_r d_r 254^-(r+1), within s_j 254^-4 / 2
// (~1.2e-10 s_j) of y; the column's bytes past its 4 P digit bytes are zeroed.
// FROM_MAX: the scale comes from the column's largest |y_ij| (found here)
// rather than from its norm, and ynorm2 is output only.  NY (4 or 5) planes:
// within s_j 254^-NY / 2 of each entry.
// T threads per column, E entries per thread (P <= T E; the host picks the
// smallest shape that holds the column, so short feature columns do not
// occupy 1024-thread workgroups).
constexpr int kDigitizeEntriesF64 = (int)(kDigitRowsMaxF64 / kDigitizeThreads);  // P <= 16384
template <bool FROM_MAX, int NY, int T = kDigitizeThreads, int E = kDigitizeEntriesF64>
__global__ __launch_bounds__(T) void digitize_columns_f64_kernel(double* __restrict__ y, int64_t m, int64_t P,
                                                                 double* __restrict__ ynorm2,
                                                                 double* __restrict__ scale)
{
    __shared__ double part[T / 64];
    __shared__ double pmax[T / 64];
    const int64_t j = blockIdx.x;
    double* src = y + j * P;  // 8 P bytes
    double x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int64_t i = threadIdx.x + (int64_t)T * e;
        x[e] = i < m ? src[i] : 0.0;
    }
    double nrm;  // >= max |y_ij|
    if (FROM_MAX) {
        double mx = 0;
#pragma unroll
        for (int e = 0; e < E; ++e) mx = fmax(mx, fabs(x[e]));
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
        if ((threadIdx.x & 63) == 0) pmax[threadIdx.x / 64] = mx;
        __syncthreads();
        mx = 0;
#pragma unroll
        for (int w = 0; w < T / 64; ++w) mx = fmax(mx, pmax[w]);
        nrm = mx;
    } else {
        nrm = sqrt(ynorm2[j]);
    }
    int ex = 0;
    if (nrm > 0) frexp(nrm, &ex);
    const double sc = nrm > 0 ? ldexp(1.0, ex + 1) : 0.0;
    const double inv = nrm > 0 ? ldexp(1.0, -(ex + 1)) : 0.0;
    __syncthreads();  // every entry of the column is read before its bytes are rewritten
    int8_t* planes = reinterpret_cast<int8_t*>(src);
    double acc = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int64_t i = threadIdx.x + (int64_t)T * e;
        if (i < P) {
            double v = x[e] * inv;  // exact: power-of-two scaling, |v| < 1/2
            double dg[NY];
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                v *= 254.0;  // |v| <= 127
                dg[r] = rint(v);
                v -= dg[r];  // exact
                planes[digit_offset(i, r, NY)] = (int8_t)dg[r];
            }
            // y_hat as the QRCP rebuilds it (Horner, lowest digit first)
            double yh = 0;
#pragma unroll
            for (int r = NY - 1; r >= 0; --r) yh = (yh + dg[r]) * (1.0 / 254.0);
#pragma unroll
            for (int r = NY; r < 8; ++r) planes[(int64_t)r * P + i] = 0;  // the column's unused last bytes
            yh *= sc;
            acc += yh * yh;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
#pragma unroll
        for (int w = 0; w < T / 64; ++w) t += part[w];
        ynorm2[j] = t;
        scale[j] = sc;
    }
}

typedef int i16v __attribute__((ext_vector_type(16)));

// NR (2 or 4): row tiles of 32 in this launch's row group.  The
// workgroup has 2 NR waves; wave w computes rows 32 (w / 2) .. +32 of the
// group for the 32 columns 32 (w % 2) .. +32, all six digit levels at once
// (v_mfma_i32_32x32x32_i8: A row l % 32 and B column l % 32 take 16 K bytes
// at offset 16 (l / 32); result element i of lane l is
// D[8 (i / 4) + 4 (l / 32) + i % 4][l % 32]).  A K slice is 64 bytes, two
// MFMA K-steps: lane half h supplies K bytes 32 h + 16 s at step s, so each
// lane reads 32 contiguous bytes of a plane per slice.  Rows past b carry
// zero digits; there are no branches in the MFMA sequence (a branch would
// make the compiler drain every outstanding load at its join).
// Epilogues: kEpiFold sums W^2 per column into proj2 (and per row into
// energy); kEpiSquare stores C[row + col * ldc] = W^2, kEpiStore
// C[row + col * ldc] = W and kEpiSubtract C[row + col * ldc] -= W;
// kEpiKhatriRao takes
// the group's rows 64..127 as the partners of rows 0..63 and stores
// C[m0 / 2 + r + col * ldc] = cscale[col] W[r] W[64 + r] (r < 64), with
// proj2[col] += the column's sum of squares and non-finite entries counted
// into *bad.
constexpr int kEpiFold = 0, kEpiSquare = 1, kEpiKhatriRao = 2, kEpiStore = 3, kEpiSubtract = 4;
template <bool ENERGY, int NR, int NY, int BASE, int EPI = kEpiFold, int ND = DigitScheme<BASE>::ND,
          int NG = DigitScheme<BASE>::NG>
__global__ __launch_bounds__(128 * NR, 4 / NR) void fold_i8_32x32_kernel(const int8_t* __restrict__ qd,
                                                                   const double* __restrict__ qs, int64_t ldq,
                                                                   int64_t bpad, int64_t b, int64_t mbase,
                                                                   const int8_t* __restrict__ yd, int64_t P,
                                                                   int colbytes,
                                                                   const double* __restrict__ yscale, int64_t m,
                                                                   int64_t n, double* __restrict__ proj2,
                                                                   double* __restrict__ energy,
                                                                   double* __restrict__ cout = nullptr,
                                                                   int64_t ldc = 0,
                                                                   const double* __restrict__ cscale = nullptr,
                                                                   unsigned long long* __restrict__ bad = nullptr,
                                                                   int lower = 0,
                                                                   double* __restrict__ colpart = nullptr)
{
    constexpr int ROWS = 32 * NR;  // Q rows staged per slice
    constexpr int NTHREADS = 128 * NR;
    using L = SliceLayout<NR>;
    __shared__ __align__(16) int8_t sQ[2][ND][ROWS * L::kPitch];
    // Each plane's Y slice is padded by 64 bytes: a stash instruction's
    // 8-lane groups write two planes of one column, which then fall on
    // different banks.
    constexpr int YPLANE = FBN * L::kPitch + 64;
    __shared__ __align__(16) int8_t sY[2][NY][YPLANE];
    __shared__ double colred[NR][FBN];
    __shared__ double rowred[2][ROWS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, l32 = lane & 31;
    const int rt = wave >> 1, ct = wave & 1;
    const int64_t g0 = (int64_t)blockIdx.x * FBN;
    // This workgroup's row group: a launch covers groups of ROWS rows from
    // mbase along y.
    const int64_t m0 = mbase + (int64_t)blockIdx.y * ROWS;
    const int64_t col = g0 + ct * 32 + l32;  // this lane's Y column
    i16v acc[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[g][e] = 0;

    const bool cok = col < n;
    const double sj = cok ? yscale[col] : 0.0;
    // Raw buffer loads: an offset past a descriptor's range returns zeros.
    // Each load's offset is a per-lane constant plus a wave-uniform slice
    // offset; slices at or past the last one get 2^31 added instead (every
    // constant and range is below 2^31, so the sum is out of range) and read
    // zeros, adding nothing.
    const int64_t ycols = std::min<int64_t>(FBN, n - g0);
    const __amdgpu_buffer_rsrc_t yrsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int8_t*>(yd + g0 * colbytes * P), (short)0, (int)(ycols * colbytes * P), kBufferWord3);
    const __amdgpu_buffer_rsrc_t qrsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int8_t*>(qd), (short)0, (int)(ND * bpad * ldq), kBufferWord3);
    // Y digits: FBN columns x NY x 64 contiguous bytes per slice (slices of
    // a column back to back), 4 NY chunks of 16 bytes per column; staged
    // through LDS once per workgroup
    // and shared by the NR row-tile waves.  When the chunks do not split
    // evenly over the threads, the threads past the last chunk load and
    // store a copy of an earlier one (the same bytes to the same place).
    constexpr int YTOTAL = NY * FBN * (FBK / 16);
    constexpr int YCHUNKS = (YTOTAL + NTHREADS - 1) / NTHREADS;
    static_assert(YCHUNKS * NTHREADS - YTOTAL < NTHREADS, "Y slice chunk count");
    uint4 ryA[YCHUNKS], ryB[YCHUNKS];
#define MPISDF_LOAD_YD(k0, ry)                                                                        \
    do {                                                                                             \
        _Pragma("unroll") for (int r_ = 0; r_ < YCHUNKS; ++r_) {                                     \
            const int cr_ = tid + NTHREADS * r_, c_ = cr_ < YTOTAL ? cr_ : cr_ - NTHREADS;          \
            const uint32_t lane_ = (uint32_t)((c_ / (4 * NY)) * colbytes * P) + 16u * (uint32_t)(c_ % (4 * NY)); \
            const uint32_t off_ = lane_ + ((uint32_t)(k0) < (uint32_t)P ? (uint32_t)NY * (uint32_t)(k0) : 0x80000000u); \
            ry[r_] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(yrsrc, off_, 0, 0)); \
        }                                                                                            \
    } while (0)
#define MPISDF_STASH_Y(buf, ry)                                                                       \
    do {                                                                                             \
        _Pragma("unroll") for (int r_ = 0; r_ < YCHUNKS; ++r_) {                                     \
            const int cr_ = tid + NTHREADS * r_, c_ = cr_ < YTOTAL ? cr_ : cr_ - NTHREADS;          \
            const int col_ = c_ / (4 * NY), p_ = (c_ % (4 * NY)) >> 2;                               \
            *reinterpret_cast<uint4*>(&sY[buf][p_][L::at(col_, c_ & 3)]) = ry[r_];                    \
        }                                                                                            \
    } while (0)
    // Q digits: ND planes x ROWS rows x 64 bytes per slice (rows of one plane
    // contiguous), QCHUNKS 16-byte chunks per thread.
    constexpr int QCHUNKS = ND * ROWS * (FBK / 16) / NTHREADS;
    static_assert(QCHUNKS * NTHREADS == ND * ROWS * (FBK / 16), "Q slice must split evenly over the threads");
    uint4 rqA[QCHUNKS], rqB[QCHUNKS];
#define MPISDF_LOAD_Q(k0, rq)                                                                         \
    do {                                                                                             \
        const uint32_t k_ = (uint32_t)(k0);                                                          \
        _Pragma("unroll") for (int r_ = 0; r_ < QCHUNKS; ++r_) {                                     \
            const int c_ = tid + NTHREADS * r_;                                                      \
            const int p_ = c_ / (ROWS * 4), rem_ = c_ - p_ * ROWS * 4, row_ = rem_ >> 2, kc_ = (rem_ & 3) * 16; \
            const uint32_t lane_ = (uint32_t)(((int64_t)p_ * bpad + m0 + row_) * FBK) + (uint32_t)kc_; \
            const uint32_t off_ = lane_ + (k_ < (uint32_t)ldq ? (k_ / FBK) * (uint32_t)(ND * bpad * FBK) : 0x80000000u); \
            rq[r_] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(qrsrc, off_, 0, 0)); \
        }                                                                                            \
    } while (0)
#define MPISDF_STASH_Q(buf, rq)                                                                       \
    do {                                                                                             \
        _Pragma("unroll") for (int r_ = 0; r_ < QCHUNKS; ++r_) {                                     \
            const int c_ = tid + NTHREADS * r_;                                                      \
            const int p_ = c_ / (ROWS * 4), rem_ = c_ - p_ * ROWS * 4, row_ = rem_ >> 2, kc_ = (rem_ & 3) * 16; \
            *reinterpret_cast<uint4*>(&sQ[buf][p_][L::at(row_, kc_ >> 4)]) = rq[r_];                    \
        }                                                                                            \
    } while (0)
    // MFMA K-step s (0 or 1) of a slice: 32 K bytes at 32 h + 16 s per lane.
#define MPISDF_MMA(buf, s)                                                                            \
    do {                                                                                             \
        i4v yv[NY];                                                                                  \
        _Pragma("unroll") for (int r = 0; r < NY; ++r)                                               \
            yv[r] = *reinterpret_cast<const i4v*>(&sY[buf][r][L::at(ct * 32 + l32, 2 * half + (s))]); \
        _Pragma("unroll") for (int p = 0; p < ND; ++p) {                                             \
            const i4v a = *reinterpret_cast<const i4v*>(&sQ[buf][p][L::at(rt * 32 + l32, 2 * half + (s))]); \
            _Pragma("unroll") for (int r = 0; r + p < NG && r < NY; ++r)                             \
                acc[p + r] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, yv[r], acc[p + r], 0, 0, 0);   \
        }                                                                                            \
    } while (0)
    // The operands of K-step s into registers, and the MFMAs of a K-step.
#define MPISDF_READ_OPS(buf, s, yv, qv)                                                               \
    do {                                                                                             \
        _Pragma("unroll") for (int r = 0; r < NY; ++r)                                               \
            yv[r] = *reinterpret_cast<const i4v*>(&sY[buf][r][L::at(ct * 32 + l32, 2 * half + (s))]); \
        _Pragma("unroll") for (int p = 0; p < ND; ++p)                                               \
            qv[p] = *reinterpret_cast<const i4v*>(&sQ[buf][p][L::at(rt * 32 + l32, 2 * half + (s))]); \
    } while (0)
#define MPISDF_MFMA_OPS(yv, qv)                                                                       \
    do {                                                                                             \
        _Pragma("unroll") for (int p = 0; p < ND; ++p)                                               \
            _Pragma("unroll") for (int r = 0; r + p < NG && r < NY; ++r)                             \
                acc[p + r] = __builtin_amdgcn_mfma_i32_32x32x32_i8(qv[p], yv[r], acc[p + r], 0, 0, 0); \
    } while (0)
    // One slice: the digits of the slice two ahead are loaded into one
    // register set, this slice is multiplied from LDS, and the other register
    // set (the next slice, loaded one step earlier) is staged into the other
    // buffer; each load has more than a slice of compute to arrive in.
    // 8-wave workgroups: K-step 0's operands are read first, and the stash
    // writes, K-step 1's reads and the next loads are each issued between
    // MFMAs, so the MFMA pipe runs while the LDS and the loads proceed (the
    // compiler otherwise hoists every LDS access ahead of the first MFMA and
    // waits for all of them: 12% slower at 128 rows x 262,144 columns).
    constexpr int kOps = NY + ND, kMoves = QCHUNKS + YCHUNKS, kMfma = 2 * DigitPairs<ND, NY, NG>::value;
    static_assert(2 * kMoves + kOps <= kMfma, "the interleave needs an MFMA per LDS write, read and load");
#define MPISDF_STEP(sl, buf, rqL, ryL, rqS, ryS)                                                      \
    do {                                                                                             \
        if constexpr (NR == 4) {                                                                     \
            i4v y0_[NY], q0_[ND], y1_[NY], q1_[ND];                                                  \
            MPISDF_READ_OPS(buf, 0, y0_, q0_);                                                       \
            MPISDF_STASH_Q((buf) ^ 1, rqS);                                                          \
            MPISDF_STASH_Y((buf) ^ 1, ryS);                                                          \
            MPISDF_READ_OPS(buf, 1, y1_, q1_);                                                       \
            MPISDF_MFMA_OPS(y0_, q0_);                                                               \
            MPISDF_MFMA_OPS(y1_, q1_);                                                               \
            MPISDF_LOAD_Q(((sl) + 2) * FBK, rqL);                                                    \
            MPISDF_LOAD_YD(((sl) + 2) * FBK, ryL);                                                   \
            __builtin_amdgcn_sched_group_barrier(kSchedDsRead, kOps, 0);                             \
            _Pragma("unroll") for (int i_ = 0; i_ < kMoves; ++i_) {                                  \
                __builtin_amdgcn_sched_group_barrier(kSchedMfma, 1, 0);                              \
                __builtin_amdgcn_sched_group_barrier(kSchedDsWrite, 1, 0);                           \
            }                                                                                        \
            _Pragma("unroll") for (int i_ = 0; i_ < kOps; ++i_) {                                    \
                __builtin_amdgcn_sched_group_barrier(kSchedMfma, 1, 0);                              \
                __builtin_amdgcn_sched_group_barrier(kSchedDsRead, 1, 0);                            \
            }                                                                                        \
            _Pragma("unroll") for (int i_ = 0; i_ < kMoves; ++i_) {                                  \
                __builtin_amdgcn_sched_group_barrier(kSchedMfma, 1, 0);                              \
                __builtin_amdgcn_sched_group_barrier(kSchedVmemRead, 1, 0);                          \
            }                                                                                        \
            if constexpr (kMfma > 2 * kMoves + kOps)                                                 \
                __builtin_amdgcn_sched_group_barrier(kSchedMfma, kMfma - 2 * kMoves - kOps, 0);      \
        } else { /* 4 waves: loads, MFMAs, then the stash (measured best there) */                  \
            MPISDF_LOAD_Q(((sl) + 2) * FBK, rqL);                                                    \
            MPISDF_LOAD_YD(((sl) + 2) * FBK, ryL);                                                   \
            __builtin_amdgcn_sched_barrier(0);                                                       \
            MPISDF_MMA(buf, 0);                                                                      \
            MPISDF_MMA(buf, 1);                                                                      \
            MPISDF_STASH_Q((buf) ^ 1, rqS);                                                          \
            MPISDF_STASH_Y((buf) ^ 1, ryS);                                                          \
        }                                                                                            \
        __syncthreads();                                                                             \
    } while (0)

    // K extent: lower-triangular rows have no entries past the group's last row.
    const int64_t mk = lower ? std::min<int64_t>(m, m0 + std::min<int64_t>(ROWS, b - m0)) : m;
    const int64_t nsl = (mk + FBK - 1) / FBK;
    MPISDF_LOAD_Q(0, rqA);
    MPISDF_LOAD_YD(0, ryA);
    MPISDF_LOAD_Q(FBK, rqB);
    MPISDF_LOAD_YD(FBK, ryB);
    MPISDF_STASH_Q(0, rqA);
    MPISDF_STASH_Y(0, ryA);
    __syncthreads();
    for (int64_t sl = 0; sl < nsl; sl += 2) {
        MPISDF_STEP(sl, 0, rqA, ryA, rqB, ryB);
        MPISDF_STEP(sl + 1, 1, rqB, ryB, rqA, ryA);
    }
#undef MPISDF_STEP
#undef MPISDF_MFMA_OPS
#undef MPISDF_READ_OPS
#undef MPISDF_STASH_Y
#undef MPISDF_MMA
#undef MPISDF_LOAD_YD
#undef MPISDF_LOAD_Q
#undef MPISDF_STASH_Q

    // W[t][j] = u_t s_j sum_g BASE^-(g+2) acc_g.
    if constexpr (EPI == kEpiSquare || EPI == kEpiStore || EPI == kEpiSubtract) {
        // lane (l32, half) owns column col and, per e / 4, four consecutive rows
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4) {
            double w4[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int e = 4 * q4 + i;
                double v = 0;
#pragma unroll
                for (int g = NG - 1; g >= 0; --g) v = v * (1.0 / BASE) + (double)acc[g][e];
                const int64_t row = m0 + rt * 32 + 8 * q4 + 4 * half + i;
                const double wv = row < b ? v * (1.0 / ((double)BASE * BASE)) * qs[row] * sj : 0.0;
                w4[i] = EPI == kEpiSquare ? wv * wv : wv;
            }
            const int64_t row0 = m0 + rt * 32 + 8 * q4 + 4 * half;
            if (cok && row0 < b) {
                double* dst = cout + row0 + col * ldc;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (row0 + i < b) dst[i] = EPI == kEpiSubtract ? dst[i] - w4[i] : w4[i];
            }
        }
        return;
    }
    if constexpr (EPI == kEpiKhatriRao) {
        static_assert(NR == 4, "the Khatri-Rao epilogue pairs rows r and 64 + r of a 128-row group");
        // waves rt = 2, 3 (rows 64..127) hand their W to waves rt = 0, 1
        // through LDS (the operand buffers are free after the last barrier)
        double* xw = reinterpret_cast<double*>(&sQ[0][0][0]);
        double wv[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            double v = 0;
#pragma unroll
            for (int g = NG - 1; g >= 0; --g) v = v * (1.0 / BASE) + (double)acc[g][e];
            const int64_t row = m0 + rt * 32 + 8 * (e / 4) + 4 * half + (e % 4);
            wv[e] = row < b ? v * (1.0 / ((double)BASE * BASE)) * qs[row] * sj : 0.0;
        }
        if (rt >= 2) {
#pragma unroll
            for (int e = 0; e < 16; ++e) xw[(((rt - 2) * 2 + ct) * 16 + e) * 64 + lane] = wv[e];
        }
        __syncthreads();
        double colsq = 0;
        if (rt < 2) {
            const double cs = cok ? cscale[col] : 0.0;
            unsigned long long nbad = 0;
#pragma unroll
            for (int q4 = 0; q4 < 4; ++q4) {
                double y4[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int e = 4 * q4 + i;
                    const double y = cs * wv[e] * xw[((rt * 2 + ct) * 16 + e) * 64 + lane];
                    if (!isfinite(y)) ++nbad;
                    y4[i] = y;
                    colsq += y * y;
                }
                if (cok) {
                    double* dst = cout + m0 / 2 + rt * 32 + 8 * q4 + 4 * half + col * ldc;
#pragma unroll
                    for (int i = 0; i < 4; ++i) dst[i] = y4[i];
                }
            }
            if (nbad) atomicAdd(bad, nbad);
            colsq += __shfl_xor(colsq, 32);
            if (half == 0) colred[rt][ct * 32 + l32] = colsq;
        }
        __syncthreads();
        if (tid < FBN) {
            const int64_t c = g0 + tid;
            const double sum = colred[0][tid] + colred[1][tid];
            if (c < n) {
                // one partial per row group when the launch batches them (added
                // in group order afterwards); else row groups run in order
                if (colpart) colpart[(int64_t)blockIdx.y * n + c] = sum;
                else proj2[c] += sum;
            }
        }
        return;
    }
    double colsq = 0;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        double v = 0;
#pragma unroll
        for (int g = NG - 1; g >= 0; --g) v = v * (1.0 / BASE) + (double)acc[g][e];
        const int lrow = rt * 32 + 8 * (e / 4) + 4 * half + (e % 4);
        const int64_t row = m0 + lrow;
        const double wv = row < b ? v * (1.0 / ((double)BASE * BASE)) * qs[row] * sj : 0.0;
        colsq += wv * wv;
        if (ENERGY) {
            double s = wv * wv;
#pragma unroll
            for (int off = 1; off < 32; off <<= 1) s += __shfl_xor(s, off);
            if (l32 == 0) rowred[ct][lrow] = s;
        }
    }
    colsq += __shfl_xor(colsq, 32);
    // Column sums over the row tiles in tile order (one partial per wave),
    // so the folded sums do not depend on the order the waves finish in.
    if (half == 0) colred[rt][ct * 32 + l32] = colsq;
    __syncthreads();
    if (tid < FBN) {
        const int64_t c = g0 + tid;
        double sum = 0;
#pragma unroll
        for (int t = 0; t < NR; ++t) sum += colred[t][tid];
        if (c < n) {
            if (colpart) colpart[(int64_t)blockIdx.y * n + c] = sum;  // added in group order afterwards
            else proj2[c] += sum;  // row groups run in order
        }
    }
    if (ENERGY && tid < ROWS) {
        // energy holds this launch's per-column-block partials, [block][rows]
        // with rows = min(ROWS, b - m0); the launcher adds the blocks in order.
        const int64_t rows = b - m0 < ROWS ? b - m0 : ROWS;
        if (tid < rows) energy[(int64_t)blockIdx.x * rows + tid] = rowred[0][tid] + rowred[1][tid];
    }
}

// rows: Q^T rows of this launch's group (1..128).  Groups of more than 64
// rows run on 8-wave workgroups (NR = 4, one per CU); smaller groups on
// 4-wave workgroups (NR = 2, two per CU), which multiply no zero rows.
template <int NY, int BASE, int ND = DigitScheme<BASE>::ND>
void launch_fold_i8(int64_t rows, bool energy, dim3 grid, hipStream_t s, const int8_t* qd, const double* qs,
                    int64_t ldq, int64_t bpad, int64_t b, int64_t m0, const int8_t* yd, int64_t P, int colbytes,
                    const double* yscale, int64_t m, int64_t n, double* proj2, double* en)
{
    if (rows <= 0 || rows > FBM) throw std::logic_error("INT8 fold row group out of range");
    double* const e = energy ? en : nullptr;
#define MPISDF_FOLD(E, NR)                                                                                    \
    hipLaunchKernelGGL((fold_i8_32x32_kernel<E, NR, NY, BASE, kEpiFold, ND>), grid, dim3(128 * NR), 0, s, qd, qs, ldq, bpad, b, m0, yd, P, \
                       colbytes, yscale, m, n, proj2, e)
    if (rows > 64) {
        if (energy) MPISDF_FOLD(true, 4);
        else MPISDF_FOLD(false, 4);
    } else {
        if (energy) MPISDF_FOLD(true, 2);
        else MPISDF_FOLD(false, 2);
    }
#undef MPISDF_FOLD
    MPISDF_HIP(hipGetLastError());
}

}  // namespace

// One workgroup per column, shaped to the pitch (P <= 16384).
template <bool FROM_MAX, int NY>
static void launch_digitize_f64(double* y, int64_t m, int64_t n, int64_t P, double* ynorm2, double* scale,
                                hipStream_t stream)
{
    if (P <= 256 * 6)
        hipLaunchKernelGGL((digitize_columns_f64_kernel<FROM_MAX, NY, 256, 6>), dim3((unsigned)n), dim3(256), 0,
                           stream, y, m, P, ynorm2, scale);
    else if (P <= 256 * 16)
        hipLaunchKernelGGL((digitize_columns_f64_kernel<FROM_MAX, NY, 256, 16>), dim3((unsigned)n), dim3(256), 0,
                           stream, y, m, P, ynorm2, scale);
    else
        hipLaunchKernelGGL((digitize_columns_f64_kernel<FROM_MAX, NY>), dim3((unsigned)n), dim3(kDigitizeThreads), 0,
                           stream, y, m, P, ynorm2, scale);
}

void digitize_columns_f64(double* y, int64_t m, int64_t n, int64_t P, double* ynorm2, double* scale,
                          hipStream_t stream)
{
    if (n <= 0) return;
    if (P % kDigitSlice || P < m) throw std::invalid_argument("digit planes need a pitch that is a multiple of 64 and >= m");
    if (P > kDigitRowsMaxF64)
        throw std::invalid_argument("digitize_columns_f64 holds one column in registers (pitch <= 16384)");
    if (n > 0x7fffffff) throw std::overflow_error("too many columns to digitise in one launch");
    launch_digitize_f64<false, 4>(y, m, n, P, ynorm2, scale, stream);
    MPISDF_HIP(hipGetLastError());
}

void digitize_columns_f64_max(double* y, int64_t m, int64_t n, int64_t P, double* ynorm2, double* scale,
                              hipStream_t stream, int planes)
{
    if (planes != 4 && planes != 5) throw std::invalid_argument("columns are digitised into four or five planes");
    if (n <= 0) return;
    if (P % kDigitSlice || P < m) throw std::invalid_argument("digit planes need a pitch that is a multiple of 64 and >= m");
    if (P > kDigitRowsMaxF64)
        throw std::invalid_argument("digitize_columns_f64 holds one column in registers (pitch <= 16384)");
    if (n > 0x7fffffff) throw std::overflow_error("too many columns to digitise in one launch");
    if (planes == 4)
        launch_digitize_f64<true, 4>(y, m, n, P, ynorm2, scale, stream);
    else
        launch_digitize_f64<true, 5>(y, m, n, P, ynorm2, scale, stream);
    MPISDF_HIP(hipGetLastError());
}

void split_rows_i8(const double* a, int64_t lda, int64_t rows, int64_t kdim, FoldI8Rows& out, hipStream_t stream,
                   int nd)
{
    if (rows <= 0 || kdim <= 0 || lda < kdim) throw std::invalid_argument("invalid row digit operand");
    if (nd != 4 && nd != 5) throw std::invalid_argument("row digits come in four or five base-254 digits");
    if (kdim > 32767) throw std::invalid_argument("the base-254 fold needs at most 32767 entries per row");
    out.rows = rows;
    out.kdim = kdim;
    out.bpad = (rows + FBM - 1) / FBM * FBM;
    out.ldq = (kdim + FBK - 1) / FBK * FBK;
    out.nd = nd;
    const size_t need = (size_t)(nd * out.bpad * out.ldq);
    if (need >= (size_t(1) << 31)) throw std::invalid_argument("the INT8 fold's row digits exceed 2 GiB");
    if (out.digits.count < need) out.digits.resize(need);
    if (out.scales.count < (size_t)out.bpad) out.scales.resize((size_t)out.bpad);
    if (out.bpad > 0x7fffffff) throw std::overflow_error("too many rows to digitise");
    if (nd == 4)
        hipLaunchKernelGGL((split_q_kernel<254, 4>), dim3((unsigned)out.bpad), dim3(256), 0, stream, a, lda, kdim, rows,
                           out.ldq, out.bpad, out.digits.ptr, out.scales.ptr);
    else
        hipLaunchKernelGGL((split_q_kernel<254, 5>), dim3((unsigned)out.bpad), dim3(256), 0, stream, a, lda, kdim, rows,
                           out.ldq, out.bpad, out.digits.ptr, out.scales.ptr);
    MPISDF_HIP(hipGetLastError());
}

// out[(128 g + r) * d + a] = u[a * ell + 64 g + r], out[(128 g + 64 + r) * d + a]
// = v[a * ell + 64 g + r] (r < 64; zero past ell).
__global__ void stack_khatri_rao_kernel(const double* __restrict__ u, const double* __restrict__ v, int64_t d,
                                        int64_t ell, int64_t rows, double* __restrict__ out)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= rows * d) return;
    const int64_t row = idx / d, a = idx % d;
    const int64_t g = row / 128, r = row % 128, sk = 64 * g + (r & 63);
    const double* src = r < 64 ? u : v;
    out[idx] = sk < ell ? src[a * ell + sk] : 0.0;
}

void split_khatri_rao_rows_i8(const double* u, const double* v, int64_t d, int64_t ell, FoldI8Rows& out,
                              hipStream_t stream)
{
    const int64_t rows = 2 * ((ell + 63) / 64 * 64);
    DeviceBuffer<double> stacked((size_t)(rows * d));
    hipLaunchKernelGGL(stack_khatri_rao_kernel, dim3((unsigned)((rows * d + 255) / 256)), dim3(256), 0, stream, u, v,
                       d, ell, rows, stacked.ptr);
    MPISDF_HIP(hipGetLastError());
    split_rows_i8(stacked.ptr, d, rows, d, out, stream, kSketchDigits254);
    MPISDF_HIP(hipStreamSynchronize(stream));  // before the stacked rows are freed
}

// The 128-row groups of q that run on 8-wave workgroups in one launch (grid.y
// walks the groups): every group but a last one of at most 64 rows, which
// runs on its own 4-wave launch.
static int64_t batched_groups(int64_t rows)
{
    const int64_t g = rows / FBM + (rows % FBM > 64 ? 1 : 0);
    if (g > 65535) throw std::overflow_error("too many 128-row groups for one launch");
    return g;
}

// Room for one column partial per row group and column of a batched launch.
static double* group_partials(const FoldI8Rows& q, int64_t groups, int64_t n)
{
    const size_t need = (size_t)(groups * n);
    if (q.part.count < need) q.part.resize(need);
    return q.part.ptr;
}

void fold_rows_square_i8(const FoldI8Rows& q, const int8_t* Yd, int64_t P, const double* yscale, int64_t n, double* c,
                         int64_t ldc, hipStream_t stream)
{
    if (n <= 0) return;
    if (P % kDigitSlice || P < q.kdim) throw std::invalid_argument("digit planes need a pitch that is a multiple of 64 and >= m");
    if (q.ldq > 26624) throw std::invalid_argument("the five-plane base-254 fold takes at most 26624 entries per row");
    if (q.nd != 5) throw std::invalid_argument("this product takes row digits split into five digits");
    if (ldc < q.rows) throw std::invalid_argument("the product's leading dimension is shorter than its rows");
    const int64_t gx = (n + FBN - 1) / FBN;
    if (gx > 0x7fffffff) throw std::overflow_error("fold grid too large");
    const int64_t g = batched_groups(q.rows);
    if (g > 0)
        hipLaunchKernelGGL((fold_i8_32x32_kernel<false, 4, 5, 254, kEpiSquare>), dim3((unsigned)gx, (unsigned)g),
                           dim3(512), 0, stream, q.digits.ptr, q.scales.ptr, q.ldq, q.bpad, q.rows, (int64_t)0, Yd, P,
                           8, yscale, q.kdim, n, nullptr, nullptr, c, ldc, nullptr, nullptr, 0, nullptr);
    if (g * FBM < q.rows)
        hipLaunchKernelGGL((fold_i8_32x32_kernel<false, 2, 5, 254, kEpiSquare>), dim3((unsigned)gx), dim3(256), 0,
                           stream, q.digits.ptr, q.scales.ptr, q.ldq, q.bpad, q.rows, g * FBM, Yd, P, 8, yscale,
                           q.kdim, n, nullptr, nullptr, c, ldc, nullptr, nullptr, 0, nullptr);
    MPISDF_HIP(hipGetLastError());
}

void fold_rows_subtract_i8(const FoldI8Rows& q, const int8_t* Yd, int64_t P, const double* yscale, int64_t n,
                           double* c, int64_t ldc, hipStream_t stream)
{
    if (n <= 0) return;
    if (P % kDigitSlice || P < q.kdim) throw std::invalid_argument("digit planes need a pitch that is a multiple of 64 and >= m");
    if (q.ldq > 26624) throw std::invalid_argument("the five-plane base-254 fold takes at most 26624 entries per row");
    if (q.nd != 5) throw std::invalid_argument("this product takes row digits split into five digits");
    if (ldc < q.rows) throw std::invalid_argument("the product's leading dimension is shorter than its rows");
    const int64_t gx = (n + FBN - 1) / FBN;
    if (gx > 0x7fffffff) throw std::overflow_error("fold grid too large");
    const int64_t g = batched_groups(q.rows);
    if (g > 0)
        hipLaunchKernelGGL((fold_i8_32x32_kernel<false, 4, 5, 254, kEpiSubtract>), dim3((unsigned)gx, (unsigned)g),
                           dim3(512), 0, stream, q.digits.ptr, q.scales.ptr, q.ldq, q.bpad, q.rows, (int64_t)0, Yd, P,
                           8, yscale, q.kdim, n, nullptr, nullptr, c, ldc, nullptr, nullptr, 0, nullptr);
    if (g * FBM < q.rows)
        hipLaunchKernelGGL((fold_i8_32x32_kernel<false, 2, 5, 254, kEpiSubtract>), dim3((unsigned)gx), dim3(256), 0,
                           stream, q.digits.ptr, q.scales.ptr, q.ldq, q.bpad, q.rows, g * FBM, Yd, P, 8, yscale,
                           q.kdim, n, nullptr, nullptr, c, ldc, nullptr, nullptr, 0, nullptr);
    MPISDF_HIP(hipGetLastError());
}

void fold_rows_khatri_rao_i8(const FoldI8Rows& q, const int8_t* Yd, int64_t P, const double* yscale, int64_t n,
                             const double* cscale, double* y, int64_t ldy, double* ynorm2, unsigned long long* bad,
                             hipStream_t stream)
{
    if (n <= 0) return;
    if (q.rows % FBM) throw std::invalid_argument("Khatri-Rao row digits come in groups of 64 + 64 rows");
    if (P % kDigitSlice || P < q.kdim) throw std::invalid_argument("digit planes need a pitch that is a multiple of 64 and >= m");
    if (q.ldq > 26624) throw std::invalid_argument("the five-plane base-254 fold takes at most 26624 entries per row");
    if (q.nd != kSketchDigits254) throw std::invalid_argument("Khatri-Rao row digits come from split_khatri_rao_rows_i8");
    if (ldy < q.rows / 2) throw std::invalid_argument("the sketch's leading dimension is shorter than its rows");
    const int64_t gx = (n + FBN - 1) / FBN;
    if (gx > 0x7fffffff) throw std::overflow_error("fold grid too large");
    const int64_t g = batched_groups(q.rows);
    if (g <= 0) return;
    double* part = group_partials(q, g, n);
    hipLaunchKernelGGL((fold_i8_32x32_kernel<false, 4, kSketchDigits254, 254, kEpiKhatriRao, kSketchDigits254,
                                             kSketchDigits254>),
                       dim3((unsigned)gx, (unsigned)g),
                       dim3(512), 0, stream, q.digits.ptr, q.scales.ptr, q.ldq, q.bpad, q.rows, (int64_t)0, Yd, P, 8,
                       yscale, q.kdim, n, ynorm2, nullptr, y, ldy, cscale, bad, 0, part);
    MPISDF_HIP(hipGetLastError());
    launch_sum_partials(part, g, n, ynorm2, true, stream);  // the groups' column sums in group order
}

void fold_rows_i8(const FoldI8Rows& q, bool lower, const int8_t* Yd, int64_t P, const double* yscale, int64_t n,
                  double* proj2, hipStream_t stream)
{
    if (n <= 0) return;
    if (P % kDigitSlice || P < q.kdim) throw std::invalid_argument("digit planes need a pitch that is a multiple of 64 and >= m");
    // five Y planes: up to five pairs share an accumulator, 5 * 127^2 * ldq < 2^31
    if (q.ldq > 26624) throw std::invalid_argument("the five-plane base-254 fold takes at most 26624 entries per row");
    if (q.nd != 5) throw std::invalid_argument("this product takes row digits split into five digits");
    const int64_t gx = (n + FBN - 1) / FBN;
    if (gx > 0x7fffffff) throw std::overflow_error("fold grid too large");
    const int64_t g = batched_groups(q.rows);
    if (g > 0) {
        // each group folds its own K extent (lower: up to its last row)
        double* part = group_partials(q, g, n);
        hipLaunchKernelGGL((fold_i8_32x32_kernel<false, 4, 5, 254, kEpiFold>), dim3((unsigned)gx, (unsigned)g),
                           dim3(512), 0, stream, q.digits.ptr, q.scales.ptr, q.ldq, q.bpad, q.rows, (int64_t)0, Yd,
                           P, 8, yscale, q.kdim, n, proj2, nullptr, nullptr, 0, nullptr, nullptr, lower ? 1 : 0, part);
        MPISDF_HIP(hipGetLastError());
        launch_sum_partials(part, g, n, proj2, true, stream);  // the groups' column sums in group order
    }
    if (g * FBM < q.rows) {
        const int64_t m0 = g * FBM, rows = q.rows - m0;
        // row t of a lower-triangular operand has no entries past t
        const int64_t kext = lower ? std::min<int64_t>(q.kdim, m0 + rows) : q.kdim;
        launch_fold_i8<5, 254>(rows, false, dim3((unsigned)gx), stream, q.digits.ptr, q.scales.ptr, q.ldq, q.bpad,
                               q.rows, m0, Yd, P, 8, yscale, kext, n, proj2, nullptr);
    }
}

void digitize_columns(float* y, int64_t m, int64_t n, int64_t P, int planes, double* ynorm2, double* scale,
                      hipStream_t stream)
{
    if (planes != 3 && planes != 4) throw std::invalid_argument("digit planes must be 3 or 4");
    if (n <= 0) return;
    if (P % kDigitSlice || P < m) throw std::invalid_argument("digit planes need a pitch that is a multiple of 64 and >= m");
    if (P > kDigitRowsMax) throw std::invalid_argument("digitize_columns holds one column in registers (pitch <= 36864)");
    if (n > 0x7fffffff) throw std::overflow_error("too many columns to digitise in one launch");
    if (planes == 4)
        hipLaunchKernelGGL(digitize_columns_kernel<4>, dim3((unsigned)n), dim3(kDigitizeThreads), 0, stream, y, m, P,
                           ynorm2, scale);
    else
        hipLaunchKernelGGL(digitize_columns_kernel<3>, dim3((unsigned)n), dim3(kDigitizeThreads), 0, stream, y, m, P,
                           ynorm2, scale);
    MPISDF_HIP(hipGetLastError());
}

// Candidate digit columns side by side: column c of out (4 P bytes) is the
// first 4 P bytes of column cand[c] of Yd (pitch colbytes P), scale oscale[c].
__global__ void gather_digit_columns_kernel(const int8_t* __restrict__ yd, int64_t P, int colbytes,
                                            const double* __restrict__ yscale, const int64_t* __restrict__ cand,
                                            int8_t* __restrict__ out, double* __restrict__ oscale)
{
    const int64_t c = blockIdx.x, j = cand[c];
    const uint4* src = reinterpret_cast<const uint4*>(yd + j * colbytes * P);
    uint4* dst = reinterpret_cast<uint4*>(out + c * 4 * P);
    for (int64_t i = threadIdx.x; i < P / 4; i += blockDim.x) dst[i] = src[i];  // 4 P bytes in 16-byte chunks
    if (threadIdx.x == 0) oscale[c] = yscale[j];
}

void q_digits_extend(const double* Q, int64_t ldq, int64_t m, int64_t k, int64_t kcap, QDigitStore& st,
                     hipStream_t stream)
{
    if (m > 32767) throw std::invalid_argument("the base-254 digits of q need at most 32767 entries");
    if (st.cap == 0) {
        st.cap = std::max<int64_t>(kcap, 1);
        st.bpad = (st.cap + FBM - 1) / FBM * FBM;
        st.ldq = (m + FBK - 1) / FBK * FBK;
        const size_t bytes = (size_t)(kFoldQDigits254 * st.bpad * st.ldq);
        if (bytes >= (size_t(1) << 31)) throw std::invalid_argument("the digits of Q exceed 2 GiB");
        st.digits.resize(bytes);
        st.scales.resize((size_t)st.bpad);
        // rows past the digitised ones read as zero digits
        MPISDF_HIP(hipMemsetAsync(st.digits.ptr, 0, bytes, stream));
        MPISDF_HIP(hipMemsetAsync(st.scales.ptr, 0, st.bpad * sizeof(double), stream));
        st.rows = 0;
    }
    if (k > st.cap) throw std::invalid_argument("more pivot vectors than the digit store holds");
    if (k <= st.rows) return;
    // rows st.rows .. k-1: row t of this launch lands at st.rows + t of the layout
    hipLaunchKernelGGL((split_q_kernel<254, kFoldQDigits254>), dim3((unsigned)(k - st.rows)), dim3(256), 0, stream,
                       Q + st.rows * ldq, ldq, m, k - st.rows, st.ldq, st.bpad, st.digits.ptr + st.rows * FBK,
                       st.scales.ptr + st.rows);
    MPISDF_HIP(hipGetLastError());
    st.rows = k;
}

void project_candidates_i8(QDigitStore& st, int64_t m, const int8_t* Yd, int64_t P, int colbytes,
                           const double* yscale, const int64_t* cand, int64_t nc, double* H, int64_t ldh,
                           hipStream_t stream)
{
    const int64_t k = st.rows;
    if (k <= 0 || nc <= 0) return;
    if (P % kDigitSlice || P < m || st.ldq > P) throw std::invalid_argument("candidate digit columns need a pitch >= the digits of q");
    if (colbytes != 4 && colbytes != 8) throw std::invalid_argument("digit columns take 4 or 8 bytes per entry");
    if (ldh < k) throw std::invalid_argument("the projection's leading dimension is shorter than the pivots");
    if (nc > 0x7fffffff) throw std::overflow_error("too many candidate columns");
    const size_t need = (size_t)(nc * 4 * P);
    if (st.gathered.count < need) st.gathered.resize(need);
    if (st.gscales.count < (size_t)nc) st.gscales.resize((size_t)nc);
    hipLaunchKernelGGL(gather_digit_columns_kernel, dim3((unsigned)nc), dim3(256), 0, stream, Yd, P, colbytes, yscale,
                       cand, st.gathered.ptr, st.gscales.ptr);
    MPISDF_HIP(hipGetLastError());
    const int64_t gx = (nc + FBN - 1) / FBN;
    const int64_t g = batched_groups(k);
    if (g > 0)
        hipLaunchKernelGGL((fold_i8_32x32_kernel<false, 4, 4, 254, kEpiStore, kFoldQDigits254>),
                           dim3((unsigned)gx, (unsigned)g), dim3(512), 0, stream, st.digits.ptr, st.scales.ptr, st.ldq,
                           st.bpad, k, (int64_t)0, st.gathered.ptr, P, 4, st.gscales.ptr, m, nc, nullptr, nullptr, H,
                           ldh, nullptr, nullptr, 0, nullptr);
    if (g * FBM < k)
        hipLaunchKernelGGL((fold_i8_32x32_kernel<false, 2, 4, 254, kEpiStore, kFoldQDigits254>), dim3((unsigned)gx),
                           dim3(256), 0, stream, st.digits.ptr, st.scales.ptr, st.ldq, st.bpad, k, g * FBM,
                           st.gathered.ptr, P, 4, st.gscales.ptr, m, nc, nullptr, nullptr, H, ldh, nullptr, nullptr,
                           0, nullptr);
    MPISDF_HIP(hipGetLastError());
}

void fold_projections_i8(int64_t b, int64_t n, int64_t m, const double* Qb, int64_t ldqb, const int8_t* Yd,
                         int64_t P, int planes, int base, int colbytes, const double* yscale, double* proj2,
                         double* energy, FoldI8Scratch& scratch, hipStream_t stream)
{
    if (base == 128 ? (planes != 3 && planes != 4) : (base != 254 || planes != 4))
        throw std::invalid_argument("digit planes must be 3 or 4 of base 128, or 4 of base 254");
    if (colbytes != 4 && colbytes != 8)
        throw std::invalid_argument("digit columns take the 4 or 8 bytes per entry of the entries they replace");
    if (base == 254 && m > 32767) throw std::invalid_argument("the base-254 fold needs at most 32767 sketch rows");
    if (b <= 0 || n <= 0) return;
    if (m >= (int64_t(1) << 16)) throw std::invalid_argument("the INT8 fold needs fewer than 65536 sketch rows");
    if (P % kDigitSlice || P < m) throw std::invalid_argument("digit planes need a pitch that is a multiple of 64 and >= m");
    const int64_t bpad = (b + FBM - 1) / FBM * FBM;
    const int64_t ldq = (m + FBK - 1) / FBK * FBK;
    const int nd = base == 254 ? (energy ? DigitScheme<254>::ND : kFoldQDigits254) : DigitScheme<128>::ND;
    const size_t need = (size_t)(nd * bpad * ldq);
    if (need >= (size_t(1) << 31)) throw std::invalid_argument("the INT8 fold's Q digits exceed 2 GiB");
    if (scratch.digits.count < need) scratch.digits.resize(need);
    if (scratch.scales.count < (size_t)bpad) scratch.scales.resize(bpad);
    if (base == 254 && !energy)
        hipLaunchKernelGGL((split_q_kernel<254, kFoldQDigits254>), dim3((unsigned)bpad), dim3(256), 0, stream, Qb, ldqb,
                           m, b, ldq, bpad, scratch.digits.ptr, scratch.scales.ptr);
    else if (base == 254)
        hipLaunchKernelGGL(split_q_kernel<254>, dim3((unsigned)bpad), dim3(256), 0, stream, Qb, ldqb, m, b, ldq, bpad,
                           scratch.digits.ptr, scratch.scales.ptr);
    else
        hipLaunchKernelGGL(split_q_kernel<128>, dim3((unsigned)bpad), dim3(256), 0, stream, Qb, ldqb, m, b, ldq, bpad,
                           scratch.digits.ptr, scratch.scales.ptr);
    MPISDF_HIP(hipGetLastError());
    const int64_t gx = (n + FBN - 1) / FBN;
    if (gx > 0x7fffffff) throw std::overflow_error("fold grid too large");
    if (!energy && b > FBM) {
        // Many row groups (a fold of every pivot so far): the groups of more
        // than 64 rows in one launch, each writing its column partial, added
        // in group order, then a last group of at most 64 rows; the columns in
        // chunks that bound the partials.
        const int64_t g = batched_groups(b);
        const int64_t chunk = std::max<int64_t>(FBN, (int64_t(1) << 25) / std::max<int64_t>(g, 1) / FBN * FBN);
        const size_t need_part = (size_t)(g * std::min(chunk, n));
        if (scratch.colpart.count < need_part) scratch.colpart.resize(need_part);
        for (int64_t c0 = 0; c0 < n; c0 += chunk) {
            const int64_t nc = std::min(chunk, n - c0);
            const int8_t* yc = Yd + c0 * colbytes * P;
            const dim3 grid((unsigned)((nc + FBN - 1) / FBN), (unsigned)g);
#define MPISDF_FOLD_BATCH(NY, BASE, ND)                                                                          \
    hipLaunchKernelGGL((fold_i8_32x32_kernel<false, 4, NY, BASE, kEpiFold, ND>), grid, dim3(512), 0, stream,       \
                       scratch.digits.ptr, scratch.scales.ptr, ldq, bpad, b, (int64_t)0, yc, P, colbytes,          \
                       yscale + c0, m, nc, proj2 + c0, nullptr, nullptr, 0, nullptr, nullptr, 0, scratch.colpart.ptr)
            if (g > 0) {
                if (base == 254) MPISDF_FOLD_BATCH(4, 254, kFoldQDigits254);
                else if (planes == 4) MPISDF_FOLD_BATCH(4, 128, 5);
                else MPISDF_FOLD_BATCH(3, 128, 5);
                MPISDF_HIP(hipGetLastError());
                launch_sum_partials(scratch.colpart.ptr, g, nc, proj2 + c0, true, stream);
            }
#undef MPISDF_FOLD_BATCH
            if (g * FBM < b) {
                const int64_t m0 = g * FBM, rows = b - m0;
                const dim3 g1((unsigned)((nc + FBN - 1) / FBN));
                if (base == 254)
                    launch_fold_i8<4, 254, kFoldQDigits254>(rows, false, g1, stream, scratch.digits.ptr,
                                                            scratch.scales.ptr, ldq, bpad, b, m0, yc, P, colbytes,
                                                            yscale + c0, m, nc, proj2 + c0, nullptr);
                else if (planes == 4)
                    launch_fold_i8<4, 128>(rows, false, g1, stream, scratch.digits.ptr, scratch.scales.ptr, ldq, bpad,
                                           b, m0, yc, P, colbytes, yscale + c0, m, nc, proj2 + c0, nullptr);
                else
                    launch_fold_i8<3, 128>(rows, false, g1, stream, scratch.digits.ptr, scratch.scales.ptr, ldq, bpad,
                                           b, m0, yc, P, colbytes, yscale + c0, m, nc, proj2 + c0, nullptr);
            }
        }
        return;
    }
    if (energy && scratch.part.count < (size_t)(gx * FBM)) scratch.part.resize((size_t)(gx * FBM));
    // One launch per group of 128 rows (Q^T rows), each with its own tile
    // count; the row energies of a group are summed over column blocks in
    // block order.
    for (int64_t m0 = 0; m0 < b; m0 += FBM) {
        const int64_t rows = std::min<int64_t>(FBM, b - m0);
        if (base == 254 && !energy)
            launch_fold_i8<4, 254, kFoldQDigits254>(rows, false, dim3((unsigned)gx), stream, scratch.digits.ptr,
                                                    scratch.scales.ptr, ldq, bpad, b, m0, Yd, P, colbytes, yscale, m, n,
                                                    proj2, scratch.part.ptr);
        else if (base == 254)
            launch_fold_i8<4, 254>(rows, energy != nullptr, dim3((unsigned)gx), stream, scratch.digits.ptr,
                                   scratch.scales.ptr, ldq, bpad, b, m0, Yd, P, colbytes, yscale, m, n, proj2,
                                   scratch.part.ptr);
        else if (planes == 4)
            launch_fold_i8<4, 128>(rows, energy != nullptr, dim3((unsigned)gx), stream, scratch.digits.ptr,
                              scratch.scales.ptr, ldq, bpad, b, m0, Yd, P, colbytes, yscale, m, n, proj2,
                              scratch.part.ptr);
        else
            launch_fold_i8<3, 128>(rows, energy != nullptr, dim3((unsigned)gx), stream, scratch.digits.ptr,
                              scratch.scales.ptr, ldq, bpad, b, m0, Yd, P, colbytes, yscale, m, n, proj2,
                              scratch.part.ptr);
        if (energy) launch_sum_partials(scratch.part.ptr, gx, rows, energy + m0, true, stream);
    }
}

}  // namespace mpisdf
