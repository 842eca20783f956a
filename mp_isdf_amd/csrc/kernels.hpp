// Kernel launchers of the mp_isdf_amd HIP library (host declarations).
#pragma once

#include "common.hpp"

namespace mpisdf {

struct Context;

// ---- AO evaluation (ao_eval.hip) -----------------------------------------
// out[comp * cstride + g * ld + ao]; comp = 0 (values) / 0..2 (gradient) /
// 0..3 (values, then gradient) depending on ``mode``.
void launch_ao_eval(Context& ctx, const double* xyz, int64_t n, AoMode mode, DType dtype,
                    void* out, int64_t ld, int64_t cstride, hipStream_t stream);

// Adjoint of gradient evaluation (FP64): gout[c * nao + mu] += sum_p
// abar[p * ldab + mu] * d phi_mu(r_p) / d r_c, gradients recomputed on the fly;
// per-tile partials (``part``, resized as needed) are added in tile order.
void launch_ao_adjoint(Context& ctx, const double* xyz, int64_t n, const double* abar, int64_t ldab, double* gout,
                       DeviceBuffer<double>& part, hipStream_t stream);

// One axis block of the concatenated feature matrix.
struct AxisSegment {
    int kind;        // 0 AO, 1 dX, 2 dY, 3 dZ, 4 occupied
    int64_t offset;  // first feature column
    int64_t width;   // nao or nocc
    double scale;
};
void launch_assemble_axis(const double* ao, const double* occ, int64_t n, int64_t nao, int64_t nocc,
                          const AxisSegment& seg, double* features, int64_t width, hipStream_t stream);

// ---- Dense FP64 GEMM (gemm.hip) -------------------------------------------
// Strided view: element (i, j) lives at p[i * rs + j * cs].
struct MatView {
    const double* p;
    int64_t rs, cs;
};
struct MutView {
    double* p;
    int64_t rs, cs;
};
// C = alpha * A[M,K] * B[K,N] + beta * C with FP64 accumulation.
void gemm_f64(int64_t M, int64_t N, int64_t K, double alpha, MatView A, MatView B, double beta,
              MutView C, hipStream_t stream);
// Same with an FP32 B operand (element (i, j) at B[i * brs + j * bcs]).
void gemm_f64_f32b(int64_t M, int64_t N, int64_t K, double alpha, MatView A, const float* B,
                   int64_t brs, int64_t bcs, double beta, MutView C, hipStream_t stream);

// TN product with both operands contiguous along K (gemm_tn.hip):
// C[t + j * ldc] = alpha * sum_k A[k + t * lda] B[k + j * ldb] + beta * C[...].
// triangular = 1 (A lower: A[t][k] = 0 for k > t) or 2 (A upper: zero for
// k < t) skips the zero part of K.
void gemm_tn_f64(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda, const double* B,
                 int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t stream, int triangular = 0);
// gemm_tn_f64 with a fused epilogue (C column-major M x N, beta = 0):
//   Square:            C = (A^T B)^2 elementwise (alpha ignored);
//   ColumnSumSquares:  C = alpha A^T B and colacc[g * N + n] = sum over the
//                      rows m of row group g (128 rows each) of C[m, n]^2;
//   RowDot:            C = alpha A^T B and rowacc[x * M + m] = sum over the
//                      columns n of column block x (tn_column_blocks(N)
//                      blocks of 128) of C[m, n] colw[n]; the caller adds
//                      the blocks in order (launch_sum_partials).
enum class TnFused { Square, ColumnSumSquares, RowDot };
void gemm_tn_f64_fused(TnFused mode, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                       const double* B, int64_t ldb, double* C, int64_t ldc, double* colacc, double* rowacc,
                       const double* colw, hipStream_t stream, int triangular = 0);

int64_t tn_column_blocks(int64_t N);

// gemm_tn_f64 (no triangular skip; beta 0 or 1) with K split over the grid
// when the product has too few 128 x 128 tiles to fill the chip: partial
// products over K chunks in ``part``, added to C in chunk order (entries
// i < M, j < N of C only).
void gemm_tn_f64_split(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda, const double* B,
                       int64_t ldb, double beta, double* C, int64_t ldc, DeviceBuffer<double>& part,
                       hipStream_t stream);

// FP64 Khatri-Rao sketch on the TN kernel: ut/vt are ell x d row-major
// (ut[s * d + a]), phi[g * ldphi + a];  y[g * ldy + s] = scale[g] (ut_s . phi_g)
// (vt_s . phi_g) for g < count, ynorm2[g] += ||y_g||^2 (when given; summed
// over row groups in a fixed order through ``part``), non-finite outputs
// counted into *bad.
void kr_sketch_f64_tn(const double* ut, const double* vt, int64_t ell, int64_t d, const double* phi, int64_t ldphi,
                      int64_t count, const double* scale, double* y, int64_t ldy, double* ynorm2,
                      DeviceBuffer<double>& part, unsigned long long* bad, hipStream_t stream);

// QRCP fold: W = Qb^T Y (Qb: m x b column-major with leading dimension ldq,
// Y: m x n with leading dimension ldy, FP32 or FP64); proj2[j] += sum_t
// W[t,j]^2 and, when energy is given, energy[t] += sum_j W[t,j]^2.  FP64
// MFMA, W never stored; every sum is added in a fixed order (``part`` holds
// the energies' per-column-block partials).
void fold_projections(int64_t b, int64_t n, int64_t m, const double* Qb, int64_t ldq, const void* Y, bool y_f32,
                      int64_t ldy, double* proj2, double* energy, DeviceBuffer<double>& part, hipStream_t stream);

// Same fold for FP32 Y with split-BF16 MFMA (hi*hi + hi*lo + lo*hi) and FP64
// accumulation across 64-wide K chunks; error ~1e-6 ||q|| ||y_j|| per entry.
struct FoldScratch {
    DeviceBuffer<uint16_t> hi, lo;
};
// Per-entry error scale of the split fold relative to ||q|| ||y_j|| (FP32
// chunk accumulation over 64 K values, dropped lo*lo term, split residuals).
constexpr double kFoldEpsW = 2e-6;
// When ynorm2/err2 are given, err2[j] accumulates a bound on the error of the
// folded sum: 2 eps ||y_j|| sqrt(s_j) + b eps^2 ||y_j||^2 per fold.
void fold_projections_bf16x3(int64_t b, int64_t n, int64_t m, const double* Qb, int64_t ldqb, const float* Y,
                             int64_t ldy,
                             double* proj2, double* energy, FoldScratch& scratch, hipStream_t stream,
                             const double* ynorm2 = nullptr, double* err2 = nullptr);

// Sketch in digit-plane form (fold_i8.hip): in place, FP32 column j
// (pitch 4 P bytes, P a multiple of 64 >= m) becomes ``planes`` (3 or 4)
// signed base-128 digits per entry, y_hat = scale[j] sum_r d_r 128^-(r+1),
// stored slice by slice (digit_offset); ynorm2[j] (in: of the FP32 column,
// out: of y_hat).  m <= 36864.
void digitize_columns(float* y, int64_t m, int64_t n, int64_t P, int planes, double* ynorm2, double* scale,
                      hipStream_t stream);
// The same for an FP64 sketch (column pitch 8 P bytes, m <= P <= 16384):
// four signed base-254 planes (|d| <= 127), y_hat = scale[j] sum_r d_r
// 254^-(r+1) within scale[j] 254^-4 / 2 (~1e-10 s_j) of each entry, so its
// greedy pivots are the FP64 sketch's; ynorm2[j] (in: of the FP64 column,
// out: of y_hat).
void digitize_columns_f64(double* y, int64_t m, int64_t n, int64_t P, double* ynorm2, double* scale,
                          hipStream_t stream);
// Five (or four) base-254 planes (within scale[j] 254^-5 / 2 ~ 4.8e-13
// scale[j], or 254^-4 / 2 ~ 1.2e-10 scale[j], of each entry), the scale taken
// from each column's largest entry (found by the kernel); ynorm2 (out)
// receives ||y_hat_j||^2.
void digitize_columns_f64_max(double* y, int64_t m, int64_t n, int64_t P, double* ynorm2, double* scale,
                              hipStream_t stream, int planes = 5);
constexpr int64_t kDigitRowsMaxF64 = 16384;  // digitize_columns_f64 holds one FP64 column in registers
constexpr int64_t kDigitRowsMax = 36864;  // digitize_columns holds one FP32 column in registers
constexpr int64_t kDigitSlice = 64;       // entries per slice; P is a multiple of it
// Digits per entry.  Three base-128 digits round an entry to 2^-22 s_j, 3e-5
// to 6e-5 of the column norm in RMS over 13,200 rows, far below the noise of
// a BF16 or FP8 sketch.  Four (2^-29 s_j, ~2e-7 of the norm) keep a split-BF16
// sketch (~1e-7) intact: its greedy pivots then agree with the FP64 form's on
// >= 98% of the anchors, where three digits alone drop the agreement to ~93%.
constexpr int kDigitPlanes = 3;
constexpr int kDigitPlanesMax = 4;
// Byte offset, within its column's 4 P bytes, of digit r of entry i: each
// slice of 64 entries keeps its digit planes in one run of 64 planes bytes and
// the runs follow each other, so a column's digits are its first planes * P
// bytes (the rest, if any, zero) and the fold streams them contiguously.
__host__ __device__ inline int64_t digit_offset(int64_t i, int r, int planes)
{
    return (i / kDigitSlice) * (planes * kDigitSlice) + r * kDigitSlice + (i % kDigitSlice);
}

// QRCP fold on the digit planes with INT8 MFMA (digit-pair products in exact
// INT32, combined in FP64); m < 65536.
struct FoldI8Scratch {
    DeviceBuffer<int8_t> digits;
    DeviceBuffer<double> scales;
    DeviceBuffer<double> part;     // per-column-block row energies of one launch
    DeviceBuffer<double> colpart;  // per-row-group column sums of a fold of many groups
};
// Relative bound (of ||y_j||^2) on the error a folded sum accumulates over a
// whole factorisation with the INT8 fold; enters the keys as a margin.
constexpr double kFoldI8Margin = 1e-9;
// base: 128 (three or four planes) or 254 (four planes, m <= 32767);
// colbytes: bytes per pitch entry of a digit column (4 for planes held in
// place of FP32 entries, 8 in place of FP64 ones).
void fold_projections_i8(int64_t b, int64_t n, int64_t m, const double* Qb, int64_t ldqb, const int8_t* Yd,
                         int64_t P, int planes, int base, int colbytes, const double* yscale, double* proj2,
                         double* energy, FoldI8Scratch& scratch, hipStream_t stream);

// The QRCP's pivot vectors q_0 .. q_{rows-1} as INT8 digits (four base-254
// digits per entry under one power of two per vector, the fold's layout with
// room for cap vectors), extended as pivots are added; for products of Q^T
// with candidate columns of a base-254 digit sketch.
struct QDigitStore {
    DeviceBuffer<int8_t> digits;
    DeviceBuffer<double> scales;
    int64_t rows = 0, cap = 0, bpad = 0, ldq = 0;
    DeviceBuffer<int8_t> gathered;   // candidate digit columns of one product
    DeviceBuffer<double> gscales;
};
// Digitise the columns rows .. k-1 of Q (m x k, leading dimension ldq, FP64)
// into st (sized for kcap vectors on first use).
void q_digits_extend(const double* Q, int64_t ldq, int64_t m, int64_t k, int64_t kcap, QDigitStore& st,
                     hipStream_t stream);
// H[t + c * ldh] = q_t . y_hat_{cand[c]} for t < st.rows, c < nc: Q^T times
// candidate columns of the four-plane base-254 digit sketch Yd (column j at
// byte offset colbytes P j, scales yscale[j]), on INT8 MFMA (13 digit pairs;
// within ~1e-11 ||y_hat|| of the exact products, the truncation of q).
void project_candidates_i8(QDigitStore& st, int64_t m, const int8_t* Yd, int64_t P, int colbytes,
                           const double* yscale, const int64_t* cand, int64_t nc, double* H, int64_t ldh,
                           hipStream_t stream);

// The rows of a row-major FP64 matrix a (rows x kdim, leading dimension lda)
// as the fold's Q operand: nd (five, or four for the Khatri-Rao sketch)
// base-254 digits per entry under one power of two per row (split once,
// folded against many digit batches).
struct FoldI8Rows {
    DeviceBuffer<int8_t> digits;
    DeviceBuffer<double> scales;
    int64_t rows = 0, kdim = 0, bpad = 0, ldq = 0;
    int nd = 5;
    // per-row-group column sums of a launch that batches the row groups
    mutable DeviceBuffer<double> part;
};
void split_rows_i8(const double* a, int64_t lda, int64_t rows, int64_t kdim, FoldI8Rows& out, hipStream_t stream,
                   int nd = 5);
// c[t + j * ldc] -= a_t . y_hat_j for the rows t of q and the n five-plane
// base-254 digit columns of Yd (pitch 8 P bytes, scales yscale[j]): the 15
// digit pairs p + r <= 4 (within ~1e-11 of |a_t| |y_j|).  kdim <= 26624.
void fold_rows_subtract_i8(const FoldI8Rows& q, const int8_t* Yd, int64_t P, const double* yscale, int64_t n,
                           double* c, int64_t ldc, hipStream_t stream);
// proj2[j] += sum_t (a_t . y_hat_j)^2 over the rows t of q, for the n
// base-254 digit columns of Yd (five planes in place of FP64 entries, from
// digitize_columns_f64_max; pitch 8 P bytes, scales yscale[j]); the 15 digit
// pairs p + r <= 4.  lower: a is lower triangular, and each 128-row group
// folds only the entries up to its last row.  kdim <= 26624.
void fold_rows_i8(const FoldI8Rows& q, bool lower, const int8_t* Yd, int64_t P, const double* yscale, int64_t n,
                  double* proj2, hipStream_t stream);
// c[t + j * ldc] = (a_t . y_hat_j)^2 for every row t of q (same digits and
// limits as fold_rows_i8).
void fold_rows_square_i8(const FoldI8Rows& q, const int8_t* Yd, int64_t P, const double* yscale, int64_t n, double* c,
                         int64_t ldc, hipStream_t stream);
// Khatri-Rao products: q's rows come in groups of 128, rows 64 g + r of U
// then the same rows of V (zero past the sketch), and
//   y[s + j * ldy] = cscale[j] (u_s . y_hat_j)(v_s . y_hat_j),
// ynorm2[j] += ||y_j||^2 (summed per group in group order), non-finite
// entries counted into *bad.  U, V and the features (digitize_columns_f64_max
// with kSketchDigits254 planes) come in four base-254 digits, multiplied in
// the ten digit pairs p + r <= 3: each factor u_s . z lands within ~2e-10 of
// its FP64 value relative to |u_s| |z| (RMS), below the ~5e-10 of the column
// norm at which the QRCP then holds each sketch column (four planes,
// digitize_columns_f64), so a fifth digit and the five pairs p + r = 4 would
// buy accuracy that the next step drops.
constexpr int kSketchDigits254 = 4;
// The row digits fold_rows_khatri_rao_i8 takes, from u, v [d][ell] (u[a * ell + s]).
void split_khatri_rao_rows_i8(const double* u, const double* v, int64_t d, int64_t ell, FoldI8Rows& out,
                              hipStream_t stream);
void fold_rows_khatri_rao_i8(const FoldI8Rows& q, const int8_t* Yd, int64_t P, const double* yscale, int64_t n,
                             const double* cscale, double* y, int64_t ldy, double* ynorm2, unsigned long long* bad,
                             hipStream_t stream);

}  // namespace mpisdf
