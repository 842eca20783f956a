/*
 * mp_isdf_amd: C ABI of the MI355X (gfx950) anchor-selection library.
 *
 * Every entry point returns a status (0 = ok; 1 invalid argument, 2 overflow,
 * 3 runtime error, 4 logic error, 5 no HIP device) and leaves a thread-local
 * message in mpisdf_last_error().  The pybind11 module ``mp_isdf`` maps these
 * onto ValueError / OverflowError / RuntimeError exactly as the reference's
 * pybind11 exception translation does.  Host pointers are plain C arrays in
 * row-major order; ``*_device`` functions take device pointers and a
 * hipStream_t passed as ``void*`` (NULL = default stream).
 *
 * Each declaration names the reference interface it replaces (file:line in
 * mikovtun/mp_isdf).
 */
#ifndef MP_ISDF_AMD_H
#define MP_ISDF_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpisdf_context mpisdf_context;

/* Message of the last failed call on this thread. */
const char* mpisdf_last_error(void);

/* Number of HIP devices visible (0 when none).  Replaces the reference's
 * compile-time ``mp_isdf.cuda_enabled`` flag (cpp/src/ao_eval_pybind.cpp:73-77)
 * with a runtime query. */
int mpisdf_device_count(int* count);

/* ---- Basis context: LibcintContext(atm, bas, env) ---------------------- */
/* cpp/src/ao_eval_pybind.cpp:33-46, :88-90 -> cpp/src/ao_eval.cpp:78-158.
 * atm[natm*6], bas[nbas*8] (int32), env[nenv] (float64) are copied.
 * cart != 0 selects cartesian shells (no reference counterpart). */
int mpisdf_context_create(const int32_t* atm, int64_t natm, const int32_t* bas, int64_t nbas,
                          const double* env, int64_t nenv, int cart, mpisdf_context** out);
int mpisdf_context_destroy(mpisdf_context* ctx);
/* The randomized selectors keep their sketch storage (up to ell * n * 8
 * bytes) in the context between calls, so that repeated selections do not
 * free and reallocate it; this gives it back to the device (no reference
 * counterpart: the reference allocates per call). */
int mpisdf_context_release_workspace(mpisdf_context* ctx);
/* LibcintContext.nao (cpp/src/ao_eval_pybind.cpp:91). */
int mpisdf_context_nao(const mpisdf_context* ctx, int64_t* nao);

/* Analytic AO overlap s[nao*nao] (host array) of the context's functions:
 * replaces PySCF's mol.intor("int1e_ovlp") in the reference's overlap check
 * (python/isdf.py:665-671).  Obara-Saika cartesian overlaps, contracted and
 * transformed exactly as the AO evaluation does. */
int mpisdf_overlap(const mpisdf_context* ctx, double* s);

/* ---- AO evaluation --------------------------------------------------------- */
/* LibcintContext.eval_ao / eval_ao_gradient (cpp/src/ao_eval_pybind.cpp:48-64,
 * :92-101 -> cpp/src/ao_eval.cpp:274-310).  Host buffers, FP64.
 * gradient == 0: out[n][nao]; gradient != 0: out[3][n][nao]. */
int mpisdf_eval_ao(const mpisdf_context* ctx, const double* coords, int64_t n, int gradient,
                   double* out);

/* Device form of kernel ``isdf_eval_ao`` (cpp/src/ao_eval.cpp:205-250).
 * mode 0 values (1 comp), 1 gradient (3), 2 values + gradient (4);
 * dtype 0 f64, 1 f32, 2 bf16.  out[comp * cstride + g * ld + ao]. */
int mpisdf_ao_eval_device(const mpisdf_context* ctx, const double* xyz, int64_t n, int mode,
                          int dtype, void* out, int64_t ld, int64_t cstride, void* stream);

/* ---- Axis features ----------------------------------------------------------- */
/* AnchorSelection.evaluate_axis (cpp/src/anchor_select_pybind.cpp:211-222 ->
 * cpp/include/anchor_select_impl.hpp:11-81).  axes: 0 AO, 1 dX, 2 dY, 3 dZ,
 * 4 occupied (AxisType order, cpp/include/enums.hpp:8-14).  mo_coeff is the
 * row-major (nao, nao) matrix or NULL.  Call with out == NULL to get *width. */
int mpisdf_evaluate_axis(const mpisdf_context* ctx, const double* points, int64_t n,
                         const int32_t* axes, int32_t naxes, const double* mo_coeff, int64_t nocc,
                         double* out, int64_t* width);

/* ---- Anchor selection ---------------------------------------------------------- */
/* AnchorSelection.select (cpp/src/anchor_select_pybind.cpp:162-210 ->
 * cpp/src/anchor_select.cpp:79-173).  points[n*3], weights[n] on the host.
 * algorithm: 0 QRColumnPivoting (qrcp_dense), 1 OnlineQRColumnPivoting
 * (qrcp_online), 2 RandomizedQuadratureKhatriRao (rand_sketch); 3 (CVT)
 * returns status 1 (see mpisdf_select_cvt); 4 and 5 (density sketches)
 * need mpisdf_select_importance.
 * precision (rand_sketch only): 0 fp64, 1 fp32 (rejected), 2 bf16, 3 fp8 e4m3
 * (two-term phi), 4 split bf16 (hi.hi + lo.hi + hi.lo, sketch digits in four
 * planes), 5 fp64i8 (the FP64 sketch on four base-254 digit planes with the
 * INT8 fold; at most 16384 sketch rows).
 * axis_weights: NULL or naxis_weights == naxes finite positive factors.
 * debug: bit 0 prints the reference's progress lines; bit 1 (value 2) runs
 * the dense audit of rand_sketch and density-sketch pivots
 * (anchor_select_randomized_id.cpp:19-88: projection residuals of the
 * selected and of equally many dense QRCP pivots on the actual weighted pair
 * matrix, printed to stderr; limited to 2 million packed pair-matrix
 * elements, else status 1); other algorithms ignore it, as the reference does.
 * Writes original grid indices in pivot order to out[0..*count).
 * diagnostics: NULL or double[28] (see mp_isdf_amd/csrc/select.hpp; slots
 * 24-27 hold the audit residuals, -1 when no audit ran). */
int mpisdf_select(const mpisdf_context* ctx, const double* points, const double* weights, int64_t n,
                  const int32_t* axes, int32_t naxes, const double* mo_coeff, int64_t nocc,
                  const double* axis_weights, int64_t naxis_weights, double tolerance,
                  int64_t max_rank, int32_t algorithm, int64_t batch_size, int32_t debug,
                  int64_t sketch_size, int64_t oversampling, uint64_t seed, int32_t precision,
                  int64_t* out, int64_t capacity, int64_t* count, double* diagnostics);

/* AnchorSelection.select with point_importance and ao_screening_threshold
 * (cpp/src/anchor_select_pybind.cpp:186-209 -> cpp/src/anchor_select_randomized_id.cpp:92-374):
 * algorithm 4 RandomizedAODensityKhatriRao, 5 RandomizedAODensityNormalizedKhatriRao
 * (AO axis only).  point_importance[n] (finite, nonnegative) replaces |w| as
 * the sketch column weight; points with zero weight or zero importance, and
 * points whose AO squared norm is <= ao_screening_threshold, are not
 * candidates; algorithm 5 divides each pair column by that squared norm.
 * Other algorithms ignore point_importance, as the reference does. */
int mpisdf_select_importance(const mpisdf_context* ctx, const double* points, const double* weights, int64_t n,
                             const int32_t* axes, int32_t naxes, const double* mo_coeff, int64_t nocc,
                             const double* axis_weights, int64_t naxis_weights, double tolerance,
                             int64_t max_rank, int32_t algorithm, int64_t batch_size, int32_t debug,
                             int64_t sketch_size, int64_t oversampling, uint64_t seed, int32_t precision,
                             const double* point_importance, double ao_screening_threshold,
                             int64_t* out, int64_t capacity, int64_t* count, double* diagnostics);

/* AnchorSelection.select with algorithm=CentroidalVoronoi
 * (cpp/src/anchor_select_pybind.cpp:162-210 -> cpp/src/anchor_select_cvt.cpp:172-204):
 * points[n*3], weights[n], cvt_weights[n] on the host (finite; cvt_weights
 * nonnegative; a point's mass is cvt_weights where weights != 0).  Writes the
 * rank selected original indices, sorted and distinct, to out. */
int mpisdf_select_cvt(const double* points, const double* weights, const double* cvt_weights, int64_t n,
                      int64_t rank, uint64_t seed, int64_t max_iter, double tol, int64_t* out, int64_t capacity,
                      int64_t* count);

/* AO-only robust blockwise random pivoting with the optional interpolative
 * decomposition: mp_isdf.rbrp_ao_cpu (cpp/src/anchor_select_pybind.cpp:95-115 ->
 * cpp/src/anchor_select_rbrp.cpp:186-318), with the pair-row products on the
 * GPU.  ao[n*nao] row-major (nonzero-weight grid rows), weights[n] nonzero.
 * tau_block NaN selects 1/block_size.  Outputs (capacity >= the rank cap
 * min(nao(nao+1)/2, n, max_rank or n)): indices (rows, selection order),
 * interp_vecs[k*n] and transform[k*nao*nao] when compute_id (may be NULL),
 * history[4 per block] (may be NULL; every block adds at least one anchor,
 * so 4*capacity doubles always suffice), *nblocks, scalars[3] = relative
 * residual, condition, retained rank; *status 0 rank_cap, 1 tolerance,
 * 2 no_candidates, 3 numerical_rank, 4 roundoff_limited.  batch_size must be
 * positive, as in the reference; the pair rows are formed once in one device
 * buffer (n * nao(nao+1)/2 doubles), so it does not bound memory here. */
int mpisdf_rbrp_ao(const double* ao, const double* weights, int64_t n, int64_t nao, int64_t batch_size,
                   int64_t block_size, double tolerance, int64_t max_rank, uint64_t seed, double tau_block,
                   double rank_rtol, int64_t refresh_every, double svd_rtol, int compute_id, int64_t* indices,
                   int64_t capacity, int64_t* count, double* interp_vecs, double* transform, double* history,
                   int64_t* nblocks, double* scalars, int32_t* status);
/* The same on grid points: mp_isdf.rbrp_ao_device
 * (cpp/src/anchor_select_pybind.cpp:117-140 -> cpp/src/anchor_select_rbrp_device.cpp:327-491);
 * AO values come from the context's evaluation kernel. */
int mpisdf_rbrp_points(const mpisdf_context* ctx, const double* points, const double* weights, int64_t n,
                       int64_t batch_size, int64_t block_size, double tolerance, int64_t max_rank, uint64_t seed,
                       double tau_block, double rank_rtol, int64_t refresh_every, double svd_rtol, int compute_id,
                       int64_t* indices, int64_t capacity, int64_t* count, double* interp_vecs, double* transform,
                       double* history, int64_t* nblocks, double* scalars, int32_t* status);

/* ---- Khatri-Rao sketch stages (device pointers) ---------------------------- */
/* Gaussian factors u, v [d][ell] row-major from std::mt19937_64(seed) and
 * (seed ^ 0x9e3779b97f4a7c15) with std::normal_distribution<double>(0, 1),
 * filled in index order (cpp/src/anchor_select_randomized_id.cpp:269-279).
 * Host buffers. */
int mpisdf_omega(int64_t d, int64_t ell, uint64_t seed, double* u, double* v);

/* FP64 sketch (the reference's id_factored_sketch, randomized_id.cpp:299-309):
 * y[g*ldy + s] = ell^-1/2 sw[g] (sum_a phi[g*d+a] u[a*ell+s]) (sum_a phi[g*d+a] v[a*ell+s]);
 * *bad (device, may be NULL) counts non-finite outputs. */
int mpisdf_kr_sketch_f64_device(const double* phi, int64_t count, int64_t d, const double* u, const double* v,
                                int64_t ell, const double* sw, double* y, int64_t ldy,
                                unsigned long long* bad, void* stream);

/* Reduced-precision Omega operands from u/v, in segments of ldo values (ldo
 * >= d, a multiple of 16, zero-padded past d): dtype 2 bf16 ut[s*ldo + a];
 * 3 fp8 (e4m3 times 32) [u | u], row pitch 2 ldo; 4 split bf16 [hi | hi | lo]
 * (u = hi + lo), row pitch 3 ldo. */
int mpisdf_sketch_prepare_omega(const double* u, const double* v, int64_t d, int64_t ell, int64_t ldo, int dtype,
                                void* ut, void* vt, void* stream);
/* Split-BF16 rows out[g*3*ldo + {0, ldo, 2 ldo} + a] = {hi, lo, hi} of
 * phi[g*ld + a] (FP64), so that mpisdf_kr_sketch_device (dtype 2, d = 3 ldo)
 * with the dtype-4 Omega sums hi.hi + lo.hi + hi.lo. */
int mpisdf_sketch_split_bf16x3(const double* phi, int64_t ld, int64_t count, int64_t d, void* out, int64_t ldo,
                               void* stream);
/* BF16 point scales scale[g] = sw[g] / sqrt(ell). */
int mpisdf_sketch_scales_bf16(const double* sw, int64_t count, int64_t ell, float* scale, void* stream);
/* FP8 rows [hi | lo] (row pitch 2 ldo; hi = e4m3(x), lo = e4m3(x - hi) on the
 * point's one power-of-two scale); scale[g] absorbs the scales.  Entries past
 * d of each segment are not written. */
int mpisdf_sketch_quantize_fp8(const float* phi, int64_t ld, int64_t count, int64_t d, const double* sw,
                               int64_t ell, uint8_t* out, int64_t ldo, float* scale, void* stream);
/* MFMA sketch (dtype 2 bf16, 3 fp8): FP32 y[g*ldy + s], ynorm2[g] += ||y_g||^2;
 * d is the K length (all segments of the operands' rows). */
int mpisdf_kr_sketch_device(int dtype, const void* ut, const void* vt, int64_t ell, int64_t d, const void* phi,
                            int64_t ldphi, int64_t count, const float* scale, float* y, int64_t ldy,
                            double* ynorm2, unsigned long long* bad, void* stream);
/* Allocation-free form of mpisdf_kr_sketch_device, with the same arguments
 * and results bit for bit.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  It calls no hipMalloc/hipFree and does not
 * synchronise the stream: the per-Omega-row-block partial sums of ||y_g||^2
 * live in a static device array of 2^22 doubles (32 MiB) that each GPU holds
 * until the process ends, so count * ceil(ell / 256) must not exceed 2^22
 * (status 1 otherwise).  All calls on one GPU share that array: calls that
 * overlap in time on one GPU race. */
int mpisdf_kr_sketch_device_noalloc(int dtype, const void* ut, const void* vt, int64_t ell, int64_t d,
                                    const void* phi, int64_t ldphi, int64_t count, const float* scale, float* y,
                                    int64_t ldy, double* ynorm2, unsigned long long* bad, void* stream);

/* ---- QRCP fold (device pointers) --------------------------------------------- */
/* Block update of the column-pivoted QR's residual norms (replaces the
 * norm downdating inside LAPACK/MAGMA dgeqp3, cpp/src/anchor_select_qrcp.cpp:36-90):
 * W = Qb^T Y with Qb[m x b] column-major FP64 and Y[m x n] (ldy);
 * y_f32 = 0: FP64 Y, FP64 MFMA; 1: FP32 Y, FP64 MFMA (exact products);
 * 2: FP32 Y, split-BF16 MFMA with FP64 chunk accumulation;
 * 3 / 4: FP32 Y, INT8 MFMA on three / four base-128 digit planes of it
 * (~1e-10 relative; m < 65536); 5: FP64 Y, INT8 MFMA on four base-254
 * digit planes of it (m <= 16384).
 * proj2[j] += sum_t W[t,j]^2 and, if energy != NULL, energy[t] += sum_j W[t,j]^2.
 * W is never stored. */
int mpisdf_qrcp_fold_device(int64_t b, int64_t n, int64_t m, const double* Qb, const void* Y, int y_f32,
                            int64_t ldy, double* proj2, double* energy, void* stream);

/* In place: FP32 column j of y (pitch 4 P bytes = P floats, P a multiple of
 * 64 >= m, m <= 36864) becomes planes (3 or 4) signed base-128 digits per
 * entry, y_hat_j[i] = scale[j] sum_{r<planes} d_r[i] 128^-(r+1) with scale[j]
 * a power of two >= 2 ||y_j|| (each entry within 2^-(7 planes + 1) scale[j]
 * of y), digit r of entry i at byte 64 planes (i / 64) + 64 r + i % 64 of the
 * column (zero past m and in the column's last (4 - planes) P bytes);
 * ynorm2[j] holds ||y_j||^2 on entry and
 * ||y_hat_j||^2 on return.  The digit form of the sketch that the reduced-precision selection
 * factorises (no reference counterpart: the reference factorises an FP64
 * sketch, cpp/src/anchor_select_randomized_id.cpp:320-326). */
int mpisdf_digitize_columns_device(float* y, int64_t m, int64_t n, int64_t P, int32_t planes, double* ynorm2,
                                   double* scale, void* stream);
/* The same in place of FP64 column j (pitch 8 P bytes = P doubles, P a
 * multiple of 64, m <= P <= 16384): four signed base-254 planes (|d| <= 127),
 * y_hat_j[i] = scale[j] sum_r d_r[i] 254^-(r+1), each entry within
 * scale[j] 254^-4 / 2 of y, digit r of entry i at byte 256 (i / 64) + 64 r +
 * i % 64 of the column (zero past m and in its last 4 P bytes).  The digit
 * form of the FP64 sketch that precision fp64i8 factorises (no reference
 * counterpart). */
int mpisdf_digitize_columns_f64_device(double* y, int64_t m, int64_t n, int64_t P, double* ynorm2, double* scale,
                                       void* stream);

/* Column-pivoted QR pivots of a device-resident sketch (replaces qrcp_cpu /
 * qrcp_magma, cpp/src/anchor_select_qrcp.cpp:36-90, as called from
 * cpp/src/anchor_select_randomized_id.cpp:320-326): dgeqp3's greedy pivot
 * order (max residual norm, lowest index on ties) of Y[m x n] held as
 * ytype 0 FP64 or 1 FP32 column-major (ldy elements; ynorm2 and yscale may
 * be NULL), or 4 / 5 digits (three / four planes) from
 * mpisdf_digitize_columns_device (ldy = P, the column pitch in floats; ynorm2
 * and yscale as it wrote them; folds on INT8 MFMA), or 6 (four base-254
 * planes) from mpisdf_digitize_columns_f64_device (ldy = P in doubles).
 * Stops before a pivot with |R_ii| <= stop_rel |R_00| or after kmax pivots.
 * Writes 0-based pivots and |R_ii| (host arrays of kmax entries; rdiag may
 * be NULL) and their count. */
int mpisdf_qrcp_device(const void* y, int ytype, int64_t m, int64_t n, int64_t ldy, const double* ynorm2,
                       const double* yscale, int64_t kmax, double stop_rel, int64_t* pivots, double* rdiag,
                       int64_t* count, void* stream);

/* mpisdf_qrcp_device for columns the caller has reordered (no reference
 * counterpart: the selection's own column order for large grids, exposed for
 * tests and callers that order their columns).  tie_rank: NULL or a device
 * permutation of 0 .. n-1; ties between equal norms go to the lowest rank
 * (the column's position before reordering) instead of the lowest column.
 * dormant != 0: columns whose ||y_j||^2 stays far below the certification
 * bound are not folded (their keys stay ||y_j||^2, an upper bound), which
 * pays when the columns come in roughly descending norm; the pivots are the
 * same.  1024 candidate columns per block.  folded_columns (may be NULL):
 * the columns folded by the end (n without dormant columns). */
int mpisdf_qrcp_device_ordered(const void* y, int ytype, int64_t m, int64_t n, int64_t ldy, const double* ynorm2,
                               const double* yscale, const int64_t* tie_rank, int32_t dormant, int64_t kmax,
                               double stop_rel, int64_t* pivots, double* rdiag, int64_t* count,
                               int64_t* folded_columns, void* stream);

/* ---- ISDF fit on selected anchors ------------------------------------------- */
/* Replaces the driver refit and pair-space error of the reference
 * (python/isdf.py:238-263 _physical_pair_refit, :176-191 pair error) with the
 * normal-equations form on the GPU: G2 = (Phi_a Phi_a^T)^2 = L L^T (FP64),
 * theta_g = L^-T L^-1 (Phi_a phi_g)^2 / w_anchor, and
 * error_g^2 = ||phi_g||^4 - ||L^-1 (Phi_a phi_g)^2||^2.
 * Host arrays: anchor_xyz[k*3], anchor_w[k], xyz[n*3], w[n] (grid points with
 * nonzero weight).  theta (optional) receives row-major [k][n]; tsum
 * (optional) sum_g theta[:, g] w_g.  results[10]: 0 l2 error, 1 relative error,
 * 2 condition estimate of G2, 3 max diagonal of G2, 4 seconds (factor),
 * 5 seconds (grid pass), 6 summed HIP-event seconds of the grid pass's FP64
 * TN products (S, Z = L^-1 S and, when theta is formed, theta = L^-T Z),
 * 7 their algorithmic flops (2 k d + k (k+1) per point, + k (k+1) with
 * theta), 8 retained rank, 9 1 when the anchor pairs were numerically
 * dependent and the pseudo-inverse of G2 over its eigenvalues above 1e-14 of
 * the largest replaced L^-1 (the reference's truncated-SVD refit, which keeps
 * singular values above 1e-12 of the largest; k <= 2048), else 0. */
int mpisdf_isdf_fit(const mpisdf_context* ctx, const double* anchor_xyz, const double* anchor_w, int64_t k,
                    const double* xyz, const double* w, int64_t n, int64_t batch, double* theta, double* tsum,
                    double* results);
/* Same with device-resident grid (xyz_dev[n*3], w_dev[n]) and anchors
 * (anchor_xyz_dev[k*3]); anchor_w stays on the host.  tsum (host, k entries,
 * optional) = sum_g theta[:, g] w_g, formed as L^-T (sum_g w_g z_g) / w_anchor
 * with z_g = L^-1 (Phi_a phi_g)^2: theta itself is not formed, and the grid
 * pass runs only the S and Z products (a consumer that needs theta forms
 * L^-T Z per batch, or calls mpisdf_isdf_fit). */
int mpisdf_isdf_fit_device(const mpisdf_context* ctx, const double* anchor_xyz_dev, const double* anchor_w, int64_t k,
                           const double* xyz_dev, const double* w_dev, int64_t n, int64_t batch, double* tsum,
                           double* results, void* stream);

/* The fit of mpisdf_isdf_fit_device with a choice of arithmetic for the
 * grid pass's Z = L^-1 (Phi_a phi_g)^2 product, the same refit and error as
 * python/isdf.py:238-263 and :176-191.  form 0: FP64 MFMA (as
 * mpisdf_isdf_fit_device).  form 1 (no reference counterpart): each column
 * s_g and each row of L^-1 as five base-254 digit planes under a power of two
 * (within ~4.8e-13 of its largest entry); ||L^-1 s_g||^2 comes from the 15
 * digit-pair products on INT8 MFMA, exact INT32 sums combined in FP64,
 * without forming Z (k <= 16384; a pseudo-inverse fit, results[9] = 1, keeps
 * FP64 products).  results[12]: 0-9 as for
 * mpisdf_isdf_fit, 10 HIP-event seconds of the Z product alone, 11 its
 * launches. */
int mpisdf_isdf_fit_device_form(const mpisdf_context* ctx, const double* anchor_xyz_dev, const double* anchor_w,
                                int64_t k, const double* xyz_dev, const double* w_dev, int64_t n, int64_t batch,
                                int32_t form, double* tsum, double* results, void* stream);

/* Pair-space error of given interpolation vectors theta (host, row-major
 * [k][n]) over the grid points xyz[n*3], w[n] (host), as the reference's
 * _stream_pair_space_error_norms (python/isdf.py:176-191) evaluates it, in
 * O(k) per point: ||phi_g||^4 - 2 c.s_g + c.(G2 c), c = anchor_w o theta[:, g],
 * s_g = (Phi_a phi_g)^2.  results[2]: 0 l2 error, 1 relative error. */
int mpisdf_pair_space_error(const mpisdf_context* ctx, const double* anchor_xyz, const double* anchor_w, int64_t k,
                            const double* xyz, const double* w, int64_t n, const double* theta, int64_t batch,
                            double* results);
/* Grid-quadrature AO overlap out[nao*nao] = sum_g w_g phi_g phi_g^T (host
 * arrays; the AO values never leave the GPU), the quadrature side of the
 * reference's overlap check (python/isdf.py:665-671). */
int mpisdf_quadrature_overlap(const mpisdf_context* ctx, const double* xyz, const double* w, int64_t n, int64_t batch,
                              double* out);

/* ---- Adjoint (no reference counterpart: the reference has no nuclear
 * derivatives; it provides the dx/dy/dz AO derivatives the adjoint is checked
 * against, cpp/src/ao_eval.cpp:215-244) ----------------------------------- */
/* gout[c*nao + mu] += sum_p abar[p*ldab + mu] * d phi_mu(r_p)/d r_c for the
 * device points xyz[n*3]; the AO gradients are recomputed, never stored. */
int mpisdf_ao_adjoint_device(const mpisdf_context* ctx, const double* xyz, int64_t n, const double* abar,
                             int64_t ldab, double* gout, void* stream);
/* Host-array form: xyz[n*3], abar[n*nao] row-major, gout[3*nao] overwritten. */
int mpisdf_ao_adjoint_host(const mpisdf_context* ctx, const double* xyz, int64_t n, const double* abar, double* gout);
/* Pair-reconstruction loss L = sum_g |w_g| ||phi_g phi_g^T - sum_k c_gk phi_k phi_k^T||^2
 * of the FP64 fit on fixed anchor and grid points (host arrays) and its
 * gradient with respect to the nuclear coordinates of the context:
 * *loss, grad_ao[3*nao] (optional), grad_atoms[natm*3] (optional). */
int mpisdf_isdf_fit_gradient(const mpisdf_context* ctx, const double* anchor_xyz, const double* anchor_w, int64_t k,
                             const double* xyz, const double* w, int64_t n, int64_t batch, double* loss,
                             double* grad_ao, double* grad_atoms);

/* Coulomb energy of the fitted density on fixed anchor and grid points (host
 * arrays) and its gradient with respect to the nuclear coordinates of the
 * context: E_J = tr(J dm) / 2 for the ISDF J of evaluate_isdf_jk
 * (python/isdf.py:130-151) with the grid metric of the FP64 refit
 * (isdf.py:48-67, :238-263, :674-689), evaluated as u^T W K W u / 2 with
 * u = C^T rho the fitted density on the grid (no theta, no metric formed).
 * dm[nao*nao] symmetric; w nonzero.  *energy, grad_ao[3*nao] and
 * grad_atoms[natm*3] (each optional; a gradient needs independent anchor
 * pairs).  No counterpart in the reference, which has no nuclear gradients. */
int mpisdf_coulomb_energy_gradient(const mpisdf_context* ctx, const double* anchor_xyz, int64_t k, const double* xyz,
                                   const double* w, int64_t n, const double* dm, int64_t batch, double* energy,
                                   double* grad_ao, double* grad_atoms);

/* ---- Coulomb metric and J/K ---------------------------------------------- */
/* Grid Coulomb metric, python/isdf.py:48-67 (grid_coulomb_metric) and :76-92
 * (grid_cross_metric):  metric[a*kr + b] = sum_ij left[a][i] w_i K_ij w_j
 * right[b][j] over n grid points, K_ij = 1/|r_i - r_j| (0 for distinct
 * coincident points), K_ii = (32/15) pi^2 (3/(4 pi))^(5/3) |w_i|^(-1/3).
 * Host arrays: left[kl*n], right[kr*n] (interpolation vectors, not weight
 * scaled), xyz[n*3], w[n] (finite, nonzero), metric[kl*kr] overwritten.
 * symmetrize != 0 (kl == kr) returns (M + M^T)/2 as grid_coulomb_metric does;
 * with right == left this is grid_coulomb_metric(left, xyz, w). */
int mpisdf_grid_coulomb_metric(const double* left, int64_t kl, const double* right, int64_t kr, const double* xyz,
                               const double* w, int64_t n, int symmetrize, double* metric);
/* Device form: left/right rows at stride ldl/ldr (>= n; multiples of 4 take
 * the vector load path), device xyz, w and metric.  block_rows (a multiple
 * of 128, or 0 for the automatic size) sets how many kernel rows one block
 * materialises in device memory (block_rows * n doubles), the role of the
 * reference's metric_batch_size.  Values are not checked; a zero weight makes
 * the self term infinite.  Returns after the stream has drained. */
int mpisdf_grid_coulomb_metric_device(const double* left, int64_t kl, int64_t ldl, const double* right, int64_t kr,
                                      int64_t ldr, const double* xyz, const double* w, int64_t n, int symmetrize,
                                      double* metric, int64_t block_rows, void* stream);
/* ISDF J and K, python/isdf.py:130-151 (evaluate_isdf_jk): x[k*nao] anchor AO
 * values, z[k*k] the anchor metric (both anchor weights included), dm[nao*nao]
 * a finite symmetric density; P = x dm x^T, J = x^T diag(z diag(P)) x,
 * K = x^T (z o P) x, both symmetrised into j[nao*nao], kmat[nao*nao].
 * Invalid dm raises ValueError with the reference's messages. */
int mpisdf_isdf_jk(const double* x, int64_t k, int64_t nao, const double* z, const double* dm, double* j,
                   double* kmat);
/* Device form (no value checks; returns after the stream has drained). */
int mpisdf_isdf_jk_device(const double* x, int64_t k, int64_t nao, const double* z, const double* dm, double* j,
                          double* kmat, void* stream);

/* ---- Grid front end (no reference counterpart: the reference takes its grid
 * from PySCF's dft.gen_grid.Grids, python/isdf.py:488-496) ------------------ */
/* Becke-partitioned atom-centred grid with PySCF's defaults: Treutler radial
 * nodes, Lebedev angular rules of PySCF's order per level and period,
 * NWChem pruning.  Levels 0..3 (and higher levels whose rules are
 * tabulated; status 1 otherwise).  Call with xyz == NULL to get *count; then
 * xyz[3*count], w[count].  device != 0 runs the partition on the GPU when one
 * is present; the weights equal the host loop's bit for bit. */
/* The npoints-point Lebedev-Laikov rule (PySCF's dft.gen_grid angular
 * grids): xyz[3*npoints] unit vectors, weights[npoints] summing to 1.
 * Status 1 when no rule of that size is tabulated. */
int mpisdf_lebedev(int32_t npoints, double* xyz, double* weights);
int mpisdf_becke_grid(const int32_t* charges, const double* coords, int64_t natm, int32_t level, int32_t device,
                      double* xyz, double* weights, int64_t capacity, int64_t* count);

/* ---- Integral-direct Cholesky Coulomb metric ---------------------------- */

/* Pivoted Cholesky of the packed AO-pair Coulomb matrix V[(xy),(zw)] =
 * (xy|zw) over x >= y (index x(x+1)/2 + y), columns computed
 * integral-direct per pivot on the GPU: mp_isdf.cholesky_eri
 * (cpp/src/coulomb_metric_pybind.cpp:25-45 -> cpp/src/cholesky_eri_greedy.cpp:53-175,
 * cpp/src/cholesky_eri_uniform.cpp:4-10).  algorithm 0 unblocked_greedy
 * (largest residual, first index on ties), 1 unblocked_uniform (uniform
 * among residuals above tolerance * the largest initial diagonal, drawn from
 * std::mt19937_64(seed) as cpp/include/cholesky_pivot.hpp:16-36 draws);
 * tolerance relative to the largest initial diagonal, max_rank 0 for no
 * cap; s, p and d shells.  The handle holds the factors on the device until
 * mpisdf_eri_cholesky_destroy. */
typedef struct mpisdf_eri_cholesky mpisdf_eri_cholesky;
int mpisdf_cholesky_eri(const mpisdf_context* ctx, double tolerance, int64_t max_rank, int32_t algorithm,
                        uint64_t seed, int32_t debug, mpisdf_eri_cholesky** out);
/* rank and npair = nao(nao+1)/2 of a decomposition. */
int mpisdf_eri_cholesky_size(const mpisdf_eri_cholesky* h, int64_t* rank, int64_t* npair);
/* factors[rank*npair] (row-major, one factor per row), pivots[rank] (packed
 * pair indices in selection order), residual[npair]; any may be NULL. */
int mpisdf_eri_cholesky_copy(const mpisdf_eri_cholesky* h, double* factors, int64_t* pivots, double* residual);
int mpisdf_eri_cholesky_destroy(mpisdf_eri_cholesky* h);
/* out[nvec*nvec] = P P^T, P = packed(transform) F^T with transform[nvec][nao*nao]
 * row-major (packed folds T[v,y,x] into pair (x,y)): mp_isdf.coulomb_metric
 * (cpp/src/coulomb_metric_pybind.cpp:73-92 -> cpp/src/coulomb_metric.cpp:20-98),
 * without anchor weights.  batch_size must be positive (the reference's host
 * batching; the products here run whole on the GPU). */
int mpisdf_coulomb_metric(const mpisdf_context* ctx, const double* transform, int64_t nvec, int32_t algorithm,
                          double tolerance, int64_t max_rank, int64_t batch_size, int32_t debug, uint64_t seed,
                          double* out);

#ifdef __cplusplus
}
#endif

#endif /* MP_ISDF_AMD_H */
