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
/* LibcintContext.nao (cpp/src/ao_eval_pybind.cpp:91). */
int mpisdf_context_nao(const mpisdf_context* ctx, int64_t* nao);

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
 * (qrcp_online), 2 RandomizedQuadratureKhatriRao (rand_sketch); the other
 * reference values (CVT, density sketches) return status 1.
 * precision (rand_sketch only): 0 fp64, 1 fp32, 2 bf16, 3 fp8 e4m3.
 * axis_weights: NULL or naxis_weights == naxes finite positive factors.
 * Writes original grid indices in pivot order to out[0..*count).
 * diagnostics: NULL or double[16] (see mp_isdf_amd/csrc/select.hpp). */
int mpisdf_select(const mpisdf_context* ctx, const double* points, const double* weights, int64_t n,
                  const int32_t* axes, int32_t naxes, const double* mo_coeff, int64_t nocc,
                  const double* axis_weights, int64_t naxis_weights, double tolerance,
                  int64_t max_rank, int32_t algorithm, int64_t batch_size, int32_t debug,
                  int64_t sketch_size, int64_t oversampling, uint64_t seed, int32_t precision,
                  int64_t* out, int64_t capacity, int64_t* count, double* diagnostics);

#ifdef __cplusplus
}
#endif

#endif /* MP_ISDF_AMD_H */
