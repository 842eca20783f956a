// Reduced-precision Khatri-Rao sketch on MFMA (see sketch_mfma.hip).
#pragma once

#include "common.hpp"

namespace mpisdf {

// Omega entries are stored in FP8 as 32 * u (keeps |u| < 6 well inside e4m3).
constexpr float kOmegaFp8Scale = 32.0f;

// Operand layouts of the reduced-precision sketch forms along K, in
// segments of ldo values (ldo >= d, a multiple of 16; the row pitch is
// sketch_segments(form) * ldo):
//   kSketchBF16:   phi [x],            Omega [u]             (BF16)
//   kSketchFP8:    phi [hi | lo],      Omega [u | u]         (e4m3: hi = e4m3(x),
//                  lo = e4m3(x - hi) on the point's one power-of-two scale;
//                  Omega times kOmegaFp8Scale)
//   kSketchBF16x3: phi [hi | lo | hi], Omega [hi | hi | lo]  (BF16 splits
//                  x = hi + lo: the K sum is hi.hi + lo.hi + hi.lo, ~2^-16 of
//                  the FP64 product)
enum SketchForm : int { kSketchBF16 = 2, kSketchFP8 = 3, kSketchBF16x3 = 4 };
inline int sketch_segments(int form) { return form == kSketchBF16 ? 1 : form == kSketchFP8 ? 2 : 3; }

// ut/vt rows of the form's layout from u/v[a * ell + s], zero-padded for
// d <= a < ldo in every segment.
void prepare_omega_lowp(const double* u, const double* v, int64_t d, int64_t ell, int64_t ldo, int form,
                        void* ut, void* vt, hipStream_t s);

// scale[g] = sw[g] / sqrt(ell) for the BF16 form.
void bf16_point_scales(const double* sw, int64_t count, int64_t ell, float* scale, hipStream_t s);

// FP8 rows [hi | lo] (row pitch 2 ldo) with per-point power-of-two scales;
// scale[g] absorbs them, the omega scale and sw[g] / sqrt(ell).  Entries
// d <= a < ldo of each segment are left as they are (callers zero them once).
void quantize_fp8_rows(const float* phi, int64_t ld, int64_t count, int64_t d, const double* sw, int64_t ell,
                       uint8_t* out, int64_t ldo, float* scale, hipStream_t s);

// Split-BF16 rows [hi | lo | hi] (row pitch 3 ldo) of FP64 in[g * ld + a].
void split_rows_bf16x3(const double* in, int64_t ld, int64_t count, int64_t d, int64_t ldo, void* out, hipStream_t s);

// out[g * ldo + a] = cast(in[g * d + a]) with zero padding (bf16 or f32).
void cast_rows(const double* in, int64_t count, int64_t d, int64_t ldo, DType dtype, void* out, hipStream_t s);

// Y[g * ldy + s] (FP32) and ynorm2[g] += ||y_g||^2 for g < count (per Omega
// row block partials in ``part``, added in block order); d must be a
// multiple of 8 (bf16) or 16 (fp8) -- callers zero-pad.
void kr_sketch_lowp(DType dtype, const void* ut, const void* vt, int64_t ell, int64_t d, const void* phi, int64_t ldphi,
                    int64_t count, const float* scale, float* y, int64_t ldy, double* ynorm2, unsigned long long* bad,
                    DeviceBuffer<double>& part, hipStream_t s);
// The same with caller-provided partials (kr_sketch_lowp_partials(ell, count)
// doubles, every one written before it is read); allocates nothing.
void kr_sketch_lowp_into(DType dtype, const void* ut, const void* vt, int64_t ell, int64_t d, const void* phi,
                         int64_t ldphi, int64_t count, const float* scale, float* y, int64_t ldy, double* ynorm2,
                         unsigned long long* bad, double* part, hipStream_t s);
int64_t kr_sketch_lowp_partials(int64_t ell, int64_t count);

}  // namespace mpisdf
