// Reduced-precision Khatri-Rao sketch on MFMA (see sketch_mfma.hip).
#pragma once

#include "common.hpp"

namespace mpisdf {

// Omega entries are stored in FP8 as 32 * u (keeps |u| < 6 well inside e4m3).
constexpr float kOmegaFp8Scale = 32.0f;

// ut/vt[s * ldo + a] = cast(u/v[a * ell + s]) (times kOmegaFp8Scale for FP8),
// zero-padded for d <= a < ldo.
void prepare_omega_lowp(const double* u, const double* v, int64_t d, int64_t ell, int64_t ldo, DType dtype,
                        void* ut, void* vt, hipStream_t s);

// scale[g] = sw[g] / sqrt(ell) for the BF16 form.
void bf16_point_scales(const double* sw, int64_t count, int64_t ell, float* scale, hipStream_t s);

// FP8 rows with per-point power-of-two scales; scale[g] absorbs them, the
// omega scale and sw[g] / sqrt(ell).
void quantize_fp8_rows(const float* phi, int64_t ld, int64_t count, int64_t d, const double* sw, int64_t ell,
                       uint8_t* out, int64_t ldo, float* scale, hipStream_t s);

// out[g * ldo + a] = cast(in[g * d + a]) with zero padding (bf16 or f32).
void cast_rows(const double* in, int64_t count, int64_t d, int64_t ldo, DType dtype, void* out, hipStream_t s);

// Y[g * ldy + s] (FP32) and ynorm2[g] += ||y_g||^2 for g < count (per Omega
// row block partials in ``part``, added in block order); d must be a
// multiple of 8 (bf16) or 16 (fp8) -- callers zero-pad.
void kr_sketch_lowp(DType dtype, const void* ut, const void* vt, int64_t ell, int64_t d, const void* phi, int64_t ldphi,
                    int64_t count, const float* scale, float* y, int64_t ldy, double* ynorm2, unsigned long long* bad,
                    DeviceBuffer<double>& part, hipStream_t s);

}  // namespace mpisdf
