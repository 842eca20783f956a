// Axis specification and device feature assembly
// (cpp/src/anchor_select.cpp:18-52, cpp/include/anchor_select_impl.hpp:11-81).
#pragma once

#include "context.hpp"
#include "kernels.hpp"

namespace mpisdf {

struct AxisSpec {
    int64_t width = 0;
    bool need_ao = false, need_grad = false, need_occ = false;
    std::vector<int> kinds;
    std::vector<int64_t> widths;
};

// Validates the axis list as AnchorSelection::axis_spec does.
AxisSpec axis_spec(const Context& ctx, const int32_t* axes, int naxes, const double* mo_coeff,
                   int64_t nocc);

// Device scratch reused across batches.
struct AxisWorkspace {
    DeviceBuffer<double> ao;   // [4][n][nao]
    DeviceBuffer<double> occ;  // [n][nocc]
    DeviceBuffer<double> mo;   // (nao, nao)
    bool mo_ready = false;     // mo holds this call's coefficients
};

// features[g * width + col] for ``n`` device points; scales: one per axis or
// nullptr for all ones.  ``mo_host`` is uploaded once into ws.mo.
void evaluate_axis_device(Context& ctx, const AxisSpec& spec, const double* xyz, int64_t n,
                          const double* scales, const double* mo_host, int64_t nocc,
                          double* features, AxisWorkspace& ws, hipStream_t stream);

}  // namespace mpisdf
