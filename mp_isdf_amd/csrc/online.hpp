// Online (bounded-memory, multipass) QRCP selection on the GPU
// (cpp/src/anchor_select_online_qr.cpp:127-375).
#pragma once

#include "axis.hpp"
#include "select.hpp"

namespace mpisdf {

std::vector<int64_t> select_online(Context& ctx, const SelectArgs& a, const AxisSpec& spec,
                                   const std::vector<double>& scales, double* diag, hipStream_t s);

}  // namespace mpisdf
