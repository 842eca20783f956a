#include "online.hpp"

namespace mpisdf {

std::vector<int64_t> select_online(Context&, const SelectArgs&, const AxisSpec&, const std::vector<double>&,
                                   double*, hipStream_t)
{
    throw std::logic_error("online QRCP is not available yet");
}

}  // namespace mpisdf
