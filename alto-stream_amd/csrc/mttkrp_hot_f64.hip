// Hot-row MTTKRP instances: FP64 factors.
#define ALTO_HOT_INSTANTIATE
#include "mttkrp_hot.hpp"

namespace alto {
template int hot_rank<double, double, double>(int, const HotLaunch &, hipStream_t);
}  // namespace alto
