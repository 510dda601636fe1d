// Hot-row MTTKRP instances: FP32 factors.
#define ALTO_HOT_INSTANTIATE
#include "mttkrp_hot.hpp"

namespace alto {
template int hot_rank<float, float, float>(int, const HotLaunch &, hipStream_t);
template int hot_rank<float, float, double>(int, const HotLaunch &, hipStream_t);
}  // namespace alto
