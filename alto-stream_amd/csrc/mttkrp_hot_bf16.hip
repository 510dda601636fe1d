// Hot-row MTTKRP instances: BF16-stored factors.
#define ALTO_HOT_INSTANTIATE
#include "mttkrp_hot.hpp"

namespace alto {
template int hot_rank<float, uint16_t, float>(int, const HotLaunch &, hipStream_t);
template int hot_rank<float, uint16_t, double>(int, const HotLaunch &, hipStream_t);
}  // namespace alto
