// Device helpers shared by the MTTKRP kernels.
#pragma once
#include <type_traits>

#include "common.hpp"

namespace alto {

// Sum of a row that this range owns (output-ordered schedule): added to the
// output, or stored when the output held zeros on entry, which saves the
// dependent load of a read-modify-write.
template <typename OT, typename CT>
__device__ __forceinline__ void flush_owned(OT *p, CT v, bool zeroed) {
    if (zeroed) *p = (OT)v;
    else *p += (OT)v;
}

// Workgroup b of nb is dispatched to XCD b mod 8; this maps it to a logical
// index so that each XCD walks one contiguous slice (its L2 is shared by the
// neighbouring ranges).
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nb) {
    const unsigned q = nb / 8, r = nb % 8, x = b % 8, slot = b / 8;
    return x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

// Compile-time loop: f(integral_constant<int, I>) for I in [B, E).
template <int B, int E, typename F> __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

}  // namespace alto
