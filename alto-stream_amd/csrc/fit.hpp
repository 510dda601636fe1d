// cpd_fit_alto (include/cpd.hpp:193-275) on the device, shared by the
// single and the grouped CP-ALS.
#pragma once
#include <math.h>

#include "dense.hpp"

namespace alto {

// inner = sum over blocks/columns of u*M; ||M||^2 exact and diagonal-only.
// Writes (fit, exact fit) to fitd[2k], fitd[2k+1] with k = *cnt (then
// incremented) when cnt != nullptr, else k = slot.
__device__ __forceinline__ void k_fit_body(const double *__restrict__ grams, int nmode, int R,
                                           const double *__restrict__ lambda,
                                           const double *__restrict__ partials, int nblk,
                                           const double *__restrict__ normX, int fit_mode,
                                           double *__restrict__ fitd, int *__restrict__ cnt, int slot) {
    extern __shared__ double sh[];  // R*R + R
    double *tg = sh, *ipc = sh + R * R;
    const int t = threadIdx.x, RR = R * R;
    for (int e = t; e < RR; e += blockDim.x) {
        int i = e / R, j = e % R;
        double v = lambda[i] * lambda[j];
        for (int m = 0; m < nmode; ++m) v *= grams[(size_t)m * RR + e];
        tg[e] = v;
    }
    for (int c = t >> 6; c < R; c += blockDim.x >> 6) {
        double s = 0.0;
        const int lane = t & 63;
        for (int b = lane; b < nblk; b += 64) s += partials[(size_t)b * 3 * R + 2 * R + c];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        if (lane == 0) ipc[c] = s;
    }
    __syncthreads();
    if (t == 0) {
        double inner = 0.0;
        for (int c = 0; c < R; ++c) inner += ipc[c];
        double nx = *normX;
        double ue = 0.0, ud = 0.0;
        for (int e = 0; e < RR; ++e) ue += tg[e];
        for (int i = 0; i < R; ++i) ud += tg[i * R + i];
        double f[2];
        double nu[2] = {fabs(ue), fabs(ud)};
        for (int k = 0; k < 2; ++k) {
            double r = nx + nu[k] - 2.0 * inner;
            if (r > 0.0) r = sqrt(r);
            f[k] = r / sqrt(nx);
        }
        const int k = cnt ? *cnt : slot;
        fitd[2 * k] = fit_mode == ALTO_FIT_EXACT ? f[0] : f[1];
        fitd[2 * k + 1] = f[0];
        if (cnt) *cnt = k + 1;
    }
}

}  // namespace alto
