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

// a * b + c with one rounding, in the type of the operands
__device__ __forceinline__ float fused_madd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fused_madd(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Compile-time loop: f(integral_constant<int, I>) for I in [B, E).
template <int B, int E, typename F> __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}


// Sums boundary slots -- (row, R values) pairs in schedule order, rows
// non-decreasing, empty slots (row ~0) skipped -- as a fixed tree, so the
// result does not depend on scheduling: at each level, work item (chunk of
// kSlotChunk slots, column) adds the runs of equal rows of its chunk in slot
// order, writes the runs that start and end inside the chunk to the output
// (stored when the output held zeros, else added), and passes its first and
// last runs -- which may continue in the neighbouring chunks -- to the next
// level as two FP64 slots (a chunk of one run passes the row twice, the
// second with zero).  At the last level (one chunk) every run is complete.
// A hot output row spans thousands of workgroups, so one serial walk of its
// run would take milliseconds.
constexpr int kSlotChunk = 32;
constexpr uint32_t kNoRow = 0xFFFFFFFFu;

template <typename OT, int R, typename VT>
__device__ __forceinline__ void bnd_level_item(uint64_t t, const uint32_t *rows, const VT *vals, uint64_t n,
                                               uint32_t *orow, double *oval, OT *out, bool zeroed, bool last) {
    const uint64_t c = t / R;
    const int col = (int)(t % R);
    const uint64_t a = c * kSlotChunk;
    if (a >= n) return;
    const uint64_t b = a + kSlotChunk < n ? a + kSlotChunk : n;
    auto put = [&](uint32_t row, double v) {
        OT *p = out + (uint64_t)row * R + col;
        *p = zeroed ? (OT)v : (OT)((double)*p + v);
    };
    // the chunk's rows, then its values, as independent loads (one memory
    // round trip each), walked from registers
    uint32_t rr[kSlotChunk];
    double xv[kSlotChunk];
#pragma unroll
    for (int k = 0; k < kSlotChunk; ++k) rr[k] = a + k < b ? rows[a + k] : kNoRow;
#pragma unroll
    for (int k = 0; k < kSlotChunk; ++k) xv[k] = rr[k] != kNoRow ? (double)vals[(a + k) * R + col] : 0.0;
    uint32_t r0 = kNoRow, cur = kNoRow;
    double s0 = 0, cs = 0;
    bool first = true;
#pragma unroll
    for (int k = 0; k < kSlotChunk; ++k) {
        const uint32_t r = rr[k];
        if (r == kNoRow) continue;
        const double x = xv[k];
        if (cur == kNoRow) {
            r0 = cur = r;
            cs = x;
        } else if (r == cur) {
            cs += x;
        } else {
            if (first) {  // the chunk's first run ends here
                s0 = cs;
                first = false;
            } else {      // a run inside the chunk: complete
                put(cur, cs);
            }
            cur = r;
            cs = x;
        }
    }
    if (r0 == kNoRow) {  // no row in the chunk
        if (!last && col == 0) orow[2 * c] = orow[2 * c + 1] = kNoRow;
        return;
    }
    if (first) {  // one run over the chunk
        s0 = cs;
        cs = 0;
    }
    if (last) {
        put(r0, s0);
        if (!first) put(cur, cs);
        return;
    }
    if (col == 0) {
        orow[2 * c] = r0;
        orow[2 * c + 1] = cur;
    }
    oval[2 * c * R + col] = s0;
    oval[(2 * c + 1) * R + col] = cs;
}

// First level over all slots, one work item per thread.
template <typename OT, int R, typename VT>
__global__ __launch_bounds__(256) void k_bnd_level(const uint32_t *__restrict__ rows, const VT *__restrict__ vals,
                                                   uint64_t n, uint32_t *__restrict__ orow,
                                                   double *__restrict__ oval, OT *__restrict__ out, int zeroed,
                                                   int last) {
    bnd_level_item<OT, R, VT>((uint64_t)blockIdx.x * 256 + threadIdx.x, rows, vals, n, orow, oval, out,
                              zeroed != 0, last != 0);
}

// The remaining levels in one workgroup (a barrier between levels), reading
// level buffer 0 first and alternating.
template <typename OT, int R>
__global__ __launch_bounds__(1024) void k_bnd_tail(uint64_t n, uint32_t *lrow0, double *lval0, uint32_t *lrow1,
                                                   double *lval1, OT *__restrict__ out, int zeroed) {
    uint32_t *lr[2] = {lrow0, lrow1};
    double *lv[2] = {lval0, lval1};
    int k = 0;
    for (;;) {
        const uint64_t nch = (n + kSlotChunk - 1) / kSlotChunk;
        const bool last = nch <= 1;
        for (uint64_t t = threadIdx.x; t < nch * R; t += 1024)
            bnd_level_item<OT, R, double>(t, lr[k], lv[k], n, lr[k ^ 1], lv[k ^ 1], out, zeroed != 0, last);
        if (last) return;
        __syncthreads();
        n = 2 * nch;
        k ^= 1;
    }
}

// Level buffer slots for a merge of ns slots: 2 * ceil(ns / kSlotChunk).
inline uint64_t bnd_level_slots(uint64_t ns) { return 2 * ((ns + kSlotChunk - 1) / kSlotChunk); }

// The merge of ns slots (values VT: float for FP32 workgroup sums, double
// otherwise): levels over the whole grid while more than kTailSlots slots
// remain, then the rest in one workgroup.  lrow / lval: two level buffers of
// bnd_level_slots(ns).
constexpr uint64_t kTailSlots = 2048;

template <typename OT, int R, typename VT>
inline void launch_bnd_merge(const uint32_t *rows, const VT *vals, uint64_t ns, uint32_t *const *lrow,
                             double *const *lval, OT *out, int zeroed, hipStream_t s) {
    uint64_t nch = (ns + kSlotChunk - 1) / kSlotChunk;
    k_bnd_level<OT, R, VT><<<(unsigned)((nch * R + 255) / 256), 256, 0, s>>>(rows, vals, ns, lrow[0], lval[0],
                                                                             out, zeroed, nch <= 1);
    if (nch <= 1) return;
    ns = 2 * nch;
    int k = 0;  // the level buffer holding the current slots
    while (ns > kTailSlots) {
        nch = (ns + kSlotChunk - 1) / kSlotChunk;
        k_bnd_level<OT, R, double><<<(unsigned)((nch * R + 255) / 256), 256, 0, s>>>(
            lrow[k], lval[k], ns, lrow[k ^ 1], lval[k ^ 1], out, zeroed, 0);
        ns = 2 * nch;
        k ^= 1;
    }
    k_bnd_tail<OT, R><<<1, 1024, 0, s>>>(ns, lrow[k], lval[k], lrow[k ^ 1], lval[k ^ 1], out, zeroed);
}

}  // namespace alto
