// Stable LSD radix sort of 64/128-bit keys with a 32-bit payload, for the
// ALTO key sort (the reference uses std::sort of (key, value) pairs,
// include/alto.hpp:1059-1090).  Stability makes ties (duplicate coordinates)
// keep input order; distinct keys get the unique ascending order.
//
// One pass per 8-bit digit over the bits the ALTO masks use:
//   rs_hist    : per-tile digit histogram -> counts[digit * ntiles + tile]
//   scan       : exclusive scan of counts (digit-major) -> global offsets
//   rs_scatter : per-tile stable ranking (wave ballots), digit-grouped
//                staging in LDS, coalesced write-out.
#include "common.hpp"
#include "sort.hpp"

namespace alto {

constexpr int RS_THREADS = 256;
// keys per thread: 16 (4096 keys per workgroup), or 4 below kRsSmall keys,
// where 4096-key tiles would leave most CUs idle (a 300K-key CP-stream slice
// schedule made 73 workgroups)
constexpr uint64_t kRsSmall = 1ULL << 21;
constexpr int SCAN_CHUNK = 4096;

__device__ __forceinline__ uint32_t digit_of(uint64_t w, int sh) {
    return (uint32_t)((w >> sh) & 0xFFu);
}

template <int RS_IPT>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t *__restrict__ kw,
                                                        uint64_t n, int sh,
                                                        uint32_t *__restrict__ counts,
                                                        uint32_t ntiles) {
    constexpr int RS_TILE = RS_THREADS * RS_IPT;
    __shared__ uint32_t hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int j = 0; j < RS_IPT; ++j) {
        uint64_t i = base + (uint64_t)j * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&hist[digit_of(kw[i], sh)], 1u);
    }
    __syncthreads();
    counts[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = hist[threadIdx.x];
}

// ---- exclusive scan of a u32 array (3 kernels) ----
__device__ __forceinline__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t *total,
                                                            uint32_t *sh) {
    // sh: 256 + 4 words of LDS
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
    for (int w = 0; w < 4; ++w) {
        uint32_t s = sh[w];
        if (w < wid) wbase += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wbase + x - v;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t *__restrict__ a,
                                                     uint64_t n, uint32_t *__restrict__ sums) {
    __shared__ uint32_t sh[8];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK;
    uint32_t s = 0;
    for (int j = 0; j < SCAN_CHUNK / 256; ++j) {
        uint64_t i = base + (uint64_t)j * 256 + threadIdx.x;
        if (i < n) s += a[i];
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ __launch_bounds__(256) void k_scan_top(uint32_t *__restrict__ sums, uint32_t m) {
    __shared__ uint32_t sh[8];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < m; base += 256) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < m ? sums[i] : 0, tot;
        uint32_t ex = block_exclusive_scan_256(v, &tot, sh);
        if (i < m) sums[i] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(256) void k_scan_down(uint32_t *__restrict__ a, uint64_t n,
                                                   const uint32_t *__restrict__ sums) {
    __shared__ uint32_t sh[8];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK;
    uint32_t carry = sums[blockIdx.x];
    for (int j = 0; j < SCAN_CHUNK / 256; ++j) {
        uint64_t i = base + (uint64_t)j * 256 + threadIdx.x;
        uint32_t v = i < n ? a[i] : 0, tot;
        uint32_t ex = block_exclusive_scan_256(v, &tot, sh);
        if (i < n) a[i] = carry + ex;
        carry += tot;
    }
}

int exclusive_scan_u32(uint32_t *a, uint64_t n, DevBuf &tmp, hipStream_t s) {
    if (n == 0) return ALTO_OK;
    unsigned nch = cdiv(n, SCAN_CHUNK);
    ALTO_TRY(tmp.alloc(sizeof(uint32_t) * nch));
    k_scan_reduce<<<nch, 256, 0, s>>>(a, n, tmp.as<uint32_t>());
    k_scan_top<<<1, 256, 0, s>>>(tmp.as<uint32_t>(), nch);
    k_scan_down<<<nch, 256, 0, s>>>(a, n, tmp.as<uint32_t>());
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

// ---- stable scatter ----
template <int KW, int RS_IPT>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(
    const uint64_t *__restrict__ k0, const uint64_t *__restrict__ k1,
    const uint32_t *__restrict__ pin, uint64_t *__restrict__ o0, uint64_t *__restrict__ o1,
    uint32_t *__restrict__ pout, uint64_t n, int word, int sh,
    const uint32_t *__restrict__ offs, uint32_t ntiles) {
    constexpr int RS_TILE = RS_THREADS * RS_IPT;
    __shared__ uint32_t wcnt[4][256];
    __shared__ uint32_t boff[256];
    __shared__ uint32_t goff[256];
    __shared__ uint32_t scan_sh[8];
    __shared__ uint64_t s0[RS_TILE];
    __shared__ uint64_t s1[KW == 2 ? RS_TILE : 1];
    __shared__ uint32_t sp[RS_TILE];
    __shared__ uint8_t sd[RS_TILE];

    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int q = 0; q < 4; ++q) wcnt[q][t] = 0;
    __syncthreads();

    const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    const uint64_t lt_mask = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    uint64_t r0[RS_IPT], r1[RS_IPT];
    uint32_t rp[RS_IPT];
    uint32_t rr[RS_IPT];
    uint32_t rd[RS_IPT];
#pragma unroll
    for (int j = 0; j < RS_IPT; ++j) {
        uint64_t i = base + (uint64_t)w * (RS_IPT * 64) + (uint64_t)j * 64 + lane;
        bool valid = i < n;
        uint64_t a0 = valid ? k0[i] : 0, a1 = 0;
        if (KW == 2) a1 = valid ? k1[i] : 0;
        uint32_t p = valid ? pin[i] : 0;
        uint32_t d = digit_of(word == 0 ? a0 : a1, sh);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bb = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        uint32_t rin = (uint32_t)__popcll(peers & lt_mask);
        uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t c0 = valid ? wcnt[w][d] : 0;
        __builtin_amdgcn_wave_barrier();
        if (valid && rin == 0) wcnt[w][d] = c0 + cnt;
        __builtin_amdgcn_wave_barrier();
        r0[j] = a0;
        r1[j] = a1;
        rp[j] = p;
        rr[j] = valid ? c0 + rin : 0xFFFFFFFFu;
        rd[j] = d;
    }
    __syncthreads();
    // digit t: exclusive prefix over waves, then over digits
    uint32_t run = 0;
    for (int q = 0; q < 4; ++q) {
        uint32_t c = wcnt[q][t];
        wcnt[q][t] = run;
        run += c;
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_scan_256(run, &tot, scan_sh);
    boff[t] = ex;
    goff[t] = offs[(uint64_t)t * ntiles + blockIdx.x];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_IPT; ++j) {
        if (rr[j] != 0xFFFFFFFFu) {
            uint32_t d = rd[j];
            uint32_t pos = boff[d] + wcnt[w][d] + rr[j];
            s0[pos] = r0[j];
            if (KW == 2) s1[pos] = r1[j];
            sp[pos] = rp[j];
            sd[pos] = (uint8_t)d;
        }
    }
    __syncthreads();
    uint32_t nvalid = (uint32_t)((n - base) < (uint64_t)RS_TILE ? (n - base) : RS_TILE);
#pragma unroll
    for (int j = 0; j < RS_IPT; ++j) {
        uint32_t slot = (uint32_t)j * RS_THREADS + t;
        if (slot < nvalid) {
            uint32_t d = sd[slot];
            uint64_t dst = (uint64_t)goff[d] + (slot - boff[d]);
            o0[dst] = s0[slot];
            if (KW == 2) o1[dst] = s1[slot];
            pout[dst] = sp[slot];
        }
    }
}

__global__ void k_iota_u32(uint32_t *p, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}

// Sorts (k0[, k1]) keys ascending over bits [0, bits) stably, carrying the
// payload.  k1 == nullptr for 64-bit keys.  On return the sorted data is in
// the caller's arrays.  n < 2^32.
int radix_sort_pairs(uint64_t *k0, uint64_t *k1, uint32_t *pay, uint64_t n, int bits,
                     SortScratch &ws, hipStream_t s) {
    if (n <= 1 || bits <= 0) return ALTO_OK;
    const int KW = k1 ? 2 : 1;
    const bool small = n < kRsSmall;
    unsigned ntiles = cdiv(n, RS_THREADS * (small ? 4 : 16));
    ALTO_TRY(ws.t0.alloc(sizeof(uint64_t) * n));
    if (KW == 2) ALTO_TRY(ws.t1.alloc(sizeof(uint64_t) * n));
    ALTO_TRY(ws.tp.alloc(sizeof(uint32_t) * n));
    ALTO_TRY(ws.counts.alloc(sizeof(uint32_t) * 256 * (uint64_t)ntiles));
    uint64_t *a0 = k0, *a1 = k1, *b0 = ws.t0.as<uint64_t>(), *b1 = ws.t1.as<uint64_t>();
    uint32_t *ap = pay, *bp = ws.tp.as<uint32_t>();
    int passes = (bits + 7) / 8;
    for (int ps = 0; ps < passes; ++ps) {
        int bit = ps * 8;
        int word = bit / 64, sh = bit % 64;
        const uint64_t *src = word == 0 ? a0 : a1;
        (small ? k_rs_hist<4> : k_rs_hist<16>)<<<ntiles, RS_THREADS, 0, s>>>(src, n, sh,
                                                                          ws.counts.as<uint32_t>(), ntiles);
        ALTO_LAUNCH_CHECK();
        ALTO_TRY(exclusive_scan_u32(ws.counts.as<uint32_t>(), 256ULL * ntiles, ws.scan_tmp, s));
        if (KW == 1)
            (small ? k_rs_scatter<1, 4> : k_rs_scatter<1, 16>)<<<ntiles, RS_THREADS, 0, s>>>(
                a0, nullptr, ap, b0, nullptr, bp, n, word, sh, ws.counts.as<uint32_t>(), ntiles);
        else
            (small ? k_rs_scatter<2, 4> : k_rs_scatter<2, 16>)<<<ntiles, RS_THREADS, 0, s>>>(
                a0, a1, ap, b0, b1, bp, n, word, sh, ws.counts.as<uint32_t>(), ntiles);
        ALTO_LAUNCH_CHECK();
        std::swap(a0, b0);
        std::swap(a1, b1);
        std::swap(ap, bp);
    }
    if (a0 != k0) {
        ALTO_HIP_TRY(hipMemcpyAsync(k0, a0, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, s));
        if (KW == 2)
            ALTO_HIP_TRY(hipMemcpyAsync(k1, a1, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, s));
        ALTO_HIP_TRY(hipMemcpyAsync(pay, ap, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    }
    return ALTO_OK;
}

int iota_u32(uint32_t *p, uint64_t n, hipStream_t s) {
    if (n == 0) return ALTO_OK;
    k_iota_u32<<<cdiv(n, 256), 256, 0, s>>>(p, n);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

}  // namespace alto
