// ALTO tensor construction on the GPU: masks, linearisation, key sort,
// line partitions, and the per-mode output-ordered schedules.
//
// Reference: create_alto (include/alto.hpp:436-582), setup_packed_alto
// (954-1057), sort_alto (1059-1090), prtn_alto (1128-1226).
#include <algorithm>
#include <string.h>

#include "tensor.hpp"

namespace alto {

int mode_bits(uint64_t dim) {
    // util.hpp:50-59: 64 - clz(dim - 1); a size-1 mode gets 0 bits.
    if (dim <= 1) return 0;
    return 64 - __builtin_clzll(dim - 1);
}

int build_masks(int nmode, const uint64_t *dims, int kw, uint64_t *masks, BitPos *bp) {
    int bits[kMaxModes], order[kMaxModes];
    int total = 0, maxb = 0;
    for (int n = 0; n < nmode; ++n) {
        bits[n] = mode_bits(dims[n]);
        total += bits[n];
        maxb = std::max(maxb, bits[n]);
        order[n] = n;
    }
    if (total > 64 * kw) return -1;
    // Modes ordered by bit count, ties in mode order (SHORT_FIRST).
    std::stable_sort(order, order + nmode, [&](int a, int b) { return bits[a] < bits[b]; });
    memset(masks, 0, sizeof(uint64_t) * nmode * kw);
    memset(bp, 0, sizeof(*bp));
    bp->nmode = nmode;
    int next = 0;  // next key bit, LSB first
    for (int level = 0; level < maxb; ++level)
        for (int k = 0; k < nmode; ++k) {
            int m = order[k];
            if (level < bits[m]) {
                masks[m * kw + next / 64] |= 1ULL << (next % 64);
                bp->pos[m][level] = (uint8_t)next;
                ++next;
            }
        }
    for (int m = 0; m < nmode; ++m) bp->bits[m] = bits[m];
    return total;
}

struct IdxPtrs {
    const uint64_t *p[kMaxModes];
};

// key = OR_m pdep(cidx[m][i], mask_m); flags out-of-range coordinates.
__global__ __launch_bounds__(256) void k_linearize(IdxPtrs ci, uint64_t n, BitPos bp,
                                                   DimsArg dims, uint64_t *__restrict__ k0,
                                                   uint64_t *__restrict__ k1,
                                                   uint32_t *__restrict__ bad) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t lo = 0, hi = 0;
    bool oob = false;
    for (int m = 0; m < bp.nmode; ++m) {
        uint64_t x = ci.p[m][i];
        oob |= x >= dims.d[m];
        for (int b = 0; b < bp.bits[m]; ++b) {
            if ((x >> b) & 1ULL) {
                int p = bp.pos[m][b];
                if (p < 64) lo |= 1ULL << p;
                else hi |= 1ULL << (p - 64);
            }
        }
    }
    k0[i] = lo;
    if (k1) k1[i] = hi;
    if (oob) atomicOr(bad, 1u);
}

__device__ __forceinline__ uint64_t pext_pos(uint64_t lo, uint64_t hi, const BitPos &bp, int m) {
    uint64_t x = 0;
    for (int b = 0; b < bp.bits[m]; ++b) {
        int p = bp.pos[m][b];
        uint64_t bit = p < 64 ? (lo >> p) & 1ULL : (hi >> (p - 64)) & 1ULL;
        x |= bit << b;
    }
    return x;
}

// Interleaved ALTO keys, packed keys and gathered FP64 values in sorted order.
__global__ __launch_bounds__(256) void k_finish(const uint64_t *__restrict__ k0,
                                                const uint64_t *__restrict__ k1,
                                                const uint32_t *__restrict__ perm,
                                                const double *__restrict__ vin, uint64_t n,
                                                BitPos bp, PackInfo pk, int kw, int pkw,
                                                uint64_t *__restrict__ keys,
                                                uint64_t *__restrict__ pkeys,
                                                double *__restrict__ vout) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t lo = k0[i], hi = k1 ? k1[i] : 0;
    keys[i * kw] = lo;
    if (kw == 2) keys[i * kw + 1] = hi;
    uint64_t plo = 0, phi = 0;
    for (int m = 0; m < bp.nmode; ++m) {
        uint64_t x = pext_pos(lo, hi, bp, m);
        int s = pk.shift[m], w = pk.width[m];
        if (w == 0) continue;
        if (s >= 64) {
            phi |= x << (s - 64);
        } else {
            plo |= x << s;
            if (s + w > 64) phi |= x >> (64 - s);
        }
    }
    pkeys[i * pkw] = plo;
    if (pkw == 2) pkeys[i * pkw + 1] = phi;
    vout[i] = vin[perm[i]];
}

struct Chunk {
    uint64_t start, end;
    uint32_t part;
};

// Per-partition [min, max] of every mode index (prtn_alto, alto.hpp:1199-1226).
__global__ __launch_bounds__(256) void k_intervals(const uint64_t *__restrict__ pkeys, int pkw,
                                                   PackInfo pk, const Chunk *__restrict__ chunks,
                                                   unsigned long long *__restrict__ iv) {
    __shared__ unsigned long long smin[kMaxModes][4], smax[kMaxModes][4];
    Chunk c = chunks[blockIdx.x];
    unsigned long long lo[kMaxModes], hi[kMaxModes];
    for (int m = 0; m < pk.nmode; ++m) {
        lo[m] = ~0ULL;
        hi[m] = 0;
    }
    for (uint64_t i = c.start + threadIdx.x; i < c.end; i += blockDim.x) {
        uint64_t a = pkeys[i * pkw], b = pkw == 2 ? pkeys[i * pkw + 1] : 0;
        for (int m = 0; m < pk.nmode; ++m) {
            unsigned long long x = extract_field(a, b, pk.shift[m], pk.width[m]);
            lo[m] = x < lo[m] ? x : lo[m];
            hi[m] = x > hi[m] ? x : hi[m];
        }
    }
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int m = 0; m < pk.nmode; ++m) {
        unsigned long long a = lo[m], b = hi[m];
        for (int o = 32; o > 0; o >>= 1) {
            unsigned long long a2 = __shfl_down(a, o, 64), b2 = __shfl_down(b, o, 64);
            a = a2 < a ? a2 : a;
            b = b2 > b ? b2 : b;
        }
        if (lane == 0) {
            smin[m][wid] = a;
            smax[m][wid] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)pk.nmode) {
        int m = threadIdx.x;
        unsigned long long a = smin[m][0], b = smax[m][0];
        for (int w = 1; w < 4; ++w) {
            a = smin[m][w] < a ? smin[m][w] : a;
            b = smax[m][w] > b ? smax[m][w] : b;
        }
        atomicMin(&iv[((uint64_t)c.part * pk.nmode + m) * 2], a);
        atomicMax(&iv[((uint64_t)c.part * pk.nmode + m) * 2 + 1], b);
    }
}

__global__ void k_init_intervals(unsigned long long *iv, uint64_t nprtn, int nmode) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nprtn * nmode) {
        iv[2 * i] = ~0ULL;
        iv[2 * i + 1] = 0;
    }
}

// Conflict-resolution bits (include/alto.hpp:1250-1346).  Per CR mode c the
// lowest and highest partition that touches each row; a row is a conflict
// when they differ (the reference's owner array ends at -1 exactly then).
struct CrArgs {
    int n;                  // CR modes
    int shift[kMaxModes];   // packed field of the mode
    int width[kMaxModes];
    int bit[kMaxModes];     // its bit in the key (same position in both layouts)
    uint64_t off[kMaxModes];  // its rows' offset in the owner arrays
};

__global__ __launch_bounds__(256) void k_cr_owner(const uint64_t *__restrict__ pkeys, int pkw, uint64_t nnz,
                                                  uint64_t per, CrArgs a, uint32_t *__restrict__ plo,
                                                  uint32_t *__restrict__ phi) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    const uint32_t p = (uint32_t)(i / per);  // partition p holds [p * per, (p + 1) * per)
    const uint64_t lo = pkeys[i * pkw], hi = pkw == 2 ? pkeys[i * pkw + 1] : 0;
    for (int c = 0; c < a.n; ++c) {
        const uint64_t row = a.off[c] + extract_field(lo, hi, a.shift[c], a.width[c]);
        atomicMin(plo + row, p);
        atomicMax(phi + row, p);
    }
}

__global__ __launch_bounds__(256) void k_cr_set(uint64_t *__restrict__ keys, int kw, uint64_t *__restrict__ pkeys,
                                                int pkw, uint64_t nnz, CrArgs a, const uint32_t *__restrict__ plo,
                                                const uint32_t *__restrict__ phi) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    const uint64_t lo = pkeys[i * pkw], hi = pkw == 2 ? pkeys[i * pkw + 1] : 0;
    uint64_t set[2] = {0, 0};
    for (int c = 0; c < a.n; ++c) {
        const uint64_t row = a.off[c] + extract_field(lo, hi, a.shift[c], a.width[c]);
        if (plo[row] != phi[row]) set[a.bit[c] >> 6] |= 1ULL << (a.bit[c] & 63);
    }
    if (set[0] | set[1]) {
        pkeys[i * pkw] = lo | set[0];
        if (pkw == 2) pkeys[i * pkw + 1] = hi | set[1];
        keys[i * kw] |= set[0];
        if (kw == 2) keys[i * kw + 1] |= set[1];
    }
}

__global__ void k_to_f32(float *d, const double *s, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}

int ensure_vals32(TensorImpl *t, hipStream_t s) {
    if (t->vals32_valid) return ALTO_OK;
    t->vals32_valid = true;
    ALTO_TRY(t->vals32.alloc(sizeof(float) * t->nnz));
    if (t->nnz) {
        k_to_f32<<<cdiv(t->nnz, 256), 256, 0, s>>>(t->vals32.as<float>(), t->vals64.as<double>(),
                                                  t->nnz);
        ALTO_LAUNCH_CHECK();
    }
    return ALTO_OK;
}

struct FiberOrder {
    int n;
    int mode[kMaxModes];  // most significant first
};

// Composite sort key of a nonzero: its fields concatenated in `ord` order
// (first mode in the top bits), at most 128 bits.
__global__ __launch_bounds__(256) void k_fiber_key(const uint64_t *__restrict__ pkeys, int pkw,
                                                   PackInfo pk, FiberOrder ord, uint64_t n,
                                                   uint64_t *__restrict__ k0,
                                                   uint64_t *__restrict__ k1) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t a = pkeys[i * pkw], b = pkw == 2 ? pkeys[i * pkw + 1] : 0;
    unsigned __int128 c = 0;
    for (int q = 0; q < ord.n; ++q) {
        int m = ord.mode[q], w = pk.width[m];
        if (w == 0) continue;
        c = (c << w) | extract_field(a, b, pk.shift[m], w);
    }
    k0[i] = (uint64_t)c;
    if (k1) k1[i] = (uint64_t)(c >> 64);
}

__global__ __launch_bounds__(256) void k_gather_sched(const uint64_t *__restrict__ pkeys, int pkw,
                                                      const double *__restrict__ v64,
                                                      const uint32_t *__restrict__ perm, uint64_t n,
                                                      uint64_t *__restrict__ opk,
                                                      double *__restrict__ o64) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t j = perm[i];
    opk[i * pkw] = pkeys[(uint64_t)j * pkw];
    if (pkw == 2) opk[i * pkw + 1] = pkeys[(uint64_t)j * pkw + 1];
    o64[i] = v64[j];
}

__device__ __forceinline__ bool same_key(const uint64_t *k, int pkw, uint64_t i, uint64_t j) {
    return k[i * pkw] == k[j * pkw] && (pkw == 1 || k[i * pkw + 1] == k[j * pkw + 1]);
}

// head[i] = 1 where a run of equal coordinates starts.
__global__ __launch_bounds__(256) void k_run_heads(const uint64_t *__restrict__ k, int pkw,
                                                   uint64_t n, uint32_t *__restrict__ head) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || !same_key(k, pkw, i, i - 1)) ? 1u : 0u;
}

// One thread per run head: the values of a run of at most kShortRun equal
// coordinates are summed in order in FP64 (the run keeps the ALTO order of
// the duplicates); longer runs are appended to `longs` for k_run_merge_long.
// pos = exclusive scan of the head flags.
constexpr int kShortRun = 64;

__device__ __forceinline__ void put_merged(const uint64_t *k, int pkw, uint64_t i, uint64_t o,
                                           double sum, uint64_t *ok, double *ov64, float *ov32) {
    ok[o * pkw] = k[i * pkw];
    if (pkw == 2) ok[o * pkw + 1] = k[i * pkw + 1];
    ov32[o] = (float)sum;
    if (ov64) ov64[o] = sum;
}

__global__ __launch_bounds__(256) void k_run_merge(const uint64_t *__restrict__ k, int pkw,
                                                   const double *__restrict__ v, uint64_t n,
                                                   const uint32_t *__restrict__ pos,
                                                   uint64_t *__restrict__ ok, double *__restrict__ ov64,
                                                   float *__restrict__ ov32,
                                                   uint64_t *__restrict__ longs,
                                                   unsigned long long *__restrict__ nlong) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (i > 0 && same_key(k, pkw, i, i - 1))) return;
    double sum = v[i];
    uint64_t j = i + 1;
    for (; j < n && j < i + kShortRun && same_key(k, pkw, j, i); ++j) sum += v[j];
    if (j < n && j == i + kShortRun && same_key(k, pkw, j, i)) {
        longs[atomicAdd(nlong, 1ULL)] = i;
        return;
    }
    put_merged(k, pkw, i, pos[i], sum, ok, ov64, ov32);
}

// One wave per long run: lane l sums elements head + l, head + l + 64, ...
// of the run, then a fixed shuffle tree (deterministic).
__global__ __launch_bounds__(256) void k_run_merge_long(const uint64_t *__restrict__ k, int pkw,
                                                        const double *__restrict__ v, uint64_t n,
                                                        const uint32_t *__restrict__ pos,
                                                        uint64_t *__restrict__ ok,
                                                        double *__restrict__ ov64,
                                                        float *__restrict__ ov32,
                                                        const uint64_t *__restrict__ longs,
                                                        const unsigned long long *__restrict__ nlong) {
    const uint64_t nl = *nlong;
    const int lane = threadIdx.x & 63;
    for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64; w < nl;
         w += ((uint64_t)gridDim.x * blockDim.x) / 64) {
        const uint64_t h = longs[w];
        double a = 0.0;
        bool more = true;
        for (uint64_t b = h; more; b += 64) {
            const uint64_t j = b + lane;
            const bool in = j < n && same_key(k, pkw, j, h);
            if (in) a += v[j];
            more = __any(in);
        }
        for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
        if (lane == 0) put_merged(k, pkw, h, pos[h], a, ok, ov64, ov32);
    }
}

// Output-ordered copy of the tensor for one mode (the row order the
// reference's op_sched variant walks through oidx, alto.hpp:1708-1801).
// Inside a row the nonzeros are in fiber order: sorted by the other modes'
// indices, smallest mode first (ties by mode number), so nonzeros that share
// a prefix of those indices are adjacent and the MTTKRP gathers the prefix
// rows once per fiber.  Nonzeros with equal coordinates (duplicates kept by
// the reference's ALTO tensor) are merged into one entry whose value is
// their FP64 sum in ALTO order: the MTTKRP is linear in the values, so this
// only reorders the additions.
void fiber_levels(const TensorImpl *t, int mode, int *lvl) {
    int k = 0;
    for (int m = 0; m < t->nmode; ++m)
        if (m != mode) lvl[k++] = m;
    std::stable_sort(lvl, lvl + k, [&](int a, int b) { return t->dims[a] < t->dims[b]; });
}

int ensure_schedule(TensorImpl *t, int mode, bool need64, hipStream_t s) {
    ModeSchedule &sc = t->sched[mode];
    if (sc.built && (!need64 || sc.has64)) return ALTO_OK;
    sc.hot.built = false;
    const uint64_t n = t->nnz;
    const int pkw = t->pkw;
    fiber_levels(t, mode, sc.lvl);
    sc.n = 0;
    sc.neff = -1;  // set below; an empty schedule is never copied
    if (n == 0) {
        ALTO_TRY(sc.pkeys.alloc(sizeof(uint64_t) * pkw));
        ALTO_TRY(sc.vals32.alloc(sizeof(float)));
        if (need64) ALTO_TRY(sc.vals64.alloc(sizeof(double)));
        sc.built = true;
        sc.has64 = need64;
        return ALTO_OK;
    }
    FiberOrder ord{};
    ord.n = t->nmode;
    ord.mode[0] = mode;
    for (int q = 0; q < t->nmode - 1; ++q) ord.mode[q + 1] = sc.lvl[q];
    sc.neff = 0;
    for (int q = 0; q < ord.n; ++q)
        if (t->pack.width[ord.mode[q]] > 0) sc.eff[sc.neff++] = ord.mode[q];
    // A mode of size one adds no sort field: a mode whose remaining fields
    // come in the same order as those of a built schedule (the stream mode of
    // a one-step time slice and the smallest other mode) gets a copy of it.
    for (int m2 = 0; m2 < t->nmode; ++m2) {
        const ModeSchedule &o = t->sched[m2];
        if (m2 == mode || !o.built || (need64 && !o.has64) || o.neff != sc.neff ||
            memcmp(o.eff, sc.eff, sizeof(int) * sc.neff) != 0)
            continue;
        const uint64_t m = o.n;
        ALTO_TRY(sc.pkeys.alloc(sizeof(uint64_t) * pkw * std::max<uint64_t>(m, 1)));
        ALTO_TRY(sc.vals32.alloc(sizeof(float) * (m + kSchedPad)));
        ALTO_HIP_TRY(hipMemcpyAsync(sc.pkeys.p, o.pkeys.p, sizeof(uint64_t) * pkw * m, hipMemcpyDeviceToDevice, s));
        ALTO_HIP_TRY(hipMemcpyAsync(sc.vals32.p, o.vals32.p, sizeof(float) * m, hipMemcpyDeviceToDevice, s));
        if (need64) {
            ALTO_TRY(sc.vals64.alloc(sizeof(double) * (m + kSchedPad)));
            ALTO_HIP_TRY(hipMemcpyAsync(sc.vals64.p, o.vals64.p, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
        }
        sc.n = m;
        sc.built = true;
        sc.has64 = need64;
        return ALTO_OK;
    }
    const bool wide = t->total_bits > 64;
    DevBuf lk0, lk1, lperm, lsk, lsv, llongs;
    const bool keep = t->keep_sched_scratch;
    DevBuf &k0 = keep ? t->sch_k0 : lk0, &k1 = keep ? t->sch_k1 : lk1, &perm = keep ? t->sch_perm : lperm,
           &sk = keep ? t->sch_sk : lsk, &sv = keep ? t->sch_sv : lsv, &longs = keep ? t->sch_longs : llongs;
    ALTO_TRY(k0.alloc(sizeof(uint64_t) * n));
    if (wide) ALTO_TRY(k1.alloc(sizeof(uint64_t) * n));
    ALTO_TRY(perm.alloc(sizeof(uint32_t) * n));
    k_fiber_key<<<cdiv(n, 256), 256, 0, s>>>(t->pkeys.as<uint64_t>(), pkw, t->pack, ord, n,
                                             k0.as<uint64_t>(), wide ? k1.as<uint64_t>() : nullptr);
    ALTO_LAUNCH_CHECK();
    ALTO_TRY(iota_u32(perm.as<uint32_t>(), n, s));
    ALTO_TRY(radix_sort_pairs(k0.as<uint64_t>(), wide ? k1.as<uint64_t>() : nullptr,
                              perm.as<uint32_t>(), n, t->total_bits, t->sort_ws, s));
    ALTO_TRY(sk.alloc(sizeof(uint64_t) * pkw * n));
    ALTO_TRY(sv.alloc(sizeof(double) * n));
    k_gather_sched<<<cdiv(n, 256), 256, 0, s>>>(t->pkeys.as<uint64_t>(), pkw,
                                                t->vals64.as<double>(), perm.as<uint32_t>(), n,
                                                sk.as<uint64_t>(), sv.as<double>());
    ALTO_LAUNCH_CHECK();
    // merge duplicate coordinates (perm reused as the head flags / positions)
    uint32_t *pos = perm.as<uint32_t>();
    k_run_heads<<<cdiv(n, 256), 256, 0, s>>>(sk.as<uint64_t>(), pkw, n, pos);
    ALTO_LAUNCH_CHECK();
    uint32_t last_head = 0, last_pos = 0;
    ALTO_HIP_TRY(hipMemcpyAsync(&last_head, pos + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ALTO_TRY(exclusive_scan_u32(pos, n, t->sort_ws.scan_tmp, s));
    ALTO_HIP_TRY(hipMemcpyAsync(&last_pos, pos + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ALTO_HIP_TRY(hipStreamSynchronize(s));
    const uint64_t m = (uint64_t)last_pos + last_head;
    ALTO_TRY(sc.pkeys.alloc(sizeof(uint64_t) * pkw * m));
    ALTO_TRY(sc.vals32.alloc(sizeof(float) * (m + kSchedPad)));
    if (need64) ALTO_TRY(sc.vals64.alloc(sizeof(double) * (m + kSchedPad)));
    ALTO_TRY(longs.alloc(sizeof(uint64_t) * (n / kShortRun + 2)));
    unsigned long long *nlong = (unsigned long long *)(longs.as<uint64_t>() + n / kShortRun + 1);
    ALTO_HIP_TRY(hipMemsetAsync(nlong, 0, sizeof(unsigned long long), s));
    k_run_merge<<<cdiv(n, 256), 256, 0, s>>>(sk.as<uint64_t>(), pkw, sv.as<double>(), n, pos,
                                             sc.pkeys.as<uint64_t>(),
                                             need64 ? sc.vals64.as<double>() : nullptr,
                                             sc.vals32.as<float>(), longs.as<uint64_t>(), nlong);
    ALTO_LAUNCH_CHECK();
    k_run_merge_long<<<1024, 256, 0, s>>>(sk.as<uint64_t>(), pkw, sv.as<double>(), n, pos,
                                          sc.pkeys.as<uint64_t>(),
                                          need64 ? sc.vals64.as<double>() : nullptr,
                                          sc.vals32.as<float>(), longs.as<uint64_t>(), nlong);
    ALTO_LAUNCH_CHECK();
    if (!keep) ALTO_HIP_TRY(hipStreamSynchronize(s));  // the local scratch is released on return
    sc.n = m;
    sc.built = true;
    sc.has64 = need64;
    return ALTO_OK;
}


int build_alto(TensorImpl *t, const alto_hip_coo *coo, int nprtn, int key_words, bool check,
               hipStream_t s, bool alto_order) {
    if (!coo || !coo->dims || !coo->cidx || !coo->vals) return ALTO_ERR_ARG;
    if (coo->nmode < 1 || coo->nmode > kMaxModes || nprtn < 0 || (!alto_order && nprtn != 0)) return ALTO_ERR_ARG;
    if (coo->nnz >= (1ULL << 32)) {
        set_error("nnz %llu exceeds 2^32 - 1", (unsigned long long)coo->nnz);
        return ALTO_ERR_UNSUPPORTED;
    }
    for (int m = 0; m < coo->nmode; ++m)
        if (coo->dims[m] == 0 || (coo->nnz && !coo->cidx[m])) return ALTO_ERR_ARG;
    t->nmode = coo->nmode;
    t->nnz = coo->nnz;
    t->nprtn = nprtn;
    memset(t->cr_masks, 0, sizeof(t->cr_masks));  // update_salto clears them too (alto.hpp:228, 337)
    t->vals32_valid = false;
    for (int m = 0; m < kMaxModes; ++m) {
        t->sched[m].built = false;
        t->sched[m].hot.built = false;
    }
    t->hot.built = false;
    memcpy(t->dims, coo->dims, sizeof(uint64_t) * t->nmode);
    int total = 0;
    for (int m = 0; m < t->nmode; ++m) total += mode_bits(t->dims[m]);
    int kw = key_words == 0 ? (total <= 64 ? 1 : 2) : key_words;
    if (kw != 1 && kw != 2) return ALTO_ERR_ARG;
    if (build_masks(t->nmode, t->dims, kw, t->masks, &t->bp) < 0) {
        set_error("ALTO key needs %d bits, key holds %d", total, 64 * kw);
        return ALTO_ERR_KEY_WIDTH;
    }
    t->kw = kw;
    t->total_bits = total;
    t->pkw = total <= 64 ? 1 : 2;
    t->pack.nmode = t->nmode;
    int sh = 0;
    for (int m = 0; m < t->nmode; ++m) {
        t->pack.shift[m] = sh;
        t->pack.width[m] = mode_bits(t->dims[m]);
        sh += t->pack.width[m];
    }
    const uint64_t n = t->nnz;
    ALTO_TRY(t->keys.alloc(sizeof(uint64_t) * kw * n));
    ALTO_TRY(t->pkeys.alloc(sizeof(uint64_t) * t->pkw * n));
    ALTO_TRY(t->perm.alloc(sizeof(uint32_t) * n));
    ALTO_TRY(t->vals64.alloc(sizeof(double) * n));
    DevBuf dci[kMaxModes], dvals;
    IdxPtrs ip{};
    const double *vin = coo->vals;
    if (!coo->on_device) {
        for (int m = 0; m < t->nmode; ++m) {
            ALTO_TRY(dci[m].alloc(sizeof(uint64_t) * n));
            if (n)
                ALTO_HIP_TRY(hipMemcpyAsync(dci[m].p, coo->cidx[m], sizeof(uint64_t) * n,
                                            hipMemcpyHostToDevice, s));
            ip.p[m] = dci[m].as<uint64_t>();
        }
        ALTO_TRY(dvals.alloc(sizeof(double) * n));
        if (n) ALTO_HIP_TRY(hipMemcpyAsync(dvals.p, coo->vals, sizeof(double) * n, hipMemcpyHostToDevice, s));
        vin = dvals.as<double>();
    } else {
        for (int m = 0; m < t->nmode; ++m) {
            if (n) ALTO_TRY(check_device_access(coo->cidx[m], "ALTO tensor coordinates"));
            ip.p[m] = coo->cidx[m];
        }
        if (n) ALTO_TRY(check_device_access(coo->vals, "ALTO tensor values"));
    }
    ALTO_TRY(t->k0.alloc(sizeof(uint64_t) * n));
    if (kw == 2) ALTO_TRY(t->k1.alloc(sizeof(uint64_t) * n));
    ALTO_TRY(t->bad.alloc(sizeof(uint32_t)));
    if (check) ALTO_HIP_TRY(hipMemsetAsync(t->bad.p, 0, sizeof(uint32_t), s));
    DimsArg da{};
    for (int m = 0; m < t->nmode; ++m) da.d[m] = t->dims[m];
    if (n) {
        k_linearize<<<cdiv(n, 256), 256, 0, s>>>(ip, n, t->bp, da, t->k0.as<uint64_t>(),
                                                 kw == 2 ? t->k1.as<uint64_t>() : nullptr,
                                                 t->bad.as<uint32_t>());
        ALTO_LAUNCH_CHECK();
        ALTO_TRY(iota_u32(t->perm.as<uint32_t>(), n, s));
        if (alto_order)
            ALTO_TRY(radix_sort_pairs(t->k0.as<uint64_t>(), kw == 2 ? t->k1.as<uint64_t>() : nullptr,
                                      t->perm.as<uint32_t>(), n, total, t->sort_ws, s));
        k_finish<<<cdiv(n, 256), 256, 0, s>>>(t->k0.as<uint64_t>(),
                                              kw == 2 ? t->k1.as<uint64_t>() : nullptr,
                                              t->perm.as<uint32_t>(), vin, n, t->bp, t->pack, kw,
                                              t->pkw, t->keys.as<uint64_t>(),
                                              t->pkeys.as<uint64_t>(), t->vals64.as<double>());
        ALTO_LAUNCH_CHECK();
    }
    if (check) {
        uint32_t hbad = 0;
        ALTO_HIP_TRY(hipMemcpyAsync(&hbad, t->bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        ALTO_HIP_TRY(hipStreamSynchronize(s));
        if (hbad) {
            set_error("a coordinate is not below its mode size");
            return ALTO_ERR_ARG;
        }
    } else if (!coo->on_device) {
        ALTO_HIP_TRY(hipStreamSynchronize(s));  // host-staged inputs die with this call
    }
    if (nprtn == 0) {
        t->prtn_ptr.assign(1, 0);
        t->intervals.clear();
        return ALTO_OK;
    }
    // prtn_alto: nnz_ptrn = ceil(nnz / nprtn) (alto.hpp:1134-1161)
    t->prtn_ptr.assign(nprtn + 1, 0);
    uint64_t per = (n + nprtn - 1) / nprtn;
    std::vector<Chunk> chunks;
    const uint64_t CH = 1 << 16;
    for (int p = 0; p < nprtn; ++p) {
        uint64_t a = std::min<uint64_t>((uint64_t)p * per, n);
        uint64_t b = std::min<uint64_t>((uint64_t)(p + 1) * per, n);
        t->prtn_ptr[p + 1] = b;
        for (uint64_t c = a; c < b; c += CH) chunks.push_back({c, std::min(c + CH, b), (uint32_t)p});
    }
    ALTO_TRY(t->ivbuf.alloc(sizeof(uint64_t) * 2 * nprtn * t->nmode));
    k_init_intervals<<<cdiv((uint64_t)nprtn * t->nmode, 256), 256, 0, s>>>(
        t->ivbuf.as<unsigned long long>(), nprtn, t->nmode);
    if (!chunks.empty()) {
        ALTO_TRY(t->chunkbuf.alloc(sizeof(Chunk) * chunks.size()));
        ALTO_HIP_TRY(hipMemcpyAsync(t->chunkbuf.p, chunks.data(), sizeof(Chunk) * chunks.size(),
                                    hipMemcpyHostToDevice, s));
        k_intervals<<<(unsigned)chunks.size(), 256, 0, s>>>(t->pkeys.as<uint64_t>(), t->pkw, t->pack,
                                                            t->chunkbuf.as<Chunk>(),
                                                            t->ivbuf.as<unsigned long long>());
    }
    t->intervals.assign((size_t)2 * nprtn * t->nmode, 0);
    ALTO_HIP_TRY(hipMemcpyAsync(t->intervals.data(), t->ivbuf.p, sizeof(uint64_t) * t->intervals.size(),
                                hipMemcpyDeviceToHost, s));
    ALTO_HIP_TRY(hipStreamSynchronize(s));
    for (int p = 0; p < nprtn; ++p)
        for (int m = 0; m < t->nmode; ++m) {
            uint64_t *lo = &t->intervals[((size_t)p * t->nmode + m) * 2];
            if (t->prtn_ptr[p] == t->prtn_ptr[p + 1]) {
                lo[0] = 0;  // empty partition: [0, 0] (alto.hpp:1220-1225)
                lo[1] = 0;
            } else if (lo[0] > lo[1]) {
                lo[0] = lo[1];
            }
        }
    return ALTO_OK;
}

// GENERATE_CR_BITS (include/alto.hpp:1250-1346).  A mode takes a bit when
// it has more than MIN_NUM_FIBERS = 512 rows, at most MIN_FIBER_REUSE = 4
// nonzeros per row (integer nnz / dims) and a free key bit remains; the
// bits are handed out from the lowest free position (total_bits) upwards in
// mode order, within the key width (64 * kw bits, the reference's LIT).
// Idempotent: the fields, and with them the result, do not change.
int apply_cr_bits(TensorImpl *t, hipStream_t s) {
    if (t->nprtn < 1 || t->prtn_ptr.size() != (size_t)t->nprtn + 1) {
        set_error("conflict-resolution bits need the reference partitions");
        return ALTO_ERR_ARG;
    }
    constexpr uint64_t kMinNumFibers = 512, kMinFiberReuse = 4;
    int free_bits = 64 * t->kw - t->total_bits;
    CrArgs a{};
    uint64_t rows = 0, cm[kMaxModes * 2] = {};
    for (int m = 0; m < t->nmode; ++m) {
        const uint64_t reuse = t->nnz / t->dims[m];
        if (t->dims[m] > kMinNumFibers && reuse <= kMinFiberReuse && free_bits > 0) {
            if (t->dims[m] >= (1ULL << 32)) {  // the owner arrays hold one entry per row
                set_error("conflict-resolution bits: mode %d has %llu rows (at most 2^32 - 1)", m,
                          (unsigned long long)t->dims[m]);
                return ALTO_ERR_UNSUPPORTED;
            }
            const int bit = 64 * t->kw - free_bits;
            if (bit >= 64 * t->pkw) {
                set_error("conflict-resolution bit %d of mode %d lies beyond the %d-bit packed keys", bit, m,
                          64 * t->pkw);
                return ALTO_ERR_UNSUPPORTED;
            }
            cm[m * t->kw + (bit >> 6)] = 1ULL << (bit & 63);
            a.shift[a.n] = t->pack.shift[m];
            a.width[a.n] = t->pack.width[m];
            a.bit[a.n] = bit;
            a.off[a.n] = rows;
            rows += t->dims[m];
            ++a.n;
            --free_bits;
        }
    }
    if (a.n == 0 || t->nnz == 0) {
        memcpy(t->cr_masks, cm, sizeof(cm));
        return ALTO_OK;
    }
    DevBuf own;
    ALTO_TRY(own.alloc(sizeof(uint32_t) * 2 * rows));
    uint32_t *plo = own.as<uint32_t>(), *phi = plo + rows;
    ALTO_HIP_TRY(hipMemsetAsync(plo, 0xFF, sizeof(uint32_t) * rows, s));
    ALTO_HIP_TRY(hipMemsetAsync(phi, 0, sizeof(uint32_t) * rows, s));
    const uint64_t per = (t->nnz + t->nprtn - 1) / t->nprtn;
    k_cr_owner<<<cdiv(t->nnz, 256), 256, 0, s>>>(t->pkeys.as<uint64_t>(), t->pkw, t->nnz, per, a, plo, phi);
    ALTO_LAUNCH_CHECK();
    k_cr_set<<<cdiv(t->nnz, 256), 256, 0, s>>>(t->keys.as<uint64_t>(), t->kw, t->pkeys.as<uint64_t>(), t->pkw,
                                               t->nnz, a, plo, phi);
    ALTO_LAUNCH_CHECK();
    ALTO_HIP_TRY(hipStreamSynchronize(s));  // the owner arrays are released on return
    memcpy(t->cr_masks, cm, sizeof(cm));
    return ALTO_OK;
}

}  // namespace alto

using namespace alto;

extern "C" {

int alto_hip_tensor_cr_bits(alto_hip_tensor t, uint64_t *cr_masks, void *stream) {
    if (!t) return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    ALTO_TRY(apply_cr_bits(&t->impl, to_stream(stream)));
    if (cr_masks) memcpy(cr_masks, t->impl.cr_masks, sizeof(uint64_t) * t->impl.nmode * t->impl.kw);
    return ALTO_OK;
}

int alto_hip_tensor_update(alto_hip_tensor t, const alto_hip_coo *coo, int nprtn, void *stream) {
    if (!t || !coo || nprtn < 1) return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    return build_alto(&t->impl, coo, nprtn, 0, true, to_stream(stream));
}

int alto_hip_tensor_create(const alto_hip_coo *coo, int nprtn, int key_words, void *stream,
                           alto_hip_tensor *out) {
    if (!coo || !out || nprtn < 1) return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    auto holder = new alto_hip_tensor_s;
    int st = build_alto(&holder->impl, coo, nprtn, key_words, true, to_stream(stream));
    if (st != ALTO_OK) {
        delete holder;
        return st;
    }
    *out = holder;
    return ALTO_OK;
}

int alto_hip_tensor_destroy(alto_hip_tensor t) {
    delete t;
    return ALTO_OK;
}

int alto_hip_tensor_info(alto_hip_tensor t, int *nmode, uint64_t *nnz, int *key_words, int *nprtn,
                         int *total_bits) {
    if (!t) return ALTO_ERR_ARG;
    if (nmode) *nmode = t->impl.nmode;
    if (nnz) *nnz = t->impl.nnz;
    if (key_words) *key_words = t->impl.kw;
    if (nprtn) *nprtn = t->impl.nprtn;
    if (total_bits) *total_bits = t->impl.total_bits;
    return ALTO_OK;
}

int alto_hip_tensor_dims(alto_hip_tensor t, uint64_t *dims) {
    if (!t || !dims) return ALTO_ERR_ARG;
    memcpy(dims, t->impl.dims, sizeof(uint64_t) * t->impl.nmode);
    return ALTO_OK;
}

int alto_hip_tensor_masks(alto_hip_tensor t, uint64_t *masks) {
    if (!t || !masks) return ALTO_ERR_ARG;
    memcpy(masks, t->impl.masks, sizeof(uint64_t) * t->impl.nmode * t->impl.kw);
    return ALTO_OK;
}

int alto_hip_tensor_keys(alto_hip_tensor t, uint64_t *keys) {
    if (!t || !keys) return ALTO_ERR_ARG;
    return alto_hip_memcpy_d2h(keys, t->impl.keys.p, sizeof(uint64_t) * t->impl.kw * t->impl.nnz,
                               nullptr);
}

int alto_hip_tensor_packed_keys(alto_hip_tensor t, uint64_t *keys, int *pos) {
    if (!t) return ALTO_ERR_ARG;
    const TensorImpl &ti = t->impl;
    if (pos)
        for (int m = 0; m < ti.nmode; ++m) pos[m] = ti.pack.shift[m];
    if (!keys || ti.nnz == 0) return ALTO_OK;
    if (ti.pkw == ti.kw)
        return alto_hip_memcpy_d2h(keys, ti.pkeys.p, sizeof(uint64_t) * ti.kw * ti.nnz, nullptr);
    // forced two-word keys over a one-word layout: zero high words
    std::vector<uint64_t> lo(ti.nnz);
    ALTO_TRY(alto_hip_memcpy_d2h(lo.data(), ti.pkeys.p, sizeof(uint64_t) * ti.nnz, nullptr));
    for (uint64_t i = 0; i < ti.nnz; ++i) {
        keys[2 * i] = lo[i];
        keys[2 * i + 1] = 0;
    }
    return ALTO_OK;
}

int alto_hip_tensor_vals(alto_hip_tensor t, double *vals) {
    if (!t || !vals) return ALTO_ERR_ARG;
    return alto_hip_memcpy_d2h(vals, t->impl.vals64.p, sizeof(double) * t->impl.nnz, nullptr);
}

int alto_hip_tensor_perm(alto_hip_tensor t, uint32_t *perm) {
    if (!t || !perm) return ALTO_ERR_ARG;
    return alto_hip_memcpy_d2h(perm, t->impl.perm.p, sizeof(uint32_t) * t->impl.nnz, nullptr);
}

int alto_hip_tensor_partitions(alto_hip_tensor t, uint64_t *prtn_ptr, uint64_t *intervals) {
    if (!t) return ALTO_ERR_ARG;
    if (prtn_ptr) memcpy(prtn_ptr, t->impl.prtn_ptr.data(), sizeof(uint64_t) * t->impl.prtn_ptr.size());
    if (intervals)
        memcpy(intervals, t->impl.intervals.data(), sizeof(uint64_t) * t->impl.intervals.size());
    return ALTO_OK;
}

int alto_hip_tensor_device_arrays(alto_hip_tensor t, const uint64_t **keys, const double **vals) {
    if (!t) return ALTO_ERR_ARG;
    if (keys) *keys = t->impl.keys.as<uint64_t>();
    if (vals) *vals = t->impl.vals64.as<double>();
    return ALTO_OK;
}

}  // extern "C"
