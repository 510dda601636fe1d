// Hot-row MTTKRP: one 1024-thread workgroup per CU with the most popular
// factor rows of the gathered modes resident in LDS.
//
// Reference: the output-ordered (op_sched / oidx) MTTKRP variant of
// mttkrp_alto_par (include/alto.hpp:826-908, 1708-1801); per nonzero
// row = val * prod_{m != n} A_m[idx_m], accumulated into A_n[idx_n].
//
// Why: on skewed tensors the MTTKRP is bound by the rate at which the vector
// memory path serves factor-row gathers, and most gathers go to a few hundred
// rows per mode (Zipf(1.1) over 500,000 rows puts ~60% of the draws on the
// 256 most popular rows).  Serving those from LDS leaves the vector memory
// path to the rest.
//
// Layout: per mode m, the rows are ranked by how many nonzeros use them
// (ensure_hot_modes).  For a target mode n the LDS budget is split greedily
// over the gathered modes by that count, giving K_m hot slots per mode, and
// the schedule's packed keys are re-encoded once (ensure_hot_schedule): a
// gathered index i with popularity rank r < K_m becomes the code I_m + r, any
// other index keeps its value, so `code >= I_m` selects LDS slot code - I_m.
// Fields that need one more bit for the codes are widened (the re-encoded
// keys have their own PackInfo).
//
// Kernel: grid = one workgroup per CU, each walking a contiguous share of
// the schedule (XCD-contiguous), split into one contiguous range per lane
// group of L = R / VPL lanes.  A group loads 16 entries at a time (coalesced
// keys and values, the next 16 prefetched), each lane decodes 16 / L of them,
// and the decoded fields are broadcast inside the group with DPP (L <= 4) or
// ds_swizzle (L = 8, 16), so no LDS is spent on records.  Runs of one output
// row accumulate in registers; rows that start and end inside a range are
// owned and stored directly.  The first and last runs of every group are
// combined across the workgroup in FP64 in LDS (the hot tables' space, after
// a barrier); rows that start and end inside the workgroup are owned, and the
// workgroup's first and last rows go to a boundary buffer that launch_bnd_merge
// sums in schedule order in FP64.  No atomics: the result does not depend on
// scheduling.
#include <algorithm>
#include <vector>

#include "mttkrp_hot.hpp"

namespace alto {

// ---------------------------------------------------------------------------
// Popularity of every mode's rows.

// cnt[row] += number of nonzeros with that index; one atomic per distinct
// row in a wave (hot rows repeat inside a wave).
__global__ __launch_bounds__(256) void k_hot_count(const uint64_t *__restrict__ pkeys, int pkw,
                                                   uint64_t n, int sh, int w,
                                                   uint32_t *__restrict__ cnt) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < n;
    uint32_t row = 0;
    if (act) {
        const uint64_t lo = pkeys[i * pkw], hi = pkw == 2 ? pkeys[i * pkw + 1] : 0;
        row = (uint32_t)extract_field(lo, hi, sh, w);
    }
    const int lane = threadIdx.x & 63;
    uint64_t pending = __ballot(act);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)row, leader);
        const uint64_t same = __ballot(row == r) & pending;
        if (lane == leader) atomicAdd(cnt + r, (uint32_t)__popcll(same));
        pending &= ~same;
    }
}

// Sort key: descending count (ties: ascending row, the sort is stable).
__global__ __launch_bounds__(256) void k_hot_keys(const uint32_t *__restrict__ cnt, uint64_t n,
                                                  uint64_t *__restrict__ key, uint32_t *__restrict__ row) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = 0xFFFFFFFFull - cnt[i];
    row[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// The k most popular rows: list[s] = row of rank s, rank[row] = s, and the
// counts in popularity order.
__global__ __launch_bounds__(256) void k_hot_select(const uint64_t *__restrict__ key,
                                                    const uint32_t *__restrict__ row, uint32_t k,
                                                    uint32_t *__restrict__ list, uint32_t *__restrict__ rank,
                                                    uint32_t *__restrict__ cnt) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= k) return;
    const uint32_t r = row[s];
    list[s] = r;
    rank[r] = s;
    cnt[s] = (uint32_t)(0xFFFFFFFFull - key[s]);
}

struct HotEnc {
    int nmode;
    int shift[kMaxModes], width[kMaxModes];    // source fields (tensor layout)
    int nshift[kMaxModes], nwidth[kMaxModes];  // re-encoded fields
    uint32_t K[kMaxModes];                     // hot slots per mode (0: none)
    uint32_t dim[kMaxModes];
    const uint32_t *rank[kMaxModes];
};

__global__ __launch_bounds__(256) void k_hot_encode(const uint64_t *__restrict__ src, int pkw, uint64_t n,
                                                    HotEnc he, uint64_t *__restrict__ dst) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t lo = src[i * pkw], hi = pkw == 2 ? src[i * pkw + 1] : 0;
    uint64_t olo = 0, ohi = 0;
    for (int m = 0; m < he.nmode; ++m) {
        uint64_t v = extract_field(lo, hi, he.shift[m], he.width[m]);
        if (he.K[m]) {
            const uint32_t r = he.rank[m][v];
            if (r < he.K[m]) v = (uint64_t)he.dim[m] + r;
        }
        const int sh = he.nshift[m], w = he.nwidth[m];
        if (sh < 64) {
            olo |= v << sh;
            if (sh > 0 && sh + w > 64) ohi |= v >> (64 - sh);
        } else {
            ohi |= v << (sh - 64);
        }
    }
    dst[i * pkw] = olo;
    if (pkw == 2) dst[i * pkw + 1] = ohi;
}

// ---------------------------------------------------------------------------
// Host side.

static int num_cus() {
    static int n = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
            return 256;
        return v;
    }();
    return n;
}

static int elem_bytes(int prec) { return prec == ALTO_PREC_F64 ? 8 : prec == ALTO_PREC_F32 ? 4 : 2; }

bool hot_supported(const TensorImpl *t, int prec, int rank) {
    if (t->nmode < 3 || t->nmode > 5) return false;
    if (rank != 16 && rank != 32 && rank != 64) return false;
    const uint64_t rb = (uint64_t)rank * elem_bytes(prec);
    for (int m = 0; m < t->nmode; ++m)
        if (t->dims[m] * rb >= (1ULL << 31) || t->dims[m] + kHotMax >= (1ULL << 31)) return false;
    return true;
}

int ensure_hot_modes(TensorImpl *t, hipStream_t s) {
    if (t->hot.built) return ALTO_OK;
    DevBuf cnt, key, row, cnt_out;
    for (int m = 0; m < t->nmode; ++m) {
        const uint64_t dim = t->dims[m];
        const uint32_t k = (uint32_t)std::min<uint64_t>(kHotMax, dim);
        ALTO_TRY(cnt.alloc(sizeof(uint32_t) * dim));
        ALTO_TRY(key.alloc(sizeof(uint64_t) * dim));
        ALTO_TRY(row.alloc(sizeof(uint32_t) * dim));
        ALTO_TRY(cnt_out.alloc(sizeof(uint32_t) * k));
        ALTO_TRY(t->hot.list[m].alloc(sizeof(uint32_t) * k));
        ALTO_TRY(t->hot.rank[m].alloc(sizeof(uint32_t) * dim));
        k_fill_u32<<<cdiv(dim, 256), 256, 0, s>>>(cnt.as<uint32_t>(), dim, 0u);
        k_fill_u32<<<cdiv(dim, 256), 256, 0, s>>>(t->hot.rank[m].as<uint32_t>(), dim, kNone);
        ALTO_LAUNCH_CHECK();
        if (t->nnz) {
            k_hot_count<<<cdiv(t->nnz, 256), 256, 0, s>>>(t->pkeys.as<uint64_t>(), t->pkw, t->nnz,
                                                          t->pack.shift[m], t->pack.width[m],
                                                          cnt.as<uint32_t>());
            ALTO_LAUNCH_CHECK();
        }
        k_hot_keys<<<cdiv(dim, 256), 256, 0, s>>>(cnt.as<uint32_t>(), dim, key.as<uint64_t>(),
                                                  row.as<uint32_t>());
        ALTO_LAUNCH_CHECK();
        ALTO_TRY(radix_sort_pairs(key.as<uint64_t>(), nullptr, row.as<uint32_t>(), dim, 32, t->sort_ws, s));
        k_hot_select<<<cdiv(k, 256), 256, 0, s>>>(key.as<uint64_t>(), row.as<uint32_t>(), k,
                                                  t->hot.list[m].as<uint32_t>(),
                                                  t->hot.rank[m].as<uint32_t>(), cnt_out.as<uint32_t>());
        ALTO_LAUNCH_CHECK();
        t->hot.cnt[m].assign(k, 0);
        ALTO_HIP_TRY(hipMemcpyAsync(t->hot.cnt[m].data(), cnt_out.p, sizeof(uint32_t) * k,
                                    hipMemcpyDeviceToHost, s));
        ALTO_HIP_TRY(hipStreamSynchronize(s));  // the scratch buffers are reused / released
    }
    t->hot.built = true;
    return ALTO_OK;
}

static int bits_for(uint64_t x) {  // bits to hold values 0..x
    int b = 0;
    while (b < 64 && (x >> b) != 0) ++b;
    return b;
}

// Greedy split of the LDS rows over the gathered modes by entry count, then
// the re-encoded key layout (widened fields for hot codes; modes whose codes
// would not fit in the key words get no slots).
static void pick_slots(const TensorImpl *t, int target, int rb, HotSched *h) {
    h->nq = 0;
    for (int m = 0; m < t->nmode; ++m)
        if (m != target) h->qmode[h->nq++] = m;
    uint32_t p[4] = {0, 0, 0, 0};
    const uint32_t S = (uint32_t)(kHotLds / rb);
    for (uint32_t s = 0; s < S; ++s) {
        int best = -1;
        uint32_t bc = 0;
        for (int q = 0; q < h->nq; ++q) {
            const std::vector<uint32_t> &c = t->hot.cnt[h->qmode[q]];
            if (p[q] < c.size() && c[p[q]] > bc) {
                bc = c[p[q]];
                best = q;
            }
        }
        if (best < 0) break;
        ++p[best];
    }
    for (int q = 0; q < h->nq; ++q) h->K[q] = p[q];
    for (;;) {
        int w[kMaxModes], tot = 0;
        for (int m = 0; m < t->nmode; ++m) w[m] = t->pack.width[m];
        for (int q = 0; q < h->nq; ++q) {
            const int m = h->qmode[q];
            if (h->K[q]) w[m] = std::max(w[m], bits_for(t->dims[m] + h->K[q] - 1));
        }
        for (int m = 0; m < t->nmode; ++m) tot += w[m];
        if (tot <= 64 * t->pkw) {
            h->hpk.nmode = t->nmode;
            int sh = 0;
            for (int m = 0; m < t->nmode; ++m) {
                h->hpk.shift[m] = sh;
                h->hpk.width[m] = w[m];
                sh += w[m];
            }
            break;
        }
        // drop the mode whose slots serve the fewest gathers
        int drop = -1;
        uint64_t least = ~0ULL;
        for (int q = 0; q < h->nq; ++q) {
            if (!h->K[q]) continue;
            uint64_t sum = 0;
            for (uint32_t i = 0; i < h->K[q]; ++i) sum += t->hot.cnt[h->qmode[q]][i];
            if (sum < least) {
                least = sum;
                drop = q;
            }
        }
        h->K[drop] = 0;
    }
    h->rb = rb;
}

int ensure_hot_schedule(TensorImpl *t, int target, int rb, hipStream_t s) {
    ModeSchedule &sc = t->sched[target];
    HotSched &h = sc.hot;
    if (h.built && h.rb == rb) return ALTO_OK;
    h.built = false;
    ALTO_TRY(ensure_hot_modes(t, s));
    pick_slots(t, target, rb, &h);
    HotEnc he{};
    he.nmode = t->nmode;
    for (int m = 0; m < t->nmode; ++m) {
        he.shift[m] = t->pack.shift[m];
        he.width[m] = t->pack.width[m];
        he.nshift[m] = h.hpk.shift[m];
        he.nwidth[m] = h.hpk.width[m];
        he.dim[m] = (uint32_t)t->dims[m];
    }
    for (int q = 0; q < h.nq; ++q) {
        const int m = h.qmode[q];
        he.K[m] = h.K[q];
        he.rank[m] = t->hot.rank[m].as<uint32_t>();
    }
    ALTO_TRY(h.keys.alloc(sizeof(uint64_t) * t->pkw * (sc.n + kSchedPad)));
    {
        const uint64_t words = 2 * (uint64_t)t->pkw * kSchedPad;  // the zero padding, as u32 words
        k_fill_u32<<<cdiv(words, 256), 256, 0, s>>>(h.keys.as<uint32_t>() + 2 * (uint64_t)t->pkw * sc.n, words,
                                                     0u);
        ALTO_LAUNCH_CHECK();
    }
    if (sc.n) {
        k_hot_encode<<<cdiv(sc.n, 256), 256, 0, s>>>(sc.pkeys.as<uint64_t>(), t->pkw, sc.n, he,
                                                     h.keys.as<uint64_t>());
        ALTO_LAUNCH_CHECK();
    }
    h.built = true;
    return ALTO_OK;
}

int run_hot(TensorImpl *t, int target, int prec, int rank, const void *const *factors, void *out,
            bool out64, bool zeroed, int shape, hipStream_t s) {
    if (!hot_supported(t, prec, rank)) return ALTO_ERR_UNSUPPORTED;
    if ((uintptr_t)out % 16) {
        set_error("the hot-row MTTKRP needs a 16-byte aligned output");
        return ALTO_ERR_UNSUPPORTED;
    }
    ModeSchedule &sc = t->sched[target];
    if (!sc.built || (prec == ALTO_PREC_F64 && !sc.has64)) return ALTO_ERR_ARG;
    if (sc.n == 0) return ALTO_OK;
    const int rb = rank * elem_bytes(prec);
    ALTO_TRY(ensure_hot_schedule(t, target, rb, s));
    const HotSched &hs = sc.hot;
    HotLaunch h{};
    h.keys = hs.keys.as<uint64_t>();
    h.vals = prec == ALTO_PREC_F64 ? sc.vals64.p : sc.vals32.p;
    h.n = sc.n;
    h.out = out;
    h.nwg = num_cus();
    h.pkw = t->pkw;
    h.nm1 = hs.nq;
    h.zeroed = zeroed;
    h.shape = shape;
    uint32_t lds = 0;
    for (int q = 0; q < hs.nq; ++q) {
        const int m = hs.qmode[q];
        h.ha.fsh[q] = hs.hpk.shift[m];
        h.ha.fw[q] = hs.hpk.width[m];
        h.ha.dim[q] = (uint32_t)t->dims[m];
        h.ha.lds[q] = lds;
        h.ha.K[q] = hs.K[q];
        h.ha.list[q] = t->hot.list[m].as<uint32_t>();
        h.ha.fac[q] = factors[m];
        lds += hs.K[q] * (uint32_t)rb;
    }
    h.ha.tsh = hs.hpk.shift[target];
    h.ha.tw = hs.hpk.width[target];
    h.hot_bytes = lds;
    const size_t nslot = 2 * (size_t)h.nwg, nl = bnd_level_slots(nslot);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t srb = al(nslot * sizeof(uint32_t)), svb = al(nslot * rank * sizeof(double));
    const size_t lrb = al(nl * sizeof(uint32_t)), lvb = al(nl * rank * sizeof(double));
    ALTO_TRY(t->hot_bnd.alloc(srb + svb + 2 * (lrb + lvb)));
    char *base = (char *)t->hot_bnd.p;
    h.bnd_row = (uint32_t *)base;
    h.bnd_val = (double *)(base + srb);
    for (int k = 0; k < 2; ++k) {
        h.lrow[k] = (uint32_t *)(base + srb + svb + k * (lrb + lvb));
        h.lval[k] = (double *)(base + srb + svb + k * (lrb + lvb) + lrb);
    }
    if (prec == ALTO_PREC_F64) return hot_rank<double, double, double>(rank, h, s);
    if (prec == ALTO_PREC_F32)
        return out64 ? hot_rank<float, float, double>(rank, h, s) : hot_rank<float, float, float>(rank, h, s);
    return out64 ? hot_rank<float, uint16_t, double>(rank, h, s) : hot_rank<float, uint16_t, float>(rank, h, s);
}

}  // namespace alto

