// StreamingSparseTensor on the device (src/streaming_sptensor.cpp:69-264 with
// tt_sort, src/sort.cpp:477-519, and next_dynamic_batch at timeslice limit 1):
//   1. order the nonzeros lexicographically by (s, s+1, ..., s-1) mod N, ties
//      in input order: stable LSD radix sorts of composite keys of at most 128
//      bits, the least significant group of modes first, carrying the
//      permutation;
//   2. relabel every non-stream mode by first appearance in that order: a
//      stable sort of the mode's values carrying positions gives each value's
//      first position (its run head); marking first positions and an
//      exclusive scan over positions gives label = number of distinct values
//      that appear earlier;
//   3. slices: one per time index (sorted, so contiguous); a time index below
//      the largest one without nonzeros is an error, as the reference asserts.
//      The model dims after slice k are the running maximum of label + 1,
//      i.e. the number of first appearances before the slice's end (the same
//      scan, read at the slice ends).
#include <algorithm>
#include <vector>

#include "sort.hpp"
#include "stream.hpp"

namespace alto {

struct ModePtrs {
    const uint64_t *p[kMaxModes];
};
struct ModeDims {
    uint64_t d[kMaxModes];
};

__global__ __launch_bounds__(256) void k_st_check(ModePtrs ci, int N, uint64_t n, ModeDims dims,
                                                  int *__restrict__ bad) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int m = 0; m < N; ++m)
        if (ci.p[m][i] >= dims.d[m]) *bad = 1;
}

// Composite key of the modes order[0..cnt) (most significant first) of
// nonzero perm[i]: field q occupies `width[q]` bits.
struct KeyGroup {
    int cnt;
    int mode[kMaxModes];
    int width[kMaxModes];
};

__global__ __launch_bounds__(256) void k_st_key(ModePtrs ci, KeyGroup g, const uint32_t *__restrict__ perm,
                                                uint64_t n, uint64_t *__restrict__ k0,
                                                uint64_t *__restrict__ k1) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t src = perm[i];
    unsigned __int128 key = 0;
    for (int q = 0; q < g.cnt; ++q) {
        key <<= g.width[q];
        key |= (unsigned __int128)ci.p[g.mode[q]][src];
    }
    k0[i] = (uint64_t)key;
    if (k1) k1[i] = (uint64_t)(key >> 64);
}

__global__ __launch_bounds__(256) void k_st_gather_u64(const uint64_t *__restrict__ src,
                                                       const uint32_t *__restrict__ perm, uint64_t n,
                                                       uint64_t *__restrict__ dst) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

__global__ __launch_bounds__(256) void k_st_gather_f64(const double *__restrict__ src,
                                                       const uint32_t *__restrict__ perm, uint64_t n,
                                                       double *__restrict__ dst) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// Over the values sorted stably with their positions: head[j] = 1 at the
// first element of each run of equal values; F[pos[j]] = head[j] (F has
// n + 1 entries, F[n] = 0).
__global__ __launch_bounds__(256) void k_st_heads(const uint64_t *__restrict__ v,
                                                  const uint32_t *__restrict__ pos, uint64_t n,
                                                  uint32_t *__restrict__ head, uint32_t *__restrict__ F) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > n) return;
    if (j == n) {
        F[n] = 0u;
        return;
    }
    const uint32_t h = (j == 0 || v[j] != v[j - 1]) ? 1u : 0u;
    head[j] = h;
    F[pos[j]] = h;
}

// runs[r] = first position of run r (the head's position); after the
// exclusive scan of head, run id of element j = scan[j] + head[j] - 1.
__global__ __launch_bounds__(256) void k_st_run_first(const uint32_t *__restrict__ head_scan,
                                                      const uint64_t *__restrict__ v,
                                                      const uint32_t *__restrict__ pos, uint64_t n,
                                                      uint32_t *__restrict__ run_first) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (j == 0 || v[j] != v[j - 1]) run_first[head_scan[j]] = pos[j];
}

// label of the element at sorted index j = Fs[first position of its run];
// written at its position.
__global__ __launch_bounds__(256) void k_st_label(const uint32_t *__restrict__ head_scan,
                                                  const uint64_t *__restrict__ v,
                                                  const uint32_t *__restrict__ pos,
                                                  const uint32_t *__restrict__ run_first,
                                                  const uint32_t *__restrict__ Fs, uint64_t n,
                                                  uint64_t *__restrict__ out) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t head = (j == 0 || v[j] != v[j - 1]) ? 1u : 0u;
    const uint32_t run = head_scan[j] + head - 1u;
    out[pos[j]] = Fs[run_first[run]];
}

// start[t] = first position of time index t in the sorted time indices (run
// heads); time indices without nonzeros keep the fill value.
__global__ __launch_bounds__(256) void k_st_time_starts(const uint64_t *__restrict__ t, uint64_t n,
                                                        uint64_t *__restrict__ start) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (i == 0 || t[i] != t[i - 1])) start[t[i]] = i;
}

__global__ __launch_bounds__(256) void k_st_fill_u64(uint64_t *__restrict__ p, uint64_t n, uint64_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_st_at(const uint32_t *__restrict__ Fs,
                                               const uint64_t *__restrict__ ends, uint64_t k,
                                               uint32_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) out[i] = Fs[ends[i]];
}

__global__ void k_st_zero_i(int *p) {
    if (threadIdx.x == 0) *p = 0;
}

static int num_bits(uint64_t d) {
    int b = 0;
    while (b < 64 && (1ULL << b) < d) ++b;
    return b;
}

int stream_prepare(const alto_hip_coo *coo, int s, StreamInput *out, hipStream_t st) {
    const int N = coo->nmode;
    const uint64_t n = coo->nnz;
    if (n == 0) {
        out->slices.clear();
        return ALTO_OK;
    }
    if (n >= (1ULL << 32)) return ALTO_ERR_UNSUPPORTED;
    // inputs on the device
    DevBuf up[kMaxModes], upv;
    ModePtrs ci{};
    const double *vin = coo->vals;
    for (int m = 0; m < N; ++m) {
        ci.p[m] = coo->cidx[m];
        if (!coo->on_device) {
            ALTO_TRY(up[m].alloc(sizeof(uint64_t) * n));
            ALTO_HIP_TRY(hipMemcpyAsync(up[m].p, coo->cidx[m], sizeof(uint64_t) * n, hipMemcpyHostToDevice, st));
            ci.p[m] = up[m].as<uint64_t>();
        }
    }
    if (!coo->on_device) {
        ALTO_TRY(upv.alloc(sizeof(double) * n));
        ALTO_HIP_TRY(hipMemcpyAsync(upv.p, coo->vals, sizeof(double) * n, hipMemcpyHostToDevice, st));
        vin = upv.as<double>();
    }
    {
        DevBuf bad;
        ModeDims md{};
        for (int m = 0; m < N; ++m) md.d[m] = coo->dims[m];
        ALTO_TRY(bad.alloc(sizeof(int)));
        k_st_zero_i<<<1, 64, 0, st>>>(bad.as<int>());
        k_st_check<<<cdiv(n, 256), 256, 0, st>>>(ci, N, n, md, bad.as<int>());
        ALTO_LAUNCH_CHECK();
        int hb = 0;
        ALTO_HIP_TRY(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
        ALTO_HIP_TRY(hipStreamSynchronize(st));
        if (hb) {
            set_error("a coordinate is not below its mode size");
            return ALTO_ERR_ARG;
        }
    }
    // 1. lexicographic order by (s, s+1, ...) mod N: groups of <= 128 key
    // bits, least significant group sorted first
    DevBuf perm, k0, k1;
    SortScratch ws;
    ALTO_TRY(perm.alloc(sizeof(uint32_t) * n));
    ALTO_TRY(iota_u32(perm.as<uint32_t>(), n, st));
    ALTO_TRY(k0.alloc(sizeof(uint64_t) * n));
    {
        int lex[kMaxModes], w[kMaxModes];
        for (int q = 0; q < N; ++q) {
            lex[q] = (s + q) % N;
            w[q] = num_bits(coo->dims[lex[q]]);
        }
        int hi = N;  // modes lex[lo .. hi) form the next group
        while (hi > 0) {
            int lo = hi, bits = 0;
            while (lo > 0 && bits + w[lo - 1] <= 128) bits += w[--lo];
            if (lo == hi) return ALTO_ERR_UNSUPPORTED;  // cannot happen: widths <= 64
            if (bits > 0) {
                KeyGroup g{};
                g.cnt = hi - lo;
                for (int q = 0; q < g.cnt; ++q) {
                    g.mode[q] = lex[lo + q];
                    g.width[q] = w[lo + q];
                }
                const bool wide = bits > 64;
                if (wide && !k1.p) ALTO_TRY(k1.alloc(sizeof(uint64_t) * n));
                k_st_key<<<cdiv(n, 256), 256, 0, st>>>(ci, g, perm.as<uint32_t>(), n, k0.as<uint64_t>(),
                                                       wide ? k1.as<uint64_t>() : nullptr);
                ALTO_LAUNCH_CHECK();
                ALTO_TRY(radix_sort_pairs(k0.as<uint64_t>(), wide ? k1.as<uint64_t>() : nullptr,
                                          perm.as<uint32_t>(), n, bits, ws, st));
            }
            hi = lo;
        }
    }
    // sorted coordinates and values
    for (int m = 0; m < N; ++m) {
        ALTO_TRY(out->cid[m].alloc(sizeof(uint64_t) * n));
        k_st_gather_u64<<<cdiv(n, 256), 256, 0, st>>>(ci.p[m], perm.as<uint32_t>(), n, out->cid[m].as<uint64_t>());
    }
    ALTO_TRY(out->vals.alloc(sizeof(double) * n));
    k_st_gather_f64<<<cdiv(n, 256), 256, 0, st>>>(vin, perm.as<uint32_t>(), n, out->vals.as<double>());
    ALTO_LAUNCH_CHECK();
    for (int m = 0; m < N; ++m) up[m].release();
    upv.release();
    k1.release();

    // 3a. time slices from the sorted time indices
    uint64_t tmax = 0;
    ALTO_HIP_TRY(hipMemcpyAsync(&tmax, out->cid[s].as<uint64_t>() + n - 1, sizeof(uint64_t),
                                hipMemcpyDeviceToHost, st));
    ALTO_HIP_TRY(hipStreamSynchronize(st));
    const uint64_t T = tmax + 1;
    if (T > n) {  // more time indices than nonzeros: one of them is empty
        set_error("a time slice has no nonzeros (%llu time indices, %llu nonzeros)", (unsigned long long)T,
                  (unsigned long long)n);
        return ALTO_ERR_EMPTY_SLICE;
    }
    std::vector<uint64_t> hs(T);
    {
        DevBuf starts;
        ALTO_TRY(starts.alloc(sizeof(uint64_t) * T));
        k_st_fill_u64<<<cdiv(T, 256), 256, 0, st>>>(starts.as<uint64_t>(), T, UINT64_MAX);
        k_st_time_starts<<<cdiv(n, 256), 256, 0, st>>>(out->cid[s].as<uint64_t>(), n, starts.as<uint64_t>());
        ALTO_LAUNCH_CHECK();
        ALTO_HIP_TRY(hipMemcpyAsync(hs.data(), starts.p, sizeof(uint64_t) * T, hipMemcpyDeviceToHost, st));
        ALTO_HIP_TRY(hipStreamSynchronize(st));
    }
    std::vector<uint32_t> hc(T);
    for (uint64_t t = 0; t < T; ++t) {
        if (hs[t] == UINT64_MAX) {
            set_error("time slice %llu has no nonzeros", (unsigned long long)t);
            return ALTO_ERR_EMPTY_SLICE;
        }
        hc[t] = (uint32_t)((t + 1 < T ? hs[t + 1] : n) - hs[t]);
    }
    std::vector<uint64_t> ends(T);
    out->slices.assign(T, Slice{});
    {
        uint64_t a = 0;
        for (uint64_t t = 0; t < T; ++t) {
            out->slices[t].start = a;
            a += hc[t];
            out->slices[t].end = a;
            ends[t] = a;
            out->slices[t].dims[s] = 1;
            out->slices[t].t0 = t;
        }
    }
    DevBuf dends, dcnt;
    ALTO_TRY(dends.alloc(sizeof(uint64_t) * T));
    ALTO_TRY(dcnt.alloc(sizeof(uint32_t) * T));
    ALTO_HIP_TRY(hipMemcpyAsync(dends.p, ends.data(), sizeof(uint64_t) * T, hipMemcpyHostToDevice, st));

    // 2. first-appearance labels of every non-stream mode; 3b. slice dims
    DevBuf v, pos, head, F, run_first, scan_tmp;
    ALTO_TRY(v.alloc(sizeof(uint64_t) * n));
    ALTO_TRY(pos.alloc(sizeof(uint32_t) * n));
    ALTO_TRY(head.alloc(sizeof(uint32_t) * n));
    ALTO_TRY(F.alloc(sizeof(uint32_t) * (n + 1)));
    ALTO_TRY(run_first.alloc(sizeof(uint32_t) * n));
    std::vector<uint32_t> hd(T);
    for (int m = 0; m < N; ++m) {
        if (m == s) continue;
        ALTO_HIP_TRY(hipMemcpyAsync(v.p, out->cid[m].p, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, st));
        ALTO_TRY(iota_u32(pos.as<uint32_t>(), n, st));
        ALTO_TRY(radix_sort_pairs(v.as<uint64_t>(), nullptr, pos.as<uint32_t>(), n, num_bits(coo->dims[m]), ws, st));
        k_st_heads<<<cdiv(n + 1, 256), 256, 0, st>>>(v.as<uint64_t>(), pos.as<uint32_t>(), n, head.as<uint32_t>(),
                                                     F.as<uint32_t>());
        ALTO_LAUNCH_CHECK();
        ALTO_TRY(exclusive_scan_u32(head.as<uint32_t>(), n, scan_tmp, st));
        ALTO_TRY(exclusive_scan_u32(F.as<uint32_t>(), n + 1, scan_tmp, st));
        k_st_run_first<<<cdiv(n, 256), 256, 0, st>>>(head.as<uint32_t>(), v.as<uint64_t>(), pos.as<uint32_t>(), n,
                                                     run_first.as<uint32_t>());
        k_st_label<<<cdiv(n, 256), 256, 0, st>>>(head.as<uint32_t>(), v.as<uint64_t>(), pos.as<uint32_t>(),
                                                 run_first.as<uint32_t>(), F.as<uint32_t>(), n,
                                                 out->cid[m].as<uint64_t>());
        k_st_at<<<cdiv(T, 256), 256, 0, st>>>(F.as<uint32_t>(), dends.as<uint64_t>(), T, dcnt.as<uint32_t>());
        ALTO_LAUNCH_CHECK();
        ALTO_HIP_TRY(hipMemcpyAsync(hd.data(), dcnt.p, sizeof(uint32_t) * T, hipMemcpyDeviceToHost, st));
        ALTO_HIP_TRY(hipStreamSynchronize(st));
        for (uint64_t t = 0; t < T; ++t) out->slices[t].dims[m] = hd[t];
    }
    return ALTO_OK;
}

}  // namespace alto

using namespace alto;

extern "C" int alto_hip_stream_prepare(const alto_hip_coo *coo, int stream_mode, uint64_t *const *out_cidx,
                                       double *out_vals, uint64_t max_slices, uint64_t *out_nslices,
                                       uint64_t *out_slice_ptr, uint64_t *out_slice_dims, void *stream) {
    if (!coo || !coo->dims || !coo->cidx || !coo->vals || !out_nslices) return ALTO_ERR_ARG;
    const int N = coo->nmode;
    if (N < 2 || N > kMaxModes || stream_mode < 0 || stream_mode >= N) return ALTO_ERR_ARG;
    for (int m = 0; m < N; ++m)
        if (coo->dims[m] == 0) return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    hipStream_t st = to_stream(stream);
    StreamInput sin;
    ALTO_TRY(stream_prepare(coo, stream_mode, &sin, st));
    const uint64_t n = coo->nnz, T = sin.slices.size();
    *out_nslices = T;
    if (n) {
        for (int m = 0; m < N; ++m)
            if (out_cidx && out_cidx[m])
                ALTO_HIP_TRY(hipMemcpyAsync(out_cidx[m], sin.cid[m].p, sizeof(uint64_t) * n,
                                            coo->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
        if (out_vals)
            ALTO_HIP_TRY(hipMemcpyAsync(out_vals, sin.vals.p, sizeof(double) * n,
                                        coo->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    }
    ALTO_HIP_TRY(hipStreamSynchronize(st));
    if (T > max_slices) return ALTO_OK;  // slice arrays too small: the count only
    for (uint64_t k = 0; k < T; ++k) {
        if (out_slice_ptr) {
            out_slice_ptr[k] = sin.slices[k].start;
            out_slice_ptr[k + 1] = sin.slices[k].end;
        }
        if (out_slice_dims)
            for (int m = 0; m < N; ++m) out_slice_dims[k * N + m] = sin.slices[k].dims[m];
    }
    return ALTO_OK;
}

namespace alto {

// next_dynamic_batch (src/streaming_sptensor.cpp:196-264): from the first
// time index not yet taken, take whole time slices while fewer than
// nnz_threshold nonzeros are taken, stopping after timeslice_limit of them.
int graph_signature(hipGraph_t g, GraphSig *sig) {
    size_t n = 0;
    ALTO_HIP_TRY(hipGraphGetNodes(g, nullptr, &n));
    std::vector<hipGraphNode_t> nodes(n);
    if (n) ALTO_HIP_TRY(hipGraphGetNodes(g, nodes.data(), &n));
    sig->nodes.clear();
    sig->nodes.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        hipGraphNodeType ty;
        ALTO_HIP_TRY(hipGraphNodeGetType(nodes[i], &ty));
        const void *fn = nullptr;
        if (ty == hipGraphNodeTypeKernel) {
            hipKernelNodeParams kp{};
            ALTO_HIP_TRY(hipGraphKernelNodeGetParams(nodes[i], &kp));
            fn = kp.func;
        }
        sig->nodes.emplace_back((int)ty, fn);
    }
    return ALTO_OK;
}

std::vector<Slice> stream_batches(const std::vector<Slice> &slices, int s, uint64_t nnz_threshold,
                                  uint64_t timeslice_limit) {
    const uint64_t thr = nnz_threshold ? nnz_threshold : ~0ULL;
    const uint64_t lim = timeslice_limit ? timeslice_limit : 1;
    std::vector<Slice> out;
    size_t k = 0;
    while (k < slices.size()) {
        Slice b = slices[k];
        uint64_t taken = 0, cnt = 0;
        do {
            taken += slices[k].end - slices[k].start;
            ++cnt;
            ++k;
        } while (k < slices.size() && taken < thr && cnt < lim);
        const Slice &last = slices[k - 1];
        b.end = last.end;
        for (int m = 0; m < kMaxModes; ++m) b.dims[m] = last.dims[m];
        b.dims[s] = cnt;
        out.push_back(b);
    }
    return out;
}

}  // namespace alto
