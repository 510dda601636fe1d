// Synthetic sparse tensors for the benchmark workloads, generated directly
// in device memory.  No reference counterpart: the reference's own generator
// (src/poisson_generator.cpp) makes Poisson CP-model tensors, while the
// measured workloads are defined by per-mode index and value distributions.
//
// Counter-based: every random draw is splitmix64(seed, stream, counter), so a
// nonzero's coordinates and value depend only on (seed, nonzero index) and the
// result is the same for any launch shape.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <numeric>
#include <random>
#include <vector>

#include "common.hpp"

namespace alto {

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t ctr) {
    return mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL)) + ctr);
}
// uniform double in [0, 1) with 53 random bits
__device__ inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

struct GenMode {
    uint64_t dim;
    int dist;              // 0 uniform, 1 zipf (table)
    const double *cdf;     // zipf cumulative probabilities (dim entries)
    const uint64_t *perm;  // zipf rank -> index
    uint64_t block_off;    // sub-block offset for the block fraction
};

struct GenArgs {
    int nmode;
    GenMode m[kMaxModes];
    uint64_t nnz, seed;
    int vdist;             // 0 U(0,1), 1 lognormal(mu, sigma), 2 N(0,1), 3 Poisson(lam)+1
    double p1, p2;
    double block_frac;     // fraction of nonzeros drawn inside the sub-block
    uint64_t block;        // sub-block edge
    const double *pois_cdf;
    int pois_n;
};

__global__ __launch_bounds__(256) void k_generate(GenArgs a, uint64_t **cidx, double *vals) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nnz) return;
    bool in_block = a.block_frac > 0.0 && u01(rnd(a.seed, 99, i)) < a.block_frac;
    for (int m = 0; m < a.nmode; ++m) {
        const GenMode &g = a.m[m];
        uint64_t r = rnd(a.seed, 1 + m, i);
        uint64_t x;
        if (in_block) {
            x = g.block_off + __umul64hi(r, a.block);
        } else if (g.dist == 1) {
            double u = u01(r);
            uint64_t lo = 0, hi = g.dim - 1;  // first k with cdf[k] > u
            while (lo < hi) {
                uint64_t mid = (lo + hi) >> 1;
                if (g.cdf[mid] > u) hi = mid;
                else lo = mid + 1;
            }
            x = g.perm[lo];
        } else {
            x = __umul64hi(r, g.dim);
        }
        cidx[m][i] = x;
    }
    double v;
    uint64_t r1 = rnd(a.seed, 50, i), r2 = rnd(a.seed, 51, i);
    if (a.vdist == 0) {
        v = u01(r1);
    } else if (a.vdist == 1 || a.vdist == 2) {
        double u1 = 1.0 - u01(r1), u2 = u01(r2);
        double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        v = a.vdist == 1 ? exp(a.p1 + a.p2 * z) : z;
    } else {
        double u = u01(r1);
        int k = 0;
        while (k < a.pois_n - 1 && a.pois_cdf[k] <= u) ++k;
        v = (double)(k + 1);
    }
    vals[i] = v;
}

// Row-gather rate: each lane group reads one 128-byte row per step at the
// row the next entry of `idx` names (a coalesced 4-byte stream, as the
// schedule's keys are).  LG = 8: 8 lanes x 16 B; LG = 4: 4 lanes x 2 x 16 B,
// the fiber kernel's FP32 rank-32 row.
template <int LG>
__global__ __launch_bounds__(256) void k_gather_probe(const float4 *__restrict__ tab,
                                                      const uint32_t *__restrict__ idx, uint64_t n,
                                                      int per, float *__restrict__ out) {
    constexpr int NL = 8 / LG;
    const uint64_t g = ((uint64_t)blockIdx.x * 256 + threadIdx.x) / LG;
    const int l = threadIdx.x % LG;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int k = 0; k < per; ++k) {
        const uint64_t i = g * per + k;
        if (i >= n) break;
        const float4 *p = tab + (uint64_t)idx[i] * 8 + l * NL;
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            const float4 v = p[j];
            a.x += v.x;
            a.y += v.y;
            a.z += v.z;
            a.w += v.w;
        }
    }
    out[g * LG + l] = a.x + a.y + a.z + a.w;
}

// Row indices of the probe: uniform over `rows`, or Zipf by the cumulative
// table cdf (rows entries) mapped through perm.
__global__ __launch_bounds__(256) void k_probe_idx(uint32_t *__restrict__ idx, uint64_t n, uint64_t rows,
                                                   const double *__restrict__ cdf,
                                                   const uint32_t *__restrict__ perm, uint64_t seed) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t r = rnd(seed, 7, i);
    if (!cdf) {
        idx[i] = (uint32_t)__umul64hi(r, rows);
        return;
    }
    const double u = u01(r);
    uint64_t lo = 0, hi = rows - 1;  // first k with cdf[k] > u
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (cdf[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    idx[i] = perm[lo];
}

}  // namespace alto

using namespace alto;

extern "C" int alto_hip_generate(int nmode, const uint64_t *dims, uint64_t nnz, uint64_t seed,
                                 const int *mode_dist, const double *zipf_s, int value_dist,
                                 double p1, double p2, double block_frac, uint64_t block,
                                 uint64_t *const *cidx_dev, double *vals_dev, void *stream) {
    if (nmode < 1 || nmode > kMaxModes || !dims || !mode_dist || !cidx_dev || !vals_dev)
        return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    hipStream_t s = to_stream(stream);
    GenArgs a{};
    a.nmode = nmode;
    a.nnz = nnz;
    a.seed = seed;
    a.vdist = value_dist;
    a.p1 = p1;
    a.p2 = p2;
    a.block_frac = block_frac;
    a.block = block;
    DevBuf cdfs[kMaxModes], perms[kMaxModes], pcdf, ptrs;
    for (int m = 0; m < nmode; ++m) {
        if (dims[m] == 0) return ALTO_ERR_ARG;
        a.m[m].dim = dims[m];
        a.m[m].dist = mode_dist[m];
        if (block_frac > 0.0) {
            if (block > dims[m]) return ALTO_ERR_ARG;
            a.m[m].block_off = rnd(seed, 200 + m, 0) % (dims[m] - block + 1);
        }
        if (mode_dist[m] == 1) {
            double sexp = zipf_s ? zipf_s[m] : 1.1;
            std::vector<double> cdf(dims[m]);
            double acc = 0.0;
            for (uint64_t k = 0; k < dims[m]; ++k) {
                acc += pow((double)(k + 1), -sexp);
                cdf[k] = acc;
            }
            for (auto &c : cdf) c /= acc;
            cdf.back() = 2.0;  // every u < 1 lands inside the table
            std::vector<uint64_t> perm(dims[m]);
            std::iota(perm.begin(), perm.end(), 0ULL);
            for (uint64_t k = dims[m] - 1; k > 0; --k) {  // seeded Fisher-Yates
                uint64_t j = rnd(seed, 300 + m, k) % (k + 1);
                std::swap(perm[k], perm[j]);
            }
            ALTO_TRY(cdfs[m].alloc(sizeof(double) * dims[m]));
            ALTO_TRY(perms[m].alloc(sizeof(uint64_t) * dims[m]));
            ALTO_HIP_TRY(hipMemcpyAsync(cdfs[m].p, cdf.data(), sizeof(double) * dims[m],
                                        hipMemcpyHostToDevice, s));
            ALTO_HIP_TRY(hipMemcpyAsync(perms[m].p, perm.data(), sizeof(uint64_t) * dims[m],
                                        hipMemcpyHostToDevice, s));
            ALTO_HIP_TRY(hipStreamSynchronize(s));
            a.m[m].cdf = cdfs[m].as<double>();
            a.m[m].perm = perms[m].as<uint64_t>();
        }
    }
    if (value_dist == 3) {
        std::vector<double> c;
        double lam = p1, pk = exp(-lam), acc = 0.0;
        for (int k = 0; k < 64; ++k) {
            acc += pk;
            c.push_back(acc);
            pk *= lam / (k + 1);
        }
        c.back() = 2.0;
        ALTO_TRY(pcdf.alloc(sizeof(double) * c.size()));
        ALTO_HIP_TRY(hipMemcpyAsync(pcdf.p, c.data(), sizeof(double) * c.size(),
                                    hipMemcpyHostToDevice, s));
        a.pois_cdf = pcdf.as<double>();
        a.pois_n = (int)c.size();
    }
    ALTO_TRY(ptrs.alloc(sizeof(uint64_t *) * nmode));
    ALTO_HIP_TRY(hipMemcpyAsync(ptrs.p, cidx_dev, sizeof(uint64_t *) * nmode, hipMemcpyHostToDevice, s));
    if (nnz) {
        k_generate<<<cdiv(nnz, 256), 256, 0, s>>>(a, ptrs.as<uint64_t *>(), vals_dev);
        ALTO_LAUNCH_CHECK();
    }
    ALTO_HIP_TRY(hipStreamSynchronize(s));
    return ALTO_OK;
}

extern "C" int alto_hip_gather_probe(uint64_t rows, uint64_t gathers, int dist, double zipf_s,
                                     int layout, uint64_t seed, float *ms, void *stream) {
    using namespace alto;
    if (!ms || rows < 1 || rows >= (1ULL << 32) || gathers < 64 || dist < 0 || dist > 2 ||
        (layout != 0 && layout != 1) || (dist >= 1 && !(zipf_s > 0.0)))
        return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    hipStream_t s = to_stream(stream);
    const int per = 64;
    const uint64_t groups = gathers / per, n = groups * per;
    const int LG = layout == 0 ? 8 : 4;
    const unsigned blocks = cdiv(groups * LG, 256);
    DevBuf tab, idx, out, cdf, perm;
    ALTO_TRY(tab.alloc(rows * 128));
    ALTO_TRY(idx.alloc(sizeof(uint32_t) * n));
    ALTO_TRY(out.alloc(sizeof(float) * (size_t)blocks * 256));
    ALTO_HIP_TRY(hipMemsetAsync(tab.p, 0, rows * 128, s));
    if (dist >= 1) {
        std::vector<double> c(rows);
        double acc = 0.0;
        for (uint64_t k = 0; k < rows; ++k) {
            acc += pow((double)(k + 1), -zipf_s);
            c[k] = acc;
        }
        for (auto &v : c) v /= acc;
        c.back() = 2.0;
        std::vector<uint32_t> pm(rows);
        std::iota(pm.begin(), pm.end(), 0u);
        if (dist == 1)  // popular rows scattered over the table; dist 2 keeps them first
            for (uint64_t k = rows - 1; k > 0; --k) std::swap(pm[k], pm[rnd(seed, 301, k) % (k + 1)]);
        ALTO_TRY(cdf.alloc(sizeof(double) * rows));
        ALTO_TRY(perm.alloc(sizeof(uint32_t) * rows));
        ALTO_HIP_TRY(hipMemcpyAsync(cdf.p, c.data(), sizeof(double) * rows, hipMemcpyHostToDevice, s));
        ALTO_HIP_TRY(hipMemcpyAsync(perm.p, pm.data(), sizeof(uint32_t) * rows, hipMemcpyHostToDevice, s));
        ALTO_HIP_TRY(hipStreamSynchronize(s));
    }
    k_probe_idx<<<cdiv(n, 256), 256, 0, s>>>(idx.as<uint32_t>(), n, rows, dist ? cdf.as<double>() : nullptr,
                                             dist ? perm.as<uint32_t>() : nullptr, seed);
    ALTO_LAUNCH_CHECK();
    auto launch = [&]() {
        if (LG == 8)
            k_gather_probe<8><<<blocks, 256, 0, s>>>(tab.as<float4>(), idx.as<uint32_t>(), n, per, out.as<float>());
        else
            k_gather_probe<4><<<blocks, 256, 0, s>>>(tab.as<float4>(), idx.as<uint32_t>(), n, per, out.as<float>());
    };
    hipEvent_t e0, e1;
    ALTO_HIP_TRY(hipEventCreate(&e0));
    ALTO_HIP_TRY(hipEventCreate(&e1));
    launch();
    ALTO_HIP_TRY(hipEventRecord(e0, s));
    const int reps = 5;
    for (int r = 0; r < reps; ++r) launch();
    ALTO_HIP_TRY(hipEventRecord(e1, s));
    ALTO_HIP_TRY(hipEventSynchronize(e1));
    ALTO_LAUNCH_CHECK();
    float t = 0.f;
    ALTO_HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / reps;
    return ALTO_OK;
}
