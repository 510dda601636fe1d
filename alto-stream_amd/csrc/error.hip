// cpd_error (src/streaming_cpd.cpp:368-464) with kruskal_norm
// (src/kruskal_model.cpp:400-442): the relative residual of a Kruskal model
// (all weights 1, as cpd_error sets them) against a sparse tensor,
//   err = sqrt(||X||^2 + ||M||^2 - 2 <X, M>) / ||X||,
// where <X, M> = sum_ir U_min[i, r] * MTTKRP_min[i, r] over the mode with the
// fewest rows (the first such mode), ||X||^2 = sum of squared stored values,
// and ||M||^2 = sum over i <= j of (2 - [i == j]) * prod_m (U_m^T U_m)[i, j].
// The MTTKRP runs over an ALTO tensor of the input (the reference uses its
// COO mttkrp_par); the sums are fixed-order two-level reductions.
#include <math.h>

#include <algorithm>
#include <mutex>

#include "dense.hpp"
#include "tensor.hpp"

namespace alto {

constexpr int kDotBlocks = 1024;

struct IdxArgs {
    const uint64_t *p[kMaxModes];
};

__global__ __launch_bounds__(256) void k_dot_part(const double *__restrict__ a,
                                                  const double *__restrict__ b, uint64_t n,
                                                  double *__restrict__ part) {
    __shared__ double sh[4];
    double s = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        s += a[i] * b[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

struct Rows {
    uint64_t r[kMaxModes];
};

__global__ __launch_bounds__(256) void k_check_coords(IdxArgs ci, int N, uint64_t nnz, Rows rows,
                                                      int *__restrict__ bad) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    for (int m = 0; m < N; ++m)
        if (ci.p[m][i] >= rows.r[m]) *bad = 1;
}

// One wave: inner and ||X||^2 from their partials, ||M||^2 from the Grams.
__global__ void k_error_final(const double *__restrict__ dot_part, const double *__restrict__ xx,
                              const double *__restrict__ grams, int nmode, int R,
                              double *__restrict__ out) {
    const int lane = threadIdx.x;
    double inner = wave_reduce_entry(dot_part, kDotBlocks, 1, 0, 0);
    double z = 0.0;
    const int RR = R * R;
    for (int e = lane; e < RR; e += 64) {
        const int i = e / R, j = e % R;
        if (j < i) continue;
        double v = 1.0;
        for (int m = 0; m < nmode; ++m) v *= grams[(size_t)m * RR + e];
        z += (i == j) ? v : 2.0 * v;
    }
    for (int o = 32; o > 0; o >>= 1) z += __shfl_down(z, o, 64);
    if (lane == 0) {
        const double X = *xx, Z = fabs(z);
        const double resid = sqrt(X + Z - 2.0 * inner);
        out[0] = resid / sqrt(X);
        out[1] = X;
        out[2] = Z;
        out[3] = inner;
    }
}

// factors: device FP64, factor m has rows[m] rows.  cidx/vals: device COO.
// out[4] (host): err, ||X||^2, ||M||^2, <X, M>.
int cpd_error_dev(int N, const uint64_t *rows, uint64_t nnz, const uint64_t *const *cidx,
                  const double *vals, const double *const *factors, int R, hipStream_t s,
                  double *out) {
    if (N < 1 || N > kMaxModes || R < 1 || R > kMaxRank) return ALTO_ERR_ARG;
    int mn = 0;
    for (int m = 1; m < N; ++m)
        if (rows[m] < rows[mn]) mn = m;
    const uint64_t I = rows[mn];
    DevBuf mt, dpart, sq, xx, G, gpart, res, bad;
    {  // every coordinate below its mode's row count
        IdxArgs ci{};
        Rows rr{};
        for (int m = 0; m < N; ++m) {
            ci.p[m] = cidx[m];
            rr.r[m] = rows[m];
        }
        ALTO_TRY(bad.alloc(sizeof(int)));
        ALTO_HIP_TRY(hipMemsetAsync(bad.p, 0, sizeof(int), s));
        if (nnz) k_check_coords<<<cdiv(nnz, 256), 256, 0, s>>>(ci, N, nnz, rr, bad.as<int>());
        ALTO_LAUNCH_CHECK();
        int hb = 0;
        ALTO_HIP_TRY(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
        ALTO_HIP_TRY(hipStreamSynchronize(s));
        if (hb) {
            set_error("a coordinate is not below its mode's factor rows");
            return ALTO_ERR_ARG;
        }
    }
    ALTO_TRY(mt.alloc(sizeof(double) * (I ? I : 1) * R));
    ALTO_TRY(dpart.alloc(sizeof(double) * kDotBlocks));
    ALTO_TRY(sq.alloc(sizeof(double) * kSumsqBlocks));
    ALTO_TRY(xx.alloc(sizeof(double)));
    ALTO_TRY(G.alloc(sizeof(double) * N * R * R));
    ALTO_TRY(gpart.alloc(sizeof(double) * kMaxBlocks * R * R));
    ALTO_TRY(res.alloc(sizeof(double) * 4));
    ALTO_TRY(launch_fill_zero(mt.as<double>(), I * R, s));
    {
        // the MTTKRP through an ALTO tensor and the mode's output-ordered
        // schedule: the few rows of the shortest mode would serialise the
        // reference's per-nonzero row updates
        alto_hip_coo c{N, rows, nnz, cidx, vals, 1};
        TensorImpl t;
        ALTO_TRY(build_alto(&t, &c, 0, 0, false, s));
        const void *fp[kMaxModes];
        for (int m = 0; m < N; ++m) fp[m] = factors[m];
        ALTO_TRY(run_mttkrp(&t, mn, ALTO_PREC_F64, R, fp, mt.p, 1, 2, 0, -1, s, false, true));
        ALTO_HIP_TRY(hipStreamSynchronize(s));
    }
    k_dot_part<<<kDotBlocks, 256, 0, s>>>(factors[mn], mt.as<double>(), I * R, dpart.as<double>());
    ALTO_LAUNCH_CHECK();
    ALTO_TRY(launch_sumsq(vals, nnz, xx.as<double>(), sq.as<double>(), s));
    for (int m = 0; m < N; ++m) {
        const int nb = row_blocks(rows[m]);
        ALTO_TRY(launch_gram(factors[m], ALTO_PREC_F64, rows[m], R, gpart.as<double>(), nb, s));
        ALTO_TRY(launch_gram_reduce(gpart.as<double>(), nb, R, G.as<double>() + (size_t)m * R * R, s));
    }
    k_error_final<<<1, 64, 0, s>>>(dpart.as<double>(), xx.as<double>(), G.as<double>(), N, R,
                                   res.as<double>());
    ALTO_LAUNCH_CHECK();
    ALTO_HIP_TRY(hipMemcpyAsync(out, res.p, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
    ALTO_HIP_TRY(hipStreamSynchronize(s));
    return ALTO_OK;
}

}  // namespace alto

using namespace alto;

extern "C" int alto_hip_cpd_error(const alto_hip_coo *coo, int rank, const double *const *factors,
                                  double *err, double *parts, void *stream) {
    if (!coo || !coo->dims || !coo->cidx || !coo->vals || !factors || !err) return ALTO_ERR_ARG;
    const int N = coo->nmode;
    if (N < 1 || N > kMaxModes || rank < 1 || rank > kMaxRank) return ALTO_ERR_ARG;
    for (int m = 0; m < N; ++m)
        if (!factors[m] || coo->dims[m] == 0) return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    hipStream_t s = to_stream(stream);
    const uint64_t nnz = coo->nnz;
    DevBuf dc[kMaxModes], dv, df[kMaxModes];
    const uint64_t *cp[kMaxModes];
    const double *fp[kMaxModes];
    const double *vp = coo->vals;
    if (coo->on_device) {
        for (int m = 0; m < N; ++m) {
            ALTO_TRY(check_device_access(factors[m], "cpd_error factor"));
            if (nnz) ALTO_TRY(check_device_access(coo->cidx[m], "cpd_error coordinates"));
        }
        if (nnz) ALTO_TRY(check_device_access(coo->vals, "cpd_error values"));
    }
    for (int m = 0; m < N; ++m) {
        cp[m] = coo->cidx[m];
        fp[m] = factors[m];
        if (!coo->on_device) {
            ALTO_TRY(dc[m].alloc(sizeof(uint64_t) * (nnz ? nnz : 1)));
            if (nnz)
                ALTO_HIP_TRY(hipMemcpyAsync(dc[m].p, coo->cidx[m], sizeof(uint64_t) * nnz,
                                            hipMemcpyHostToDevice, s));
            cp[m] = dc[m].as<uint64_t>();
            const uint64_t n = coo->dims[m] * (uint64_t)rank;
            ALTO_TRY(df[m].alloc(sizeof(double) * n));
            ALTO_HIP_TRY(hipMemcpyAsync(df[m].p, factors[m], sizeof(double) * n, hipMemcpyHostToDevice, s));
            fp[m] = df[m].as<double>();
        }
    }
    if (!coo->on_device) {
        ALTO_TRY(dv.alloc(sizeof(double) * (nnz ? nnz : 1)));
        if (nnz) ALTO_HIP_TRY(hipMemcpyAsync(dv.p, coo->vals, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
        vp = dv.as<double>();
    }
    double out[4];
    ALTO_TRY(cpd_error_dev(N, coo->dims, nnz, cp, vp, fp, rank, s, out));
    *err = out[0];
    if (parts) {
        parts[0] = out[1];
        parts[1] = out[2];
        parts[2] = out[3];
    }
    return ALTO_OK;
}

// ---------------------------------------------------------------------------
// alto_hip_cpd_error_noalloc: the same quantities with no device allocation,
// no host synchronisation and no host read, for callers that cannot allocate
// (graph capture, tight serving loops).  <X, M> is summed per nonzero,
// val * sum_r prod_m U_m[i_m, r] (the same sum as the MTTKRP dot product,
// in another order), so no MTTKRP output is needed; the scratch holds the
// reduction partials, the Grams and a coordinate-check flag.

namespace alto {

constexpr int kNaMaxRank = 64;
constexpr int kNaGramBlocks = 256;
constexpr size_t kNaFlag = 0;                                 // int flag in the first slot
constexpr size_t kNaDot = 32;                                 // kDotBlocks partials
constexpr size_t kNaSq = kNaDot + kDotBlocks;                 // kSumsqBlocks partials
constexpr size_t kNaXX = kNaSq + kSumsqBlocks;                // ||X||^2
constexpr size_t kNaG = kNaXX + 32;                           // kMaxModes R x R Grams
constexpr size_t kNaGpart = kNaG + (size_t)kMaxModes * kNaMaxRank * kNaMaxRank;
constexpr size_t kNaDoubles = kNaGpart + (size_t)kNaGramBlocks * kNaMaxRank * kNaMaxRank;

// The scratch array (about 8.4 MB per GPU, held until the process ends).  In
// host code its name is only a handle: kernels get the device address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[kNaDoubles];

constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mu;
static double *g_na_ptr[kNaMaxDevices] = {};

// Device address of `scratch` on the current device, resolved once per device.
static int na_scratch(double **out) {
    int dev = 0;
    ALTO_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDevices) return ALTO_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> g(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        ALTO_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_na_ptr[dev] = (double *)p;
    }
    *out = g_na_ptr[dev];
    return ALTO_OK;
}

__global__ void k_na_clear_flag(int *flag) {
    if (threadIdx.x == 0) *flag = 0;
}

struct FacArgs {
    const double *p[kMaxModes];
};

// part[b] = sum over block b's nonzeros of val * sum_r prod_m U_m[i_m, r];
// a coordinate not below its mode's rows sets the flag and is skipped.
__global__ __launch_bounds__(256) void k_na_inner(IdxArgs ci, const double *__restrict__ vals, uint64_t nnz,
                                                  FacArgs F, Rows rows, int N, int R, double *__restrict__ part,
                                                  int *__restrict__ flag) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nnz; i += (uint64_t)gridDim.x * 256) {
        bool ok = true;
        for (int m = 0; m < N; ++m) ok = ok && ci.p[m][i] < rows.r[m];
        if (!ok) {
            *flag = 1;
            continue;
        }
        double srow = 0.0;
        for (int r = 0; r < R; ++r) {
            double v = 1.0;
            for (int m = 0; m < N; ++m) v *= F.p[m][ci.p[m][i] * R + r];
            srow += v;
        }
        acc += vals[i] * srow;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// As k_error_final, into the caller's device output; NaN when the flag is set.
__global__ void k_na_final(const double *__restrict__ dot_part, const double *__restrict__ xx,
                           const double *__restrict__ grams, int nmode, int R, const int *__restrict__ flag,
                           double *__restrict__ out) {
    const int lane = threadIdx.x;
    double inner = wave_reduce_entry(dot_part, kDotBlocks, 1, 0, 0);
    double z = 0.0;
    const int RR = R * R;
    for (int e = lane; e < RR; e += 64) {
        const int i = e / R, j = e % R;
        if (j < i) continue;
        double v = 1.0;
        for (int m = 0; m < nmode; ++m) v *= grams[(size_t)m * RR + e];
        z += (i == j) ? v : 2.0 * v;
    }
    for (int o = 32; o > 0; o >>= 1) z += __shfl_down(z, o, 64);
    if (lane == 0) {
        if (*flag) {
            // NaN made at run time (0 / 0): as a constant, this compiler
            // materialises it with a 64-bit scalar move whose encoding keeps
            // only the low 32 bits of the literal, which stores 0
            const double q = (double)*flag - 1.0;
            const double nan_v = q / q;
            for (int k = 0; k < 4; ++k) out[k] = nan_v;
            return;
        }
        const double X = *xx, Z = fabs(z);
        out[0] = sqrt(X + Z - 2.0 * inner) / sqrt(X);
        out[1] = X;
        out[2] = Z;
        out[3] = inner;
    }
}

}  // namespace alto

extern "C" int alto_hip_cpd_error_noalloc(const alto_hip_coo *coo, int rank, const double *const *factors,
                                          double *out, void *stream) {
    if (!coo || !coo->dims || !coo->cidx || !coo->vals || !factors || !out) return ALTO_ERR_ARG;
    if (!coo->on_device) {
        set_error("cpd_error_noalloc takes device arrays (coo->on_device = 1)");
        return ALTO_ERR_ARG;
    }
    const int N = coo->nmode;
    if (N < 1 || N > kMaxModes || rank < 1) return ALTO_ERR_ARG;
    if (rank > kNaMaxRank) {
        set_error("cpd_error_noalloc: rank %d exceeds its scratch limit of %d", rank, kNaMaxRank);
        return ALTO_ERR_UNSUPPORTED;
    }
    for (int m = 0; m < N; ++m)
        if (!factors[m] || coo->dims[m] == 0 || (coo->nnz && !coo->cidx[m])) return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    hipStream_t s = to_stream(stream);
    double *sc = nullptr;
    ALTO_TRY(na_scratch(&sc));
    int *flag = (int *)(sc + kNaFlag);
    IdxArgs ci{};
    Rows rr{};
    FacArgs F{};
    for (int m = 0; m < N; ++m) {
        ci.p[m] = coo->cidx[m];
        rr.r[m] = coo->dims[m];
        F.p[m] = factors[m];
    }
    k_na_clear_flag<<<1, 64, 0, s>>>(flag);
    k_na_inner<<<kDotBlocks, 256, 0, s>>>(ci, coo->vals, coo->nnz, F, rr, N, rank, sc + kNaDot, flag);
    ALTO_LAUNCH_CHECK();
    ALTO_TRY(launch_sumsq(coo->vals, coo->nnz, sc + kNaXX, sc + kNaSq, s));
    const size_t RR = (size_t)rank * rank;
    for (int m = 0; m < N; ++m) {
        const int nb = std::min(row_blocks(coo->dims[m]), kNaGramBlocks);
        ALTO_TRY(launch_gram(factors[m], ALTO_PREC_F64, coo->dims[m], rank, sc + kNaGpart, nb, s));
        ALTO_TRY(launch_gram_reduce(sc + kNaGpart, nb, rank, sc + kNaG + (size_t)m * RR, s));
    }
    k_na_final<<<1, 64, 0, s>>>(sc + kNaDot, sc + kNaXX, sc + kNaG, N, rank, flag, out);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}
