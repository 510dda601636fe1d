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
