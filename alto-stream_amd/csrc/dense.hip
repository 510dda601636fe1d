// Dense FP64 kernels of the ALS / CP-stream step (gfx950).
//
// Reference operations: hadamard_product (src/matrix.cpp:269-294), the Gram
// Hadamard in pseudo_inverse (include/cpd.hpp:383-411), dpotrf/dpotrs
// (cpd.hpp:423-433, matrix.cpp:333-372), the dgelsy fallback (cpd.hpp:438-476),
// KruskalModelNorm (src/kruskal_model.cpp:227-353), update_gram
// (src/gram.cpp:29-42).
//
// Every reduction is two-level with a fixed order (per-block partials, then
// one wave per entry summing the partials in a fixed lane/shuffle order), so
// results do not depend on scheduling.
#include <float.h>
#include <math.h>

#include <mutex>

#include "dense.hpp"

namespace alto {

// 4 doubles: the accumulator of v_mfma_f64_16x16x4f64
typedef double v4d __attribute__((ext_vector_type(4)));


__global__ void k_fill_zero(double *p, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

int launch_fill_zero(double *p, uint64_t n, hipStream_t s) {
    if (!n) return ALTO_OK;
    k_fill_zero<<<cdiv(n, 256), 256, 0, s>>>(p, n);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

int row_blocks(uint64_t rows) {
    uint64_t b = (rows + 63) / 64;
    if (b < 1) b = 1;
    if (b > (uint64_t)kMaxBlocks) b = kMaxBlocks;
    return (int)b;
}

int gram_blocks(uint64_t rows, int R) {
    const int b = row_blocks(rows);
    const int cap = R > 32 ? 512 : R > 16 ? 1280 : kMaxBlocks;
    return b < cap ? b : cap;
}

// One workgroup.  Phi = Hadamard of G_m (m != mode, in mode order); A = Phi +
// diag_add I; Cholesky A = L L^T (lower, right-looking); L^{-1} in place
// (column sweep from the right, as LAPACK dtrti2); vinv = L^{-T} L^{-1}.
__global__ __launch_bounds__(256) void k_chol_inv(const double *__restrict__ grams, int nmode,
                                                  int mode, int R, double diag_add,
                                                  double *__restrict__ vinv,
                                                  int *__restrict__ flag,
                                                  double *__restrict__ phi_out, int fail_mode,
                                                  double *__restrict__ rho_out) {
    extern __shared__ double lds[];
    double *A = lds;              // R*R
    double *tmp = lds + R * R;    // R
    __shared__ int fail;
    __shared__ double dmax;
    const int t = threadIdx.x;
    const int RR = R * R;
    for (int e = t; e < RR; e += blockDim.x) {
        double v = 1.0;
        for (int m = 0; m < nmode; ++m)
            if (m != mode) v *= grams[(size_t)m * RR + e];
        if (phi_out) phi_out[e] = v;
        A[e] = v;
    }
    __syncthreads();
    if (t == 0) {
        fail = 0;
        if (rho_out) {  // ADMM penalty: rho = tr(Phi) / R (admm.cpp:254)
            double tr = 0.0;
            for (int i = 0; i < R; ++i) tr += A[i * R + i];
            dmax = tr / (double)R;
            *rho_out = dmax;
        } else {
            dmax = diag_add;
        }
    }
    __syncthreads();
    for (int i = t; i < R; i += blockDim.x) A[i * R + i] += dmax;
    __syncthreads();
    for (int j = 0; j < R; ++j) {
        double d = A[j * R + j];
        if (!(d > 0.0)) {
            if (t == 0) fail = 1;
            break;
        }
        double l = sqrt(d);
        __syncthreads();
        if (t == 0) A[j * R + j] = l;
        for (int i = j + 1 + t; i < R; i += blockDim.x) A[i * R + j] /= l;
        __syncthreads();
        int n = R - j - 1;
        for (int e = t; e < n * n; e += blockDim.x) {
            int a = j + 1 + e / n, b = j + 1 + e % n;
            if (a >= b) A[a * R + b] -= A[a * R + j] * A[b * R + j];
        }
        __syncthreads();
    }
    __syncthreads();
    if (fail && fail_mode == 1) {
        // pseudo_inverse(A, B) (matrix.cpp:226-250) leaves B unchanged on a
        // failed factorisation: identity.
        for (int e = t; e < RR; e += blockDim.x) vinv[e] = (e / R == e % R) ? 1.0 : 0.0;
        if (t == 0 && flag) *flag = 1;
        return;
    }
    if (fail) {
        // Reference fallback (cpd.hpp:438-476): the unregularised matrix,
        // symmetrised from its lower triangle, which the reference never
        // writes off the diagonal; minimum-norm solve of that diagonal.
        if (t == 0) {
            double mx = 0.0;
            for (int r = 0; r < R; ++r) {
                double v = 1.0;
                for (int m = 0; m < nmode; ++m)
                    if (m != mode) v *= grams[(size_t)m * RR + r * R + r];
                mx = fabs(v) > mx ? fabs(v) : mx;
            }
            dmax = mx;
        }
        __syncthreads();
        double tol = dmax * R * DBL_EPSILON;
        for (int e = t; e < RR; e += blockDim.x) {
            int i = e / R, j = e % R;
            double v = 0.0;
            if (i == j) {
                double d = 1.0;
                for (int m = 0; m < nmode; ++m)
                    if (m != mode) d *= grams[(size_t)m * RR + e];
                v = fabs(d) > tol ? 1.0 / d : 0.0;
            }
            vinv[e] = v;
        }
        if (t == 0 && flag) *flag = 1;
        return;
    }
    // In-place inverse of the lower factor.
    for (int j = R - 1; j >= 0; --j) {
        if (t == 0) A[j * R + j] = 1.0 / A[j * R + j];
        for (int k = j + 1 + t; k < R; k += blockDim.x) tmp[k] = A[k * R + j];
        __syncthreads();
        double ajj = A[j * R + j];
        for (int i = j + 1 + t; i < R; i += blockDim.x) {
            double s = 0.0;
            for (int k = j + 1; k <= i; ++k) s += A[i * R + k] * tmp[k];
            A[i * R + j] = -ajj * s;
        }
        __syncthreads();
    }
    for (int e = t; e < RR; e += blockDim.x) {
        int a = e / R, b = e % R;
        int k0 = a > b ? a : b;
        double s = 0.0;
        for (int k = k0; k < R; ++k) s += A[k * R + a] * A[k * R + b];
        vinv[e] = s;
    }
    if (t == 0 && flag) *flag = 0;
}

// One-wave variant of k_chol_inv for R <= 64 (lane i owns row i): the same
// Cholesky arithmetic (A[a][b] -= A[a][j] * A[b][j], j ascending), then each
// lane k forms column k of L^{-1} by forward substitution and row k of
// vinv = L^{-T} L^{-1}.  Barriers of a one-wave workgroup are nearly free.
__global__ __launch_bounds__(64) void k_chol_inv_w(const double *__restrict__ grams, int nmode,
                                                   int mode, int R, double diag_add,
                                                   double *__restrict__ vinv,
                                                   int *__restrict__ flag,
                                                   double *__restrict__ phi_out, int fail_mode,
                                                   double *__restrict__ rho_out) {
    extern __shared__ double lds[];
    double *A = lds;           // R*R
    double *X = lds + R * R;   // R*R, L^{-1}
    __shared__ int fail;
    __shared__ double dshift;
    const int t = threadIdx.x;
    const int RR = R * R;
    for (int e = t; e < RR; e += 64) {
        double v = 1.0;
        for (int m = 0; m < nmode; ++m)
            if (m != mode) v *= grams[(size_t)m * RR + e];
        if (phi_out) phi_out[e] = v;
        A[e] = v;
    }
    __syncthreads();
    if (t == 0) {
        fail = 0;
        if (rho_out) {  // ADMM penalty: rho = tr(Phi) / R (admm.cpp:254)
            double tr = 0.0;
            for (int i = 0; i < R; ++i) tr += A[i * R + i];
            dshift = tr / (double)R;
            *rho_out = dshift;
        } else {
            dshift = diag_add;
        }
    }
    __syncthreads();
    if (t < R) A[t * R + t] += dshift;
    __syncthreads();
    for (int j = 0; j < R; ++j) {
        const double d = A[j * R + j];
        if (!(d > 0.0)) {
            if (t == 0) fail = 1;
            break;
        }
        const double l = sqrt(d);
        __syncthreads();
        if (t == j) A[j * R + j] = l;
        if (t > j && t < R) A[t * R + j] /= l;
        __syncthreads();
        if (t > j && t < R) {
            const double aij = A[t * R + j];
            for (int b = j + 1; b <= t; ++b) A[t * R + b] -= aij * A[b * R + j];
        }
        __syncthreads();
    }
    __syncthreads();
    if (fail && fail_mode == 1) {
        for (int e = t; e < RR; e += 64) vinv[e] = (e / R == e % R) ? 1.0 : 0.0;
        if (t == 0 && flag) *flag = 1;
        return;
    }
    if (fail) {
        if (t == 0) {
            double mx = 0.0;
            for (int r = 0; r < R; ++r) {
                double v = 1.0;
                for (int m = 0; m < nmode; ++m)
                    if (m != mode) v *= grams[(size_t)m * RR + r * R + r];
                mx = fabs(v) > mx ? fabs(v) : mx;
            }
            dshift = mx;
        }
        __syncthreads();
        const double tol = dshift * R * DBL_EPSILON;
        for (int e = t; e < RR; e += 64) {
            int i = e / R, jj = e % R;
            double v = 0.0;
            if (i == jj) {
                double dd = 1.0;
                for (int m = 0; m < nmode; ++m)
                    if (m != mode) dd *= grams[(size_t)m * RR + e];
                v = fabs(dd) > tol ? 1.0 / dd : 0.0;
            }
            vinv[e] = v;
        }
        if (t == 0 && flag) *flag = 1;
        return;
    }
    if (t < R) {  // column t of L^{-1}
        const int k = t;
        for (int i = 0; i < k; ++i) X[i * R + k] = 0.0;
        for (int i = k; i < R; ++i) {
            double sacc = i == k ? 1.0 : 0.0;
            for (int p = k; p < i; ++p) sacc -= A[i * R + p] * X[p * R + k];
            X[i * R + k] = sacc / A[i * R + i];
        }
    }
    __syncthreads();
    if (t < R) {  // row t of L^{-T} L^{-1}
        const int a = t;
        for (int b = 0; b < R; ++b) {
            const int k0 = a > b ? a : b;
            double sacc = 0.0;
            for (int k = k0; k < R; ++k) sacc += X[k * R + a] * X[k * R + b];
            vinv[a * R + b] = sacc;
        }
    }
    if (t == 0 && flag) *flag = 0;
}

// v_readlane of a double: lane `l` (wave-uniform) of x.
__device__ __forceinline__ double readlane_f64(double x, int l) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Inverse of Phi + shift I (Phi = Hadamard of the Grams of the modes other
// than `mode`) for R in {8, 16, 32, 64}: Gauss-Jordan elimination without
// pivoting on the symmetric matrix built from Phi's row-major upper triangle
// (the triangle the reference's Cholesky reads).  For a symmetric positive
// definite matrix the pivots are the squares of the Cholesky diagonal, so the
// elimination fails (pivot not > 0) exactly where the Cholesky would; every
// thread reads the same pivot, so the decision is uniform.  The matrix stays
// in registers: thread t < min(R^2, 256) owns column t % R for the rows
// i = t / R (mod C), C = min(R^2, 256) / R.  Step k needs row k (for the
// thread's column) and column k (for the thread's rows).  Row k's owners
// publish it to LDS, double-buffered by the parity of k, so each step has one
// barrier; column k comes from the lane of the same wave that owns column k
// for the same rows (v_readlane, no LDS).  Fallbacks as k_chol_inv_w; the
// result is symmetrised.
template <int R>
__device__ __forceinline__ void k_inv_gj_body(const double *__restrict__ grams, int nmode, int mode,
                                                double diag_add, double *__restrict__ vinv,
                                                int *__restrict__ flag, double *__restrict__ phi_out,
                                                int fail_mode, double *__restrict__ rho_out) {
    constexpr int RR = R * R, LD = R + 1;
    constexpr int NT = RR < 256 ? RR : 256, C = NT / R, NR = R / C, G = 64 / R;
    __shared__ double full[R * LD];
    __shared__ double rowb[2][R];
    __shared__ double shift_s;
    const int t = threadIdx.x;
    const bool own = t < NT;
    const int j = t % R, cls = t / R;  // rows cls, cls + C, ...
    const int grp = (t & 63) / R;
    // Phi from the upper triangles, mode by mode; each thread's entries load
    // together (one round trip per mode)
    constexpr int PE = (RR + 255) / 256;
    double v[PE], w[PE];
#pragma unroll
    for (int q = 0; q < PE; ++q) v[q] = w[q] = 1.0;
    for (int m = 0; m < nmode; ++m) {
        if (m == mode) continue;
        const double *g = grams + (size_t)m * RR;
#pragma unroll
        for (int q = 0; q < PE; ++q) {
            const int e = t + 256 * q;
            if (e < RR) {
                const int i = e / R, jj = e % R;
                const int lo = i < jj ? i : jj, hi = i < jj ? jj : i;
                v[q] *= g[lo * R + hi];
            }
        }
    }
    if (phi_out) {
        for (int m = 0; m < nmode; ++m) {
            if (m == mode) continue;
            const double *g = grams + (size_t)m * RR;
#pragma unroll
            for (int q = 0; q < PE; ++q)
                if (t + 256 * q < RR) w[q] *= g[t + 256 * q];
        }
    }
#pragma unroll
    for (int q = 0; q < PE; ++q) {
        const int e = t + 256 * q;
        if (e < RR) {
            full[(e / R) * LD + e % R] = v[q];
            if (phi_out) phi_out[e] = w[q];
        }
    }
    __syncthreads();
    if (t == 0) {
        double shift = diag_add;
        if (rho_out) {  // ADMM penalty: rho = tr(Phi) / R (admm.cpp:254), summed in row order
            double tr = 0.0;
            for (int k = 0; k < R; ++k) tr += full[k * LD + k];
            shift = tr / (double)R;
            *rho_out = shift;
        }
        shift_s = shift;
    }
    __syncthreads();
    double a[NR];
#pragma unroll
    for (int n = 0; n < NR; ++n) {
        const int i = cls + C * n;
        a[n] = own ? full[i * LD + j] + (i == j ? shift_s : 0.0) : 0.0;
    }
    bool bad = false;
    for (int k = 0; k < R; ++k) {
        const int b = k & 1;
        if (own) {
            if (k % C == cls) {
                double v = 0.0;
#pragma unroll
                for (int n = 0; n < NR; ++n) v = n == k / C ? a[n] : v;
                rowb[b][j] = v;
            }
        }
        __syncthreads();
        const double p = rowb[b][k];
        if (!(p > 0.0)) {
            bad = true;
            break;
        }
        if (own) {
            const double ip = 1.0 / p;
            const double rk = (j == k ? 1.0 : rowb[b][j]) * ip;  // new row k
#pragma unroll
            for (int n = 0; n < NR; ++n) {
                // column k of these rows: lane g R + k of this wave, g = this
                // lane's group of R lanes (R < 64: 64 / R groups per wave)
                double akn = 0.0;
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const double r = readlane_f64(a[n], g * R + k);
                    akn = grp == g ? r : akn;
                }
                const int i = cls + C * n;
                a[n] = i == k ? rk : (j == k ? 0.0 : a[n]) - akn * rk;
            }
        }
    }
    if (bad) {
        if (fail_mode == 1) {
            for (int e = t; e < RR; e += 256) vinv[e] = (e / R == e % R) ? 1.0 : 0.0;
        } else {
            // reference fallback (cpd.hpp:438-476): minimum-norm solve of the
            // unregularised diagonal
            double mx = 0.0;
            for (int r = 0; r < R; ++r) {
                double v = 1.0;
                for (int m = 0; m < nmode; ++m)
                    if (m != mode) v *= grams[(size_t)m * RR + r * R + r];
                mx = fabs(v) > mx ? fabs(v) : mx;
            }
            const double tol = mx * R * DBL_EPSILON;
            for (int e = t; e < RR; e += 256) {
                const int i = e / R, jj = e % R;
                double v = 0.0;
                if (i == jj) {
                    double dd = 1.0;
                    for (int m = 0; m < nmode; ++m)
                        if (m != mode) dd *= grams[(size_t)m * RR + e];
                    v = fabs(dd) > tol ? 1.0 / dd : 0.0;
                }
                vinv[e] = v;
            }
        }
        if (t == 0 && flag) *flag = 1;
        return;
    }
    if (own) {
#pragma unroll
        for (int n = 0; n < NR; ++n) full[(cls + C * n) * LD + j] = a[n];
    }
    __syncthreads();
    for (int e = t; e < RR; e += 256) {
        const int i = e / R, jj = e % R;
        vinv[e] = 0.5 * (full[i * LD + jj] + full[jj * LD + i]);
    }
    if (t == 0 && flag) *flag = 0;
}

template <int R>
__global__ __launch_bounds__(256) void k_inv_gj(const double *__restrict__ grams, int nmode, int mode,
                                               double diag_add, double *__restrict__ vinv,
                                               int *__restrict__ flag, double *__restrict__ phi_out,
                                               int fail_mode, double *__restrict__ rho_out) {
    k_inv_gj_body<R>(grams, nmode, mode, diag_add, vinv, flag, phi_out, fail_mode, rho_out);
}

// The same for R in {8, 16, 32} in one wave, with no barrier: lane l owns
// column l % R for the rows i = l / R (mod C), C = 64 / R.  Step k takes row
// k (for the lane's column) and column k (for its rows) from the lanes that
// own them (ds_bpermute), the pivot by v_readlane; the loop is unrolled, so
// every register index is static.  Same arithmetic, fallbacks and output as
// k_inv_gj.
template <int R>
__device__ __forceinline__ void k_inv_gj1_body(const double *__restrict__ grams, int nmode, int mode,
                                               double diag_add, double *__restrict__ vinv,
                                               int *__restrict__ flag, double *__restrict__ phi_out,
                                               int fail_mode, double *__restrict__ rho_out) {
    static_assert(R == 8 || R == 16 || R == 32, "one wave");
    constexpr int RR = R * R, C = 64 / R, NR = R / C, LD = R + 1;
    __shared__ double full[R * LD];
    const int t = threadIdx.x, j = t % R, cls = t / R;
    // Phi from the upper triangles, mode by mode, straight into registers
    double a[NR];
#pragma unroll
    for (int n = 0; n < NR; ++n) a[n] = 1.0;
    for (int m = 0; m < nmode; ++m) {
        if (m == mode) continue;
        const double *g = grams + (size_t)m * RR;
#pragma unroll
        for (int n = 0; n < NR; ++n) {
            const int i = cls + C * n;
            const int lo = i < j ? i : j, hi = i < j ? j : i;
            a[n] *= g[lo * R + hi];
        }
    }
    if (phi_out) {
        for (int e = t; e < RR; e += 64) {
            double w = 1.0;
            for (int m = 0; m < nmode; ++m)
                if (m != mode) w *= grams[(size_t)m * RR + e];
            phi_out[e] = w;
        }
    }
    double shift = diag_add;
    if (rho_out) {  // ADMM penalty: rho = tr(Phi) / R (admm.cpp:254), summed in row order
        double tr = 0.0;
        for (int k = 0; k < R; ++k) {
            double v = 1.0;
            for (int m = 0; m < nmode; ++m)
                if (m != mode) v *= grams[(size_t)m * RR + k * R + k];
            tr += v;
        }
        shift = tr / (double)R;
        if (t == 0) *rho_out = shift;
    }
#pragma unroll
    for (int n = 0; n < NR; ++n)
        if (cls + C * n == j) a[n] += shift;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int kn = k / C, kc = k % C;
        const double p = readlane_f64(a[kn], kc * R + k);
        if (!(p > 0.0)) {
            bad = true;
            break;
        }
        const double ip = 1.0 / p;
        const double rkj = __shfl(a[kn], kc * R + j, 64);
        const double rk = (j == k ? 1.0 : rkj) * ip;  // new row k
        double akn[NR];
#pragma unroll
        for (int n = 0; n < NR; ++n) akn[n] = __shfl(a[n], cls * R + k, 64);
#pragma unroll
        for (int n = 0; n < NR; ++n) {
            const int i = cls + C * n;
            a[n] = i == k ? rk : (j == k ? 0.0 : a[n]) - akn[n] * rk;
        }
    }
    if (bad) {
        if (fail_mode == 1) {
            for (int e = t; e < RR; e += 64) vinv[e] = (e / R == e % R) ? 1.0 : 0.0;
        } else {
            // reference fallback (cpd.hpp:438-476): minimum-norm solve of the
            // unregularised diagonal
            double mx = 0.0;
            for (int r = 0; r < R; ++r) {
                double v = 1.0;
                for (int m = 0; m < nmode; ++m)
                    if (m != mode) v *= grams[(size_t)m * RR + r * R + r];
                mx = fabs(v) > mx ? fabs(v) : mx;
            }
            const double tol = mx * R * DBL_EPSILON;
            for (int e = t; e < RR; e += 64) {
                const int i = e / R, jj = e % R;
                double v = 0.0;
                if (i == jj) {
                    double dd = 1.0;
                    for (int m = 0; m < nmode; ++m)
                        if (m != mode) dd *= grams[(size_t)m * RR + e];
                    v = fabs(dd) > tol ? 1.0 / dd : 0.0;
                }
                vinv[e] = v;
            }
        }
        if (t == 0 && flag) *flag = 1;
        return;
    }
#pragma unroll
    for (int n = 0; n < NR; ++n) full[(cls + C * n) * LD + j] = a[n];
    __syncthreads();
    for (int e = t; e < RR; e += 64) {
        const int i = e / R, jj = e % R;
        vinv[e] = 0.5 * (full[i * LD + jj] + full[jj * LD + i]);
    }
    if (t == 0 && flag) *flag = 0;
}

template <int R>
__global__ __launch_bounds__(64) void k_inv_gj1(const double *__restrict__ grams, int nmode, int mode,
                                               double diag_add, double *__restrict__ vinv,
                                               int *__restrict__ flag, double *__restrict__ phi_out,
                                               int fail_mode, double *__restrict__ rho_out) {
    k_inv_gj1_body<R>(grams, nmode, mode, diag_add, vinv, flag, phi_out, fail_mode, rho_out);
}

int launch_chol_inv(const double *grams, int nmode, int mode, int R, double diag_add,
                    bool copy_first, double *vinv, int *flag, double *phi_out, hipStream_t s,
                    int fail_mode, double *rho_out) {
    (void)copy_first;  // 1.0 * G == G exactly: both Hadamard orders agree bitwise
    if (R < 1 || R > kMaxRank) return ALTO_ERR_UNSUPPORTED;
    static std::once_flag attr;
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute((const void *)k_chol_inv,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256);
        (void)hipFuncSetAttribute((const void *)k_chol_inv_w,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256);
    });
    if (R == 8 || R == 16 || R == 32 || R == 64) {
        if (R == 8)
            k_inv_gj1<8><<<1, 64, 0, s>>>(grams, nmode, mode, diag_add, vinv, flag, phi_out, fail_mode, rho_out);
        else if (R == 16)
            k_inv_gj1<16><<<1, 64, 0, s>>>(grams, nmode, mode, diag_add, vinv, flag, phi_out, fail_mode, rho_out);
        else if (R == 32)
            k_inv_gj1<32><<<1, 64, 0, s>>>(grams, nmode, mode, diag_add, vinv, flag, phi_out, fail_mode, rho_out);
        else
            k_inv_gj<64><<<1, 256, 0, s>>>(grams, nmode, mode, diag_add, vinv, flag, phi_out, fail_mode, rho_out);
    } else if (R <= 64) {
        const size_t lds = sizeof(double) * 2 * (size_t)R * R;
        k_chol_inv_w<<<1, 64, lds, s>>>(grams, nmode, mode, R, diag_add, vinv, flag, phi_out,
                                        fail_mode, rho_out);
    } else {
        const size_t lds = sizeof(double) * ((size_t)R * R + R);
        k_chol_inv<<<1, 256, lds, s>>>(grams, nmode, mode, R, diag_add, vinv, flag, phi_out,
                                       fail_mode, rho_out);
    }
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

static inline int pow2_at_least(int r) {
    int p = 1;
    while (p < r) p <<= 1;
    return p;
}

// u = M * vinv (in place), with per-block column partials.
__global__ __launch_bounds__(256) void k_solve(double *__restrict__ M, uint64_t rows, int R,
                                               int RP, const double *__restrict__ vinv,
                                               double *__restrict__ partials) {
    extern __shared__ double lds[];
    double *W = lds;              // R*R
    double *T = lds + R * R;      // tile: TR rows x R
    const int t = threadIdx.x;
    const int lanes = 256 / RP;   // rows in flight per pass
    const int TR = lanes * 4;
    const int c = t % RP, rl = t / RP;
    for (int e = t; e < R * R; e += 256) W[e] = vinv[e];
    double ss = 0.0, mx = -INFINITY, ip = 0.0;
    uint64_t ntiles = (rows + TR - 1) / TR;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint64_t r0 = tile * TR;
        int nr = (int)((rows - r0) < (uint64_t)TR ? (rows - r0) : TR);
        __syncthreads();
        for (int e = t; e < nr * R; e += 256) T[e] = M[r0 * R + e];
        __syncthreads();
        if (c < R) {
            for (int q = 0; q < 4; ++q) {
                int rr = rl + q * lanes;
                if (rr < nr) {
                    const double *x = T + rr * R;
                    double s = 0.0;
                    for (int k = 0; k < R; ++k) s += x[k] * W[k * R + c];
                    M[(r0 + rr) * R + c] = s;
                    ss += s * s;
                    mx = s > mx ? s : mx;
                    ip += s * x[c];
                }
            }
        }
    }
    __syncthreads();
    double *red = T;  // reuse: 3 * 256 doubles
    red[t] = ss;
    red[256 + t] = mx;
    red[512 + t] = ip;
    __syncthreads();
    if (rl == 0 && c < R) {
        for (int l = 1; l < lanes; ++l) {
            ss += red[l * RP + c];
            double m2 = red[256 + l * RP + c];
            mx = m2 > mx ? m2 : mx;
            ip += red[512 + l * RP + c];
        }
        double *p = partials + (size_t)blockIdx.x * 3 * R;
        p[c] = ss;
        p[R + c] = mx;
        p[2 * R + c] = ip;
    }
}

// k_solve for R in {8, 16, 32, 64}: thread (rl, c) keeps column c of vinv in
// registers, so the inner product reads only the (broadcast) tile row from
// LDS.  Same tiles, arithmetic order and partials as k_solve.
template <int R, typename MT>
__device__ __forceinline__ void k_solve_r_body(unsigned bid_, unsigned nblk_, const MT *M, double *U, uint64_t rows,
                                                 const double *__restrict__ vinv,
                                                 double *__restrict__ partials) {
    constexpr int lanes = 256 / R, TR = lanes * 4;
    __shared__ double T[TR * R > 768 ? TR * R : 768];
    const int t = threadIdx.x;
    const int c = t % R, rl = t / R;
    double w[R];
#pragma unroll
    for (int k = 0; k < R; ++k) w[k] = vinv[k * R + c];
    double ss = 0.0, mx = -INFINITY, ip = 0.0;
    const uint64_t ntiles = (rows + TR - 1) / TR;
    // the next tile's 4 elements per thread load while this tile is solved
    static_assert(TR * R == 1024, "four elements per thread and tile");
    double pre[4];
    auto fetch = [&](uint64_t tile) {
        const uint64_t r0 = tile * TR;
        const int n = (int)((rows - r0) < (uint64_t)TR ? (rows - r0) : TR) * R;
#pragma unroll
        for (int q = 0; q < 4; ++q) pre[q] = t + 256 * q < n ? (double)M[r0 * R + t + 256 * q] : 0.0;
    };
    if (bid_ < ntiles) fetch(bid_);
    for (uint64_t tile = bid_; tile < ntiles; tile += nblk_) {
        const uint64_t r0 = tile * TR;
        const int nr = (int)((rows - r0) < (uint64_t)TR ? (rows - r0) : TR);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) T[t + 256 * q] = pre[q];
        __syncthreads();
        if (tile + nblk_ < ntiles) fetch(tile + nblk_);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rr = rl + q * lanes;
            if (rr < nr) {
                const double *x = T + rr * R;
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < R; ++k) s += x[k] * w[k];
                U[(r0 + rr) * R + c] = s;
                ss += s * s;
                mx = s > mx ? s : mx;
                ip += s * x[c];
            }
        }
    }
    __syncthreads();
    double *red = T;  // 3 * 256 doubles
    red[t] = ss;
    red[256 + t] = mx;
    red[512 + t] = ip;
    __syncthreads();
    if (rl == 0) {
        for (int l = 1; l < lanes; ++l) {
            ss += red[l * R + c];
            const double m2 = red[256 + l * R + c];
            mx = m2 > mx ? m2 : mx;
            ip += red[512 + l * R + c];
        }
        double *p = partials + (size_t)bid_ * 3 * R;
        p[c] = ss;
        p[R + c] = mx;
        p[2 * R + c] = ip;
    }
}

template <int R, typename MT>
__global__ __launch_bounds__(256) void k_solve_r(const MT *M, double *U, uint64_t rows,
                                                 const double *__restrict__ vinv,
                                                 double *__restrict__ partials) {
    k_solve_r_body<R, MT>(blockIdx.x, gridDim.x, M, U, rows, vinv, partials);
}

// u = M * vinv in place on the FP64 matrix cores, R in {16, 32, 64}, with
// the per-block column partials of k_solve_r.  A wave takes 16 rows at a
// time: it loads them coalesced (the next 16 while the current ones are
// multiplied), stages them in its own LDS tile, and forms the 16 x R product
// block by block as v_mfma_f64_16x16x4f64(M[i][4s + l/16], vinv[4s + l/16][16J
// + l%16]) over the R/4 k-steps; vinv sits in LDS.  D item q of lane l is row
// l/16 + 4q, column 16J + l%16 of the block.  Dynamic LDS: R^2 + 64 R doubles.
template <int R>
__device__ __forceinline__ void k_solve_mfma_body(unsigned bid_, unsigned nblk_, double *M, uint64_t rows, const double *__restrict__ vinv,
                                                    double *__restrict__ partials) {
    static_assert(R % 16 == 0 && R <= 64, "16-column blocks");
    constexpr int NB = R / 16, KS = R / 4, PL = R / 4;  // column blocks, k-steps, tile doubles per lane
    extern __shared__ double lds[];
    double *V = lds, *T = lds + R * R;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int i16 = lane & 15, k4 = lane >> 4;
    double *Tw = T + wave * 16 * R;
    for (int e = t; e < R * R; e += 256) V[e] = vinv[e];
    __syncthreads();
    double ss[NB], mx[NB], ip[NB];
#pragma unroll
    for (int J = 0; J < NB; ++J) {
        ss[J] = 0.0;
        mx[J] = -INFINITY;
        ip[J] = 0.0;
    }
    const uint64_t ntiles = (rows + 15) / 16, stride = (uint64_t)nblk_ * 4, lim = rows * R;
    double nx[PL];
    auto load = [&](uint64_t tile) {
        const uint64_t base = tile * 16 * R;
#pragma unroll
        for (int p = 0; p < PL; ++p) {
            const uint64_t e = base + lane + 64 * p;
            nx[p] = e < lim ? M[e] : 0.0;
        }
    };
    uint64_t tile = (uint64_t)bid_ * 4 + wave;
    if (tile < ntiles) load(tile);
    for (; tile < ntiles; tile += stride) {
        // the wave's LDS operations run in order: these stores land before
        // the loads below, and the loads before the next tile's stores
#pragma unroll
        for (int p = 0; p < PL; ++p) Tw[lane + 64 * p] = nx[p];
        if (tile + stride < ntiles) load(tile + stride);
        __builtin_amdgcn_wave_barrier();
        v4d acc[NB];
#pragma unroll
        for (int J = 0; J < NB; ++J) acc[J] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const double a = Tw[i16 * R + 4 * ks + k4];
#pragma unroll
            for (int J = 0; J < NB; ++J) {
                const double b = V[(4 * ks + k4) * R + 16 * J + i16];
                acc[J] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[J], 0, 0, 0);
            }
        }
        const uint64_t r0 = tile * 16;
#pragma unroll
        for (int J = 0; J < NB; ++J)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = k4 + 4 * q, c = 16 * J + i16;
                if (r0 + rr < rows) {
                    const double u = acc[J][q];
                    M[(r0 + rr) * R + c] = u;
                    ss[J] += u * u;
                    mx[J] = u > mx[J] ? u : mx[J];
                    ip[J] += u * Tw[rr * R + c];
                }
            }
        __builtin_amdgcn_wave_barrier();
    }
    // column partials: the four row groups of the wave, then the four waves
    // in order (the tiles' LDS is reused)
    __syncthreads();
    double *red = T;  // [wave][3][R]
#pragma unroll
    for (int J = 0; J < NB; ++J) {
        double a = ss[J], m = mx[J], p = ip[J];
        a += __shfl_xor(a, 16, 64);
        a += __shfl_xor(a, 32, 64);
        const double m1 = __shfl_xor(m, 16, 64);
        m = m1 > m ? m1 : m;
        const double m2 = __shfl_xor(m, 32, 64);
        m = m2 > m ? m2 : m;
        p += __shfl_xor(p, 16, 64);
        p += __shfl_xor(p, 32, 64);
        if (k4 == 0) {
            red[(wave * 3 + 0) * R + 16 * J + i16] = a;
            red[(wave * 3 + 1) * R + 16 * J + i16] = m;
            red[(wave * 3 + 2) * R + 16 * J + i16] = p;
        }
    }
    __syncthreads();
    if (t < R) {
        double a = red[t], m = red[R + t], p = red[2 * R + t];
        for (int w = 1; w < 4; ++w) {
            a += red[(w * 3 + 0) * R + t];
            const double mw = red[(w * 3 + 1) * R + t];
            m = mw > m ? mw : m;
            p += red[(w * 3 + 2) * R + t];
        }
        double *out = partials + (size_t)bid_ * 3 * R;
        out[t] = a;
        out[R + t] = m;
        out[2 * R + t] = p;
    }
}

template <int R>
__global__ __launch_bounds__(256) void k_solve_mfma(double *M, uint64_t rows, const double *__restrict__ vinv,
                                                    double *__restrict__ partials) {
    k_solve_mfma_body<R>(blockIdx.x, gridDim.x, M, rows, vinv, partials);
}

int launch_solve(double *M, uint64_t rows, int R, const double *vinv, double *partials, int nblk,
                 hipStream_t s) {
    if (R > kMaxRank) return ALTO_ERR_UNSUPPORTED;
    int RP = pow2_at_least(R);
    int TR = (256 / RP) * 4;
    size_t tile = (size_t)TR * R;
    if (tile < 768) tile = 768;
    size_t lds = sizeof(double) * ((size_t)R * R + tile);
    static std::once_flag attr;
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute((const void *)k_solve, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024 - 256);
        (void)hipFuncSetAttribute((const void *)k_solve_mfma<64>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024 - 256);
    });
    const size_t lds_mfma = sizeof(double) * ((size_t)R * R + 64 * (size_t)R);
    if (R == 8) k_solve_r<8, double><<<nblk, 256, 0, s>>>(M, M, rows, vinv, partials);
    else if (R == 16) k_solve_mfma<16><<<nblk, 256, lds_mfma, s>>>(M, rows, vinv, partials);
    else if (R == 32) k_solve_mfma<32><<<nblk, 256, lds_mfma, s>>>(M, rows, vinv, partials);
    else if (R == 64) k_solve_mfma<64><<<nblk, 256, lds_mfma, s>>>(M, rows, vinv, partials);
    else k_solve<<<nblk, 256, lds, s>>>(M, rows, R, RP, vinv, partials);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

__device__ __forceinline__ void k_lambda_body(unsigned bid_, unsigned nblk_, const double *__restrict__ partials, int nblk,
                                                int R, int which, double *__restrict__ lambda) {
    const int c = bid_ * 4 + (threadIdx.x >> 6);
    if (c >= R) return;
    if (which == 0) {
        double s = wave_reduce_entry(partials, nblk, 3 * (size_t)R, c, 0);
        if ((threadIdx.x & 63) == 0) lambda[c] = sqrt(s > 1e-12 ? s : 1e-12);
    } else {
        double m = wave_reduce_entry(partials, nblk, 3 * (size_t)R, R + c, 1);
        if ((threadIdx.x & 63) == 0) lambda[c] = m > 1.0 ? m : 1.0;
    }
}

__global__ __launch_bounds__(256) void k_lambda(const double *__restrict__ partials, int nblk,
                                                int R, int which, double *__restrict__ lambda) {
    k_lambda_body(blockIdx.x, gridDim.x, partials, nblk, R, which, lambda);
}

int launch_lambda(const double *partials, int nblk, int R, int which, double *lambda,
                  hipStream_t s) {
    k_lambda<<<cdiv(R, 4), 256, 0, s>>>(partials, nblk, R, which, lambda);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

template <int MODE, int NQ>  // MODE 0: scale f64 u by 1/lambda and store; 1: read stored factor
__device__ __forceinline__ void k_scale_gram_body(unsigned bid_, unsigned nblk_, const double *u,
                                                    const double *__restrict__ lambda,
                                                    const void *__restrict__ Uin,
                                                    void *__restrict__ Uout, int prec,
                                                    uint64_t rows, int R,
                                                    double *__restrict__ gpart, double *zero,
                                                    uint64_t zero_n) {
    extern __shared__ double T[];  // TR x R
    constexpr int TR = 32;
    const int t = threadIdx.x;
    const int RR = R * R;
    double g[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) g[q] = 0.0;
    uint64_t ntiles = (rows + TR - 1) / TR;
    for (uint64_t tile = bid_; tile < ntiles; tile += nblk_) {
        uint64_t r0 = tile * TR;
        int nr = (int)((rows - r0) < (uint64_t)TR ? (rows - r0) : TR);
        __syncthreads();
        for (int e = t; e < nr * R; e += 256) {
            uint64_t gi = r0 * R + e;
            double v;
            if (MODE == 0) {
                double x = u[gi];
                if (lambda) x = x / lambda[e % R];
                if (prec == ALTO_PREC_F64) {
                    ((double *)Uout)[gi] = x;
                    v = x;
                } else if (prec == ALTO_PREC_F32) {
                    float f = (float)x;
                    ((float *)Uout)[gi] = f;
                    v = (double)f;
                } else {
                    uint16_t b = f32_to_bf16((float)x);
                    ((uint16_t *)Uout)[gi] = b;
                    v = (double)bf16_to_f32(b);
                }
                if (zero) zero[gi] = 0.0;  // the element was read above
            } else {
                if (prec == ALTO_PREC_F64) v = ((const double *)Uin)[gi];
                else if (prec == ALTO_PREC_F32) v = (double)((const float *)Uin)[gi];
                else v = (double)bf16_to_f32(((const uint16_t *)Uin)[gi]);
            }
            T[e] = v;
        }
        __syncthreads();
        // thread t owns entries t * NQ .. t * NQ + NQ - 1; when they share
        // one row a (R a multiple of NQ), T[r][a] is read once per row
        const int e0 = t * NQ;
        if (R % NQ == 0 && e0 < RR) {
            const int a = e0 / R, b0 = e0 % R;
            for (int r = 0; r < nr; ++r) {
                const double xa = T[r * R + a];
#pragma unroll
                for (int q = 0; q < NQ; ++q) g[q] += xa * T[r * R + b0 + q];
            }
        } else if (R % NQ != 0) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int e = e0 + q;
                if (e < RR) {
                    const int a = e / R, b = e % R;
                    for (int r = 0; r < nr; ++r) g[q] += T[r * R + a] * T[r * R + b];
                }
            }
        }
    }
    if (zero)  // beyond this factor's rows: nobody reads them here
        for (uint64_t e = rows * R + (uint64_t)bid_ * 256 + t; e < zero_n; e += (uint64_t)nblk_ * 256)
            zero[e] = 0.0;
    double *p = gpart + (size_t)bid_ * RR;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = t * NQ + q;
        if (e < RR) p[e] = g[q];
    }
}

template <int MODE, int NQ>
__global__ __launch_bounds__(256) void k_scale_gram(const double *u, const double *__restrict__ lambda,
                                                    const void *__restrict__ Uin, void *__restrict__ Uout,
                                                    int prec, uint64_t rows, int R, double *__restrict__ gpart,
                                                    double *zero, uint64_t zero_n) {
    k_scale_gram_body<MODE, NQ>(blockIdx.x, gridDim.x, u, lambda, Uin, Uout, prec, rows, R, gpart, zero, zero_n);
}

// The same on the FP64 matrix cores, R a multiple of 16.  A wave takes four
// rows at a time: lane l holds U[4c + l / 16][16 I + l % 16] for each
// 16-column block I, which is both the A operand (U^T, 16 x 4) and the B
// operand (U, 4 x 16) of v_mfma_f64_16x16x4f64, so block (I, J) of U^T U
// accumulates as mfma(v[I], v[J]) with no LDS traffic.  Only the blocks
// I <= J are formed; the partial is mirrored from them, so it is exactly
// symmetric.  Rows are dealt out four at a time to the grid's waves; the
// next four rows are loaded while the current ones are multiplied.  The four
// waves' sums are added in LDS in wave order (deterministic).  MODE 0 also
// scales u by 1 / lambda and stores the factor in PT (FP64, FP32 or BF16
// bits), taking the Gram of the stored values; MODE 1 reads U in PT.  u and
// Uout may be the same array (FP64 in place): each element is read and
// written by one lane, the read first.

template <int MODE, int R, typename PT>
__device__ __forceinline__ void k_scale_gram_mfma_body(unsigned bid_, unsigned nblk_, const double *u, const double *__restrict__ lambda,
                                                         const PT *__restrict__ Uin, PT *Uout, uint64_t rows,
                                                         double *__restrict__ gpart, double *zero, uint64_t zero_n) {
    static_assert(R % 16 == 0 && R <= 64, "16-column blocks");
    constexpr int NB = R / 16, NP = NB * (NB + 1) / 2, RR = R * R;
    using RawT = typename std::conditional<MODE == 0, double, PT>::type;
    __shared__ double G[RR];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int kr = lane >> 4, c16 = lane & 15;
    v4d acc[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) acc[p] = v4d{0.0, 0.0, 0.0, 0.0};
    double lam[NB];
#pragma unroll
    for (int I = 0; I < NB; ++I) lam[I] = (MODE == 0 && lambda) ? lambda[16 * I + c16] : 1.0;
    const uint64_t nchunks = (rows + 3) / 4, stride = (uint64_t)nblk_ * 4;
    auto load = [&](uint64_t c, RawT *raw) {
        const uint64_t row = c * 4 + kr;
#pragma unroll
        for (int I = 0; I < NB; ++I) {
            const uint64_t gi = row * R + 16 * I + c16;
            if constexpr (MODE == 0) raw[I] = row < rows ? u[gi] : 0.0;
            else raw[I] = row < rows ? Uin[gi] : (PT)0;
        }
    };
    uint64_t c = (uint64_t)bid_ * 4 + wave;
    RawT nxt[NB];
    if (c < nchunks) load(c, nxt);
    for (; c < nchunks; c += stride) {
        RawT raw[NB];
#pragma unroll
        for (int I = 0; I < NB; ++I) raw[I] = nxt[I];
        if (c + stride < nchunks) load(c + stride, nxt);
        const uint64_t row = c * 4 + kr;
        double v[NB];
#pragma unroll
        for (int I = 0; I < NB; ++I) {
            if constexpr (MODE == 0) {
                double x = raw[I];
                if (lambda) x = x / lam[I];
                const uint64_t gi = row * R + 16 * I + c16;
                if constexpr (sizeof(PT) == 8) {
                    if (row < rows) Uout[gi] = x;
                    v[I] = x;
                } else if constexpr (sizeof(PT) == 4) {
                    const float f = (float)x;
                    if (row < rows) Uout[gi] = f;
                    v[I] = (double)f;
                } else {
                    const uint16_t b = f32_to_bf16((float)x);
                    if (row < rows) Uout[gi] = b;
                    v[I] = (double)bf16_to_f32(b);
                }
                if (row >= rows) v[I] = 0.0;
                else if (zero) zero[gi] = 0.0;  // its value is consumed above
            } else {
                if constexpr (sizeof(PT) == 2) v[I] = (double)bf16_to_f32(raw[I]);
                else v[I] = (double)raw[I];
            }
        }
        int p = 0;
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int J = I; J < NB; ++J, ++p) acc[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[I], v[J], acc[p], 0, 0, 0);
    }
    if (zero)  // beyond this factor's rows: nobody reads them here
        for (uint64_t e = rows * R + (uint64_t)bid_ * 256 + threadIdx.x; e < zero_n;
             e += (uint64_t)nblk_ * 256)
            zero[e] = 0.0;
    // D item q of lane l is element (l / 16 + 4 q, l % 16) of the block
    for (int w = 0; w < 4; ++w) {
        if (wave == w) {
            int p = 0;
#pragma unroll
            for (int I = 0; I < NB; ++I)
#pragma unroll
                for (int J = I; J < NB; ++J, ++p)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int i = kr + 4 * q, j = c16;
                        if (I == J && i > j) continue;
                        const int a = 16 * I + i, b = 16 * J + j;
                        const double val = acc[p][q];
                        if (w == 0) {
                            G[a * R + b] = val;
                            G[b * R + a] = val;
                        } else {
                            G[a * R + b] += val;
                            if (a != b) G[b * R + a] += val;
                        }
                    }
        }
        __syncthreads();
    }
    double *out = gpart + (size_t)bid_ * RR;
    for (int e = threadIdx.x; e < RR; e += 256) out[e] = G[e];
}

template <int MODE, int R, typename PT>
__global__ __launch_bounds__(256) void k_scale_gram_mfma(const double *u, const double *__restrict__ lambda,
                                                         const PT *__restrict__ Uin, PT *Uout, uint64_t rows,
                                                         double *__restrict__ gpart, double *zero, uint64_t zero_n) {
    k_scale_gram_mfma_body<MODE, R, PT>(blockIdx.x, gridDim.x, u, lambda, Uin, Uout, rows, gpart, zero, zero_n);
}

template <int MODE, int R>
static void launch_sg_mfma(const double *u, const double *lambda, const void *Uin, void *Uout, int prec,
                           uint64_t rows, double *gpart, int nblk, hipStream_t s, double *zero, uint64_t zero_n) {
    if (prec == ALTO_PREC_F64)
        k_scale_gram_mfma<MODE, R, double><<<nblk, 256, 0, s>>>(u, lambda, (const double *)Uin, (double *)Uout,
                                                                 rows, gpart, zero, zero_n);
    else if (prec == ALTO_PREC_F32)
        k_scale_gram_mfma<MODE, R, float><<<nblk, 256, 0, s>>>(u, lambda, (const float *)Uin, (float *)Uout,
                                                                rows, gpart, zero, zero_n);
    else
        k_scale_gram_mfma<MODE, R, uint16_t><<<nblk, 256, 0, s>>>(u, lambda, (const uint16_t *)Uin,
                                                                   (uint16_t *)Uout, rows, gpart, zero, zero_n);
}

template <int MODE>
static void launch_sg(const double *u, const double *lambda, const void *Uin, void *Uout, int prec,
                      uint64_t rows, int R, double *gpart, int nblk, hipStream_t s, double *zero = nullptr,
                      uint64_t zero_n = 0) {
    if (R == 16) return launch_sg_mfma<MODE, 16>(u, lambda, Uin, Uout, prec, rows, gpart, nblk, s, zero, zero_n);
    if (R == 32) return launch_sg_mfma<MODE, 32>(u, lambda, Uin, Uout, prec, rows, gpart, nblk, s, zero, zero_n);
    if (R == 64) return launch_sg_mfma<MODE, 64>(u, lambda, Uin, Uout, prec, rows, gpart, nblk, s, zero, zero_n);
    size_t lds = sizeof(double) * 32 * R;
    int nq = (R * R + 255) / 256;
    if (nq <= 1)
        k_scale_gram<MODE, 1><<<nblk, 256, lds, s>>>(u, lambda, Uin, Uout, prec, rows, R, gpart, zero, zero_n);
    else if (nq <= 4)
        k_scale_gram<MODE, 4><<<nblk, 256, lds, s>>>(u, lambda, Uin, Uout, prec, rows, R, gpart, zero, zero_n);
    else if (nq <= 16)
        k_scale_gram<MODE, 16><<<nblk, 256, lds, s>>>(u, lambda, Uin, Uout, prec, rows, R, gpart, zero, zero_n);
    else
        k_scale_gram<MODE, 64><<<nblk, 256, lds, s>>>(u, lambda, Uin, Uout, prec, rows, R, gpart, zero, zero_n);
}

int launch_scale_gram(const double *u, const double *lambda, uint64_t rows, int R, void *U,
                      int prec, double *gpart, int nblk, hipStream_t s, double *zero, uint64_t zero_n) {
    if (R > kMaxRank) return ALTO_ERR_UNSUPPORTED;
    if (zero && (zero == U || zero_n < rows * (uint64_t)R)) return ALTO_ERR_ARG;
    launch_sg<0>(u, lambda, nullptr, U, prec, rows, R, gpart, nblk, s, zero, zero_n);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

int launch_gram(const void *U, int prec, uint64_t rows, int R, double *gpart, int nblk,
                hipStream_t s) {
    if (R > kMaxRank) return ALTO_ERR_UNSUPPORTED;
    launch_sg<1>(nullptr, nullptr, U, nullptr, prec, rows, R, gpart, nblk, s);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

// Per-block partials of P^T U (R x R) over rows (matmul(prev_M->U[m], true,
// M->U[m], false, ...), streaming_cpd.cpp:1387-1388).
template <int NQ>
__global__ __launch_bounds__(256) void k_cross_gram(const double *__restrict__ P,
                                                    const double *__restrict__ U, uint64_t rows,
                                                    int R, double *__restrict__ gpart) {
    extern __shared__ double T[];  // 2 x 32 x R
    double *TP = T, *TU = T + 32 * R;
    const int t = threadIdx.x, RR = R * R;
    double g[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) g[q] = 0.0;
    uint64_t ntiles = (rows + 31) / 32;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint64_t r0 = tile * 32;
        int nr = (int)((rows - r0) < 32 ? (rows - r0) : 32);
        __syncthreads();
        for (int e = t; e < nr * R; e += 256) {
            TP[e] = P[r0 * R + e];
            TU[e] = U[r0 * R + e];
        }
        __syncthreads();
        // thread t owns entries t * NQ .. t * NQ + NQ - 1; when they share
        // one row a (R a multiple of NQ), TP[r][a] is read once per row
        const int e0 = t * NQ;
        if (R % NQ == 0 && e0 < RR) {
            const int a = e0 / R, b0 = e0 % R;
            for (int r = 0; r < nr; ++r) {
                const double pa = TP[r * R + a];
#pragma unroll
                for (int q = 0; q < NQ; ++q) g[q] += pa * TU[r * R + b0 + q];
            }
        } else if (R % NQ != 0) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int e = e0 + q;
                if (e < RR) {
                    const int a = e / R, b = e % R;
                    for (int r = 0; r < nr; ++r) g[q] += TP[r * R + a] * TU[r * R + b];
                }
            }
        }
    }
    double *p = gpart + (size_t)blockIdx.x * RR;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = t * NQ + q;
        if (e < RR) p[e] = g[q];
    }
}

int launch_cross_gram(const double *P, const double *U, uint64_t rows, int R, double *gpart,
                      int nblk, hipStream_t s) {
    if (R > kMaxRank) return ALTO_ERR_UNSUPPORTED;
    size_t lds = sizeof(double) * 64 * R;
    int nq = (R * R + 255) / 256;
    if (nq <= 1)
        k_cross_gram<1><<<nblk, 256, lds, s>>>(P, U, rows, R, gpart);
    else if (nq <= 4)
        k_cross_gram<4><<<nblk, 256, lds, s>>>(P, U, rows, R, gpart);
    else if (nq <= 16)
        k_cross_gram<16><<<nblk, 256, lds, s>>>(P, U, rows, R, gpart);
    else
        k_cross_gram<64><<<nblk, 256, lds, s>>>(P, U, rows, R, gpart);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

// Column sums of squares in the partials layout of k_solve (sum at [0, R)).
__global__ __launch_bounds__(256) void k_colsumsq(const double *__restrict__ U, uint64_t rows,
                                                  int R, int RP, double *__restrict__ partials) {
    __shared__ double red[256];
    const int t = threadIdx.x, c = t % RP, rl = t / RP, lanes = 256 / RP;
    double ss = 0.0;
    if (c < R)
        for (uint64_t r = (uint64_t)blockIdx.x * lanes + rl; r < rows; r += (uint64_t)gridDim.x * lanes) {
            double v = U[r * R + c];
            ss += v * v;
        }
    red[t] = ss;
    __syncthreads();
    if (rl == 0 && c < R) {
        for (int l = 1; l < lanes; ++l) ss += red[l * RP + c];
        partials[(size_t)blockIdx.x * 3 * R + c] = ss;
    }
}

int launch_colsumsq(const double *U, uint64_t rows, int R, double *partials, int nblk,
                    hipStream_t s) {
    int RP = pow2_at_least(R);
    k_colsumsq<<<nblk, 256, 0, s>>>(U, rows, R, RP, partials);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

__device__ __forceinline__ void k_gram_reduce_body(unsigned bid_, unsigned nblk_, const double *__restrict__ gpart, int nblk,
                                                     int RR, double *__restrict__ G) {
    const int e = bid_ * 4 + (threadIdx.x >> 6);
    if (e >= RR) return;
    double s = wave_reduce_entry(gpart, nblk, RR, e, 0);
    if ((threadIdx.x & 63) == 0) G[e] = s;
}

__global__ __launch_bounds__(256) void k_gram_reduce(const double *__restrict__ gpart, int nblk,
                                                     int RR, double *__restrict__ G) {
    k_gram_reduce_body(blockIdx.x, gridDim.x, gpart, nblk, RR, G);
}

int launch_gram_reduce(const double *gpart, int nblk, int R, double *G, hipStream_t s) {
    int RR = R * R;
    k_gram_reduce<<<cdiv(RR, 4), 256, 0, s>>>(gpart, nblk, RR, G);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

__global__ __launch_bounds__(256) void k_sumsq_part(const double *__restrict__ v, uint64_t n,
                                                    double *__restrict__ part) {
    __shared__ double sh[4];
    double s = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        s += v[i] * v[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// One wave: lane j adds part[j], part[j + 64], ... in order, then a fixed
// shuffle tree (the loads of all lanes are in flight together; a single
// thread walking 1024 partials took ~57 us).
__global__ __launch_bounds__(64) void k_sum_final(const double *__restrict__ part, int n, double *__restrict__ out) {
    double s = 0.0;
    for (int b = threadIdx.x; b < n; b += 64) s += part[b];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (threadIdx.x == 0) *out = s;
}

int launch_sumsq(const double *v, uint64_t n, double *out, double *part, hipStream_t s) {
    // about 4096 values per block: a small tensor takes a few blocks
    const int nb = (int)std::min<uint64_t>(kSumsqBlocks, std::max<uint64_t>(1, cdiv(n, 4096)));
    k_sumsq_part<<<nb, 256, 0, s>>>(v, n, part);
    k_sum_final<<<1, 64, 0, s>>>(part, nb, out);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

// ---- grouped launches ----

bool group_rank_supported(int R) { return R == 8 || R == 16 || R == 32 || R == 64; }

template <int R>
__global__ __launch_bounds__(R == 64 ? 256 : 64) void k_inv_group(const BatchJob *__restrict__ jobs, int nmode,
                                                                   int mode, double diag_add) {
    const BatchJob &J = jobs[blockIdx.x];
    if (!*J.active) return;
    if constexpr (R == 64) k_inv_gj_body<64>(J.grams, nmode, mode, diag_add, J.vinv, J.flag, nullptr, 0, nullptr);
    else k_inv_gj1_body<R>(J.grams, nmode, mode, diag_add, J.vinv, J.flag, nullptr, 0, nullptr);
}

int launch_inv_group(const BatchJob *jobs, int njobs, int nmode, int mode, int R, double diag_add,
                     hipStream_t s) {
    if (njobs <= 0) return ALTO_OK;
    switch (R) {
    case 8: k_inv_group<8><<<njobs, 64, 0, s>>>(jobs, nmode, mode, diag_add); break;
    case 16: k_inv_group<16><<<njobs, 64, 0, s>>>(jobs, nmode, mode, diag_add); break;
    case 32: k_inv_group<32><<<njobs, 64, 0, s>>>(jobs, nmode, mode, diag_add); break;
    case 64: k_inv_group<64><<<njobs, 256, 0, s>>>(jobs, nmode, mode, diag_add); break;
    default: return ALTO_ERR_UNSUPPORTED;
    }
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

// launch_solve's kernel choice per rank: the register kernel at rank 8, the
// matrix-core kernel above
template <int R>
__global__ __launch_bounds__(256) void k_solve_group(const BatchJob *__restrict__ jobs,
                                                     const unsigned *__restrict__ first, int njobs) {
    const int j = find_job(first, njobs, blockIdx.x);
    const BatchJob &J = jobs[j];
    if (!*J.active) return;
    const unsigned b = blockIdx.x - first[j], nb = first[j + 1] - first[j];
    if constexpr (R == 8) k_solve_r_body<8, double>(b, nb, J.mt, J.mt, J.rows, J.vinv, J.part);
    else k_solve_mfma_body<R>(b, nb, J.mt, J.rows, J.vinv, J.part);
}

int launch_solve_group(const BatchJob *jobs, const unsigned *first, int njobs, unsigned nblocks, int R,
                       hipStream_t s) {
    if (njobs <= 0 || nblocks == 0) return ALTO_OK;
    static std::once_flag attr;
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute((const void *)k_solve_group<64>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024 - 256);
    });
    const size_t lds_mfma = sizeof(double) * ((size_t)R * R + 64 * (size_t)R);
    switch (R) {
    case 8: k_solve_group<8><<<nblocks, 256, 0, s>>>(jobs, first, njobs); break;
    case 16: k_solve_group<16><<<nblocks, 256, lds_mfma, s>>>(jobs, first, njobs); break;
    case 32: k_solve_group<32><<<nblocks, 256, lds_mfma, s>>>(jobs, first, njobs); break;
    case 64: k_solve_group<64><<<nblocks, 256, lds_mfma, s>>>(jobs, first, njobs); break;
    default: return ALTO_ERR_UNSUPPORTED;
    }
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

__global__ __launch_bounds__(256) void k_lambda_group(const BatchJob *__restrict__ jobs, int per_job, int R,
                                                      int which) {
    const BatchJob &J = jobs[blockIdx.x / per_job];
    if (!*J.active) return;
    k_lambda_body(blockIdx.x % per_job, per_job, J.part, J.nb_solve, R, which, J.lam);
}

int launch_lambda_group(const BatchJob *jobs, int njobs, int R, int which, hipStream_t s) {
    if (njobs <= 0) return ALTO_OK;
    const int per = (int)cdiv(R, 4);
    k_lambda_group<<<njobs * per, 256, 0, s>>>(jobs, per, R, which);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

// launch_sg<0>'s kernel choice per rank: the LDS kernel at rank 8, the
// matrix-core kernel above
template <int R, typename PT>
__global__ __launch_bounds__(256) void k_scale_gram_group(const BatchJob *__restrict__ jobs,
                                                          const unsigned *__restrict__ first, int njobs,
                                                          int prec, int lambda_on) {
    const int j = find_job(first, njobs, blockIdx.x);
    const BatchJob &J = jobs[j];
    if (!*J.active) return;
    const unsigned b = blockIdx.x - first[j], nb = first[j + 1] - first[j];
    const double *lam = lambda_on ? J.lam : nullptr;
    double *zero = lambda_on ? J.zero : nullptr;
    const uint64_t zn = lambda_on ? J.zero_n : 0;
    if constexpr (R == 8)
        k_scale_gram_body<0, 1>(b, nb, J.u, lam, nullptr, J.U, prec, J.rows, 8, J.gpart, zero, zn);
    else
        k_scale_gram_mfma_body<0, R, PT>(b, nb, J.u, lam, nullptr, (PT *)J.U, J.rows, J.gpart, zero, zn);
}

template <int R>
static void launch_sgg(const BatchJob *jobs, const unsigned *first, int njobs, unsigned nblocks, int prec,
                       int lambda_on, hipStream_t s) {
    const size_t lds = R == 8 ? sizeof(double) * 32 * R : 0;
    if (prec == ALTO_PREC_F64)
        k_scale_gram_group<R, double><<<nblocks, 256, lds, s>>>(jobs, first, njobs, prec, lambda_on);
    else if (prec == ALTO_PREC_F32)
        k_scale_gram_group<R, float><<<nblocks, 256, lds, s>>>(jobs, first, njobs, prec, lambda_on);
    else
        k_scale_gram_group<R, uint16_t><<<nblocks, 256, lds, s>>>(jobs, first, njobs, prec, lambda_on);
}

int launch_scale_gram_group(const BatchJob *jobs, const unsigned *first, int njobs, unsigned nblocks, int R,
                            int prec, int lambda_on, hipStream_t s) {
    if (njobs <= 0 || nblocks == 0) return ALTO_OK;
    switch (R) {
    case 8: launch_sgg<8>(jobs, first, njobs, nblocks, prec, lambda_on, s); break;
    case 16: launch_sgg<16>(jobs, first, njobs, nblocks, prec, lambda_on, s); break;
    case 32: launch_sgg<32>(jobs, first, njobs, nblocks, prec, lambda_on, s); break;
    case 64: launch_sgg<64>(jobs, first, njobs, nblocks, prec, lambda_on, s); break;
    default: return ALTO_ERR_UNSUPPORTED;
    }
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

// Gram partials of the stored factors (launch_gram's kernels): u unused
template <int R, typename PT>
__global__ __launch_bounds__(256) void k_gram_group(const BatchJob *__restrict__ jobs,
                                                    const unsigned *__restrict__ first, int njobs, int prec) {
    const int j = find_job(first, njobs, blockIdx.x);
    const BatchJob &J = jobs[j];
    if (!*J.active) return;
    const unsigned b = blockIdx.x - first[j], nb = first[j + 1] - first[j];
    if constexpr (R == 8)
        k_scale_gram_body<1, 1>(b, nb, nullptr, nullptr, J.U, nullptr, prec, J.rows, 8, J.gpart, nullptr, 0);
    else
        k_scale_gram_mfma_body<1, R, PT>(b, nb, nullptr, nullptr, (const PT *)J.U, nullptr, J.rows, J.gpart, nullptr,
                                         0);
}

template <int R>
static void launch_gg(const BatchJob *jobs, const unsigned *first, int njobs, unsigned nblocks, int prec,
                      hipStream_t s) {
    const size_t lds = R == 8 ? sizeof(double) * 32 * R : 0;
    if (prec == ALTO_PREC_F64) k_gram_group<R, double><<<nblocks, 256, lds, s>>>(jobs, first, njobs, prec);
    else if (prec == ALTO_PREC_F32) k_gram_group<R, float><<<nblocks, 256, lds, s>>>(jobs, first, njobs, prec);
    else k_gram_group<R, uint16_t><<<nblocks, 256, lds, s>>>(jobs, first, njobs, prec);
}

int launch_gram_group(const BatchJob *jobs, const unsigned *first, int njobs, unsigned nblocks, int R, int prec,
                      hipStream_t s) {
    if (njobs <= 0 || nblocks == 0) return ALTO_OK;
    switch (R) {
    case 8: launch_gg<8>(jobs, first, njobs, nblocks, prec, s); break;
    case 16: launch_gg<16>(jobs, first, njobs, nblocks, prec, s); break;
    case 32: launch_gg<32>(jobs, first, njobs, nblocks, prec, s); break;
    case 64: launch_gg<64>(jobs, first, njobs, nblocks, prec, s); break;
    default: return ALTO_ERR_UNSUPPORTED;
    }
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

__global__ __launch_bounds__(256) void k_gram_reduce_group(const BatchJob *__restrict__ jobs, int per_job,
                                                           int RR) {
    const BatchJob &J = jobs[blockIdx.x / per_job];
    if (!*J.active) return;
    k_gram_reduce_body(blockIdx.x % per_job, per_job, J.gpart, J.nb_gram, RR, J.G);
}

int launch_gram_reduce_group(const BatchJob *jobs, int njobs, int R, hipStream_t s) {
    if (njobs <= 0) return ALTO_OK;
    const int RR = R * R, per = (int)cdiv(RR, 4);
    k_gram_reduce_group<<<njobs * per, 256, 0, s>>>(jobs, per, RR);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

}  // namespace alto
