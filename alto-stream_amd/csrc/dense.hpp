// Dense R x R / I x R kernels shared by CP-ALS and CP-stream.
#pragma once
#include <math.h>

#include "common.hpp"

namespace alto {

constexpr int kDenseThreads = 256;
constexpr int kMaxBlocks = 2048;  // row blocks per reduction (partials buffer size)
constexpr int kMaxRank = 128;

// One wave per output entry: lane l sums (op 0) or maxes (op 1)
// part[b * stride + off] over blocks b = l, l + 64, ... into four
// accumulators taken in turn (four loads in flight), combined as
// (a0 + a1) + (a2 + a3), then a fixed shuffle tree; the result is valid in
// lane 0.  Deterministic for a given nblk.
__device__ __forceinline__ double wave_reduce_entry(const double *__restrict__ part, int nblk,
                                                    size_t stride, size_t off, int op) {
    const int lane = threadIdx.x & 63;
    const double init = op == 0 ? 0.0 : -INFINITY;
    double a0 = init, a1 = init, a2 = init, a3 = init;
    int b = lane;
    for (; b + 192 < nblk; b += 256) {
        const double v0 = part[(size_t)b * stride + off];
        const double v1 = part[(size_t)(b + 64) * stride + off];
        const double v2 = part[(size_t)(b + 128) * stride + off];
        const double v3 = part[(size_t)(b + 192) * stride + off];
        if (op == 0) {
            a0 += v0;
            a1 += v1;
            a2 += v2;
            a3 += v3;
        } else {
            a0 = v0 > a0 ? v0 : a0;
            a1 = v1 > a1 ? v1 : a1;
            a2 = v2 > a2 ? v2 : a2;
            a3 = v3 > a3 ? v3 : a3;
        }
    }
    for (; b < nblk; b += 64) {
        const double v = part[(size_t)b * stride + off];
        a0 = op == 0 ? a0 + v : (v > a0 ? v : a0);
    }
    double a;
    if (op == 0) {
        a = (a0 + a1) + (a2 + a3);
    } else {
        const double m01 = a1 > a0 ? a1 : a0, m23 = a3 > a2 ? a3 : a2;
        a = m23 > m01 ? m23 : m01;
    }
    for (int o = 32; o > 0; o >>= 1) {
        double v = __shfl_down(a, o, 64);
        a = op == 0 ? a + v : (v > a ? v : a);
    }
    return a;
}

// Launch helpers (all FP64 R x R math, row-major).
// grams: nmode consecutive R x R matrices.
int launch_fill_zero(double *p, uint64_t n, hipStream_t s);
// vinv <- inverse of (Hadamard_{m != mode} G_m + diag_add I) via Cholesky; on a
// failed pivot, the diagonal minimum-norm inverse of the unregularised matrix
// (reference fallback); flag[0] = 1 when the fallback ran.
// fail_mode 1: identity on failure (pseudo_inverse(A, B) leaves B unchanged,
// matrix.cpp:226-250).  rho_out != nullptr: add tr(Phi)/R instead of
// diag_add and store it there (admm.cpp:254-264).
int launch_chol_inv(const double *grams, int nmode, int mode, int R, double diag_add,
                    bool copy_first, double *vinv, int *flag, double *phi_out, hipStream_t s,
                    int fail_mode = 0, double *rho_out = nullptr);
int launch_cross_gram(const double *P, const double *U, uint64_t rows, int R, double *gpart,
                      int nblk, hipStream_t s);
int launch_colsumsq(const double *U, uint64_t rows, int R, double *partials, int nblk,
                    hipStream_t s);
int row_blocks(uint64_t rows);
// Row blocks of the Gram partials: row_blocks capped so that the blocks of the
// matrix-core Gram kernel just fill the device at its occupancy (rank 64: 2
// waves per SIMD, rank 32: 5) and the R x R partials stay small.
int gram_blocks(uint64_t rows, int R);
// u = M * vinv in place; partials[b][0..R) = sum u^2, [R..2R) = max u,
// [2R..3R) = sum u * M (per row block b).
int launch_solve(double *M, uint64_t rows, int R, const double *vinv, double *partials,
                 int nblk, hipStream_t s);
// lambda from partials: which 0 = 2-norm, 1 = max-norm.
int launch_lambda(const double *partials, int nblk, int R, int which, double *lambda,
                  hipStream_t s);
// Stores U = u / lambda (or u when lambda == nullptr) in precision prec,
// and writes Gram partials of the stored values.
// zero (optional, never U): u's storage, left zero in [0, zero_n) afterwards
// (each element is zeroed after its value is read; zero_n >= rows R), so
// the next MTTKRP can write into it without a separate fill.
int launch_scale_gram(const double *u, const double *lambda, uint64_t rows, int R, void *U,
                      int prec, double *gpart, int nblk, hipStream_t s, double *zero = nullptr,
                      uint64_t zero_n = 0);
// Gram partials of an existing factor stored in precision prec.
int launch_gram(const void *U, int prec, uint64_t rows, int R, double *gpart, int nblk,
                hipStream_t s);
int launch_gram_reduce(const double *gpart, int nblk, int R, double *G, hipStream_t s);
// ---- grouped launches: one launch runs a step for many CP-ALS instances ----
// One instance's buffers for the dense steps of one mode update.
struct BatchJob {
    const double *u;      // rows x R FP64 input of the scale/Gram step (mt, or the initial factor)
    const double *grams;  // the instance's nmode Gram matrices
    double *G;            // Gram output: grams + mode R x R
    double *vinv;         // R x R
    int *flag;
    double *mt;           // MTTKRP output, solved in place, rows x R FP64
    uint64_t rows;
    double *part;         // solve partials, nb_solve x 3R
    double *lam;          // R
    void *U;              // factor storage (the instance's precision)
    double *gpart;        // Gram partials, nb_gram x R x R
    double *zero;         // zeroed in [0, zero_n) by the scale step (the next MTTKRP's output)
    uint64_t zero_n;
    const int *active;    // 0: the instance has stopped; every grouped kernel skips it
    int nb_solve, nb_gram;
};

// One instance's fit and stopping state in a grouped CP-ALS.
struct FitJob {
    const double *grams, *lam, *part, *normX;
    double *fitd;  // (fit, exact fit) per iteration
    int *cnt;      // iterations done
    int *active;
    int nblk;      // solve partials of the last mode
    int max_iters;
    double epsilon;
};

// Grouped forms of launch_chol_inv (diag_add, no phi/rho), launch_solve,
// launch_lambda, launch_scale_gram (lambda_on 0: no scaling, no zeroing; u
// from the job) and launch_gram_reduce, for R in {8, 16, 32, 64}.  jobs and
// first are device arrays; first has njobs + 1 entries (prefix sums of
// nb_solve for the solve, of nb_gram for the scale step).
int launch_inv_group(const BatchJob *jobs, int njobs, int nmode, int mode, int R, double diag_add,
                     hipStream_t s);
int launch_solve_group(const BatchJob *jobs, const unsigned *first, int njobs, unsigned nblocks, int R,
                       hipStream_t s);
int launch_lambda_group(const BatchJob *jobs, int njobs, int R, int which, hipStream_t s);
int launch_scale_gram_group(const BatchJob *jobs, const unsigned *first, int njobs, unsigned nblocks, int R,
                            int prec, int lambda_on, hipStream_t s);
int launch_gram_reduce_group(const BatchJob *jobs, int njobs, int R, hipStream_t s);
// Gram partials of the stored factors U (launch_gram), first: prefix sums of nb_gram.
int launch_gram_group(const BatchJob *jobs, const unsigned *first, int njobs, unsigned nblocks, int R, int prec,
                      hipStream_t s);
bool group_rank_supported(int R);

constexpr int kSumsqBlocks = 1024;
// out = sum of v^2 (fixed order); part: kSumsqBlocks doubles of scratch.
int launch_sumsq(const double *v, uint64_t n, double *out, double *part, hipStream_t s);

}  // namespace alto
