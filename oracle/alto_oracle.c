/*
 * alto_oracle.c -- CPU FP64 restatement of the ALTO-stream sparse CP hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle.  The parity tests,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it as a
 * checker (and, for the baseline, as the timed CPU port of the reference's
 * OpenMP algorithm).  No product code links, loads or calls it.
 *
 * Every function cites the reference file:line it restates; paths are
 * relative to the reference repository root (jeewhanchoi/ALTO-stream).
 *
 * Deliberate, documented choices where the reference is under-specified:
 *  - sort_alto uses std::sort (unstable; alto.hpp:1081).  The oracle sorts
 *    stably, so ties (duplicate coordinates) keep input order.  On distinct
 *    keys the two orders are identical; with duplicates the key sequence and
 *    the per-key multiset of values are identical.
 *  - Gram lower triangles are never written by the reference
 *    (gram.cpp:29-42 writes the upper triangle; matrix.cpp:3-11 leaves
 *    memory uninitialised).  The oracle tracks full R x R arrays exactly as
 *    the reference writes them, with the lower triangles starting at zero.
 *  - glibc rand() initialisation (util.cpp:28-43) is replaced by explicit
 *    initial factor matrices passed by the caller.
 */
#include "alto_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

void orc_set_threads(int nthreads) { g_threads = nthreads > 0 ? nthreads : 1; }
int orc_get_threads(void) { return g_threads; }

/* util.hpp:50-59 get_num_bits, per mode: 64 - clz(dim - 1).  A mode of size
 * 1 has clz(0), undefined in C; LZCNT gives 64, i.e. 0 bits. */
int orc_num_bits(uint64_t dim)
{
    if (dim <= 1) return 0;
    return 64 - __builtin_clzll(dim - 1);
}

/* alto.hpp:954-1057 setup_packed_alto(at, LSB_FIRST, SHORT_FIRST).
 * Modes are ordered by bit count ascending; std::sort on <= 16 elements is an
 * insertion sort, so ties keep mode order (alto.hpp:1013).  At each level,
 * every mode that still has bits receives the next key bit, LSB first
 * (alto.hpp:1022-1040).  Returns the total bit count, -1 if it exceeds the
 * key width (alto.hpp:995 asserts). */
int orc_alto_masks(int nmode, const uint64_t *dims, int kw, uint64_t *masks)
{
    int bits[ORC_MAX_MODES], order[ORC_MAX_MODES];
    int total = 0, maxb = 0;
    for (int n = 0; n < nmode; ++n) {
        bits[n] = orc_num_bits(dims[n]);
        total += bits[n];
        if (bits[n] > maxb) maxb = bits[n];
        order[n] = n;
    }
    if (total > 64 * kw) return -1;
    for (int i = 1; i < nmode; ++i) {            /* stable insertion sort */
        int v = order[i], j = i - 1;
        while (j >= 0 && bits[order[j]] > bits[v]) { order[j + 1] = order[j]; --j; }
        order[j + 1] = v;
    }
    memset(masks, 0, sizeof(uint64_t) * (size_t)nmode * kw);
    int shift = 0;
    for (int level = 0; level < maxb; ++level) {
        for (int k = 0; k < nmode; ++k) {
            int m = order[k];
            if (level < bits[m]) {
                masks[m * kw + shift / 64] |= 1ULL << (shift % 64);
                ++shift;
            }
        }
    }
    return total;
}

/* bitops.hpp:9-15 / 60-79: pdep/pext, as defined for BMI2 (Intel SDM): the
 * b-th set bit of the mask (counting from bit 0 of word 0) receives bit b of
 * x, and pext is its inverse.  The 128-bit form in the reference splits at
 * popcount(mask_lo), which is the same bit numbering. */
void orc_pdep(uint64_t x, const uint64_t *mask, int kw, uint64_t *out)
{
    int b = 0;
    for (int w = 0; w < kw; ++w) {
        uint64_t m = mask[w], r = 0;
        while (m) {
            uint64_t low = m & (~m + 1);
            if (b < 64 && ((x >> b) & 1ULL)) r |= low;
            m ^= low;
            ++b;
        }
        out[w] = r;
    }
}

uint64_t orc_pext(const uint64_t *key, const uint64_t *mask, int kw)
{
    uint64_t r = 0;
    int b = 0;
    for (int w = 0; w < kw; ++w) {
        uint64_t m = mask[w];
        while (m) {
            uint64_t low = m & (~m + 1);
            if (key[w] & low) r |= 1ULL << b;
            m ^= low;
            ++b;
        }
    }
    return r;
}

/* alto.hpp:515-531 (create_alto) / 241-250 (update_salto): key_i =
 * OR_m pdep(cidx[m][i], mask_m). */
void orc_linearize(int nmode, uint64_t nnz, const uint64_t *const *cidx,
                   const uint64_t *masks, int kw, uint64_t *keys)
{
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (uint64_t i = 0; i < nnz; ++i) {
        uint64_t acc[2] = {0, 0}, part[2];
        for (int m = 0; m < nmode; ++m) {
            orc_pdep(cidx[m][i], masks + (size_t)m * kw, kw, part);
            for (int w = 0; w < kw; ++w) acc[w] |= part[w];
        }
        for (int w = 0; w < kw; ++w) keys[i * kw + w] = acc[w];
    }
}

static int key_less(const uint64_t *keys, int kw, uint64_t a, uint64_t b)
{
    for (int w = kw - 1; w >= 0; --w) {
        uint64_t x = keys[a * kw + w], y = keys[b * kw + w];
        if (x != y) return x < y;
    }
    return 0;
}

/* Stable bottom-up merge sort of an index array by key. */
static void merge_sort_idx(uint64_t n, uint64_t *idx,
                           int (*less)(const void *, uint64_t, uint64_t),
                           const void *ctx)
{
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    uint64_t *src = idx, *dst = tmp;
    for (uint64_t width = 1; width < n; width *= 2) {
        #pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
        for (uint64_t lo = 0; lo < n; lo += 2 * width) {
            uint64_t mid = lo + width < n ? lo + width : n;
            uint64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            uint64_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) {
                if (less(ctx, src[j], src[i])) dst[k++] = src[j++];
                else dst[k++] = src[i++];
            }
            while (i < mid) dst[k++] = src[i++];
            while (j < hi) dst[k++] = src[j++];
        }
        uint64_t *t = src; src = dst; dst = t;
    }
    if (src != idx) memcpy(idx, src, sizeof(uint64_t) * n);
    free(tmp);
}

struct key_ctx { const uint64_t *keys; int kw; };
static int key_ctx_less(const void *c, uint64_t a, uint64_t b)
{
    const struct key_ctx *k = (const struct key_ctx *)c;
    return key_less(k->keys, k->kw, a, b);
}

/* alto.hpp:1059-1090 sort_alto: sort (key, val) pairs by key.  The oracle
 * sort is stable (see file header).  perm (optional) receives, for each
 * sorted position, the original nonzero index. */
void orc_sort_alto(uint64_t nnz, int kw, uint64_t *keys, double *vals,
                   uint64_t *perm)
{
    uint64_t *idx = (uint64_t *)malloc(sizeof(uint64_t) * (nnz ? nnz : 1));
    for (uint64_t i = 0; i < nnz; ++i) idx[i] = i;
    struct key_ctx ctx = {keys, kw};
    merge_sort_idx(nnz, idx, key_ctx_less, &ctx);
    uint64_t *k2 = (uint64_t *)malloc(sizeof(uint64_t) * (nnz ? nnz : 1) * kw);
    double *v2 = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
    for (uint64_t i = 0; i < nnz; ++i) {
        for (int w = 0; w < kw; ++w) k2[i * kw + w] = keys[idx[i] * kw + w];
        v2[i] = vals ? vals[idx[i]] : 0.0;
    }
    memcpy(keys, k2, sizeof(uint64_t) * nnz * kw);
    if (vals) memcpy(vals, v2, sizeof(double) * nnz);
    if (perm) memcpy(perm, idx, sizeof(uint64_t) * nnz);
    free(k2); free(v2); free(idx);
}

/* alto.hpp:1128-1226 prtn_alto: nnz_ptrn = ceil(nnz/nprtn); partition p covers
 * [min(p*nnz_ptrn, nnz), min((p+1)*nnz_ptrn, nnz)); intervals[p][m] = [min,
 * max] of pext(key, mask_m) over the partition, [0,0] when empty. */
void orc_prtn_alto(int nmode, const uint64_t *dims, uint64_t nnz, int kw,
                   const uint64_t *keys, const uint64_t *masks, int nprtn,
                   uint64_t *prtn_ptr, uint64_t *intervals)
{
    uint64_t per = (nnz + nprtn - 1) / nprtn;
    prtn_ptr[0] = 0;
    for (int p = 0; p < nprtn; ++p) {
        uint64_t s = (uint64_t)p * per, e = (uint64_t)(p + 1) * per;
        if (e > nnz) e = nnz;
        if (s > e) s = e;
        (void)s;
        prtn_ptr[p + 1] = e;
    }
    #pragma omp parallel for num_threads(g_threads) schedule(static, 1)
    for (int p = 0; p < nprtn; ++p) {
        uint64_t lo[ORC_MAX_MODES], hi[ORC_MAX_MODES];
        for (int n = 0; n < nmode; ++n) { lo[n] = dims[n]; hi[n] = 0; }
        for (uint64_t i = prtn_ptr[p]; i < prtn_ptr[p + 1]; ++i) {
            for (int n = 0; n < nmode; ++n) {
                uint64_t x = orc_pext(keys + i * kw, masks + (size_t)n * kw, kw);
                if (x < lo[n]) lo[n] = x;
                if (x > hi[n]) hi[n] = x;
            }
        }
        for (int n = 0; n < nmode; ++n) {
            if (lo[n] > hi[n]) lo[n] = hi[n];
            intervals[((size_t)p * nmode + n) * 2 + 0] = lo[n];
            intervals[((size_t)p * nmode + n) * 2 + 1] = hi[n];
        }
    }
}

/* alto.hpp:1250-1346 GENERATE_CR_BITS (at the end of prtn_alto).  Free bits
 * = 64*kw - alto_bits (alto_bits = popcount of all masks).  Mode n takes the
 * bit 1 << (64*kw - free_bits) when dims[n] > MIN_NUM_FIBERS (512),
 * nnz / dims[n] <= MIN_FIBER_REUSE (4) and free_bits > 0, then free_bits--.
 * Per such mode an owner array over its rows: 0 untouched, p+1 touched only
 * by partition p, -1 touched by more than one (the reference's CAS loop
 * reaches the same state in any order); every nonzero whose row owner is -1
 * gets the mode's bit OR'ed into its key.  keys (nnz x kw) are updated in
 * place; cr_masks (nmode x kw words) receive each mode's bit (zero: none).
 * Returns the number of modes with a bit. */
int orc_cr_bits(int nmode, const uint64_t *dims, uint64_t nnz, int kw, uint64_t *keys,
                const uint64_t *masks, int nprtn, const uint64_t *prtn_ptr, uint64_t *cr_masks)
{
    int alto_bits = 0;
    for (int n = 0; n < nmode; ++n)
        for (int w = 0; w < kw; ++w) alto_bits += __builtin_popcountll(masks[(size_t)n * kw + w]);
    int free_bits = 64 * kw - alto_bits, ncr = 0;
    memset(cr_masks, 0, sizeof(uint64_t) * (size_t)nmode * kw);
    for (int n = 0; n < nmode; ++n) {
        uint64_t reuse = nnz / dims[n];
        if (dims[n] > 512 && reuse <= 4 && free_bits > 0) {
            int bit = 64 * kw - free_bits;
            cr_masks[(size_t)n * kw + bit / 64] = 1ULL << (bit % 64);
            free_bits--;
            ncr++;
        }
    }
    for (int n = 0; n < nmode; ++n) {
        int has = 0;
        for (int w = 0; w < kw; ++w) has |= cr_masks[(size_t)n * kw + w] != 0;
        if (!has) continue;
        int *owner = (int *)calloc(dims[n], sizeof(int));
        for (int p = 0; p < nprtn; ++p)
            for (uint64_t i = prtn_ptr[p]; i < prtn_ptr[p + 1]; ++i) {
                uint64_t row = orc_pext(keys + i * kw, masks + (size_t)n * kw, kw);
                if (owner[row] == 0) owner[row] = p + 1;
                else if (owner[row] != p + 1) owner[row] = -1;
            }
        for (uint64_t i = 0; i < nnz; ++i) {
            uint64_t row = orc_pext(keys + i * kw, masks + (size_t)n * kw, kw);
            if (owner[row] == -1)
                for (int w = 0; w < kw; ++w) keys[i * kw + w] |= cr_masks[(size_t)n * kw + w];
        }
        free(owner);
    }
    return ncr;
}

/* util.hpp:30-48 get_num_ptrn (partition count used by CP-stream). */
int orc_get_num_ptrn(uint64_t nnz, int max_threads)
{
    int np = max_threads;
    uint64_t per = (nnz + np - 1) / np;
    if (nnz < (uint64_t)np) {
        np = 1;
    } else {
        while (per < (uint64_t)max_threads) {
            np /= 2;
            per = (nnz + np - 1) / np;
        }
    }
    return np;
}

/* mttkrp.cpp:122-156 mttkrp: sequential COO MTTKRP, out += (the `-c` oracle
 * of the reference, main.cpp:563-578). */
void orc_mttkrp_coo(int nmode, uint64_t nnz, const uint64_t *const *cidx,
                    const double *vals, double *const *factors, int rank,
                    int mode, double *out)
{
    double *row = (double *)malloc(sizeof(double) * rank);
    for (uint64_t i = 0; i < nnz; ++i) {
        for (int r = 0; r < rank; ++r) row[r] = vals[i];
        for (int m = 0; m < nmode; ++m) {
            if (m == mode) continue;
            const double *f = factors[m] + cidx[m][i] * (uint64_t)rank;
            for (int r = 0; r < rank; ++r) row[r] *= f[r];
        }
        double *o = out + cidx[mode][i] * (uint64_t)rank;
        for (int r = 0; r < rank; ++r) o[r] += row[r];
    }
    free(row);
}

/* alto.hpp:835-846 (and 902-908): variant choice of mttkrp_alto_par with
 * MIN_NUM_FIBERS = 512, MIN_FIBER_REUSE = 4 (alto.hpp:118-121); no oidx and
 * no conflict-resolution masks on the default build. */
int orc_mttkrp_variant(int nmode, const uint64_t *dims, uint64_t nnz,
                       int stream_mode, int target)
{
    (void)nmode;
    uint64_t reuse = nnz / dims[target];
    if (dims[target] > 512 && reuse <= 4) return ORC_VAR_ATOMIC;
    if (target == stream_mode) return ORC_VAR_STREAM;
    return ORC_VAR_DA_MEM_PULL;
}

static inline void nnz_row(int nmode, int kw, const uint64_t *key, double val,
                           const uint64_t *masks, double *const *factors,
                           int rank, int target, double *row)
{
    for (int r = 0; r < rank; ++r) row[r] = val;
    for (int m = 0; m < nmode; ++m) {
        if (m == target) continue;
        uint64_t id = orc_pext(key, masks + (size_t)m * kw, kw);
        const double *f = factors[m] + id * (uint64_t)rank;
        for (int r = 0; r < rank; ++r) row[r] *= f[r];
    }
}

/* alto.hpp:1349-1407 (_atomic), 1636-1706 (_stream), 1544-1634
 * (_da_mem_pull): out += MTTKRP(target).  `out` plays the role of
 * factors[target] in the reference.  With more than one oracle thread the
 * atomic variant is order-nondeterministic, as in the reference. */
int orc_mttkrp_alto(int nmode, const uint64_t *dims, uint64_t nnz, int kw,
                    const uint64_t *keys, const double *vals,
                    const uint64_t *masks, int nprtn, const uint64_t *prtn_ptr,
                    const uint64_t *intervals, int stream_mode,
                    double *const *factors, int rank, int target, double *out)
{
    int var = orc_mttkrp_variant(nmode, dims, nnz, stream_mode, target);
    if (var == ORC_VAR_ATOMIC) {
        #pragma omp parallel for num_threads(g_threads) schedule(static, 1)
        for (int p = 0; p < nprtn; ++p) {
            double *row = (double *)malloc(sizeof(double) * rank);
            for (uint64_t i = prtn_ptr[p]; i < prtn_ptr[p + 1]; ++i) {
                nnz_row(nmode, kw, keys + i * kw, vals[i], masks, factors, rank,
                        target, row);
                uint64_t id = orc_pext(keys + i * kw, masks + (size_t)target * kw, kw);
                double *o = out + id * (uint64_t)rank;
                for (int r = 0; r < rank; ++r) {
                    #pragma omp atomic update
                    o[r] += row[r];
                }
            }
            free(row);
        }
    } else if (var == ORC_VAR_STREAM) {
        double *part = (double *)calloc((size_t)nprtn * rank, sizeof(double));
        #pragma omp parallel for num_threads(g_threads) schedule(static, 1)
        for (int p = 0; p < nprtn; ++p) {
            double *row = (double *)malloc(sizeof(double) * rank);
            double *acc = part + (size_t)p * rank;
            for (uint64_t i = prtn_ptr[p]; i < prtn_ptr[p + 1]; ++i) {
                nnz_row(nmode, kw, keys + i * kw, vals[i], masks, factors, rank,
                        target, row);
                for (int r = 0; r < rank; ++r) acc[r] += row[r];
            }
            free(row);
        }
        for (int p = 0; p < nprtn; ++p)
            for (int r = 0; r < rank; ++r) out[r] += part[(size_t)p * rank + r];
        free(part);
    } else {
        double **ofibs = (double **)malloc(sizeof(double *) * nprtn);
        #pragma omp parallel for num_threads(g_threads) schedule(static, 1)
        for (int p = 0; p < nprtn; ++p) {
            uint64_t lo = intervals[((size_t)p * nmode + target) * 2];
            uint64_t hi = intervals[((size_t)p * nmode + target) * 2 + 1];
            double *o = (double *)calloc((hi - lo + 1) * rank, sizeof(double));
            double *row = (double *)malloc(sizeof(double) * rank);
            for (uint64_t i = prtn_ptr[p]; i < prtn_ptr[p + 1]; ++i) {
                nnz_row(nmode, kw, keys + i * kw, vals[i], masks, factors, rank,
                        target, row);
                uint64_t id = orc_pext(keys + i * kw, masks + (size_t)target * kw, kw) - lo;
                for (int r = 0; r < rank; ++r) o[id * rank + r] += row[r];
            }
            free(row);
            ofibs[p] = o;
        }
        uint64_t I = dims[target];
        #pragma omp parallel for num_threads(g_threads) schedule(static)
        for (uint64_t i = 0; i < I; ++i) {
            for (int p = 0; p < nprtn; ++p) {
                uint64_t lo = intervals[((size_t)p * nmode + target) * 2];
                uint64_t hi = intervals[((size_t)p * nmode + target) * 2 + 1];
                if (i >= lo && i <= hi) {
                    const double *o = ofibs[p] + (i - lo) * rank;
                    for (int r = 0; r < rank; ++r) out[i * rank + r] += o[r];
                }
            }
        }
        for (int p = 0; p < nprtn; ++p) free(ofibs[p]);
        free(ofibs);
    }
    return var;
}

/* gram.cpp:29-42 update_gram: row-major cblas_dsyrk(Upper, Trans), G = U^T U;
 * only the upper triangle (j >= i) is written. */
void orc_gram_upper(uint64_t rows, int rank, const double *U, double *G)
{
    for (int i = 0; i < rank; ++i)
        for (int j = i; j < rank; ++j) {
            double s = 0.0;
            for (uint64_t k = 0; k < rows; ++k)
                s += U[k * rank + i] * U[k * rank + j];
            G[i * rank + j] = s;
        }
}

/* LAPACK dpotrf('L') on a column-major array is a Cholesky factorisation that
 * reads and writes the row-major upper triangle: A = U^T U with U stored in
 * A[i*R + j], j >= i.  Returns 0 on success, k+1 when pivot k fails. */
static int chol_upper(int R, double *A)
{
    for (int j = 0; j < R; ++j) {
        double s = A[j * R + j];
        for (int k = 0; k < j; ++k) s -= A[k * R + j] * A[k * R + j];
        if (!(s > 0.0)) return j + 1;
        double d = sqrt(s);
        A[j * R + j] = d;
        for (int i = j + 1; i < R; ++i) {
            double t = A[j * R + i];
            for (int k = 0; k < j; ++k) t -= A[k * R + j] * A[k * R + i];
            A[j * R + i] = t / d;
        }
    }
    return 0;
}

/* dpotrs('L') on one right-hand side with the factor from chol_upper. */
static void chol_solve_row(int R, const double *A, double *x)
{
    for (int i = 0; i < R; ++i) {
        double t = x[i];
        for (int k = 0; k < i; ++k) t -= A[k * R + i] * x[k];
        x[i] = t / A[i * R + i];
    }
    for (int i = R - 1; i >= 0; --i) {
        double t = x[i];
        for (int k = i + 1; k < R; ++k) t -= A[i * R + k] * x[k];
        x[i] = t / A[i * R + i];
    }
}

/* Minimum-norm least-squares solve of S x = b for symmetric S (the result
 * LAPACK dgelsy returns when the numerical rank is clear-cut), by a cyclic
 * Jacobi eigendecomposition; eigenvalues below R*eps*max|lambda| are
 * treated as zero, as dgelsy does with rcond = -1 (machine precision). */
static void minnorm_solve_sym(int R, const double *S, uint64_t rows, double *B)
{
    double *A = (double *)malloc(sizeof(double) * R * R);
    double *V = (double *)calloc((size_t)R * R, sizeof(double));
    memcpy(A, S, sizeof(double) * R * R);
    for (int i = 0; i < R; ++i) V[i * R + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < R; ++p)
            for (int q = p + 1; q < R; ++q) off += A[p * R + q] * A[p * R + q];
        if (off < 1e-300) break;
        for (int p = 0; p < R; ++p)
            for (int q = p + 1; q < R; ++q) {
                double apq = A[p * R + q];
                if (fabs(apq) < 1e-300) continue;
                double theta = (A[q * R + q] - A[p * R + p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) /
                           (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < R; ++k) {
                    double akp = A[k * R + p], akq = A[k * R + q];
                    A[k * R + p] = c * akp - s * akq;
                    A[k * R + q] = s * akp + c * akq;
                }
                for (int k = 0; k < R; ++k) {
                    double apk = A[p * R + k], aqk = A[q * R + k];
                    A[p * R + k] = c * apk - s * aqk;
                    A[q * R + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < R; ++k) {
                    double vkp = V[k * R + p], vkq = V[k * R + q];
                    V[k * R + p] = c * vkp - s * vkq;
                    V[k * R + q] = s * vkp + c * vkq;
                }
            }
    }
    double amax = 0.0;
    for (int i = 0; i < R; ++i) if (fabs(A[i * R + i]) > amax) amax = fabs(A[i * R + i]);
    double tol = amax * R * 2.220446049250313e-16;
    double *y = (double *)malloc(sizeof(double) * R);
    for (uint64_t r = 0; r < rows; ++r) {
        double *b = B + r * R;
        for (int k = 0; k < R; ++k) {
            double s = 0.0;
            for (int i = 0; i < R; ++i) s += V[i * R + k] * b[i];
            y[k] = fabs(A[k * R + k]) > tol ? s / A[k * R + k] : 0.0;
        }
        for (int i = 0; i < R; ++i) {
            double s = 0.0;
            for (int k = 0; k < R; ++k) s += V[i * R + k] * y[k];
            b[i] = s;
        }
    }
    free(y); free(A); free(V);
}

/* cpd.hpp:377-480 pseudo_inverse(grams, M, mode): grams[mode] <- Hadamard of
 * the other Grams over the full R x R arrays (copy of grams[m0], then
 * multiplied in mode order), +1e-3 on the diagonal, Cholesky; on success
 * every row of U is solved in place.  On failure the unregularised V is
 * restored, its upper triangle is overwritten from its lower triangle, and a
 * minimum-norm solve is used (dgelsy, cpd.hpp:438-476).  Returns 0 on the
 * Cholesky path, 1 on the fallback. */
int orc_pseudo_inverse_als(int nmode, int rank, double *const *grams,
                           int mode, uint64_t rows, double *U)
{
    int R = rank;
    int m0 = mode == 0 ? 1 : 0;
    double *G = grams[mode];
    memcpy(G, grams[m0], sizeof(double) * R * R);
    for (int i = m0 + 1; i < nmode; ++i)
        if (i != mode)
            for (int j = 0; j < R * R; ++j) G[j] *= grams[i][j];
    double *saved = (double *)malloc(sizeof(double) * R * R);
    memcpy(saved, G, sizeof(double) * R * R);
    for (int r = 0; r < R; ++r) G[r * R + r] += 1e-3;
    int fallback = 0;
    if (chol_upper(R, G) == 0) {
        #pragma omp parallel for num_threads(g_threads) schedule(static)
        for (uint64_t i = 0; i < rows; ++i) chol_solve_row(R, G, U + i * R);
    } else {
        fallback = 1;
        memcpy(G, saved, sizeof(double) * R * R);
        for (int i = 0; i < R; ++i)
            for (int j = i; j < R; ++j) G[i * R + j] = G[j * R + i];
        minnorm_solve_sym(R, G, rows, U);
    }
    free(saved);
    return fallback;
}

/* kruskal_model.cpp:227-353 KruskalModelNorm.  which = 0 (MAT_NORM_2):
 * lambda_j = sqrt(max(sum_i U_ij^2, 1e-12)); which = 1 (MAT_NORM_MAX):
 * lambda_j = max(1, max_i U_ij) (signed maximum).  U /= lambda; lambda is
 * overwritten. */
void orc_kruskal_norm(uint64_t rows, int rank, double *U, double *lambda,
                      int which)
{
    for (int r = 0; r < rank; ++r) lambda[r] = which == 0 ? 0.0 : 1.0;
    for (uint64_t i = 0; i < rows; ++i)
        for (int r = 0; r < rank; ++r) {
            double v = U[i * rank + r];
            if (which == 0) lambda[r] += v * v;
            else if (v > lambda[r]) lambda[r] = v;
        }
    if (which == 0)
        for (int r = 0; r < rank; ++r) {
            double l = lambda[r] > 1e-12 ? lambda[r] : 1e-12;
            lambda[r] = sqrt(l);
        }
    for (uint64_t i = 0; i < rows; ++i)
        for (int r = 0; r < rank; ++r) U[i * rank + r] /= lambda[r];
}

/* cpd.hpp:193-275 cpd_fit_alto.  <X,M> = sum_r lambda_r sum_i
 * U_last[i,r] * MTTKRP_last[i,r].  ||M||^2: ORC_FIT_REFERENCE_DIAG reads the
 * row-major lower triangle of the tracked Gram arrays, as the reference does
 * (cpd.hpp:228-259); ORC_FIT_EXACT uses the symmetric Grams built from their
 * upper triangles.  Returns r/||X|| with r = ||X||^2 + ||M||^2 - 2<X,M>,
 * square-rooted only when positive (cpd.hpp:264-268). */
double orc_cpd_fit(int nmode, const uint64_t *dims, int rank,
                   double *const *factors, const double *lambda,
                   double *const *grams, const double *last_mttkrp,
                   double normX, int fit_mode)
{
    int R = rank;
    uint64_t I = dims[nmode - 1];
    const double *U = factors[nmode - 1];
    double *acc = (double *)calloc(R, sizeof(double));
    for (uint64_t i = 0; i < I; ++i)
        for (int r = 0; r < R; ++r) acc[r] += U[i * R + r] * last_mttkrp[i * R + r];
    double inner = 0.0;
    for (int r = 0; r < R; ++r) inner += acc[r] * lambda[r];
    double normU = 0.0;
    for (int i = 0; i < R; ++i) {
        int jmax = fit_mode == ORC_FIT_EXACT ? R - 1 : i;
        for (int j = 0; j <= jmax; ++j) {
            double t = lambda[i] * lambda[j];
            for (int m = 0; m < nmode; ++m) {
                const double *G = grams[m];
                double g = fit_mode == ORC_FIT_EXACT
                               ? (j >= i ? G[i * R + j] : G[j * R + i])
                               : G[i * R + j];
                t *= g;
            }
            if (fit_mode == ORC_FIT_EXACT || j == i) normU += t;
            else normU += 2.0 * t;
        }
    }
    normU = fabs(normU);
    double res = normX + normU - 2.0 * inner;
    if (res > 0.0) res = sqrt(res);
    free(acc);
    return res / sqrt(normX);
}

/* cpd.hpp:47-191 cpd_alto: the CP-ALS loop.  factors/lambda hold the initial
 * model on entry and the final model on return; fits[i] / fits_exact[i]
 * receive the fit of iteration i in the requested convention and in the
 * exact one.  Returns the number of iterations run. */
int orc_cpd_als(int nmode, const uint64_t *dims, uint64_t nnz, int kw,
                const uint64_t *keys, const double *vals,
                const uint64_t *masks, int nprtn, const uint64_t *prtn_ptr,
                const uint64_t *intervals, int rank, int max_iters,
                double eps, int fit_mode, double *const *factors,
                double *lambda, double *fits, double *fits_exact)
{
    int R = rank;
    double normX = 0.0;
    for (uint64_t i = 0; i < nnz; ++i) normX += vals[i] * vals[i];
    double *grams[ORC_MAX_MODES];
    for (int m = 0; m < nmode; ++m) {
        grams[m] = (double *)calloc((size_t)R * R, sizeof(double));
        orc_gram_upper(dims[m], R, factors[m], grams[m]);
    }
    uint64_t Ilast = dims[nmode - 1];
    double *scratch = (double *)malloc(sizeof(double) * Ilast * R);
    double fit = 0.0, prev = 0.0;
    int it_run = max_iters;
    for (int it = 0; it < max_iters; ++it) {
        for (int j = 0; j < nmode; ++j) {
            memset(factors[j], 0, sizeof(double) * dims[j] * R);
            orc_mttkrp_alto(nmode, dims, nnz, kw, keys, vals, masks, nprtn,
                            prtn_ptr, intervals, -1, factors, R, j, factors[j]);
            if (j == nmode - 1)
                memcpy(scratch, factors[j], sizeof(double) * dims[j] * R);
            orc_pseudo_inverse_als(nmode, R, grams, j, dims[j], factors[j]);
            orc_kruskal_norm(dims[j], R, factors[j], lambda, it == 0 ? 0 : 1);
            orc_gram_upper(dims[j], R, factors[j], grams[j]);
        }
        fit = orc_cpd_fit(nmode, dims, R, factors, lambda, grams, scratch, normX,
                          fit_mode);
        if (fits) fits[it] = fit;
        if (fits_exact)
            fits_exact[it] = orc_cpd_fit(nmode, dims, R, factors, lambda, grams,
                                         scratch, normX, ORC_FIT_EXACT);
        if (it > 0 && fabs(prev - fit) < eps) { it_run = it + 1; break; }
        prev = fit;
    }
    for (int m = 0; m < nmode; ++m) free(grams[m]);
    free(scratch);
    return it_run;
}

/* admm.cpp:216-429 admm_cf_ (called through admm_cf, admm.cpp:46-66).
 * Phi = Hadamard of grams[m], m != mode (matrix.cpp:269-294).
 * Unconstrained: (Phi + 1e-12 I) x = mttkrp row by row; primal <- x.
 * Non-negative: rho = tr(Phi)/R, Cholesky of Phi + rho I, then the AO-ADMM
 * primal/dual sweep with at most 25 iterations and the float tolerance
 * 1e-6f (admm.cpp:409-413).  Row chunking (64 rows) does not change the
 * arithmetic per row.  Returns the iteration count. */
double orc_admm_cf(uint64_t rows, int rank, int nmode, int mode,
                   double *const *grams, int nonneg, double *primal,
                   double *dual, double *mttkrp)
{
    int R = rank;
    double *Phi = (double *)malloc(sizeof(double) * R * R);
    for (int i = 0; i < R * R; ++i) Phi[i] = 1.0;
    for (int m = 0; m < nmode; ++m)
        if (m != mode)
            for (int j = 0; j < R * R; ++j) Phi[j] *= grams[m][j];
    if (!nonneg) {
        for (int i = 0; i < R; ++i) Phi[i * R + i] += 1e-12;
        if (chol_upper(R, Phi) == 0) {
            for (uint64_t i = 0; i < rows; ++i) chol_solve_row(R, Phi, mttkrp + i * R);
        }
        memcpy(primal, mttkrp, sizeof(double) * rows * R);
        free(Phi);
        return 0.0;
    }
    double rho = 0.0;
    for (int i = 0; i < R; ++i) rho += Phi[i * R + i];
    rho /= (double)R;
    for (int i = 0; i < R; ++i) Phi[i * R + i] += rho;
    if (chol_upper(R, Phi) != 0) { free(Phi); return -1.0; }
    size_t n = rows * (size_t)R;
    double *aux = (double *)malloc(sizeof(double) * (n ? n : 1));
    double *init = (double *)malloc(sizeof(double) * (n ? n : 1));
    for (size_t idx = 0; idx < n; ++idx) aux[idx] = mttkrp[idx] + rho * (primal[idx] + dual[idx]);
    for (uint64_t i = 0; i < rows; ++i) chol_solve_row(R, Phi, aux + i * R);
    for (size_t idx = 0; idx < n; ++idx) {
        init[idx] = primal[idx];
        double x = aux[idx] - dual[idx];
        primal[idx] = x < 0.0 ? 0.0 : x;
    }
    const double tol = (double)1e-6f;
    int it;
    for (it = 0; it < 25; ++it) {
        double p_res = 0, d_res = 0, p_norm = 0, d_norm = 0;
        for (size_t idx = 0; idx < n; ++idx) {
            double x = primal[idx];
            double dd = x - init[idx];
            d_res += dd * dd;
            p_norm += x * x;
            double y = x - aux[idx];
            double di = dual[idx] + y;
            dual[idx] = di;
            d_norm += di * di;
            p_res += y * y;
            aux[idx] = mttkrp[idx] + rho * (x + di);
        }
        for (uint64_t i = 0; i < rows; ++i) chol_solve_row(R, Phi, aux + i * R);
        for (size_t idx = 0; idx < n; ++idx) {
            double x = aux[idx] - dual[idx];
            x = x > 0.0 ? x : 0.0;
            init[idx] = x;
            primal[idx] = x;
        }
        if (p_res <= tol * p_norm && d_res <= tol * d_norm) { ++it; break; }
    }
    free(aux); free(init); free(Phi);
    return (double)it;
}

/* ---------------- CP-stream ---------------- */

struct lex_ctx { const uint64_t *const *cidx; const int *order; int nmode; };
static int lex_less(const void *c, uint64_t a, uint64_t b)
{
    const struct lex_ctx *k = (const struct lex_ctx *)c;
    for (int t = 0; t < k->nmode; ++t) {
        uint64_t x = k->cidx[k->order[t]][a], y = k->cidx[k->order[t]][b];
        if (x != y) return x < y;
    }
    return 0;
}

static double frob(const double *a, size_t n)
{
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * a[i];
    return sqrt(s);
}

/* next_dynamic_batch (streaming_sptensor.cpp:196-264) over the sorted and
 * relabelled stream: from the first nonzero of time index *batch, take whole
 * time slices while fewer than thr nonzeros are taken, stopping once lim
 * slices are taken; the batch's stream index is the time index minus the
 * batch's first (bc[s]), the other modes' indices are copied (bc[m]) and
 * their dims are the running max of (index + 1).  thr 0 = no threshold,
 * lim 0 counts as 1.  Returns the batch's nonzero count (0 at the end). */
static uint64_t next_batch(int nmode, int s, uint64_t nnz, uint64_t *const *cidx, uint64_t thr,
                           uint64_t lim, uint64_t *ptr, uint64_t *batch, uint64_t *prev_dim,
                           uint64_t *const *bc, uint64_t *bd, uint64_t *start_out, uint64_t *t0)
{
    if (thr == 0) thr = UINT64_MAX;
    if (lim == 0) lim = 1;
    const uint64_t prev_batch = *batch;
    if (*ptr >= nnz) return 0;
    uint64_t start = *ptr;
    while (start < nnz && cidx[s][start] < *batch) ++start;
    uint64_t end = start;
    while (end - start < thr && end < nnz) {
        while (end < nnz && cidx[s][end] < *batch + 1) {
            bc[s][end - start] = *batch - prev_batch;
            ++end;
        }
        ++*batch;
        if (*batch - prev_batch >= lim) break;
    }
    const uint64_t bn = end - start;
    for (int m = 0; m < nmode; ++m) {
        if (m == s) { bd[m] = *batch - prev_batch; continue; }
        uint64_t mx = 0;
        for (uint64_t i = 0; i < bn; ++i) {
            bc[m][i] = cidx[m][start + i];
            if (bc[m][i] > mx) mx = bc[m][i];
        }
        bd[m] = mx + 1 > prev_dim[m] ? mx + 1 : prev_dim[m];
        prev_dim[m] = bd[m];
    }
    *ptr = end;
    *start_out = start;
    *t0 = prev_batch;
    return bn;
}

/* streaming_cpd.cpp:26-366 cpstream with model CPSTREAM_ALTO,
 * driving cpstream_alto_iter (streaming_cpd.cpp:1198-1519).  Batches as
 * next_batch (timeslice_limit / nnz_threshold; the reference defaults 1 and
 * unlimited, main.cpp:130-131, are lim 1 and thr 0).
 *  - StreamingSparseTensor (streaming_sptensor.cpp:69-125): sort by
 *    (s, s+1, ...) mod N (sort.cpp:477-519; stable here), relabel every
 *    non-stream index by first appearance.
 *  - next_dynamic_batch (streaming_sptensor.cpp:196-264): next_batch above.
 *  - Model growth takes rows [old, new) of init_factors[m] (replacing the
 *    reference's rand() fill, kruskal_model.cpp:101-133); prev_M grows with
 *    zeros.
 * Outputs: final factors (out_factors[m], out_dims[m] rows), the time-mode
 * rows (one per time index), the outer-iteration count per batch, the ADMM
 * iteration total per batch, the final Grams (nmode*R*R) and the number of
 * time indices processed (out_ntime).  Returns the number of batches
 * processed (at most max_slices when > 0). */
int orc_cpstream(int nmode, const uint64_t *dims, uint64_t nnz,
                 const uint64_t *const *cidx_in, const double *vals_in,
                 int s, int rank, int max_iters, double eps,
                 int nonneg, int max_slices, uint64_t nnz_threshold, uint64_t timeslice_limit,
                 double *const *init_factors,
                 double *const *out_factors, uint64_t *out_dims,
                 double *out_time_rows, int *out_outer_iters,
                 double *out_admm_iters, double *out_grams, uint64_t *out_ntime)
{
    int R = rank;
    int order[ORC_MAX_MODES];
    for (int t = 0; t < nmode; ++t) order[t] = (s + t) % nmode;
    uint64_t *idx = (uint64_t *)malloc(sizeof(uint64_t) * nnz);
    for (uint64_t i = 0; i < nnz; ++i) idx[i] = i;
    struct lex_ctx lc = {cidx_in, order, nmode};
    merge_sort_idx(nnz, idx, lex_less, &lc);
    uint64_t *cidx[ORC_MAX_MODES];
    double *vals = (double *)malloc(sizeof(double) * nnz);
    for (int m = 0; m < nmode; ++m) {
        cidx[m] = (uint64_t *)malloc(sizeof(uint64_t) * nnz);
        for (uint64_t i = 0; i < nnz; ++i) cidx[m][i] = cidx_in[m][idx[i]];
    }
    for (uint64_t i = 0; i < nnz; ++i) vals[i] = vals_in[idx[i]];
    free(idx);
    for (int m = 0; m < nmode; ++m) {
        if (m == s) continue;
        uint64_t *perm = (uint64_t *)malloc(sizeof(uint64_t) * dims[m]);
        for (uint64_t i = 0; i < dims[m]; ++i) perm[i] = UINT64_MAX;
        uint64_t seen = 0;
        for (uint64_t i = 0; i < nnz; ++i) {
            uint64_t v = cidx[m][i];
            if (perm[v] == UINT64_MAX) perm[v] = seen++;
            cidx[m][i] = perm[v];
        }
        free(perm);
    }

    uint64_t cur[ORC_MAX_MODES] = {0}, prev_dim[ORC_MAX_MODES] = {0};
    uint64_t maxd[ORC_MAX_MODES];
    for (int m = 0; m < nmode; ++m) maxd[m] = dims[m];  /* a batch holds at most dims[s] time rows */
    double *U[ORC_MAX_MODES], *P[ORC_MAX_MODES], *G[ORC_MAX_MODES];
    for (int m = 0; m < nmode; ++m) {
        U[m] = (double *)calloc(maxd[m] * R, sizeof(double));
        P[m] = (double *)calloc(maxd[m] * R, sizeof(double));
        G[m] = (double *)calloc((size_t)R * R, sizeof(double));
    }
    double lambda[256];
    uint64_t maxall = 1;
    for (int m = 0; m < nmode; ++m) if (maxd[m] > maxall) maxall = maxd[m];
    double *buf = (double *)malloc(sizeof(double) * maxall * R);
    double *dual = (double *)malloc(sizeof(double) * maxall * R);
    double *duals[ORC_MAX_MODES];
    for (int m = 0; m < nmode; ++m) duals[m] = (double *)calloc(maxd[m] * R, sizeof(double));
    double *old_gram = (double *)malloc(sizeof(double) * R * R);
    double *hist = (double *)malloc(sizeof(double) * R * R);
    double *ata = (double *)malloc(sizeof(double) * R * R);
    double *Phi = (double *)malloc(sizeof(double) * R * R);
    (void)dual;

    uint64_t ptr = 0;
    uint64_t batch = 0;
    int slices = 0;
    uint64_t *bc[ORC_MAX_MODES];
    for (int m = 0; m < nmode; ++m) bc[m] = (uint64_t *)malloc(sizeof(uint64_t) * (nnz ? nnz : 1));
    double *bv = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * 2 * (nnz ? nnz : 1));

    /* init_salto's stream mode (streaming_cpd.cpp:62-65): the single-row
     * stream MTTKRP variant only at timeslice limit 1 */
    const int sdisp = timeslice_limit > 1 ? -1 : s;
    while (ptr < nnz && (max_slices <= 0 || slices < max_slices)) {
        uint64_t bd[ORC_MAX_MODES], start = 0, t0 = 0;
        const uint64_t bn = next_batch(nmode, s, nnz, cidx, nnz_threshold, timeslice_limit, &ptr, &batch,
                                       prev_dim, bc, bd, &start, &t0);
        if (bn == 0) break;                    /* reference asserts nnz > 0 */
        for (uint64_t i = 0; i < bn; ++i) bv[i] = vals[start + i];

        if (slices == 0) {
            for (int m = 0; m < nmode; ++m) {
                memcpy(U[m], init_factors[m], sizeof(double) * bd[m] * R);
                memset(P[m], 0, sizeof(double) * maxd[m] * R);
            }
            for (int r = 0; r < R; ++r) lambda[r] = 1.0;
            memset(U[s], 0, sizeof(double) * bd[s] * R);
            for (int m = 0; m < nmode; ++m) orc_gram_upper(bd[m], R, U[m], G[m]);
            memset(G[s], 0, sizeof(double) * R * R);
        } else {
            for (int m = 0; m < nmode; ++m) {
                if (bd[m] > cur[m]) {
                    memcpy(U[m] + cur[m] * R, init_factors[m] + cur[m] * R,
                           sizeof(double) * (bd[m] - cur[m]) * R);
                    memset(P[m] + cur[m] * R, 0, sizeof(double) * (bd[m] - cur[m]) * R);
                }
            }
            for (int m = 0; m < nmode; ++m)
                if (m != s) orc_gram_upper(bd[m], R, U[m], G[m]);
        }
        for (int m = 0; m < nmode; ++m) {
            cur[m] = bd[m];
            memset(duals[m], 0, sizeof(double) * bd[m] * R);
        }

        /* update_salto (alto.hpp:195-299) with nprtn = get_num_ptrn. */
        int kw = 0;
        int total = 0;
        for (int m = 0; m < nmode; ++m) total += orc_num_bits(bd[m]);
        kw = total > 64 ? 2 : 1;
        uint64_t masks[ORC_MAX_MODES * 2];
        orc_alto_masks(nmode, bd, kw, masks);
        orc_linearize(nmode, bn, (const uint64_t *const *)bc, masks, kw, keys);
        orc_sort_alto(bn, kw, keys, bv, NULL);
        int nprtn = orc_get_num_ptrn(bn, g_threads);
        uint64_t *pp = (uint64_t *)malloc(sizeof(uint64_t) * (nprtn + 1));
        uint64_t *iv = (uint64_t *)malloc(sizeof(uint64_t) * nprtn * nmode * 2);
        orc_prtn_alto(nmode, bd, bn, kw, keys, masks, nprtn, pp, iv);

        /* cpstream_alto_iter (streaming_cpd.cpp:1198-1519). */
        if (slices == 0) {
            for (int m = 0; m < nmode; ++m) {
                if (m == s) continue;
                orc_kruskal_norm(bd[m], R, U[m], lambda, 0);
                orc_gram_upper(bd[m], R, U[m], G[m]);
            }
            for (int r = 0; r < R; ++r) lambda[r] = 1.0;
        }
        memset(old_gram, 0, sizeof(double) * R * R);
        double delta = 0.0, prev_delta = 0.0;
        int outer = 0;
        double admm_total = 0.0;
        for (int i = 0; i < max_iters; ++i) {
            delta = 0.0;
            memset(U[s], 0, sizeof(double) * R * bd[s]);
            orc_mttkrp_alto(nmode, bd, bn, kw, keys, bv, masks, nprtn, pp, iv, sdisp,
                            U, R, s, U[s]);
            for (int q = 0; q < R * R; ++q) Phi[q] = 1.0;
            for (int m = 0; m < nmode; ++m)
                if (m != s) for (int q = 0; q < R * R; ++q) Phi[q] *= G[m][q];
            for (int q = 0; q < R; ++q) Phi[q * R + q] += 1e-12;
            if (chol_upper(R, Phi) == 0)
                for (uint64_t r = 0; r < bd[s]; ++r) chol_solve_row(R, Phi, U[s] + r * R);
            for (int a = 1; a < R; ++a)          /* copy_upper_tri (matrix.cpp:207-218) */
                for (int b = 0; b < a; ++b) G[s][b + a * R] = G[s][a + b * R];
            memcpy(old_gram, G[s], sizeof(double) * R * R);
            for (uint64_t r = 0; r < bd[s]; ++r)
                for (int n2 = 0; n2 < R; ++n2)
                    for (int m2 = 0; m2 < R; ++m2)
                        G[s][m2 + n2 * R] += U[s][r * R + m2] * U[s][r * R + n2];
            for (int j = 0; j < nmode; ++j) {
                if (j == s) continue;
                memset(buf, 0, sizeof(double) * bd[j] * R);
                orc_mttkrp_alto(nmode, bd, bn, kw, keys, bv, masks, nprtn, pp, iv,
                                sdisp, U, R, j, buf);
                memcpy(ata, old_gram, sizeof(double) * R * R);
                for (int m = 0; m < nmode; ++m) {
                    if (m == j || m == s) continue;
                    for (int a = 0; a < R; ++a)     /* prevU_m^T U_m (matrix.cpp:63-91) */
                        for (int b = 0; b < R; ++b) {
                            double t = 0.0;
                            for (uint64_t k = 0; k < bd[m]; ++k)
                                t += P[m][k * R + a] * U[m][k * R + b];
                            hist[a * R + b] = t;
                        }
                    for (int q = 0; q < R * R; ++q) ata[q] *= hist[q];
                }
                for (uint64_t k = 0; k < bd[j]; ++k)  /* buf += prevU_j * ata */
                    for (int b = 0; b < R; ++b) {
                        double t = 0.0;
                        for (int a = 0; a < R; ++a) t += P[j][k * R + a] * ata[a * R + b];
                        buf[k * R + b] += t;
                    }
                admm_total += orc_admm_cf(bd[j], R, nmode, j, G, nonneg, U[j],
                                          duals[j], buf);
                orc_gram_upper(bd[j], R, U[j], G[j]);
                size_t sz = (size_t)bd[j] * R;
                double dn = 0.0;
                for (size_t q = 0; q < sz; ++q) {
                    double d = P[j][q] - U[j][q];
                    dn += d * d;
                }
                delta += sqrt(dn) / (frob(U[j], sz) + 1e-12);
            }
            outer = i + 1;
            if (i > 0 && fabs(prev_delta - delta) < eps) break;
            prev_delta = delta;
        }
        for (int q = 0; q < R * R; ++q) G[s][q] *= 0.99;
        free(pp); free(iv);

        if (out_time_rows) memcpy(out_time_rows + (size_t)t0 * R, U[s], sizeof(double) * bd[s] * R);
        if (out_outer_iters) out_outer_iters[slices] = outer;
        if (out_admm_iters) out_admm_iters[slices] = admm_total;
        for (int m = 0; m < nmode; ++m) memcpy(P[m], U[m], sizeof(double) * bd[m] * R);
        ++slices;
    }
    for (int m = 0; m < nmode; ++m) {
        if (out_dims) out_dims[m] = cur[m];
        if (out_factors && out_factors[m]) memcpy(out_factors[m], U[m], sizeof(double) * cur[m] * R);
        if (out_grams) memcpy(out_grams + (size_t)m * R * R, G[m], sizeof(double) * R * R);
    }
    if (out_ntime) *out_ntime = batch;
    for (int m = 0; m < nmode; ++m) {
        free(U[m]); free(P[m]); free(G[m]); free(duals[m]); free(cidx[m]); free(bc[m]);
    }
    free(vals); free(buf); free(dual); free(old_gram); free(hist); free(ata); free(Phi);
    free(bv); free(keys);
    return slices;
}

/* ---------------- spCP-stream ---------------- */

/* C = op(A) B for R x R row-major arrays (matmul, matrix.cpp:92-119). */
static void rr_matmul(int R, const double *A, int ta, const double *B, double *C)
{
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < R; ++j) {
            double t = 0.0;
            for (int k = 0; k < R; ++k) t += (ta ? A[k * R + i] : A[i * R + k]) * B[k * R + j];
            C[i * R + j] = t;
        }
}

/* C = A^T B over `rows` rows of rows x R row-major arrays (matmul with
 * transA, matrix.cpp:63-91; rsp_mataTb, rowsparse_matrix.cpp:56-66). */
static void tall_atb(uint64_t rows, int R, const double *A, const double *B, double *C)
{
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < R; ++j) {
            double t = 0.0;
            for (uint64_t k = 0; k < rows; ++k) t += A[k * R + i] * B[k * R + j];
            C[i * R + j] = t;
        }
}

/* mat_form_gram (matrix.cpp:141-156): out = Hadamard of mats[m], m != mode. */
static void hadamard_excl(int nmode, int R, double *const *mats, int mode, double *out)
{
    for (int q = 0; q < R * R; ++q) out[q] = 1.0;
    for (int m = 0; m < nmode; ++m)
        if (m != mode)
            for (int q = 0; q < R * R; ++q) out[q] *= mats[m][q];
}

/* streaming_cpd.cpp:26-366 cpstream with model SPCPSTREAM_ALTO
 * (use_spcpstream, main.cpp:163, 397), batches as next_batch, driving
 * spcpstream_alto_iter (streaming_cpd.cpp:466-1085).  The stream handling,
 * model growth and outputs are those of orc_cpstream; constraints are not
 * used (spcpstream_alto_iter solves with pseudo_inverse only).  Per slice:
 *  - nonzero_slices (streaming_cpd.cpp:1124-1161): the ascending distinct
 *    indices of each mode in the slice (nz rows), the others (z rows) and
 *    the compact map ridx;
 *  - A_nz / A_nz_prev = the nz rows of U / prev U (convert_to_rspmat,
 *    rowsparse_matrix.cpp:28-54); c = U^T U (upper triangle, mat_aTa,
 *    matrix.cpp:252-267), h = prevU^T U; after the first slice
 *    c_nz_prev = A_nz_prev^T A_nz_prev (mataTa_idx_based,
 *    rowsparse_matrix.cpp:131-150) and c_z_prev = c_prev - c_nz_prev;
 *  - update_salto_rowsparse (alto.hpp:301-434): the slice's ALTO tensor over
 *    compact coordinates (dims = nz row counts);
 *  - outer iterations (streaming_cpd.cpp:712-1000): s_t by MTTKRP over the
 *    compact factors and a Cholesky solve with the Hadamard of the c's, the
 *    temporal Gram update, then per non-stream mode m:
 *      Q = Had_{k != m} h_k, Phi = Had_{k != m} c_k + 1e-12 I,
 *      A_nz = (MTTKRP_m + A_nz_prev Q) Phi^-1,
 *      c_nz = A_nz^T A_nz, h_nz = A_nz_prev^T A_nz, QPi = Q Phi^-1,
 *      h_z = c_z_prev QPi, c_z = QPi^T h_z, c = c_nz + c_z, h = h_nz + h_z,
 *      delta += sqrt(|tr c + tr c_prev - 2 tr h| / (tr c + 1e-12));
 *  - z rows become prevU QPi, nz rows A_nz (streaming_cpd.cpp:1017-1045);
 *    G_s *= 0.99; c_prev = c, c_nz_prev = c_nz.
 * A failed Cholesky leaves the right-hand side unchanged (pseudo_inverse,
 * matrix.cpp:226-250).  max_iters must be >= 1 (with 0 the reference reads
 * an uninitialised Q Phi^-1).  out_grams: the reference's grams array (the
 * non-stream modes' upper-triangle Grams from the start of the last slice,
 * and G_s), the time rows one per time index, out_ntime the number of time
 * indices.  Returns the number of batches, or -1 on bad arguments. */
int orc_spcpstream(int nmode, const uint64_t *dims, uint64_t nnz,
                   const uint64_t *const *cidx_in, const double *vals_in,
                   int s, int rank, int max_iters, double eps, int max_slices,
                   uint64_t nnz_threshold, uint64_t timeslice_limit,
                   double *const *init_factors, double *const *out_factors,
                   uint64_t *out_dims, double *out_time_rows, int *out_outer_iters,
                   double *out_grams, uint64_t *out_ntime)
{
    if (max_iters < 1) return -1;
    int R = rank;
    const size_t RR = (size_t)R * R;
    int order[ORC_MAX_MODES];
    for (int t = 0; t < nmode; ++t) order[t] = (s + t) % nmode;
    uint64_t *idx = (uint64_t *)malloc(sizeof(uint64_t) * (nnz ? nnz : 1));
    for (uint64_t i = 0; i < nnz; ++i) idx[i] = i;
    struct lex_ctx lc = {cidx_in, order, nmode};
    merge_sort_idx(nnz, idx, lex_less, &lc);
    uint64_t *cidx[ORC_MAX_MODES];
    double *vals = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
    for (int m = 0; m < nmode; ++m) {
        cidx[m] = (uint64_t *)malloc(sizeof(uint64_t) * (nnz ? nnz : 1));
        for (uint64_t i = 0; i < nnz; ++i) cidx[m][i] = cidx_in[m][idx[i]];
    }
    for (uint64_t i = 0; i < nnz; ++i) vals[i] = vals_in[idx[i]];
    free(idx);
    for (int m = 0; m < nmode; ++m) {
        if (m == s) continue;
        uint64_t *perm = (uint64_t *)malloc(sizeof(uint64_t) * dims[m]);
        for (uint64_t i = 0; i < dims[m]; ++i) perm[i] = UINT64_MAX;
        uint64_t seen = 0;
        for (uint64_t i = 0; i < nnz; ++i) {
            uint64_t v = cidx[m][i];
            if (perm[v] == UINT64_MAX) perm[v] = seen++;
            cidx[m][i] = perm[v];
        }
        free(perm);
    }

    uint64_t cur[ORC_MAX_MODES] = {0}, prev_dim[ORC_MAX_MODES] = {0}, maxd[ORC_MAX_MODES];
    for (int m = 0; m < nmode; ++m) maxd[m] = dims[m];  /* a batch holds at most dims[s] time rows */
    double *U[ORC_MAX_MODES], *P[ORC_MAX_MODES], *G[ORC_MAX_MODES];
    double *c[ORC_MAX_MODES], *h[ORC_MAX_MODES], *cnz[ORC_MAX_MODES], *hnz[ORC_MAX_MODES];
    double *cz[ORC_MAX_MODES], *hz[ORC_MAX_MODES], *cprev[ORC_MAX_MODES];
    double *czprev[ORC_MAX_MODES], *cnzprev[ORC_MAX_MODES], *Q[ORC_MAX_MODES], *QPi[ORC_MAX_MODES];
    double *Anz[ORC_MAX_MODES], *Anzp[ORC_MAX_MODES];
    int64_t *ridx[ORC_MAX_MODES];
    uint64_t *nzr[ORC_MAX_MODES], nnzr[ORC_MAX_MODES], *cc[ORC_MAX_MODES], *bc[ORC_MAX_MODES];
    for (int m = 0; m < nmode; ++m) {
        U[m] = (double *)calloc(maxd[m] * R, sizeof(double));
        P[m] = (double *)calloc(maxd[m] * R, sizeof(double));
        Anz[m] = (double *)calloc(maxd[m] * R, sizeof(double));
        Anzp[m] = (double *)calloc(maxd[m] * R, sizeof(double));
        ridx[m] = (int64_t *)malloc(sizeof(int64_t) * maxd[m]);
        nzr[m] = (uint64_t *)malloc(sizeof(uint64_t) * maxd[m]);
        cc[m] = (uint64_t *)malloc(sizeof(uint64_t) * (nnz ? nnz : 1));
        bc[m] = (uint64_t *)malloc(sizeof(uint64_t) * (nnz ? nnz : 1));
        G[m] = (double *)calloc(RR, sizeof(double));
        c[m] = (double *)calloc(RR, sizeof(double));     /* InitSparseCPGrams: zeros */
        h[m] = (double *)calloc(RR, sizeof(double));
        cnz[m] = (double *)calloc(RR, sizeof(double));
        hnz[m] = (double *)calloc(RR, sizeof(double));
        cz[m] = (double *)calloc(RR, sizeof(double));
        hz[m] = (double *)calloc(RR, sizeof(double));
        cprev[m] = (double *)calloc(RR, sizeof(double));
        czprev[m] = (double *)calloc(RR, sizeof(double));
        cnzprev[m] = (double *)calloc(RR, sizeof(double));
        Q[m] = (double *)calloc(RR, sizeof(double));
        QPi[m] = (double *)calloc(RR, sizeof(double));
    }
    double lambda[256];
    double *Phi = (double *)malloc(sizeof(double) * RR);
    double *old_gram = (double *)malloc(sizeof(double) * RR);
    double *tmp = (double *)malloc(sizeof(double) * R);
    double *bv = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * 2 * (nnz ? nnz : 1));

    uint64_t ptr = 0, batch = 0;
    int slices = 0;
    const int sdisp = timeslice_limit > 1 ? -1 : s;  /* init_salto's stream mode, as orc_cpstream */
    while (ptr < nnz && (max_slices <= 0 || slices < max_slices)) {
        uint64_t bd[ORC_MAX_MODES], start = 0, t0 = 0;
        const uint64_t bn = next_batch(nmode, s, nnz, cidx, nnz_threshold, timeslice_limit, &ptr, &batch,
                                       prev_dim, bc, bd, &start, &t0);
        if (bn == 0) break;
        for (uint64_t i = 0; i < bn; ++i) bv[i] = vals[start + i];

        /* model creation / growth (streaming_cpd.cpp:106-128) */
        if (slices == 0) {
            for (int m = 0; m < nmode; ++m) {
                memcpy(U[m], init_factors[m], sizeof(double) * bd[m] * R);
                memset(P[m], 0, sizeof(double) * maxd[m] * R);
            }
            memset(U[s], 0, sizeof(double) * bd[s] * R);
            for (int m = 0; m < nmode; ++m) orc_gram_upper(bd[m], R, U[m], G[m]);
            memset(G[s], 0, sizeof(double) * RR);
        } else {
            for (int m = 0; m < nmode; ++m)
                if (bd[m] > cur[m]) {
                    memcpy(U[m] + cur[m] * R, init_factors[m] + cur[m] * R,
                           sizeof(double) * (bd[m] - cur[m]) * R);
                    memset(P[m] + cur[m] * R, 0, sizeof(double) * (bd[m] - cur[m]) * R);
                }
            for (int m = 0; m < nmode; ++m)
                if (m != s) orc_gram_upper(bd[m], R, U[m], G[m]);
        }
        for (int m = 0; m < nmode; ++m) cur[m] = bd[m];

        /* spcpstream_alto_iter: first slice normalisation (585-598) */
        if (slices == 0) {
            for (int m = 0; m < nmode; ++m) {
                if (m == s) continue;
                orc_kruskal_norm(bd[m], R, U[m], lambda, 0);
                orc_gram_upper(bd[m], R, U[m], G[m]);
            }
        }
        /* nonzero_slices and the row-sparse matrices (607-641) */
        for (int m = 0; m < nmode; ++m) {
            for (uint64_t i = 0; i < bd[m]; ++i) ridx[m][i] = -1;
            for (uint64_t e = 0; e < bn; ++e) ridx[m][bc[m][e]] = 0;
            uint64_t k = 0;
            for (uint64_t i = 0; i < bd[m]; ++i)
                if (ridx[m][i] == 0) { nzr[m][k] = i; ridx[m][i] = (int64_t)k; ++k; }
                else ridx[m][i] = -1;
            nnzr[m] = k;
            for (uint64_t r = 0; r < k; ++r) {
                memcpy(Anzp[m] + r * R, P[m] + nzr[m][r] * R, sizeof(double) * R);
                memcpy(Anz[m] + r * R, U[m] + nzr[m][r] * R, sizeof(double) * R);
            }
            orc_gram_upper(bd[m], R, U[m], c[m]);
            tall_atb(bd[m], R, P[m], U[m], h[m]);
        }
        if (slices > 0) {  /* 644-656 */
            for (int m = 0; m < nmode; ++m) {
                if (m == s) continue;
                tall_atb(nnzr[m], R, Anzp[m], Anzp[m], cnzprev[m]);
                for (size_t q = 0; q < RR; ++q) czprev[m][q] = cprev[m][q] - cnzprev[m][q];
            }
        }
        /* update_salto_rowsparse: compact coordinates (alto.hpp:301-434) */
        uint64_t cdims[ORC_MAX_MODES];
        for (int m = 0; m < nmode; ++m) {
            cdims[m] = nnzr[m];
            for (uint64_t e = 0; e < bn; ++e) cc[m][e] = (uint64_t)ridx[m][bc[m][e]];
        }
        int total = 0;
        for (int m = 0; m < nmode; ++m) total += orc_num_bits(cdims[m]);
        int kw = total > 64 ? 2 : 1;
        uint64_t masks[ORC_MAX_MODES * 2];
        orc_alto_masks(nmode, cdims, kw, masks);
        orc_linearize(nmode, bn, (const uint64_t *const *)cc, masks, kw, keys);
        double *sv = (double *)malloc(sizeof(double) * bn);
        memcpy(sv, bv, sizeof(double) * bn);
        orc_sort_alto(bn, kw, keys, sv, NULL);
        int nprtn = orc_get_num_ptrn(bn, g_threads);
        uint64_t *pp = (uint64_t *)malloc(sizeof(uint64_t) * (nprtn + 1));
        uint64_t *iv = (uint64_t *)malloc(sizeof(uint64_t) * nprtn * nmode * 2);
        orc_prtn_alto(nmode, cdims, bn, kw, keys, masks, nprtn, pp, iv);

        double delta = 0.0, prev_delta = 0.0;
        int outer = 0;
        for (int i = 0; i < max_iters; ++i) {
            delta = 0.0;
            /* s_t (712-742) */
            memset(Anz[s], 0, sizeof(double) * nnzr[s] * R);
            orc_mttkrp_alto(nmode, cdims, bn, kw, keys, sv, masks, nprtn, pp, iv, sdisp, Anz, R, s,
                            Anz[s]);
            hadamard_excl(nmode, R, c, s, Phi);
            for (int q = 0; q < R; ++q) Phi[q * R + q] += 1e-12;
            if (chol_upper(R, Phi) == 0)
                for (uint64_t r = 0; r < nnzr[s]; ++r) chol_solve_row(R, Phi, Anz[s] + r * R);
            /* temporal Gram (744-790) */
            for (int a = 1; a < R; ++a)
                for (int b = 0; b < a; ++b) G[s][b + a * R] = G[s][a + b * R];
            memcpy(old_gram, G[s], sizeof(double) * RR);
            for (uint64_t r = 0; r < bd[s]; ++r)
                for (int n2 = 0; n2 < R; ++n2)
                    for (int m2 = 0; m2 < R; ++m2)
                        G[s][m2 + n2 * R] += Anz[s][r * R + m2] * Anz[s][r * R + n2];
            memcpy(h[s], old_gram, sizeof(double) * RR);
            memcpy(c[s], G[s], sizeof(double) * RR);
            for (int m = 0; m < nmode; ++m) {  /* 792-970 */
                if (m == s) continue;
                hadamard_excl(nmode, R, h, m, Q[m]);
                hadamard_excl(nmode, R, c, m, Phi);
                for (int q = 0; q < R; ++q) Phi[q * R + q] += 1e-12;
                memset(Anz[m], 0, sizeof(double) * nnzr[m] * R);
                orc_mttkrp_alto(nmode, cdims, bn, kw, keys, sv, masks, nprtn, pp, iv, sdisp, Anz, R, m,
                                Anz[m]);
                for (uint64_t r = 0; r < nnzr[m]; ++r) {  /* rsp_mat_mul + rsp_mat_add */
                    for (int b = 0; b < R; ++b) {
                        double t = 0.0;
                        for (int a = 0; a < R; ++a) t += Anzp[m][r * R + a] * Q[m][a * R + b];
                        tmp[b] = t;
                    }
                    for (int b = 0; b < R; ++b) Anz[m][r * R + b] += tmp[b];
                }
                int ok = chol_upper(R, Phi) == 0;
                if (ok)
                    for (uint64_t r = 0; r < nnzr[m]; ++r) chol_solve_row(R, Phi, Anz[m] + r * R);
                tall_atb(nnzr[m], R, Anz[m], Anz[m], cnz[m]);
                tall_atb(nnzr[m], R, Anzp[m], Anz[m], hnz[m]);
                memcpy(QPi[m], Q[m], sizeof(double) * RR);
                if (ok)
                    for (int r = 0; r < R; ++r) chol_solve_row(R, Phi, QPi[m] + (size_t)r * R);
                rr_matmul(R, czprev[m], 0, QPi[m], hz[m]);
                rr_matmul(R, QPi[m], 1, hz[m], cz[m]);
                double trc = 0.0, trh = 0.0, trcp = 0.0;
                for (size_t q = 0; q < RR; ++q) {
                    c[m][q] = cnz[m][q] + cz[m][q];
                    h[m][q] = hnz[m][q] + hz[m][q];
                }
                for (int q = 0; q < R; ++q) {
                    trc += c[m][q * R + q];
                    trh += h[m][q * R + q];
                    trcp += cprev[m][q * R + q];
                }
                delta += sqrt(fabs((trc + trcp - 2.0 * trh) / (trc + 1e-12)));
            }
            outer = i + 1;
            if (i > 0 && fabs(prev_delta - delta) < eps) break;
            prev_delta = delta;
        }
        /* z rows <- prevU QPi, nz rows <- A_nz (1017-1045) */
        for (int m = 0; m < nmode; ++m) {
            for (uint64_t i = 0; i < bd[m]; ++i) {
                if (ridx[m][i] >= 0) continue;
                for (int b = 0; b < R; ++b) {
                    double t = 0.0;
                    for (int a = 0; a < R; ++a) t += P[m][i * R + a] * QPi[m][a * R + b];
                    tmp[b] = t;
                }
                memcpy(U[m] + i * R, tmp, sizeof(double) * R);
            }
            for (uint64_t r = 0; r < nnzr[m]; ++r)
                memcpy(U[m] + nzr[m][r] * R, Anz[m] + r * R, sizeof(double) * R);
        }
        for (size_t q = 0; q < RR; ++q) G[s][q] *= 0.99;  /* 1051-1058 */
        for (int m = 0; m < nmode; ++m) {
            if (m == s) continue;
            memcpy(cprev[m], c[m], sizeof(double) * RR);
            memcpy(cnzprev[m], cnz[m], sizeof(double) * RR);
        }
        free(pp); free(iv); free(sv);
        if (out_time_rows) memcpy(out_time_rows + (size_t)t0 * R, U[s], sizeof(double) * bd[s] * R);
        if (out_outer_iters) out_outer_iters[slices] = outer;
        for (int m = 0; m < nmode; ++m) memcpy(P[m], U[m], sizeof(double) * bd[m] * R);
        ++slices;
    }
    for (int m = 0; m < nmode; ++m) {
        if (out_dims) out_dims[m] = cur[m];
        if (out_factors && out_factors[m]) memcpy(out_factors[m], U[m], sizeof(double) * cur[m] * R);
        if (out_grams) memcpy(out_grams + (size_t)m * RR, G[m], sizeof(double) * RR);
    }
    if (out_ntime) *out_ntime = batch;
    for (int m = 0; m < nmode; ++m) {
        free(U[m]); free(P[m]); free(Anz[m]); free(Anzp[m]); free(ridx[m]); free(nzr[m]);
        free(cc[m]); free(bc[m]); free(G[m]); free(c[m]); free(h[m]); free(cnz[m]); free(hnz[m]);
        free(cz[m]); free(hz[m]); free(cprev[m]); free(czprev[m]); free(cnzprev[m]); free(Q[m]);
        free(QPi[m]); free(cidx[m]);
    }
    free(vals); free(Phi); free(old_gram); free(tmp); free(bv); free(keys);
    return slices;
}
