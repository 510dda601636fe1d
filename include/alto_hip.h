/*
 * alto_hip.h -- C ABI of the MI355X (gfx950) ALTO sparse CP-decomposition
 * library, libalto_amd.so.
 *
 * Plain pointers, sizes and opaque handles only.  Every call returns an int
 * status (ALTO_OK = 0, negative on error) instead of asserting or exiting as
 * the reference does (alto.hpp:995, matrix.cpp:347, cpd.hpp:471).
 * `stream` arguments are hipStream_t values passed as void* (NULL = the
 * library's default stream).
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the jeewhanchoi/ALTO-stream repository root).
 */
#ifndef ALTO_HIP_H_
#define ALTO_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALTO_MAX_MODES 8

/* Status codes. */
#define ALTO_OK 0
#define ALTO_ERR_ARG -1          /* bad argument (null pointer, bad mode, rank, ...) */
#define ALTO_ERR_KEY_WIDTH -2    /* coordinates need more bits than the key holds */
#define ALTO_ERR_HIP -3          /* HIP runtime error (details: alto_hip_last_error) */
#define ALTO_ERR_NOMEM -4        /* device allocation failed */
#define ALTO_ERR_NO_DEVICE -5    /* no usable GPU */
#define ALTO_ERR_UNSUPPORTED -6  /* configuration not built (e.g. rank > 128) */
#define ALTO_ERR_NUMERIC -7      /* Cholesky failure where the reference exits */
#define ALTO_ERR_EMPTY_SLICE -8  /* CP-stream time slice with no nonzeros */

/* Arithmetic of the MTTKRP / factor storage. */
#define ALTO_PREC_F64 0   /* FP64 values, factors and output (plain path) */
#define ALTO_PREC_F32 1   /* FP32 values and factors, FP32 products */
#define ALTO_PREC_BF16 2  /* FP32 values, BF16-stored factors widened to FP32 */

/* Fit convention of cpd_fit_alto (cpd.hpp:193-275).  The reference reads the
 * Gram lower triangles, which it never writes (gram.cpp:29-42); with that
 * memory zero its ||M||^2 keeps only the diagonal terms. */
#define ALTO_FIT_EXACT 0           /* ||M||^2 from the symmetric Grams */
#define ALTO_FIT_REFERENCE_DIAG 1  /* ||M||^2 from the Gram lower triangles as
                                      the reference reads them (zero off-diagonal) */

typedef struct alto_hip_tensor_s *alto_hip_tensor;

/* A COO sparse tensor; mirrors SparseTensor (include/sptensor.hpp:8-25):
 * per-mode index arrays cidx[m][i] (0-based, uint64) and FP64 values. */
typedef struct {
    int nmode;
    const uint64_t *dims;            /* host array, nmode entries */
    uint64_t nnz;
    const uint64_t *const *cidx;     /* host array of nmode pointers */
    const double *vals;
    int on_device;                   /* 1: cidx[m] and vals are device pointers */
} alto_hip_coo;

/* ---- runtime ---- */
const char *alto_hip_error_string(int status);
const char *alto_hip_last_error(void);
int alto_hip_init(int device);
int alto_hip_device_name(char *buf, size_t len);
int alto_hip_malloc(void **ptr, size_t bytes);
int alto_hip_free(void *ptr);
int alto_hip_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int alto_hip_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int alto_hip_memset(void *ptr, int value, size_t bytes, void *stream);
int alto_hip_stream_create(void **stream);
int alto_hip_stream_destroy(void *stream);
int alto_hip_stream_sync(void *stream);
/* Converts n elements between precisions (F64 <-> F32 <-> BF16 bits). */
int alto_hip_convert(void *dst, int dst_prec, const void *src, int src_prec,
                     size_t n, void *stream);
/* Event timing on `stream`: start records, stop records and synchronises,
 * returning the elapsed milliseconds. */
int alto_hip_timer_start(void **timer, void *stream);
int alto_hip_timer_stop(void *timer, void *stream, float *ms);

/* ---- ALTO tensor: create_alto / destroy_alto (include/alto.hpp:436-604) ----
 * Masks (setup_packed_alto, LSB_FIRST + SHORT_FIRST, alto.hpp:954-1057),
 * linearisation (alto.hpp:515-531), sort (alto.hpp:1059-1090) and the
 * nprtn line partitions with per-mode intervals (prtn_alto,
 * alto.hpp:1128-1226) are computed on the GPU.  key_words: 0 = smallest
 * that fits (64-bit when the coordinates need <= 64 bits), 1 = 64-bit
 * (ALTO_MASK_LENGTH 64), 2 = 128-bit. */
int alto_hip_tensor_create(const alto_hip_coo *coo, int nprtn, int key_words,
                           void *stream, alto_hip_tensor *out);
int alto_hip_tensor_destroy(alto_hip_tensor t);
/* update_salto (include/alto.hpp:195-299): rebuilds t in place from a new
 * COO tensor (new dims and nonzeros, same key-width rule as create with
 * key_words 0), reusing its device buffers; the per-mode schedules are
 * rebuilt on the next MTTKRP. */
int alto_hip_tensor_update(alto_hip_tensor t, const alto_hip_coo *coo, int nprtn, void *stream);
int alto_hip_tensor_info(alto_hip_tensor t, int *nmode, uint64_t *nnz,
                         int *key_words, int *nprtn, int *total_bits);
int alto_hip_tensor_dims(alto_hip_tensor t, uint64_t *dims);
/* Getters for the bit-exact checks: masks (nmode*key_words words), sorted
 * keys (nnz*key_words words, word 0 = low 64 bits), sorted FP64 values,
 * the sort permutation (sorted position -> input index), partitions. */
int alto_hip_tensor_masks(alto_hip_tensor t, uint64_t *masks);
int alto_hip_tensor_keys(alto_hip_tensor t, uint64_t *keys);
int alto_hip_tensor_vals(alto_hip_tensor t, double *vals);
/* The sorted keys in the ALT_PEXT layout (alto.hpp:260-292 with mode_pos,
 * alto.hpp:982-988): mode m's index in bits [pos[m], pos[m] + bits_m),
 * pos = prefix sum of the modes' bit counts in mode order; key_words words
 * per key (word 0 = low 64 bits).  The MTTKRP kernels read this layout.
 * pos: nmode ints (may be NULL). */
int alto_hip_tensor_packed_keys(alto_hip_tensor t, uint64_t *keys, int *pos);
int alto_hip_tensor_perm(alto_hip_tensor t, uint32_t *perm);
/* GENERATE_CR_BITS (include/alto.hpp:1250-1346, a build option of
 * create_alto, off by default at alto.hpp:127): a mode with more than 512
 * rows and at most 4 nonzeros per row (integer nnz / dims) takes the next
 * free key bit, from bit total_bits upwards in mode order while the key
 * (64 * key_words bits) has one free; every nonzero whose row of that mode
 * is touched by more than one partition gets the bit, in the keys of both
 * layouts (alto_hip_tensor_keys, alto_hip_tensor_packed_keys).  cr_masks
 * (nmode * key_words words, may be NULL) receives each mode's bit (zero:
 * none).  The MTTKRP ignores the bits (its schedules need no atomics per
 * nonzero); alto_hip_tensor_update clears them, as update_salto does.
 * ALTO_ERR_UNSUPPORTED when a bit would lie beyond the packed keys (forced
 * two-word keys over at most 64 used bits). */
int alto_hip_tensor_cr_bits(alto_hip_tensor t, uint64_t *cr_masks, void *stream);
int alto_hip_tensor_partitions(alto_hip_tensor t, uint64_t *prtn_ptr,
                               uint64_t *intervals);
/* Device pointers of the sorted keys / FP64 values (for device-side users). */
int alto_hip_tensor_device_arrays(alto_hip_tensor t, const uint64_t **keys,
                                  const double **vals);

/* ---- MTTKRP: mttkrp_alto_par (include/alto.hpp:902-908) ----
 * out[target] += MTTKRP.  factors: host array of nmode DEVICE pointers,
 * row-major I_m x rank, element type by `prec` (double / float / bf16 bits).
 * out: device, dims[target] x rank, double when out_f64 != 0 (or prec is
 * F64), else float -- never BF16, whatever the factor storage.  out_bytes:
 * the capacity of out; below dims[target] * rank * 8 (FP64 output) or * 4
 * (FP32 output) the call fails with ALTO_ERR_ARG and writes nothing.  The
 * output is accumulated, as the reference does, so callers zero it first
 * (cpd.hpp:110). */
int alto_hip_mttkrp(alto_hip_tensor t, int target, int prec, int rank,
                    const void *const *factors, void *out, int out_f64, size_t out_bytes,
                    void *stream);
/* Same, selecting the kernel explicitly (for tests / the autotuner):
 * algo 0 = automatic, 1 = ALTO-order segments (atomic flush),
 * 2 = output-ordered segments (per-mode row order). */
int alto_hip_mttkrp_algo(alto_hip_tensor t, int target, int prec, int rank,
                         const void *const *factors, void *out, int out_f64, size_t out_bytes,
                         int algo, void *stream);
/* Same, with the kernel's tuning knobs, which the defaults of the calls
 * above pick (autotuned once per shape bucket for >= 2^20 schedule entries,
 * unless ALTO_AUTOTUNE=0 is in the environment):
 * segment = entries per lane group (0 = from the entry count, 4..64);
 * flags bit 0 = non-temporal key/value stream loads, bit 1 = reuse the
 * gathered product across equal consecutive keys, bit 3 = no decode-once
 * kernel, bit 4 = no fiber kernel, bit 5 = the output holds zeros on entry
 * (algo 2: rows owned by one range are stored instead of added; the caller
 * guarantees the zeros), bit 6 = fiber kernel also for BF16 factors,
 * bit 7 = the hot-row kernel (algo 2; one workgroup per CU with the most
 * popular factor rows of the gathered modes in LDS; ranks 16, 32, 64, 3 to
 * 5 modes, 16-byte aligned output; its workgroup shape in bits 9-10:
 * 0 = 1024 threads, 1 = 768, 2 = 512), bit 8 = the tuner leaves the hot-row
 * kernel out (-1 = defaults: bits 0 and 1).
 * No reference counterpart (the reference fixes its launch shape). */
int alto_hip_mttkrp_config(alto_hip_tensor t, int target, int prec, int rank,
                           const void *const *factors, void *out, int out_f64, size_t out_bytes,
                           int algo, uint64_t segment, int flags, void *stream);
/* Reference COO MTTKRP, mttkrp (src/mttkrp.cpp:122-156), on the GPU:
 * out += MTTKRP over a device COO tensor, FP64. */
int alto_hip_mttkrp_coo(const alto_hip_coo *coo, int target, int rank,
                        const double *const *factors, double *out, void *stream);

/* ---- CP-ALS: cpd_alto (include/cpd.hpp:47-191) with pseudo_inverse
 * (cpd.hpp:377-480), update_gram (src/gram.cpp:29-42), KruskalModelNorm
 * (src/kruskal_model.cpp:319-353) and cpd_fit_alto (cpd.hpp:193-275). */
typedef struct {
    int rank;
    int max_iters;
    double epsilon;       /* stop when |fit - prev_fit| < epsilon, it > 0 */
    int prec;             /* ALTO_PREC_F64 / F32 / BF16 */
    int fit_mode;         /* ALTO_FIT_REFERENCE_DIAG (default printout) or EXACT */
    int verbose;          /* 1: print the reference's per-iteration lines */
} alto_hip_als_opts;

/* factors: host, nmode pointers to I_m x rank FP64 (in: initial model, out:
 * final model); lambda: host rank doubles (out).  fits / fits_exact: host
 * arrays of max_iters (may be NULL).  timings (may be NULL): 8 doubles,
 * seconds: total, MTTKRP, pseudo-inverse, copy, normalise, Gram, fit, and
 * per-iteration total; the per-phase entries are measured only with
 * verbose = 1 (the reference's printout), otherwise they are 0 and the
 * iterations after the first replay one captured hipGraph (with
 * epsilon <= 0 every iteration runs and the fits are read once at the end).
 * Returns the iteration count in *iters. */
int alto_hip_cpd_als(alto_hip_tensor t, const alto_hip_als_opts *opts,
                     double *const *factors, double *lambda, double *fits,
                     double *fits_exact, int *iters, double *timings,
                     void *stream);

/* ---- cpd_fit_alto (include/cpd.hpp:193-275), on the GPU ----
 * fit = sqrt(normX + ||M||^2 - 2 <X, M>) / sqrt(normX) (the unsquared value
 * when the residual is negative, as cpd.hpp:264-268), with
 * <X, M> = sum_r lambda_r sum_i U_last[i, r] * mttkrp_last[i, r] and ||M||^2
 * from lambda and the nmode Gram matrices (grams: nmode consecutive R x R
 * row-major; their upper triangles are read, as update_gram writes them).
 * fit_mode ALTO_FIT_REFERENCE_DIAG keeps the diagonal terms of ||M||^2 (the
 * reference's reading of never-written lower triangles as zeros),
 * ALTO_FIT_EXACT all of them.  Host arrays: U_last and mttkrp_last rows x
 * rank, lambda rank. */
int alto_hip_cpd_fit(int nmode, int rank, uint64_t rows, const double *U_last,
                     const double *mttkrp_last, const double *lambda, const double *grams,
                     double normX, int fit_mode, double *fit, void *stream);

/* ---- batched CP-ALS: cpd_alto (cpd.hpp:47-191) on many independent tensors ----
 * Runs alto_hip_cpd_als on ntensors distinct tensors concurrently: nstreams
 * (1..64) worker threads, each with its own HIP stream and a device
 * workspace reused from tensor to tensor, take the tensors in order of
 * decreasing nnz * max_iters.  The output-ordered schedules of every tensor
 * are built first.  Per tensor k: opts[k] (verbose must be 0), factors[k][m]
 * (host FP64 dims x rank: initial model in, final model out), lambdas[k]
 * (rank), fits[k] (max_iters) and iters[k]; each may be NULL.  Each tensor's
 * results are those of alto_hip_cpd_als on it alone.  timings (may be NULL,
 * 2 doubles): wall seconds of the CP-ALS runs, then of the schedule builds
 * that precede them.  Returns the first failing status. */
int alto_hip_cpd_als_batch(int ntensors, const alto_hip_tensor *tensors,
                           const alto_hip_als_opts *opts, double *const *const *factors,
                           double *const *lambdas, double *const *fits, int *iters,
                           int nstreams, double *timings);
/* The same with device-resident models: factors[k][m] are device FP64
 * arrays (dims x rank; initial model in, final model out), converted into
 * and out of the working precision on the device, so no model crosses
 * PCIe (lambdas, fits and iters stay host arrays).  A tensor outside the
 * grouped classes (see alto_hip_cpd_als_batch) takes a host round trip.
 * A factor array the GPU cannot read (pageable host memory) is
 * ALTO_ERR_ARG.  No reference counterpart (the batched workload has none);
 * for callers whose models already live in HBM. */
int alto_hip_cpd_als_batch_dev(int ntensors, const alto_hip_tensor *tensors,
                               const alto_hip_als_opts *opts, double *const *const *factors,
                               double *const *lambdas, double *const *fits, int *iters,
                               int nstreams, double *timings);

/* ---- pseudo_inverse (include/cpd.hpp:377-480): U <- U V^{-1} row by row with
 * V = Hadamard of grams[m], m != mode (nmode consecutive rank x rank
 * row-major matrices; the row-major upper triangle is read, as DPOTRF('L')
 * does) plus 1e-3 on the diagonal; when V is not positive definite, the
 * minimum-norm solve of the unregularised diagonal and *fallback = 1.
 * Host arrays: grams, U (rows x rank, in place). */
int alto_hip_pseudo_inverse(int nmode, int rank, const double *grams, int mode, uint64_t rows,
                            double *U, int *fallback, void *stream);

/* ---- model error: cpd_error (src/streaming_cpd.cpp:368-464) with
 * kruskal_norm (src/kruskal_model.cpp:400-442) ----
 * err = sqrt(||X||^2 + ||M||^2 - 2 <X, M>) / ||X|| for the Kruskal model of
 * the given factors with all weights 1 (cpd_error sets them to 1):
 * <X, M> from the COO MTTKRP of the mode with the fewest rows (dims[m] rows
 * per factor, rank columns, FP64), ||X||^2 = sum of squared stored values.
 * The reference prints err^2 as "final fit error".  parts (may be NULL):
 * ||X||^2, ||M||^2, <X, M>.  Arrays (coordinates, values and factors) host
 * or device as coo->on_device says; with on_device, an array the GPU cannot
 * read (pageable host memory) is ALTO_ERR_ARG, as is a coordinate >= dims[m]. */
int alto_hip_cpd_error(const alto_hip_coo *coo, int rank, const double *const *factors,
                       double *err, double *parts, void *stream);

/* ---- alto_hip_cpd_error without allocation ----
 * No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  The quantities of
 * alto_hip_cpd_error, enqueued on `stream` with no device allocation, no
 * host synchronisation and no host read: out (DEVICE, 4 doubles) receives
 * err, ||X||^2, ||M||^2 and <X, M> when the stream reaches the call.  Device
 * inputs only (coo->on_device = 1, factors[m] device FP64, dims[m] x rank).
 * rank <= 64, nmode <= 8 (its scratch, a static device array of about
 * 8.4 MB per GPU, is sized for that; larger ranks are ALTO_ERR_UNSUPPORTED).
 * A coordinate not below its mode's size sets all four outputs to NaN
 * (alto_hip_cpd_error returns ALTO_ERR_ARG).  Calls on one GPU share that
 * scratch: calls that overlap on one GPU (other streams or threads) race. */
int alto_hip_cpd_error_noalloc(const alto_hip_coo *coo, int rank, const double *const *factors,
                               double *out, void *stream);

/* ---- CP-stream: cpstream + cpstream_alto_iter (src/streaming_cpd.cpp:26-366,
 * 1198-1519) with admm_cf_ (src/admm.cpp:216-429), CPSTREAM_ALTO model.
 * Time slices are grouped into batches as next_dynamic_batch does
 * (src/streaming_sptensor.cpp:196-264): consecutive time indices, no
 * further one once nnz_threshold nonzeros are taken (0 = no threshold, the
 * reference's ULLONG_MAX default), at most timeslice_limit of them (0 or 1 =
 * one per batch, the reference's default); a batch's stream mode has one row
 * per time index.  init_factors: host, nmode arrays of dims[m] x rank rows
 * used when the model grows (row order = first-appearance labels).
 * Outputs (host, may be NULL): final factors (dims[m] x rank each; the
 * stream mode holds the last batch's rows), final dims, the time rows (one
 * per time index), outer iterations per batch, ADMM iterations per batch,
 * final Grams (nmode*rank*rank), and the final error of cpstream's closing
 * cpd_error call (streaming_cpd.cpp:307-345; see alto_hip_cpd_error) over the
 * processed nonzeros in the relabeled index space, the time factor being the
 * time rows.  timings (may be NULL): 4 doubles: slice-loop seconds,
 * slice-build seconds, MTTKRP seconds, time indices processed.  max_slices >
 * 0 caps the number of batches.  prec: ALTO_PREC_F64 or ALTO_PREC_F32
 * (MTTKRP arithmetic; the solve, ADMM and Grams are FP64 either way).
 * rank <= 64.  ALTO_ADMM_CACHE=0 in the environment makes the non-negative
 * ADMM stream its row state from memory instead of keeping it on chip (same
 * results).  Returns the number of batches processed, or a negative
 * status. */
int alto_hip_cpstream(const alto_hip_coo *coo, int stream_mode, int rank,
                      int max_iters, double epsilon, int nonneg, int max_slices,
                      uint64_t nnz_threshold, uint64_t timeslice_limit,
                      int prec, const double *const *init_factors, double *const *out_factors,
                      uint64_t *out_dims, double *out_time_rows,
                      int *out_outer_iters, double *out_admm_iters,
                      double *out_grams, double *out_final_error, double *timings,
                      void *stream);

/* ---- StreamingSparseTensor (src/streaming_sptensor.cpp:69-264 with tt_sort,
 * src/sort.cpp:477-519; next_dynamic_batch at timeslice limit 1), on the
 * device: the nonzeros stably sorted by (s, s+1, ...) mod N (ties keep input
 * order), every non-stream index relabelled by first appearance in that
 * order, one slice per time index.  out_cidx[m] / out_vals (nnz each; host
 * or device as coo->on_device; entries may be NULL) receive the prepared
 * stream.  *out_nslices = number of slices; when it is <= max_slices,
 * out_slice_ptr (nslices + 1) receives the slice boundaries and
 * out_slice_dims (nslices x nmode, row-major) the model dims after each
 * slice (running max of label + 1; 1 for the stream mode).  nnz < 2^32;
 * coordinates >= dims are ALTO_ERR_ARG; a time index without nonzeros below
 * the largest is ALTO_ERR_EMPTY_SLICE (the reference asserts).  CP-stream and
 * spCP-stream prepare their input with it. */
int alto_hip_stream_prepare(const alto_hip_coo *coo, int stream_mode, uint64_t *const *out_cidx,
                            double *out_vals, uint64_t max_slices, uint64_t *out_nslices,
                            uint64_t *out_slice_ptr, uint64_t *out_slice_dims, void *stream);

/* ---- spCP-stream: cpstream with use_spcpstream (model SPCPSTREAM_ALTO,
 * src/streaming_cpd.cpp:26-366, main.cpp:163, 397) driving
 * spcpstream_alto_iter (streaming_cpd.cpp:466-1085), batches as
 * alto_hip_cpstream.  Per batch only the rows that occur in it are updated
 * inside the outer loop (row-sparse MTTKRP over compact coordinates,
 * update_salto_rowsparse, alto.hpp:301-434); the other rows take
 * prevU * (Q Phi^-1) at the end.  Constraints do not apply (the reference's
 * spCP-stream solves unconstrained).  Arguments and outputs as
 * alto_hip_cpstream, without the ADMM counts; out_grams holds the
 * reference's grams array (non-stream modes: U^T U at the start of the last
 * batch; stream mode: G_s).  max_iters >= 1, rank <= 64.  Returns the number
 * of batches processed, or a negative status. */
int alto_hip_spcpstream(const alto_hip_coo *coo, int stream_mode, int rank,
                        int max_iters, double epsilon, int max_slices,
                        uint64_t nnz_threshold, uint64_t timeslice_limit, int prec,
                        const double *const *init_factors, double *const *out_factors,
                        uint64_t *out_dims, double *out_time_rows, int *out_outer_iters,
                        double *out_grams, double *out_final_error, double *timings,
                        void *stream);

/* ---- row-sparse MTTKRP: rowsparse_mttkrp (include/rowsparse_mttkrp.hpp:18-41,
 * src/rowsparse_mttkrp.cpp:97-187) ----
 * Factor matrices are row-sparse, as the reference's RowSparseMatrix
 * (include/rowsparse_matrix.hpp:21-27): mode m stores rsp_rows[m] rows
 * (rsp_vals[m], row-major rsp_rows[m] x rank, FP64) and ridx[m][i] maps index
 * i of mode m (i < dims[m]) to its stored row; -1 = no stored row (the
 * nonzero then contributes zero; the reference would read out of bounds).
 * stream_mode (or -1): that mode's factor is its single stored row, used for
 * every nonzero (ridx[stream_mode] is ignored), as the reference's
 * m_idx = 0.  Entries rsp_vals[mode] / ridx[mode] are ignored.
 * Output, row-sparse: the distinct indices of `mode` that occur, ascending
 * (out_rowind, capacity dims[mode]), their MTTKRP rows (out_vals, capacity
 * dims[mode] x rank, overwritten) and *out_nnzr.  All arrays are host or
 * device memory as coo->on_device says. */
int alto_hip_rowsparse_mttkrp(const alto_hip_coo *coo, int mode, int stream_mode, int rank,
                              const double *const *rsp_vals, const uint64_t *rsp_rows,
                              const int32_t *const *ridx, uint64_t *out_rowind, double *out_vals,
                              uint64_t *out_nnzr, void *stream);

/* ---- synthetic workloads (no reference counterpart: benchmark input) ----
 * Fills device arrays cidx_dev[m] (nnz uint64 each) and vals_dev (nnz FP64).
 * mode_dist[m]: 0 = uniform over the mode, 1 = Zipf(zipf_s[m]) over a seeded
 * permutation of the mode.  value_dist: 0 = U(0,1), 1 = lognormal(p1, p2),
 * 2 = N(0,1), 3 = Poisson(p1) + 1.  A fraction block_frac of the nonzeros is
 * drawn uniformly inside a seeded block^nmode sub-block.  Deterministic in
 * (seed, nonzero index). */
int alto_hip_generate(int nmode, const uint64_t *dims, uint64_t nnz, uint64_t seed,
                      const int *mode_dist, const double *zipf_s, int value_dist,
                      double p1, double p2, double block_frac, uint64_t block,
                      uint64_t *const *cidx_dev, double *vals_dev, void *stream);

/* ---- gather-rate probe (no reference counterpart: roofline calibration) ----
 * Average ms of one launch gathering `gathers` 128-byte rows of a device
 * table of `rows` rows, the access pattern of the MTTKRP's factor-row
 * gathers: row indices from a coalesced 4-byte stream, uniform (dist 0),
 * Zipf(zipf_s) over a seeded permutation (dist 1) or Zipf(zipf_s) with the
 * popular rows first (dist 2: a table relabelled by popularity); layout 0 = 8 lanes x
 * 16 B per row, layout 1 = 4 lanes x 2 x 16 B (the fiber kernel's FP32
 * rank-32 row). */
int alto_hip_gather_probe(uint64_t rows, uint64_t gathers, int dist, double zipf_s, int layout,
                          uint64_t seed, float *ms, void *stream);

#ifdef __cplusplus
}
#endif
#endif
