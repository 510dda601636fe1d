// C++ host interface of the MI355X ALTO library, mirroring the reference's
// entry points so that reference-style host code (src/main.cpp) ports by
// changing includes only:
//   SparseTensor / ImportSparseTensor      include/sptensor.hpp:8-59
//   KruskalModel / Create / RandomInit     include/kruskal_model.hpp:9-57
//   AltoTensor<LIT>, create_alto,          include/alto.hpp:35-111
//     destroy_alto, mttkrp_alto_par
//   init_salto, update_salto,              include/alto.hpp:66-69, 78-82
//     create_da_mem, destroy_da_mem
//   mttkrp, mttkrp_par                     include/mttkrp.hpp:9-11
//   cpd_alto<LIT>, cpd_fit_alto<LIT>       include/cpd.hpp:33-38
//   Matrix                                 include/matrix.hpp:14-20
//   cpd_constraint and its setters         include/constraints.hpp:6-32
//   cpstream, cpd_error                    include/streaming_cpd.hpp:21-31, 88
// Everything runs through the C ABI of libalto_amd.so (include/alto_hip.h);
// no function here computes on the CPU except file parsing and the glibc
// rand() initialisation the reference uses.
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <omp.h>  // omp_lock_t of the reference signatures (no OpenMP call is made)

#include <string>
#include <vector>

#include "../../include/alto_hip.h"

typedef unsigned long long IType;
typedef double FType;

// ALTO key type (include/common.hpp:64-73): 64-bit keys unless
// ALTO_MASK_LENGTH is 128 (or ALTO_MASK_128 is defined).
#if (defined(ALTO_MASK_LENGTH) && ALTO_MASK_LENGTH == 128) || defined(ALTO_MASK_128)
typedef unsigned long long RLIType;
typedef unsigned __int128 LIType;
#else
typedef unsigned long long LIType;
#endif

// Decomposition models of the command line (include/common.hpp:57).
typedef enum Model_ { ALS, CPSTREAM, CPSTREAM_ALTO, SPCPSTREAM, SPCPSTREAM_ALTO } Model;

namespace altoamd {

[[noreturn]] inline void fail(const char *what, int st) {
    fprintf(stderr, "ALTO: %s failed: %s (%s)\n", what, alto_hip_error_string(st), alto_hip_last_error());
    exit(1);
}
inline void check(int st, const char *what) {
    if (st != ALTO_OK) fail(what, st);
}

}  // namespace altoamd

// ---- COO tensor (include/sptensor.hpp:8-25) ----
struct SparseTensor {
    int nmodes;
    IType *dims;
    IType nnz;
    IType **cidx;
    FType *vals;
};
enum FileFormat { TEXT_FORMAT = 0, BINARY_FORMAT = 1 };

// sptensor.cpp:160-249: text .tns (1-based unless zero_based) or .bin.
void ImportSparseTensor(const char *file_path, FileFormat f, SparseTensor **X, IType zero_based);
void DestroySparseTensor(SparseTensor *X);

// ---- Kruskal model (include/kruskal_model.hpp:9-16) ----
struct KruskalModel {
    int mode;
    IType *dims;
    IType rank;
    FType **U;
    FType *lambda;
};
void CreateKruskalModel(int mode, IType *dims, IType rank, KruskalModel **M);
// util.cpp:28-43 / kruskal_model.cpp:163-178: +-3 * rand()/RAND_MAX per
// element, modes in order (srand is the caller's, as main.cpp:188 does).
void KruskalModelRandomInit(KruskalModel *M, unsigned int seed);
void DestroyKruskalModel(KruskalModel *M);

// ---- ALTO tensor (include/alto.hpp:35-62); LIT selects the key width ----
template <typename LIT> struct AltoTensor {
    alto_hip_tensor handle;
    int nmode;
    int nprtn;
    IType nnz;
    IType *dims;
    LIT *cr_masks;     // conflict-resolution bit per mode (zero: none)
    LIT alto_cr_mask;  // their union
};

template <typename LIT> inline int alto_key_words() { return sizeof(LIT) <= 8 ? 1 : 2; }

namespace altoamd {
// create_alto's body; cr_bits: prtn_alto's conflict-resolution bits
// (alto.hpp:1250-1346), which the streaming tensor never takes.
template <typename LIT> void make_alto(SparseTensor *spt, AltoTensor<LIT> **at, int nprtn, bool cr_bits) {
    alto_hip_coo c{spt->nmodes, (const uint64_t *)spt->dims, spt->nnz,
                   (const uint64_t *const *)spt->cidx, spt->vals, 0};
    auto *a = new AltoTensor<LIT>;
    altoamd::check(alto_hip_tensor_create(&c, nprtn, alto_key_words<LIT>(), nullptr, &a->handle),
                   "create_alto");
    a->nmode = spt->nmodes;
    a->nprtn = nprtn;
    a->nnz = spt->nnz;
    a->dims = (IType *)malloc(sizeof(IType) * a->nmode);
    for (int m = 0; m < a->nmode; ++m) a->dims[m] = spt->dims[m];
    a->cr_masks = (LIT *)calloc(a->nmode, sizeof(LIT));
    a->alto_cr_mask = 0;
    if (cr_bits) {
        const int kw = alto_key_words<LIT>();
        std::vector<uint64_t> w((size_t)a->nmode * kw);
        check(alto_hip_tensor_cr_bits(a->handle, w.data(), nullptr), "create_alto (GENERATE_CR_BITS)");
        for (int m = 0; m < a->nmode; ++m) {
            LIT v = (LIT)w[(size_t)m * kw];
            if (kw == 2) v |= (LIT)w[(size_t)m * kw + 1] << 32 << 32;
            a->cr_masks[m] = v;
            a->alto_cr_mask |= v;
        }
    }
    *at = a;
}
}  // namespace altoamd

// create_alto (alto.hpp:436-582): masks, linearisation, sort, partitions on
// the GPU; with GENERATE_CR_BITS defined, also the conflict-resolution bits.
template <typename LIT> void create_alto(SparseTensor *spt, AltoTensor<LIT> **at, int nprtn) {
#ifdef GENERATE_CR_BITS
    altoamd::make_alto(spt, at, nprtn, true);
#else
    altoamd::make_alto(spt, at, nprtn, false);
#endif
}

template <typename LIT> void destroy_alto(AltoTensor<LIT> *at) {
    alto_hip_tensor_destroy(at->handle);
    free(at->dims);
    free(at->cr_masks);
    delete at;
}

// init_salto (alto.hpp:151-193): the streaming ALTO tensor, here built from
// spt at once (update_salto rebuilds it).  stream_mode is kept for
// reference-style callers; the GPU MTTKRP has no separate stream variant.
template <typename LIT>
void init_salto(SparseTensor *spt, AltoTensor<LIT> **at, int max_nprtn, int stream_mode) {
    (void)stream_mode;
    altoamd::make_alto(spt, at, max_nprtn > 0 ? max_nprtn : 1, false);
}

// update_salto (alto.hpp:195-299): rebuilds the tensor from a new slice,
// reusing its device buffers.
template <typename LIT> void update_salto(SparseTensor *spt, AltoTensor<LIT> *at, int nprtn) {
    alto_hip_coo c{spt->nmodes, (const uint64_t *)spt->dims, spt->nnz,
                   (const uint64_t *const *)spt->cidx, spt->vals, 0};
    altoamd::check(alto_hip_tensor_update(at->handle, &c, nprtn > 0 ? nprtn : 1, nullptr), "update_salto");
    at->nmode = spt->nmodes;
    at->nprtn = nprtn;
    at->nnz = spt->nnz;
    at->dims = (IType *)realloc(at->dims, sizeof(IType) * at->nmode);
    for (int m = 0; m < at->nmode; ++m) at->dims[m] = spt->dims[m];
    free(at->cr_masks);  // the rebuild clears the conflict-resolution bits (alto.hpp:337)
    at->cr_masks = (LIT *)calloc(at->nmode, sizeof(LIT));
    at->alto_cr_mask = 0;
}

// create_da_mem / destroy_da_mem (alto.hpp:606-679): the reference's
// per-partition output copies.  The GPU MTTKRP needs none (its per-mode
// output-ordered schedules live with the tensor), so *ofibs is NULL and
// destroy does nothing; both exist so reference call sites compile.
template <typename LIT>
void create_da_mem(int target_mode, IType rank, AltoTensor<LIT> *at, FType ***ofibs) {
    (void)target_mode;
    (void)rank;
    (void)at;
    *ofibs = nullptr;
}
template <typename LIT>
void destroy_da_mem(AltoTensor<LIT> *at, FType **ofibs, IType rank, int target_mode) {
    (void)at;
    (void)ofibs;
    (void)rank;
    (void)target_mode;
}

// Device copies of a host model for repeated MTTKRPs (the GPU replacement
// of the reference's da_mem/ofibs scratch, alto.hpp:606-679).
struct DeviceFactors {
    int nmode = 0, prec = ALTO_PREC_F64;
    IType rank = 0;
    std::vector<void *> U;
    std::vector<IType> dims;
    DeviceFactors(int nmode_, const IType *dims_, IType rank_, int prec_ = ALTO_PREC_F64);
    ~DeviceFactors();
    void upload(int m, const FType *host);
    void download(int m, FType *host);
    void zero(int m);
    size_t bytes(int m) const;  // capacity of U[m]
    // U[target] += MTTKRP with the output type the kernels write: FP64 for
    // FP64 factors, FP32 for FP32 factors; BF16 factors accumulate through an
    // FP32 copy of U[target] that is rounded back (the kernels never write
    // BF16).
    void mttkrp(int target, alto_hip_tensor t);
    void *acc = nullptr;  // FP32 accumulation buffer of the BF16 path
    size_t acc_bytes = 0;
};

// mttkrp_alto_par (alto.hpp:902-908) on device factors: U[target] += MTTKRP.
template <typename LIT>
void mttkrp_alto_par(int target, DeviceFactors &F, IType rank, AltoTensor<LIT> *at) {
    (void)rank;
    F.mttkrp(target, at->handle);
}

// Host-pointer form: factors[target] += MTTKRP (copies in and out).
template <typename LIT>
void mttkrp_alto_par(int target, FType **factors, IType rank, AltoTensor<LIT> *at) {
    DeviceFactors F(at->nmode, at->dims, rank);
    for (int m = 0; m < at->nmode; ++m) F.upload(m, factors[m]);
    mttkrp_alto_par(target, F, rank, at);
    F.download(target, factors[target]);
}

// The reference's signature (alto.hpp:904): the write locks, the da_mem
// copies and the op_sched row lists are not needed on the GPU and ignored.
template <typename LIT>
void mttkrp_alto_par(int target_mode, FType **factors, IType rank, AltoTensor<LIT> *at, omp_lock_t *wlocks,
                     FType **ofibs, std::vector<std::vector<size_t>> &oidx) {
    (void)wlocks;
    (void)ofibs;
    (void)oidx;
    mttkrp_alto_par(target_mode, factors, rank, at);
}

// mttkrp.cpp:122-156 (COO MTTKRP, the reference's -c oracle) on the GPU:
// M->U[mode] += MTTKRP.
void mttkrp(SparseTensor *X, KruskalModel *M, IType mode);
inline void mttkrp_par(SparseTensor *X, KruskalModel *M, IType mode, void * /*writelocks*/) {
    mttkrp(X, M, mode);
}

// RowSparseMatrix (include/rowsparse_matrix.hpp:21-27): nnzr stored rows of
// an I x J matrix, row-major in `vals`, with their row indices in `rowind`.
struct RowSparseMatrix {
    IType I = 0, J = 0, nnzr = 0;
    size_t *rowind = nullptr;
    FType *vals = nullptr;
};
RowSparseMatrix *rspmat_init(IType nrows, IType ncols, IType nnzr);
void rspmat_free(RowSparseMatrix *mat);

// rowsparse_mttkrp, the streaming overload (include/rowsparse_mttkrp.hpp:36-41,
// src/rowsparse_mttkrp.cpp:97-187): the factors are row-sparse (rsp_mats[m],
// ridx[m][i] = stored row of index i or -1); the stream mode uses its stored
// row 0.  Returns a new row-sparse matrix: the distinct `mode` indices of X
// (ascending) and their MTTKRP rows.
RowSparseMatrix *rowsparse_mttkrp(SparseTensor *X, RowSparseMatrix **rsp_mats, IType mode,
                                  IType streaming_mode, std::vector<std::vector<int>> &ridx);

// ---- Matrix (include/matrix.hpp:14-20): a row-major view ----
struct Matrix {
    IType I;
    IType J;
    FType *vals;
    int rowmajor;
};

// cpd_alto (cpd.hpp:47-191): CP-ALS on the GPU, printing the reference's lines.
template <typename LIT>
void cpd_alto(AltoTensor<LIT> *AT, KruskalModel *M, int max_iters, double epsilon,
              int prec = ALTO_PREC_F64, int fit_mode = ALTO_FIT_REFERENCE_DIAG) {
    alto_hip_als_opts o{(int)M->rank, max_iters, epsilon, prec, fit_mode, 1};
    int iters = 0;
    altoamd::check(alto_hip_cpd_als(AT->handle, &o, M->U, M->lambda, nullptr, nullptr, &iters, nullptr,
                                    nullptr),
                   "cpd_alto");
}

// cpd_fit_alto (cpd.hpp:193-275) on the GPU: the reference's relative
// residual from the last mode's factor and MTTKRP, lambda and the Grams
// (their upper triangles; the lower ones are read as zeros, as the
// reference's never-written lower triangles).
template <typename LIT>
double cpd_fit_alto(AltoTensor<LIT> *AT, KruskalModel *M, Matrix **grams, FType *U_mttkrp, FType normAT) {
    const int N = AT->nmode;
    const int R = (int)M->rank;
    std::vector<double> g((size_t)N * R * R);
    for (int m = 0; m < N; ++m)
        for (int e = 0; e < R * R; ++e) g[(size_t)m * R * R + e] = grams[m]->vals[e];
    double fit = 0.0;
    altoamd::check(alto_hip_cpd_fit(N, R, AT->dims[N - 1], M->U[N - 1], U_mttkrp, M->lambda, g.data(), normAT,
                                    ALTO_FIT_REFERENCE_DIAG, &fit, nullptr),
                   "cpd_fit_alto");
    return fit;
}

// ---- cpd_constraint (include/constraints.hpp:6-32, src/constraints.cpp) ----
// The GPU ADMM applies the non-negative projection itself, so func stays
// NULL; LASSO projects like NON_NEGATIVE, as the reference's setter makes it
// (constraints.cpp:63-68).
typedef enum { UNCONSTRAINED, NON_NEGATIVE, LASSO, MAX_COL_NORM } con_solve_type;
typedef struct {
    con_solve_type solve_type;
    char description[256];
    void (*func)(FType *primal, IType nrows, IType ncols);
} cpd_constraint;
cpd_constraint *init_constraint();
void free_constraint(cpd_constraint *con);
void apply_nonneg_constraint(cpd_constraint *con);
void apply_lasso_constraint(cpd_constraint *con);

// cpstream (streaming_cpd.cpp:26-366), the reference's signature
// (streaming_cpd.hpp:21-31): CPSTREAM_ALTO, or SPCPSTREAM_ALTO when
// use_spcpstream (spcpstream_alto_iter, 466-1085; the constraint is then
// unused, as in the reference).  Time slices are batched as
// next_dynamic_batch (streaming_sptensor.cpp:196-264): nnz_threshold (the
// reference's default ULLONG_MAX) and timeslice_limit (default 1).  Rows
// are initialised with the reference's rand_val in mode order over the
// final dims.  Like the reference, it destroys X (streaming_cpd.cpp:348).
void cpstream(SparseTensor *X, int rank, int max_iters, int streaming_mode, FType epsilon, IType seed,
              cpd_constraint *con, bool use_spcpstream, IType nnz_threshold, IType timeslice_limit);
// The same without destroying X, with the MTTKRP precision and a batch cap
// (max_slices batches, 0 = all).  Returns the number of batches.
int cpstream_amd(SparseTensor *X, int rank, int max_iters, int streaming_mode, FType epsilon, bool nonneg,
                 bool use_spcpstream, IType nnz_threshold, IType timeslice_limit, int max_slices = 0,
                 int prec = ALTO_PREC_F64);

// cpd_error (streaming_cpd.cpp:368-464) on the GPU: sets M's weights to 1,
// as the reference does, and returns sqrt(||X - M||^2) / ||X|| (the
// reference prints its square as the final fit error).
double cpd_error(SparseTensor *X, KruskalModel *M);
