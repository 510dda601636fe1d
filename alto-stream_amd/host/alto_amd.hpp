// C++ host interface of the MI355X ALTO library, mirroring the reference's
// entry points so that reference-style host code (src/main.cpp) ports by
// changing includes only:
//   SparseTensor / ImportSparseTensor      include/sptensor.hpp:8-59
//   KruskalModel / Create / RandomInit     include/kruskal_model.hpp:9-57
//   AltoTensor<LIT>, create_alto,          include/alto.hpp:35-111
//     destroy_alto, mttkrp_alto_par
//   mttkrp, mttkrp_par                     include/mttkrp.hpp:9-11
//   cpd_alto<LIT>                          include/cpd.hpp:33-45
//   cpstream                               include/streaming_cpd.hpp:21-33
// Everything runs through the C ABI of libalto_amd.so (include/alto_hip.h);
// no function here computes on the CPU except file parsing and the glibc
// rand() initialisation the reference uses.
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "../../include/alto_hip.h"

typedef unsigned long long IType;
typedef double FType;

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
};

template <typename LIT> inline int alto_key_words() { return sizeof(LIT) <= 8 ? 1 : 2; }

// create_alto (alto.hpp:436-582): masks, linearisation, sort, partitions on the GPU.
template <typename LIT> void create_alto(SparseTensor *spt, AltoTensor<LIT> **at, int nprtn) {
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
    *at = a;
}

template <typename LIT> void destroy_alto(AltoTensor<LIT> *at) {
    alto_hip_tensor_destroy(at->handle);
    free(at->dims);
    delete at;
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
};

// mttkrp_alto_par (alto.hpp:902-908) on device factors: U[target] += MTTKRP.
template <typename LIT>
void mttkrp_alto_par(int target, DeviceFactors &F, IType rank, AltoTensor<LIT> *at) {
    altoamd::check(alto_hip_mttkrp(at->handle, target, F.prec, (int)rank, (const void *const *)F.U.data(),
                                   F.U[target], 1, nullptr),
                   "mttkrp_alto_par");
}

// Host-pointer form: factors[target] += MTTKRP (copies in and out).
template <typename LIT>
void mttkrp_alto_par(int target, FType **factors, IType rank, AltoTensor<LIT> *at) {
    DeviceFactors F(at->nmode, at->dims, rank);
    for (int m = 0; m < at->nmode; ++m) F.upload(m, factors[m]);
    mttkrp_alto_par(target, F, rank, at);
    F.download(target, factors[target]);
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

// cpstream (streaming_cpd.cpp:26-366), timeslice limit 1: CPSTREAM_ALTO, or
// SPCPSTREAM_ALTO when use_spcpstream (spcpstream_alto_iter, 466-1085; the
// constraint is then unused, as in the reference).  Rows are initialised
// with the reference's rand_val in mode order over the final dims.  Returns
// the number of slices.
int cpstream(SparseTensor *X, int rank, int max_iters, int streaming_mode, FType epsilon, IType seed,
             bool nonneg, int max_slices = 0, int prec = ALTO_PREC_F64, bool use_spcpstream = false);
