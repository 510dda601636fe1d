// Host-side implementation of alto_amd.hpp: tensor I/O, the Kruskal model
// container and thin wrappers over the C ABI.
#include "alto_amd.hpp"

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

using altoamd::check;

// ---- sptensor.cpp:74-242 ----
// Text import (read_tns_dims + read_tns_data, sptensor.cpp:74-158) in
// parallel: the file is read once into memory and split at line boundaries
// into one chunk per thread; a first pass counts each chunk's nonzero lines
// and index maxima, a second parses them into place.  Line rules as the
// reference's getline loop: a line counts when it has more than one
// character (newline included) and does not start with '#'; the mode count
// is the number of space/tab-separated tokens of the first such line minus
// one; each line is parsed as a NUL-terminated string (strtoull per mode,
// then strtod), so a short line reads zeros, as in the reference.
namespace {

struct Chunk {
    size_t begin, end;  // [begin, end) of the buffer, starting at a line start
    IType nnz = 0;
    IType maxidx[ALTO_MAX_MODES] = {0};
};

int io_threads() {
    unsigned hw = std::thread::hardware_concurrency();
    if (const char *e = getenv("OMP_NUM_THREADS")) {
        int v = atoi(e);
        if (v > 0) hw = (unsigned)v;
    }
    if (hw < 1) hw = 1;
    return (int)std::min(hw, 64u);
}

// Calls f(line, length incl. newline) for every line of [begin, end);
// lines are NUL-terminated in place (the newline becomes '\0').
template <typename F>
void for_lines(char *buf, size_t begin, size_t end, F f) {
    size_t a = begin;
    while (a < end) {
        size_t b = a;
        while (b < end && buf[b] != '\n') ++b;
        const size_t len = b < end ? b - a + 1 : b - a;  // getline's count
        if (b < end) buf[b] = '\0';
        f(buf + a, len);
        a = b + 1;
    }
}

}  // namespace

static void read_text(FILE *fp, SparseTensor *X, IType zero_based) {
    if (fseek(fp, 0, SEEK_END) != 0) {
        fprintf(stderr, "ALTO: cannot seek the tensor file\n");
        exit(1);
    }
    const long fsz = ftell(fp);
    rewind(fp);
    std::vector<char> buf((size_t)(fsz > 0 ? fsz : 0) + 1, '\0');
    if (fsz > 0 && fread(buf.data(), 1, (size_t)fsz, fp) != (size_t)fsz) {
        fprintf(stderr, "ALTO: short read of the tensor file\n");
        exit(1);
    }
    const size_t n = (size_t)(fsz > 0 ? fsz : 0);
    // mode count from the first counted line (split on space and tab only)
    int nmodes = 0;
    {
        size_t a = 0;
        while (a < n) {
            size_t b = a;
            while (b < n && buf[b] != '\n') ++b;
            const size_t len = b < n ? b - a + 1 : b - a;
            if (len > 1 && buf[a] != '#') {
                std::string line(buf.data() + a, len);
                char *save = nullptr;
                for (char *t = strtok_r(&line[0], " \t", &save); t; t = strtok_r(nullptr, " \t", &save)) ++nmodes;
                break;
            }
            a = b + 1;
        }
    }
    --nmodes;
    if (nmodes < 1 || nmodes > ALTO_MAX_MODES) {
        fprintf(stderr, "ALTO: bad .tns header (%d modes)\n", nmodes);
        exit(1);
    }
    // chunks at line starts
    const int T = (int)std::max<size_t>(1, std::min<size_t>((size_t)io_threads(), n / (1 << 16) + 1));
    std::vector<Chunk> ch(T);
    {
        size_t prev = 0;
        for (int c = 0; c < T; ++c) {
            size_t b = c + 1 == T ? n : std::max(prev, (size_t)((double)n * (c + 1) / T));
            while (b < n && b > 0 && buf[b - 1] != '\n') ++b;  // next line start
            ch[c].begin = prev;
            ch[c].end = b;
            prev = b;
        }
    }
    char *base = buf.data();
    auto parallel = [&](auto body) {
        std::vector<std::thread> th;
        for (int c = 1; c < T; ++c) th.emplace_back(body, c);
        body(0);
        for (auto &t : th) t.join();
    };
    // pass 1: counted lines and index maxima per chunk (read_tns_dims)
    parallel([&](int c) {
        Chunk &k = ch[c];
        for_lines(base, k.begin, k.end, [&](char *line, size_t len) {
            if (len > 1 && line[0] != '#') {
                char *p = line;
                for (int m = 0; m < nmodes; ++m) {
                    IType v = strtoull(p, &p, 10) + zero_based;
                    if (v > k.maxidx[m]) k.maxidx[m] = v;
                }
                ++k.nnz;
            }
        });
    });
    IType nnz = 0;
    std::vector<IType> off(T);
    std::vector<IType> dims(nmodes, 0);
    for (int c = 0; c < T; ++c) {
        off[c] = nnz;
        nnz += ch[c].nnz;
        for (int m = 0; m < nmodes; ++m) dims[m] = std::max(dims[m], ch[c].maxidx[m]);
    }
    X->nmodes = nmodes;
    X->nnz = nnz;
    X->dims = (IType *)malloc(sizeof(IType) * nmodes);
    X->cidx = (IType **)malloc(sizeof(IType *) * nmodes);
    for (int m = 0; m < nmodes; ++m) {
        X->dims[m] = dims[m];
        X->cidx[m] = (IType *)malloc(sizeof(IType) * (nnz ? nnz : 1));
    }
    X->vals = (FType *)malloc(sizeof(FType) * (nnz ? nnz : 1));
    // pass 2: parse into place (read_tns_data); lines are NUL-terminated now
    const IType one_based = zero_based ? 0 : 1;
    parallel([&](int c) {
        const Chunk &k = ch[c];
        IType at = off[c];
        size_t a = k.begin;
        while (a < k.end) {
            const size_t len0 = strlen(base + a);
            const bool nl = a + len0 < k.end;  // the newline was replaced by '\0'
            const size_t len = len0 + (nl ? 1 : 0);
            char *line = base + a;
            if (len > 1 && line[0] != '#') {
                char *p = line;
                for (int m = 0; m < nmodes; ++m) X->cidx[m][at] = (IType)strtoull(p, &p, 10) - one_based;
                X->vals[at] = (FType)strtod(p, &p);
                ++at;
            }
            a += len0 + 1;
        }
    });
}

void ImportSparseTensor(const char *file_path, FileFormat f, SparseTensor **X_, IType zero_based) {
    FILE *fp = fopen(file_path, "r");
    if (!fp) {
        fprintf(stderr, "ALTO: cannot open %s\n", file_path);
        exit(1);
    }
    SparseTensor *X = (SparseTensor *)calloc(1, sizeof(SparseTensor));
    if (f == TEXT_FORMAT) {
        read_text(fp, X, zero_based);
    } else {
        // .bin: u64 nmodes, u64 dims[nmodes], u64 nnz, u64 cidx per mode, f64 vals
        IType nm = 0;
        if (fread(&nm, sizeof(IType), 1, fp) != 1 || nm < 1 || nm > ALTO_MAX_MODES) {
            fprintf(stderr, "ALTO: bad .bin header in %s\n", file_path);
            exit(1);
        }
        X->nmodes = (int)nm;
        X->dims = (IType *)malloc(sizeof(IType) * nm);
        size_t ok = fread(X->dims, sizeof(IType), nm, fp);
        ok += fread(&X->nnz, sizeof(IType), 1, fp);
        X->cidx = (IType **)malloc(sizeof(IType *) * nm);
        for (IType m = 0; m < nm; ++m) {
            X->cidx[m] = (IType *)malloc(sizeof(IType) * (X->nnz ? X->nnz : 1));
            ok += fread(X->cidx[m], sizeof(IType), X->nnz, fp);
        }
        X->vals = (FType *)malloc(sizeof(FType) * (X->nnz ? X->nnz : 1));
        ok += fread(X->vals, sizeof(FType), X->nnz, fp);
        if (ok != nm + 1 + (nm + 1) * X->nnz) {
            fprintf(stderr, "ALTO: truncated .bin file %s\n", file_path);
            exit(1);
        }
    }
    fclose(fp);
    *X_ = X;
}

void DestroySparseTensor(SparseTensor *X) {
    if (!X) return;
    for (int m = 0; m < X->nmodes; ++m) free(X->cidx[m]);
    free(X->cidx);
    free(X->dims);
    free(X->vals);
    free(X);
}

// ---- kruskal_model.cpp ----
void CreateKruskalModel(int mode, IType *dims, IType rank, KruskalModel **M_) {
    KruskalModel *M = (KruskalModel *)calloc(1, sizeof(KruskalModel));
    M->mode = mode;
    M->rank = rank;
    M->dims = (IType *)malloc(sizeof(IType) * mode);
    M->U = (FType **)malloc(sizeof(FType *) * mode);
    for (int m = 0; m < mode; ++m) {
        M->dims[m] = dims[m];
        M->U[m] = (FType *)calloc(dims[m] * rank, sizeof(FType));
    }
    M->lambda = (FType *)malloc(sizeof(FType) * rank);
    for (IType r = 0; r < rank; ++r) M->lambda[r] = 1.0;
    *M_ = M;
}

static FType rand_val() {  // util.cpp:28-35
    FType v = 3.0 * ((FType)rand() / (FType)RAND_MAX);
    if (rand() % 2 == 0) v *= -1;
    return v;
}

void KruskalModelRandomInit(KruskalModel *M, unsigned int /*seed: ignored, util.cpp:37-43*/) {
    for (IType r = 0; r < M->rank; ++r) M->lambda[r] = 1.0;
    for (int m = 0; m < M->mode; ++m)
        for (IType i = 0; i < M->dims[m] * M->rank; ++i) M->U[m][i] = rand_val();
}

void DestroyKruskalModel(KruskalModel *M) {
    if (!M) return;
    for (int m = 0; m < M->mode; ++m) free(M->U[m]);
    free(M->U);
    free(M->dims);
    free(M->lambda);
    free(M);
}

// ---- device factors ----
static size_t esize(int prec) { return prec == ALTO_PREC_F64 ? 8 : prec == ALTO_PREC_F32 ? 4 : 2; }

DeviceFactors::DeviceFactors(int nmode_, const IType *dims_, IType rank_, int prec_)
    : nmode(nmode_), prec(prec_), rank(rank_), U(nmode_, nullptr), dims(dims_, dims_ + nmode_) {
    for (int m = 0; m < nmode; ++m) check(alto_hip_malloc(&U[m], esize(prec) * dims[m] * rank), "malloc");
}

DeviceFactors::~DeviceFactors() {
    for (auto p : U) alto_hip_free(p);
    if (acc) alto_hip_free(acc);
}

size_t DeviceFactors::bytes(int m) const { return esize(prec) * dims[m] * rank; }

void DeviceFactors::mttkrp(int target, alto_hip_tensor t) {
    const size_t n = dims[target] * rank;
    if (prec != ALTO_PREC_BF16) {
        check(alto_hip_mttkrp(t, target, prec, (int)rank, (const void *const *)U.data(), U[target], 0,
                              bytes(target), nullptr),
              "mttkrp_alto_par");
        return;
    }
    if (acc_bytes < 4 * n) {
        if (acc) alto_hip_free(acc);
        acc = nullptr;
        check(alto_hip_malloc(&acc, 4 * n), "malloc");
        acc_bytes = 4 * n;
    }
    check(alto_hip_convert(acc, ALTO_PREC_F32, U[target], ALTO_PREC_BF16, n, nullptr), "convert");
    check(alto_hip_mttkrp(t, target, prec, (int)rank, (const void *const *)U.data(), acc, 0, acc_bytes, nullptr),
          "mttkrp_alto_par");
    check(alto_hip_convert(U[target], ALTO_PREC_BF16, acc, ALTO_PREC_F32, n, nullptr), "convert");
}

void DeviceFactors::upload(int m, const FType *host) {
    size_t n = dims[m] * rank;
    if (prec == ALTO_PREC_F64) {
        check(alto_hip_memcpy_h2d(U[m], host, 8 * n, nullptr), "h2d");
        return;
    }
    void *tmp = nullptr;
    check(alto_hip_malloc(&tmp, 8 * n), "malloc");
    check(alto_hip_memcpy_h2d(tmp, host, 8 * n, nullptr), "h2d");
    check(alto_hip_convert(U[m], prec, tmp, ALTO_PREC_F64, n, nullptr), "convert");
    check(alto_hip_stream_sync(nullptr), "sync");
    alto_hip_free(tmp);
}

void DeviceFactors::download(int m, FType *host) {
    size_t n = dims[m] * rank;
    if (prec == ALTO_PREC_F64) {
        check(alto_hip_memcpy_d2h(host, U[m], 8 * n, nullptr), "d2h");
        return;
    }
    void *tmp = nullptr;
    check(alto_hip_malloc(&tmp, 8 * n), "malloc");
    check(alto_hip_convert(tmp, ALTO_PREC_F64, U[m], prec, n, nullptr), "convert");
    check(alto_hip_memcpy_d2h(host, tmp, 8 * n, nullptr), "d2h");
    alto_hip_free(tmp);
}

void DeviceFactors::zero(int m) { check(alto_hip_memset(U[m], 0, esize(prec) * dims[m] * rank, nullptr), "memset"); }

// ---- mttkrp.cpp:122-156 on the GPU ----
void mttkrp(SparseTensor *X, KruskalModel *M, IType mode) {
    std::vector<void *> dc(X->nmodes);
    void *dv = nullptr;
    for (int m = 0; m < X->nmodes; ++m) {
        check(alto_hip_malloc(&dc[m], 8 * X->nnz), "malloc");
        check(alto_hip_memcpy_h2d(dc[m], X->cidx[m], 8 * X->nnz, nullptr), "h2d");
    }
    check(alto_hip_malloc(&dv, 8 * X->nnz), "malloc");
    check(alto_hip_memcpy_h2d(dv, X->vals, 8 * X->nnz, nullptr), "h2d");
    DeviceFactors F(M->mode, M->dims, M->rank);
    for (int m = 0; m < M->mode; ++m) F.upload(m, M->U[m]);
    alto_hip_coo c{X->nmodes, (const uint64_t *)X->dims, X->nnz, (const uint64_t *const *)dc.data(),
                   (const double *)dv, 1};
    check(alto_hip_mttkrp_coo(&c, (int)mode, (int)M->rank, (const double *const *)F.U.data(),
                              (double *)F.U[mode], nullptr),
          "mttkrp");
    F.download((int)mode, M->U[mode]);
    for (auto p : dc) alto_hip_free(p);
    alto_hip_free(dv);
}

// ---- rowsparse_mttkrp.cpp:97-187 ----
RowSparseMatrix *rspmat_init(IType nrows, IType ncols, IType nnzr) {
    RowSparseMatrix *m = new RowSparseMatrix;
    m->I = nrows;
    m->J = ncols;
    m->nnzr = nnzr;
    m->rowind = (size_t *)calloc(nnzr ? nnzr : 1, sizeof(size_t));
    m->vals = (FType *)calloc((nnzr ? nnzr : 1) * ncols, sizeof(FType));
    return m;
}

void rspmat_free(RowSparseMatrix *mat) {
    if (!mat) return;
    free(mat->rowind);
    free(mat->vals);
    delete mat;
}

RowSparseMatrix *rowsparse_mttkrp(SparseTensor *X, RowSparseMatrix **rsp_mats, IType mode,
                                  IType streaming_mode, std::vector<std::vector<int>> &ridx) {
    const int N = X->nmodes;
    const IType R = rsp_mats[0]->J;
    std::vector<const double *> vals(N, nullptr);
    std::vector<uint64_t> rows(N, 0);
    std::vector<std::vector<int32_t>> maps(N);
    std::vector<const int32_t *> mp(N, nullptr);
    for (int m = 0; m < N; ++m) {
        if ((IType)m == mode) continue;
        vals[m] = rsp_mats[m]->vals;
        rows[m] = rsp_mats[m]->nnzr;
        if ((IType)m == streaming_mode) continue;
        maps[m].assign(ridx[m].begin(), ridx[m].end());
        maps[m].resize(X->dims[m], -1);
        mp[m] = maps[m].data();
    }
    const IType I = X->dims[mode];
    std::vector<uint64_t> rowind(I);
    std::vector<double> out(I * R);
    uint64_t nnzr = 0;
    alto_hip_coo c{X->nmodes, (const uint64_t *)X->dims, X->nnz, (const uint64_t *const *)X->cidx,
                   X->vals, 0};
    check(alto_hip_rowsparse_mttkrp(&c, (int)mode, (int)streaming_mode, (int)R, vals.data(), rows.data(),
                                    mp.data(), rowind.data(), out.data(), &nnzr, nullptr),
          "rowsparse_mttkrp");
    RowSparseMatrix *M = rspmat_init(I, R, nnzr);
    for (uint64_t i = 0; i < nnzr; ++i) M->rowind[i] = (size_t)rowind[i];
    memcpy(M->vals, out.data(), sizeof(double) * nnzr * R);
    return M;
}

// ---- constraints.cpp:42-68 ----
cpd_constraint *init_constraint() {
    cpd_constraint *con = (cpd_constraint *)calloc(1, sizeof(cpd_constraint));
    con->solve_type = UNCONSTRAINED;
    snprintf(con->description, sizeof(con->description), "UNCONSTRAINED");
    return con;
}
void free_constraint(cpd_constraint *con) { free(con); }
void apply_nonneg_constraint(cpd_constraint *con) {
    snprintf(con->description, sizeof(con->description), "NON-NEGATIVE");
    con->solve_type = NON_NEGATIVE;
}
void apply_lasso_constraint(cpd_constraint *con) {
    snprintf(con->description, sizeof(con->description), "LASSO (SPARSITY)");
    con->solve_type = LASSO;
}

// ---- streaming_cpd.cpp:368-464 ----
double cpd_error(SparseTensor *X, KruskalModel *M) {
    alto_hip_coo c{X->nmodes, (const uint64_t *)X->dims, X->nnz, (const uint64_t *const *)X->cidx, X->vals, 0};
    double err = 0.0;
    check(alto_hip_cpd_error(&c, (int)M->rank, (const double *const *)M->U, &err, nullptr, nullptr), "cpd_error");
    for (IType r = 0; r < M->rank; ++r) M->lambda[r] = 1.0;
    return err;
}

// ---- streaming_cpd.cpp:26-366 ----
void cpstream(SparseTensor *X, int rank, int max_iters, int streaming_mode, FType epsilon, IType seed,
              cpd_constraint *con, bool use_spcpstream, IType nnz_threshold, IType timeslice_limit) {
    (void)seed;  // the reference seeds once with srand (main.cpp:188)
    const bool nonneg = con && (con->solve_type == NON_NEGATIVE || con->solve_type == LASSO);
    cpstream_amd(X, rank, max_iters, streaming_mode, epsilon, nonneg, use_spcpstream, nnz_threshold,
                 timeslice_limit);
    DestroySparseTensor(X);  // as the reference (streaming_cpd.cpp:348)
}

int cpstream_amd(SparseTensor *X, int rank, int max_iters, int streaming_mode, FType epsilon, bool nonneg,
                 bool use_spcpstream, IType nnz_threshold, IType timeslice_limit, int max_slices, int prec) {
    const int N = X->nmodes;
    std::vector<std::vector<double>> init(N);
    std::vector<const double *> ip(N);
    for (int m = 0; m < N; ++m) {
        init[m].resize(X->dims[m] * rank);
        for (auto &v : init[m]) v = rand_val();
        ip[m] = init[m].data();
    }
    alto_hip_coo c{N, (const uint64_t *)X->dims, X->nnz, (const uint64_t *const *)X->cidx, X->vals, 0};
    std::vector<double> trows(X->dims[streaming_mode] * rank);
    std::vector<int> outer(X->dims[streaming_mode]);
    std::vector<double> admm(X->dims[streaming_mode]);
    double tim[4] = {0, 0, 0, 0};
    double err = 0.0;
    printf("Processing Streaming Sparse Tensor (nnz threshold: %llu, timeslice_limit: %llu)\n",
           (unsigned long long)nnz_threshold, (unsigned long long)timeslice_limit);
    fprintf(stderr, "==== Executing %s (ALTO) ====\n", use_spcpstream ? "spCPSTREAM" : "CPSTREAM");
    // the C ABI reads 0 as "no threshold"; the reference's default is ULLONG_MAX
    const uint64_t thr = nnz_threshold == ~0ULL ? 0 : (uint64_t)nnz_threshold;
    int n = use_spcpstream
                ? alto_hip_spcpstream(&c, streaming_mode, rank, max_iters, epsilon, max_slices, thr,
                                      (uint64_t)timeslice_limit, prec, ip.data(), nullptr, nullptr, trows.data(),
                                      outer.data(), nullptr, &err, tim, nullptr)
                : alto_hip_cpstream(&c, streaming_mode, rank, max_iters, epsilon, nonneg ? 1 : 0, max_slices, thr,
                                    (uint64_t)timeslice_limit, prec, ip.data(), nullptr, nullptr, trows.data(),
                                    outer.data(), admm.data(), nullptr, &err, tim, nullptr);
    if (n < 0) altoamd::fail(use_spcpstream ? "spcpstream" : "cpstream", n);
    for (int k = 0; k < n; ++k) {
        if (use_spcpstream) fprintf(stderr, "batch %d: outer iterations %d\n", k, outer[k]);
        else fprintf(stderr, "batch %d: outer iterations %d, ADMM iterations %.0f\n", k, outer[k], admm[k]);
    }
    const int nt = (int)tim[3];
    fprintf(stderr, "Total CPSTREAM time (#it, total, iter)\n%d\t%f\t%f\n", nt, tim[0], tim[0]);
    printf("Total CPSTREAM time: %f s (%d time slices in %d batches, %.2f slices/s)\n", tim[0], nt, n,
           nt / (tim[0] > 0 ? tim[0] : 1));
    printf("final fit error: %f\n", err * err);  // streaming_cpd.cpp:345
    return n;
}
