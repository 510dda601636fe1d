// cpd64 / cpd128: command line of the reference (src/main.cpp:68-410) on the
// MI355X library.  Same flags and printed lines; the work runs on the GPU.
//   -l als            CP-ALS (cpd_alto)                 main.cpp:335-372
//   -l cpstream_alto  CP-stream (cpstream)              main.cpp:376-407
//   -l spcpstream_alto  spCP-stream (cpstream, use_spcpstream)  main.cpp:163, 397
//   -p                MTTKRP benchmark (BenchmarkAlto)  main.cpp:412-530
//   -c                ALTO (GPU) vs sequential COO MTTKRP (host) check  main.cpp:532-649
//   -g / -k           dynamic time batching of CP-stream: timeslice limit and
//                     nnz threshold (next_dynamic_batch, main.cpp:130-131, 235-238)
// Extra long option: --prec f64|f32|bf16 (MTTKRP / factor storage precision).
// -d / -s / -f drive the reference's synthetic Poisson generator, which this
// build does not include: they are refused.  -o is accepted and unused, as in
// the reference (its model export is commented out, main.cpp:355).
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <chrono>
#include <climits>
#include <string>
#include <vector>

#include "alto_amd.hpp"  // LIType: 128-bit keys with ALTO_MASK_128

static const struct option long_opt[] = {
    {"help", 0, NULL, 'h'},          {"version", 0, NULL, 'v'},      {"input", 1, NULL, 'i'},
    {"output", 1, NULL, 'o'},        {"bin", 1, NULL, 'b'},          {"model", 1, NULL, 'l'},
    {"constraint", 1, NULL, 'z'},    {"rank", 1, NULL, 'r'},         {"max-iter", 1, NULL, 'm'},
    {"seed", 1, NULL, 'x'},          {"dims", 1, NULL, 'd'},         {"target-mode", 1, NULL, 't'},
    {"sparsity", 1, NULL, 's'},      {"epsilon", 1, NULL, 'e'},      {"file", 1, NULL, 'f'},
    {"zero-based", 0, NULL, 'u'},    {"check", 0, NULL, 'c'},        {"bench", 0, NULL, 'p'},
    {"streaming-mode", 1, NULL, 'a'}, {"timeslice-limit", 1, NULL, 'g'}, {"nnz-threshold", 1, NULL, 'k'},
    {"prec", 1, NULL, 'P'},          {"partitions", 1, NULL, 'n'},   {NULL, 0, NULL, 0}};
static const char *short_opt = "hvi:o:b:l:z:r:m:x:d:t:s:e:f:ucpa:g:k:n:";

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void Usage(const char *call) {
    fprintf(stderr, "Usage: %s [OPTIONS]\n", call);
    fprintf(stderr, "Options:\n");
    fprintf(stderr, "\t-h or --help         Display this information\n");
    fprintf(stderr, "\t-v or --version      Display version information\n");
    fprintf(stderr, "\t-i or --input        Input tensor file in text\n");
    fprintf(stderr, "\t-b or --bin          Input tensor file in binary\n");
    fprintf(stderr, "\t-l or --model        als | cpstream_alto | spcpstream_alto\n");
    fprintf(stderr, "\t-z or --constraint   nonneg | lasso (CP-stream)\n");
    fprintf(stderr, "\t-r or --rank         Rank\n");
    fprintf(stderr, "\t-m or --max-iter     Maximum outer iterations\n");
    fprintf(stderr, "\t-t or --target-mode  Target mode of tensor (-1: all)\n");
    fprintf(stderr, "\t-e or --epsilon      Convergence criteria\n");
    fprintf(stderr, "\t-x or --seed         Random value seed\n");
    fprintf(stderr, "\t-a or --streaming-mode  Streaming mode (CP-stream)\n");
    fprintf(stderr, "\t-g or --timeslice-limit Time slices per CP-stream batch (default 1)\n");
    fprintf(stderr, "\t-k or --nnz-threshold   Nonzeros that end a CP-stream batch (default: none)\n");
    fprintf(stderr, "\t-u or --zero-based   Assume input tensor is zero-based (default: one-based)\n");
    fprintf(stderr, "\t-c or --check        Run ALTO (GPU) validation against the sequential COO MTTKRP\n");
    fprintf(stderr, "\t-p or --bench        Run ALTO (GPU) MTTKRP benchmark\n");
    fprintf(stderr, "\t-n or --partitions   Number of ALTO line partitions (default 1)\n");
    fprintf(stderr, "\t--prec               f64 | f32 | bf16 (default f64; CP-stream: f64 | f32)\n");
}

static void PrintTensorInfo(IType rank, int max_iters, SparseTensor *X) {
    double tot = 1.0;
    printf("Tensor:");
    for (int m = 0; m < X->nmodes; ++m) {
        printf(" %llu", X->dims[m]);
        tot *= (double)X->dims[m];
    }
    printf("\nNNZ: %llu (density %g)\nRank: %llu, max iterations %d\n", X->nnz, X->nnz / tot, rank, max_iters);
}

static void BenchmarkAlto(SparseTensor *X, int max_iters, IType rank, int target_mode, int nprtn, int prec) {
    PrintTensorInfo(rank, max_iters, X);
    KruskalModel *M;
    CreateKruskalModel(X->nmodes, X->dims, rank, &M);
    KruskalModelRandomInit(M, 0);
    AltoTensor<LIType> *AT;
    double t0 = now();
    create_alto(X, &AT, nprtn);
    printf("ALTO creation time:   %f\n", now() - t0);
    DeviceFactors F(X->nmodes, X->dims, rank, prec);
    for (int m = 0; m < X->nmodes; ++m) F.upload(m, M->U[m]);
    // U[mode] += MTTKRP in the factors' own storage (BF16 accumulates
    // through an FP32 copy), as the reference's BenchmarkAlto does in FP64
    auto run = [&](int mode) { F.mttkrp(mode, AT->handle); };
    int tm = target_mode < 0 ? 0 : target_mode;
    run(tm);  // warm-up
    altoamd::check(alto_hip_stream_sync(nullptr), "sync");
    void *timer = nullptr;
    altoamd::check(alto_hip_timer_start(&timer, nullptr), "timer");
    for (int i = 0; i < max_iters; ++i) {
        if (target_mode == -1)
            for (int m = 0; m < X->nmodes; ++m) run(m);
        else
            run(target_mode);
    }
    float ms = 0;
    altoamd::check(alto_hip_timer_stop(timer, nullptr, &ms), "timer");
    printf("ALTO runtime:   %f\n", ms * 1e-3);
    destroy_alto(AT);
    DestroyKruskalModel(M);
}

// mttkrp (src/mttkrp.cpp:122-156): the sequential COO MTTKRP the reference's
// -c check compares against, on the host.
static void SeqCooMttkrp(SparseTensor *X, FType **U, IType rank, int mode, FType *out) {
    std::vector<FType> row(rank);
    for (IType i = 0; i < X->nnz; ++i) {
        for (IType r = 0; r < rank; ++r) row[r] = X->vals[i];
        for (int m = 0; m < X->nmodes; ++m) {
            if (m == mode) continue;
            const FType *u = U[m] + X->cidx[m][i] * rank;
            for (IType r = 0; r < rank; ++r) row[r] *= u[r];
        }
        FType *o = out + X->cidx[mode][i] * rank;
        for (IType r = 0; r < rank; ++r) o[r] += row[r];
    }
}

static void RunAltoCheck(SparseTensor *X, IType rank, int target_mode, int nprtn) {
    printf("===Check for mode %d===\n", target_mode);
    PrintTensorInfo(rank, 1, X);
    KruskalModel *M;
    CreateKruskalModel(X->nmodes, X->dims, rank, &M);
    KruskalModelRandomInit(M, 0);
    std::vector<std::vector<FType>> truth(X->nmodes), alto(X->nmodes);
    std::vector<FType *> fac(X->nmodes);
    for (int m = 0; m < X->nmodes; ++m) fac[m] = M->U[m];
    int m0 = target_mode < 0 ? 0 : target_mode, m1 = target_mode < 0 ? X->nmodes : target_mode + 1;
    AltoTensor<LIType> *AT;
    create_alto(X, &AT, nprtn);
    for (int m = m0; m < m1; ++m) {
        // truth: the reference's sequential COO mttkrp, on the host
        std::vector<FType> saved(M->U[m], M->U[m] + X->dims[m] * rank);
        truth[m] = saved;
        SeqCooMttkrp(X, M->U, rank, m, truth[m].data());
        mttkrp_alto_par(m, fac.data(), rank, AT);
        alto[m].assign(M->U[m], M->U[m] + X->dims[m] * rank);
        memcpy(M->U[m], saved.data(), sizeof(FType) * saved.size());
        IType cnt = 0;
        for (size_t i = 0; i < truth[m].size(); ++i)
            if (fabs(truth[m][i] - alto[m][i]) > 1.0e-6) ++cnt;
        printf("mode %d: ", m);
        if (cnt == 0) printf("Results of mttkrp_alto is correct\n");
        else printf("Results of mttkrp_alto is incorrect by %llu\n", cnt);
    }
    destroy_alto(AT);
    DestroyKruskalModel(M);
}

int main(int argc, char **argv) {
    int max_iters = 10;
    IType rank = 16;
    int target_mode = -1;
    std::string text_file, binary_file, model_string, con_string;
    IType zero_based = 0;
    double epsilon = 5e-2;
    bool do_check = false, do_bench = false;
    int streaming_mode = -1;
    int nprtn = 1;
    int prec = ALTO_PREC_F64;
    int model = -1;  // 0 als, 1 cpstream_alto, 2 spcpstream_alto
    bool nonneg = false;
    IType nnz_threshold = ULLONG_MAX;  // main.cpp:130-131
    IType timeslice_limit = 1;
    int c;
    while ((c = getopt_long(argc, argv, short_opt, long_opt, NULL)) != -1) {
        switch (c) {
        case 'h': Usage(argv[0]); return 0;
        case 'v': printf("%s version 0.1.1 (MI355X)\n", argv[0]); return 0;
        case 'i': text_file = optarg; break;
        case 'b': binary_file = optarg; break;
        case 'o': break;
        case 'l':
            model_string = optarg;
            if (model_string == "ALS" || model_string == "als") model = 0;
            else if (model_string == "CPSTREAM_ALTO" || model_string == "cpstream_alto") model = 1;
            else if (model_string == "SPCPSTREAM_ALTO" || model_string == "spcpstream_alto") model = 2;
            else fprintf(stderr, "Invalid -model: %s.\n", optarg);
            break;
        case 'z':
            con_string = optarg;
            nonneg = con_string == "nonneg" || con_string == "lasso";  // constraints.cpp:61-66
            break;
        case 'r': rank = (IType)atoll(optarg); break;
        case 'm': max_iters = atoi(optarg); break;
        case 'x': srand(atoi(optarg)); break;
        case 'd': case 's': case 'f':
            fprintf(stderr, "-%c drives the synthetic tensor generator, which this build does not include; "
                            "pass a tensor with -i or -b\n", c);
            return -1;
        case 'g': timeslice_limit = (IType)strtoull(optarg, nullptr, 10); break;
        case 'k': nnz_threshold = (IType)strtoull(optarg, nullptr, 10); break;
        case 't': target_mode = atoi(optarg); break;
        case 'e': epsilon = atof(optarg); break;
        case 'u': zero_based = 1; break;
        case 'c': do_check = true; break;
        case 'p': do_bench = true; break;
        case 'a': streaming_mode = atoi(optarg); break;
        case 'n': nprtn = atoi(optarg); break;
        case 'P':
            if (!strcmp(optarg, "f32")) prec = ALTO_PREC_F32;
            else if (!strcmp(optarg, "bf16")) prec = ALTO_PREC_BF16;
            else if (!strcmp(optarg, "f64")) prec = ALTO_PREC_F64;
            else {
                fprintf(stderr, "Invalid --prec: %s (f64 | f32 | bf16)\n", optarg);
                return -1;
            }
            break;
        default: Usage(argv[0]); return -1;
        }
    }
    SparseTensor *X = nullptr;
    double t0 = now();
    if (!binary_file.empty()) ImportSparseTensor(binary_file.c_str(), BINARY_FORMAT, &X, zero_based);
    else if (!text_file.empty()) ImportSparseTensor(text_file.c_str(), TEXT_FORMAT, &X, zero_based);
    else {
        fprintf(stderr, "No input tensor (-i or -b)... exiting\n");
        Usage(argv[0]);
        return -1;
    }
    fprintf(stderr, "Reading tensor: %f s\n", now() - t0);
    // The library initialises the GPU on first use, after the model is drawn
    // from rand(): the HIP runtime must not shift the srand(-x) sequence.
    if (do_check) {
        RunAltoCheck(X, rank, target_mode, nprtn);
        DestroySparseTensor(X);
        return 0;
    }
    if (do_bench) {
        BenchmarkAlto(X, max_iters, rank, target_mode, nprtn, prec);
        DestroySparseTensor(X);
        return 0;
    }
    if (model == 0) {
        PrintTensorInfo(rank, max_iters, X);
        KruskalModel *M;
        CreateKruskalModel(X->nmodes, X->dims, rank, &M);
        KruskalModelRandomInit(M, 0);
        AltoTensor<LIType> *AT;
        create_alto(X, &AT, nprtn);
        double w = now();
        cpd_alto(AT, M, max_iters, epsilon, prec);
        printf("CPD (ALTO): %f s\n", now() - w);
        DestroyKruskalModel(M);
        destroy_alto(AT);
    } else if (model == 1 || model == 2) {
        if (streaming_mode < 0 || streaming_mode >= X->nmodes) {
            fprintf(stderr, "Need to provide -streaming_mode: %d\n", streaming_mode);
            DestroySparseTensor(X);
            return -1;
        }
        if (prec == ALTO_PREC_BF16) {
            fprintf(stderr, "CP-stream runs its MTTKRP in f64 or f32 (--prec bf16 is for -l als and -p)\n");
            DestroySparseTensor(X);
            return -1;
        }
        if (timeslice_limit < 1) {
            fprintf(stderr, "Invalid -timeslice-limit: %llu\n", timeslice_limit);
            DestroySparseTensor(X);
            return -1;
        }
        double w = now();
        cpstream_amd(X, (int)rank, max_iters, streaming_mode, epsilon, nonneg, model == 2, nnz_threshold,
                     timeslice_limit, 0, prec);
        printf("%s (ALTO): %f s\n", model == 2 ? "spCPSTREAM" : "CPSTREAM", now() - w);
    } else {
        fprintf(stderr, "No -l model given (als | cpstream_alto | spcpstream_alto)\n");
        DestroySparseTensor(X);
        return -1;
    }
    DestroySparseTensor(X);
    return 0;
}
