// CP-ALS on the GPU: cpd_alto (include/cpd.hpp:47-191).
//
// Per iteration and mode n (all on one stream, no host round trip except the
// fit scalar the convergence test needs):
//   out = 0; out += MTTKRP_n (FP64 accumulation of FP32/BF16/FP64 products)
//   vinv = (Hadamard_{m != n} G_m + 1e-3 I)^-1      (FP64 Cholesky, 1 WG)
//   u = out * vinv, column partials (sum u^2, max u, sum u*out)
//   lambda = 2-norm (iteration 0) or max-norm (later iterations)
//   U_n = u / lambda stored in the factor precision, G_n = U_n^T U_n (FP64)
// then the fit (cpd_fit_alto, cpd.hpp:193-275) in FP64.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <thread>
#include <vector>

#include "dense.hpp"
#include "fit.hpp"
#include "tensor.hpp"

namespace alto {

__global__ __launch_bounds__(256) void k_fit(const double *__restrict__ grams, int nmode, int R,
                                             const double *__restrict__ lambda,
                                             const double *__restrict__ partials, int nblk,
                                             const double *__restrict__ normX, int fit_mode,
                                             double *__restrict__ fitd, int *__restrict__ cnt,
                                             int slot) {
    k_fit_body(grams, nmode, R, lambda, partials, nblk, normX, fit_mode, fitd, cnt, slot);
}

static size_t prec_size(int prec) { return prec == ALTO_PREC_F64 ? 8 : prec == ALTO_PREC_F32 ? 4 : 2; }

struct Events {
    std::vector<hipEvent_t> ev;
    ~Events() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    hipEvent_t make() {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ev.push_back(e);
        return e;
    }
};

static float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

}  // namespace alto

using namespace alto;

namespace alto {

// Device buffers of one CP-ALS run; reused across runs (buffers only grow).
struct AlsWork {
    DevBuf Uf[kMaxModes], mt, grams, vinv, flag, part, gpart, lam, normX, fitd, sumsq, cnt;
};

static int check_als_args(alto_hip_tensor th, const alto_hip_als_opts *o, double *const *factors) {
    if (!th || !o || !factors) return ALTO_ERR_ARG;
    const int R = o->rank, prec = o->prec;
    if (R < 1 || R > kMaxRank || o->max_iters < 0) return ALTO_ERR_ARG;
    if (prec != ALTO_PREC_F64 && prec != ALTO_PREC_F32 && prec != ALTO_PREC_BF16) return ALTO_ERR_ARG;
    for (int m = 0; m < th->impl.nmode; ++m)
        if (!factors[m]) return ALTO_ERR_ARG;
    return ALTO_OK;
}

__global__ void k_zero_int(int *p) {
    if (threadIdx.x == 0) *p = 0;
}

// Final model back to the host as FP64.
static int download_model(TensorImpl *t, int R, int prec, DevBuf *Uf, DevBuf &mt, DevBuf &lam,
                          double *const *factors, double *lambda, hipStream_t s) {
    for (int m = 0; m < t->nmode; ++m) {
        uint64_t n = t->dims[m] * R;
        ALTO_TRY(alto_hip_convert(mt.p, ALTO_PREC_F64, Uf[m].p, prec, n, s));
        ALTO_HIP_TRY(hipMemcpyAsync(factors[m], mt.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        ALTO_HIP_TRY(hipStreamSynchronize(s));
    }
    if (lambda) {
        ALTO_HIP_TRY(hipMemcpyAsync(lambda, lam.p, sizeof(double) * R, hipMemcpyDeviceToHost, s));
        ALTO_HIP_TRY(hipStreamSynchronize(s));
    }
    return ALTO_OK;
}

// cpd_alto on stream s with workspace w (arguments already checked).
static int als_run(TensorImpl *t, const alto_hip_als_opts *o, double *const *factors,
                   double *lambda, double *fits, double *fits_exact, int *iters, double *timings,
                   AlsWork &w, hipStream_t s) {
    const int R = o->rank, N = t->nmode, prec = o->prec;
    uint64_t maxI = 1;
    for (int m = 0; m < N; ++m) maxI = t->dims[m] > maxI ? t->dims[m] : maxI;
    DevBuf *Uf = w.Uf;
    DevBuf &mt = w.mt, &grams = w.grams, &vinv = w.vinv, &flag = w.flag, &part = w.part,
           &gpart = w.gpart, &lam = w.lam, &normX = w.normX, &fitd = w.fitd;
    const void *fptr[kMaxModes];
    for (int m = 0; m < N; ++m) {
        ALTO_TRY(Uf[m].alloc(prec_size(prec) * t->dims[m] * R));
        fptr[m] = Uf[m].p;
    }
    ALTO_TRY(mt.alloc(sizeof(double) * maxI * R));
    ALTO_TRY(grams.alloc(sizeof(double) * N * R * R));
    ALTO_TRY(vinv.alloc(sizeof(double) * R * R));
    ALTO_TRY(flag.alloc(sizeof(int) * 4));
    ALTO_TRY(part.alloc(sizeof(double) * kMaxBlocks * 3 * R));
    ALTO_TRY(gpart.alloc(sizeof(double) * kMaxBlocks * R * R));
    ALTO_TRY(lam.alloc(sizeof(double) * R));
    ALTO_TRY(normX.alloc(sizeof(double)));
    int maxit = o->max_iters > 0 ? o->max_iters : 1;
    ALTO_TRY(fitd.alloc(sizeof(double) * 2 * maxit));

    // Initial model -> factor storage; initial Grams (init_grams, gram.cpp:11-27).
    for (int m = 0; m < N; ++m) {
        uint64_t n = t->dims[m] * R;
        ALTO_HIP_TRY(hipMemcpyAsync(mt.p, factors[m], sizeof(double) * n, hipMemcpyHostToDevice, s));
        int gb = gram_blocks(t->dims[m], R);
        ALTO_TRY(launch_scale_gram(mt.as<double>(), nullptr, t->dims[m], R, Uf[m].p, prec,
                                   gpart.as<double>(), gb, s));
        ALTO_TRY(launch_gram_reduce(gpart.as<double>(), gb, R, grams.as<double>() + (size_t)m * R * R, s));
    }
    ALTO_TRY(w.sumsq.alloc(sizeof(double) * kSumsqBlocks));
    ALTO_TRY(launch_sumsq(t->vals64.as<double>(), t->nnz, normX.as<double>(), w.sumsq.as<double>(), s));
    // Build the per-mode output-ordered schedules and tune their launch
    // shapes outside the timed loop (tensor preparation, like create_alto).
    for (int m = 0; m < N; ++m) ALTO_TRY(prepare_mttkrp(t, m, prec, R, fptr, 1, s));
    ALTO_HIP_TRY(hipStreamSynchronize(s));

    if (!o->verbose && o->max_iters > 0) {
        // Iteration 0 is enqueued directly; the later ones, identical but
        // for the max-norm, replay one captured graph.  With epsilon > 0 the
        // fit is read back after every iteration and the loop stops as the
        // reference's does; with epsilon <= 0 (|delta fit| < epsilon never
        // holds) the fits are read once at the end.
        const bool check = o->epsilon > 0.0;
        ALTO_TRY(w.cnt.alloc(sizeof(int)));
        k_zero_int<<<1, 64, 0, s>>>(w.cnt.as<int>());
        ALTO_LAUNCH_CHECK();
        auto enqueue = [&](bool first) -> int {
            for (int n = 0; n < N; ++n) {
                const uint64_t I = t->dims[n];
                const int nb = row_blocks(I);
                // mt is zeroed by the previous mode's Gram kernel (which
                // reads the solved factor out of it), except before the
                // first MTTKRP of the run
                if (first && n == 0) ALTO_TRY(launch_fill_zero(mt.as<double>(), I * R, s));
                ALTO_TRY(run_mttkrp(t, n, prec, R, fptr, mt.p, 1, 2, 0, -1, s, false, true));
                ALTO_TRY(launch_chol_inv(grams.as<double>(), N, n, R, 1e-3, true, vinv.as<double>(),
                                         flag.as<int>(), nullptr, s));
                ALTO_TRY(launch_solve(mt.as<double>(), I, R, vinv.as<double>(), part.as<double>(), nb, s));
                ALTO_TRY(launch_lambda(part.as<double>(), nb, R, first ? 0 : 1, lam.as<double>(), s));
                const uint64_t Inext = std::max<uint64_t>(I, t->dims[(n + 1) % N]);
                ALTO_TRY(launch_scale_gram(mt.as<double>(), lam.as<double>(), I, R, Uf[n].p, prec,
                                           gpart.as<double>(), gram_blocks(I, R), s, mt.as<double>(),
                                           Inext * R));
                ALTO_TRY(launch_gram_reduce(gpart.as<double>(), gram_blocks(I, R), R,
                                            grams.as<double>() + (size_t)n * R * R, s));
            }
            const int nb = row_blocks(t->dims[N - 1]);
            const size_t lds = sizeof(double) * ((size_t)R * R + R);
            k_fit<<<1, 256, lds, s>>>(grams.as<double>(), N, R, lam.as<double>(), part.as<double>(), nb,
                                      normX.as<double>(), o->fit_mode, fitd.as<double>(),
                                      w.cnt.as<int>(), 0);
            ALTO_LAUNCH_CHECK();
            return ALTO_OK;
        };
        // The executable graph is destroyed only after the stream has
        // drained: its launches may still be queued or running until then.
        struct GraphExec {
            hipGraphExec_t ge = nullptr;
            hipStream_t s = nullptr;
            ~GraphExec() {
                if (ge) {
                    (void)hipStreamSynchronize(s);
                    (void)hipGraphExecDestroy(ge);
                }
            }
        } gx;
        gx.s = s;
        Events evs;
        hipEvent_t e0 = evs.make(), e1 = evs.make();
        std::vector<double> fd(2 * (size_t)o->max_iters);
        int done = o->max_iters;
        double prev = 0.0;
        ALTO_HIP_TRY(hipEventRecord(e0, s));
        for (int it = 0; it < o->max_iters; ++it) {
            if (it == 0) {
                ALTO_TRY(enqueue(true));
            } else if (!graphs_enabled()) {
                ALTO_TRY(enqueue(false));
            } else {
                if (!gx.ge) {
                    CaptureScope cs;  // no device allocation or release meanwhile
                    hipGraph_t g = nullptr;
                    ALTO_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    const int st = enqueue(false);
                    const hipError_t ce = hipStreamEndCapture(s, &g);
                    if (st != ALTO_OK || ce != hipSuccess) {
                        if (g) (void)hipGraphDestroy(g);
                        if (st != ALTO_OK) return st;
                        ALTO_HIP_TRY(ce);
                    }
                    const hipError_t ie = hipGraphInstantiate(&gx.ge, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    ALTO_HIP_TRY(ie);
                }
                ALTO_HIP_TRY(hipGraphLaunch(gx.ge, s));
            }
            if (check) {
                ALTO_HIP_TRY(hipMemcpyAsync(&fd[2 * it], fitd.as<double>() + 2 * it, sizeof(double) * 2,
                                            hipMemcpyDeviceToHost, s));
                ALTO_HIP_TRY(hipStreamSynchronize(s));
                const double fit = fd[2 * it];
                if (it > 0 && fabs(prev - fit) < o->epsilon) {
                    done = it + 1;
                    break;
                }
                prev = fit;
            }
        }
        ALTO_HIP_TRY(hipEventRecord(e1, s));
        if (!check)
            ALTO_HIP_TRY(hipMemcpyAsync(fd.data(), fitd.p, sizeof(double) * fd.size(), hipMemcpyDeviceToHost, s));
        ALTO_HIP_TRY(hipStreamSynchronize(s));
        for (int it = 0; it < done; ++it) {
            if (fits) fits[it] = fd[2 * it];
            if (fits_exact) fits_exact[it] = fd[2 * it + 1];
        }
        const double tot = elapsed(e0, e1) * 1e-3;
        if (timings) {
            double tv[8] = {tot, 0, 0, 0, 0, 0, 0, tot / done};
            memcpy(timings, tv, sizeof(tv));
        }
        if (iters) *iters = done;
        return download_model(t, R, prec, Uf, mt, lam, factors, lambda, s);
    }

    if (o->verbose)
        printf("Running ALTO CP-ALS with %d max iterations and %.2e epsilon\n", o->max_iters,
               o->epsilon);
    Events evs;
    double t_tot = 0, t_mttkrp = 0, t_pinv = 0, t_copy = 0, t_norm = 0, t_upd = 0, t_fit = 0;
    double fit = 0.0, prev = 0.0;
    int done = o->max_iters;
    hipEvent_t e_it0 = evs.make(), e_it1 = evs.make();
    std::vector<hipEvent_t> em(6 * N + 2);
    for (auto &e : em) e = evs.make();
    for (int it = 0; it < o->max_iters; ++it) {
        ALTO_HIP_TRY(hipEventRecord(e_it0, s));
        for (int n = 0; n < N; ++n) {
            uint64_t I = t->dims[n];
            int nb = row_blocks(I);
            hipEvent_t *E = &em[6 * n];
            ALTO_HIP_TRY(hipEventRecord(E[0], s));
            ALTO_TRY(launch_fill_zero(mt.as<double>(), I * R, s));
            ALTO_HIP_TRY(hipEventRecord(E[1], s));
            ALTO_TRY(run_mttkrp(t, n, prec, R, fptr, mt.p, 1, 2, 0, -1, s, false, true));
            ALTO_HIP_TRY(hipEventRecord(E[2], s));
            ALTO_TRY(launch_chol_inv(grams.as<double>(), N, n, R, 1e-3, true, vinv.as<double>(),
                                     flag.as<int>(), nullptr, s));
            ALTO_TRY(launch_solve(mt.as<double>(), I, R, vinv.as<double>(), part.as<double>(), nb, s));
            ALTO_HIP_TRY(hipEventRecord(E[3], s));
            ALTO_TRY(launch_lambda(part.as<double>(), nb, R, it == 0 ? 0 : 1, lam.as<double>(), s));
            ALTO_HIP_TRY(hipEventRecord(E[4], s));
            ALTO_TRY(launch_scale_gram(mt.as<double>(), lam.as<double>(), I, R, Uf[n].p, prec,
                                       gpart.as<double>(), gram_blocks(I, R), s));
            ALTO_TRY(launch_gram_reduce(gpart.as<double>(), gram_blocks(I, R), R,
                                        grams.as<double>() + (size_t)n * R * R, s));
            ALTO_HIP_TRY(hipEventRecord(E[5], s));
        }
        hipEvent_t ef0 = em[6 * N], ef1 = em[6 * N + 1];
        ALTO_HIP_TRY(hipEventRecord(ef0, s));
        {
            int nb = row_blocks(t->dims[N - 1]);
            size_t lds = sizeof(double) * ((size_t)R * R + R);
            k_fit<<<1, 256, lds, s>>>(grams.as<double>(), N, R, lam.as<double>(), part.as<double>(),
                                      nb, normX.as<double>(), o->fit_mode, fitd.as<double>(), nullptr,
                                      it);
            ALTO_LAUNCH_CHECK();
        }
        ALTO_HIP_TRY(hipEventRecord(ef1, s));
        ALTO_HIP_TRY(hipEventRecord(e_it1, s));
        double fpair[2];
        ALTO_HIP_TRY(hipMemcpyAsync(fpair, fitd.as<double>() + 2 * it, sizeof(fpair),
                                    hipMemcpyDeviceToHost, s));
        ALTO_HIP_TRY(hipStreamSynchronize(s));
        double it_s = elapsed(e_it0, e_it1) * 1e-3, mt_s = 0;
        for (int n = 0; n < N; ++n) {
            hipEvent_t *E = &em[6 * n];
            t_copy += elapsed(E[0], E[1]) * 1e-3;
            mt_s += elapsed(E[1], E[2]) * 1e-3;
            t_pinv += elapsed(E[2], E[3]) * 1e-3;
            t_norm += elapsed(E[3], E[4]) * 1e-3;
            t_upd += elapsed(E[4], E[5]) * 1e-3;
        }
        t_fit += elapsed(ef0, ef1) * 1e-3;
        t_mttkrp += mt_s;
        t_tot += it_s;
        fit = fpair[0];
        if (fits) fits[it] = fpair[0];
        if (fits_exact) fits_exact[it] = fpair[1];
        if (o->verbose)
            printf("it: %d\t fit: %g\t fit-delta: %g\ttime(for MTTKRP): %.4f s (%.4f s)\n", it, fit,
                   fabs(prev - fit), it_s, mt_s);
        if (it > 0 && fabs(prev - fit) < o->epsilon) {
            done = it + 1;
            break;
        }
        prev = fit;
    }
    int di = done > 0 ? done : 1;
    if (o->verbose) {
        printf("Total time (for MTTKRP):\t %.4f s (%.4f s)\n", t_tot, t_mttkrp);
        printf("Total     MTTKRP    PseudoInv MemCopy   Normalize Update    Fit\n");
        printf("%07.4f   %07.4f   %07.4f   %07.4f   %07.4f   %07.4f   %07.4f\n", t_tot, t_mttkrp,
               t_pinv, t_copy, t_norm, t_upd, t_fit);
        printf("Per iteration\n%07.4f   %07.4f   %07.4f   %07.4f   %07.4f   %07.4f   %07.4f\n",
               t_tot / di, t_mttkrp / di, t_pinv / di, t_copy / di, t_norm / di, t_upd / di,
               t_fit / di);
    }
    if (timings) {
        double tv[8] = {t_tot, t_mttkrp, t_pinv, t_copy, t_norm, t_upd, t_fit, t_tot / di};
        memcpy(timings, tv, sizeof(tv));
    }
    if (iters) *iters = o->max_iters > 0 ? done : 0;
    return download_model(t, R, prec, Uf, mt, lam, factors, lambda, s);
}

}  // namespace alto

extern "C" int alto_hip_cpd_als(alto_hip_tensor th, const alto_hip_als_opts *o,
                                double *const *factors, double *lambda, double *fits,
                                double *fits_exact, int *iters, double *timings, void *stream) {
    ALTO_TRY(check_als_args(th, o, factors));
    ALTO_TRY(ensure_init());
    AlsWork w;
    return als_run(&th->impl, o, factors, lambda, fits, fits_exact, iters, timings, w,
                   to_stream(stream));
}

// The batched CP-ALS over host (dev_io false) or device (true) FP64 models.
static int cpd_als_batch(int ntensors, const alto_hip_tensor *tensors, const alto_hip_als_opts *opts,
                         double *const *const *factors, double *const *lambdas, double *const *fits, int *iters,
                         int nstreams, double *timings, bool dev_io) {
    if (ntensors < 0 || (ntensors > 0 && (!tensors || !opts || !factors)) || nstreams < 1 ||
        nstreams > 64)
        return ALTO_ERR_ARG;
    for (int k = 0; k < ntensors; ++k) {
        ALTO_TRY(check_als_args(tensors[k], &opts[k], factors[k]));
        if (opts[k].verbose) return ALTO_ERR_ARG;
    }
    ALTO_TRY(ensure_init());
    int dev = 0;
    ALTO_HIP_TRY(hipGetDevice(&dev));
    const auto t0 = std::chrono::steady_clock::now();
    // Schedules and launch tuning first, on one stream: they allocate and
    // free device memory, which would serialise the concurrent runs, and the
    // tuner times kernels, which concurrent runs would disturb.  The tuner
    // reads zero factors of each mode's shape (timing does not depend on the
    // values).
    hipStream_t s0 = default_stream();
    uint64_t fmax = 1;
    for (int k = 0; k < ntensors; ++k)
        for (int m = 0; m < tensors[k]->impl.nmode; ++m)
            fmax = std::max<uint64_t>(fmax, tensors[k]->impl.dims[m] * (uint64_t)opts[k].rank);
    DevBuf zf;
    ALTO_TRY(zf.alloc(sizeof(double) * fmax));
    ALTO_TRY(launch_fill_zero(zf.as<double>(), fmax, s0));
    for (int k = 0; k < ntensors; ++k) {
        const void *zp[kMaxModes];
        for (int m = 0; m < kMaxModes; ++m) zp[m] = zf.p;
        for (int m = 0; m < tensors[k]->impl.nmode; ++m)
            ALTO_TRY(prepare_mttkrp(&tensors[k]->impl, m, opts[k].prec, opts[k].rank, zp, 1, s0));
    }
    ALTO_HIP_TRY(hipStreamSynchronize(s0));
    const auto t1 = std::chrono::steady_clock::now();
    // Largest tensors first.
    std::vector<int> order(ntensors);
    for (int k = 0; k < ntensors; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return tensors[a]->impl.nnz * opts[a].max_iters > tensors[b]->impl.nnz * opts[b].max_iters;
    });
    // Tensors of a shape class the grouped CP-ALS takes run class by class,
    // one launch per ALS step and class; the others on the worker threads.
    std::vector<int> rest;
    if (grouped_als_enabled()) {
        std::map<std::tuple<int, int, int, int>, std::vector<int>> cls;
        for (int k : order) {
            const TensorImpl *t = &tensors[k]->impl;
            if (als_group_eligible(t, &opts[k]))
                cls[std::make_tuple(t->nmode, opts[k].rank, opts[k].prec, opts[k].fit_mode)].push_back(k);
            else
                rest.push_back(k);
        }
        std::vector<std::vector<int>> classes;
        for (auto &kv : cls) classes.push_back(kv.second);
        ALTO_TRY(als_groups_run(classes, tensors, opts, factors, lambdas, fits, iters, dev_io));
    } else {
        rest = order;
    }
    std::atomic<int> next{0};
    std::atomic<int> status{ALTO_OK};
    std::mutex err_mu;
    std::string err_msg;
    auto worker = [&]() {
        if (hipSetDevice(dev) != hipSuccess) {
            status = ALTO_ERR_HIP;
            return;
        }
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            status = ALTO_ERR_HIP;
            return;
        }
        {
            AlsWork w;
            for (int i = next++; i < (int)rest.size() && status == ALTO_OK; i = next++) {
                const int k = rest[i];
                const TensorImpl *t = &tensors[k]->impl;
                // a device model outside the grouped classes: a host round trip
                std::vector<std::vector<double>> hf(dev_io ? t->nmode : 0);
                double *hp[kMaxModes] = {};
                int st = ALTO_OK;
                for (int m = 0; m < (int)hf.size() && st == ALTO_OK; ++m) {
                    const size_t cnt = (size_t)t->dims[m] * opts[k].rank;
                    hf[m].resize(cnt);
                    hp[m] = hf[m].data();
                    if (hipMemcpyAsync(hp[m], factors[k][m], 8 * cnt, hipMemcpyDeviceToHost, s) != hipSuccess ||
                        hipStreamSynchronize(s) != hipSuccess)
                        st = ALTO_ERR_HIP;
                }
                if (st == ALTO_OK)
                    st = als_run(&tensors[k]->impl, &opts[k], dev_io ? hp : factors[k], lambdas ? lambdas[k] : nullptr,
                                 fits ? fits[k] : nullptr, nullptr, iters ? iters + k : nullptr, nullptr, w, s);
                for (int m = 0; m < (int)hf.size() && st == ALTO_OK; ++m)
                    if (hipMemcpyAsync(factors[k][m], hp[m], 8 * hf[m].size(), hipMemcpyHostToDevice, s) !=
                            hipSuccess ||
                        hipStreamSynchronize(s) != hipSuccess)
                        st = ALTO_ERR_HIP;
                if (st != ALTO_OK) {
                    std::lock_guard<std::mutex> g(err_mu);
                    if (status == ALTO_OK) {
                        status = st;
                        err_msg = last_error();
                    }
                }
            }
            (void)hipStreamSynchronize(s);
        }
        (void)hipStreamDestroy(s);
    };
    const int nrest = (int)rest.size();
    const int nw = nrest == 0 ? 0 : nstreams < nrest ? nstreams : nrest;
    std::vector<std::thread> pool;
    for (int i = 0; i < nw; ++i) pool.emplace_back(worker);
    for (auto &th : pool) th.join();
    if (status != ALTO_OK) {
        set_error("%s", err_msg.c_str());
        return status;
    }
    if (timings) {
        timings[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
        timings[1] = std::chrono::duration<double>(t1 - t0).count();
    }
    return ALTO_OK;
}

extern "C" int alto_hip_cpd_als_batch(int ntensors, const alto_hip_tensor *tensors,
                                      const alto_hip_als_opts *opts, double *const *const *factors,
                                      double *const *lambdas, double *const *fits, int *iters,
                                      int nstreams, double *timings) {
    return cpd_als_batch(ntensors, tensors, opts, factors, lambdas, fits, iters, nstreams, timings, false);
}

extern "C" int alto_hip_cpd_als_batch_dev(int ntensors, const alto_hip_tensor *tensors,
                                          const alto_hip_als_opts *opts, double *const *const *factors,
                                          double *const *lambdas, double *const *fits, int *iters,
                                          int nstreams, double *timings) {
    if (ntensors > 0 && (!tensors || !factors)) return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    for (int k = 0; k < ntensors; ++k) {  // device models: refuse memory the GPU cannot read
        if (!tensors[k] || !factors[k]) return ALTO_ERR_ARG;
        for (int m = 0; m < tensors[k]->impl.nmode; ++m) {
            if (!factors[k][m]) return ALTO_ERR_ARG;
            ALTO_TRY(check_device_access(factors[k][m], "batched CP-ALS device model"));
        }
    }
    return cpd_als_batch(ntensors, tensors, opts, factors, lambdas, fits, iters, nstreams, timings, true);
}

// pseudo_inverse(grams, M, mode) (cpd.hpp:377-480): U <- U (Had_{m != mode}
// G_m + 1e-3 I)^{-1} row by row; when the factorisation fails, the
// minimum-norm solve of the unregularised diagonal.  Host arrays.
extern "C" int alto_hip_pseudo_inverse(int nmode, int rank, const double *grams, int mode, uint64_t rows,
                                       double *U, int *fallback, void *stream) {
    if (nmode < 1 || nmode > kMaxModes || rank < 1 || rank > kMaxRank || mode < 0 || mode >= nmode || !grams ||
        (!U && rows))
        return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    hipStream_t s = to_stream(stream);
    const size_t RR = (size_t)rank * rank;
    DevBuf g, u, vinv, flag, part;
    ALTO_TRY(g.alloc(sizeof(double) * nmode * RR));
    ALTO_TRY(u.alloc(sizeof(double) * (rows ? rows : 1) * rank));
    ALTO_TRY(vinv.alloc(sizeof(double) * RR));
    ALTO_TRY(flag.alloc(sizeof(int)));
    const int nb = row_blocks(rows ? rows : 1);
    ALTO_TRY(part.alloc(sizeof(double) * 3 * rank * nb));
    ALTO_HIP_TRY(hipMemcpyAsync(g.p, grams, sizeof(double) * nmode * RR, hipMemcpyHostToDevice, s));
    if (rows) ALTO_HIP_TRY(hipMemcpyAsync(u.p, U, sizeof(double) * rows * rank, hipMemcpyHostToDevice, s));
    ALTO_TRY(launch_chol_inv(g.as<double>(), nmode, mode, rank, 1e-3, true, vinv.as<double>(), flag.as<int>(),
                             nullptr, s));
    if (rows) {
        ALTO_TRY(launch_solve(u.as<double>(), rows, rank, vinv.as<double>(), part.as<double>(), nb, s));
        ALTO_HIP_TRY(hipMemcpyAsync(U, u.p, sizeof(double) * rows * rank, hipMemcpyDeviceToHost, s));
    }
    int hf = 0;
    ALTO_HIP_TRY(hipMemcpyAsync(&hf, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    ALTO_HIP_TRY(hipStreamSynchronize(s));
    if (fallback) *fallback = hf;
    return ALTO_OK;
}

namespace alto {

// partials[b * 3R + 2R + c] = lambda_c * sum over the rows of block b of
// U[i, c] * MT[i, c] (the layout k_fit reads).
__global__ __launch_bounds__(256) void k_fit_inner(const double *__restrict__ U, const double *__restrict__ MT,
                                                   const double *__restrict__ lambda, uint64_t rows, int R,
                                                   double *__restrict__ part) {
    const int c = threadIdx.x;
    if (c >= R) return;
    const uint64_t r0 = (uint64_t)blockIdx.x * 64, r1 = r0 + 64 < rows ? r0 + 64 : rows;
    double s = 0.0;
    for (uint64_t i = r0; i < r1; ++i) s += U[i * R + c] * MT[i * R + c];
    part[(size_t)blockIdx.x * 3 * R + 2 * R + c] = lambda[c] * s;
}

// Symmetrises nmode R x R matrices from their row-major upper triangles.
__global__ void k_sym_upper(double *__restrict__ g, int nmode, int R) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nmode * R * R) return;
    const int m = e / (R * R), ij = e % (R * R), i = ij / R, j = ij % R;
    if (i > j) g[e] = g[(size_t)m * R * R + j * R + i];
}

}  // namespace alto

extern "C" int alto_hip_cpd_fit(int nmode, int rank, uint64_t rows, const double *U_last,
                                const double *mttkrp_last, const double *lambda, const double *grams,
                                double normX, int fit_mode, double *fit, void *stream) {
    if (nmode < 2 || nmode > kMaxModes || rank < 1 || rank > kMaxRank || !U_last || !mttkrp_last ||
        !lambda || !grams || !fit || (rows && !U_last))
        return ALTO_ERR_ARG;
    if (fit_mode != ALTO_FIT_EXACT && fit_mode != ALTO_FIT_REFERENCE_DIAG) return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    hipStream_t s = to_stream(stream);
    const int R = rank;
    const size_t RR = (size_t)R * R;
    const uint64_t nb = rows ? (rows + 63) / 64 : 1;
    if (nb > (1ULL << 31)) return ALTO_ERR_UNSUPPORTED;
    DevBuf dU, dM, dl, dg, dp, dn, df;
    ALTO_TRY(dU.alloc(sizeof(double) * (rows ? rows : 1) * R));
    ALTO_TRY(dM.alloc(sizeof(double) * (rows ? rows : 1) * R));
    ALTO_TRY(dl.alloc(sizeof(double) * R));
    ALTO_TRY(dg.alloc(sizeof(double) * nmode * RR));
    ALTO_TRY(dp.alloc(sizeof(double) * nb * 3 * R));
    ALTO_TRY(dn.alloc(sizeof(double)));
    ALTO_TRY(df.alloc(sizeof(double) * 2));
    if (rows) {
        ALTO_HIP_TRY(hipMemcpyAsync(dU.p, U_last, sizeof(double) * rows * R, hipMemcpyHostToDevice, s));
        ALTO_HIP_TRY(hipMemcpyAsync(dM.p, mttkrp_last, sizeof(double) * rows * R, hipMemcpyHostToDevice, s));
    }
    ALTO_HIP_TRY(hipMemcpyAsync(dl.p, lambda, sizeof(double) * R, hipMemcpyHostToDevice, s));
    ALTO_HIP_TRY(hipMemcpyAsync(dg.p, grams, sizeof(double) * nmode * RR, hipMemcpyHostToDevice, s));
    ALTO_HIP_TRY(hipMemcpyAsync(dn.p, &normX, sizeof(double), hipMemcpyHostToDevice, s));
    ALTO_TRY(launch_fill_zero(dp.as<double>(), nb * 3 * R, s));
    if (rows) k_fit_inner<<<(unsigned)nb, 256, 0, s>>>(dU.as<double>(), dM.as<double>(), dl.as<double>(), rows, R,
                                                      dp.as<double>());
    k_sym_upper<<<cdiv(nmode * RR, 256), 256, 0, s>>>(dg.as<double>(), nmode, R);
    k_fit<<<1, 256, sizeof(double) * (RR + R), s>>>(dg.as<double>(), nmode, R, dl.as<double>(), dp.as<double>(),
                                                   (int)nb, dn.as<double>(), fit_mode, df.as<double>(), nullptr, 0);
    ALTO_LAUNCH_CHECK();
    double out[2];
    ALTO_HIP_TRY(hipMemcpyAsync(out, df.p, sizeof(out), hipMemcpyDeviceToHost, s));
    ALTO_HIP_TRY(hipStreamSynchronize(s));
    *fit = out[0];
    return ALTO_OK;
}
