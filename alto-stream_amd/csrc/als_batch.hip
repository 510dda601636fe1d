// Grouped CP-ALS: cpd_alto (include/cpd.hpp:47-191) on many independent
// tensors of one class (mode count, rank, precision, fit convention) at once.
//
// Every step of an ALS iteration is one launch for the whole class: the
// MTTKRP of mode n over all tensors' output-ordered schedules, the R x R
// inverses (one workgroup per tensor), the row solve, the lambda, the
// scaling with the Gram partials and their reduction, then the fits and
// the stopping rule.  A launch finds its tensor from the workgroup index
// (prefix sums of the tensors' workgroup counts) and runs the same code the
// single-tensor kernels run on that tensor's buffers, so each tensor gets
// the result of cpd_alto on it alone.  A tensor that has stopped (its
// max_iters reached or its fit converged) is skipped by every launch.
// Iterations after the first replay one captured graph.
//
// Host side: the initial models (FP64) are converted to the factor
// precision on the host by a thread pool into pinned staging buffers, whose
// layout mirrors the device layout, and copied in 64 MB pieces; the final
// models come back the same way.  Several classes run on their own streams,
// so one class's copies overlap another's iterations.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "dense.hpp"
#include "fit.hpp"
#include "tensor.hpp"

namespace alto {

// Rows of a factor per solve / Gram workgroup in a grouped launch.
constexpr uint64_t kGroupRows = 1024;

__global__ void k_set_active(int *__restrict__ active, const int *__restrict__ max_iters, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) active[i] = max_iters[i] > 0 ? 1 : 0;
}

// Grouped fit (one workgroup per instance) and stopping rule: after the fit
// of iteration c - 1 an instance stops when c reaches its max_iters or, for
// c >= 2, when its fit moved by less than epsilon (cpd.hpp:162-166).
__global__ __launch_bounds__(256) void k_fit_group(const FitJob *__restrict__ jobs, int nmode, int R,
                                                   int fit_mode) {
    const FitJob &J = jobs[blockIdx.x];
    if (!*J.active) return;
    k_fit_body(J.grams, nmode, R, J.lam, J.part, J.nblk, J.normX, fit_mode, J.fitd, J.cnt, 0);
}

__global__ void k_stop_group(const FitJob *__restrict__ jobs, int njobs, int *__restrict__ nactive) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= njobs) return;
    const FitJob &J = jobs[i];
    if (!*J.active) return;
    const int c = *J.cnt;
    bool stop = c >= J.max_iters;
    if (!stop && c >= 2) stop = fabs(J.fitd[2 * (c - 1)] - J.fitd[2 * (c - 2)]) < J.epsilon;
    if (stop) *J.active = 0;
    else atomicAdd(nactive, 1);
}

__global__ void k_zero_one(int *p) {
    if (threadIdx.x == 0) *p = 0;
}

// Fixed set of host threads running one function on every thread id.
class HostPool {
  public:
    explicit HostPool(int n) : n_(n < 1 ? 1 : n) {
        for (int i = 1; i < n_; ++i) th_.emplace_back([this, i] { loop(i); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    int size() const { return n_; }
    void run(const std::function<void(int)> &f) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            pending_ = n_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }

  private:
    void loop(int id) {
        int seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return quit_ || gen_ != seen; });
            if (quit_) return;
            seen = gen_;
            const std::function<void(int)> *f = job_;
            lk.unlock();
            (*f)(id);
            lk.lock();
            if (--pending_ == 0) done_.notify_all();
        }
    }
    int n_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *job_ = nullptr;
    int gen_ = 0, pending_ = 0;
    bool quit_ = false;
};

// Pinned staging for the model copies: two 64 MB buffers used in turn.
struct Staging {
    static constexpr size_t kBytes = 64ull << 20;
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    ~Staging() {
        for (int b = 0; b < 2; ++b) {
            if (ev[b]) {
                (void)hipEventSynchronize(ev[b]);
                (void)hipEventDestroy(ev[b]);
            }
            if (buf[b]) (void)hipHostFree(buf[b]);
        }
    }
    int init() {
        for (int b = 0; b < 2; ++b) {
            ALTO_HIP_TRY(hipHostMalloc(&buf[b], kBytes, hipHostMallocDefault));
            ALTO_HIP_TRY(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
        }
        return ALTO_OK;
    }
};

// One host FP64 matrix and its place in the device factor region.
struct Seg {
    size_t off;  // bytes from the region's start
    size_t n;    // elements
    double *host;
};

// The overlap of segment g with the byte window [w0, w1) of the region, in
// elements of size esz: [e0, e1).
static inline bool seg_window(const Seg &g, size_t esz, size_t w0, size_t w1, size_t *e0, size_t *e1) {
    const size_t a = g.off, b = g.off + g.n * esz;
    if (b <= w0 || a >= w1) return false;
    *e0 = a >= w0 ? 0 : (w0 - a) / esz;
    *e1 = b <= w1 ? g.n : (w1 - a) / esz;
    return *e1 > *e0;
}

// Converts the segments' host FP64 values into the region at dev (FP64 or
// FP32), window by window through the staging buffers, on stream s.
static int upload_region(const std::vector<Seg> &segs, size_t bytes, int prec, char *dev, Staging &st,
                         HostPool &pool, hipStream_t s) {
    const size_t esz = prec == ALTO_PREC_F64 ? 8 : 4;
    size_t first = 0;
    for (size_t w0 = 0, c = 0; w0 < bytes; w0 += Staging::kBytes, ++c) {
        const int b = (int)(c & 1);
        const size_t w1 = std::min(bytes, w0 + Staging::kBytes);
        if (st.used[b]) ALTO_HIP_TRY(hipEventSynchronize(st.ev[b]));
        char *hb = (char *)st.buf[b];
        while (first < segs.size() && segs[first].off + segs[first].n * esz <= w0) ++first;
        const int T = pool.size();
        pool.run([&](int id) {
            for (size_t g = first; g < segs.size() && segs[g].off < w1; ++g) {
                size_t e0, e1;
                if (!seg_window(segs[g], esz, w0, w1, &e0, &e1)) continue;
                const size_t len = e1 - e0, lo = e0 + len * id / T, hi = e0 + len * (id + 1) / T;
                if (hi <= lo) continue;
                const double *src = segs[g].host + lo;
                char *dst = hb + (segs[g].off + lo * esz - w0);  // element lo's byte in the window
                if (prec == ALTO_PREC_F64) {
                    memcpy(dst, src, (hi - lo) * 8);
                } else {
                    float *d = (float *)dst;
                    for (size_t e = 0; e < hi - lo; ++e) d[e] = (float)src[e];
                }
            }
        });
        ALTO_HIP_TRY(hipMemcpyAsync(dev + w0, hb, w1 - w0, hipMemcpyHostToDevice, s));
        ALTO_HIP_TRY(hipEventRecord(st.ev[b], s));
        st.used[b] = true;
    }
    return ALTO_OK;
}

// The reverse of upload_region: device values (FP64 or FP32) to the host
// FP64 matrices; the next window's copy runs while one is converted.
static int download_region(const std::vector<Seg> &segs, size_t bytes, int prec, const char *dev, Staging &st,
                           HostPool &pool, hipStream_t s) {
    const size_t esz = prec == ALTO_PREC_F64 ? 8 : 4;
    const size_t nw = (bytes + Staging::kBytes - 1) / Staging::kBytes;
    auto issue = [&](size_t c) -> int {
        const int b = (int)(c & 1);
        const size_t w0 = c * Staging::kBytes, w1 = std::min(bytes, w0 + Staging::kBytes);
        if (st.used[b]) ALTO_HIP_TRY(hipEventSynchronize(st.ev[b]));
        ALTO_HIP_TRY(hipMemcpyAsync(st.buf[b], dev + w0, w1 - w0, hipMemcpyDeviceToHost, s));
        ALTO_HIP_TRY(hipEventRecord(st.ev[b], s));
        st.used[b] = true;
        return ALTO_OK;
    };
    if (nw) ALTO_TRY(issue(0));
    size_t first = 0;
    for (size_t c = 0; c < nw; ++c) {
        const int b = (int)(c & 1);
        const size_t w0 = c * Staging::kBytes, w1 = std::min(bytes, w0 + Staging::kBytes);
        ALTO_HIP_TRY(hipEventSynchronize(st.ev[b]));
        st.used[b] = false;
        if (c + 1 < nw) ALTO_TRY(issue(c + 1));
        const char *hb = (const char *)st.buf[b];
        while (first < segs.size() && segs[first].off + segs[first].n * esz <= w0) ++first;
        const int T = pool.size();
        pool.run([&](int id) {
            for (size_t g = first; g < segs.size() && segs[g].off < w1; ++g) {
                size_t e0, e1;
                if (!seg_window(segs[g], esz, w0, w1, &e0, &e1)) continue;
                const size_t len = e1 - e0, lo = e0 + len * id / T, hi = e0 + len * (id + 1) / T;
                if (hi <= lo) continue;
                double *dst = segs[g].host + lo;
                const char *src = hb + (segs[g].off + lo * esz - w0);  // element lo's byte in the window
                if (prec == ALTO_PREC_F64) {
                    memcpy(dst, src, (hi - lo) * 8);
                } else {
                    const float *f = (const float *)src;
                    for (size_t e = 0; e < hi - lo; ++e) dst[e] = (double)f[e];
                }
            }
        });
    }
    return ALTO_OK;
}

// Device-resident models: the segments' FP64 matrices are device arrays,
// converted into / out of the region on the stream (no host copy).
static int upload_region_dev(const std::vector<Seg> &segs, int prec, char *dev, hipStream_t s) {
    for (const Seg &g : segs) ALTO_TRY(alto_hip_convert(dev + g.off, prec, g.host, ALTO_PREC_F64, g.n, s));
    return ALTO_OK;
}
static int download_region_dev(const std::vector<Seg> &segs, int prec, const char *dev, hipStream_t s) {
    for (const Seg &g : segs) ALTO_TRY(alto_hip_convert(g.host, ALTO_PREC_F64, dev + g.off, prec, g.n, s));
    return ALTO_OK;
}

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// One class of tensors in flight: device memory, jobs, graph.
struct AlsGroup {
    std::vector<int> ks;  // tensor indices
    int N = 0, R = 0, prec = 0, fit_mode = 0, maxit = 0;
    bool any_eps = false;
    bool dev_io = false;  // the models are device arrays
    hipStream_t s = nullptr;
    DevBuf arena, jobs;
    char *Ubase = nullptr;
    size_t Ubytes = 0;
    std::vector<Seg> segs;
    // device pointers into jobs
    const BatchJob *bj[kMaxModes] = {};
    // MTTKRP jobs per mode and kernel (0: fiber, 1: decode-once)
    const MttkrpJob *mj[kMaxModes][2] = {};
    const unsigned *fm[kMaxModes][2] = {};
    unsigned tm[kMaxModes][2] = {};
    int nm[kMaxModes][2] = {};
    const unsigned *fs[kMaxModes] = {}, *fg[kMaxModes] = {};
    unsigned ts[kMaxModes] = {}, tg[kMaxModes] = {};
    const FitJob *fj = nullptr;
    double *lam = nullptr, *fitd = nullptr;
    int *cnt = nullptr, *nactive = nullptr;
    hipGraphExec_t ge = nullptr;
    ~AlsGroup() {
        if (ge) {
            (void)hipStreamSynchronize(s);
            (void)hipGraphExecDestroy(ge);
        }
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }

    int enqueue(bool first_iter) {
        const int n = (int)ks.size();
        k_zero_one<<<1, 64, 0, s>>>(nactive);
        for (int m = 0; m < N; ++m) {
            for (int d = 0; d < 2; ++d)
                ALTO_TRY(launch_mttkrp_group(mj[m][d], fm[m][d], nm[m][d], tm[m][d], prec, R, N, d == 1, s));
            ALTO_TRY(launch_inv_group(bj[m], n, N, m, R, 1e-3, s));
            ALTO_TRY(launch_solve_group(bj[m], fs[m], n, ts[m], R, s));
            ALTO_TRY(launch_lambda_group(bj[m], n, R, first_iter ? 0 : 1, s));
            ALTO_TRY(launch_scale_gram_group(bj[m], fg[m], n, tg[m], R, prec, 1, s));
            ALTO_TRY(launch_gram_reduce_group(bj[m], n, R, s));
        }
        const size_t lds = sizeof(double) * ((size_t)R * R + R);
        k_fit_group<<<n, 256, lds, s>>>(fj, N, R, fit_mode);
        k_stop_group<<<cdiv(n, 256), 256, 0, s>>>(fj, n, nactive);
        ALTO_LAUNCH_CHECK();
        return ALTO_OK;
    }
};

// Device layout, jobs, model upload and the initial Grams of one class, then
// its iterations (all enqueued on the class's stream; with epsilon > 0 the
// host reads the count of running tensors after every iteration).
static int group_start(AlsGroup &G, const alto_hip_tensor *tensors, const alto_hip_als_opts *opts,
                       double *const *const *factors, Staging &st, HostPool &pool) {
    const int n = (int)G.ks.size();
    const alto_hip_als_opts &o0 = opts[G.ks[0]];
    G.N = tensors[G.ks[0]]->impl.nmode;
    G.R = o0.rank;
    G.prec = o0.prec;
    G.fit_mode = o0.fit_mode;
    const int N = G.N, R = G.R;
    const size_t RR = (size_t)R * R, esz = G.prec == ALTO_PREC_F64 ? 8 : 4;
    G.maxit = 1;
    for (int k : G.ks) {
        G.maxit = std::max(G.maxit, opts[k].max_iters);
        G.any_eps = G.any_eps || opts[k].epsilon > 0.0;
    }
    ALTO_HIP_TRY(hipStreamCreateWithFlags(&G.s, hipStreamNonBlocking));
    hipStream_t s = G.s;

    // ---- layout: the factor region first (mirrored by the host staging) ----
    struct Mem {
        size_t U[kMaxModes], mt, grams, vinv, part, gpart;
        int nbs[kMaxModes], nbg[kMaxModes];
        uint64_t maxI;
    };
    std::vector<Mem> mem(n);
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
        const TensorImpl *t = &tensors[G.ks[i]]->impl;
        for (int m = 0; m < N; ++m) {
            mem[i].U[m] = off;
            off = align256(off + esz * t->dims[m] * R);
        }
    }
    G.Ubytes = off;
    size_t maxs = 1, maxg = 1;
    for (int i = 0; i < n; ++i) {
        const TensorImpl *t = &tensors[G.ks[i]]->impl;
        Mem &M = mem[i];
        M.maxI = 1;
        int ns = 1, ng = 1;
        for (int m = 0; m < N; ++m) {
            M.maxI = std::max<uint64_t>(M.maxI, t->dims[m]);
            // the class's tensors fill the device together, so each takes
            // one block per kGroupRows rows of a factor (not the single-run
            // count of one block per 64 rows): fewer and smaller partials
            const int nb = (int)std::max<uint64_t>(1, (t->dims[m] + kGroupRows - 1) / kGroupRows);
            M.nbs[m] = std::min(row_blocks(t->dims[m]), nb);
            M.nbg[m] = std::min(gram_blocks(t->dims[m], R), nb);
            ns = std::max(ns, M.nbs[m]);
            ng = std::max(ng, M.nbg[m]);
        }
        maxs = std::max<size_t>(maxs, ns);
        maxg = std::max<size_t>(maxg, ng);
        M.mt = off;
        off = align256(off + 8 * M.maxI * R);
        M.grams = off;
        off = align256(off + 8 * N * RR);
        M.vinv = off;
        off = align256(off + 8 * RR);
        M.part = off;
        off = align256(off + 8 * (size_t)ns * 3 * R);
        M.gpart = off;
        off = align256(off + 8 * (size_t)ng * RR);
    }
    (void)maxs;
    (void)maxg;
    const size_t o_lam = off;
    off = align256(off + 8 * (size_t)n * R);
    const size_t o_normX = off;
    off = align256(off + 8 * (size_t)n);
    const size_t o_fitd = off;
    off = align256(off + 8 * (size_t)n * 2 * G.maxit);
    const size_t o_flag = off;
    off = align256(off + 16 * (size_t)n);
    const size_t o_cnt = off;
    off = align256(off + 4 * (size_t)n);
    const size_t o_active = off;
    off = align256(off + 4 * (size_t)n);
    const size_t o_maxit = off;
    off = align256(off + 4 * (size_t)n);
    const size_t o_nactive = off;
    off = align256(off + 4);
    const size_t o_sumsq = off;
    off = align256(off + 8 * (size_t)kSumsqBlocks);
    ALTO_TRY(G.arena.alloc(off));
    char *A = (char *)G.arena.p;
    G.Ubase = A;
    // everything after the factors starts at zero (counters, the first MTTKRP's outputs)
    ALTO_TRY(launch_fill_zero((double *)(A + G.Ubytes), (off - G.Ubytes) / 8, s));
    G.lam = (double *)(A + o_lam);
    G.fitd = (double *)(A + o_fitd);
    G.cnt = (int *)(A + o_cnt);
    G.nactive = (int *)(A + o_nactive);
    int *active = (int *)(A + o_active);
    {
        std::vector<int> mi(n);
        for (int i = 0; i < n; ++i) mi[i] = opts[G.ks[i]].max_iters;
        ALTO_HIP_TRY(hipMemcpyAsync(A + o_maxit, mi.data(), 4 * (size_t)n, hipMemcpyHostToDevice, s));
        k_set_active<<<cdiv(n, 256), 256, 0, s>>>(active, (const int *)(A + o_maxit), n);
        ALTO_LAUNCH_CHECK();
        ALTO_HIP_TRY(hipStreamSynchronize(s));  // mi leaves scope
    }
    // ||X||^2 of each tensor (the sum of squared stored values, cpd.hpp:80-83)
    for (int i = 0; i < n; ++i) {
        TensorImpl *t = &tensors[G.ks[i]]->impl;
        ALTO_TRY(launch_sumsq(t->vals64.as<double>(), t->nnz, (double *)(A + o_normX) + i,
                              (double *)(A + o_sumsq), s));
    }

    // ---- jobs ----
    std::vector<BatchJob> bj((size_t)N * n);
    std::vector<MttkrpJob> mj((size_t)N * n);  // per mode: fiber jobs, then decode-once jobs
    std::vector<unsigned> fs((size_t)N * (n + 1)), fg((size_t)N * (n + 1)), fm((size_t)N * (n + 2));
    std::vector<MttkrpJob> tmp(n);
    std::vector<FitJob> fj(n);
    for (int i = 0; i < n; ++i) {
        const int k = G.ks[i];
        TensorImpl *t = &tensors[k]->impl;
        const Mem &M = mem[i];
        double *mt = (double *)(A + M.mt), *grams = (double *)(A + M.grams);
        for (int m = 0; m < N; ++m) {
            BatchJob &J = bj[(size_t)m * n + i];
            J.u = mt;
            J.grams = grams;
            J.G = grams + m * RR;
            J.vinv = (double *)(A + M.vinv);
            J.flag = (int *)(A + o_flag) + 4 * i;
            J.mt = mt;
            J.rows = t->dims[m];
            J.part = (double *)(A + M.part);
            J.lam = G.lam + (size_t)i * R;
            J.U = A + M.U[m];
            J.gpart = (double *)(A + M.gpart);
            J.zero = mt;
            J.zero_n = std::max<uint64_t>(t->dims[m], t->dims[(m + 1) % N]) * R;
            J.active = active + i;
            J.nb_solve = M.nbs[m];
            J.nb_gram = M.nbg[m];
            fs[(size_t)m * (n + 1) + i + 1] = fs[(size_t)m * (n + 1) + i] + M.nbs[m];
            fg[(size_t)m * (n + 1) + i + 1] = fg[(size_t)m * (n + 1) + i] + M.nbg[m];
        }
        FitJob &F = fj[i];
        F.grams = grams;
        F.lam = G.lam + (size_t)i * R;
        F.part = (double *)(A + M.part);
        F.normX = (double *)(A + o_normX) + i;
        F.fitd = G.fitd + (size_t)i * 2 * G.maxit;
        F.cnt = G.cnt + i;
        F.active = active + i;
        F.nblk = M.nbs[N - 1];
        F.max_iters = opts[k].max_iters;
        F.epsilon = opts[k].epsilon;
    }
    // MTTKRP jobs: per mode the fiber-kernel tensors, then the decode-once
    // ones, each list with its prefix sums (n + 2 entries per mode hold both)
    for (int m = 0; m < N; ++m) {
        for (int i = 0; i < n; ++i) {
            TensorImpl *t = &tensors[G.ks[i]]->impl;
            const void *fptr[kMaxModes] = {};
            for (int q = 0; q < N; ++q) fptr[q] = A + mem[i].U[q];
            ALTO_TRY(mttkrp_group_job(t, m, G.prec, R, fptr, (double *)(A + mem[i].mt), active + i, &tmp[i], s));
        }
        MttkrpJob *dst = mj.data() + (size_t)m * n;
        unsigned *pf = fm.data() + (size_t)m * (n + 2);
        int c = 0;
        for (int d = 0; d < 2; ++d) {
            const int c0 = c;
            pf[c0 + d] = 0;
            for (int i = 0; i < n; ++i)
                if (tmp[i].dec == d) {
                    dst[c] = tmp[i];
                    pf[c + d + 1] = pf[c + d] + tmp[i].nblocks;
                    ++c;
                }
            G.nm[m][d] = c - c0;
            G.tm[m][d] = pf[c + d];
        }
    }
    // one device copy of all job tables
    size_t jo = 0;
    auto place = [&](size_t bytes) {
        const size_t o = jo;
        jo = align256(jo + bytes);
        return o;
    };
    const size_t j_bj = place(sizeof(BatchJob) * bj.size()), j_mj = place(sizeof(MttkrpJob) * mj.size());
    const size_t j_fs = place(4 * fs.size()), j_fg = place(4 * fg.size()), j_fm = place(4 * fm.size());
    const size_t j_fj = place(sizeof(FitJob) * fj.size());
    std::vector<char> hj(jo);
    memcpy(hj.data() + j_bj, bj.data(), sizeof(BatchJob) * bj.size());
    memcpy(hj.data() + j_mj, mj.data(), sizeof(MttkrpJob) * mj.size());
    memcpy(hj.data() + j_fs, fs.data(), 4 * fs.size());
    memcpy(hj.data() + j_fg, fg.data(), 4 * fg.size());
    memcpy(hj.data() + j_fm, fm.data(), 4 * fm.size());
    memcpy(hj.data() + j_fj, fj.data(), sizeof(FitJob) * fj.size());
    ALTO_TRY(G.jobs.alloc(jo));
    ALTO_HIP_TRY(hipMemcpyAsync(G.jobs.p, hj.data(), jo, hipMemcpyHostToDevice, s));
    char *J = (char *)G.jobs.p;
    for (int m = 0; m < N; ++m) {
        G.bj[m] = (const BatchJob *)(J + j_bj) + (size_t)m * n;
        G.mj[m][0] = (const MttkrpJob *)(J + j_mj) + (size_t)m * n;
        G.mj[m][1] = G.mj[m][0] + G.nm[m][0];
        G.fs[m] = (const unsigned *)(J + j_fs) + (size_t)m * (n + 1);
        G.fg[m] = (const unsigned *)(J + j_fg) + (size_t)m * (n + 1);
        G.fm[m][0] = (const unsigned *)(J + j_fm) + (size_t)m * (n + 2);
        G.fm[m][1] = G.fm[m][0] + G.nm[m][0] + 1;
        G.ts[m] = fs[(size_t)m * (n + 1) + n];
        G.tg[m] = fg[(size_t)m * (n + 1) + n];
    }
    G.fj = (const FitJob *)(J + j_fj);
    ALTO_HIP_TRY(hipStreamSynchronize(s));  // hj leaves scope

    // ---- initial model: factor storage and Grams (init_grams, gram.cpp:11-27) ----
    G.segs.clear();
    for (int i = 0; i < n; ++i) {
        const int k = G.ks[i];
        const TensorImpl *t = &tensors[k]->impl;
        for (int m = 0; m < N; ++m) G.segs.push_back(Seg{mem[i].U[m], (size_t)t->dims[m] * R, factors[k][m]});
    }
    if (G.dev_io) ALTO_TRY(upload_region_dev(G.segs, G.prec, G.Ubase, s));
    else ALTO_TRY(upload_region(G.segs, G.Ubytes, G.prec, G.Ubase, st, pool, s));
    for (int m = 0; m < N; ++m) {
        ALTO_TRY(launch_gram_group(G.bj[m], G.fg[m], n, G.tg[m], R, G.prec, s));
        ALTO_TRY(launch_gram_reduce_group(G.bj[m], n, R, s));
    }

    // ---- iterations ----
    ALTO_TRY(G.enqueue(true));
    for (int it = 1; it < G.maxit; ++it) {
        if (G.any_eps) {
            int na = 0;
            ALTO_HIP_TRY(hipMemcpyAsync(&na, G.nactive, 4, hipMemcpyDeviceToHost, s));
            ALTO_HIP_TRY(hipStreamSynchronize(s));
            if (na == 0) break;
        }
        if (!graphs_enabled()) {
            ALTO_TRY(G.enqueue(false));
            continue;
        }
        if (!G.ge) {
            CaptureScope cs;
            hipGraph_t g = nullptr;
            ALTO_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int e = G.enqueue(false);
            const hipError_t ce = hipStreamEndCapture(s, &g);
            if (e != ALTO_OK || ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                if (e != ALTO_OK) return e;
                ALTO_HIP_TRY(ce);
            }
            const hipError_t ie = hipGraphInstantiate(&G.ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            ALTO_HIP_TRY(ie);
        }
        ALTO_HIP_TRY(hipGraphLaunch(G.ge, s));
    }
    return ALTO_OK;
}

// Final models, lambdas, fits and iteration counts of one class to the host.
static int group_finish(AlsGroup &G, double *const *lambdas, double *const *fits, int *iters, Staging &st,
                        HostPool &pool) {
    const int n = (int)G.ks.size(), R = G.R;
    if (G.dev_io) ALTO_TRY(download_region_dev(G.segs, G.prec, G.Ubase, G.s));
    else ALTO_TRY(download_region(G.segs, G.Ubytes, G.prec, G.Ubase, st, pool, G.s));
    std::vector<double> lam((size_t)n * R), fd((size_t)n * 2 * G.maxit);
    std::vector<int> cnt(n);
    ALTO_HIP_TRY(hipMemcpyAsync(lam.data(), G.lam, 8 * lam.size(), hipMemcpyDeviceToHost, G.s));
    ALTO_HIP_TRY(hipMemcpyAsync(fd.data(), G.fitd, 8 * fd.size(), hipMemcpyDeviceToHost, G.s));
    ALTO_HIP_TRY(hipMemcpyAsync(cnt.data(), G.cnt, 4 * (size_t)n, hipMemcpyDeviceToHost, G.s));
    ALTO_HIP_TRY(hipStreamSynchronize(G.s));
    for (int i = 0; i < n; ++i) {
        const int k = G.ks[i];
        if (lambdas && lambdas[k]) memcpy(lambdas[k], lam.data() + (size_t)i * R, 8 * (size_t)R);
        if (iters) iters[k] = cnt[i];
        if (fits && fits[k])
            for (int it = 0; it < cnt[i]; ++it) fits[k][it] = fd[(size_t)i * 2 * G.maxit + 2 * it];
    }
    return ALTO_OK;
}

// ALTO_HIP_GROUP_SERIAL=1: one class after another (for per-kernel timing)
static bool concurrent_classes() {
    static const bool on = [] {
        const char *e = getenv("ALTO_HIP_GROUP_SERIAL");
        return !(e && e[0] == '1' && e[1] == 0);
    }();
    return on;
}

bool grouped_als_enabled() {
    static const bool on = [] {
        const char *e = getenv("ALTO_HIP_GROUPED");
        return !(e && e[0] == '0' && e[1] == 0);
    }();
    return on;
}

bool als_group_eligible(const TensorImpl *t, const alto_hip_als_opts *o) {
    return !o->verbose && group_rank_supported(o->rank) && mttkrp_group_supported(t, o->prec, o->rank);
}

int als_groups_run(const std::vector<std::vector<int>> &classes, const alto_hip_tensor *tensors,
                   const alto_hip_als_opts *opts, double *const *const *factors, double *const *lambdas,
                   double *const *fits, int *iters, bool dev_io) {
    if (classes.empty()) return ALTO_OK;
    Staging st;  // pinned staging and host converter threads: host models only
    if (!dev_io) ALTO_TRY(st.init());
    unsigned hc = std::thread::hardware_concurrency();
    HostPool pool(dev_io ? 1 : (int)std::min<unsigned>(hc ? hc : 1, 16));
    std::vector<AlsGroup> groups(classes.size());
    for (size_t g = 0; g < classes.size(); ++g) {
        groups[g].ks = classes[g];
        groups[g].dev_io = dev_io;
        ALTO_TRY(group_start(groups[g], tensors, opts, factors, st, pool));
        if (!concurrent_classes()) ALTO_HIP_TRY(hipStreamSynchronize(groups[g].s));
    }
    for (auto &G : groups) ALTO_TRY(group_finish(G, lambdas, fits, iters, st, pool));
    return ALTO_OK;
}

}  // namespace alto
