// spCP-stream on the GPU: cpstream with use_spcpstream (model SPCPSTREAM_ALTO,
// src/streaming_cpd.cpp:26-366, main.cpp:163, 397) driving
// spcpstream_alto_iter (streaming_cpd.cpp:466-1085).  Time slices are grouped
// into batches as next_dynamic_batch does (streaming_sptensor.cpp:196-264);
// a batch's stream mode has one row per time index.
//
// Per slice the update touches only the rows that occur in the slice:
//   nonzero_slices (streaming_cpd.cpp:1124-1161) -> per mode the ascending
//     occurring indices (nz rows) and the compact map; here a mark pass, an
//     exclusive scan and a compaction;
//   A_nz / A_nz_prev = nz rows of U / prev U (convert_to_rspmat,
//     rowsparse_matrix.cpp:28-54);
//   update_salto_rowsparse (alto.hpp:301-434) -> the slice as an ALTO tensor
//     over compact coordinates; every MTTKRP runs over it with the compact
//     factors (the dense ALTO MTTKRP over an output-ordered schedule);
//   per outer iteration: s_t, the temporal Gram, then per mode
//     A_nz = (MTTKRP + A_nz_prev Q) Phi^-1 and the R x R bookkeeping of
//     c = U^T U and h = prevU^T U split into nz and z parts (streaming_cpd.cpp
//     :792-970), so the z rows are never touched inside the loop;
//   z rows <- prevU (Q Phi^-1), nz rows <- A_nz (streaming_cpd.cpp:1017-1045).
// Constraints are not used (spcpstream_alto_iter only calls pseudo_inverse).
// R x R math, the solves and all factors are FP64; the MTTKRP runs in FP64 or
// FP32 (compact factors converted after each update).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "dense.hpp"
#include "sort.hpp"
#include "stream.hpp"
#include "tensor.hpp"

namespace alto {

__global__ __launch_bounds__(256) void k_sp_mark(const uint64_t *__restrict__ idx, uint64_t n,
                                                 uint32_t *__restrict__ flag) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[idx[i]] = 1u;
}

// Stream-mode index of a batch's nonzeros (time index - the batch's first)
// and the identity row list of its time rows.
__global__ __launch_bounds__(256) void k_sp_local_time(uint64_t *__restrict__ out,
                                                       const uint64_t *__restrict__ t, uint64_t n,
                                                       uint64_t t0) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = t[i] - t0;
}
__global__ __launch_bounds__(256) void k_sp_iota_u64(uint64_t *__restrict__ out, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = i;
}

__global__ __launch_bounds__(256) void k_sp_zero_u32(uint32_t *__restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// rowind[pos[i]] = i for the marked rows
__global__ __launch_bounds__(256) void k_sp_compact(const uint32_t *__restrict__ flag,
                                                    const uint32_t *__restrict__ pos, uint64_t rows,
                                                    uint64_t *__restrict__ rowind) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows && flag[i]) rowind[pos[i]] = i;
}

// compact coordinate of every slice nonzero (the ridx map)
__global__ __launch_bounds__(256) void k_sp_map(const uint64_t *__restrict__ idx, uint64_t n,
                                                const uint32_t *__restrict__ pos,
                                                uint64_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = pos[idx[i]];
}

// dst[r] = src[rowind[r]]
__global__ __launch_bounds__(256) void k_sp_gather(const double *__restrict__ src,
                                                   const uint64_t *__restrict__ rowind,
                                                   uint64_t nnzr, int R, double *__restrict__ dst) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnzr * R) return;
    const uint64_t r = e / R;
    dst[e] = src[rowind[r] * R + (e % R)];
}

// out = Hadamard of mats[m], m != mode, in mode order (mat_form_gram,
// matrix.cpp:141-156)
__global__ void k_sp_hadamard(const double *__restrict__ mats, int N, int mode, int R,
                              double *__restrict__ out) {
    const int RR = R * R;
    for (int e = threadIdx.x; e < RR; e += blockDim.x) {
        double v = 1.0;
        for (int m = 0; m < N; ++m)
            if (m != mode) v *= mats[(size_t)m * RR + e];
        out[e] = v;
    }
}

__global__ __launch_bounds__(256) void k_sp_sub(const double *__restrict__ a,
                                                const double *__restrict__ b, int n,
                                                double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}

__global__ __launch_bounds__(256) void k_sp_scale(double *__restrict__ p, double f, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] *= f;
}

__global__ void k_sp_zero1(double *p) {
    if (threadIdx.x == 0) p[0] = 0.0;
}

// Temporal Gram (streaming_cpd.cpp:744-790): G_s <- upper-triangle copy
// (copy_upper_tri, matrix.cpp:207-218); old = G_s; G_s += sum_r s_r^T s_r;
// h_s = old; c_s = G_s.  One block.
__global__ __launch_bounds__(256) void k_sp_stream_gram(double *__restrict__ Gs,
                                                        const double *__restrict__ st,
                                                        uint64_t rows, int R,
                                                        double *__restrict__ hs,
                                                        double *__restrict__ cs) {
    extern __shared__ double sym[];
    const int RR = R * R;
    for (int e = threadIdx.x; e < RR; e += blockDim.x) {
        const int a = e / R, b = e % R;
        sym[e] = a > b ? Gs[b * R + a] : Gs[e];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < RR; e += blockDim.x) {
        const int n2 = e / R, m2 = e % R;
        double v = sym[e];
        for (uint64_t r = 0; r < rows; ++r) v += st[r * R + m2] * st[r * R + n2];
        Gs[e] = v;
        cs[e] = v;
        hs[e] = sym[e];
    }
}

// A[r] = (A[r] + Ap[r] Q) V for the nnzr compact rows (rsp_mat_mul +
// rsp_mat_add + pseudo_inverse, streaming_cpd.cpp:855-876).
__global__ __launch_bounds__(256) void k_sp_rows(double *__restrict__ A, const double *__restrict__ Ap,
                                                 const double *__restrict__ Q,
                                                 const double *__restrict__ V, uint64_t nnzr, int R,
                                                 int TR) {
    extern __shared__ double lds[];
    double *sQ = lds, *sV = lds + R * R, *T = lds + 2 * R * R;
    const int RR = R * R;
    for (int e = threadIdx.x; e < RR; e += blockDim.x) {
        sQ[e] = Q[e];
        sV[e] = V[e];
    }
    const uint64_t ntiles = (nnzr + TR - 1) / TR;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t r0 = tile * TR;
        const int nr = (int)((nnzr - r0) < (uint64_t)TR ? (nnzr - r0) : TR);
        __syncthreads();
        for (int e = threadIdx.x; e < nr * R; e += blockDim.x) {
            const int r = e / R, c = e % R;
            const double *ap = Ap + (r0 + r) * R;
            double t = 0.0;
            for (int a = 0; a < R; ++a) t += ap[a] * sQ[a * R + c];
            T[e] = A[(r0 + r) * R + c] + t;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nr * R; e += blockDim.x) {
            const int r = e / R, c = e % R;
            double u = 0.0;
            for (int a = 0; a < R; ++a) u += T[r * R + a] * sV[a * R + c];
            A[(r0 + r) * R + c] = u;
        }
    }
}

// R x R bookkeeping of one mode (streaming_cpd.cpp:878-960): QPi = Q V,
// h_z = c_z_prev QPi, c_z = QPi^T h_z, c = c_nz + c_z, h = h_nz + h_z, and
// delta += sqrt(|tr c + tr c_prev - 2 tr h| / (tr c + 1e-12)).  One block.
__global__ __launch_bounds__(256) void k_sp_rr(const double *__restrict__ Q, const double *__restrict__ V,
                                               const double *__restrict__ czp, const double *__restrict__ cnz,
                                               const double *__restrict__ hnz, const double *__restrict__ cprev,
                                               int R, double *__restrict__ QPi, double *__restrict__ hz,
                                               double *__restrict__ cz, double *__restrict__ c,
                                               double *__restrict__ h, double *__restrict__ delta) {
    extern __shared__ double lds[];
    const int RR = R * R;
    double *sP = lds, *sH = lds + RR, *dc = lds + 2 * RR, *dh = dc + R;
    for (int e = threadIdx.x; e < RR; e += blockDim.x) {
        const int a = e / R, b = e % R;
        double t = 0.0;
        for (int k = 0; k < R; ++k) t += Q[a * R + k] * V[k * R + b];
        sP[e] = t;
        QPi[e] = t;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < RR; e += blockDim.x) {
        const int a = e / R, b = e % R;
        double t = 0.0;
        for (int k = 0; k < R; ++k) t += czp[a * R + k] * sP[k * R + b];
        sH[e] = t;
        hz[e] = t;
        const double hv = hnz[e] + t;
        h[e] = hv;
        if (a == b) dh[a] = hv;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < RR; e += blockDim.x) {
        const int a = e / R, b = e % R;
        double t = 0.0;
        for (int k = 0; k < R; ++k) t += sP[k * R + a] * sH[k * R + b];
        cz[e] = t;
        const double cv = cnz[e] + t;
        c[e] = cv;
        if (a == b) dc[a] = cv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double trc = 0.0, trh = 0.0, trp = 0.0;
        for (int a = 0; a < R; ++a) {
            trc += dc[a];
            trh += dh[a];
            trp += cprev[a * R + a];
        }
        delta[0] += sqrt(fabs((trc + trp - 2.0 * trh) / (trc + 1e-12)));
    }
}

// U[i] = A_nz[pos[i]] for the slice's rows, prevU[i] QPi for the others
// (streaming_cpd.cpp:1017-1045).
__global__ __launch_bounds__(256) void k_sp_final(double *__restrict__ U, const double *__restrict__ P,
                                                  const uint32_t *__restrict__ flag,
                                                  const uint32_t *__restrict__ pos,
                                                  const double *__restrict__ A,
                                                  const double *__restrict__ QPi, uint64_t rows, int R) {
    extern __shared__ double sQ[];
    for (int e = threadIdx.x; e < R * R; e += blockDim.x) sQ[e] = QPi[e];
    __syncthreads();
    const uint64_t n = rows * R;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = e / R;
        const int c = (int)(e % R);
        double v;
        if (flag[i]) {
            v = A[(uint64_t)pos[i] * R + c];
        } else {
            v = 0.0;
            for (int a = 0; a < R; ++a) v += P[i * R + a] * sQ[a * R + c];
        }
        U[e] = v;
    }
}

}  // namespace alto

using namespace alto;

extern "C" int alto_hip_spcpstream(const alto_hip_coo *coo, int s, int R, int max_iters,
                                   double epsilon, int max_slices, uint64_t nnz_threshold,
                                   uint64_t timeslice_limit, int prec,
                                   const double *const *init_factors, double *const *out_factors,
                                   uint64_t *out_dims, double *out_time_rows, int *out_outer_iters,
                                   double *out_grams, double *out_final_error, double *timings,
                                   void *stream) {
    if (!coo || !coo->dims || !coo->cidx || !coo->vals || !init_factors) return ALTO_ERR_ARG;
    const int N = coo->nmode;
    if (N < 2 || N > kMaxModes || s < 0 || s >= N || R < 1 || R > 64 || max_iters < 1)
        return ALTO_ERR_ARG;
    if (prec != ALTO_PREC_F64 && prec != ALTO_PREC_F32) return ALTO_ERR_ARG;
    const uint64_t nnz = coo->nnz;
    if (nnz >= (1ULL << 32)) return ALTO_ERR_UNSUPPORTED;
    for (int m = 0; m < N; ++m) {
        if (!init_factors[m] || coo->dims[m] == 0) return ALTO_ERR_ARG;
        if (coo->dims[m] >= (1ULL << 32)) return ALTO_ERR_UNSUPPORTED;  // 32-bit compact positions
    }
    ALTO_TRY(ensure_init());
    hipStream_t st = to_stream(stream);

    StreamInput sin;
    ALTO_TRY(stream_prepare(coo, s, &sin, st));
    const std::vector<Slice> slices = stream_batches(sin.slices, s, nnz_threshold, timeslice_limit);
    int nsl = (int)slices.size();  // batches
    if (max_slices > 0 && nsl > max_slices) nsl = max_slices;
    const uint64_t ntime = nsl ? slices[nsl - 1].t0 + slices[nsl - 1].dims[s] : 0;  // time indices processed

    // ---- device state ----
    uint64_t maxd[kMaxModes], maxslice = 1;
    for (int m = 0; m < N; ++m) {
        maxd[m] = 1;
        for (int k = 0; k < nsl; ++k) maxd[m] = std::max(maxd[m], slices[k].dims[m]);
    }
    for (int k = 0; k < nsl; ++k) maxslice = std::max(maxslice, slices[k].end - slices[k].start);
    const uint64_t used = nsl ? slices[nsl - 1].end : 0;
    const size_t RR = (size_t)R * R;
    DevBuf *dci = sin.cid;
    DevBuf &dvals = sin.vals;
    DevBuf zidx, U[kMaxModes], P[kMaxModes], A[kMaxModes], Ap[kMaxModes],
        Af[kMaxModes], flag[kMaxModes], pos[kMaxModes], rowind[kMaxModes], cco[kMaxModes];
    DevBuf G, c, h, cnz, hnz, cz, hz, cprev, czprev, cnzprev, Q, QPi, vinv, gpart, part, colp, lam,
        delta, iflag, scan_tmp, trows;
    for (int m = 0; m < N; ++m) {
        if (m == s) continue;
        ALTO_TRY(flag[m].alloc(sizeof(uint32_t) * maxd[m]));
        ALTO_TRY(pos[m].alloc(sizeof(uint32_t) * maxd[m]));
        ALTO_TRY(rowind[m].alloc(sizeof(uint64_t) * maxd[m]));
        ALTO_TRY(cco[m].alloc(sizeof(uint64_t) * maxslice));
    }
    DevBuf sidx;
    ALTO_TRY(zidx.alloc(sizeof(uint64_t) * maxslice));
    ALTO_TRY(sidx.alloc(sizeof(uint64_t) * maxd[s]));
    k_sp_iota_u64<<<cdiv(maxd[s], 256), 256, 0, st>>>(sidx.as<uint64_t>(), maxd[s]);
    for (int m = 0; m < N; ++m) {
        ALTO_TRY(U[m].alloc(sizeof(double) * maxd[m] * R));
        ALTO_TRY(P[m].alloc(sizeof(double) * maxd[m] * R));
        ALTO_TRY(A[m].alloc(sizeof(double) * maxd[m] * R));
        ALTO_TRY(Ap[m].alloc(sizeof(double) * maxd[m] * R));
        if (prec == ALTO_PREC_F32) ALTO_TRY(Af[m].alloc(sizeof(float) * maxd[m] * R));
    }
    for (DevBuf *b : {&G, &c, &h, &cnz, &hnz, &cz, &hz, &cprev, &czprev, &cnzprev, &Q, &QPi}) {
        ALTO_TRY(b->alloc(sizeof(double) * N * RR));
        ALTO_TRY(launch_fill_zero(b->as<double>(), N * RR, st));  // InitSparseCPGrams
    }
    ALTO_TRY(vinv.alloc(sizeof(double) * RR));
    ALTO_TRY(gpart.alloc(sizeof(double) * kMaxBlocks * RR));
    ALTO_TRY(part.alloc(sizeof(double) * 3 * R * kMaxBlocks));
    ALTO_TRY(colp.alloc(sizeof(double) * 3 * R * kMaxBlocks));
    ALTO_TRY(lam.alloc(sizeof(double) * R));
    ALTO_TRY(delta.alloc(sizeof(double)));
    ALTO_TRY(iflag.alloc(sizeof(int) * 4));
    ALTO_TRY(trows.alloc(sizeof(double) * std::max<uint64_t>(ntime, 1) * R));
    ALTO_TRY(launch_fill_zero(trows.as<double>(), std::max<uint64_t>(ntime, 1) * R, st));
    const int TR = std::max(1, 256 / R);
    const size_t lds_rows = sizeof(double) * (2 * RR + (size_t)TR * R);
    const size_t lds_rr = sizeof(double) * (2 * RR + 2 * (size_t)R);
    {
        static std::once_flag attr;
        std::call_once(attr, [] {
            (void)hipFuncSetAttribute((const void *)k_sp_rows,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256);
            (void)hipFuncSetAttribute((const void *)k_sp_rr,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256);
        });
    }
    TensorImpl at;
    at.keep_sched_scratch = true;  // rebuilt for every slice
    EventSet evs;
    hipEvent_t ev[4];
    for (auto &e : ev) ALTO_TRY(evs.make(&e));
    double t_build = 0;
    ALTO_HIP_TRY(hipEventRecord(ev[0], st));

    auto mat = [&](DevBuf &b, int m) { return b.as<double>() + (size_t)m * RR; };
    auto atb = [&](const double *X, const double *Y, uint64_t rows, double *out) -> int {
        const int nb = row_blocks(rows);
        ALTO_TRY(launch_cross_gram(X, Y, rows, R, gpart.as<double>(), nb, st));
        return launch_gram_reduce(gpart.as<double>(), nb, R, out, st);
    };
    uint64_t cur[kMaxModes] = {0}, nnzr[kMaxModes];
    const void *fptr[kMaxModes];
    for (int m = 0; m < N; ++m) fptr[m] = prec == ALTO_PREC_F32 ? Af[m].p : A[m].p;
    auto refresh_f32 = [&](int m) -> int {
        if (prec == ALTO_PREC_F32)
            return alto_hip_convert(Af[m].p, ALTO_PREC_F32, A[m].p, ALTO_PREC_F64, nnzr[m] * R, st);
        return ALTO_OK;
    };
    std::vector<int> outer_v(std::max(nsl, 1), 0);

    IterGraph ig(st);  // the outer-iteration graph, reused from slice to slice
    for (int k = 0; k < nsl; ++k) {
        const Slice &sl = slices[k];
        const uint64_t *bd = sl.dims;
        const uint64_t bn = sl.end - sl.start;
        // ---- model creation / growth (streaming_cpd.cpp:106-128) ----
        if (k == 0) {
            for (int m = 0; m < N; ++m) {
                ALTO_HIP_TRY(hipMemcpyAsync(U[m].p, init_factors[m], sizeof(double) * bd[m] * R,
                                            hipMemcpyHostToDevice, st));
                ALTO_TRY(launch_fill_zero(P[m].as<double>(), maxd[m] * R, st));
            }
            ALTO_TRY(launch_fill_zero(U[s].as<double>(), bd[s] * R, st));
            ALTO_TRY(launch_fill_zero(mat(G, s), RR, st));
            // first-slice normalisation (streaming_cpd.cpp:585-598): 2-norm
            // columns of the non-stream modes, Grams of the normalised factors
            for (int m = 0; m < N; ++m) {
                if (m == s) continue;
                const int nb = row_blocks(bd[m]);
                ALTO_TRY(launch_colsumsq(U[m].as<double>(), bd[m], R, colp.as<double>(), nb, st));
                ALTO_TRY(launch_lambda(colp.as<double>(), nb, R, 0, lam.as<double>(), st));
                ALTO_TRY(launch_scale_gram(U[m].as<double>(), lam.as<double>(), bd[m], R, U[m].p,
                                           ALTO_PREC_F64, gpart.as<double>(), nb, st));
                ALTO_TRY(launch_gram_reduce(gpart.as<double>(), nb, R, mat(G, m), st));
            }
        } else {
            for (int m = 0; m < N; ++m)
                if (bd[m] > cur[m]) {
                    ALTO_HIP_TRY(hipMemcpyAsync(U[m].as<double>() + cur[m] * R, init_factors[m] + cur[m] * R,
                                                sizeof(double) * (bd[m] - cur[m]) * R,
                                                hipMemcpyHostToDevice, st));
                    ALTO_TRY(launch_fill_zero(P[m].as<double>() + cur[m] * R, (bd[m] - cur[m]) * R, st));
                }
            for (int m = 0; m < N; ++m)
                if (m != s) ALTO_TRY(atb(U[m].as<double>(), U[m].as<double>(), bd[m], mat(G, m)));
        }
        for (int m = 0; m < N; ++m) cur[m] = bd[m];

        // ---- nonzero_slices: occurring rows, compact map, compact coordinates ----
        ALTO_HIP_TRY(hipEventRecord(ev[2], st));
        std::vector<uint32_t> lastf(N, 0), lastp(N, 0);
        for (int m = 0; m < N; ++m) {
            if (m == s) continue;
            const uint64_t *ci = dci[m].as<uint64_t>() + sl.start;
            k_sp_zero_u32<<<cdiv(bd[m], 256), 256, 0, st>>>(flag[m].as<uint32_t>(), bd[m]);
            k_sp_mark<<<cdiv(bn, 256), 256, 0, st>>>(ci, bn, flag[m].as<uint32_t>());
            ALTO_LAUNCH_CHECK();
            ALTO_HIP_TRY(hipMemcpyAsync(pos[m].p, flag[m].p, sizeof(uint32_t) * bd[m],
                                        hipMemcpyDeviceToDevice, st));
            ALTO_TRY(exclusive_scan_u32(pos[m].as<uint32_t>(), bd[m], scan_tmp, st));
            ALTO_HIP_TRY(hipMemcpyAsync(&lastf[m], flag[m].as<uint32_t>() + bd[m] - 1, sizeof(uint32_t),
                                        hipMemcpyDeviceToHost, st));
            ALTO_HIP_TRY(hipMemcpyAsync(&lastp[m], pos[m].as<uint32_t>() + bd[m] - 1, sizeof(uint32_t),
                                        hipMemcpyDeviceToHost, st));
            k_sp_compact<<<cdiv(bd[m], 256), 256, 0, st>>>(flag[m].as<uint32_t>(), pos[m].as<uint32_t>(),
                                                           bd[m], rowind[m].as<uint64_t>());
            k_sp_map<<<cdiv(bn, 256), 256, 0, st>>>(ci, bn, pos[m].as<uint32_t>(), cco[m].as<uint64_t>());
            ALTO_LAUNCH_CHECK();
        }
        ALTO_HIP_TRY(hipStreamSynchronize(st));
        for (int m = 0; m < N; ++m) nnzr[m] = m == s ? bd[s] : (uint64_t)lastp[m] + lastf[m];
        // the batch over compact coordinates (update_salto_rowsparse); every
        // time index of a batch has nonzeros, so its stream rows are all used
        k_sp_local_time<<<cdiv(bn, 256), 256, 0, st>>>(zidx.as<uint64_t>(), dci[s].as<uint64_t>() + sl.start,
                                                       bn, sl.t0);
        ALTO_LAUNCH_CHECK();
        {
            const uint64_t *cp[kMaxModes];
            for (int m = 0; m < N; ++m) cp[m] = m == s ? zidx.as<uint64_t>() : cco[m].as<uint64_t>();
            alto_hip_coo sc{N, nnzr, bn, cp, dvals.as<double>() + sl.start, 1};
            ALTO_TRY(build_alto(&at, &sc, 0, 0, false, st, false));  // only the schedules are read
        }
        ALTO_HIP_TRY(hipEventRecord(ev[3], st));
        // ---- row-sparse matrices and the c / h Grams (streaming_cpd.cpp:607-656) ----
        for (int m = 0; m < N; ++m) {
            const uint64_t *ri = m == s ? sidx.as<uint64_t>() : rowind[m].as<uint64_t>();
            k_sp_gather<<<cdiv(nnzr[m] * R, 256), 256, 0, st>>>(P[m].as<double>(), ri, nnzr[m], R,
                                                                 Ap[m].as<double>());
            k_sp_gather<<<cdiv(nnzr[m] * R, 256), 256, 0, st>>>(U[m].as<double>(), ri, nnzr[m], R,
                                                                 A[m].as<double>());
            ALTO_LAUNCH_CHECK();
            if (m == s) continue;  // c_s and h_s are set before their first use
            ALTO_HIP_TRY(hipMemcpyAsync(mat(c, m), mat(G, m), sizeof(double) * RR, hipMemcpyDeviceToDevice, st));
            ALTO_TRY(atb(P[m].as<double>(), U[m].as<double>(), bd[m], mat(h, m)));
            if (k > 0) {
                ALTO_TRY(atb(Ap[m].as<double>(), Ap[m].as<double>(), nnzr[m], mat(cnzprev, m)));
                k_sp_sub<<<cdiv(RR, 256), 256, 0, st>>>(mat(cprev, m), mat(cnzprev, m), (int)RR,
                                                        mat(czprev, m));
                ALTO_LAUNCH_CHECK();
            }
        }
        for (int m = 0; m < N; ++m) ALTO_TRY(refresh_f32(m));

        // ---- outer iterations (streaming_cpd.cpp:712-1000) ----
        double prev_delta = 0.0;
        int outer = 0;
        // one outer iteration, no host syncs: replayed as a graph from the
        // second iteration on
        auto outer_iter = [&]() -> int {
            k_sp_zero1<<<1, 64, 0, st>>>(delta.as<double>());
            // s_t: MTTKRP of the stream mode, solve with Had_{m != s} c_m + 1e-12 I
            ALTO_TRY(launch_fill_zero(A[s].as<double>(), nnzr[s] * R, st));
            ALTO_TRY(run_mttkrp(&at, s, prec, R, fptr, A[s].p, 1, 2, 0, -1, st, false, true));
            ALTO_TRY(launch_chol_inv(c.as<double>(), N, s, R, 1e-12, false, vinv.as<double>(),
                                     iflag.as<int>(), nullptr, st, 1, nullptr));
            ALTO_TRY(launch_solve(A[s].as<double>(), nnzr[s], R, vinv.as<double>(), part.as<double>(),
                                  1, st));
            k_sp_stream_gram<<<1, 256, sizeof(double) * RR, st>>>(mat(G, s), A[s].as<double>(), nnzr[s], R,
                                                                  mat(h, s), mat(c, s));
            ALTO_LAUNCH_CHECK();
            ALTO_TRY(refresh_f32(s));
            for (int m = 0; m < N; ++m) {
                if (m == s) continue;
                const uint64_t n = nnzr[m];
                k_sp_hadamard<<<1, 256, 0, st>>>(h.as<double>(), N, m, R, mat(Q, m));
                ALTO_TRY(launch_chol_inv(c.as<double>(), N, m, R, 1e-12, false, vinv.as<double>(),
                                         iflag.as<int>(), nullptr, st, 1, nullptr));
                ALTO_TRY(launch_fill_zero(A[m].as<double>(), n * R, st));
                ALTO_TRY(run_mttkrp(&at, m, prec, R, fptr, A[m].p, 1, 2, 0, -1, st, false, true));
                const unsigned nb = (unsigned)std::min<uint64_t>(cdiv(n, TR), 2048);
                k_sp_rows<<<nb, 256, lds_rows, st>>>(A[m].as<double>(), Ap[m].as<double>(), mat(Q, m),
                                                     vinv.as<double>(), n, R, TR);
                ALTO_LAUNCH_CHECK();
                ALTO_TRY(atb(A[m].as<double>(), A[m].as<double>(), n, mat(cnz, m)));
                ALTO_TRY(atb(Ap[m].as<double>(), A[m].as<double>(), n, mat(hnz, m)));
                k_sp_rr<<<1, 256, lds_rr, st>>>(mat(Q, m), vinv.as<double>(), mat(czprev, m), mat(cnz, m),
                                                mat(hnz, m), mat(cprev, m), R, mat(QPi, m), mat(hz, m),
                                                mat(cz, m), mat(c, m), mat(h, m), delta.as<double>());
                ALTO_LAUNCH_CHECK();
                ALTO_TRY(refresh_f32(m));
            }
            return ALTO_OK;
        };
        for (int i = 0; i < max_iters; ++i) {
            ALTO_TRY(ig.run(i, outer_iter));
            double hd = 0.0;
            ALTO_HIP_TRY(hipMemcpyAsync(&hd, delta.p, sizeof(double), hipMemcpyDeviceToHost, st));
            ALTO_HIP_TRY(hipStreamSynchronize(st));
            float bms = 0;
            (void)hipEventElapsedTime(&bms, ev[2], ev[3]);
            if (i == 0) t_build += bms * 1e-3;
            outer = i + 1;
            if (i > 0 && fabs(prev_delta - hd) < epsilon) break;
            prev_delta = hd;
        }
        // ---- z rows, nz rows, forgetting, previous Grams (1017-1058) ----
        for (int m = 0; m < N; ++m) {
            if (m == s) {
                ALTO_HIP_TRY(hipMemcpyAsync(U[s].p, A[s].p, sizeof(double) * bd[s] * R, hipMemcpyDeviceToDevice,
                                            st));
                continue;
            }
            const uint64_t n = bd[m] * R;
            const unsigned nb = (unsigned)std::min<uint64_t>(cdiv(n, 256), 1024);
            k_sp_final<<<nb, 256, sizeof(double) * RR, st>>>(U[m].as<double>(), P[m].as<double>(),
                                                             flag[m].as<uint32_t>(), pos[m].as<uint32_t>(),
                                                             A[m].as<double>(), mat(QPi, m), bd[m], R);
            ALTO_LAUNCH_CHECK();
        }
        k_sp_scale<<<cdiv(RR, 256), 256, 0, st>>>(mat(G, s), 0.99, (int)RR);
        ALTO_LAUNCH_CHECK();
        ALTO_HIP_TRY(hipMemcpyAsync(cprev.p, c.p, sizeof(double) * N * RR, hipMemcpyDeviceToDevice, st));
        ALTO_HIP_TRY(hipMemcpyAsync(cnzprev.p, cnz.p, sizeof(double) * N * RR, hipMemcpyDeviceToDevice, st));
        for (int m = 0; m < N; ++m)
            ALTO_HIP_TRY(hipMemcpyAsync(P[m].p, U[m].p, sizeof(double) * bd[m] * R, hipMemcpyDeviceToDevice, st));
        if (out_time_rows)
            ALTO_HIP_TRY(hipMemcpyAsync(out_time_rows + (size_t)sl.t0 * R, U[s].p, sizeof(double) * bd[s] * R,
                                        hipMemcpyDeviceToHost, st));
        ALTO_HIP_TRY(hipMemcpyAsync(trows.as<double>() + (size_t)sl.t0 * R, U[s].p, sizeof(double) * bd[s] * R,
                                    hipMemcpyDeviceToDevice, st));
        outer_v[k] = outer;
    }
    ALTO_HIP_TRY(hipEventRecord(ev[1], st));
    ALTO_HIP_TRY(hipEventSynchronize(ev[1]));
    float tot_ms = 0;
    (void)hipEventElapsedTime(&tot_ms, ev[0], ev[1]);
    if (out_outer_iters)
        for (int k = 0; k < nsl; ++k) out_outer_iters[k] = outer_v[k];
    if (out_final_error) {
        // cpd_error over the processed nonzeros (relabelled), time factor = s_t rows
        DevBuf &tidx = sin.cid[s];
        const uint64_t *cp[kMaxModes];
        const double *fp[kMaxModes];
        uint64_t rows[kMaxModes];
        for (int m = 0; m < N; ++m) {
            cp[m] = m == s ? tidx.as<uint64_t>() : dci[m].as<uint64_t>();
            fp[m] = m == s ? trows.as<double>() : U[m].as<double>();
            rows[m] = m == s ? std::max<uint64_t>(ntime, 1) : std::max<uint64_t>(cur[m], 1);
        }
        double eo[4];
        ALTO_TRY(cpd_error_dev(N, rows, used, cp, dvals.as<double>(), fp, R, st, eo));
        *out_final_error = eo[0];
    }
    for (int m = 0; m < N; ++m) {
        if (out_dims) out_dims[m] = cur[m];
        if (out_factors && out_factors[m] && cur[m])
            ALTO_HIP_TRY(hipMemcpyAsync(out_factors[m], U[m].p, sizeof(double) * cur[m] * R,
                                        hipMemcpyDeviceToHost, st));
    }
    if (out_grams)
        ALTO_HIP_TRY(hipMemcpyAsync(out_grams, G.p, sizeof(double) * N * RR, hipMemcpyDeviceToHost, st));
    ALTO_HIP_TRY(hipStreamSynchronize(st));
    if (timings) {
        timings[0] = tot_ms * 1e-3;
        timings[1] = t_build;
        timings[2] = 0.0;
        timings[3] = (double)ntime;
    }
    return nsl;
}
