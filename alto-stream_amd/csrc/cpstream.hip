// CP-stream on the GPU: cpstream (src/streaming_cpd.cpp:26-366) with the
// CPSTREAM_ALTO model, driving cpstream_alto_iter (streaming_cpd.cpp:
// 1198-1519) and admm_cf_ (src/admm.cpp:216-429).  Time slices are grouped
// into batches as next_dynamic_batch does (streaming_sptensor.cpp:196-264):
// consecutive time indices until timeslice_limit of them or at least
// nnz_threshold nonzeros; a batch's stream mode has one row per time index.
//
// Host side (C++): StreamingSparseTensor (streaming_sptensor.cpp:69-264):
// stable lexicographic sort by (s, s+1, ...) mod N, first-appearance
// relabelling of every non-stream mode, slice boundaries and running-max
// slice dims.  Device side, per slice: ALTO rebuild of the slice into
// reused buffers, stream-mode MTTKRP + FP64 solve, temporal Gram update,
// per-mode MTTKRP + historical term + non-negative AO-ADMM (FP64 primal,
// dual and solve), Gram refresh and the convergence quantity; the host reads
// one scalar per outer iteration.  Factor rows that appear when the model
// grows come from caller-supplied initial matrices (the reference fills
// them with rand(), kruskal_model.cpp:101-133).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <vector>

#include "dense.hpp"
#include "stream.hpp"
#include "tensor.hpp"

namespace alto {

static inline int pow2(int r) {
    int p = 1;
    while (p < r) p <<= 1;
    return p;
}

// Row tiles for the ADMM kernels: 256 threads = (256 / RP) row lanes x RP
// columns, 4 rows per lane per tile.
struct TileShape {
    int RP, lanes, TR;
};
static TileShape tile_shape(int R) {
    TileShape t;
    t.RP = pow2(R);
    t.lanes = 256 / t.RP;
    t.TR = t.lanes * 4;
    return t;
}

__device__ inline double block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// First ADMM pass (admm.cpp:282-325) fused with the historical term
// (streaming_cpd.cpp:1411-1412: mttkrp += prevU_j * ata):
//   mt += P * ata;  rhs = mt + rho (x + d)  (unconstrained: rhs = mt)
//   aux = rhs * vinv;  new primal = max(aux - d, 0)
// Writes mt and aux (non-negative) or the solved factor U (unconstrained),
// and the per-block sum of (new primal - x)^2.
__global__ __launch_bounds__(256) void k_admm_first(double *__restrict__ mt,
                                                    const double *__restrict__ P,
                                                    const double *__restrict__ ata,
                                                    double *__restrict__ U,
                                                    const double *__restrict__ dual,
                                                    const double *__restrict__ vinv,
                                                    const double *__restrict__ rho_p,
                                                    uint64_t rows, int R, int RP, int nonneg,
                                                    double *__restrict__ aux,
                                                    double *__restrict__ part) {
    extern __shared__ double lds[];
    const int lanes = 256 / RP, TR = lanes * 4;
    double *Wa = lds, *Wv = lds + R * R, *T = Wv + R * R, *T2 = T + TR * R;
    __shared__ double red[4];
    const int t = threadIdx.x, c = t % RP, rl = t / RP;
    for (int e = t; e < R * R; e += 256) {
        Wa[e] = ata[e];
        Wv[e] = vinv[e];
    }
    const double rho = nonneg ? *rho_p : 0.0;
    double dres = 0.0;
    uint64_t ntiles = (rows + TR - 1) / TR;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint64_t r0 = tile * TR;
        int nr = (int)((rows - r0) < (uint64_t)TR ? (rows - r0) : TR);
        __syncthreads();
        for (int e = t; e < nr * R; e += 256) T[e] = P[r0 * R + e];
        __syncthreads();
        if (c < R)
            for (int q = 0; q < 4; ++q) {
                int rr = rl + q * lanes;
                if (rr < nr) {
                    uint64_t g = (r0 + rr) * R + c;
                    double h = 0.0;
                    for (int k = 0; k < R; ++k) h += T[rr * R + k] * Wa[k * R + c];
                    double m = mt[g] + h;
                    mt[g] = m;
                    T2[rr * R + c] = nonneg ? m + rho * (U[g] + dual[g]) : m;
                }
            }
        __syncthreads();
        if (c < R)
            for (int q = 0; q < 4; ++q) {
                int rr = rl + q * lanes;
                if (rr < nr) {
                    uint64_t g = (r0 + rr) * R + c;
                    double a = 0.0;
                    for (int k = 0; k < R; ++k) a += T2[rr * R + k] * Wv[k * R + c];
                    if (nonneg) {
                        aux[g] = a;
                        double x = a - dual[g];
                        x = x < 0.0 ? 0.0 : x;
                        double dd = x - U[g];
                        dres += dd * dd;
                    } else {
                        U[g] = a;
                    }
                }
            }
    }
    double tot = block_sum(dres, red);
    if (t == 0) part[blockIdx.x] = tot;
}

// Cross-workgroup traffic of the ADMM loop goes only through device-scope
// atomic accesses (stores and loads carry the device coherence scope, so a
// store is acknowledged once it is visible device-wide); the wait for that
// acknowledgement orders it before the arrival count.  No cache writeback or
// invalidation is needed: each workgroup's rows are its own.
__device__ __forceinline__ void store_dev(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_dev(unsigned *p, unsigned v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_dev(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned load_dev(const unsigned *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void wait_stores() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Sums over workgroups of the three device-scope partials (wave 0; at most
// 1024 workgroups): lane l adds the partials of workgroups l, l + 64, ...
// in that order, loading four workgroups' partials at a time, then the
// lanes' sums are combined by a shuffle tree.
__device__ void sum_partials(const double *part, unsigned nblk, double *out) {
    const int lane = threadIdx.x & 63;
    double a[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j0 = 0; j0 < 16; j0 += 4) {
        if ((unsigned)(64 * j0) >= nblk) break;  // uniform
        double v[4][3];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const unsigned b = lane + 64u * (j0 + j);
#pragma unroll
            for (int k = 0; k < 3; ++k) v[j][k] = b < nblk ? load_dev(part + (size_t)b * 3 + k) : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) a[k] += v[j][k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double x = a[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        out[k] = x;
    }
}

// The AO-ADMM loop (admm.cpp:331-413) of one factor in one launch: at most
// max_iter iterations on the state (aux, dual), each
//   x = max(aux - d, 0) (the current primal); y = x - aux; d += y;
//   aux = (mt + rho (x + d)) * vinv,
// with p_norm, d_norm and p_res summed per workgroup.  The convergence test
// (float tolerance 1e-6f; d_res is the first-pass value at iteration 0 and 0
// afterwards: the reference sets init = primal in every later prox step) is
// a grid-wide decision: every workgroup publishes its partials and arrives
// at bar[0]; the last to arrive sums the partials in workgroup order,
// counts the iteration, stores the decision in bar[2] and advances the
// generation bar[1], which the others wait for.  Then U = max(aux - dual,
// 0) (the new primal).  The grid must be resident at once (the launch sizes
// it from the occupancy); a wait longer than `limit` wall-clock ticks sets
// *err and ends the kernel (its results are then not used).
//
// CT = 0: the loop alone, after k_admm_first, streaming aux, dual and mt
// from memory every iteration.  CT > 0 (the launch guarantees that every
// workgroup owns at most CT row tiles): the first pass of k_admm_first and
// the loop in one launch, on chip: a workgroup keeps its rows of
// mt + P * ata in LDS (after Wv and T2: CT tiles of TR x R, then ata and a
// staging tile of P) and a thread its elements of aux and dual in registers;
// P, mt, U and dual are read once, dual and U written once at the end (mt and
// aux are scratch), instead of writing two and reading three arrays in the
// first pass and every iteration.  The first pass's residual partial goes to
// part0 through device-scope stores, read by the workgroup that decides the
// first iteration.  The operations and the order of every sum are those of
// k_admm_first and the streamed loop.
constexpr int kAdmmCacheTiles = 7;

// part0 summed in the order of wave_reduce_entry (four strided accumulators,
// then the lanes), with device-scope loads: the partials were stored by other
// workgroups of the same launch.
__device__ double sum_part0_dev(const double *part0, int nblk) {
    const int lane = threadIdx.x & 63;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int b = lane;
    for (; b + 192 < nblk; b += 256) {
        const double v0 = load_dev(part0 + b), v1 = load_dev(part0 + b + 64);
        const double v2 = load_dev(part0 + b + 128), v3 = load_dev(part0 + b + 192);
        a0 += v0;
        a1 += v1;
        a2 += v2;
        a3 += v3;
    }
    for (; b < nblk; b += 64) a0 += load_dev(part0 + b);
    double a = (a0 + a1) + (a2 + a3);
    for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
    return a;
}

template <int CT>
__global__ __launch_bounds__(256) void k_admm_loop(const double *__restrict__ mt,
                                                   double *__restrict__ aux,
                                                   double *__restrict__ dual,
                                                   const double *__restrict__ vinv,
                                                   const double *__restrict__ rho_p,
                                                   double *__restrict__ U, uint64_t rows, int R, int RP,
                                                   double *part, double *part0,
                                                   unsigned *bar, int *err, long long limit,
                                                   int max_iter, double *iters,
                                                   const double *__restrict__ P,
                                                   const double *__restrict__ ata) {
    extern __shared__ double lds[];
    const int lanes = 256 / RP, TR = lanes * 4;
    double *Wv = lds, *T2 = lds + R * R;
    __shared__ double red[4];
    __shared__ int conv, last;
    const int t = threadIdx.x, c = t % RP, rl = t / RP;
    const unsigned nblk = gridDim.x;
    for (int e = t; e < R * R; e += 256) Wv[e] = vinv[e];
    const double rho = *rho_p;
    const uint64_t ntiles = (rows + TR - 1) / TR;
    constexpr bool cached = CT > 0;
    constexpr int NC = CT > 0 ? CT : 1;
    double ca[NC][4], cd[NC][4];
    double *Mt = T2 + TR * R;  // on-chip form: tile i's rows of mt + P ata at Mt + i TR R
    if constexpr (cached) {
        // the first pass (k_admm_first, non-negative), tile by tile
        double *Wa = Mt + CT * TR * R, *T = Wa + R * R;
        for (int e = t; e < R * R; e += 256) Wa[e] = ata[e];
        double dres = 0.0;
#pragma unroll
        for (int i = 0; i < CT; ++i) {
#pragma unroll
            for (int q = 0; q < 4; ++q) ca[i][q] = cd[i][q] = 0.0;
            const uint64_t tile = blockIdx.x + (uint64_t)i * nblk;
            if (tile >= ntiles) continue;  // uniform in the workgroup
            const uint64_t r0 = tile * TR;
            const int nr = (int)((rows - r0) < (uint64_t)TR ? (rows - r0) : TR);
            __syncthreads();
            for (int e = t; e < nr * R; e += 256) T[e] = P[r0 * R + e];
            __syncthreads();
            double uu[4] = {0.0, 0.0, 0.0, 0.0};
            if (c < R)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int rr = rl + q * lanes;
                    if (rr < nr) {
                        const uint64_t g = (r0 + rr) * R + c;
                        double h = 0.0;
                        for (int k = 0; k < R; ++k) h += T[rr * R + k] * Wa[k * R + c];
                        const double m = mt[g] + h;
                        Mt[i * TR * R + rr * R + c] = m;
                        uu[q] = U[g];
                        cd[i][q] = dual[g];
                        T2[rr * R + c] = m + rho * (uu[q] + cd[i][q]);
                    }
                }
            __syncthreads();
            if (c < R)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int rr = rl + q * lanes;
                    if (rr < nr) {
                        double a = 0.0;
                        for (int k = 0; k < R; ++k) a += T2[rr * R + k] * Wv[k * R + c];
                        ca[i][q] = a;
                        double x = a - cd[i][q];
                        x = x < 0.0 ? 0.0 : x;
                        const double dd = x - uu[q];
                        dres += dd * dd;
                    }
                }
        }
        const double tot = block_sum(dres, red);
        if (t == 0) store_dev(part0 + blockIdx.x, tot);  // ordered before the first arrival
    }
    for (int it = 0; it < max_iter; ++it) {
        double pn = 0.0, dn = 0.0, pr = 0.0;
        if constexpr (cached) {
#pragma unroll
            for (int i = 0; i < CT; ++i) {
                const uint64_t tile = blockIdx.x + (uint64_t)i * nblk;
                if (tile >= ntiles) break;  // uniform in the workgroup
                const uint64_t r0 = tile * TR;
                const int nr = (int)((rows - r0) < (uint64_t)TR ? (rows - r0) : TR);
                __syncthreads();
                if (c < R)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int rr = rl + q * lanes;
                        if (rr < nr) {
                            const double a = ca[i][q], d = cd[i][q];
                            double x = a - d;
                            x = x < 0.0 ? 0.0 : x;
                            pn += x * x;
                            const double y = x - a;
                            const double di = d + y;
                            dn += di * di;
                            pr += y * y;
                            cd[i][q] = di;
                            T2[rr * R + c] = Mt[i * TR * R + rr * R + c] + rho * (x + di);
                        }
                    }
                __syncthreads();
                if (c < R)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int rr = rl + q * lanes;
                        if (rr < nr) {
                            double a = 0.0;
                            for (int k = 0; k < R; ++k) a += T2[rr * R + k] * Wv[k * R + c];
                            ca[i][q] = a;
                        }
                    }
            }
        }
        for (uint64_t tile = blockIdx.x; !cached && tile < ntiles; tile += nblk) {
            uint64_t r0 = tile * TR;
            int nr = (int)((rows - r0) < (uint64_t)TR ? (rows - r0) : TR);
            __syncthreads();
            double di_keep[4];
            if (c < R)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    int rr = rl + q * lanes;
                    if (rr < nr) {
                        uint64_t g = (r0 + rr) * R + c;
                        double a = aux[g], d = dual[g];
                        double x = a - d;
                        x = x < 0.0 ? 0.0 : x;
                        pn += x * x;
                        double y = x - a;
                        double di = d + y;
                        dn += di * di;
                        pr += y * y;
                        di_keep[q] = di;
                        T2[rr * R + c] = mt[g] + rho * (x + di);
                    }
                }
            __syncthreads();
            if (c < R)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    int rr = rl + q * lanes;
                    if (rr < nr) {
                        uint64_t g = (r0 + rr) * R + c;
                        double a = 0.0;
                        for (int k = 0; k < R; ++k) a += T2[rr * R + k] * Wv[k * R + c];
                        aux[g] = a;
                        dual[g] = di_keep[q];
                    }
                }
        }
        double s0 = block_sum(pn, red);
        double s1 = block_sum(dn, red);
        double s2 = block_sum(pr, red);
        unsigned gen = 0;
        if (t == 0) {
            store_dev(part + blockIdx.x * 3 + 0, s0);
            store_dev(part + blockIdx.x * 3 + 1, s1);
            store_dev(part + blockIdx.x * 3 + 2, s2);
            gen = load_dev(bar + 1);
            wait_stores();
            last = atomicAdd(bar, 1u) == nblk - 1;
        }
        __syncthreads();
        if (last) {
            if (t < 64) {
                double sums[3];
                sum_partials(part, nblk, sums);
                const double spn = sums[0], sdn = sums[1], spr = sums[2];
                const double sdr =
                    it != 0 ? 0.0 : cached ? sum_part0_dev(part0, nblk) : wave_reduce_entry(part0, nblk, 1, 0, 0);
                if (t == 0) {
                    const double tol = (double)1e-6f;
                    conv = spr <= tol * spn && sdr <= tol * sdn;
                    atomicAdd(iters, 1.0);
                    store_dev(bar + 2, (unsigned)conv);
                    store_dev(bar, 0u);
                    wait_stores();
                    atomicAdd(bar + 1, 1u);
                }
            }
        } else if (t == 0) {
            const long long t0 = wall_clock64();
            int ok = 1;
            while (load_dev(bar + 1) == gen) {
                if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                    wall_clock64() - t0 > limit) {
                    atomicOr(err, 1);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            conv = ok ? (int)load_dev(bar + 2) : -1;
        }
        __syncthreads();
        if (conv < 0) return;
        if (conv) break;
    }
    if constexpr (cached) {
#pragma unroll
        for (int i = 0; i < CT; ++i) {
            const uint64_t tile = blockIdx.x + (uint64_t)i * nblk;
            if (tile >= ntiles) break;
            const uint64_t r0 = tile * TR;
            const int nr = (int)((rows - r0) < (uint64_t)TR ? (rows - r0) : TR);
            if (c < R)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int rr = rl + q * lanes;
                    if (rr < nr) {
                        const uint64_t g = (r0 + rr) * R + c;
                        const double x = ca[i][q] - cd[i][q];
                        U[g] = x > 0.0 ? x : 0.0;
                        dual[g] = cd[i][q];
                    }
                }
        }
        return;
    }
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += nblk) {
        uint64_t r0 = tile * TR;
        int nr = (int)((rows - r0) < (uint64_t)TR ? (rows - r0) : TR);
        if (c < R)
            for (int q = 0; q < 4; ++q) {
                int rr = rl + q * lanes;
                if (rr < nr) {
                    uint64_t g = (r0 + rr) * R + c;
                    double x = aux[g] - dual[g];
                    U[g] = x > 0.0 ? x : 0.0;
                }
            }
    }
}

__global__ void k_admm_final(const double *__restrict__ aux, const double *__restrict__ dual,
                             double *__restrict__ U, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double x = aux[i] - dual[i];
        U[i] = x > 0.0 ? x : 0.0;
    }
}

// G_s[m + n R] += sum_r U_s[r, m] U_s[r, n] (streaming_cpd.cpp:1321-1327).
__global__ void k_gram_add(double *__restrict__ G, const double *__restrict__ U, uint64_t rows,
                           int R) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= R * R) return;
    int n = e / R, m = e % R;
    double g = G[e];
    for (uint64_t r = 0; r < rows; ++r) g += U[r * R + m] * U[r * R + n];
    G[e] = g;
}

// ata = old_gram (*) cross_0 (*) cross_1 ... (streaming_cpd.cpp:1378-1392).
__global__ void k_ata(double *__restrict__ ata, const double *__restrict__ old_gram,
                      const double *__restrict__ cross, int ncross, int RR) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= RR) return;
    double v = old_gram[e];
    for (int q = 0; q < ncross; ++q) v *= cross[(size_t)q * RR + e];
    ata[e] = v;
}

__global__ void k_scale(double *__restrict__ p, double f, int n) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] *= f;
}

// Per-block sums of (P - U)^2 and U^2.
__global__ __launch_bounds__(256) void k_diff_norms(const double *__restrict__ P,
                                                    const double *__restrict__ U, uint64_t n,
                                                    double *__restrict__ part) {
    __shared__ double red[4];
    double a = 0.0, b = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        double d = P[i] - U[i];
        a += d * d;
        b += U[i] * U[i];
    }
    double sa = block_sum(a, red);
    double sb = block_sum(b, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sa;
        part[2 * blockIdx.x + 1] = sb;
    }
}

// delta += ||P - U|| / (||U|| + 1e-12) (streaming_cpd.cpp:1461).
__global__ void k_delta_add(const double *__restrict__ part, int nblk, double *__restrict__ delta) {
    double a = wave_reduce_entry(part, nblk, 2, 0, 0);
    double b = wave_reduce_entry(part, nblk, 2, 1, 0);
    if (threadIdx.x == 0) *delta += sqrt(a) / (sqrt(b) + 1e-12);
}

// Stream-mode index of a batch's nonzeros: time index - the batch's first.
__global__ void k_local_time(uint64_t *__restrict__ out, const uint64_t *__restrict__ t, uint64_t n,
                             uint64_t t0) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = t[i] - t0;
}

static constexpr int kAdmmMaxIter = 25;

// Workgroups of the ADMM kernels: one per row tile, at most `resident` (the
// ADMM loop's grid barrier needs the whole grid resident at once; at most
// 4 per CU and 1024 in all, the halves of the partials buffer).
static int admm_blocks(uint64_t rows, const TileShape &ts, int resident) {
    uint64_t n = (rows + ts.TR - 1) / ts.TR;
    return (int)std::min<uint64_t>(std::max<uint64_t>(n, 1), (uint64_t)resident);
}

}  // namespace alto

using namespace alto;

extern "C" int alto_hip_cpstream(const alto_hip_coo *coo, int s, int R, int max_iters,
                                 double epsilon, int nonneg, int max_slices, uint64_t nnz_threshold,
                                 uint64_t timeslice_limit, int prec,
                                 const double *const *init_factors, double *const *out_factors,
                                 uint64_t *out_dims, double *out_time_rows, int *out_outer_iters,
                                 double *out_admm_iters, double *out_grams,
                                 double *out_final_error, double *timings,
                                 void *stream) {
    if (!coo || !coo->dims || !coo->cidx || !coo->vals || !init_factors) return ALTO_ERR_ARG;
    const int N = coo->nmode;
    if (N < 2 || N > kMaxModes || s < 0 || s >= N || R < 1 || R > kMaxRank || max_iters < 0)
        return ALTO_ERR_ARG;
    if (prec != ALTO_PREC_F64 && prec != ALTO_PREC_F32) return ALTO_ERR_ARG;
    const uint64_t nnz = coo->nnz;
    if (nnz >= (1ULL << 32)) return ALTO_ERR_UNSUPPORTED;
    for (int m = 0; m < N; ++m)
        if (!init_factors[m] || coo->dims[m] == 0) return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    hipStream_t st = to_stream(stream);

    // ---- StreamingSparseTensor: sort by (s, s+1, ...), relabel, slices ----
    StreamInput sin;
    ALTO_TRY(stream_prepare(coo, s, &sin, st));
    const std::vector<Slice> slices = stream_batches(sin.slices, s, nnz_threshold, timeslice_limit);
    int nsl = (int)slices.size();  // batches
    if (max_slices > 0 && nsl > max_slices) nsl = max_slices;
    const uint64_t ntime = nsl ? slices[nsl - 1].t0 + slices[nsl - 1].dims[s] : 0;  // time indices processed

    // ---- device state ----
    uint64_t maxd[kMaxModes], maxall = 1, maxslice = 1;
    for (int m = 0; m < N; ++m) {
        maxd[m] = m == s ? 1 : 0;
        for (int k = 0; k < nsl; ++k) maxd[m] = std::max(maxd[m], slices[k].dims[m]);
        maxall = std::max(maxall, maxd[m]);
    }
    for (int k = 0; k < nsl; ++k) maxslice = std::max(maxslice, slices[k].end - slices[k].start);
    uint64_t used = nsl ? slices[nsl - 1].end : 0;
    DevBuf zidx, U[kMaxModes], P[kMaxModes], Uf[kMaxModes], duals[kMaxModes];
    DevBuf G, oldG, ata, cross, gpart, vinv, flag, bar, rho, buf, aux, part0, part, dpart, delta,
        admm_it, colp, lam;
    DevBuf *dci = sin.cid;
    DevBuf &dvals = sin.vals;
    ALTO_TRY(zidx.alloc(sizeof(uint64_t) * maxslice));
    for (int m = 0; m < N; ++m) {
        ALTO_TRY(U[m].alloc(sizeof(double) * maxd[m] * R));
        ALTO_TRY(P[m].alloc(sizeof(double) * maxd[m] * R));
        ALTO_TRY(duals[m].alloc(sizeof(double) * maxd[m] * R));
        if (prec == ALTO_PREC_F32) ALTO_TRY(Uf[m].alloc(sizeof(float) * maxd[m] * R));
    }
    const size_t RR = (size_t)R * R;
    ALTO_TRY(G.alloc(sizeof(double) * N * RR));
    ALTO_TRY(oldG.alloc(sizeof(double) * RR));
    ALTO_TRY(ata.alloc(sizeof(double) * RR));
    ALTO_TRY(cross.alloc(sizeof(double) * N * RR));
    ALTO_TRY(gpart.alloc(sizeof(double) * kMaxBlocks * RR));
    ALTO_TRY(vinv.alloc(sizeof(double) * RR));
    ALTO_TRY(flag.alloc(sizeof(int) * 4));
    ALTO_TRY(bar.alloc(sizeof(unsigned) * 4));
    ALTO_HIP_TRY(hipMemsetAsync(flag.p, 0, sizeof(int) * 4, st));
    ALTO_HIP_TRY(hipMemsetAsync(bar.p, 0, sizeof(unsigned) * 4, st));
    ALTO_TRY(rho.alloc(sizeof(double)));
    ALTO_TRY(buf.alloc(sizeof(double) * maxall * R));
    ALTO_TRY(aux.alloc(sizeof(double) * maxall * R));
    ALTO_TRY(part0.alloc(sizeof(double) * kMaxBlocks));
    ALTO_TRY(part.alloc(sizeof(double) * 3 * kMaxBlocks));
    ALTO_TRY(dpart.alloc(sizeof(double) * 2 * kMaxBlocks));
    ALTO_TRY(delta.alloc(sizeof(double)));
    ALTO_TRY(admm_it.alloc(sizeof(double)));
    ALTO_TRY(colp.alloc(sizeof(double) * 3 * R * kMaxBlocks));
    ALTO_TRY(lam.alloc(sizeof(double) * R));
    DevBuf trows;
    ALTO_TRY(trows.alloc(sizeof(double) * std::max<uint64_t>(ntime, 1) * R));
    ALTO_TRY(launch_fill_zero(trows.as<double>(), std::max<uint64_t>(ntime, 1) * R, st));
    TensorImpl at;
    at.keep_sched_scratch = true;  // rebuilt for every slice
    const TileShape ts = tile_shape(R);
    const size_t lds_first = sizeof(double) * (2 * RR + 2 * (size_t)ts.TR * R);
    const size_t lds_iter = sizeof(double) * (RR + (size_t)ts.TR * R);
    const size_t lds_iter_c =
        lds_iter + sizeof(double) * (RR + (kAdmmCacheTiles + 1) * (size_t)ts.TR * R);  // + Mt, ata, P tile
    if (lds_first > 160 * 1024 - 512) {
        set_error("CP-stream supports rank <= 64 (rank %d)", R);
        return ALTO_ERR_UNSUPPORTED;
    }
    {
        static std::once_flag attr;
        std::call_once(attr, [] {
            (void)hipFuncSetAttribute((const void *)k_admm_first,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256);
            (void)hipFuncSetAttribute((const void *)k_admm_loop<0>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256);
            (void)hipFuncSetAttribute((const void *)k_admm_loop<kAdmmCacheTiles>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256);
        });
    }
    // ADMM grids, all resident at once (occupancy x CUs, at most 4 per CU and
    // 1024 in all): the first pass alone (unconstrained), and the streamed
    // and the cached form of the ADMM loop, whose grid barrier gives up after
    // 50 ms of wall clock (the barrier measured slower at 4 workgroups per CU
    // than at 2: 262 against 279 slices/s over the 1000-slice stream, so the
    // streamed loop takes at most 2)
    int admm_resident = 1, admm_resident_s = 1, admm_resident_c = 0;
    long long bar_limit = 0;
    // ALTO_ADMM_CACHE=0 in the environment streams the ADMM state from memory
    // every iteration (the on-chip form is the default where it fits)
    bool admm_cache = true;
    {
        const char *e = getenv("ALTO_ADMM_CACHE");
        if (e && e[0] == '0') admm_cache = false;
    }
    {
        int dev = 0, cus = 0, occ = 0, khz = 0;
        ALTO_HIP_TRY(hipGetDevice(&dev));
        ALTO_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        ALTO_HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        int occ_c = 0;
        ALTO_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_admm_loop<0>, 256, lds_iter));
        if (lds_iter_c <= 160 * 1024 - 512)
            ALTO_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_c, k_admm_loop<kAdmmCacheTiles>, 256,
                                                                      lds_iter_c));
        admm_resident_c = std::min(cus * std::min(occ_c, 4), 1024);
        if (occ < 1 || cus < 1 || khz < 1) {
            set_error("ADMM loop kernel: no resident workgroup (occupancy %d, %d CUs)", occ, cus);
            return ALTO_ERR_UNSUPPORTED;
        }
        admm_resident = std::min(cus * std::min(occ, 4), 1024);
        admm_resident_s = std::min(cus * std::min(occ, 2), 1024);
        bar_limit = (long long)khz * 50;
    }
    EventSet evs;
    hipEvent_t ev[4];
    for (auto &e : ev) ALTO_TRY(evs.make(&e));
    double t_build = 0, t_mttkrp = 0;
    ALTO_HIP_TRY(hipEventRecord(ev[0], st));

    uint64_t cur[kMaxModes] = {0};
    const void *fptr[kMaxModes];
    auto gather_ptrs = [&]() {
        for (int m = 0; m < N; ++m)
            fptr[m] = prec == ALTO_PREC_F32 ? Uf[m].p : U[m].p;
    };
    gather_ptrs();
    auto refresh_f32 = [&](int m, uint64_t rows) -> int {
        if (prec == ALTO_PREC_F32 && rows)
            return alto_hip_convert(Uf[m].p, ALTO_PREC_F32, U[m].p, ALTO_PREC_F64, rows * R, st);
        return ALTO_OK;
    };
    auto gram = [&](int m, uint64_t rows) -> int {
        int nb = row_blocks(rows);
        ALTO_TRY(launch_gram(U[m].p, ALTO_PREC_F64, rows, R, gpart.as<double>(), nb, st));
        return launch_gram_reduce(gpart.as<double>(), nb, R, G.as<double>() + m * RR, st);
    };

    IterGraph ig(st);  // the outer-iteration graph, reused from slice to slice
    for (int k = 0; k < nsl; ++k) {
        const Slice &sl = slices[k];
        const uint64_t *bd = sl.dims;
        if (k == 0) {
            for (int m = 0; m < N; ++m) {
                ALTO_HIP_TRY(hipMemcpyAsync(U[m].p, init_factors[m], sizeof(double) * bd[m] * R,
                                            hipMemcpyHostToDevice, st));
                ALTO_TRY(launch_fill_zero(P[m].as<double>(), maxd[m] * R, st));
            }
            ALTO_TRY(launch_fill_zero(U[s].as<double>(), bd[s] * R, st));
            for (int m = 0; m < N; ++m) ALTO_TRY(gram(m, bd[m]));
            ALTO_TRY(launch_fill_zero(G.as<double>() + s * RR, RR, st));
        } else {
            for (int m = 0; m < N; ++m)
                if (bd[m] > cur[m]) {
                    ALTO_HIP_TRY(hipMemcpyAsync(U[m].as<double>() + cur[m] * R, init_factors[m] + cur[m] * R,
                                                sizeof(double) * (bd[m] - cur[m]) * R,
                                                hipMemcpyHostToDevice, st));
                    ALTO_TRY(launch_fill_zero(P[m].as<double>() + cur[m] * R, (bd[m] - cur[m]) * R, st));
                }
            for (int m = 0; m < N; ++m)
                if (m != s) ALTO_TRY(gram(m, bd[m]));
        }
        for (int m = 0; m < N; ++m) {
            cur[m] = bd[m];
            ALTO_TRY(launch_fill_zero(duals[m].as<double>(), bd[m] * R, st));
        }
        // update_salto (alto.hpp:195-299) into reused buffers
        {
            const uint64_t *cp[kMaxModes];
            k_local_time<<<cdiv(sl.end - sl.start, 256), 256, 0, st>>>(
                zidx.as<uint64_t>(), dci[s].as<uint64_t>() + sl.start, sl.end - sl.start, sl.t0);
            ALTO_LAUNCH_CHECK();
            for (int m = 0; m < N; ++m)
                cp[m] = m == s ? zidx.as<uint64_t>() : dci[m].as<uint64_t>() + sl.start;
            alto_hip_coo sc{N, bd, sl.end - sl.start, cp, dvals.as<double>() + sl.start, 1};
            ALTO_HIP_TRY(hipEventRecord(ev[2], st));
            ALTO_TRY(build_alto(&at, &sc, 0, 0, false, st, false));  // only the schedules are read
            ALTO_HIP_TRY(hipEventRecord(ev[3], st));
        }
        if (k == 0) {  // cpstream_alto_iter, iter == 0 (streaming_cpd.cpp:1248-1261)
            for (int m = 0; m < N; ++m) {
                if (m == s) continue;
                int nb = row_blocks(bd[m]);
                ALTO_TRY(launch_colsumsq(U[m].as<double>(), bd[m], R, colp.as<double>(), nb, st));
                ALTO_TRY(launch_lambda(colp.as<double>(), nb, R, 0, lam.as<double>(), st));
                ALTO_TRY(launch_scale_gram(U[m].as<double>(), lam.as<double>(), bd[m], R, U[m].p,
                                           ALTO_PREC_F64, gpart.as<double>(), nb, st));
                ALTO_TRY(launch_gram_reduce(gpart.as<double>(), nb, R, G.as<double>() + m * RR, st));
            }
        }
        for (int m = 0; m < N; ++m) ALTO_TRY(refresh_f32(m, bd[m]));
        ALTO_TRY(launch_fill_zero(oldG.as<double>(), RR, st));
        ALTO_TRY(launch_fill_zero(admm_it.as<double>(), 1, st));
        double prev_delta = 0.0;
        int outer = 0;
        // one outer iteration (streaming_cpd.cpp:1269-1470), no host syncs:
        // replayed as a graph from the second iteration on
        auto outer_iter = [&]() -> int {
            ALTO_TRY(launch_fill_zero(delta.as<double>(), 1, st));
            // s_t: stream-mode MTTKRP and solve (streaming_cpd.cpp:1272-1299)
            ALTO_TRY(launch_fill_zero(U[s].as<double>(), bd[s] * R, st));
            ALTO_TRY(run_mttkrp(&at, s, prec, R, fptr, U[s].p, 1, 2, 0, -1, st, false, true));
            ALTO_TRY(launch_chol_inv(G.as<double>(), N, s, R, 1e-12, false, vinv.as<double>(),
                                     flag.as<int>() + 1, nullptr, st, 1, nullptr));
            ALTO_TRY(launch_solve(U[s].as<double>(), bd[s], R, vinv.as<double>(), colp.as<double>(),
                                  1, st));
            ALTO_TRY(refresh_f32(s, bd[s]));
            // temporal Gram (streaming_cpd.cpp:1301-1327)
            ALTO_HIP_TRY(hipMemcpyAsync(oldG.p, G.as<double>() + s * RR, sizeof(double) * RR,
                                        hipMemcpyDeviceToDevice, st));
            k_gram_add<<<cdiv(RR, 256), 256, 0, st>>>(G.as<double>() + s * RR, U[s].as<double>(), bd[s], R);
            for (int j = 0; j < N; ++j) {
                if (j == s) continue;
                const uint64_t I = bd[j];
                const int nb = row_blocks(I);
                ALTO_TRY(launch_fill_zero(buf.as<double>(), I * R, st));
                ALTO_TRY(run_mttkrp(&at, j, prec, R, fptr, buf.p, 1, 2, 0, -1, st, false, true));
                int nc = 0;
                for (int m = 0; m < N; ++m) {
                    if (m == j || m == s) continue;
                    int nbm = row_blocks(bd[m]);
                    ALTO_TRY(launch_cross_gram(P[m].as<double>(), U[m].as<double>(), bd[m], R,
                                               gpart.as<double>(), nbm, st));
                    ALTO_TRY(launch_gram_reduce(gpart.as<double>(), nbm, R, cross.as<double>() + nc * RR, st));
                    ++nc;
                }
                k_ata<<<cdiv(RR, 256), 256, 0, st>>>(ata.as<double>(), oldG.as<double>(),
                                                     cross.as<double>(), nc, (int)RR);
                if (nonneg) {
                    ALTO_TRY(launch_chol_inv(G.as<double>(), N, j, R, 0.0, false, vinv.as<double>(),
                                             flag.as<int>() + 2, nullptr, st, 1, rho.as<double>()));
                } else {
                    ALTO_TRY(launch_chol_inv(G.as<double>(), N, j, R, 1e-12, false, vinv.as<double>(),
                                             flag.as<int>() + 1, nullptr, st, 1, nullptr));
                }
                // the ADMM loop keeps its rows on chip when every workgroup of
                // the cached form's resident grid owns at most kAdmmCacheTiles
                // row tiles; the first pass runs on the same grid (the loop
                // reads its per-workgroup partials)
                const bool cache = nonneg && admm_cache && admm_resident_c > 0 &&
                                   (I + ts.TR - 1) / ts.TR <= (uint64_t)kAdmmCacheTiles * admm_resident_c;
                const int nba =
                    admm_blocks(I, ts, !nonneg ? admm_resident : cache ? admm_resident_c : admm_resident_s);
                if (!cache) {  // the on-chip loop form runs the first pass itself
                    k_admm_first<<<nba, 256, lds_first, st>>>(buf.as<double>(), P[j].as<double>(),
                                                             ata.as<double>(), U[j].as<double>(),
                                                             duals[j].as<double>(), vinv.as<double>(),
                                                             rho.as<double>(), I, R, ts.RP, nonneg,
                                                             aux.as<double>(), part0.as<double>());
                    ALTO_LAUNCH_CHECK();
                }
                if (nonneg) {
                    (cache ? k_admm_loop<kAdmmCacheTiles> : k_admm_loop<0>)<<<nba, 256,
                                                                           cache ? lds_iter_c : lds_iter, st>>>(
                        buf.as<double>(), aux.as<double>(), duals[j].as<double>(), vinv.as<double>(),
                        rho.as<double>(), U[j].as<double>(), I, R, ts.RP, part.as<double>(), part0.as<double>(),
                        bar.as<unsigned>(), flag.as<int>() + 3, bar_limit, kAdmmMaxIter, admm_it.as<double>(),
                        P[j].as<double>(), ata.as<double>());
                    ALTO_LAUNCH_CHECK();
                }
                ALTO_TRY(gram(j, I));
                ALTO_TRY(refresh_f32(j, I));
                k_diff_norms<<<nb, 256, 0, st>>>(P[j].as<double>(), U[j].as<double>(), I * R,
                                                 dpart.as<double>());
                k_delta_add<<<1, 64, 0, st>>>(dpart.as<double>(), nb, delta.as<double>());
                ALTO_LAUNCH_CHECK();
            }
            return ALTO_OK;
        };
        // The iterations after the first replay one captured graph (kept
        // across batches, updated in place); the host reads delta and the
        // Cholesky flags once per outer iteration.
        for (int i = 0; i < max_iters; ++i) {
            ALTO_TRY(ig.run(i, outer_iter));
            double hd[2];
            int hf[4];
            ALTO_HIP_TRY(hipMemcpyAsync(hd, delta.p, sizeof(double), hipMemcpyDeviceToHost, st));
            ALTO_HIP_TRY(hipMemcpyAsync(hf, flag.p, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
            ALTO_HIP_TRY(hipStreamSynchronize(st));
            if (nonneg && hf[2]) {
                set_error("Cholesky factorisation failed in ADMM (a mode's Gram Hadamard, batch %d)", k);
                return ALTO_ERR_NUMERIC;
            }
            if (nonneg && hf[3]) {
                set_error("ADMM loop: its workgroups were not all resident at once (the device is shared "
                          "with other work), batch %d",
                          k);
                return ALTO_ERR_HIP;
            }
            float bms = 0;
            (void)hipEventElapsedTime(&bms, ev[2], ev[3]);
            if (i == 0) t_build += bms * 1e-3;
            outer = i + 1;
            if (i > 0 && fabs(prev_delta - hd[0]) < epsilon) break;
            prev_delta = hd[0];
        }
        k_scale<<<cdiv(RR, 256), 256, 0, st>>>(G.as<double>() + s * RR, 0.99, (int)RR);
        for (int m = 0; m < N; ++m)
            ALTO_HIP_TRY(hipMemcpyAsync(P[m].p, U[m].p, sizeof(double) * bd[m] * R,
                                        hipMemcpyDeviceToDevice, st));
        double ait = 0.0;
        if (out_time_rows)
            ALTO_HIP_TRY(hipMemcpyAsync(out_time_rows + (size_t)sl.t0 * R, U[s].p, sizeof(double) * bd[s] * R,
                                        hipMemcpyDeviceToHost, st));
        if (out_final_error)
            ALTO_HIP_TRY(hipMemcpyAsync(trows.as<double>() + (size_t)sl.t0 * R, U[s].p,
                                        sizeof(double) * bd[s] * R, hipMemcpyDeviceToDevice, st));
        ALTO_HIP_TRY(hipMemcpyAsync(&ait, admm_it.p, sizeof(double), hipMemcpyDeviceToHost, st));
        ALTO_HIP_TRY(hipStreamSynchronize(st));
        if (out_outer_iters) out_outer_iters[k] = outer;
        if (out_admm_iters) out_admm_iters[k] = ait;
    }
    ALTO_HIP_TRY(hipEventRecord(ev[1], st));
    ALTO_HIP_TRY(hipEventSynchronize(ev[1]));
    float tot_ms = 0;
    (void)hipEventElapsedTime(&tot_ms, ev[0], ev[1]);
    if (out_final_error) {
        // cpd_error(X, factored) (streaming_cpd.cpp:307-345): the processed
        // nonzeros in relabeled coordinates, the time factor = the s_t rows.
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
        timings[2] = t_mttkrp;
        timings[3] = (double)ntime;
    }
    return nsl;
}
