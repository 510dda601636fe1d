// ALTO MTTKRP on gfx950.
//
// Reference: mttkrp_alto_par and its variants (include/alto.hpp:826-908,
// 1349-1407 _atomic, 1544-1634 _da_mem_pull, 1636-1706 _stream,
// 1708-1801 _op_sched).  Per nonzero: row = val * prod_{m != n}
// A_m[idx_m(key)], accumulated into A_n[idx_n(key)] (out +=).
//
// GPU mapping: a lane group of L = R / V lanes owns one output row slice of
// V elements per lane (V = 16 bytes / factor element size), so every factor
// row gather is one 16-byte load per lane and a wave moves 64/L rows per
// instruction.  Each group walks a contiguous range of nonzeros in batches
// of B, issuing the B gathers of a mode back to back, and keeps a running
// sum while the output index repeats; the sum is flushed with a hardware
// atomic add only when the output index changes and at the end of the range.
//   algo 1: nonzeros in ALTO order (runs come from duplicates / dense regions)
//   algo 2: nonzeros in the mode's output order (ALTO order inside each row;
//           the oidx order of the op_sched variant) -> one flush per row per
//           group, atomics only where a row crosses a group boundary.
#include <stdlib.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "mttkrp.hpp"
#include "tensor.hpp"

namespace alto {

template <typename FT> struct FVec;
template <> struct FVec<double> {
    static constexpr int V = 2;
    using CT = double;
    __device__ __forceinline__ static void load(const double *p, double *o) {
        double2 v = *(const double2 *)p;
        o[0] = v.x;
        o[1] = v.y;
    }
};
template <> struct FVec<float> {
    static constexpr int V = 4;
    using CT = float;
    __device__ __forceinline__ static void load(const float *p, float *o) {
        float4 v = *(const float4 *)p;
        o[0] = v.x;
        o[1] = v.y;
        o[2] = v.z;
        o[3] = v.w;
    }
};
template <> struct FVec<uint16_t> {  // BF16 storage, FP32 compute
    static constexpr int V = 8;
    using CT = float;
    __device__ __forceinline__ static void load(const uint16_t *p, float *o) {
        uint4 v = *(const uint4 *)p;
        uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            o[2 * q] = __uint_as_float(w[q] << 16);
            o[2 * q + 1] = __uint_as_float(w[q] & 0xffff0000u);
        }
    }
};

template <typename FT> struct FPtrs {
    const FT *p[kMaxModes];
};

__device__ __forceinline__ void atomic_add_out(double *p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add_out(float *p, float v) { unsafeAtomicAdd(p, v); }

template <int PKW>
__device__ __forceinline__ uint64_t field(uint64_t lo, uint64_t hi, int sh, int w) {
    if (PKW == 1) return w >= 64 ? lo >> sh : (lo >> sh) & ((1ULL << w) - 1ULL);
    return extract_field(lo, hi, sh, w);
}

template <typename T> __device__ __forceinline__ T stream_load(const T *p, bool nt) {
    return nt ? __builtin_nontemporal_load(p) : *p;
}

// flags: bit 0 = non-temporal loads of the key/value stream (keeps factor
// rows resident in L2), bit 1 = skip the gathers of a nonzero whose key
// equals the previous one (duplicate coordinates reuse the gathered product),
// bit 2 = exclusive runs: the nonzeros are in output-row order, so a run that
// starts and ends inside one lane group's range owns its row (plain
// read-add-store), and only the runs at range boundaries are combined -- in
// LDS across the workgroup, then with atomics for the workgroup's first and
// last rows; bit 3 = no decode-once kernel; bit 4 = no fiber kernel (output
// order only).
// Rows owned by one range are flushed with flush_owned (mttkrp.hpp): stored
// when the output held zeros on entry (flags bit 5), else added.

template <typename VT, typename FT, typename OT, int R, int PKW, int B>
__global__ __launch_bounds__(256) void k_mttkrp_seg(const uint64_t *__restrict__ pkeys,
                                                    const VT *__restrict__ vals, uint64_t nnz,
                                                    FPtrs<FT> F, OT *__restrict__ out,
                                                    PackInfo pk, int target, uint64_t per_group,
                                                    int flags) {
    using Vec = FVec<FT>;
    using CT = typename Vec::CT;
    constexpr int V = Vec::V;
    constexpr int L = R / V;
    constexpr int G = 256 / L;  // lane groups per workgroup
    static_assert(R % V == 0 && L >= 1 && L <= 64 && (64 % L) == 0, "bad rank tiling");
    __shared__ CT s_part[2][G][R];       // first / last run partial of each group
    __shared__ uint64_t s_row[2][G];     // their rows (~0 = none)

    const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t group = gtid / L;
    const int gib = threadIdx.x / L;     // group in block
    const int lig = (int)(gtid % L);
    const bool excl = flags & 4;
    uint64_t start = group * per_group;
    uint64_t end = start + per_group < nnz ? start + per_group : nnz;
    if (start > nnz) start = nnz;

    const int nmode = pk.nmode;
    const int tsh = pk.shift[target], tw = pk.width[target];
    OT *outp = out + lig * V;

    const bool nt = flags & 1, dedup = flags & 2, zeroed = flags & 32;
    uint64_t cur = ~0ULL, first_row = ~0ULL;
    bool in_first = true;
    CT acc[V], facc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        acc[v] = 0;
        facc[v] = 0;
    }
    uint64_t plo = ~0ULL, phi = ~0ULL;
    CT pg[V];
#pragma unroll
    for (int v = 0; v < V; ++v) pg[v] = 0;

    for (uint64_t base = start; base < end; base += B) {
        uint64_t klo[B], khi[B];
        CT val[B];
        bool dup[B];
        CT g[B][V];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            uint64_t i = base + k < end ? base + k : end - 1;
            klo[k] = stream_load(pkeys + i * PKW, nt);
            khi[k] = PKW == 2 ? stream_load(pkeys + i * PKW + 1, nt) : 0;
            val[k] = (CT)stream_load(vals + i, nt);
        }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            uint64_t pl = k == 0 ? plo : klo[k - 1], ph = k == 0 ? phi : khi[k - 1];
            dup[k] = dedup && (k > 0 || base > start) && klo[k] == pl && khi[k] == ph;
#pragma unroll
            for (int v = 0; v < V; ++v) g[k][v] = 1;
        }
        for (int m = 0; m < nmode; ++m) {
            if (m == target) continue;
            const int sh = pk.shift[m], w = pk.width[m];
            const FT *fm = F.p[m] + lig * V;
            CT f[B][V];
#pragma unroll
            for (int k = 0; k < B; ++k) {
                if (!dup[k]) {
                    uint64_t row = field<PKW>(klo[k], khi[k], sh, w);
                    Vec::load(fm + row * R, f[k]);
                } else {
#pragma unroll
                    for (int v = 0; v < V; ++v) f[k][v] = 1;
                }
            }
#pragma unroll
            for (int k = 0; k < B; ++k)
#pragma unroll
                for (int v = 0; v < V; ++v) g[k][v] *= f[k][v];
        }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            if (dup[k]) {
#pragma unroll
                for (int v = 0; v < V; ++v) g[k][v] = k == 0 ? pg[v] : g[k - 1][v];
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v) pg[v] = g[B - 1][v];
        plo = klo[B - 1];
        phi = khi[B - 1];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            if (base + k < end) {
                uint64_t row = field<PKW>(klo[k], khi[k], tsh, tw);
                if (row != cur) {
                    if (cur != ~0ULL) {
                        if (!excl) {
#pragma unroll
                            for (int v = 0; v < V; ++v) atomic_add_out(outp + cur * R + v, (OT)acc[v]);
                        } else if (in_first) {
                            first_row = cur;
#pragma unroll
                            for (int v = 0; v < V; ++v) facc[v] = acc[v];
                            in_first = false;
                        } else {
#pragma unroll
                            for (int v = 0; v < V; ++v) flush_owned(outp + cur * R + v, acc[v], zeroed);
                        }
                    }
                    cur = row;
#pragma unroll
                    for (int v = 0; v < V; ++v) acc[v] = val[k] * g[k][v];
                } else {
#pragma unroll
                    for (int v = 0; v < V; ++v) acc[v] += val[k] * g[k][v];
                }
            }
        }
    }
    if (!excl) {
        if (cur != ~0ULL) {
#pragma unroll
            for (int v = 0; v < V; ++v) atomic_add_out(outp + cur * R + v, (OT)acc[v]);
        }
        return;
    }
    // Boundary runs: [first run] ... [last run] of each group, in order.
    if (in_first) {  // zero or one run in this group: it is the first run
        first_row = cur;
#pragma unroll
        for (int v = 0; v < V; ++v) facc[v] = acc[v];
        cur = ~0ULL;
    }
#pragma unroll
    for (int v = 0; v < V; ++v) {
        s_part[0][gib][lig * V + v] = facc[v];
        s_part[1][gib][lig * V + v] = acc[v];
    }
    if (lig == 0) {
        s_row[0][gib] = first_row;
        s_row[1][gib] = cur;
    }
    __syncthreads();
    if (threadIdx.x < R) {
        const int r = threadIdx.x;
        OT *o = out + r;
        uint64_t crow = ~0ULL;
        CT a = 0;
        bool first_merged = true;
        for (int e = 0; e < 2 * G; ++e) {
            uint64_t row = s_row[e & 1][e >> 1];
            if (row == ~0ULL) continue;
            CT x = s_part[e & 1][e >> 1][r];
            if (row == crow) {
                a += x;
            } else {
                if (crow != ~0ULL) {
                    if (first_merged) atomic_add_out(o + crow * R, (OT)a);
                    else flush_owned(o + crow * R, a, zeroed);
                    first_merged = false;
                }
                crow = row;
                a = x;
            }
        }
        if (crow != ~0ULL) atomic_add_out(o + crow * R, (OT)a);
    }
}

// Decode-once variant of k_mttkrp_seg (same ranges, same run merging and
// boundary combining).  Each lane of a group decodes one nonzero of the
// group's next L (coalesced key/value loads) into an LDS record holding the
// byte offsets of the gathered rows; the group reads the records back with
// broadcast LDS reads, so per nonzero a lane only issues its gathers (uniform
// base address + 32-bit offset, VPL elements in 16-byte loads) and its
// multiplies.  The keys and values of the next L nonzeros are loaded while
// the current gathers are in flight.  NM1 = number of gathered modes.
// Workgroups are remapped so that each XCD (round-robin dispatch over 8
// XCDs) walks one contiguous slice of the nonzeros: neighbouring ranges share
// that XCD's L2.  Needs I_m * R * sizeof(factor) < 2^32 bytes.
template <typename VT> struct __align__(16) DecRec {
    uint32_t off[4];  // byte offsets of the gathered rows (non-target modes)
    uint32_t trow;    // output row
    uint32_t pad;
    VT val;
};

template <typename FT, int VPL> struct RowLoad {
    using CT = typename FVec<FT>::CT;
    static constexpr int E16 = 16 / (int)sizeof(FT);  // elements per 16-byte load
    static_assert(VPL % E16 == 0, "VPL must be a multiple of a 16-byte load");
    __device__ __forceinline__ static void load(const char *p, CT *o) {
#pragma unroll
        for (int q = 0; q < VPL / E16; ++q) FVec<FT>::load((const FT *)(p + 16 * q), o + q * E16);
    }
};


// Boundary slots of an FP32-output MTTKRP: workgroup lb writes its FP32 sum
// of its first row to slot 2 lb and of its last row to slot 2 lb + 1 (a
// workgroup of one row repeats it with zero), in schedule order, and
// launch_bnd_merge adds the runs of equal rows in slot order in FP64.  With
// row == nullptr the boundary rows take hardware atomics in the output type.
struct BndSlots {
    uint32_t *row;
    float *val;         // a workgroup's partial is an FP32 sum: exact in FP32
    uint32_t *lrow[2];  // the merge's level buffers (FP64 sums)
    double *lval[2];
};

// Combines one workgroup's range partials: slot e = 0 .. 2G-1 holds group
// e >> 1's first (e even) or last (e odd) row partial, in output order, with
// row 0xFFFFFFFF for an empty slot.  Consecutive equal rows are summed; the
// workgroup's first and last rows may continue in the neighbouring
// workgroups (atomic adds), the rows between are owned by this workgroup.
// Two levels, so no thread walks all 2G slots: P = 256 / R threads per
// column each reduce 2G / P consecutive slots and flush the runs that start
// and end inside their chunk (owned: neither the workgroup's first nor its
// last row); one thread per column then merges the chunks' first and last
// runs in order.  Requires a 256-thread workgroup.
template <typename OT, typename CT, int R, int G>
__device__ __forceinline__ void block_combine(const uint32_t (*s_row)[G], const CT (*s_part)[G][R], OT *out,
                                              bool zeroed, BndSlots bs, unsigned lb) {
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    constexpr int P = 256 / R, S = 2 * G, C = S / P;
    static_assert(256 % R == 0 && S % P == 0, "combine tiling");
    __shared__ uint32_t m_row[2][R][P];
    __shared__ CT m_sum[2][R][P];
    const int col = threadIdx.x / P, j = threadIdx.x % P;
    OT *o = out + col;
    uint32_t frow = NONE, crow = NONE;
    CT fsum = 0, csum = 0;
    int runs = 0;
#pragma unroll
    for (int k = 0; k < C; ++k) {
        const int e = j * C + k;
        const uint32_t row = s_row[e & 1][e >> 1];
        if (row == NONE) continue;
        const CT x = s_part[e & 1][e >> 1][col];
        if (row == crow) {
            csum += x;
            continue;
        }
        if (runs == 1) {
            frow = crow;
            fsum = csum;
        } else if (runs > 1) {
            flush_owned(o + (uint64_t)crow * R, csum, zeroed);
        }
        crow = row;
        csum = x;
        ++runs;
    }
    if (runs == 1) {  // a single run: it is the chunk's first
        frow = crow;
        fsum = csum;
        crow = NONE;
    }
    m_row[0][col][j] = frow;
    m_sum[0][col][j] = fsum;
    m_row[1][col][j] = crow;
    m_sum[1][col][j] = csum;
    __syncthreads();
    if (j != 0) return;
    uint32_t mrow = NONE, r0 = NONE;
    CT a = 0, a0 = 0;
    bool first = true;
    for (int k = 0; k < 2 * P; ++k) {
        const uint32_t row = m_row[k & 1][col][k >> 1];
        if (row == NONE) continue;
        const CT x = m_sum[k & 1][col][k >> 1];
        if (row == mrow) {
            a += x;
        } else {
            if (mrow != NONE) {
                if (first) {
                    if (bs.row) {
                        r0 = mrow;
                        a0 = a;
                    } else {
                        atomic_add_out(o + (uint64_t)mrow * R, (OT)a);
                    }
                } else {
                    flush_owned(o + (uint64_t)mrow * R, a, zeroed);
                }
                first = false;
            }
            mrow = row;
            a = x;
        }
    }
    if (bs.row) {
        if (first) {  // one row in the whole workgroup
            r0 = mrow;
            a0 = a;
            a = 0;
        }
        if (col == 0) {
            bs.row[2 * (uint64_t)lb] = r0;
            bs.row[2 * (uint64_t)lb + 1] = mrow;
        }
        bs.val[2 * (uint64_t)lb * R + col] = (float)a0;
        bs.val[(2 * (uint64_t)lb + 1) * R + col] = (float)a;
    } else if (mrow != NONE) {
        atomic_add_out(o + (uint64_t)mrow * R, (OT)a);
    }
}

template <typename VT, typename FT, typename OT, int R, int PKW, int NM1, int VPL, bool ZEROED>
__device__ __forceinline__ void k_mttkrp_dec_body(unsigned lb, const uint64_t *__restrict__ pkeys,
                                                  const VT *__restrict__ vals, uint64_t nnz,
                                                  const FPtrs<FT> &F, OT *__restrict__ out,
                                                  const PackInfo &pk, int target, uint64_t per_group,
                                                  int flags, BndSlots bs) {
    using CT = typename FVec<FT>::CT;
    constexpr int L = R / VPL;
    constexpr int G = 256 / L;
    constexpr int CB0 = VPL >= 8 ? 2 : 4;
    constexpr int CB = L < CB0 ? L : CB0;  // nonzeros whose gathers are issued together
    static_assert(R % VPL == 0 && L >= 1 && L <= 64 && (64 % L) == 0, "bad rank tiling");
    using Rec = DecRec<VT>;
    __shared__ Rec s_rec[256];
    __shared__ CT s_part[2][G][R];
    __shared__ uint32_t s_row[2][G];
    constexpr uint32_t NONE = 0xFFFFFFFFu;

    const uint64_t gtid = (uint64_t)lb * blockDim.x + threadIdx.x;
    const uint64_t group = gtid / L;
    const int gib = threadIdx.x / L;
    const int lig = threadIdx.x % L;
    const bool nt = flags & 1, excl = flags & 4;
    constexpr bool zeroed = ZEROED;  // output zero on entry: owned rows are stored
    uint64_t start = group * per_group;
    uint64_t end = start + per_group < nnz ? start + per_group : nnz;
    if (start > nnz) start = nnz;

    int osh[NM1], ow[NM1];
    const char *fb[NM1];
#pragma unroll
    for (int q = 0; q < NM1; ++q) {
        int m = q < target ? q : q + 1;  // q-th mode other than the target
        osh[q] = pk.shift[m];
        ow[q] = pk.width[m];
        fb[q] = (const char *)F.p[m];
    }
    const int tsh = pk.shift[target], tw = pk.width[target];
    const uint32_t lane_b = (uint32_t)(lig * VPL * sizeof(FT));
    OT *outp = out + lig * VPL;
    Rec *myrec = s_rec + gib * L;

    uint32_t cur = NONE, first_row = NONE;
    bool in_first = true;
    CT acc[VPL], facc[VPL];
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
        acc[v] = 0;
        facc[v] = 0;
    }
    uint64_t klo = 0, khi = 0;
    VT kval = 0;
    if (start < end) {
        uint64_t i = start + lig < end ? start + lig : end - 1;
        klo = stream_load(pkeys + i * PKW, nt);
        if (PKW == 2) khi = stream_load(pkeys + i * PKW + 1, nt);
        kval = stream_load(vals + i, nt);
    }
    for (uint64_t base = start; base < end; base += L) {
        {
            Rec r;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                r.off[q] = q < NM1 ? (uint32_t)field<PKW>(klo, khi, osh[q < NM1 ? q : 0], ow[q < NM1 ? q : 0]) *
                                         (uint32_t)(R * sizeof(FT))
                                   : 0u;
            r.trow = (uint32_t)field<PKW>(klo, khi, tsh, tw);
            r.pad = 0;
            r.val = kval;
            myrec[lig] = r;
        }
        const uint64_t nb = base + L;
        if (nb < end) {  // prefetch the next step
            uint64_t i = nb + lig < end ? nb + lig : end - 1;
            klo = stream_load(pkeys + i * PKW, nt);
            if (PKW == 2) khi = stream_load(pkeys + i * PKW + 1, nt);
            kval = stream_load(vals + i, nt);
        }
        __builtin_amdgcn_wave_barrier();
        const int nh = (end - base) < (uint64_t)L ? (int)(end - base) : L;
#pragma unroll
        for (int c0 = 0; c0 < L; c0 += CB) {
            uint32_t ob[CB][NM1], tr[CB];
            CT vv[CB];
#pragma unroll
            for (int k = 0; k < CB; ++k) {
                const Rec &r = myrec[c0 + k];
#pragma unroll
                for (int q = 0; q < NM1; ++q) ob[k][q] = r.off[q] + lane_b;
                tr[k] = r.trow;
                vv[k] = (CT)r.val;
            }
            CT f[NM1][CB][VPL];
#pragma unroll
            for (int q = 0; q < NM1; ++q)
#pragma unroll
                for (int k = 0; k < CB; ++k) RowLoad<FT, VPL>::load(fb[q] + ob[k][q], f[q][k]);
#pragma unroll
            for (int k = 0; k < CB; ++k) {
                // entries past the range's end add zero to the current run
                const bool live = c0 + k < nh;
                const uint32_t row = live ? tr[k] : cur;
                const CT vk = live ? vv[k] : (CT)0;
                if (row != cur) {  // a new run: hand the finished one on, restart from zero
                    if (cur != NONE) {
                        if (!excl) {
#pragma unroll
                            for (int v = 0; v < VPL; ++v)
                                atomic_add_out(outp + (uint64_t)cur * R + v, (OT)acc[v]);
                        } else if (in_first) {
                            first_row = cur;
#pragma unroll
                            for (int v = 0; v < VPL; ++v) facc[v] = acc[v];
                            in_first = false;
                        } else {
#pragma unroll
                            for (int v = 0; v < VPL; ++v) flush_owned(outp + (uint64_t)cur * R + v, acc[v], zeroed);
                        }
                    }
                    cur = row;
#pragma unroll
                    for (int v = 0; v < VPL; ++v) acc[v] = 0;
                }
                // acc += value * f_0 * ... * f_{NM1-1}, the last product fused
                // with the add (its first term, acc = 0, is the rounded product)
                CT g[VPL];
#pragma unroll
                for (int v = 0; v < VPL; ++v) g[v] = vk * f[0][k][v];
#pragma unroll
                for (int q = 1; q + 1 < NM1; ++q)
#pragma unroll
                    for (int v = 0; v < VPL; ++v) g[v] *= f[q][k][v];
#pragma unroll
                for (int v = 0; v < VPL; ++v) acc[v] = fused_madd(g[v], f[NM1 - 1][k][v], acc[v]);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (!excl) {
        if (cur != NONE) {
#pragma unroll
            for (int v = 0; v < VPL; ++v) atomic_add_out(outp + (uint64_t)cur * R + v, (OT)acc[v]);
        }
        return;
    }
    if (in_first) {
        first_row = cur;
#pragma unroll
        for (int v = 0; v < VPL; ++v) facc[v] = acc[v];
        cur = NONE;
    }
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
        s_part[0][gib][lig * VPL + v] = facc[v];
        s_part[1][gib][lig * VPL + v] = acc[v];
    }
    if (lig == 0) {
        s_row[0][gib] = first_row;
        s_row[1][gib] = cur;
    }
    __syncthreads();
    block_combine<OT, CT, R, G>(s_row, s_part, out, zeroed, bs, lb);
}

// Fiber variant of k_mttkrp_dec for output-ordered schedules whose rows are
// sorted by the other modes' indices (levels 1..NM1, leaf last).  Nonzeros
// that share the index prefix (target, l1, ..., lq) are adjacent, so the
// product factors like a CSF tree walk:
//   s      += val                       (equal coordinates: leaf gathered once)
//   t[leaf] += s * A_leaf[l_leaf]        at the end of a coordinate
//   t[q]    += t[q+1] * A_q[l_q]         at the end of a level-q fiber
//   out[n]  += t[1]                      at the end of the row
// Each nonzero carries the depth d of the first level whose index differs in
// the next nonzero (0: the row ends, NM1 + 1: same coordinates), computed
// from the two packed keys; a level's row is gathered only when its fiber
// ends, so gathers per nonzero drop from NM1 to (#fibers of all levels) /
// nnz.  Range boundaries end every fiber.  Row flushing and the boundary
// combine are those of k_mttkrp_dec in exclusive mode.

template <typename VT, typename FT, typename OT, int R, int PKW, int NM1, int VPL, bool ZEROED>
__global__ __launch_bounds__(256) void k_mttkrp_dec(const uint64_t *__restrict__ pkeys,
                                                    const VT *__restrict__ vals, uint64_t nnz,
                                                    FPtrs<FT> F, OT *__restrict__ out,
                                                    PackInfo pk, int target, uint64_t per_group,
                                                    int flags, BndSlots bs) {
    k_mttkrp_dec_body<VT, FT, OT, R, PKW, NM1, VPL, ZEROED>(xcd_remap(blockIdx.x, gridDim.x), pkeys, vals, nnz, F,
                                                             out, pk, target, per_group, flags, bs);
}

template <int PKW, int NM1>
__device__ __forceinline__ uint32_t fib_depth(uint64_t alo, uint64_t ahi, uint64_t blo,
                                              uint64_t bhi, const FibInfo &fi) {
    const uint64_t xl = alo ^ blo, xh = ahi ^ bhi;
    uint32_t d = NM1 + 1;
#pragma unroll
    for (int q = NM1; q >= 0; --q) {
        uint64_t x = xl & fi.mlo[q];
        if (PKW == 2) x |= xh & fi.mhi[q];
        if (x) d = q;
    }
    return d;
}

template <typename VT, typename FT, typename OT, int R, int PKW, int NM1, int VPL, int CB, bool ZEROED>
__device__ __forceinline__ void k_mttkrp_fib_body(unsigned lb, const uint64_t *__restrict__ pkeys,
                                                  const VT *__restrict__ vals, uint64_t nnz,
                                                  const FPtrs<FT> &F, OT *__restrict__ out,
                                                  const PackInfo &pk, int target, const FibInfo &fi,
                                                  uint64_t per_group, int flags, BndSlots bs) {
    using CT = typename FVec<FT>::CT;
    constexpr int L = R / VPL;
    constexpr int G = 256 / L;
    static_assert(R % VPL == 0 && L >= 1 && L <= 64 && (64 % L) == 0, "bad rank tiling");
    static_assert(L % CB == 0, "batch must divide the group");
    using Rec = DecRec<VT>;
    __shared__ Rec s_rec[256];
    __shared__ CT s_part[2][G][R];
    __shared__ uint32_t s_row[2][G];
    constexpr uint32_t NONE = 0xFFFFFFFFu;

    const uint64_t gtid = (uint64_t)lb * blockDim.x + threadIdx.x;
    const uint64_t group = gtid / L;
    const int gib = threadIdx.x / L;
    const int lig = threadIdx.x % L;
    const bool nt = flags & 1;
    constexpr bool zeroed = ZEROED;  // output zero on entry: owned rows are stored
    uint64_t start = group * per_group;
    uint64_t end = start + per_group < nnz ? start + per_group : nnz;
    if (start > nnz) start = nnz;

    int osh[NM1], ow[NM1];
    const char *fb[NM1];
#pragma unroll
    for (int q = 0; q < NM1; ++q) {
        const int m = fi.mode[q];
        osh[q] = pk.shift[m];
        ow[q] = pk.width[m];
        fb[q] = (const char *)F.p[m];
    }
    const int tsh = pk.shift[target], tw = pk.width[target];
    const uint32_t lane_b = (uint32_t)(lig * VPL * sizeof(FT));
    OT *outp = out + lig * VPL;
    Rec *myrec = s_rec + gib * L;

    // The group's first and last row partials go straight to LDS (slot 0:
    // first flushed row, slot 1: the row still open at the range end).
    bool in_first = true;
    if (lig == 0) {
        s_row[0][gib] = NONE;
        s_row[1][gib] = NONE;
    }
    CT t[NM1][VPL];
    CT sv = 0;
#pragma unroll
    for (int v = 0; v < VPL; ++v)
#pragma unroll
        for (int q = 0; q < NM1; ++q) t[q][v] = 0;
    uint64_t klo = 0, khi = 0, nlo = 0, nhi = 0;
    VT kval = 0;
    if (start < end) {
        const uint64_t i = start + lig < end ? start + lig : end - 1;
        const uint64_t j = i + 1 < nnz ? i + 1 : i;
        klo = stream_load(pkeys + i * PKW, nt);
        nlo = stream_load(pkeys + j * PKW, nt);
        if (PKW == 2) {
            khi = stream_load(pkeys + i * PKW + 1, nt);
            nhi = stream_load(pkeys + j * PKW + 1, nt);
        }
        kval = stream_load(vals + i, nt);
    }
    for (uint64_t base = start; base < end; base += L) {
        {
            Rec r;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                r.off[q] = q < NM1 ? (uint32_t)field<PKW>(klo, khi, osh[q < NM1 ? q : 0], ow[q < NM1 ? q : 0]) *
                                         (uint32_t)(R * sizeof(FT))
                                   : 0u;
            r.trow = (uint32_t)field<PKW>(klo, khi, tsh, tw);
            r.pad = base + lig + 1 >= end ? 0u : fib_depth<PKW, NM1>(klo, khi, nlo, nhi, fi);
            r.val = kval;
            myrec[lig] = r;
        }
        const uint64_t nb = base + L;
        if (nb < end) {  // prefetch the next step
            const uint64_t i = nb + lig < end ? nb + lig : end - 1;
            const uint64_t j = i + 1 < nnz ? i + 1 : i;
            klo = stream_load(pkeys + i * PKW, nt);
            nlo = stream_load(pkeys + j * PKW, nt);
            if (PKW == 2) {
                khi = stream_load(pkeys + i * PKW + 1, nt);
                nhi = stream_load(pkeys + j * PKW + 1, nt);
            }
            kval = stream_load(vals + i, nt);
        }
        __builtin_amdgcn_wave_barrier();
        const int nh = (end - base) < (uint64_t)L ? (int)(end - base) : L;
#pragma unroll
        for (int c0 = 0; c0 < L; c0 += CB) {
            uint32_t ob[CB][NM1], tr[CB], dd[CB];
            CT vv[CB];
#pragma unroll
            for (int k = 0; k < CB; ++k) {
                const Rec &r = myrec[c0 + k];
#pragma unroll
                for (int q = 0; q < NM1; ++q) ob[k][q] = r.off[q] + lane_b;
                tr[k] = r.trow;
                dd[k] = c0 + k < nh ? r.pad : (uint32_t)(NM1 + 2);
                vv[k] = (CT)r.val;
            }
            CT f[NM1][CB][VPL];
#pragma unroll
            for (int q = 0; q < NM1; ++q)
#pragma unroll
                for (int k = 0; k < CB; ++k)
                    if (dd[k] <= (uint32_t)(q + 1)) RowLoad<FT, VPL>::load(fb[q] + ob[k][q], f[q][k]);
#pragma unroll
            for (int k = 0; k < CB; ++k) {
                const uint32_t d = dd[k];
                if (d > (uint32_t)(NM1 + 1)) continue;
                sv += vv[k];
                if (d <= (uint32_t)NM1) {
#pragma unroll
                    for (int v = 0; v < VPL; ++v) t[NM1 - 1][v] += sv * f[NM1 - 1][k][v];
                    sv = 0;
                }
#pragma unroll
                for (int q = NM1 - 2; q >= 0; --q) {
                    if (d <= (uint32_t)(q + 1)) {
#pragma unroll
                        for (int v = 0; v < VPL; ++v) {
                            t[q][v] += t[q + 1][v] * f[q][k][v];
                            t[q + 1][v] = 0;
                        }
                    }
                }
                if (d == 0) {
                    const uint32_t row = tr[k];
                    if (in_first || base + c0 + k + 1 == end) {
                        const int slot = in_first ? 0 : 1;
#pragma unroll
                        for (int v = 0; v < VPL; ++v) s_part[slot][gib][lig * VPL + v] = t[0][v];
                        if (lig == 0) s_row[slot][gib] = row;
                        in_first = false;
                    } else {
#pragma unroll
                        for (int v = 0; v < VPL; ++v) flush_owned(outp + (uint64_t)row * R + v, t[0][v], zeroed);
                    }
#pragma unroll
                    for (int v = 0; v < VPL; ++v) t[0][v] = 0;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    block_combine<OT, CT, R, G>(s_row, s_part, out, zeroed, bs, lb);
}

template <typename VT, typename FT, typename OT, int R, int PKW, int NM1, int VPL, int CB, bool ZEROED>
__global__ __launch_bounds__(256) void k_mttkrp_fib(const uint64_t *__restrict__ pkeys,
                                                    const VT *__restrict__ vals, uint64_t nnz,
                                                    FPtrs<FT> F, OT *__restrict__ out,
                                                    PackInfo pk, int target, FibInfo fi,
                                                    uint64_t per_group, int flags, BndSlots bs) {
    k_mttkrp_fib_body<VT, FT, OT, R, PKW, NM1, VPL, CB, ZEROED>(xcd_remap(blockIdx.x, gridDim.x), pkeys, vals,
                                                                 nnz, F, out, pk, target, fi, per_group, flags, bs);
}

// elements per lane of the decode-once and fiber kernels (<= R): 4 FP64;
// FP32 8 from rank 16 and 4 at rank 8 (a rank-8 row is then one gather
// instruction across 2 lanes instead of two on one lane: half the distinct
// rows per instruction; 512-tensor batched CP-ALS 0.83 -> 0.81 s with ranks
// 8 and 16 both at 4, but at rank 16 the changed FP32 rounding moved
// borderline CP-stream convergence tests, so rank 16 keeps 8); BF16 8
__host__ __device__ constexpr int dec_vpl_n(int fbytes, int R) {
    const int want = fbytes == 8 ? 4 : (fbytes == 4 && R <= 8) ? 4 : 8;
    return want < R ? want : R;
}
template <typename FT, int R> constexpr int dec_vpl() { return dec_vpl_n((int)sizeof(FT), R); }

// Any rank: one wave per range, lane r covers ranks r, r+64, ...
template <typename VT, typename FT, typename OT, int PKW>
__global__ __launch_bounds__(256) void k_mttkrp_generic(const uint64_t *__restrict__ pkeys,
                                                        const VT *__restrict__ vals, uint64_t nnz,
                                                        FPtrs<FT> F, OT *__restrict__ out,
                                                        PackInfo pk, int target, int R,
                                                        uint64_t per_group) {
    using CT = typename FVec<FT>::CT;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    uint64_t start = wave * per_group;
    if (start >= nnz) return;
    uint64_t end = start + per_group < nnz ? start + per_group : nnz;
    const int nmode = pk.nmode;
    for (int r0 = 0; r0 < R; r0 += 64) {
        int r = r0 + lane;
        bool act = r < R;
        uint64_t cur = ~0ULL;
        CT acc = 0;
        for (uint64_t i = start; i < end; ++i) {
            uint64_t lo = pkeys[i * PKW], hi = PKW == 2 ? pkeys[i * PKW + 1] : 0;
            CT p = (CT)vals[i];
            for (int m = 0; m < nmode; ++m) {
                if (m == target) continue;
                uint64_t row = field<PKW>(lo, hi, pk.shift[m], pk.width[m]);
                if (act) {
                    FT fv = F.p[m][row * R + r];
                    CT c;
                    if constexpr (sizeof(FT) == 2) c = bf16_to_f32(fv);
                    else c = (CT)fv;
                    p *= c;
                }
            }
            uint64_t row = field<PKW>(lo, hi, pk.shift[target], pk.width[target]);
            if (row != cur) {
                if (cur != ~0ULL && act) atomic_add_out(out + cur * R + r, (OT)acc);
                cur = row;
                acc = p;
            } else {
                acc += p;
            }
        }
        if (cur != ~0ULL && act) atomic_add_out(out + cur * R + r, (OT)acc);
    }
}

struct MttkrpArgs {
    const uint64_t *pkeys;
    const void *vals;  // double for F64, float otherwise
    uint64_t nnz;
    const void *const *factors;
    void *out;
    PackInfo pk;
    int pkw;
    int target;
    int rank;
    int prec;
    bool out64;
    uint64_t per_group;
    int flags;
    bool decode_once;
    bool fiber;
    bool hot;       // hot-row kernel (run_hot)
    int hot_shape;  // its workgroup shape
    FibInfo fi;
    BndSlots bnd;   // FP64 boundary slots (FP32 output, output-ordered schedule)
};

template <typename VT, typename FT, typename OT, int R, int PKW, int VPL, int CB, bool Z>
static void launch_fib_z(const MttkrpArgs &a, const FPtrs<FT> &F, hipStream_t s) {
    constexpr int L = R / VPL;
    const uint64_t groups = (a.nnz + a.per_group - 1) / a.per_group;
    const unsigned blocks = cdiv(groups * L, 256);
    const int nm1 = a.pk.nmode - 1;
    if (nm1 == 2)
        k_mttkrp_fib<VT, FT, OT, R, PKW, 2, VPL, CB, Z><<<blocks, 256, 0, s>>>(
            a.pkeys, (const VT *)a.vals, a.nnz, F, (OT *)a.out, a.pk, a.target, a.fi, a.per_group,
            a.flags, a.bnd);
    else if (nm1 == 3)
        k_mttkrp_fib<VT, FT, OT, R, PKW, 3, VPL, CB, Z><<<blocks, 256, 0, s>>>(
            a.pkeys, (const VT *)a.vals, a.nnz, F, (OT *)a.out, a.pk, a.target, a.fi, a.per_group,
            a.flags, a.bnd);
    else
        k_mttkrp_fib<VT, FT, OT, R, PKW, 4, VPL, CB, Z><<<blocks, 256, 0, s>>>(
            a.pkeys, (const VT *)a.vals, a.nnz, F, (OT *)a.out, a.pk, a.target, a.fi, a.per_group,
            a.flags, a.bnd);
}

template <typename VT, typename FT, typename OT, int R, int PKW, int VPL, int CB>
static void launch_fib(const MttkrpArgs &a, const FPtrs<FT> &F, hipStream_t s) {
    if (a.flags & 32) launch_fib_z<VT, FT, OT, R, PKW, VPL, CB, true>(a, F, s);
    else launch_fib_z<VT, FT, OT, R, PKW, VPL, CB, false>(a, F, s);
}

template <typename VT, typename FT, typename OT, int R, int PKW>
static void launch_seg(const MttkrpArgs &a, hipStream_t s) {
    constexpr int V = FVec<FT>::V;
    constexpr int L = R / V;
    FPtrs<FT> F{};
    for (int m = 0; m < a.pk.nmode; ++m) F.p[m] = (const FT *)a.factors[m];
    uint64_t groups = (a.nnz + a.per_group - 1) / a.per_group;
    uint64_t threads = groups * L;
    unsigned blocks = cdiv(threads, 256);
    constexpr int B = 8;
    if (a.decode_once) {
        constexpr int VPL = dec_vpl<FT, R>();
        constexpr int LD = R / VPL;
        uint64_t dthreads = groups * LD;
        unsigned dblocks = cdiv(dthreads, 256);
        const int nm1 = a.pk.nmode - 1;
        auto dec = [&](auto zc) {
            constexpr bool Z = decltype(zc)::value;
            if (nm1 == 2)
                k_mttkrp_dec<VT, FT, OT, R, PKW, 2, VPL, Z><<<dblocks, 256, 0, s>>>(
                    a.pkeys, (const VT *)a.vals, a.nnz, F, (OT *)a.out, a.pk, a.target, a.per_group, a.flags,
                    a.bnd);
            else if (nm1 == 3)
                k_mttkrp_dec<VT, FT, OT, R, PKW, 3, VPL, Z><<<dblocks, 256, 0, s>>>(
                    a.pkeys, (const VT *)a.vals, a.nnz, F, (OT *)a.out, a.pk, a.target, a.per_group, a.flags,
                    a.bnd);
            else
                k_mttkrp_dec<VT, FT, OT, R, PKW, 4, VPL, Z><<<dblocks, 256, 0, s>>>(
                    a.pkeys, (const VT *)a.vals, a.nnz, F, (OT *)a.out, a.pk, a.target, a.per_group, a.flags,
                    a.bnd);
        };
        if (a.fiber) launch_fib<VT, FT, OT, R, PKW, VPL, 1>(a, F, s);
        else if (a.flags & 32) dec(std::true_type{});
        else dec(std::false_type{});
        if (a.bnd.row) {  // the workgroups' boundary rows, from their FP64 slots
            const uint64_t ns = 2 * (uint64_t)dblocks;
            launch_bnd_merge<OT, R, float>(a.bnd.row, a.bnd.val, ns, a.bnd.lrow, a.bnd.lval, (OT *)a.out,
                                           (a.flags & 32) ? 1 : 0, s);
        }
    } else
        k_mttkrp_seg<VT, FT, OT, R, PKW, B><<<blocks, 256, 0, s>>>(
            a.pkeys, (const VT *)a.vals, a.nnz, F, (OT *)a.out, a.pk, a.target, a.per_group, a.flags);
}

template <typename VT, typename FT, typename OT, int PKW>
static void launch_generic(const MttkrpArgs &a, hipStream_t s) {
    FPtrs<FT> F{};
    for (int m = 0; m < a.pk.nmode; ++m) F.p[m] = (const FT *)a.factors[m];
    uint64_t waves = (a.nnz + a.per_group - 1) / a.per_group;
    unsigned blocks = cdiv(waves * 64, 256);
    k_mttkrp_generic<VT, FT, OT, PKW><<<blocks, 256, 0, s>>>(
        a.pkeys, (const VT *)a.vals, a.nnz, F, (OT *)a.out, a.pk, a.target, a.rank, a.per_group);
}

template <typename VT, typename FT, typename OT, int PKW>
static int dispatch_rank(const MttkrpArgs &a, hipStream_t s) {
    constexpr int V = FVec<FT>::V;
    switch (a.rank) {
    case 8:
        if constexpr (8 % V == 0) { launch_seg<VT, FT, OT, 8, PKW>(a, s); return ALTO_OK; }
        break;
    case 16: launch_seg<VT, FT, OT, 16, PKW>(a, s); return ALTO_OK;
    case 32: launch_seg<VT, FT, OT, 32, PKW>(a, s); return ALTO_OK;
    case 64: launch_seg<VT, FT, OT, 64, PKW>(a, s); return ALTO_OK;
    default: break;
    }
    launch_generic<VT, FT, OT, PKW>(a, s);
    return ALTO_OK;
}

template <int PKW>
static int dispatch_prec(const MttkrpArgs &a, hipStream_t s) {
    if (a.prec == ALTO_PREC_F64) return dispatch_rank<double, double, double, PKW>(a, s);
    if (a.prec == ALTO_PREC_F32)
        return a.out64 ? dispatch_rank<float, float, double, PKW>(a, s)
                       : dispatch_rank<float, float, float, PKW>(a, s);
    if (a.prec == ALTO_PREC_BF16)
        return a.out64 ? dispatch_rank<float, uint16_t, double, PKW>(a, s)
                       : dispatch_rank<float, uint16_t, float, PKW>(a, s);
    return ALTO_ERR_ARG;
}

// Autotuner: for the default configuration of an output-ordered MTTKRP with
// at least kTuneMinEntries entries, the first call of each bucket (modes,
// rank, precision, output type, key words, log2 of the entry count, log2 of
// the output rows) times the candidates -- segment length {p/4, p/2, p, 2p}
// around the heuristic p, fiber or decode-once kernel -- best of five runs
// each on a scratch output.  The heuristic choice is kept unless a candidate
// is more than 1% faster.  ALTO_AUTOTUNE=0 in the environment turns it off.
constexpr uint64_t kTuneMinEntries = 1ULL << 20;

struct TuneKey {
    int nmode, rank, prec, out64, pkw, lg, lr;
    bool operator<(const TuneKey &o) const {
        return std::tie(nmode, rank, prec, out64, pkw, lg, lr) <
               std::tie(o.nmode, o.rank, o.prec, o.out64, o.pkw, o.lg, o.lr);
    }
};
struct TuneVal {
    uint64_t per_group;
    bool fiber;
    bool hot;
    int hot_shape;
};
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneVal> g_tune;

static bool autotune_enabled() {
    static const bool on = [] {
        const char *e = getenv("ALTO_AUTOTUNE");
        return !(e && e[0] == '0');
    }();
    return on;
}

// One launch of the configuration in `a`: the hot-row kernel or the range
// kernels.
static int launch_cfg(TensorImpl *t, const MttkrpArgs &a0, hipStream_t s) {
    if (a0.hot)
        return run_hot(t, a0.target, a0.prec, a0.rank, a0.factors, a0.out, a0.out64, (a0.flags & 32) != 0,
                       a0.hot_shape, s);
    MttkrpArgs a = a0;
    a.bnd = BndSlots{};
    const int R = a.rank;
    if (!a.out64 && (a.flags & 4) && a.decode_once && (R == 8 || R == 16 || R == 32 || R == 64)) {
        // FP32 output of the decode-once / fiber kernels: the workgroups'
        // boundary rows go through FP64 slots instead of FP32 atomics
        const int vpl = dec_vpl_n(a.prec == ALTO_PREC_F64 ? 8 : a.prec == ALTO_PREC_F32 ? 4 : 2, R);
        const uint64_t groups = (a.nnz + a.per_group - 1) / a.per_group;
        const uint64_t ns = 2 * (uint64_t)cdiv(groups * (R / vpl), 256);
        const uint64_t nl = bnd_level_slots(ns);  // slots of the merge's level buffers
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        const size_t rb = al(ns * sizeof(uint32_t)), vb = al(ns * R * sizeof(float));
        const size_t lrb = al(nl * sizeof(uint32_t)), lvb = al(nl * R * sizeof(double));
        ALTO_TRY(t->bnd_slots.alloc(rb + vb + 2 * (lrb + lvb)));
        char *base = (char *)t->bnd_slots.p;
        a.bnd.row = (uint32_t *)base;
        a.bnd.val = (float *)(base + rb);
        for (int k = 0; k < 2; ++k) {
            a.bnd.lrow[k] = (uint32_t *)(base + rb + vb + k * (lrb + lvb));
            a.bnd.lval[k] = (double *)(base + rb + vb + k * (lrb + lvb) + lrb);
        }
    }
    return t->pkw == 1 ? dispatch_prec<1>(a, s) : dispatch_prec<2>(a, s);
}

static int autotune(TensorImpl *t, MttkrpArgs &a, hipStream_t s) {
    int lg = 0, lr = 0;
    while ((1ULL << (lg + 1)) <= a.nnz) ++lg;
    while (lr < 63 && (1ULL << (lr + 1)) <= t->dims[a.target]) ++lr;
    const TuneKey key{t->nmode, a.rank, a.prec, a.out64 ? 1 : 0, t->pkw, lg, lr};
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            a.per_group = it->second.per_group;
            a.fiber = it->second.fiber;
            a.hot = it->second.hot;
            a.hot_shape = it->second.hot_shape;
            return ALTO_OK;
        }
    }
    if (capture_flag()) {  // timing launches cannot run inside a capture
        set_error("MTTKRP launch tuning requested inside a graph capture");
        return ALTO_ERR_HIP;
    }
    const uint64_t rows = t->dims[a.target];
    const size_t osz = a.out64 ? 8 : 4;
    DevBuf scratch;
    ALTO_TRY(scratch.alloc(osz * rows * a.rank));
    hipEvent_t e0, e1;
    ALTO_HIP_TRY(hipEventCreate(&e0));
    ALTO_HIP_TRY(hipEventCreate(&e1));
    MttkrpArgs c = a;
    c.out = scratch.p;
    const uint64_t p0 = a.per_group;
    // segment lengths p/4, p/2, p, 2p (at least 2, at most 128; repeats dropped)
    uint64_t pgs[4];
    int npg = 0;
    for (uint64_t cand : {p0 / 4, p0 / 2, p0, p0 * 2}) {
        cand = cand < 2 ? 2 : cand > 128 ? 128 : cand;
        bool seen = false;
        for (int i = 0; i < npg; ++i) seen = seen || pgs[i] == cand;
        if (!seen) pgs[npg++] = cand;
    }
    int st = ALTO_OK;
    // best of five runs of one configuration
    auto time_cfg = [&](uint64_t pg, bool fib, bool hot = false, int shape = 0) -> float {
        c.per_group = pg;
        c.fiber = fib;
        c.hot = hot;
        c.hot_shape = shape;
        float best_ms = 1e30f;
        for (int rep = 0; rep < 5 && st == ALTO_OK; ++rep) {
            float ms = 0.f;
            if (hipEventRecord(e0, s) != hipSuccess) st = ALTO_ERR_HIP;
            if (st == ALTO_OK) st = launch_cfg(t, c, s);
            if (st == ALTO_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
                st = ALTO_ERR_HIP;
            if (st == ALTO_OK) (void)hipEventElapsedTime(&ms, e0, e1);
            best_ms = ms < best_ms ? ms : best_ms;
        }
        return best_ms;
    };
    // the heuristic configuration stays unless another is more than 1% faster
    TuneVal pick{a.per_group, a.fiber, false, 0};
    const float base_ms = time_cfg(a.per_group, a.fiber);
    float best = base_ms;
    for (int fib = 0; fib < 2 && st == ALTO_OK; ++fib) {
        // BF16 rows: the fiber walk measured about 1.5x slower than decode-once
        // on a 4-mode Zipf tensor (1e8 entries, rank 32), so it is not timed
        if (fib && a.prec == ALTO_PREC_BF16) continue;
        for (int i = 0; i < npg && st == ALTO_OK; ++i) {
            if (pgs[i] == a.per_group && (fib != 0) == a.fiber) continue;
            const float ms = time_cfg(pgs[i], fib != 0);
            if (st == ALTO_OK && ms < best && ms < 0.99f * base_ms) {
                best = ms;
                pick = TuneVal{pgs[i], fib != 0, false, 0};
            }
        }
    }
    // the hot-row kernel in each workgroup shape (flags bit 8 excludes it)
    if (!(a.flags & 256) && hot_supported(t, a.prec, a.rank)) {
        for (int shape = 0; shape < 3 && st == ALTO_OK; ++shape) {
            const float ms = time_cfg(a.per_group, false, true, shape);
            if (st == ALTO_OK && ms < best && ms < 0.99f * base_ms) {
                best = ms;
                pick = TuneVal{a.per_group, false, true, shape};
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ALTO_TRY(st);
    ALTO_LAUNCH_CHECK();
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune.emplace(key, pick);
    }
    a.per_group = pick.per_group;
    a.fiber = pick.fiber;
    a.hot = pick.hot;
    a.hot_shape = pick.hot_shape;
    return ALTO_OK;
}

// Builds the output-ordered schedule and runs the autotuner for a default
// MTTKRP of `target` (algo 2) without touching any output.  Allocates nothing
// once both are done (a free would synchronise the device).
int prepare_mttkrp(TensorImpl *t, int target, int prec, int rank, const void *const *factors,
                   int out_f64, hipStream_t s) {
    return run_mttkrp(t, target, prec, rank, factors, nullptr, out_f64, 2, 0, -1, s, true, true);
}

// The launch arguments of an MTTKRP (schedule built on demand); *empty when
// there is nothing to compute; *defaults when per_group and flags are the
// defaults (the tuner may then replace them).
static int make_args(TensorImpl *t, int target, int prec, int rank, const void *const *factors, void *out,
                     int out_f64, int algo, uint64_t per_group, int flags, bool out_zeroed, MttkrpArgs &a,
                     bool *defaults_out, bool *empty, hipStream_t s) {
    *empty = false;
    if (algo == 0) algo = 2;
    a.nnz = t->nnz;
    a.factors = factors;
    a.out = out;
    a.pk = t->pack;
    a.pkw = t->pkw;
    a.target = target;
    a.rank = rank;
    a.prec = prec;
    a.out64 = prec == ALTO_PREC_F64 || out_f64;
    a.per_group = per_group;  // 0: chosen below from the entry count
    a.flags = flags < 0 ? 3 : flags;
    if (algo == 2) a.flags |= 4;  // output-ordered: runs inside a range own their row
    if (algo == 2 && out_zeroed) a.flags |= 32;
    if (algo != 2) a.flags &= ~32;  // owned rows exist only in output-ordered ranges
    // decode-once kernel: nmode <= 5 and 32-bit element offsets into every factor
    a.decode_once = t->nmode >= 3 && t->nmode <= 5 && !(a.flags & 8);
    for (int m = 0; m < t->nmode; ++m)
        if (t->dims[m] * (uint64_t)rank * (prec == ALTO_PREC_F64 ? 8 : prec == ALTO_PREC_F32 ? 4 : 2) >=
            (1ULL << 32))
            a.decode_once = false;
    if (algo == 1) {
        a.pkeys = t->pkeys.as<uint64_t>();
        if (prec == ALTO_PREC_F64) {
            a.vals = t->vals64.p;
        } else {
            ALTO_TRY(ensure_vals32(t, s));
            a.vals = t->vals32.p;
        }
    } else if (algo == 2) {
        ALTO_TRY(ensure_schedule(t, target, prec == ALTO_PREC_F64, s));
        const ModeSchedule &sc = t->sched[target];
        a.nnz = sc.n;
        if (a.nnz == 0) {
            *empty = true;
            return ALTO_OK;
        }
        a.pkeys = sc.pkeys.as<uint64_t>();
        a.vals = prec == ALTO_PREC_F64 ? sc.vals64.p : sc.vals32.p;
        // fiber kernel: FP64 and FP32 factors by default, BF16 only on request
        // (flags bit 6) or by the tuner's timing
        a.fiber = a.decode_once && !(a.flags & 16) && (prec != ALTO_PREC_BF16 || (a.flags & 64));
        auto bits = [&](int m, uint64_t *lo, uint64_t *hi) {
            const int sh = t->pack.shift[m], w = t->pack.width[m];
            unsigned __int128 mk = w ? ((((unsigned __int128)1) << w) - 1) << sh : 0;
            *lo = (uint64_t)mk;
            *hi = (uint64_t)(mk >> 64);
        };
        bits(target, &a.fi.mlo[0], &a.fi.mhi[0]);
        for (int q = 0; q < t->nmode - 1; ++q) {
            a.fi.mode[q] = sc.lvl[q];
            bits(sc.lvl[q], &a.fi.mlo[q + 1], &a.fi.mhi[q + 1]);
        }
    } else {
        return ALTO_ERR_ARG;
    }
    *defaults_out = a.per_group == 0 && flags < 0;
    if (a.per_group == 0) {
        // 64 entries per lane group, fewer when that would leave the device
        // short of about 2^18 threads (small tensors are latency bound: a
        // group walks its entries one after another)
        const uint64_t vpl = prec == ALTO_PREC_F64 ? 4 : 8;  // elements per lane
        uint64_t lanes = (uint64_t)rank / vpl;
        if (lanes < 1) lanes = 1;
        uint64_t pg = a.nnz * lanes / (1ULL << 18);
        a.per_group = pg < 4 ? 4 : pg > 64 ? 64 : pg;
    }
    return ALTO_OK;
}

int run_mttkrp(TensorImpl *t, int target, int prec, int rank, const void *const *factors,
               void *out, int out_f64, int algo, uint64_t per_group, int flags, hipStream_t s,
               bool prepare_only, bool out_zeroed) {
    if (target < 0 || target >= t->nmode || rank < 1 || rank > 4096 || !factors ||
        (!out && !prepare_only))
        return ALTO_ERR_ARG;
    if (prec != ALTO_PREC_F64 && prec != ALTO_PREC_F32 && prec != ALTO_PREC_BF16) return ALTO_ERR_ARG;
    for (int m = 0; m < t->nmode; ++m)
        if (m != target && !factors[m]) return ALTO_ERR_ARG;
    if (t->nnz == 0) return ALTO_OK;
    if (algo == 0) algo = 2;
    MttkrpArgs a{};
    bool defaults = false, empty = false;
    ALTO_TRY(make_args(t, target, prec, rank, factors, out, out_f64, algo, per_group, flags, out_zeroed, a,
                       &defaults, &empty, s));
    if (empty) return ALTO_OK;
    // flags bit 7: the hot-row kernel, workgroup shape in bits 9-10
    if (algo == 2 && (a.flags & 128)) {
        if (!hot_supported(t, prec, rank)) {
            set_error("the hot-row MTTKRP takes ranks 16, 32, 64 and 3 to 5 modes");
            return ALTO_ERR_UNSUPPORTED;
        }
        a.hot = true;
        a.hot_shape = (a.flags >> 9) & 3;
    }
    if (defaults && algo == 2 && a.decode_once && a.nnz >= kTuneMinEntries && autotune_enabled())
        ALTO_TRY(autotune(t, a, s));
    if (a.hot && out && ((uintptr_t)out % 16) && !(a.flags & 128))
        a.hot = false;  // tuned choice, output not 16-byte aligned: the range kernels take it
    if (prepare_only) {
        if (a.hot) {  // the re-encoded keys are built once, outside any timed call
            ALTO_TRY(ensure_hot_schedule(t, target, rank * (prec == ALTO_PREC_F64 ? 8 : prec == ALTO_PREC_F32 ? 4 : 2), s));
        }
        return ALTO_OK;
    }
    int st = launch_cfg(t, a, s);
    ALTO_TRY(st);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

// ---- grouped MTTKRP: one launch over many tensors' output-ordered schedules ----

// The fiber kernel of MttkrpJob j runs on grouped blocks [first[j], first[j+1]).
// DEC: the decode-once kernel, else the fiber kernel.
template <typename VT, typename FT, int R, int NM1, bool DEC>
__global__ __launch_bounds__(256) void k_mttkrp_group(const MttkrpJob *__restrict__ jobs,
                                                      const unsigned *__restrict__ first, int njobs) {
    constexpr int VPL = dec_vpl<FT, R>();
    const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
    const int j = find_job(first, njobs, b);
    // The job stays in global memory: the factor pointers and the key
    // fields are read once at the start, by mode index; the fiber levels are
    // copied to registers because the kernel reads them for every entry
    // (through a global-memory reference it would reload them after every
    // output store it cannot prove unrelated).
    const MttkrpJob &J = jobs[j];
    if (!*J.active) return;
    const FPtrs<FT> &F = *reinterpret_cast<const FPtrs<FT> *>(J.F);
    if constexpr (DEC) {
        k_mttkrp_dec_body<VT, FT, double, R, 1, NM1, VPL, true>(b - first[j], J.pkeys, (const VT *)J.vals, J.nnz, F,
                                                                (double *)J.out, J.pk, J.target, J.per_group,
                                                                J.flags, BndSlots{});
    } else {
        const FibInfo fi = J.fi;
        k_mttkrp_fib_body<VT, FT, double, R, 1, NM1, VPL, 1, true>(b - first[j], J.pkeys, (const VT *)J.vals,
                                                                   J.nnz, F, (double *)J.out, J.pk, J.target, fi,
                                                                   J.per_group, J.flags, BndSlots{});
    }
}

bool mttkrp_group_supported(const TensorImpl *t, int prec, int rank) {
    if (t->pkw != 1 || (t->nmode != 3 && t->nmode != 4)) return false;
    if (prec != ALTO_PREC_F64 && prec != ALTO_PREC_F32) return false;
    if (rank != 8 && rank != 16 && rank != 32 && rank != 64) return false;
    for (int m = 0; m < t->nmode; ++m)
        if (t->dims[m] * (uint64_t)rank * (prec == ALTO_PREC_F64 ? 8 : 4) >= (1ULL << 32)) return false;
    return true;
}

int mttkrp_group_job(TensorImpl *t, int target, int prec, int rank, const void *const *factors, void *out,
                     const int *active, MttkrpJob *job, hipStream_t s) {
    if (!mttkrp_group_supported(t, prec, rank)) return ALTO_ERR_UNSUPPORTED;
    MttkrpArgs a{};
    bool defaults = false, empty = false;
    ALTO_TRY(make_args(t, target, prec, rank, factors, out, 1, 2, 0, -1, true, a, &defaults, &empty, s));
    *job = MttkrpJob{};
    job->active = active;
    job->out = out;
    job->nblocks = 0;
    if (empty || t->nnz == 0) return ALTO_OK;
    if (!a.decode_once) return ALTO_ERR_UNSUPPORTED;
    // fiber kernel when a (target, first level) fiber holds about two or more
    // entries on average (uniform estimate), the decode-once kernel otherwise
    const double cells = (double)t->dims[target] * (double)t->dims[a.fi.mode[0]];
    job->dec = (double)a.nnz < 2.0 * cells ? 1 : 0;
    job->pkeys = a.pkeys;
    job->vals = a.vals;
    job->nnz = a.nnz;
    for (int m = 0; m < t->nmode; ++m) job->F[m] = factors[m];
    job->pk = a.pk;
    job->target = target;
    job->fi = a.fi;
    // In a grouped launch the tensors together fill the device, so each takes
    // the segment length that is fastest per entry (8 entries per lane group
    // measured best from rank 8 to 32 at 5e5 to 5e6 entries), not the one
    // that fills the device with one tensor.
    job->per_group = 8;
    job->flags = a.flags;
    const int vpl = dec_vpl_n(prec == ALTO_PREC_F64 ? 8 : prec == ALTO_PREC_F32 ? 4 : 2, rank);
    const uint64_t groups = (a.nnz + job->per_group - 1) / job->per_group;
    job->nblocks = cdiv(groups * (uint64_t)(rank / vpl), 256);
    return ALTO_OK;
}

template <typename VT, typename FT, int R, bool DEC>
static void launch_group_r(const MttkrpJob *jobs, const unsigned *first, int njobs, unsigned nblocks, int nmode,
                           hipStream_t s) {
    if (nmode == 3) k_mttkrp_group<VT, FT, R, 2, DEC><<<nblocks, 256, 0, s>>>(jobs, first, njobs);
    else k_mttkrp_group<VT, FT, R, 3, DEC><<<nblocks, 256, 0, s>>>(jobs, first, njobs);
}

template <typename VT, typename FT, bool DEC>
static int launch_group_p(const MttkrpJob *jobs, const unsigned *first, int njobs, unsigned nblocks, int rank,
                          int nmode, hipStream_t s) {
    switch (rank) {
    case 8: launch_group_r<VT, FT, 8, DEC>(jobs, first, njobs, nblocks, nmode, s); break;
    case 16: launch_group_r<VT, FT, 16, DEC>(jobs, first, njobs, nblocks, nmode, s); break;
    case 32: launch_group_r<VT, FT, 32, DEC>(jobs, first, njobs, nblocks, nmode, s); break;
    case 64: launch_group_r<VT, FT, 64, DEC>(jobs, first, njobs, nblocks, nmode, s); break;
    default: return ALTO_ERR_UNSUPPORTED;
    }
    return ALTO_OK;
}

int launch_mttkrp_group(const MttkrpJob *jobs, const unsigned *first, int njobs, unsigned nblocks, int prec,
                        int rank, int nmode, bool dec, hipStream_t s) {
    if (njobs <= 0 || nblocks == 0) return ALTO_OK;
    if (nmode != 3 && nmode != 4) return ALTO_ERR_UNSUPPORTED;
    int st;
    if (prec == ALTO_PREC_F64)
        st = dec ? launch_group_p<double, double, true>(jobs, first, njobs, nblocks, rank, nmode, s)
                 : launch_group_p<double, double, false>(jobs, first, njobs, nblocks, rank, nmode, s);
    else if (prec == ALTO_PREC_F32)
        st = dec ? launch_group_p<float, float, true>(jobs, first, njobs, nblocks, rank, nmode, s)
                 : launch_group_p<float, float, false>(jobs, first, njobs, nblocks, rank, nmode, s);
    else
        st = ALTO_ERR_UNSUPPORTED;
    ALTO_TRY(st);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

// Reference COO MTTKRP (src/mttkrp.cpp:122-156) on the device, FP64.
struct CooPtrs {
    const uint64_t *p[kMaxModes];
};

__global__ __launch_bounds__(256) void k_mttkrp_coo(CooPtrs ci, const double *__restrict__ vals,
                                                    uint64_t nnz, FPtrs<double> F, int nmode,
                                                    int target, int R, double *__restrict__ out) {
    uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    int lane = threadIdx.x & 63;
    if (w >= nnz) return;
    for (int r = lane; r < R; r += 64) {
        double p = vals[w];
        for (int m = 0; m < nmode; ++m)
            if (m != target) p *= F.p[m][ci.p[m][w] * R + r];
        unsafeAtomicAdd(out + ci.p[target][w] * R + r, p);
    }
}

}  // namespace alto

using namespace alto;

extern "C" {

// The output must hold dims[target] x rank elements of what the kernels
// write: FP64 for ALTO_PREC_F64 or out_f64, else FP32.
static int check_out_bytes(alto_hip_tensor t, int target, int prec, int rank, int out_f64,
                           size_t out_bytes) {
    if (!t || target < 0 || target >= t->impl.nmode || rank < 1) return ALTO_ERR_ARG;
    const size_t esz = (prec == ALTO_PREC_F64 || out_f64) ? 8 : 4;
    const size_t need = (size_t)t->impl.dims[target] * (size_t)rank * esz;
    if (out_bytes < need) {
        set_error("MTTKRP output holds %zu bytes, mode %d needs %llu x %d x %zu = %zu", out_bytes, target,
                  (unsigned long long)t->impl.dims[target], rank, esz, need);
        return ALTO_ERR_ARG;
    }
    return ALTO_OK;
}

int alto_hip_mttkrp_algo(alto_hip_tensor t, int target, int prec, int rank,
                         const void *const *factors, void *out, int out_f64, size_t out_bytes, int algo,
                         void *stream) {
    ALTO_TRY(check_out_bytes(t, target, prec, rank, out_f64, out_bytes));
    ALTO_TRY(ensure_init());
    return run_mttkrp(&t->impl, target, prec, rank, factors, out, out_f64, algo, 0, -1,
                      to_stream(stream));
}

int alto_hip_mttkrp_config(alto_hip_tensor t, int target, int prec, int rank,
                           const void *const *factors, void *out, int out_f64, size_t out_bytes, int algo,
                           uint64_t segment, int flags, void *stream) {
    ALTO_TRY(check_out_bytes(t, target, prec, rank, out_f64, out_bytes));
    ALTO_TRY(ensure_init());
    return run_mttkrp(&t->impl, target, prec, rank, factors, out, out_f64, algo, segment, flags,
                      to_stream(stream));
}

int alto_hip_mttkrp(alto_hip_tensor t, int target, int prec, int rank, const void *const *factors,
                    void *out, int out_f64, size_t out_bytes, void *stream) {
    return alto_hip_mttkrp_algo(t, target, prec, rank, factors, out, out_f64, out_bytes, 0, stream);
}

int alto_hip_mttkrp_coo(const alto_hip_coo *coo, int target, int rank,
                        const double *const *factors, double *out, void *stream) {
    if (!coo || !coo->on_device || !factors || !out || rank < 1) return ALTO_ERR_ARG;
    if (target < 0 || target >= coo->nmode || coo->nmode > kMaxModes) return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    if (coo->nnz == 0) return ALTO_OK;
    CooPtrs ci{};
    FPtrs<double> F{};
    for (int m = 0; m < coo->nmode; ++m) {
        ci.p[m] = coo->cidx[m];
        F.p[m] = factors[m];
    }
    k_mttkrp_coo<<<cdiv(coo->nnz * 64, 256), 256, 0, to_stream(stream)>>>(
        ci, coo->vals, coo->nnz, F, coo->nmode, target, rank, out);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

}  // extern "C"
