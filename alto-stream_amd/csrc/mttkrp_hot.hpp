// Hot-row MTTKRP kernels (templates); instantiated per precision in
// mttkrp_hot_f32.hip, mttkrp_hot_bf16.hip and mttkrp_hot_f64.hip, driven by
// mttkrp_hot.hip.
#pragma once
#include <algorithm>

#include "mttkrp.hpp"
#include "tensor.hpp"

namespace alto {

constexpr uint32_t kHotFlag = 0x80000000u;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kHotThreads = 1024;
constexpr int kHotBlock = 16;  // group ranges are multiples of this many entries
constexpr size_t kHotLds = 163840;
// ---------------------------------------------------------------------------
// The kernel.

template <typename FT> struct HotRow;  // one 16-byte chunk -> elements
template <> struct HotRow<float> {
    static constexpr int E16 = 4;
    __device__ __forceinline__ static void cvt(uint4 u, float *o) {
        o[0] = __uint_as_float(u.x);
        o[1] = __uint_as_float(u.y);
        o[2] = __uint_as_float(u.z);
        o[3] = __uint_as_float(u.w);
    }
};
template <> struct HotRow<uint16_t> {
    static constexpr int E16 = 8;
    __device__ __forceinline__ static void cvt(uint4 u, float *o) {
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            o[2 * q] = __uint_as_float(w[q] << 16);
            o[2 * q + 1] = __uint_as_float(w[q] & 0xffff0000u);
        }
    }
};
template <> struct HotRow<double> {
    static constexpr int E16 = 2;
    __device__ __forceinline__ static void cvt(uint4 u, double *o) {
        o[0] = __hiloint2double((int)u.y, (int)u.x);
        o[1] = __hiloint2double((int)u.w, (int)u.z);
    }
};

// Broadcast of lane k (compile time) of each group of L lanes to the group.
template <int L, int K> __device__ __forceinline__ uint32_t gbcast(uint32_t x) {
    static_assert(K < L, "lane");
    if constexpr (L == 1) {
        return x;
    } else if constexpr (L == 2) {
        constexpr int c = K | (K << 2) | ((2 + K) << 4) | ((2 + K) << 6);
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, c, 0xF, 0xF, false);
    } else if constexpr (L == 4) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, K * 0x55, 0xF, 0xF, false);
    } else {
        static_assert(L == 8 || L == 16, "group size");
        constexpr int pat = (0x1F & ~(L - 1)) | (K << 5);  // bitmask mode
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, pat);
    }
}

template <typename T> struct Bits32;
template <> struct Bits32<float> {
    static constexpr int N = 1;
    __device__ __forceinline__ static void split(float v, uint32_t *w) { w[0] = __float_as_uint(v); }
    __device__ __forceinline__ static float join(const uint32_t *w) { return __uint_as_float(w[0]); }
};
template <> struct Bits32<double> {
    static constexpr int N = 2;
    __device__ __forceinline__ static void split(double v, uint32_t *w) {
        const uint64_t u = (uint64_t)__double_as_longlong(v);
        w[0] = (uint32_t)u;
        w[1] = (uint32_t)(u >> 32);
    }
    __device__ __forceinline__ static double join(const uint32_t *w) {
        return __hiloint2double((int)w[1], (int)w[0]);
    }
};

struct HotArgs {
    int fsh[4], fw[4];  // gathered fields in the re-encoded keys (q-th mode other than the target)
    int tsh, tw;        // target field
    uint32_t dim[4];    // I_m of the gathered modes
    uint32_t lds[4];    // LDS byte offset of each mode's hot table
    uint32_t K[4];      // hot slots per gathered mode
    const uint32_t *list[4];
    const void *fac[4];
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// E entries' keys and values from position pos, with no bounds branch: the
// arrays carry kSchedPad entries of padding after their end (a branch here
// would leave the loads pending on one path, and the compiler then drains
// every load before the first gather batch's products).
template <int PKW, int E, bool NT>
__device__ __forceinline__ void load_keys(const uint64_t *__restrict__ p, uint64_t pos, uint64_t *k) {
    constexpr int NW = E * PKW;
    if constexpr ((NW % 2) == 0) {
        const u32x4 *q = (const u32x4 *)(p + pos * PKW);
#pragma unroll
        for (int i = 0; i < NW / 2; ++i) {
            const u32x4 u = NT ? __builtin_nontemporal_load(q + i) : q[i];
            k[2 * i] = (uint64_t)u.x | ((uint64_t)u.y << 32);
            k[2 * i + 1] = (uint64_t)u.z | ((uint64_t)u.w << 32);
        }
    } else {
#pragma unroll
        for (int w = 0; w < NW; ++w) k[w] = NT ? __builtin_nontemporal_load(p + pos * PKW + w) : p[pos * PKW + w];
    }
}

template <typename VT, int E, bool NT>
__device__ __forceinline__ void load_vals(const VT *__restrict__ p, uint64_t pos, VT *v) {
    constexpr int PER16 = 16 / (int)sizeof(VT);
    if constexpr ((E % PER16) == 0) {
        const u32x4 *q = (const u32x4 *)(p + pos);
#pragma unroll
        for (int i = 0; i < E / PER16; ++i) {
            const u32x4 u = NT ? __builtin_nontemporal_load(q + i) : q[i];
            const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int j = 0; j < PER16; ++j) {
                uint32_t b[2] = {w[j * Bits32<VT>::N], Bits32<VT>::N == 2 ? w[j * 2 + 1] : 0u};
                v[i * PER16 + j] = Bits32<VT>::join(b);
            }
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = NT ? __builtin_nontemporal_load(p + pos + e) : p[pos + e];
    }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// A lane's VPL output elements of an owned row (16-byte aligned): stored when
// the output was zero on entry, else read, added and stored, 16 bytes at a
// time; the branch is uniform.
template <typename OT, typename CT, int VPL>
__device__ __forceinline__ void flush_lane(OT *p, const CT *v, bool zeroed) {
    using V = std::conditional_t<sizeof(OT) == 8, f64x2, f32x4>;
    constexpr int PER = 16 / (int)sizeof(OT);
    static_assert(VPL % PER == 0, "vector flush");
    V *q = (V *)p;
    if (zeroed) {
#pragma unroll
        for (int i = 0; i < VPL / PER; ++i) {
            V x;
#pragma unroll
            for (int j = 0; j < PER; ++j) x[j] = (OT)v[i * PER + j];
            q[i] = x;
        }
    } else {
        V x[VPL / PER];
#pragma unroll
        for (int i = 0; i < VPL / PER; ++i) x[i] = q[i];
#pragma unroll
        for (int i = 0; i < VPL / PER; ++i) {
#pragma unroll
            for (int j = 0; j < PER; ++j) x[i][j] += (OT)v[i * PER + j];
            q[i] = x[i];
        }
    }
}

template <int PKW>
__device__ __forceinline__ uint32_t hfield(const uint64_t *k, int sh, int w) {
    if (PKW == 1) return (uint32_t)((k[0] >> sh) & ((1ULL << w) - 1ULL));
    return (uint32_t)extract_field(k[0], k[1], sh, w);
}

template <typename OT>
__device__ __forceinline__ void store_final(OT *p, double v, bool zeroed) {
    *p = zeroed ? (OT)v : (OT)((double)*p + v);
}

// Two-level run combine of 2G ordered slots (row, R partials) held in LDS by
// NT threads: P = NT / R threads per column each reduce a chunk of slots,
// storing the runs that start and end inside the chunk when `own_inner`;
// then one thread per column merges the chunks' first and last runs.
// first/last: where the overall first and last runs go (rows NONE when
// absent; row entry row_stride * column, so 0 = one shared row word); the
// runs between them are stored into out.
template <typename OT, int R, int NT>
__device__ __forceinline__ void combine_runs(const uint32_t *rows, const double *parts, int nslots,
                                             char *scratch, OT *out, bool zeroed, uint32_t *first_row,
                                             double *first_val, uint32_t *last_row, double *last_val,
                                             int row_stride) {
    constexpr int P = NT / R;
    static_assert(NT % R == 0, "combine tiling");
    uint32_t *m_row = (uint32_t *)scratch;                         // [2][R][P]
    double *m_sum = (double *)(scratch + ((8 * R * P + 15) & ~15));  // [2][R][P]
    const int col = threadIdx.x / P, j = threadIdx.x % P;
    const int C = (nslots + P - 1) / P;
    OT *o = out + col;
    uint32_t frow = kNone, crow = kNone;
    double fsum = 0, csum = 0;
    int runs = 0;
    for (int k = 0; k < C; ++k) {
        const int e = j * C + k;
        if (e >= nslots) break;
        const uint32_t row = rows[e];
        if (row == kNone) continue;
        const double x = parts[(size_t)e * R + col];
        if (row == crow) {
            csum += x;
            continue;
        }
        if (runs == 1) {
            frow = crow;
            fsum = csum;
        } else if (runs > 1) {
            store_final(o + (uint64_t)crow * R, csum, zeroed);
        }
        crow = row;
        csum = x;
        ++runs;
    }
    if (runs == 1) {
        frow = crow;
        fsum = csum;
        crow = kNone;
    }
    m_row[(0 * R + col) * P + j] = frow;
    m_sum[(0 * R + col) * P + j] = fsum;
    m_row[(1 * R + col) * P + j] = crow;
    m_sum[(1 * R + col) * P + j] = csum;
    __syncthreads();
    if (j != 0) return;
    uint32_t mrow = kNone, r0 = kNone;
    double a = 0, v0 = 0;
    for (int k = 0; k < 2 * P; ++k) {
        const int jj = k >> 1, s = k & 1;
        const uint32_t row = m_row[(s * R + col) * P + jj];
        if (row == kNone) continue;
        const double x = m_sum[(s * R + col) * P + jj];
        if (row == mrow) {
            a += x;
            continue;
        }
        if (mrow != kNone) {
            if (r0 == kNone) {
                r0 = mrow;
                v0 = a;
            } else {
                store_final(o + (uint64_t)mrow * R, a, zeroed);
            }
        }
        mrow = row;
        a = x;
    }
    if (r0 == kNone) {  // a single run: it is the first
        r0 = mrow;
        v0 = a;
        mrow = kNone;
    }
    first_row[row_stride * col] = r0;
    first_val[col] = v0;
    last_row[row_stride * col] = mrow;
    last_val[col] = a;
}

template <typename VT, typename FT, typename OT, int R, int PKW, int NM1, int NT, int CB>
__global__ __launch_bounds__(NT) void k_mttkrp_hot(const uint64_t *__restrict__ pkeys,
                                                            const VT *__restrict__ vals, uint64_t n,
                                                            HotArgs ha, OT *__restrict__ out,
                                                            uint64_t per_group, int nwg, bool zeroed,
                                                            uint32_t *__restrict__ bnd_row,
                                                            double *__restrict__ bnd_val) {
    using CT = std::conditional_t<sizeof(FT) == 8, double, float>;
    constexpr int VPL = sizeof(FT) == 8 ? 4 : 8;  // elements per lane
    constexpr int L = R / VPL;                      // lanes per group
    constexpr int G = NT / L;
    // entries a lane decodes per block (the group loads BLK = E * L at a time)
    constexpr int E0 = NM1 >= 4 ? 2 : 4;
    constexpr int E = E0 < kHotBlock / L ? E0 : kHotBlock / L;
    constexpr int BLK = E * L;
    static_assert(kHotBlock % BLK == 0, "block");
    constexpr int RB = R * (int)sizeof(FT);         // row bytes
    constexpr int C16 = RB / 16;                    // 16-byte chunks per row
    constexpr int LC = VPL * (int)sizeof(FT) / 16;  // chunks per lane
    static_assert(R % VPL == 0 && L >= 1 && L <= 16 && (64 % L) == 0 && E >= 1, "rank tiling");
    static_assert(LC >= 1 && (C16 & (C16 - 1)) == 0, "row chunks");
    extern __shared__ __align__(16) char smem[];

    const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int gib = threadIdx.x / L;
    const int lig = threadIdx.x % L;
    const uint64_t group = (uint64_t)lb * G + gib;
    uint64_t start = group * per_group;
    uint64_t end = start + per_group < n ? start + per_group : n;
    if (start > n) start = n;

    // hot tables: row s of gathered mode q at ha.lds[q] + s * RB, 16-byte
    // chunk c stored at chunk position c ^ (row index & (C16 - 1))
    {
        uint32_t tot = 0, pre[NM1 + 1];
#pragma unroll
        for (int q = 0; q < NM1; ++q) {
            pre[q] = tot;
            tot += ha.K[q] * C16;
        }
        pre[NM1] = tot;
        for (uint32_t t = threadIdx.x; t < tot; t += NT) {
            int q = 0;
#pragma unroll
            for (int qq = 1; qq < NM1; ++qq)
                if (t >= pre[qq]) q = qq;
            const uint32_t u = t - pre[q], s = u / C16, c = u % C16;
            const uint32_t row = ha.list[q][s];
            const uint4 v = *(const uint4 *)((const char *)ha.fac[q] + (uint64_t)row * RB + c * 16);
            const uint32_t base = ha.lds[q] + s * RB;
            const uint32_t swz = (base / RB) & (C16 - 1);
            *(uint4 *)(smem + base + ((c ^ swz) << 4)) = v;
        }
    }
    __syncthreads();

    const char *fb[NM1];
#pragma unroll
    for (int q = 0; q < NM1; ++q) fb[q] = (const char *)ha.fac[q];
    const uint32_t lane_b = (uint32_t)(lig * VPL * sizeof(FT));
    OT *outp = out + lig * VPL;
    // A hot code carries its LDS row address a (a multiple of RB) with the
    // row's chunk swizzle s in bits 4.. and the flag bit: kHotFlag | a | s<<4.
    // Chunk j of the row sits at a + ((j ^ s) << 4) = (a | s<<4) ^ (j << 4),
    // so a lane's chunk address is one xor with a lane constant, which also
    // clears the flag bit.
    uint32_t lds_x[LC];
#pragma unroll
    for (int c = 0; c < LC; ++c) lds_x[c] = kHotFlag | ((uint32_t)(lig * LC + c) << 4);

    uint32_t cur = kNone, first_row = kNone;
    bool in_first = true;
    CT acc[VPL], facc[VPL];
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
        acc[v] = 0;
        facc[v] = 0;
    }
    uint64_t ck[E * PKW];
    VT cv[E];
    if (start < end) {
        load_keys<PKW, E, true>(pkeys, start + (uint64_t)lig * E, ck);
        load_vals<VT, E, true>(vals, start + (uint64_t)lig * E, cv);
    }
    for (uint64_t base = start; base < end; base += BLK) {
        // decode this lane's E entries
        uint32_t roff[NM1][E], rrow[E], rval[E][Bits32<VT>::N];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint64_t *k = ck + e * PKW;
#pragma unroll
            for (int q = 0; q < NM1; ++q) {
                const uint32_t code = hfield<PKW>(k, ha.fsh[q], ha.fw[q]);
                const uint32_t la = ha.lds[q] + (code - ha.dim[q]) * (uint32_t)RB;
                roff[q][e] = code >= ha.dim[q] ? (kHotFlag | la | (((la / RB) & (C16 - 1)) << 4))
                                               : code * (uint32_t)RB;
            }
            rrow[e] = hfield<PKW>(k, ha.tsh, ha.tw);
            Bits32<VT>::split(cv[e], rval[e]);
        }
        const uint64_t nb = base + BLK;
        const int nh = (end - base) < (uint64_t)BLK ? (int)(end - base) : BLK;
        auto batch = [&](auto e0c) {
            constexpr int e0 = decltype(e0c)::value;
            uint32_t ob[CB][NM1], tr[CB];
            CT vv[CB];
            auto take = [&](auto kc) {
                constexpr int k = decltype(kc)::value;
                constexpr int e = e0 + k, owner = e / E, slot = e % E;
#pragma unroll
                for (int q = 0; q < NM1; ++q) ob[k][q] = gbcast<L, owner>(roff[q][slot]);
                tr[k] = gbcast<L, owner>(rrow[slot]);
                uint32_t w[Bits32<VT>::N];
#pragma unroll
                for (int b = 0; b < Bits32<VT>::N; ++b) w[b] = gbcast<L, owner>(rval[slot][b]);
                vv[k] = (CT)Bits32<VT>::join(w);
            };
            static_for<0, CB>(take);
            CT f[NM1][CB][VPL];
#pragma unroll
            for (int q = 0; q < NM1; ++q)
#pragma unroll
                for (int k = 0; k < CB; ++k) {
                    const uint32_t o = ob[k][q];
                    if (o & kHotFlag) {
#pragma unroll
                        for (int c = 0; c < LC; ++c) {
                            const uint4 u = *(const uint4 *)(smem + (o ^ lds_x[c]));
                            HotRow<FT>::cvt(u, f[q][k] + c * HotRow<FT>::E16);
                        }
                    } else {
                        const char *p = fb[q] + o + lane_b;
#pragma unroll
                        for (int c = 0; c < LC; ++c) {
                            const uint4 u = *(const uint4 *)(p + 16 * c);
                            HotRow<FT>::cvt(u, f[q][k] + c * HotRow<FT>::E16);
                        }
                    }
                }
            // The next block's keys and values are loaded after the first
            // batch's gathers: loads retire in order, so the first batch's
            // wait does not include them, and they get that batch's time.
            // Unconditional (a branch would leave them pending on one path):
            // past the group's last block this reads the next group's
            // entries or the padding (base < end, so at most 2 BLK - 1 <
            // kSchedPad entries past the end), which nothing uses.
            if constexpr (e0 == 0) {
                load_keys<PKW, E, true>(pkeys, nb + (uint64_t)lig * E, ck);
                load_vals<VT, E, true>(vals, nb + (uint64_t)lig * E, cv);
            }
#pragma unroll
            for (int k = 0; k < CB; ++k) {
                // entries past the group's end add zero to the current run
                // (no branch around them: every path then consumes the loads)
                const bool live = e0 + k < nh;
                const uint32_t row = live ? tr[k] : cur;
                const CT vk = live ? vv[k] : (CT)0;
                if (row != cur) {  // a new run: hand the finished one on, restart from zero
                    if (cur != kNone) {
                        if (in_first) {
                            first_row = cur;
#pragma unroll
                            for (int v = 0; v < VPL; ++v) facc[v] = acc[v];
                            in_first = false;
                        } else {
                            flush_lane<OT, CT, VPL>(outp + (uint64_t)cur * R, acc, zeroed);
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
        };
        static_for<0, BLK / CB>([&](auto i) {
            batch(std::integral_constant<int, decltype(i)::value * CB>{});
        });
    }
    if (in_first) {  // zero or one run in this group: it is the first
        first_row = cur;
#pragma unroll
        for (int v = 0; v < VPL; ++v) facc[v] = acc[v];
        cur = kNone;
    }
    __syncthreads();  // every wave is done with the hot tables
    uint32_t *s_row = (uint32_t *)smem;                                 // [2G]: slot 2g = first, 2g+1 = last
    double *s_part = (double *)(smem + ((8 * G + 15) & ~15));           // [2G][R]
    char *s_scr = smem + ((8 * G + 15) & ~15) + (size_t)16 * G * R;    // combine scratch
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
        s_part[(size_t)(2 * gib) * R + lig * VPL + v] = (double)facc[v];
        s_part[(size_t)(2 * gib + 1) * R + lig * VPL + v] = (double)acc[v];
    }
    if (lig == 0) {
        s_row[2 * gib] = first_row;
        s_row[2 * gib + 1] = cur;
    }
    __syncthreads();
    combine_runs<OT, R, NT>(s_row, s_part, 2 * G, s_scr, out, zeroed, bnd_row + 2 * lb,
                                     bnd_val + (size_t)(2 * lb) * R, bnd_row + 2 * lb + 1,
                                     bnd_val + (size_t)(2 * lb + 1) * R, 0);
}

struct HotLaunch {
    const uint64_t *keys;
    const void *vals;
    uint64_t n;
    HotArgs ha;
    void *out;
    int nwg, pkw, nm1;
    int shape;  // threads per workgroup / entries per gather batch: 0 = 1024 / 1, 1 = 768 / 2, 2 = 512 / 4
    bool zeroed;
    size_t hot_bytes;
    uint32_t *bnd_row;
    double *bnd_val;
    uint32_t *lrow[2];  // the boundary merge's level buffers
    double *lval[2];
};

// Launches the kernel and the boundary merge for ranks 16, 32, 64.
template <typename VT, typename FT, typename OT>
int hot_rank(int rank, const HotLaunch &h, hipStream_t s);

#ifdef ALTO_HOT_INSTANTIATE
template <typename VT, typename FT, typename OT, int R, int NT, int CB>
static int hot_launch_nt(const HotLaunch &h, hipStream_t s) {
    constexpr int VPL = sizeof(FT) == 8 ? 4 : 8, L = R / VPL, G = NT / L;
    using KernT = void (*)(const uint64_t *, const VT *, uint64_t, HotArgs, OT *, uint64_t, int, bool,
                           uint32_t *, double *);
    const uint64_t groups = (uint64_t)h.nwg * G;
    uint64_t pg = (h.n + groups - 1) / groups;
    pg = (pg + kHotBlock - 1) / kHotBlock * kHotBlock;
    if (pg == 0) pg = kHotBlock;
    const size_t comb = ((8 * (size_t)G + 15) & ~(size_t)15) + 16 * (size_t)G * R + 24 * (size_t)NT;
    const size_t lds = std::max(h.hot_bytes, comb);
    if (lds > kHotLds) {
        set_error("hot-row MTTKRP needs %zu bytes of LDS", lds);
        return ALTO_ERR_UNSUPPORTED;
    }
    KernT k = nullptr;
    auto pick = [&](auto pkwc, auto nm1c) {
        constexpr int PKW = decltype(pkwc)::value, NM1 = decltype(nm1c)::value;
        k = (KernT)k_mttkrp_hot<VT, FT, OT, R, PKW, NM1, NT, CB>;
    };
    auto pick_nm1 = [&](auto pkwc) {
        if (h.nm1 == 2) pick(pkwc, std::integral_constant<int, 2>{});
        else if (h.nm1 == 3) pick(pkwc, std::integral_constant<int, 3>{});
        else pick(pkwc, std::integral_constant<int, 4>{});
    };
    if (h.pkw == 1) pick_nm1(std::integral_constant<int, 1>{});
    else pick_nm1(std::integral_constant<int, 2>{});
    ALTO_HIP_TRY(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k<<<h.nwg, NT, lds, s>>>(h.keys, (const VT *)h.vals, h.n, h.ha, (OT *)h.out, pg, h.nwg, h.zeroed,
                             h.bnd_row, h.bnd_val);
    ALTO_LAUNCH_CHECK();
    launch_bnd_merge<OT, R, double>(h.bnd_row, h.bnd_val, 2 * (uint64_t)h.nwg, h.lrow, h.lval, (OT *)h.out,
                                    h.zeroed ? 1 : 0, s);
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

template <typename VT, typename FT, typename OT, int R>
static int hot_launch(const HotLaunch &h, hipStream_t s) {
    switch (h.shape) {
    case 0: return hot_launch_nt<VT, FT, OT, R, 1024, 1>(h, s);
    case 1: return hot_launch_nt<VT, FT, OT, R, 768, 2>(h, s);
    default: return hot_launch_nt<VT, FT, OT, R, 512, 4>(h, s);
    }
}

template <typename VT, typename FT, typename OT>
int hot_rank(int rank, const HotLaunch &h, hipStream_t s) {
    switch (rank) {
    case 16: return hot_launch<VT, FT, OT, 16>(h, s);
    case 32: return hot_launch<VT, FT, OT, 32>(h, s);
    case 64: return hot_launch<VT, FT, OT, 64>(h, s);
    default: return ALTO_ERR_UNSUPPORTED;
    }
}
#endif

}  // namespace alto
