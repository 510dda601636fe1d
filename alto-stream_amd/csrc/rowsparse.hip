// Row-sparse MTTKRP: rowsparse_mttkrp (src/rowsparse_mttkrp.cpp:97-187,
// streaming overload declared in include/rowsparse_mttkrp.hpp:36-41).
//
// The reference walks the nonzeros bucketed by the output index and, per
// nonzero, multiplies the value by the stored row ridx[m][i_m] of every other
// mode (the stream mode's single row for the stream mode), accumulating one
// output row per distinct output index.  The output is row-sparse: the
// distinct output indices in ascending order and their rows.
//
// Here: the row-sparse factors are expanded into dense device matrices (a
// stored row, or zeros where ridx is -1, which the reference would read out
// of bounds), the dense ALTO MTTKRP runs into a dense output, and the rows of
// the indices that occur are compacted.  Same products, FP64.
#include <vector>

#include "sort.hpp"
#include "tensor.hpp"

namespace alto {

// D[i] = rsp[ridx[i]] (zeros where ridx[i] < 0), or rsp[0] for every row
// when ridx == nullptr (the stream mode).
__global__ __launch_bounds__(256) void k_expand_rows(const double *__restrict__ rsp, uint64_t nrsp,
                                                     const int32_t *__restrict__ ridx,
                                                     uint64_t rows, int R,
                                                     double *__restrict__ D) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * R) return;
    const uint64_t i = e / R;
    const int c = (int)(e % R);
    int64_t r = ridx ? (int64_t)ridx[i] : 0;
    D[e] = (r >= 0 && (uint64_t)r < nrsp) ? rsp[(uint64_t)r * R + c] : 0.0;
}

__global__ __launch_bounds__(256) void k_mark_rows(const uint64_t *__restrict__ idx, uint64_t nnz,
                                                   uint32_t *__restrict__ flag) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnz) flag[idx[i]] = 1u;
}

// present[i] -> out_rowind[pos[i]] = i, out rows = Y[i]
__global__ __launch_bounds__(256) void k_compact_rows(const uint32_t *__restrict__ present,
                                                      const uint32_t *__restrict__ pos,
                                                      const double *__restrict__ Y, uint64_t rows,
                                                      int R, uint64_t *__restrict__ rowind,
                                                      double *__restrict__ out) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * R) return;
    const uint64_t i = e / R;
    const int c = (int)(e % R);
    if (!present[i]) return;
    const uint64_t p = pos[i];
    out[p * R + c] = Y[e];
    if (c == 0) rowind[p] = i;
}

}  // namespace alto

using namespace alto;

extern "C" int alto_hip_rowsparse_mttkrp(const alto_hip_coo *coo, int mode, int stream_mode,
                                         int rank, const double *const *rsp_vals,
                                         const uint64_t *rsp_rows, const int32_t *const *ridx,
                                         uint64_t *out_rowind, double *out_vals,
                                         uint64_t *out_nnzr, void *stream) {
    if (!coo || !coo->dims || !coo->cidx || !coo->vals || !rsp_vals || !rsp_rows || !ridx ||
        !out_rowind || !out_vals || !out_nnzr)
        return ALTO_ERR_ARG;
    const int N = coo->nmode;
    if (N < 2 || N > kMaxModes || mode < 0 || mode >= N || stream_mode < -1 || stream_mode >= N ||
        stream_mode == mode || rank < 1 || rank > 4096)
        return ALTO_ERR_ARG;
    for (int m = 0; m < N; ++m) {
        if (m == mode) continue;
        if (!rsp_vals[m] || (m != stream_mode && !ridx[m]) || (m == stream_mode && rsp_rows[m] < 1))
            return ALTO_ERR_ARG;
    }
    if (coo->dims[mode] >= (1ULL << 32)) return ALTO_ERR_UNSUPPORTED;
    ALTO_TRY(ensure_init());
    hipStream_t s = to_stream(stream);
    const bool dev = coo->on_device != 0;
    const uint64_t nnz = coo->nnz, I = coo->dims[mode];
    const int R = rank;

    // the tensor (coordinates checked against dims on the way in)
    TensorImpl t;
    ALTO_TRY(build_alto(&t, coo, 0, 0, true, s));

    // dense copies of the row-sparse factors
    DevBuf D[kMaxModes], stage_v[kMaxModes], stage_r[kMaxModes];
    const void *fptr[kMaxModes] = {};
    for (int m = 0; m < N; ++m) {
        if (m == mode) continue;
        const uint64_t rows = coo->dims[m], nrsp = rsp_rows[m];
        ALTO_TRY(D[m].alloc(sizeof(double) * rows * R));
        const double *rv = rsp_vals[m];
        const int32_t *rx = m == stream_mode ? nullptr : ridx[m];
        if (!dev) {
            ALTO_TRY(stage_v[m].alloc(sizeof(double) * (nrsp ? nrsp : 1) * R));
            if (nrsp)
                ALTO_HIP_TRY(hipMemcpyAsync(stage_v[m].p, rv, sizeof(double) * nrsp * R,
                                            hipMemcpyHostToDevice, s));
            rv = stage_v[m].as<double>();
            if (rx) {
                ALTO_TRY(stage_r[m].alloc(sizeof(int32_t) * rows));
                ALTO_HIP_TRY(hipMemcpyAsync(stage_r[m].p, ridx[m], sizeof(int32_t) * rows,
                                            hipMemcpyHostToDevice, s));
                rx = stage_r[m].as<int32_t>();
            }
        }
        if (rows)
            k_expand_rows<<<cdiv(rows * R, 256), 256, 0, s>>>(rv, nrsp, rx, rows, R, D[m].as<double>());
        ALTO_LAUNCH_CHECK();
        fptr[m] = D[m].p;
    }

    // dense MTTKRP of the target mode, then the rows whose index occurs
    DevBuf Y, present, pos, scan_tmp, rowind_d, vals_d;
    ALTO_TRY(Y.alloc(sizeof(double) * I * R));
    ALTO_HIP_TRY(hipMemsetAsync(Y.p, 0, sizeof(double) * I * R, s));
    ALTO_TRY(run_mttkrp(&t, mode, ALTO_PREC_F64, R, fptr, Y.p, 1, 2, 0, -1, s, false, true));
    ALTO_TRY(present.alloc(sizeof(uint32_t) * I));
    ALTO_TRY(pos.alloc(sizeof(uint32_t) * I));
    ALTO_HIP_TRY(hipMemsetAsync(present.p, 0, sizeof(uint32_t) * I, s));
    const uint64_t *tidx = coo->cidx[mode];
    DevBuf stage_t;
    if (!dev && nnz) {
        ALTO_TRY(stage_t.alloc(sizeof(uint64_t) * nnz));
        ALTO_HIP_TRY(hipMemcpyAsync(stage_t.p, tidx, sizeof(uint64_t) * nnz, hipMemcpyHostToDevice, s));
        tidx = stage_t.as<uint64_t>();
    }
    if (nnz) k_mark_rows<<<cdiv(nnz, 256), 256, 0, s>>>(tidx, nnz, present.as<uint32_t>());
    ALTO_LAUNCH_CHECK();
    ALTO_HIP_TRY(hipMemcpyAsync(pos.p, present.p, sizeof(uint32_t) * I, hipMemcpyDeviceToDevice, s));
    ALTO_TRY(exclusive_scan_u32(pos.as<uint32_t>(), I, scan_tmp, s));
    uint32_t last_flag = 0, last_pos = 0;
    ALTO_HIP_TRY(hipMemcpyAsync(&last_flag, present.as<uint32_t>() + I - 1, sizeof(uint32_t),
                                hipMemcpyDeviceToHost, s));
    ALTO_HIP_TRY(hipMemcpyAsync(&last_pos, pos.as<uint32_t>() + I - 1, sizeof(uint32_t),
                                hipMemcpyDeviceToHost, s));
    ALTO_HIP_TRY(hipStreamSynchronize(s));
    const uint64_t nnzr = (uint64_t)last_pos + last_flag;
    uint64_t *ri = out_rowind;
    double *ov = out_vals;
    if (!dev) {
        ALTO_TRY(rowind_d.alloc(sizeof(uint64_t) * (nnzr ? nnzr : 1)));
        ALTO_TRY(vals_d.alloc(sizeof(double) * (nnzr ? nnzr : 1) * R));
        ri = rowind_d.as<uint64_t>();
        ov = vals_d.as<double>();
    }
    k_compact_rows<<<cdiv(I * R, 256), 256, 0, s>>>(present.as<uint32_t>(), pos.as<uint32_t>(),
                                                    Y.as<double>(), I, R, ri, ov);
    ALTO_LAUNCH_CHECK();
    if (!dev && nnzr) {
        ALTO_HIP_TRY(hipMemcpyAsync(out_rowind, ri, sizeof(uint64_t) * nnzr, hipMemcpyDeviceToHost, s));
        ALTO_HIP_TRY(hipMemcpyAsync(out_vals, ov, sizeof(double) * nnzr * R, hipMemcpyDeviceToHost, s));
    }
    ALTO_HIP_TRY(hipStreamSynchronize(s));
    *out_nnzr = nnzr;
    return ALTO_OK;
}
