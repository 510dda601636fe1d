// Device-resident ALTO tensor (the GPU counterpart of AltoTensor<LIT>,
// include/alto.hpp:35-62) and the per-mode scheduling data the MTTKRP
// kernels use.
#pragma once
#include <memory>
#include <vector>

#include "common.hpp"
#include "sort.hpp"

namespace alto {

// Bit positions of every mode inside the ALTO key: bit b of the mode-m index
// sits at key bit pos[m][b] (the pdep/pext mapping of mask m).
struct BitPos {
    int nmode;
    int bits[kMaxModes];
    uint8_t pos[kMaxModes][64];
};

// Rows of every mode ranked by how many nonzeros use them (the hot-row
// MTTKRP keeps the most popular rows of the gathered modes in LDS).
constexpr uint32_t kHotMax = 4096;  // ranked rows kept per mode
struct HotModes {
    bool built = false;
    DevBuf list[kMaxModes];                // u32: the kHotMax (or fewer) most popular rows, in order
    DevBuf rank[kMaxModes];                // u32 per row: its popularity rank if listed, else ~0
    std::vector<uint32_t> cnt[kMaxModes];  // host: entry counts of the listed rows
};

// A schedule's keys re-encoded for the hot-row MTTKRP: gathered mode q keeps
// K[q] rows in LDS, and a hot index of mode m is stored as I_m + rank.
struct HotSched {
    bool built = false;
    int rb = 0;  // factor row bytes the LDS split was made for
    int nq = 0;
    int qmode[4] = {};
    uint32_t K[4] = {};
    PackInfo hpk{};
    DevBuf keys;
};

// Entries of padding after a schedule's values and the hot-row kernel's
// re-encoded keys (zero keys: row 0 of every mode): the hot-row kernel loads
// whole blocks of keys and values without a bounds branch, and its entries
// past the end add zero.
constexpr uint64_t kSchedPad = 64;

// Output-ordered schedule of one mode (algo 2): the nonzeros re-sorted by
// the mode's index, stably, so ALTO order is kept inside each row run.
struct ModeSchedule {
    bool built = false;
    bool has64 = false;
    int lvl[kMaxModes] = {};  // fiber levels: the other modes, shallow to leaf
    // The sort fields: the target and the levels that have a nonzero width.
    // Two modes with the same sequence have the same schedule.
    int eff[kMaxModes] = {};
    int neff = 0;
    uint64_t n = 0;           // entries (distinct coordinates)
    DevBuf pkeys;   // packed keys in row order
    DevBuf vals32;  // FP32 values in row order
    DevBuf vals64;  // FP64 values in row order (built on demand)
    HotSched hot;   // re-encoded keys of the hot-row MTTKRP (built on demand)
};

struct TensorImpl {
    int nmode = 0;
    int kw = 1;          // ALTO key words (1: 64-bit mask, 2: 128-bit mask)
    int pkw = 1;         // packed key words
    int total_bits = 0;
    int nprtn = 1;
    uint64_t nnz = 0;
    uint64_t dims[kMaxModes] = {};
    uint64_t masks[kMaxModes * 2] = {};
    BitPos bp{};
    PackInfo pack{};

    DevBuf keys;    // sorted ALTO keys, nnz * kw words (interleaved)
    DevBuf pkeys;   // sorted packed keys, nnz * pkw words
    DevBuf perm;    // u32: sorted position -> input nonzero index
    DevBuf vals64;  // sorted FP64 values
    DevBuf vals32;  // sorted FP32 values (built on first FP32 use)

    std::vector<uint64_t> prtn_ptr;    // nprtn + 1
    std::vector<uint64_t> intervals;   // nprtn * nmode * 2
    // Conflict-resolution bit of each mode (kw words per mode; zero: none),
    // set in keys and pkeys by apply_cr_bits, cleared by every rebuild.
    uint64_t cr_masks[kMaxModes * 2] = {};

    ModeSchedule sched[kMaxModes];
    HotModes hot;
    DevBuf hot_bnd;     // boundary runs of the hot-row MTTKRP's workgroups
    DevBuf bnd_slots;   // FP64 boundary rows of the FP32-output MTTKRP's workgroups
    SortScratch sort_ws;
    DevBuf scratch;     // generic per-call scratch
    DevBuf k0, k1, bad, ivbuf, chunkbuf;  // construction scratch, kept for rebuilds
    bool vals32_valid = false;
    // Schedule-construction scratch kept between builds (a tensor rebuilt
    // slice after slice keeps it, so a rebuild neither allocates nor frees
    // device memory, which would synchronise the device); otherwise it is
    // allocated and released per schedule.
    bool keep_sched_scratch = false;
    DevBuf sch_k0, sch_k1, sch_perm, sch_sk, sch_sv, sch_longs;
};

// (Re)builds t from a COO tensor, reusing t's device buffers (the GPU
// counterpart of init_salto + update_salto, include/alto.hpp:151-299).
// nprtn == 0 skips the reference partitions; check == false skips the
// coordinate range check; with both off the call does not synchronise.
// alto_order == false (only with nprtn == 0) keeps the nonzeros in input
// order instead of sorting them by ALTO key: for callers that use only the
// per-mode schedules, which come out the same either way (their stable sort
// is on every coordinate field, and equal coordinates keep input order in
// the ALTO sort too).
int build_alto(TensorImpl *t, const alto_hip_coo *coo, int nprtn, int key_words, bool check,
               hipStream_t s, bool alto_order = true);

// GENERATE_CR_BITS (include/alto.hpp:1250-1346): picks the modes that take
// a conflict-resolution bit and sets it in every nonzero whose row of that
// mode is touched by more than one partition.
int apply_cr_bits(TensorImpl *t, hipStream_t s);

// Host-side mask construction: LSB_FIRST + SHORT_FIRST packing
// (include/alto.hpp:954-1057).  Returns total bits or -1 if > 64*kw.
int build_masks(int nmode, const uint64_t *dims, int kw, uint64_t *masks, BitPos *bp);
int mode_bits(uint64_t dim);

int ensure_vals32(TensorImpl *t, hipStream_t s);
int ensure_schedule(TensorImpl *t, int mode, bool need64, hipStream_t s);
// out += MTTKRP of `target` (algo 1: ALTO order, 2: output-ordered schedule;
// per_group 0 and flags -1: defaults, autotuned for large tensors).
// prepare_only: build the schedule and tune, launch nothing.  out_zeroed:
// the output holds zeros (rows owned by one range are stored, not added).
int run_mttkrp(TensorImpl *t, int target, int prec, int rank, const void *const *factors,
               void *out, int out_f64, int algo, uint64_t per_group, int flags, hipStream_t s,
               bool prepare_only = false, bool out_zeroed = false);
int prepare_mttkrp(TensorImpl *t, int target, int prec, int rank, const void *const *factors,
                   int out_f64, hipStream_t s);
void fiber_levels(const TensorImpl *t, int mode, int *lvl);

// Fiber levels of an output-ordered schedule: the mode of each level after
// the target and the packed-key bits of every level.
struct FibInfo {
    int mode[4];              // mode of level q + 1, q = 0 .. NM1 - 1
    uint64_t mlo[5], mhi[5];  // packed-key bits of level 0 (target) .. NM1
};

// One tensor's part of a grouped MTTKRP: the fiber kernel over its
// output-ordered schedule of `target` into FP64 `out` (zero on entry), with
// the default segment length; skipped while *active is 0.
struct MttkrpJob {
    const uint64_t *pkeys;
    const void *vals;
    uint64_t nnz;
    const void *F[kMaxModes];
    void *out;
    PackInfo pk;
    int target;
    FibInfo fi;
    uint64_t per_group;
    int flags;
    const int *active;
    unsigned nblocks;  // workgroups of this job
    int dec;           // 1: decode-once kernel, 0: fiber kernel
};
// 3 or 4 modes, 64-bit keys, FP32 or FP64 factors, rank 8, 16, 32 or 64.
bool mttkrp_group_supported(const TensorImpl *t, int prec, int rank);
// Fills *job (building the schedule); ALTO_ERR_UNSUPPORTED outside the
// supported set.
int mttkrp_group_job(TensorImpl *t, int target, int prec, int rank, const void *const *factors, void *out,
                     const int *active, MttkrpJob *job, hipStream_t s);
// jobs and first (njobs + 1 prefix sums of nblocks) in device memory; all
// jobs share precision, rank and mode count, and take the decode-once kernel
// (dec) or the fiber kernel.
int launch_mttkrp_group(const MttkrpJob *jobs, const unsigned *first, int njobs, unsigned nblocks, int prec,
                        int rank, int nmode, bool dec, hipStream_t s);

// Hot-row MTTKRP (one workgroup per CU, popular rows in LDS): out += MTTKRP
// of `target` over its built schedule.  Ranks 16, 32, 64; 3 to 5 modes.
// shape: threads per workgroup / entries per gather batch (0: 1024 / 1,
// 1: 768 / 2, 2: 512 / 4).
bool hot_supported(const TensorImpl *t, int prec, int rank);
int ensure_hot_modes(TensorImpl *t, hipStream_t s);
int ensure_hot_schedule(TensorImpl *t, int target, int rb, hipStream_t s);
int run_hot(TensorImpl *t, int target, int prec, int rank, const void *const *factors, void *out,
            bool out64, bool zeroed, int shape, hipStream_t s);


// Grouped CP-ALS (als_batch.hip): tensors that qualify (3 or 4 modes, 64-bit
// keys, rank 8, 16, 32 or 64, FP32 or FP64) run class by class, every ALS
// step one launch per class.  classes: lists of tensor indices sharing mode
// count, rank, precision and fit convention.  dev_io: factors[k][m] are
// device FP64 arrays (else host).
bool als_group_eligible(const TensorImpl *t, const alto_hip_als_opts *o);
int als_groups_run(const std::vector<std::vector<int>> &classes, const alto_hip_tensor *tensors,
                   const alto_hip_als_opts *opts, double *const *const *factors, double *const *lambdas,
                   double *const *fits, int *iters, bool dev_io);
bool grouped_als_enabled();

}  // namespace alto

struct alto_hip_tensor_s {
    alto::TensorImpl impl;
};
