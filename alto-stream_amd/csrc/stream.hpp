// Host-side stream preparation shared by CP-stream and spCP-stream.
#pragma once
#include <stdint.h>

#include <utility>
#include <vector>

#include "tensor.hpp"

namespace alto {

// One time slice: nonzeros [start, end) of the prepared stream and the model
// dims after it (non-stream modes: running max of index + 1; stream mode 1).
struct Slice {
    uint64_t start, end;
    uint64_t dims[kMaxModes];
    uint64_t t0;  // first time index (a batch holds time indices t0 .. t0 + dims[s] - 1)
};

// StreamingSparseTensor (src/streaming_sptensor.cpp:69-264) with
// next_dynamic_batch at timeslice limit 1, built on the device: the nonzeros
// stably sorted by (s, s+1, ...) mod N, every non-stream index relabelled by
// first appearance in that order, one slice per time index.
struct StreamInput {
    DevBuf cid[kMaxModes];  // device, per mode, relabelled (stream mode: time index)
    DevBuf vals;            // device
    std::vector<Slice> slices;
};

// Groups consecutive time slices into batches as next_dynamic_batch does
// (src/streaming_sptensor.cpp:196-264): at most timeslice_limit slices (0
// counts as 1), and no further slice once nnz_threshold nonzeros are taken
// (0: no threshold).  A batch's dims are those after its last slice, with
// dims[s] = its number of time slices.
std::vector<Slice> stream_batches(const std::vector<Slice> &slices, int s, uint64_t nnz_threshold,
                                  uint64_t timeslice_limit);

// HIP events destroyed with the set (also on every early return).
struct EventSet {
    std::vector<hipEvent_t> ev;
    EventSet() = default;
    EventSet(const EventSet &) = delete;
    EventSet &operator=(const EventSet &) = delete;
    ~EventSet() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    int make(hipEvent_t *e) {
        ALTO_HIP_TRY(hipEventCreate(e));
        ev.push_back(*e);
        return ALTO_OK;
    }
};

// Host or device COO in (nnz < 2^32); rejects coordinates >= dims
// (ALTO_ERR_ARG) and a time index without nonzeros (ALTO_ERR_EMPTY_SLICE;
// the reference asserts).
int stream_prepare(const alto_hip_coo *coo, int s, StreamInput *out, hipStream_t st);

// The nodes of a captured graph in creation order: type and, for kernel
// nodes, the kernel function.
struct GraphSig {
    std::vector<std::pair<int, const void *>> nodes;
    bool operator==(const GraphSig &o) const { return nodes == o.nodes; }
};
int graph_signature(hipGraph_t g, GraphSig *sig);

// Runs the enqueue of one outer iteration directly the first time and, from
// the second time on, as one captured graph replayed on the same stream (the
// iterations of a slice are identical launches over the same buffers).  The
// executable graph is destroyed only after the stream has drained.
struct IterGraph {
    hipStream_t s = nullptr;
    hipGraphExec_t ge = nullptr;
    GraphSig sig;          // nodes of the capture ge was made from
    bool current = false;  // ge holds this sequence's capture
    explicit IterGraph(hipStream_t st) : s(st) {}
    IterGraph(const IterGraph &) = delete;
    IterGraph &operator=(const IterGraph &) = delete;
    ~IterGraph() {
        if (ge) {
            (void)hipStreamSynchronize(s);
            (void)hipGraphExecDestroy(ge);
        }
    }
    // Iteration 0 of a sequence runs directly; iteration 1 is captured and
    // iterations >= 1 replay the capture.  A sequence after the first (the
    // next slice) updates the executable graph in place
    // (hipGraphExecUpdate, cheaper than instantiating) when the new capture
    // has the same nodes calling the same kernels in the same order (only
    // arguments, grids and copy extents differ: the update's contract), and
    // instantiates a new one otherwise (the MTTKRP's kernel variant can
    // change with a slice's shape).  The capture runs under CaptureScope.
    template <typename F>
    int run(int iter, F enqueue) {
        if (iter == 0 || !graphs_enabled()) {
            current = false;
            return enqueue();
        }
        if (!current) {
            CaptureScope cs;
            hipGraph_t g = nullptr;
            ALTO_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int st = enqueue();
            const hipError_t ce = hipStreamEndCapture(s, &g);
            if (st != ALTO_OK || ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                if (st != ALTO_OK) return st;
                ALTO_HIP_TRY(ce);
            }
            GraphSig ns;
            const int gs = graph_signature(g, &ns);
            if (gs != ALTO_OK) {
                (void)hipGraphDestroy(g);
                return gs;
            }
            if (ge) {
                // no launch of the previous capture may still be running
                hipError_t ue = hipStreamSynchronize(s);
                hipGraphNode_t bad = nullptr;
                hipGraphExecUpdateResult res = hipGraphExecUpdateSuccess;
                if (ue == hipSuccess && ns == sig) ue = hipGraphExecUpdate(ge, g, &bad, &res);
                else if (ue == hipSuccess) res = hipGraphExecUpdateErrorTopologyChanged;
                if (ue != hipSuccess || res != hipGraphExecUpdateSuccess) {
                    (void)hipGetLastError();
                    (void)hipGraphExecDestroy(ge);
                    ge = nullptr;
                }
            }
            hipError_t ie = hipSuccess;
            if (!ge) ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            ALTO_HIP_TRY(ie);
            sig = std::move(ns);
            current = true;
        }
        ALTO_HIP_TRY(hipGraphLaunch(ge, s));
        return ALTO_OK;
    }
};

// cpd_error (src/streaming_cpd.cpp:368-464) over device COO and factors;
// out[4]: err, ||X||^2, ||M||^2, <X, M>.
int cpd_error_dev(int N, const uint64_t *rows, uint64_t nnz, const uint64_t *const *cidx,
                  const double *vals, const double *const *factors, int R, hipStream_t s,
                  double *out);

}  // namespace alto
