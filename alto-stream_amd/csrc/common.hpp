// Shared host/device helpers of libalto_amd.so (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <mutex>
#include <shared_mutex>
#include <string>

#include "../../include/alto_hip.h"

namespace alto {

constexpr int kMaxModes = 8;
constexpr int kWave = 64;

// Last error text (thread-local, returned by alto_hip_last_error).
void set_error(const char *fmt, ...);
const char *last_error();

#define ALTO_HIP_TRY(expr)                                                        \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ::alto::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,          \
                              hipGetErrorString(e_));                             \
            return e_ == hipErrorOutOfMemory ? ALTO_ERR_NOMEM : ALTO_ERR_HIP;     \
        }                                                                         \
    } while (0)

#define ALTO_TRY(expr)                  \
    do {                                \
        int s_ = (expr);                \
        if (s_ != ALTO_OK) return s_;   \
    } while (0)

#define ALTO_LAUNCH_CHECK() ALTO_HIP_TRY(hipGetLastError())

// Library-wide default stream (created lazily by alto_hip_init).
hipStream_t default_stream();
inline hipStream_t to_stream(void *s) { return s ? (hipStream_t)s : default_stream(); }
int ensure_init();
// A caller's array that kernels will read: device, managed or pinned host
// memory.  Pageable host memory (or an address the runtime does not know)
// would fault the GPU when a kernel reads it, so the entry points that take
// device arrays refuse it with ALTO_ERR_ARG.  Not for graph-capture paths.
int check_device_access(const void *p, const char *what);

// Graph capture and device memory.  A thread that captures a stream into a
// graph holds capture_mutex() shared from the start of the capture to the
// end of instantiation, with its capture_flag() set; device allocation and
// release take the mutex exclusively.  So no hipMalloc or hipFree (which
// waits for the whole device) runs while any thread of the process is
// capturing, and an allocation requested inside a capture (which would not
// be part of the graph) is refused with an error.
std::shared_mutex &capture_mutex();
bool &capture_flag();

// Graph replay of repeated launch sequences (CP-ALS and CP-stream
// iterations); ALTO_HIP_GRAPHS=0 in the environment turns it off (the same
// launches are then issued directly every time, e.g. for tools that do not
// follow graph launches).
bool graphs_enabled();

struct CaptureScope {
    std::shared_lock<std::shared_mutex> lk;
    CaptureScope() : lk(capture_mutex()) { capture_flag() = true; }
    ~CaptureScope() { capture_flag() = false; }
    CaptureScope(const CaptureScope &) = delete;
    CaptureScope &operator=(const CaptureScope &) = delete;
};

// Device buffer owned by RAII.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            if (capture_flag()) {
                (void)hipFree(p);
            } else {
                std::unique_lock<std::shared_mutex> lk(capture_mutex());
                (void)hipFree(p);
            }
        }
        p = nullptr;
        bytes = 0;
    }
    int alloc(size_t n) {
        if (n <= bytes && p) return ALTO_OK;
        if (capture_flag()) {
            set_error("device allocation of %zu bytes requested inside a graph capture", n);
            return ALTO_ERR_HIP;
        }
        release();
        if (n == 0) n = 16;
        std::unique_lock<std::shared_mutex> lk(capture_mutex());
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc(%zu): %s", n, hipGetErrorString(e));
            return ALTO_ERR_NOMEM;
        }
        bytes = n;
        return ALTO_OK;
    }
    template <typename T> T *as() const { return (T *)p; }
};

// Packed-key field of one mode: the mode index occupies `width` bits starting
// at bit `shift` of the (1- or 2-word) packed key.
struct DimsArg {
    uint64_t d[kMaxModes];
};

struct PackInfo {
    int nmode;
    int shift[kMaxModes];
    int width[kMaxModes];
};

__device__ __forceinline__ uint64_t extract_field(uint64_t lo, uint64_t hi, int shift,
                                                  int width) {
    uint64_t v;
    if (shift >= 64) {
        v = hi >> (shift - 64);
    } else if (shift == 0) {
        v = lo;
    } else {
        v = (lo >> shift) | (hi << (64 - shift));
    }
    return width >= 64 ? v : (v & ((1ULL << width) - 1ULL));
}

__device__ __forceinline__ float bf16_to_f32(uint16_t b) {
    return __uint_as_float(((uint32_t)b) << 16);
}

// Round-to-nearest-even FP32 -> BF16 (NaN kept quiet).
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    uint32_t rounding = 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)((u + rounding) >> 16);
}

inline unsigned cdiv(uint64_t a, uint64_t b) { return (unsigned)((a + b - 1) / b); }

// Grouped launches: the job owning block b, first[j] <= b < first[j + 1]
// (first: njobs + 1 prefix sums of the jobs' block counts, first[0] = 0).
__device__ __forceinline__ int find_job(const unsigned *__restrict__ first, int njobs, unsigned b) {
    int lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

}  // namespace alto
