// Runtime part of the C ABI: device selection, memory, streams, timers,
// precision conversion, error reporting.
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "common.hpp"

namespace alto {

static thread_local char g_err[1024] = "";
static hipStream_t g_stream = nullptr;
static int g_device = -1;

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
const char *last_error() { return g_err; }

std::shared_mutex &capture_mutex() {
    static std::shared_mutex m;
    return m;
}
bool &capture_flag() {
    static thread_local bool f = false;
    return f;
}
bool graphs_enabled() {
    static const bool on = [] {
        const char *e = getenv("ALTO_HIP_GRAPHS");
        return !(e && e[0] == '0' && e[1] == 0);
    }();
    return on;
}

int ensure_init() {
    if (g_device >= 0) return ALTO_OK;
    return alto_hip_init(0);
}

hipStream_t default_stream() { return g_stream; }

int check_device_access(const void *p, const char *what) {
    if (!p) return ALTO_OK;  // empty arrays: the callers check their own nulls
    hipPointerAttribute_t a{};
    const hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess || a.type == hipMemoryTypeUnregistered) {
        (void)hipGetLastError();  // clear the sticky query error
        set_error("%s is not memory the GPU can read (pageable host memory?): pass device arrays", what);
        return ALTO_ERR_ARG;
    }
    return ALTO_OK;
}

__global__ void k_f64_to_f32(float *d, const double *s, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}
__global__ void k_f32_to_f64(double *d, const float *s, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (double)s[i];
}
__global__ void k_f64_to_bf16(uint16_t *d, const double *s, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = f32_to_bf16((float)s[i]);
}
__global__ void k_f32_to_bf16(uint16_t *d, const float *s, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = f32_to_bf16(s[i]);
}
__global__ void k_bf16_to_f64(double *d, const uint16_t *s, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (double)bf16_to_f32(s[i]);
}
__global__ void k_bf16_to_f32(float *d, const uint16_t *s, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = bf16_to_f32(s[i]);
}

}  // namespace alto

using namespace alto;

extern "C" {

const char *alto_hip_error_string(int status) {
    switch (status) {
    case ALTO_OK: return "ok";
    case ALTO_ERR_ARG: return "invalid argument";
    case ALTO_ERR_KEY_WIDTH: return "coordinates need more bits than the ALTO key holds";
    case ALTO_ERR_HIP: return "HIP runtime error";
    case ALTO_ERR_NOMEM: return "device out of memory";
    case ALTO_ERR_NO_DEVICE: return "no usable GPU";
    case ALTO_ERR_UNSUPPORTED: return "unsupported configuration";
    case ALTO_ERR_NUMERIC: return "Cholesky factorisation failed";
    case ALTO_ERR_EMPTY_SLICE: return "empty time slice";
    default: return "unknown status";
    }
}

const char *alto_hip_last_error(void) { return last_error(); }

int alto_hip_init(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        set_error("no HIP device: %s", hipGetErrorString(e));
        return ALTO_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) {
        set_error("device %d out of range (%d devices)", device, n);
        return ALTO_ERR_ARG;
    }
    ALTO_HIP_TRY(hipSetDevice(device));
    if (g_device != device) {
        if (g_stream) (void)hipStreamDestroy(g_stream);
        g_stream = nullptr;
        ALTO_HIP_TRY(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
        g_device = device;
    }
    return ALTO_OK;
}

int alto_hip_device_name(char *buf, size_t len) {
    ALTO_TRY(ensure_init());
    hipDeviceProp_t p;
    ALTO_HIP_TRY(hipGetDeviceProperties(&p, g_device));
    snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return ALTO_OK;
}

int alto_hip_malloc(void **ptr, size_t bytes) {
    if (!ptr) return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    std::unique_lock<std::shared_mutex> lk(capture_mutex());
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        return ALTO_ERR_NOMEM;
    }
    return ALTO_OK;
}

int alto_hip_free(void *ptr) {
    std::unique_lock<std::shared_mutex> lk(capture_mutex());
    if (ptr) ALTO_HIP_TRY(hipFree(ptr));
    return ALTO_OK;
}

int alto_hip_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    ALTO_TRY(ensure_init());
    hipStream_t s = to_stream(stream);
    ALTO_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    ALTO_HIP_TRY(hipStreamSynchronize(s));
    return ALTO_OK;
}

int alto_hip_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    ALTO_TRY(ensure_init());
    hipStream_t s = to_stream(stream);
    ALTO_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    ALTO_HIP_TRY(hipStreamSynchronize(s));
    return ALTO_OK;
}

int alto_hip_memset(void *ptr, int value, size_t bytes, void *stream) {
    ALTO_TRY(ensure_init());
    ALTO_HIP_TRY(hipMemsetAsync(ptr, value, bytes, to_stream(stream)));
    return ALTO_OK;
}

int alto_hip_stream_create(void **stream) {
    if (!stream) return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    hipStream_t s;
    ALTO_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return ALTO_OK;
}

int alto_hip_stream_destroy(void *stream) {
    if (stream) ALTO_HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return ALTO_OK;
}

int alto_hip_stream_sync(void *stream) {
    ALTO_TRY(ensure_init());
    ALTO_HIP_TRY(hipStreamSynchronize(to_stream(stream)));
    return ALTO_OK;
}

int alto_hip_convert(void *dst, int dst_prec, const void *src, int src_prec, size_t n,
                     void *stream) {
    ALTO_TRY(ensure_init());
    if (!dst || !src) return ALTO_ERR_ARG;
    if (n == 0) return ALTO_OK;
    hipStream_t s = to_stream(stream);
    unsigned g = cdiv(n, 256);
    if (dst_prec == src_prec) {
        size_t el = dst_prec == ALTO_PREC_F64 ? 8 : dst_prec == ALTO_PREC_F32 ? 4 : 2;
        ALTO_HIP_TRY(hipMemcpyAsync(dst, src, n * el, hipMemcpyDeviceToDevice, s));
        return ALTO_OK;
    }
    if (src_prec == ALTO_PREC_F64 && dst_prec == ALTO_PREC_F32)
        k_f64_to_f32<<<g, 256, 0, s>>>((float *)dst, (const double *)src, n);
    else if (src_prec == ALTO_PREC_F32 && dst_prec == ALTO_PREC_F64)
        k_f32_to_f64<<<g, 256, 0, s>>>((double *)dst, (const float *)src, n);
    else if (src_prec == ALTO_PREC_F64 && dst_prec == ALTO_PREC_BF16)
        k_f64_to_bf16<<<g, 256, 0, s>>>((uint16_t *)dst, (const double *)src, n);
    else if (src_prec == ALTO_PREC_F32 && dst_prec == ALTO_PREC_BF16)
        k_f32_to_bf16<<<g, 256, 0, s>>>((uint16_t *)dst, (const float *)src, n);
    else if (src_prec == ALTO_PREC_BF16 && dst_prec == ALTO_PREC_F64)
        k_bf16_to_f64<<<g, 256, 0, s>>>((double *)dst, (const uint16_t *)src, n);
    else if (src_prec == ALTO_PREC_BF16 && dst_prec == ALTO_PREC_F32)
        k_bf16_to_f32<<<g, 256, 0, s>>>((float *)dst, (const uint16_t *)src, n);
    else
        return ALTO_ERR_ARG;
    ALTO_LAUNCH_CHECK();
    return ALTO_OK;
}

struct TimerImpl {
    hipEvent_t a, b;
};

int alto_hip_timer_start(void **timer, void *stream) {
    if (!timer) return ALTO_ERR_ARG;
    ALTO_TRY(ensure_init());
    TimerImpl *t = new TimerImpl;
    ALTO_HIP_TRY(hipEventCreate(&t->a));
    ALTO_HIP_TRY(hipEventCreate(&t->b));
    ALTO_HIP_TRY(hipEventRecord(t->a, to_stream(stream)));
    *timer = t;
    return ALTO_OK;
}

int alto_hip_timer_stop(void *timer, void *stream, float *ms) {
    if (!timer) return ALTO_ERR_ARG;
    TimerImpl *t = (TimerImpl *)timer;
    ALTO_HIP_TRY(hipEventRecord(t->b, to_stream(stream)));
    ALTO_HIP_TRY(hipEventSynchronize(t->b));
    float v = 0;
    ALTO_HIP_TRY(hipEventElapsedTime(&v, t->a, t->b));
    if (ms) *ms = v;
    (void)hipEventDestroy(t->a);
    (void)hipEventDestroy(t->b);
    delete t;
    return ALTO_OK;
}

}  // extern "C"
