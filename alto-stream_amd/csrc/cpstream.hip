// CP-stream (src/streaming_cpd.cpp:26-366, 1198-1519) -- filled in below.
#include "tensor.hpp"

extern "C" int alto_hip_cpstream(const alto_hip_coo *coo, int stream_mode, int rank, int max_iters,
                                 double epsilon, int nonneg, int max_slices,
                                 const double *const *init_factors, double *const *out_factors,
                                 uint64_t *out_dims, double *out_time_rows, int *out_outer_iters,
                                 double *out_admm_iters, double *out_grams, double *timings,
                                 void *stream) {
    (void)coo; (void)stream_mode; (void)rank; (void)max_iters; (void)epsilon; (void)nonneg;
    (void)max_slices; (void)init_factors; (void)out_factors; (void)out_dims; (void)out_time_rows;
    (void)out_outer_iters; (void)out_admm_iters; (void)out_grams; (void)timings; (void)stream;
    alto::set_error("CP-stream not built yet");
    return ALTO_ERR_UNSUPPORTED;
}
