// Kernel launchers shared between the kernel translation units and the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mv {

struct V0Args {
    int32_t *dst;            // [n_vec][rows]
    const int8_t *vecs;      // [n_vec][vec_stride]
    uint32_t n_vec, vec_stride;
    uint32_t rows, cols;
    const uint32_t *cursors;
    int8_t min_value;
    uint8_t G;
    const uint8_t *cdf, *ppf;
    const uint16_t *payload;
    uint32_t payload_size;
    int variant;             // decode kernel variant (0 reference step, 1 rotated, 2 byte tables)
};

hipError_t launch_v0_decode_matvec(const V0Args &a, hipStream_t s);

hipError_t launch_int8_matvec(int32_t *dst, const int8_t *mat, const int8_t *vecs,
                              uint32_t n_vec, uint32_t vec_stride, uint32_t rows, uint32_t cols,
                              hipStream_t s);
hipError_t launch_f32_gemv(float *y, const float *mat, const float *x, uint32_t rows,
                           uint32_t cols, float alpha, hipStream_t s);
hipError_t launch_absmax_requant(const int32_t *acc, uint32_t n, int8_t *out8, float *delta,
                                 int clamp, hipStream_t s);
hipError_t launch_scale_output(const int32_t *acc, uint32_t rows, uint32_t n_vec,
                               const float *s_w, const float *s_x, float *y, hipStream_t s);

}  // namespace mv
