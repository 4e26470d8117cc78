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

struct V1Args {
    int32_t *dst;            // [n_vec][rows]
    const int8_t *vecs;      // [n_vec][vec_stride]
    uint32_t n_vec, vec_stride;
    uint32_t rows, cols, streams, precision, table_mode, coder;
    int32_t min_value;
    uint32_t G;
    const uint32_t *cursors;
    const uint32_t *lut;     // table_mode 0: 2 * 2^P dwords
    const uint8_t *freq;     // table_mode 1: [rows][G]
    const uint16_t *payload;
    uint32_t payload_size;
    const uint32_t *out_start;   // outlier split (quad coder), null: none
    const uint32_t *out_rec;
};

hipError_t launch_v1_decode_matvec(const V1Args &a, hipStream_t s);

struct BArgs {
    int layout;              // 0 ans-warp-rows, 1 ans-warp-stripes
    int32_t *acc32;          // [n_vec][out_dim] or null
    int8_t *out8;            // [n_vec][out_dim] or null: clamp(rint(float(acc) * scale))
    float scale;
    const uint32_t *compressed;
    uint32_t compressed_words;
    const uint32_t *warp_offsets;   // uint4 units
    uint32_t end_offset;            // end of this matrix's data, uint4 units
    const uint32_t *ppf, *ppf_cum, *ppf_prob;   // 32 words each (4 entries per word)
    const int8_t *vecs;
    uint32_t n_vec, vec_stride, in_dim, out_dim;
};

hipError_t launch_b_decode_matvec(const BArgs &a, hipStream_t s);

struct V0BwdArgs {
    float *dx;               // [n_vec][cols]
    float *partial;          // workspace [n_vec][ceil(rows/64)][cols]
    const float *g;          // [n_vec][g_stride]
    uint32_t n_vec, g_stride;
    const float *s_w;        // [rows]
    const float *alpha;      // [n_vec]
    uint32_t rows, cols;
    const uint32_t *cursors;
    int8_t min_value;
    uint8_t G;
    const uint8_t *cdf, *ppf;
    const uint16_t *payload;
    uint32_t payload_size;
};

size_t v0_bwd_workspace_bytes(uint32_t rows, uint32_t cols, uint32_t n_vec);
// Whether the partial sums of a rows x cols matrix fit the 32-bit byte offsets the
// backward's group store uses: eight vector planes of ceil(rows / 64) x cols floats
// below 4 GiB.
inline bool bwd_partials_fit(uint32_t rows, uint32_t cols) {
    return 8ull * 4ull * ((rows + 63ull) / 64ull) * cols < (1ull << 32);
}
// Coder 2 keeps each row's table (S + 2 u16 words) at the payload head; the kernels read
// it by row index, so the payload must hold all of them.
inline bool lanes_tables_fit(uint32_t rows, uint32_t streams, uint32_t payload_size) {
    return uint64_t(rows) * (uint64_t(streams) + 2u) <= payload_size;
}
hipError_t launch_v0_backward(const V0BwdArgs &a, hipStream_t s);

struct QuadBwdArgs {
    float *dx;               // [n_vec][cols]
    float *partial;          // workspace [n_vec][ceil(rows/64)][cols]
    const float *g;          // [n_vec][g_stride]
    uint32_t n_vec, g_stride;
    const float *s_w;        // [rows]
    const float *alpha;      // [n_vec]
    uint32_t rows, cols, streams;
    const uint32_t *cursors;
    const uint32_t *lut;     // A / B decode tables, 2 x 256 dwords (table mode 0)
    const uint16_t *payload;
    uint32_t payload_size;
    uint32_t table_mode;     // 1: one 8-bit table per row, from freq (64 streams)
    const uint8_t *freq;     // [rows][G] (table mode 1)
    int32_t min_value;
    uint32_t G;
    const uint32_t *out_col_start;   // outlier split by column (table mode 0), null: none
    const uint32_t *out_col_rec;
    uint32_t coder;          // 1: granule layout; 2: streams back to back (table mode 0)
};

hipError_t launch_quad_backward(const QuadBwdArgs &a, hipStream_t s);
// Device address of the backward's static scratch (MATVANS_NOALLOC_SCRATCH_BYTES) on the
// current device, looked up once per device.
hipError_t backward_scratch(float **ptr);
hipError_t launch_quantize_vectors(const float *x, uint32_t n_vec, uint32_t x_stride,
                                   uint32_t cols, int8_t *xq, uint32_t xq_stride, float *s_x,
                                   hipStream_t s);

hipError_t launch_int8_matvec(int32_t *dst, const int8_t *mat, const int8_t *vecs,
                              uint32_t n_vec, uint32_t vec_stride, uint32_t rows, uint32_t cols,
                              uint32_t cus, hipStream_t s);
hipError_t launch_f32_gemv(float *y, const float *mat, const float *x, uint32_t rows,
                           uint32_t cols, float alpha, hipStream_t s);
hipError_t launch_absmax_requant(const int32_t *acc, uint32_t n, int8_t *out8, float *delta,
                                 int clamp, hipStream_t s);
hipError_t launch_f32_normalize_round(float *y, uint32_t n, float *delta, hipStream_t s);
hipError_t launch_scale_output(const int32_t *acc, uint32_t rows, uint32_t n_vec,
                               const float *s_w, const float *s_x, float *y, hipStream_t s);

}  // namespace mv
