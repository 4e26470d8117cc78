// C ABI: argument validation and dispatch to the HIP kernels.  Every entry point
// returns a matvans_status and never synchronises or allocates.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdlib>

#include "launch.h"
#include "matvans_internal.h"

using mv::fail;

namespace {
inline bool aligned(const void *p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }
inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

int hip_status(hipError_t e, const char *what) {
    if (e != hipSuccess) return fail(MATVANS_E_HIP, "%s: %s", what, hipGetErrorString(e));
    mv::clear_error();
    return MATVANS_OK;
}

// Decode kernel variant, all computing the same bits: 0 = reference form of the step,
// 1 = rotated state with dword tables read as one pair, 2 = rotated state with three
// conflict-free byte tables, 3 = rotated state with one packed dword table (half the LDS
// accesses, one more dependent op).  Unless one is forced (matvans_set_decode_variant, or
// MATVANS_DECODE_VARIANT for A/B runs), each launch picks by the waves per SIMD its grid
// gives: with few (long rows: 4096 x 14336 has two per SIMD) the step's dependent chain
// decides and 1 wins; from four on the instruction count decides and 3 wins (14336 x 4096,
// seven per SIMD: 29.8 vs 33.0 us).
std::atomic<int> g_variant{-2};   // -2 not read yet, -1 automatic, 0..3 forced

int forced_variant() {
    int v = g_variant.load(std::memory_order_relaxed);
    if (v != -2) return v;
    const char *e = std::getenv("MATVANS_DECODE_VARIANT");
    v = (e && e[0] >= '0' && e[0] <= '3' && e[1] == 0) ? e[0] - '0' : -1;
    g_variant.store(v, std::memory_order_relaxed);
    return v;
}

// Compute units of the current device, read once per device.
uint32_t current_device_cus() {
    static std::atomic<uint32_t> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    uint32_t n = cache[dev].load(std::memory_order_relaxed);
    if (n == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            v <= 0)
            v = 256;
        n = uint32_t(v);
        cache[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

int v0_variant_for(uint32_t rows) {
    const int v = forced_variant();
    if (v >= 0) return v;
    const uint64_t waves = (uint64_t(rows) + 1) / 2;   // two rows per wave
    return waves >= 4ull * 4ull * current_device_cus() ? 3 : 1;
}

int check_v0(const int32_t *dst, const int8_t *vec, uint32_t n_vec, uint32_t vec_stride,
             uint32_t rows, uint32_t cols, const uint32_t *cursors, const uint8_t *cdf,
             const uint8_t *ppf, const uint16_t *payload) {
    if (!dst || !vec || !cursors || !cdf || !ppf || !payload)
        return fail(MATVANS_E_INVALID, "null device pointer");
    if (n_vec < 1 || n_vec > 8) return fail(MATVANS_E_INVALID, "n_vec %u outside 1..8", n_vec);
    if (cols == 0 || cols % 512)
        return fail(MATVANS_E_SHAPE, "cols %u is not a multiple of 512 (32 lanes x 16)", cols);
    if (rows == 0) return fail(MATVANS_E_SHAPE, "rows == 0");
    if (!aligned(payload, 16)) return fail(MATVANS_E_ALIGN, "payload not 16-byte aligned");
    if (!aligned(vec, 16) || (n_vec > 1 && vec_stride % 16))
        return fail(MATVANS_E_ALIGN, "vectors must be 16-byte aligned (stride %u)", vec_stride);
    if (n_vec > 1 && vec_stride < cols)
        return fail(MATVANS_E_INVALID, "vec_stride %u < cols %u", vec_stride, cols);
    return MATVANS_OK;
}
}  // namespace

extern "C" int matvans_set_decode_variant(int variant) {
    if (variant < -1 || variant > 3) return fail(MATVANS_E_INVALID, "variant %d outside -1..3", variant);
    g_variant.store(variant == -1 ? -2 : variant, std::memory_order_relaxed);
    mv::clear_error();
    return MATVANS_OK;
}

extern "C" int matvans_get_decode_variant(void) { return forced_variant(); }

extern "C" int matvans_decompress_and_multiply(int32_t *dst, const int8_t *vec, uint32_t rows,
                                               uint32_t cols, float grid_spacing,
                                               const uint32_t *cursors, int8_t min_value,
                                               uint8_t G, const uint8_t *cdf_data,
                                               const uint8_t *ppf_data,
                                               const uint16_t *payload, uint32_t payload_size,
                                               void *stream) {
    (void)grid_spacing;   // unused by the reference kernel as well
    return matvans_decompress_and_multiply_nvec(dst, vec, 1, cols, rows, cols, cursors,
                                                min_value, G, cdf_data, ppf_data, payload,
                                                payload_size, stream);
}

extern "C" int matvans_decompress_and_multiply_nvec(int32_t *dst, const int8_t *vecs,
                                                    uint32_t n_vec, uint32_t vec_stride,
                                                    uint32_t rows, uint32_t cols,
                                                    const uint32_t *cursors, int8_t min_value,
                                                    uint8_t G, const uint8_t *cdf_data,
                                                    const uint8_t *ppf_data,
                                                    const uint16_t *payload,
                                                    uint32_t payload_size, void *stream) {
    if (n_vec == 1 && vec_stride < cols) vec_stride = cols;
    int st = check_v0(dst, vecs, n_vec, vec_stride, rows, cols, cursors, cdf_data, ppf_data,
                      payload);
    if (st) return st;
    mv::V0Args a{dst, vecs, n_vec, vec_stride, rows, cols, cursors, min_value, G,
                 cdf_data, ppf_data, payload, payload_size, v0_variant_for(rows)};
    return hip_status(mv::launch_v0_decode_matvec(a, as_stream(stream)), "decode_matvec launch");
}

extern "C" int matvans_v1_decompress_and_multiply(const matvans_v1_dev_matrix *m, int32_t *dst,
                                                  const int8_t *vecs, uint32_t n_vec,
                                                  uint32_t vec_stride, void *stream) {
    if (!m || !dst || !vecs || !m->cursors || !m->payload)
        return fail(MATVANS_E_INVALID, "null argument or device pointer");
    if (n_vec < 1 || n_vec > 8) return fail(MATVANS_E_INVALID, "n_vec %u outside 1..8", n_vec);
    const bool s_ok = m->streams == 32 || m->streams == 64 || m->streams == 128;
    const bool p_ok = m->precision == 8 || m->precision == 12;
    if (!s_ok || !p_ok || m->table_mode > 1)
        return fail(MATVANS_E_INVALID, "streams %u, precision %u, table mode %u", m->streams,
                    m->precision, m->table_mode);
    if (m->table_mode == 1 && (m->streams != 64 || m->precision != 8))
        return fail(MATVANS_E_UNSUPPORTED, "one table per row needs 64 streams and 8-bit tables");
    if (m->coder > 2) return fail(MATVANS_E_INVALID, "coder %u", m->coder);
    if (m->coder >= 1 && m->precision != 8)
        return fail(MATVANS_E_UNSUPPORTED, "the quad coder needs 8-bit probabilities");
    if (m->coder == 2 && m->table_mode != 0)
        return fail(MATVANS_E_UNSUPPORTED, "coder 2 has one table per matrix");
    if (m->coder == 2 && !mv::lanes_tables_fit(m->rows, m->streams, m->payload_size))
        return fail(MATVANS_E_SHAPE, "payload_size %u is shorter than coder 2's row tables",
                    m->payload_size);
    if (m->table_mode == 0 && !m->lut) return fail(MATVANS_E_INVALID, "null decode table");
    if (m->table_mode == 1 && (!m->freq || m->G < 2 || m->G > 256))
        return fail(MATVANS_E_INVALID, "per-row tables need freq and 2 <= G <= 256");
    if (m->rows == 0) return fail(MATVANS_E_SHAPE, "rows == 0");
    if (m->cols == 0 || m->cols % (16 * m->streams))
        return fail(MATVANS_E_SHAPE, "cols %u is not a multiple of %u", m->cols, 16 * m->streams);
    if (!aligned(m->payload, 16) || (m->lut && !aligned(m->lut, 16)))
        return fail(MATVANS_E_ALIGN, "payload and decode table must be 16-byte aligned");
    if (n_vec == 1 && vec_stride < m->cols) vec_stride = m->cols;
    if (!aligned(vecs, 16) || vec_stride % 16)
        return fail(MATVANS_E_ALIGN, "vectors must be 16-byte aligned (stride %u)", vec_stride);
    if (vec_stride < m->cols) return fail(MATVANS_E_INVALID, "vec_stride %u < cols %u", vec_stride, m->cols);
    if (m->n_outliers && (m->coder == 0 || m->table_mode != 0))
        return fail(MATVANS_E_UNSUPPORTED, "outliers need the quad coder with one table");
    if (m->n_outliers && (!m->out_start || !m->out_rec))
        return fail(MATVANS_E_INVALID, "outliers without device offsets or records");
    mv::V1Args a{dst, vecs, n_vec, vec_stride, m->rows, m->cols, m->streams, m->precision,
                 m->table_mode, m->coder, m->min_value, m->G, m->cursors, m->lut, m->freq, m->payload,
                 m->payload_size, m->n_outliers ? m->out_start : nullptr,
                 m->n_outliers ? m->out_rec : nullptr};
    return hip_status(mv::launch_v1_decode_matvec(a, as_stream(stream)), "v1 decode_matvec launch");
}

namespace {
int b_matvec(int layout, const uint32_t *compressed, uint64_t compressed_words,
             const uint32_t *warp_offsets, uint32_t end_offset, const uint32_t *ppf,
             const uint32_t *ppf_cum, const uint32_t *ppf_prob, const int8_t *vecs,
             uint32_t n_vec, uint32_t vec_stride, uint32_t in_dim, uint32_t out_dim, float scale,
             int8_t *out8, int32_t *acc32, void *stream) {
    if (!compressed || !warp_offsets || !ppf || !ppf_cum || !ppf_prob || !vecs || (!out8 && !acc32))
        return fail(MATVANS_E_INVALID, "null device pointer");
    if (n_vec < 1 || n_vec > 8) return fail(MATVANS_E_INVALID, "n_vec %u outside 1..8", n_vec);
    if (out_dim == 0 || in_dim == 0) return fail(MATVANS_E_SHAPE, "empty matrix");
    if (layout == 0 && in_dim % 128)
        return fail(MATVANS_E_SHAPE, "ans-warp-rows needs in_dim %% 128 == 0 (in_dim = %u)", in_dim);
    if (layout == 1 && (out_dim % 32 || in_dim % 4))
        return fail(MATVANS_E_SHAPE, "ans-warp-stripes needs out_dim %% 32 == 0, in_dim %% 4 == 0");
    if (compressed_words > 0xFFFFFFF0ull) return fail(MATVANS_E_SHAPE, "compressed data over 2^32 words");
    if (!aligned(compressed, 16)) return fail(MATVANS_E_ALIGN, "compressed data not 16-byte aligned");
    if (n_vec == 1 && vec_stride < in_dim) vec_stride = in_dim;
    // the kernels read the vectors as 32-bit words
    if (!aligned(vecs, 4) || vec_stride % 4)
        return fail(MATVANS_E_ALIGN, "vectors must be 4-byte aligned (stride %u)", vec_stride);
    if (vec_stride < in_dim) return fail(MATVANS_E_INVALID, "vec_stride %u < in_dim %u", vec_stride, in_dim);
    mv::BArgs a{layout, acc32, out8, scale, compressed, uint32_t(compressed_words), warp_offsets,
                end_offset, ppf, ppf_cum, ppf_prob, vecs, n_vec, vec_stride, in_dim, out_dim};
    return hip_status(mv::launch_b_decode_matvec(a, as_stream(stream)), "family-B decode_matvec launch");
}
}  // namespace

extern "C" int matvans_b_rows_matvec(const uint32_t *compressed, uint64_t compressed_words,
                                     const uint32_t *warp_offsets, uint32_t end_offset,
                                     const uint32_t *ppf, const uint32_t *ppf_cum,
                                     const uint32_t *ppf_prob, const int8_t *vecs, uint32_t n_vec,
                                     uint32_t vec_stride, uint32_t in_dim, uint32_t out_dim,
                                     float scale, int8_t *out8, int32_t *acc32, void *stream) {
    return b_matvec(0, compressed, compressed_words, warp_offsets, end_offset, ppf, ppf_cum,
                    ppf_prob, vecs, n_vec, vec_stride, in_dim, out_dim, scale, out8, acc32, stream);
}

extern "C" int matvans_b_stripes_matvec(const uint32_t *compressed, uint64_t compressed_words,
                                        const uint32_t *warp_offsets, uint32_t end_offset,
                                        const uint32_t *ppf, const uint32_t *ppf_cum,
                                        const uint32_t *ppf_prob, const int8_t *vecs,
                                        uint32_t n_vec, uint32_t vec_stride, uint32_t in_dim,
                                        uint32_t out_dim, float scale, int8_t *out8,
                                        int32_t *acc32, void *stream) {
    return b_matvec(1, compressed, compressed_words, warp_offsets, end_offset, ppf, ppf_cum,
                    ppf_prob, vecs, n_vec, vec_stride, in_dim, out_dim, scale, out8, acc32, stream);
}

extern "C" int matvans_decompress_and_mult(const matvans_dev_matrix *m, int8_t *d_result8,
                                           int32_t *d_result32, const int8_t *d_vector,
                                           float *d_delta, void *stream) {
    if (!m || !d_result8 || !d_result32 || !d_vector || !d_delta)
        return fail(MATVANS_E_INVALID, "null argument");
    int st = matvans_decompress_and_multiply(d_result32, d_vector, m->rows, m->cols,
                                             m->grid_spacing, m->cursors, m->min_value, m->G,
                                             m->cdf, m->ppf, m->payload, m->payload_size,
                                             stream);
    if (st) return st;
    return matvans_absmax_requant(d_result32, m->rows, d_result8, d_delta, 0, stream);
}

extern "C" int matvans_absmax_requant(const int32_t *d_acc, uint32_t n, int8_t *d_out8,
                                      float *d_delta, int clamp, void *stream) {
    if (!d_acc || !d_out8 || !d_delta) return fail(MATVANS_E_INVALID, "null argument");
    if (n == 0) return fail(MATVANS_E_SHAPE, "n == 0");
    return hip_status(mv::launch_absmax_requant(d_acc, n, d_out8, d_delta, clamp,
                                                as_stream(stream)),
                      "absmax_requant launch");
}

extern "C" int matvans_f32_normalize_round(float *d_y, uint32_t n, float *d_delta, void *stream) {
    if (!d_y || !d_delta) return fail(MATVANS_E_INVALID, "null device pointer");
    if (n == 0) return fail(MATVANS_E_SHAPE, "empty vector");
    return hip_status(mv::launch_f32_normalize_round(d_y, n, d_delta, as_stream(stream)),
                      "f32_normalize_round launch");
}

extern "C" int matvans_scale_output(const int32_t *d_acc, uint32_t rows, uint32_t n_vec,
                                    const float *d_s_w, const float *d_s_x, float *d_y,
                                    void *stream) {
    if (!d_acc || !d_s_w || !d_s_x || !d_y) return fail(MATVANS_E_INVALID, "null argument");
    if (rows == 0 || n_vec == 0 || uint64_t(rows) * n_vec > 0xFFFFFFFFull)
        return fail(MATVANS_E_SHAPE, "bad shape %u x %u", rows, n_vec);
    return hip_status(mv::launch_scale_output(d_acc, rows, n_vec, d_s_w, d_s_x, d_y,
                                              as_stream(stream)),
                      "scale_output launch");
}

extern "C" int matvans_int8_matvec(int32_t *dst, const int8_t *mat, const int8_t *vecs,
                                   uint32_t n_vec, uint32_t vec_stride, uint32_t rows,
                                   uint32_t cols, void *stream) {
    if (!dst || !mat || !vecs) return fail(MATVANS_E_INVALID, "null argument");
    if (n_vec < 1 || n_vec > 8) return fail(MATVANS_E_INVALID, "n_vec %u outside 1..8", n_vec);
    if (rows == 0 || cols == 0 || cols % 16)
        return fail(MATVANS_E_SHAPE, "cols %u must be a positive multiple of 16", cols);
    if (n_vec == 1 && vec_stride < cols) vec_stride = cols;
    if (!aligned(mat, 16) || !aligned(vecs, 16) || vec_stride % 16)
        return fail(MATVANS_E_ALIGN, "matrix and vectors must be 16-byte aligned");
    return hip_status(mv::launch_int8_matvec(dst, mat, vecs, n_vec, vec_stride, rows, cols,
                                             current_device_cus(), as_stream(stream)),
                      "int8_matvec launch");
}

extern "C" int matvans_f32_gemv(float *y, const float *mat, const float *x, uint32_t rows,
                                uint32_t cols, float alpha, void *stream) {
    if (!y || !mat || !x) return fail(MATVANS_E_INVALID, "null argument");
    if (rows == 0 || cols == 0 || cols % 4)
        return fail(MATVANS_E_SHAPE, "cols %u must be a positive multiple of 4", cols);
    if (!aligned(mat, 16) || !aligned(x, 16))
        return fail(MATVANS_E_ALIGN, "matrix and x must be 16-byte aligned");
    return hip_status(mv::launch_f32_gemv(y, mat, x, rows, cols, alpha, as_stream(stream)),
                      "f32_gemv launch");
}

extern "C" size_t matvans_backward_workspace_size(uint32_t rows, uint32_t cols,
                                                  uint32_t n_vec) {
    return mv::v0_bwd_workspace_bytes(rows, cols, n_vec);
}

extern "C" int matvans_decompress_backward(float *dx, const float *g, uint32_t n_vec,
                                           uint32_t g_stride, const float *s_w,
                                           const float *alpha, uint32_t rows, uint32_t cols,
                                           const uint32_t *cursors, int8_t min_value, uint8_t G,
                                           const uint8_t *cdf_data, const uint8_t *ppf_data,
                                           const uint16_t *payload, uint32_t payload_size,
                                           float *workspace, size_t workspace_bytes,
                                           void *stream) {
    if (!dx || !g || !s_w || !alpha || !cursors || !cdf_data || !ppf_data || !payload ||
        !workspace)
        return fail(MATVANS_E_INVALID, "null device pointer");
    if (n_vec < 1 || n_vec > 8) return fail(MATVANS_E_INVALID, "n_vec %u outside 1..8", n_vec);
    if (rows == 0 || cols == 0 || cols % 512)
        return fail(MATVANS_E_SHAPE, "shape %u x %u (cols must be a multiple of 512)", rows, cols);
    if (!mv::bwd_partials_fit(rows, cols))
        return fail(MATVANS_E_SHAPE, "%u x %u: partial sums beyond 4 GiB", rows, cols);
    if (n_vec > 1 && g_stride < rows) return fail(MATVANS_E_INVALID, "g_stride %u < rows", g_stride);
    if (!aligned(payload, 16) || !aligned(workspace, 16) || !aligned(dx, 4))
        return fail(MATVANS_E_ALIGN, "payload and workspace must be 16-byte aligned");
    size_t need = mv::v0_bwd_workspace_bytes(rows, cols, n_vec);
    if (workspace_bytes < need)
        return fail(MATVANS_E_CAPACITY, "workspace %zu bytes < %zu", workspace_bytes, need);
    mv::V0BwdArgs a{dx, workspace, g, n_vec, n_vec > 1 ? g_stride : rows, s_w, alpha, rows,
                    cols, cursors, min_value, G, cdf_data, ppf_data, payload, payload_size};
    return hip_status(mv::launch_v0_backward(a, as_stream(stream)), "backward launch");
}

namespace {
// Argument checks shared by the quad backward entry points (all but the workspace).
int check_v1_backward(const matvans_v1_dev_matrix *m, const float *dx, const float *g,
                      uint32_t n_vec, uint32_t g_stride, const float *s_w, const float *alpha) {
    if (!m || !dx || !g || !s_w || !alpha) return fail(MATVANS_E_INVALID, "null argument");
    if ((m->coder != 1 && m->coder != 2) || m->precision != 8 || m->table_mode > 1 ||
        (m->streams != 32 && m->streams != 64 && m->streams != 128) ||
        (m->table_mode == 1 && (m->streams != 64 || m->coder != 1)))
        return fail(MATVANS_E_UNSUPPORTED,
                    "backward needs the quad coder with 8-bit tables: one per matrix with 32, "
                    "64 or 128 streams, or one per row with 64 (coder 1)");
    if (!m->cursors || !m->payload || (m->table_mode == 0 && !m->lut))
        return fail(MATVANS_E_INVALID, "null device pointer");
    if (m->coder == 2 && !mv::lanes_tables_fit(m->rows, m->streams, m->payload_size))
        return fail(MATVANS_E_SHAPE, "payload_size %u is shorter than coder 2's row tables",
                    m->payload_size);
    if (m->table_mode == 1 && (!m->freq || m->G < 2 || m->G > 256))
        return fail(MATVANS_E_INVALID, "per-row tables need freq and 2 <= G <= 256");
    if (n_vec < 1 || n_vec > 8) return fail(MATVANS_E_INVALID, "n_vec %u outside 1..8", n_vec);
    if (m->rows == 0 || m->cols == 0 || m->cols % (16 * m->streams))
        return fail(MATVANS_E_SHAPE, "shape %u x %u (cols must be a multiple of %u)", m->rows,
                    m->cols, 16 * m->streams);
    if (!mv::bwd_partials_fit(m->rows, m->cols))
        return fail(MATVANS_E_SHAPE, "%u x %u: partial sums beyond 4 GiB", m->rows, m->cols);
    if (n_vec > 1 && g_stride < m->rows)
        return fail(MATVANS_E_INVALID, "g_stride %u < rows", g_stride);
    if (!aligned(m->payload, 16) || !aligned(dx, 4))
        return fail(MATVANS_E_ALIGN, "payload must be 16-byte aligned");
    if (m->n_outliers && (m->table_mode != 0 || !m->out_col_start || !m->out_col_rec))
        return fail(MATVANS_E_INVALID, "outliers need one table per matrix and their "
                    "by-column device lists");
    return MATVANS_OK;
}

mv::QuadBwdArgs v1_backward_args(const matvans_v1_dev_matrix *m, float *dx, const float *g,
                                 uint32_t n_vec, uint32_t g_stride, const float *s_w,
                                 const float *alpha, float *partial) {
    return mv::QuadBwdArgs{dx, partial, g, n_vec, n_vec > 1 ? g_stride : m->rows, s_w, alpha,
                           m->rows, m->cols, m->streams, m->cursors, m->lut, m->payload,
                           m->payload_size, m->table_mode, m->freq, int32_t(m->min_value),
                           uint32_t(m->G), m->n_outliers ? m->out_col_start : nullptr,
                           m->n_outliers ? m->out_col_rec : nullptr, m->coder};
}
}  // namespace

extern "C" int matvans_v1_decompress_backward(const matvans_v1_dev_matrix *m, float *dx,
                                              const float *g, uint32_t n_vec, uint32_t g_stride,
                                              const float *s_w, const float *alpha,
                                              float *workspace, size_t workspace_bytes,
                                              void *stream) {
    if (!workspace) return fail(MATVANS_E_INVALID, "null argument");
    int st = check_v1_backward(m, dx, g, n_vec, g_stride, s_w, alpha);
    if (st != MATVANS_OK) return st;
    if (!aligned(workspace, 16))
        return fail(MATVANS_E_ALIGN, "payload and workspace must be 16-byte aligned");
    size_t need = mv::v0_bwd_workspace_bytes(m->rows, m->cols, n_vec);
    if (workspace_bytes < need)
        return fail(MATVANS_E_CAPACITY, "workspace %zu bytes < %zu", workspace_bytes, need);
    const mv::QuadBwdArgs a = v1_backward_args(m, dx, g, n_vec, g_stride, s_w, alpha, workspace);
    return hip_status(mv::launch_quad_backward(a, as_stream(stream)), "quad backward launch");
}

extern "C" int matvans_v1_decompress_backward_noalloc(const matvans_v1_dev_matrix *m, float *dx,
                                                      const float *g, uint32_t n_vec,
                                                      uint32_t g_stride, const float *s_w,
                                                      const float *alpha, void *stream) {
    int st = check_v1_backward(m, dx, g, n_vec, g_stride, s_w, alpha);
    if (st != MATVANS_OK) return st;
    const size_t need = mv::v0_bwd_workspace_bytes(m->rows, m->cols, n_vec);
    if (need > MATVANS_NOALLOC_SCRATCH_BYTES)
        return fail(MATVANS_E_CAPACITY,
                    "%u x %u with %u vectors needs %zu bytes of partial sums; the static "
                    "scratch holds %zu (use matvans_v1_decompress_backward with a workspace)",
                    m->rows, m->cols, n_vec, need, size_t(MATVANS_NOALLOC_SCRATCH_BYTES));
    float *scratch = nullptr;
    st = hip_status(mv::backward_scratch(&scratch), "static scratch address");
    if (st != MATVANS_OK) return st;
    const mv::QuadBwdArgs a = v1_backward_args(m, dx, g, n_vec, g_stride, s_w, alpha, scratch);
    return hip_status(mv::launch_quad_backward(a, as_stream(stream)), "quad backward launch");
}

extern "C" int matvans_quantize_vectors(const float *x, uint32_t n_vec, uint32_t x_stride,
                                        uint32_t cols, int8_t *xq, uint32_t xq_stride,
                                        float *s_x, void *stream) {
    if (!x || !xq || !s_x) return fail(MATVANS_E_INVALID, "null device pointer");
    if (n_vec < 1 || cols == 0) return fail(MATVANS_E_SHAPE, "empty input");
    if (x_stride < cols || xq_stride < cols)
        return fail(MATVANS_E_INVALID, "strides must be >= cols");
    return hip_status(mv::launch_quantize_vectors(x, n_vec, x_stride, cols, xq, xq_stride, s_x,
                                                  as_stream(stream)),
                      "quantize launch");
}
