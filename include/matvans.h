/*
 * matvans.h -- C ABI of the MI355X (gfx950) fused rANS-decode x int8 matvec library.
 *
 * Every entry point takes plain pointers and sizes.  Device pointers are HIP device
 * memory owned by the caller; `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).  Launch entry points are stream-ordered, never synchronise the host,
 * never allocate and never exit the process: they return a matvans_status and leave a
 * message in matvans_last_error() (thread-local).
 *
 * Reference interfaces are cited as file:line of wildug/MatVANS.
 */
#ifndef MATVANS_H
#define MATVANS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference reports errors by fprintf + exit(-1)
 * (matvansWarpPerRow/matvansWarpPerRow.cu:416-424); this library returns them. */
typedef enum {
    MATVANS_OK = 0,
    MATVANS_E_INVALID = 1,     /* null pointer or out-of-range argument */
    MATVANS_E_SHAPE = 2,       /* shape the layout cannot represent (cols % 512 != 0, ...) */
    MATVANS_E_ALIGN = 3,       /* pointer not 16-byte aligned */
    MATVANS_E_FORMAT = 4,      /* malformed or truncated .bin data */
    MATVANS_E_IO = 5,          /* file cannot be opened / written */
    MATVANS_E_HIP = 6,         /* HIP runtime error (launch, ...) */
    MATVANS_E_RANGE = 7,       /* symbol outside the entropy model */
    MATVANS_E_UNSUPPORTED = 8, /* valid request the library does not implement */
    MATVANS_E_CAPACITY = 9     /* caller buffer too small */
} matvans_status;

/* Message of the last failing call on this thread ("" if none). */
const char *matvans_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int matvans_version(void);

/* ------------------------------------------------------------------------------
 * A1. Fused decode + int32 matvec, the drop-in for the kernel
 *     __global__ void decmpressAndMultiply(int32_t* dst, int8_t* vec, uint32_t rows,
 *         uint32_t cols, float grid_spacing, const uint32_t* cursors, int8_t min_value,
 *         uint8_t G, const uint8_t* cdf_data, const uint8_t* ppf_data,
 *         uint16_t* payload, uint32_t payload_size)
 *     (matvansWarpPerRow/matvansWarpPerRow-kernel.cu:46-50).
 * Same arguments with the same meaning, all pointers device pointers:
 *   dst[rows] = sum_c W[r,c] * vec[c] (int32, bit-exact), W decoded from the 32
 *   interleaved rANS streams of each row (payload in u16 words, row r at cursors[r]).
 * grid_spacing is accepted and unused, as in the reference.  payload_size bounds every
 * read (the reference ignores it and may read past a row's end).
 * Preconditions: cols % 512 == 0; payload and vec 16-byte aligned.
 * ----------------------------------------------------------------------------*/
int matvans_decompress_and_multiply(int32_t *dst, const int8_t *vec, uint32_t rows,
                                    uint32_t cols, float grid_spacing,
                                    const uint32_t *cursors, int8_t min_value, uint8_t G,
                                    const uint8_t *cdf_data, const uint8_t *ppf_data,
                                    const uint16_t *payload, uint32_t payload_size,
                                    void *stream);

/* Decode kernel variant of the .bin v0 launch entry points (all give the same bits):
 * 0 = reference form of the step, 1 = rotated state with dword tables, 2 = rotated state
 * with three conflict-free 256-byte tables, 3 = rotated state with one packed table dword
 * per quantile.  By default each launch picks 1 or 3 by the waves per SIMD of its grid
 * (3 from four on); set 0..3 to force one, -1 to restore the default (or
 * MATVANS_DECODE_VARIANT).  get returns the forced variant, -1 when automatic.  No
 * reference counterpart: a tuning control for measurements.  Process-wide. */
int matvans_set_decode_variant(int variant);
int matvans_get_decode_variant(void);

/* A1 with 1..8 activation vectors decoded once: vecs[v*vec_stride + c] (int8),
 * dst[v*rows + r] (int32).  No reference counterpart (the reference multiplies one
 * vector per launch, matvansWarpPerRow.cu:161-164). */
int matvans_decompress_and_multiply_nvec(int32_t *dst, const int8_t *vecs, uint32_t n_vec,
                                         uint32_t vec_stride, uint32_t rows, uint32_t cols,
                                         const uint32_t *cursors, int8_t min_value,
                                         uint8_t G, const uint8_t *cdf_data,
                                         const uint8_t *ppf_data, const uint16_t *payload,
                                         uint32_t payload_size, void *stream);

/* ------------------------------------------------------------------------------
 * A6. Host composite CompressedMatrix::decompressAndMult(int8_t* d_result8,
 *     int32_t* d_result32, int8_t* vector, float v_delta) -> float
 *     (matvansWarpPerRow/matvansWarpPerRow.cu:147,150-176).
 * A matrix already resident on the device is described by this struct of device
 * pointers (the reference keeps them in CompressedMatrix::d_*, :89-92).
 * The composite launches the fused decode-matvec and the A7 epilogue; the new
 * v_delta (absmax/127) is written to *d_delta on the device instead of being returned
 * through a host synchronisation (the reference's thrust reduce syncs, :167).
 * ----------------------------------------------------------------------------*/
typedef struct {
    uint32_t rows, cols;
    float grid_spacing;
    int8_t min_value;
    uint8_t G;
    uint16_t reserved;
    uint32_t payload_size;     /* u16 words */
    const uint32_t *cursors;   /* device, [rows] */
    const uint8_t *cdf;        /* device, [G+1] */
    const uint8_t *ppf;        /* device, [256] */
    const uint16_t *payload;   /* device, [payload_size], 16-byte aligned */
} matvans_dev_matrix;

int matvans_decompress_and_mult(const matvans_dev_matrix *m, int8_t *d_result8,
                                int32_t *d_result32, const int8_t *d_vector,
                                float *d_delta, void *stream);

/* ------------------------------------------------------------------------------
 * A7. Epilogue: a = float(max|acc|) (absMaxWithThrustDevice, matvansWarpPerRow.cu:51-60),
 *     delta = a / 127, out8[i] = (int8) rint(float(acc[i]) / delta)
 *     (normalizeAndRoundtoInt8, :62-72; clamp != 0 adds the clamp to [-128,127] of
 *     baselines/baseline_int8/baseline_matVec_custom.cu:75-86).  delta written to
 *     *d_delta (device).  All in IEEE fp32, round-half-even, no host sync.
 * ----------------------------------------------------------------------------*/
int matvans_absmax_requant(const int32_t *d_acc, uint32_t n, int8_t *d_out8, float *d_delta,
                           int clamp, void *stream);

/* A12 epilogue, the fp32 chain's: a = max|y| (absMaxWithThrustDevice,
 * baselines/cuBLAS_baseline_float/baseline_matVec_cuBLAS.cu:48-57), delta = a / 127,
 * y[i] = roundf(y[i] / delta) in place (normalizeAndRound, :59-64, launched at :129-135):
 * IEEE fp32, rounding half away from zero; delta written to *d_delta (device).  As in the
 * reference, y == 0 gives delta 0 and NaN outputs. */
int matvans_f32_normalize_round(float *d_y, uint32_t n, float *d_delta, void *stream);

/* Scaled fp32 output y[v*rows + r] = s_x[v] * s_w[r] * float(acc[v*rows + r]).
 * s_w is the per-row weight scale (device, [rows]), s_x per vector (device, [n_vec]).
 * No reference counterpart (the .bin grid_spacing is always 1.0,
 * create_matrices_WarpPerRow.py:430). */
int matvans_scale_output(const int32_t *d_acc, uint32_t rows, uint32_t n_vec,
                         const float *d_s_w, const float *d_s_x, float *d_y, void *stream);

/* ------------------------------------------------------------------------------
 * A11. Uncompressed int8 comparator, replaces vecMat_int8_warpRows
 *      (baselines/baseline_int8/baseline_matVec_custom.cu:99-155):
 *      dst[v*rows + r] = sum_c mat[r*cols + c] * vecs[v*vec_stride + c] (int32), mat
 *      row-major int8, 16-byte aligned, cols % 16 == 0, 1 <= n_vec <= 8.
 * A12. fp32 comparator, replaces cublasSgemv (baselines/cuBLAS_baseline_float/
 *      baseline_matVec_cuBLAS.cu:111-144): y[r] = alpha * sum_c mat[r*cols+c] * x[c],
 *      mat row-major fp32, cols % 4 == 0.
 * ----------------------------------------------------------------------------*/
int matvans_int8_matvec(int32_t *dst, const int8_t *mat, const int8_t *vecs, uint32_t n_vec,
                        uint32_t vec_stride, uint32_t rows, uint32_t cols, void *stream);
int matvans_f32_gemv(float *y, const float *mat, const float *x, uint32_t rows, uint32_t cols,
                     float alpha, void *stream);

/* ------------------------------------------------------------------------------
 * N3. Backward from the compressed streams (no reference counterpart):
 *     dx[v*cols + c] = alpha[v] * sum_r W[r,c] * s_w[r] * g[v*g_stride + r]
 * re-decoding the same row-ordered streams as the forward (no dense W is built).  The
 * 64-row partial sums go to the caller's workspace (matvans_backward_workspace_size
 * bytes, 16-byte aligned) and are reduced in a fixed order: results are bitwise
 * reproducible.  1 <= n_vec <= 8, cols % 512 == 0.
 * ----------------------------------------------------------------------------*/
size_t matvans_backward_workspace_size(uint32_t rows, uint32_t cols, uint32_t n_vec);
int matvans_decompress_backward(float *dx, const float *g, uint32_t n_vec, uint32_t g_stride,
                                const float *s_w, const float *alpha, uint32_t rows,
                                uint32_t cols, const uint32_t *cursors, int8_t min_value,
                                uint8_t G, const uint8_t *cdf_data, const uint8_t *ppf_data,
                                const uint16_t *payload, uint32_t payload_size,
                                float *workspace, size_t workspace_bytes, void *stream);

/* Activation quantiser of the compressed layer (no reference counterpart; the reference
 * requantises the int32 result instead, A7): per vector s_x = max|x|/127 (1 for x == 0)
 * and xq = clamp(rint(x / s_x), -127, 127), fp32 IEEE division. */
int matvans_quantize_vectors(const float *x, uint32_t n_vec, uint32_t x_stride, uint32_t cols,
                             int8_t *xq, uint32_t xq_stride, float *s_x, void *stream);

/* ------------------------------------------------------------------------------
 * A10. FxHash of an int8 array (matvansWarpPerRow.cu:31-42,
 *      create_matrices_WarpPerRow.py:58-70).  Host memory.
 * ----------------------------------------------------------------------------*/
uint32_t matvans_fxhash(const int8_t *data, uint64_t n);

/* ------------------------------------------------------------------------------
 * A5 / A9. The .bin v0 container (writer create_matrices_WarpPerRow.py:286-318,440-461;
 * reader matvansWarpPerRow.cu:100-145,225-258).  Host memory.
 * ----------------------------------------------------------------------------*/
typedef struct {
    uint32_t rows, cols;
    float grid_spacing;
    int8_t min_value;
    uint8_t G;
    uint16_t reserved;
    uint32_t payload_size;     /* u16 words */
    const uint32_t *cursors;   /* [rows] */
    const uint8_t *cdf;        /* [G+1], cdf & 0xFF */
    const uint8_t *ppf;        /* [256] */
    const uint16_t *payload;   /* [payload_size] */
} matvans_host_matrix;

typedef struct matvans_bin matvans_bin;   /* parsed file, owned by the library */

/* Parse a whole .bin file.  Validates every length against the file size and every
 * record's cursors (nondecreasing multiples of 256 inside the payload). */
int matvans_bin_read(const char *path, matvans_bin **out);
void matvans_bin_free(matvans_bin *bin);
/* Header: num_matrices, result_hash, max_word_count (:440-442); start vector (:444-453). */
int matvans_bin_header(const matvans_bin *bin, uint32_t *num_matrices, uint32_t *result_hash,
                       uint32_t *max_word_count, const int8_t **vector, uint32_t *len_v);
int matvans_bin_matrix(const matvans_bin *bin, uint32_t k, matvans_host_matrix *out);
/* Write a .bin v0 file; max_word_count is computed as the writer does (:286-288,455). */
int matvans_bin_write(const char *path, uint32_t result_hash, const int8_t *vector,
                      uint32_t len_v, uint32_t num_matrices, const matvans_host_matrix *mats);

/* ------------------------------------------------------------------------------
 * A11 harness. The uncompressed baseline container raw-matrices_*.bin (writer
 * create_matrices_WarpPerRow.py:94-129; reader Matrix::deserialize and main,
 * baselines/baseline_int8/baseline_matVec_custom.cu:371-393,480-508).  Host memory.
 * The vector's pad to 4 bytes is skipped (the reference reader assumes len_v % 4 == 0);
 * the pad byte after an odd-sized matrix is skipped as the reference does (:388-391).
 * ----------------------------------------------------------------------------*/
typedef struct matvans_raw matvans_raw;   /* parsed file, owned by the library */

int matvans_raw_read(const char *path, matvans_raw **out);
void matvans_raw_free(matvans_raw *raw);
/* Header: num_matrices, result_hash (:482-484); start vector (:487-490). */
int matvans_raw_header(const matvans_raw *raw, uint32_t *num_matrices, uint32_t *result_hash,
                       const int8_t **vector, uint32_t *len_v);
/* Matrix k: shape, grid_spacing and its row-major int8 entries (valid until raw_free). */
int matvans_raw_matrix(const matvans_raw *raw, uint32_t k, uint32_t *rows, uint32_t *cols,
                       float *grid_spacing, const int8_t **data);

/* Entropy model (create_matrices_WarpPerRow.py:131-187) and ppf (:189-208).
 * cdf16 receives G+1 <= 257 uint16 entries (cdf16[G] == 256), ppf 256 entries.
 * Counts with equal frequency are ordered by descending value (the reference inherits
 * numpy's argsort order there, which is not specified). */
int matvans_build_model(const int8_t *data, uint64_t n, int8_t *min_value, uint32_t *G,
                        uint16_t *cdf16, uint8_t *ppf);

/* Upper bound on the payload words matvans_encode_rows can produce. */
uint64_t matvans_encode_bound(uint32_t rows, uint32_t cols);
/* 32-stream row-interleaved encode (create_matrices_WarpPerRow.py:353-430), rows x cols
 * (cols % 512 == 0), threaded over rows.  Padding slots are written as 0.  Writes
 * cursors[rows] and payload; *payload_size = words used. */
int matvans_encode_rows(const int8_t *mat, uint32_t rows, uint32_t cols, int8_t min_value,
                        uint32_t G, const uint16_t *cdf16, uint32_t n_threads,
                        uint32_t *cursors, uint16_t *payload, uint64_t payload_cap,
                        uint64_t *payload_size);

/* ------------------------------------------------------------------------------
 * N2. MVX1: format extensions of the WarpPerRow codec, written by this library's encoder.
 * The reference has one variant (32 streams per row, 8-bit probabilities, one table per
 * matrix: create_matrices_WarpPerRow.py:131-272,353-430); its per-matrix constant-memory
 * CDF variant is matvansWarpPerRow_local_cdf.cu:18,196-363.  MVX1 adds
 *   streams    32, 64 or 128 interleaved streams per row.  Stream l decodes columns
 *              16*S*(k/16) + 16*l + k%16.  A row is K granules of 512 bytes in which each
 *              stream owns 256/S consecutive u16 words (8 for S = 32: the reference
 *              layout; 4 for S = 64; 2 for S = 128): word j of stream l is at
 *              cursor + 256*(j/(256/S)) + (256/S)*l + j%(256/S).
 *   precision  8 or 12 bit probabilities (decode q = x & (2^P-1);
 *              x = (x >> P)*f + q - cdf[s]; renormalise with a 16-bit word below 2^16).
 *   table_mode 0: one table per matrix, u16 cdf[G+1] (cdf[G] = 2^P);
 *              1: one 8-bit table per row (P = 8 only), u8 freq[rows][G] summing to 256;
 *                 decoded with S = 64 (one row per wave builds its own table).
 *   coder      0: one symbol per step (32-bit state, 16-bit words, as the reference);
 *              1: quad coder (P = 8, either table mode): 64-bit state x, 32-bit words, a step
 *              decodes the four bytes of the low word of x (if x < 2^32 first
 *              x = x << 32 | next word; then x = (x >> 32)*f0 f1 f2 f3 + r0 + f0(r1 +
 *              f1(r2 + f2 r3)), r_i = q_i - cdf[s_i]) -- the family-B quad coder of
 *              buffered/encoder.py:198-237 with 8-bit quantiles and one stream per coder.
 *              Words are u32 (two u16 of the payload): word j of stream l at u16 index
 *              cursor + 256*(j/(128/S)) + 2*((128/S)*l + j%(128/S)).
 *              2: the quad coder's streams (P = 8, table_mode 0) stored back to back: a
 *              row at u16 index cursor (a multiple of 8) is S u16 end offsets e[l] (the
 *              u32 words of streams 0..l), then the streams' u32 words (stream l: words
 *              e[l-1] .. e[l]-1 after the offsets, e[-1] = 0), zero words to a multiple
 *              of four.  No per-stream padding; a row holds at most 65535 stream words.
 * S = 32, P = 8, table_mode 0, coder 0 encodes exactly as the reference generator.
 * ----------------------------------------------------------------------------*/
typedef struct {
    uint32_t rows, cols;
    float grid_spacing;
    uint32_t streams, precision, table_mode, coder;
    int32_t min_value;          /* symbols min_value .. min_value + G - 1 */
    uint32_t G;
    uint32_t payload_size;      /* u16 words, a multiple of 256 (coder 2: of 8) */
    const uint32_t *cursors;    /* [rows] */
    const uint16_t *cdf;        /* table_mode 0: [G+1] */
    const uint8_t *freq;        /* table_mode 1: [rows][G] */
    const uint16_t *payload;    /* [payload_size] */
    /* outlier split (quad coder, table_mode 0; n_outliers 0: none): entries outside the
     * coded band are 0 in the streams and listed here, by row */
    uint32_t n_outliers;
    const uint32_t *out_start;  /* [rows + 1] offsets into out_rec */
    const uint32_t *out_rec;    /* [n_outliers] col | (uint8)value << 24, rows in order */
} matvans_v1_host;

/* Model for the given precision and table mode.  table_mode 0 writes cdf[G+1] (room for
 * 257 entries); table_mode 1 writes freq[rows][G] (room for rows*256 bytes).  Equal
 * counts are ordered by descending symbol, as matvans_build_model. */
int matvans_v1_model(const int8_t *mat, uint32_t rows, uint32_t cols, uint32_t precision,
                     uint32_t table_mode, int32_t *min_value, uint32_t *G, uint16_t *cdf,
                     uint8_t *freq);
uint64_t matvans_v1_encode_bound(uint32_t rows, uint32_t cols, uint32_t streams,
                                 uint32_t precision);
/* Encode rows x cols (cols % (16*streams) == 0), threaded over rows. */
int matvans_v1_encode(const int8_t *mat, uint32_t rows, uint32_t cols, uint32_t streams,
                      uint32_t precision, uint32_t table_mode, uint32_t coder,
                      int32_t min_value, uint32_t G, const uint16_t *cdf, const uint8_t *freq,
                      uint32_t n_threads, uint32_t *cursors, uint16_t *payload,
                      uint64_t payload_cap, uint64_t *payload_size);
/* Outlier split (no reference counterpart; the reference's 8-bit model gives each of a
 * long tail's rare values at least 1/256, create_matrices_WarpPerRow.py:154).  band: the
 * T in 1..127 minimising the estimated stored bits of the 8-bit quad encoding of the
 * entries in [-T, T] with every entry beyond coded as 0 plus a 32-bit record (127: none).
 * split: coded = mat with |w| > band set to 0; out_start[rows + 1] and the records
 * out_rec[] = col | (uint8)w << 24 in row-major order (cap records at most). */
int matvans_v1_outlier_band(const int8_t *mat, uint64_t n, int32_t *band);
int matvans_v1_split_outliers(const int8_t *mat, uint32_t rows, uint32_t cols, int32_t band,
                              int8_t *coded, uint32_t *out_start, uint32_t *out_rec,
                              uint64_t cap, uint64_t *n_out);
/* Decode table of a table_mode 0 matrix, 2 * 2^P entries: A[q] = f | (uint8)(min+s) << 24,
 * then B[q] = q - cdf[s]. */
int matvans_v1_lut(const uint16_t *cdf, uint32_t G, int32_t min_value, uint32_t precision,
                   uint32_t *lut);

typedef struct matvans_v1_file matvans_v1_file;
/* MVX1 container: "MVX1", u32 version (1, or 2 when a matrix has outliers: each record
 * then ends with u32 n_outliers and, when nonzero, out_start[rows + 1], out_rec[n]),
 * u32 num_matrices, u32 result_hash, u32 len_v,
 * i8 vector[len_v] (zero pad to 4 bytes), then per matrix the fields of matvans_v1_host
 * as 4-byte words, cursors, the table (zero pad to 4), the payload (zero
 * pad to 4); the record header is rows, cols, grid_spacing, streams, precision,
 * table_mode, coder, min_value, G, payload_size.  The reader validates lengths,
 * parameters, tables and cursors. */
int matvans_v1_read(const char *path, matvans_v1_file **out);
void matvans_v1_free(matvans_v1_file *f);
int matvans_v1_header(const matvans_v1_file *f, uint32_t *num_matrices, uint32_t *result_hash,
                      const int8_t **vector, uint32_t *len_v);
int matvans_v1_matrix(const matvans_v1_file *f, uint32_t k, matvans_v1_host *out);
int matvans_v1_write(const char *path, uint32_t result_hash, const int8_t *vector,
                     uint32_t len_v, uint32_t num_matrices, const matvans_v1_host *mats);

/* Device view of an MVX1 matrix (all pointers device memory). */
typedef struct {
    uint32_t rows, cols, streams, precision, table_mode, coder;
    int32_t min_value;
    uint32_t G;
    uint32_t payload_size;      /* u16 words */
    const uint32_t *cursors;    /* [rows] */
    const uint32_t *lut;        /* table_mode 0: [2 * 2^P] from matvans_v1_lut */
    const uint8_t *freq;        /* table_mode 1: [rows][G] */
    const uint16_t *payload;    /* 16-byte aligned */
    /* outlier split (zeros: none): the host lists by row, and by column for the backward */
    uint32_t n_outliers;
    const uint32_t *out_start;      /* [rows + 1] */
    const uint32_t *out_rec;        /* col | (uint8)value << 24, by row */
    const uint32_t *out_col_start;  /* [cols + 1] */
    const uint32_t *out_col_rec;    /* row | (uint8)value << 24, by column, rows ascending */
} matvans_v1_dev_matrix;

/* Upload (matvansWarpPerRow.cu:287-313, the per-matrix cudaMalloc + cudaMemcpy of
 * cursors, cdf, ppf and payload): device copies of a host matrix owned by the returned
 * handle, and the device view to pass to the decode entry points.  The payload gets 1 KiB
 * of zero slack.  Synchronous; call matvans_upload_free to release the device memory. */
typedef struct matvans_upload matvans_upload;
int matvans_matrix_upload(const matvans_host_matrix *m, matvans_upload **out,
                          matvans_dev_matrix *view);
int matvans_v1_matrix_upload(const matvans_v1_host *m, matvans_upload **out,
                             matvans_v1_dev_matrix *view);
void matvans_upload_free(matvans_upload *u);

/* dst[v][r] = sum_c W[r,c] * vecs[v][c] for v < n_vec (1..8), W decoded from an MVX1
 * matrix (int32, bit-exact).  vecs: n_vec vectors vec_stride bytes apart, 16-byte
 * aligned.  The fused decode-matvec of A1 for the extended formats. */
int matvans_v1_decompress_and_multiply(const matvans_v1_dev_matrix *m, int32_t *dst,
                                       const int8_t *vecs, uint32_t n_vec,
                                       uint32_t vec_stride, void *stream);

/* N3 on the quad coder: the same backward from an MVX1 matrix with coder 1 or 2 and 8-bit
 * tables, either one table per matrix (32, 64 or 128 streams) or, coder 1, one per row
 * (64 streams; other encodings: MATVANS_E_UNSUPPORTED):
 *     dx[v*cols + c] = alpha[v] * sum_r W[r,c] * s_w[r] * g[v*g_stride + r]
 * Workspace: matvans_backward_workspace_size(rows, cols, n_vec) bytes, 16-byte aligned.
 * Bitwise reproducible.  No reference counterpart. */
int matvans_v1_decompress_backward(const matvans_v1_dev_matrix *m, float *dx,
                                   const float *g, uint32_t n_vec, uint32_t g_stride,
                                   const float *s_w, const float *alpha, float *workspace,
                                   size_t workspace_bytes, void *stream);

/* No counterpart in the reference: an entry point that allocates nothing, for callers
 * that cannot call hipMalloc (code under HIP-graph capture, tight serving loops,
 * harnesses that forbid allocation).  The same backward as
 * matvans_v1_decompress_backward (same kernels, bitwise the same dx), with the 64-row
 * partial sums in a library-owned static device buffer of MATVANS_NOALLOC_SCRATCH_BYTES
 * per GPU instead of a caller workspace.  MATVANS_E_CAPACITY when
 * matvans_backward_workspace_size(rows, cols, n_vec) exceeds it (e.g. 4096 x 14336 and
 * 14336 x 4096 fit with up to 8 vectors).  Every call on one GPU shares that buffer, and
 * it keeps the last call's contents (each call overwrites all it reads): calls ordered on
 * one stream are safe, calls that overlap in time on one GPU race. */
#define MATVANS_NOALLOC_SCRATCH_BYTES (32u << 20)
int matvans_v1_decompress_backward_noalloc(const matvans_v1_dev_matrix *m, float *dx,
                                           const float *g, uint32_t n_vec, uint32_t g_stride,
                                           const float *s_w, const float *alpha, void *stream);


/* ------------------------------------------------------------------------------
 * A13. The reference's family-B codec ("buffered"): quad rANS with 7-bit probabilities,
 * 64-bit coder states and 32-bit words, 32 coders per warp sharing one word stream.
 * Layouts ans-warp-rows (a warp is one row; compress_ans_warp_rows, buffered/encoder.py:
 * 381-430) and ans-warp-stripes (a warp is 32 consecutive rows, one coder per row;
 * compress_ans_warp_stripes, encoder.py:323-378).  Files are safetensors containers
 * (compress_all_matrices, encoder.py:474-510).
 * ----------------------------------------------------------------------------*/

/* Fused decode x int8 matvec, the drop-in for the kernels
 *   mat_vec_ans_warp_per_row(uint4* out_vec, const uint4* compressed,
 *       const uint32_t* warp_offsets, const uint32_t* ppf, const uint32_t* ppf_cum,
 *       const uint32_t* ppf_prob, const uint4* in_vec, int in_dim, int out_dim,
 *       float scale)                       (buffered/ans-warp-rows-kernel.cu:7-17)
 *   mat_vec_ans_thread_per_row(same)      (buffered/ans-warp-stripes-kernel.cu:21-31)
 * Same arguments and meaning (device pointers; warp_offsets in 16-byte units; the three
 * tables 32 words of 4 packed bytes), plus:
 *   compressed_words  size of `compressed` in u32 words (every read stays inside),
 *   end_offset        end of this matrix's data in 16-byte units (the next matrix's first
 *                     offset, or compressed_words / 4), bounding its last warp,
 *   n_vec, vec_stride 1..8 vectors of in_dim int8, vec_stride bytes apart (4-aligned),
 *   out8              out8[v*out_dim + r] = clamp(rint(float(acc) * scale), -128, 127)
 *                     as the reference's epilogue (:212-220), or NULL,
 *   acc32             the int32 sums themselves, or NULL (at least one of out8/acc32).
 * Preconditions: rows layout in_dim % 128 == 0; stripes out_dim % 32 == 0, in_dim % 4 == 0. */
int matvans_b_rows_matvec(const uint32_t *compressed, uint64_t compressed_words,
                          const uint32_t *warp_offsets, uint32_t end_offset,
                          const uint32_t *ppf, const uint32_t *ppf_cum, const uint32_t *ppf_prob,
                          const int8_t *vecs, uint32_t n_vec, uint32_t vec_stride,
                          uint32_t in_dim, uint32_t out_dim, float scale, int8_t *out8,
                          int32_t *acc32, void *stream);
int matvans_b_stripes_matvec(const uint32_t *compressed, uint64_t compressed_words,
                             const uint32_t *warp_offsets, uint32_t end_offset,
                             const uint32_t *ppf, const uint32_t *ppf_cum,
                             const uint32_t *ppf_prob, const int8_t *vecs, uint32_t n_vec,
                             uint32_t vec_stride, uint32_t in_dim, uint32_t out_dim,
                             float scale, int8_t *out8, int32_t *acc32, void *stream);

/* 7-bit model (fit_128bit_encoder_model, encoder.py:268-306, float32 rounding and
 * first-maximum adjustment as numpy) and the decoder tables
 * (decoder_model_from_128bit_encoder_model, :309-320) packed 4 bytes per word
 * (pack_4bytes, :465-471).  n_symbols = max - min + 1 spans the value range (gaps
 * included), so cdf receives n_symbols + 1 <= 257 bytes; ppf, ppf_cum and ppf_prob 32
 * words each.  More than 128 distinct values return MATVANS_E_RANGE (the reference
 * divides by zero there). */
int matvans_b_model(const int8_t *data, uint64_t n, int32_t *min_value, uint32_t *n_symbols,
                    uint8_t *cdf, uint32_t *ppf, uint32_t *ppf_cum, uint32_t *ppf_prob);
/* Upper bound on the words matvans_b_encode writes; layout 0 rows, 1 stripes. */
uint64_t matvans_b_encode_bound(uint32_t rows, uint32_t cols, int layout);
/* Encode one matrix, word for word as encoder.py (the 32 coders carry their states from
 * warp to warp).  warp_offsets[rows or rows/32] = offset_base + start in 16-byte units. */
int matvans_b_encode(const int8_t *mat, uint32_t rows, uint32_t cols, int layout,
                     int32_t min_value, const uint8_t *cdf, uint32_t n_symbols,
                     uint32_t offset_base, uint32_t *compressed, uint64_t cap_words,
                     uint64_t *n_words, uint32_t *warp_offsets);

/* Minimal safetensors reader (the container of buffered/encoder.py:498-510, read by
 * buffered/sane_safetensors.hpp:17-68): a u64 header length, a JSON header of
 * {name: {dtype, shape, data_offsets}}, then the data.  matvans_st_tensor returns a view
 * (dtype string such as "U32", up to 4 dimensions, data pointer and byte size). */
typedef struct matvans_st matvans_st;
int matvans_st_read(const char *path, matvans_st **out);
void matvans_st_free(matvans_st *st);
int matvans_st_tensor(const matvans_st *st, const char *name, const char **dtype,
                      uint32_t *ndim, uint64_t *shape, const void **data, uint64_t *nbytes);

/* ------------------------------------------------------------------------------
 * N1. Persistent batched decode-matvec: all matrices of a batch (uneven shapes, mixed
 * encodings, 1..8 vectors each) in ONE kernel launch whose blocks pull row tiles from a
 * device work queue.  No reference counterpart: the reference launches one kernel per
 * matrix (matvansWarpPerRow.cu:156-164; buffered/ans-warp-rows.cu:143-164) with a
 * hand-set warpsPerBlock (:156).  The plan is created once (it allocates its device
 * descriptors, tile table and queue counters); matvans_batch_launch allocates nothing,
 * synchronises nothing and is one kernel launch, so it can be captured in a HIP graph.
 * The queue resets itself at the end of every launch.  Launches of one plan must not
 * overlap (they share the queue counters).
 * ----------------------------------------------------------------------------*/
enum {
    MATVANS_BATCH_V0 = 0,        /* .bin v0 record: cursors, cdf, ppf, min_value, G, payload */
    MATVANS_BATCH_QUAD = 1,      /* MVX1 quad coder, 32, 64 or 128 streams: cursors, lut,
                                    payload (batched backward: 32 or 64) */
    MATVANS_BATCH_B_ROWS = 2,    /* family-B ans-warp-rows */
    MATVANS_BATCH_B_STRIPES = 3, /* family-B ans-warp-stripes */
    MATVANS_BATCH_QUAD_LANES = 4 /* MVX1 coder 2 (quad streams back to back), 32 or 64
                                    streams: the fields of MATVANS_BATCH_QUAD */
};

typedef struct {
    uint32_t format;             /* MATVANS_BATCH_* */
    uint32_t rows, cols;
    uint32_t n_vec, vec_stride;  /* 1..8 vectors, vec_stride bytes apart */
    const int8_t *vecs;          /* device */
    int32_t *dst;                /* v0, quad: int32 [n_vec][rows], or NULL with out8 set;
                                    family B: use acc32 */
    /* v0 and quad */
    const uint32_t *cursors;
    const uint16_t *payload;
    uint32_t payload_size;
    int32_t min_value;           /* v0 */
    uint32_t G;                  /* v0 */
    const uint8_t *cdf, *ppf;    /* v0 */
    uint32_t streams;            /* quad: 32, 64 or 128 */
    const uint32_t *lut;         /* quad: the 2 * 256 dwords of matvans_v1_lut */
    /* family B (as matvans_b_rows_matvec) */
    const uint32_t *compressed;
    uint64_t compressed_words;
    const uint32_t *warp_offsets;
    uint32_t end_offset;
    const uint32_t *b_ppf, *b_cum, *b_prob;
    /* every format: the next layer's int8 input, out8[v*rows + r] =
       clamp(rint(float(sum) * scale), -128, 127) (ans-warp-rows-kernel.cu:212-220), or NULL */
    float scale;
    int32_t *acc32;              /* family B: int32 sums, or NULL */
    int8_t *out8;
    /* quad: outlier split of the matrix (matvans_v1_dev_matrix out_start / out_rec), or
       NULL */
    const uint32_t *out_start, *out_rec;
} matvans_batch_item;

typedef struct {
    uint32_t units_per_tile;     /* row groups per work-queue tile (1..64) */
    uint32_t order;              /* 0 matrix-major, 1 round-robin, 2 longest chain first */
    uint32_t blocks_per_cu;      /* persistent blocks per compute unit (1..32) */
    uint32_t schedule;           /* 0 device work queue (atomic), 1 static round-robin */
} matvans_batch_config;

typedef struct matvans_batch matvans_batch;

int matvans_batch_create(const matvans_batch_item *items, uint32_t n, matvans_batch **out);

/* N1 x N3: the backward dx = alpha * W^T (s_w * g) of every matrix of a batch in two
 * launches (a persistent launch over (64-row block, 8-stream group) units from the same
 * sharded work queue, then one fixed-order reduce over every item's columns).  Results
 * are bitwise identical to matvans_decompress_backward / matvans_v1_decompress_backward.
 * The plan owns the partial-sum workspace.  .bin v0 and quad (coder 1 or 2, 32 / 64
 * streams) items.
 * No reference counterpart. */
typedef struct {
    uint32_t format;             /* MATVANS_BATCH_V0, MATVANS_BATCH_QUAD or _QUAD_LANES */
    uint32_t rows, cols;
    uint32_t n_vec, g_stride;    /* 1..8 gradient vectors, g_stride floats apart */
    const float *g;              /* [n_vec][g_stride] */
    const float *s_w;            /* [rows] */
    const float *alpha;          /* [n_vec] */
    float *dx;                   /* [n_vec][cols] */
    const uint32_t *cursors;
    const uint16_t *payload;     /* 16-byte aligned */
    uint32_t payload_size;
    int32_t min_value;           /* v0 */
    uint32_t G;                  /* v0 */
    const uint8_t *cdf, *ppf;    /* v0 */
    uint32_t streams;            /* quad: 32 or 64 */
    const uint32_t *lut;         /* quad: the 2 * 256 dwords of matvans_v1_lut */
    /* quad: outliers by column (matvans_v1_dev_matrix out_col_start / out_col_rec), or
       NULL */
    const uint32_t *out_col_start, *out_col_rec;
} matvans_batch_bwd_item;

int matvans_batch_create_backward(const matvans_batch_bwd_item *items, uint32_t n,
                                  matvans_batch **out);
void matvans_batch_destroy(matvans_batch *plan);
int matvans_batch_configure(matvans_batch *plan, const matvans_batch_config *cfg);
int matvans_batch_get_config(const matvans_batch *plan, matvans_batch_config *cfg);
int matvans_batch_launch(const matvans_batch *plan, void *stream);
/* Autotuner: times units_per_tile x order x blocks_per_cu x schedule over `reps` launches each,
 * in two passes (the second in reverse order), keeps the configuration with the fastest of
 * its two timings and returns it (and that ms per launch).  Synchronises. */
int matvans_batch_autotune(matvans_batch *plan, void *stream, uint32_t reps,
                           matvans_batch_config *best, float *best_ms);

#ifdef __cplusplus
}
#endif
#endif /* MATVANS_H */
