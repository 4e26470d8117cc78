/*
 * CPU oracle for the WarpPerRow rANS codec -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference decoder, encoder and integer matvec.  Only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py load this
 * library, and only to check results; the product (matvans_amd) never links it.
 *
 *   oracle_fxhash      create_matrices_WarpPerRow.py:58-70, matvansWarpPerRow.cu:31-42
 *   oracle_decode_v0   matvansWarpPerRow-kernel.cu:95-174 (word order of the per-lane
 *                      FIFO, column mapping, decode step), AnsCoder.pop
 *                      create_matrices_WarpPerRow.py:250-264
 *   oracle_matvec_v0   the same decode fused with the int32 dot (kernel :133,:163-174),
 *                      threaded over row ranges with pthreads (the CPU baseline)
 *   oracle_encode_v0   create_matrices_WarpPerRow.py:353-430 (AnsCoder.push :230-248,
 *                      interrupt :267-272, interleave :390-422), rows x cols general
 *
 * Pinned in tests/test_oracle.py: decode of the golden .bin files equals the golden raw
 * matrices; encode with the golden model reproduces the golden payload outside padding
 * and the golden cursors; matvec equals the brute-force integer product.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define LANES 32
#define GRANULE 8

uint32_t oracle_fxhash(const int8_t *a, uint64_t n) {
    uint32_t h = 0;
    for (uint64_t i = 0; i < n; i++) {
        h = (h >> 27) | (h << 5);
        h = (h ^ (uint8_t)a[i]) * 0x27220A95u;
    }
    return h;
}

/* Number of u16 words that belong to row r (its granules). */
static uint64_t row_words(uint32_t rows, const uint32_t *cursors, uint64_t payload_size,
                          uint32_t r) {
    uint64_t end = (r + 1 < rows) ? cursors[r + 1] : payload_size;
    return end > cursors[r] ? end - cursors[r] : 0;
}

typedef struct {
    uint32_t rows, cols;
    const uint32_t *cursors;
    const uint16_t *payload;
    uint64_t payload_size;
    int8_t min_value;
    uint8_t lut_sym[256];
    uint8_t lut_f[256];
    uint8_t lut_b[256];
    const int8_t *vec;   /* matvec mode */
    int32_t *dst;
    int8_t *out;         /* decode mode */
    uint32_t r0, r1;
    int status;
} job_t;

static void build_tables(job_t *j, int8_t min_value, const uint8_t *cdf, const uint8_t *ppf,
                         uint32_t G) {
    for (int q = 0; q < 256; q++) {
        uint32_t s = ppf[q];
        uint32_t cs = s <= G ? cdf[s] : 0, cs1 = s + 1 <= G ? cdf[s + 1] : 0;
        j->lut_sym[q] = (uint8_t)(min_value + (int)s);          /* int8 wrap as the kernel */
        j->lut_f[q] = (uint8_t)((cs1 - cs) & 0xFF);
        j->lut_b[q] = (uint8_t)(q - cs);
    }
}

/* Decode row r; either store symbols (out) or accumulate the dot (vec). */
static int decode_row(job_t *j, uint32_t r, int32_t *acc_out) {
    const uint64_t cur = j->cursors[r];
    const uint64_t nw = row_words(j->rows, j->cursors, j->payload_size, r);
    const uint32_t steps = j->cols / LANES;
    int64_t acc = 0;
    for (uint32_t lane = 0; lane < LANES; lane++) {
        uint64_t ptr = 0;
#define WORD_AT(jj) (((uint64_t)256 * ((jj) / GRANULE) + 8 * lane + (jj) % GRANULE))
        if (WORD_AT(1) >= nw) return -1;
        uint32_t head = ((uint32_t)j->payload[cur + WORD_AT(0)] << 16) |
                        j->payload[cur + WORD_AT(1)];
        ptr = 2;
        for (uint32_t k = 0; k < steps; k++) {
            uint32_t q = head & 0xFF;
            int8_t w = (int8_t)j->lut_sym[q];
            uint32_t col = 512 * (k / 16) + 16 * lane + (k % 16);
            if (j->out) j->out[(uint64_t)r * j->cols + col] = w;
            else acc += (int32_t)w * (int32_t)j->vec[col];
            head = (head >> 8) * j->lut_f[q] + j->lut_b[q];
            if (head < (1u << 16)) {
                uint64_t p = WORD_AT(ptr);
                if (p >= nw) return -1;
                head = (head << 16) | j->payload[cur + p];
                ptr++;
            }
        }
#undef WORD_AT
    }
    if (acc_out) *acc_out = (int32_t)acc;
    return 0;
}

static void *run_rows(void *arg) {
    job_t *j = (job_t *)arg;
    for (uint32_t r = j->r0; r < j->r1; r++) {
        int32_t acc = 0;
        if (decode_row(j, r, j->out ? NULL : &acc) != 0) { j->status = -1; return NULL; }
        if (!j->out) j->dst[r] = acc;
    }
    return NULL;
}

static int run(uint32_t rows, uint32_t cols, const uint32_t *cursors, const uint16_t *payload,
               uint64_t payload_size, int8_t min_value, uint8_t G, const uint8_t *cdf,
               const uint8_t *ppf, const int8_t *vec, int32_t *dst, int8_t *out, int nthreads) {
    if (cols % 512 != 0) return -2;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((uint32_t)nthreads > rows) nthreads = rows ? (int)rows : 1;
    job_t base;
    memset(&base, 0, sizeof base);
    base.rows = rows; base.cols = cols; base.cursors = cursors; base.payload = payload;
    base.payload_size = payload_size; base.min_value = min_value;
    base.vec = vec; base.dst = dst; base.out = out;
    build_tables(&base, min_value, cdf, ppf, G);
    job_t jobs[256];
    pthread_t th[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = base;
        jobs[t].r0 = (uint32_t)((uint64_t)rows * t / nthreads);
        jobs[t].r1 = (uint32_t)((uint64_t)rows * (t + 1) / nthreads);
    }
    for (int t = 1; t < nthreads; t++) pthread_create(&th[t], NULL, run_rows, &jobs[t]);
    run_rows(&jobs[0]);
    int st = jobs[0].status;
    for (int t = 1; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        if (jobs[t].status) st = jobs[t].status;
    }
    return st;
}

int oracle_decode_v0(uint32_t rows, uint32_t cols, const uint32_t *cursors,
                     const uint16_t *payload, uint64_t payload_size, int8_t min_value,
                     uint8_t G, const uint8_t *cdf, const uint8_t *ppf, int8_t *out) {
    return run(rows, cols, cursors, payload, payload_size, min_value, G, cdf, ppf, NULL, NULL,
               out, 1);
}

int oracle_matvec_v0(uint32_t rows, uint32_t cols, const uint32_t *cursors,
                     const uint16_t *payload, uint64_t payload_size, int8_t min_value,
                     uint8_t G, const uint8_t *cdf, const uint8_t *ppf, const int8_t *vec,
                     int32_t *dst, int nthreads) {
    return run(rows, cols, cursors, payload, payload_size, min_value, G, cdf, ppf, vec, dst,
               NULL, nthreads);
}

/*
 * Encoder.  cdf16 is the full uint16 cdf (cdf16[G] == 256).  Writes cursors[rows] and the
 * interleaved payload (padding slots = 0); returns the number of words, or a negative
 * status (-2 shape, -3 capacity, -4 symbol outside the model).
 */
int64_t oracle_encode_v0(uint32_t rows, uint32_t cols, const int8_t *mat, int8_t min_value,
                         uint32_t G, const uint16_t *cdf16, uint32_t *cursors,
                         uint16_t *payload, uint64_t payload_cap) {
    if (cols % 512 != 0) return -2;
    const uint32_t steps = cols / LANES;
    /* worst case per lane: one word per symbol + 2 state words */
    uint16_t *buf = (uint16_t *)malloc(sizeof(uint16_t) * LANES * (steps + 2));
    uint32_t len[LANES];
    uint64_t pos = 0;
    for (uint32_t r = 0; r < rows; r++) {
        uint32_t mx = 0;
        for (uint32_t lane = 0; lane < LANES; lane++) {
            uint16_t *b = buf + (uint64_t)lane * (steps + 2);
            uint32_t n = 0;
            uint64_t head = 1u << 16;
            for (int64_t k = (int64_t)steps - 1; k >= 0; k--) {
                uint32_t col = 512 * (uint32_t)(k / 16) + 16 * lane + (uint32_t)(k % 16);
                int s = (int)mat[(uint64_t)r * cols + col] - (int)min_value;
                if (s < 0 || (uint32_t)s >= G) { free(buf); return -4; }
                uint32_t c0 = cdf16[s], f = cdf16[s + 1] - c0;
                if (f == 0) { free(buf); return -4; }
                if ((head >> 24) >= f) {
                    while ((head >> 16) != 0) { b[n++] = (uint16_t)(head & 0xFFFF); head >>= 16; }
                }
                head = (((head / f) << 8) | (head % f + c0)) & 0xFFFFFFFFu;
            }
            b[n++] = (uint16_t)(head & 0xFFFF);
            b[n++] = (uint16_t)(head >> 16);
            len[lane] = n;
            if (n > mx) mx = n;
        }
        uint32_t kgr = (mx + GRANULE - 1) / GRANULE;
        uint64_t words = (uint64_t)kgr * GRANULE * LANES;
        if (pos + words > payload_cap) { free(buf); return -3; }
        cursors[r] = (uint32_t)pos;
        memset(payload + pos, 0, words * sizeof(uint16_t));
        for (uint32_t lane = 0; lane < LANES; lane++) {
            const uint16_t *b = buf + (uint64_t)lane * (steps + 2);
            uint32_t n = len[lane];
            for (uint32_t jj = 0; jj < n; jj++)   /* stream = reversed emission order */
                payload[pos + 256 * (jj / 8) + 8 * lane + jj % 8] = b[n - 1 - jj];
        }
        pos += words;
    }
    free(buf);
    return (int64_t)pos;
}
