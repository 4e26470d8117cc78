// Device copies of host matrices, owned by the library: the upload step of the reference's
// benchmark loop (matvansWarpPerRow.cu:287-313, which cudaMallocs and copies cursors, cdf,
// ppf and payload per matrix and stores the pointers in CompressedMatrix::d_*), as a
// handle that owns the device buffers and a device view (matvans_dev_matrix /
// matvans_v1_dev_matrix) to pass to the decode entry points.
//
// The payload gets 1 KiB of zero slack after its last word (no kernel here reads past
// payload_size, but callers slicing a payload by rows keep 16-byte loads inside the
// allocation).  The copies are synchronous (hipMemcpy); uploading is not on the launch path.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "matvans.h"
#include "matvans_internal.h"

using mv::fail;

struct matvans_upload {
    std::vector<void *> bufs;
};

namespace {

constexpr size_t kSlackBytes = 1024;

int hip_status(hipError_t e, const char *what) {
    return e == hipSuccess ? MATVANS_OK
                           : fail(MATVANS_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

// Allocates bytes (+ slack zeroed) on the current device and copies src into it.
int put(matvans_upload *u, const void *src, size_t bytes, size_t slack, void **dst) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes + slack > 0 ? bytes + slack : 16);
    if (e != hipSuccess) return hip_status(e, "hipMalloc");
    u->bufs.push_back(p);
    if (slack) e = hipMemset(static_cast<char *>(p) + bytes, 0, slack);
    if (e == hipSuccess && bytes) e = hipMemcpy(p, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_status(e, "upload copy");
    *dst = p;
    return MATVANS_OK;
}

}  // namespace

extern "C" void matvans_upload_free(matvans_upload *u) {
    if (!u) return;
    for (void *p : u->bufs) (void)hipFree(p);
    delete u;
}

extern "C" int matvans_matrix_upload(const matvans_host_matrix *m, matvans_upload **out,
                                     matvans_dev_matrix *view) {
    if (!m || !out || !view || !m->cursors || !m->cdf || !m->ppf || (m->payload_size && !m->payload))
        return fail(MATVANS_E_INVALID, "null argument");
    *out = nullptr;
    auto *u = new matvans_upload();
    uint8_t cdf[256] = {0};
    std::memcpy(cdf, m->cdf, size_t(m->G) + 1);
    void *cur = nullptr, *dcdf = nullptr, *dppf = nullptr, *pay = nullptr;
    int st = put(u, m->cursors, size_t(m->rows) * 4, 0, &cur);
    if (!st) st = put(u, cdf, sizeof(cdf), 0, &dcdf);
    if (!st) st = put(u, m->ppf, 256, 0, &dppf);
    if (!st) st = put(u, m->payload, size_t(m->payload_size) * 2, kSlackBytes, &pay);
    if (st) {
        matvans_upload_free(u);
        return st;
    }
    *view = {m->rows, m->cols, m->grid_spacing, m->min_value, m->G, 0, m->payload_size,
             static_cast<const uint32_t *>(cur), static_cast<const uint8_t *>(dcdf),
             static_cast<const uint8_t *>(dppf), static_cast<const uint16_t *>(pay)};
    *out = u;
    mv::clear_error();
    return MATVANS_OK;
}

extern "C" int matvans_v1_matrix_upload(const matvans_v1_host *m, matvans_upload **out,
                                        matvans_v1_dev_matrix *view) {
    if (!m || !out || !view || !m->cursors || (m->payload_size && !m->payload))
        return fail(MATVANS_E_INVALID, "null argument");
    if (m->table_mode == 0 ? !m->cdf : !m->freq) return fail(MATVANS_E_INVALID, "null table");
    if (m->precision != 8 && m->precision != 12)
        return fail(MATVANS_E_INVALID, "precision %u", m->precision);
    *out = nullptr;
    std::vector<uint32_t> lut;
    if (m->table_mode == 0) {
        lut.resize(size_t(2) << m->precision);
        if (int st = matvans_v1_lut(m->cdf, m->G, m->min_value, m->precision, lut.data()))
            return st;
    }
    auto *u = new matvans_upload();
    void *cur = nullptr, *tab = nullptr, *pay = nullptr;
    int st = put(u, m->cursors, size_t(m->rows) * 4, 0, &cur);
    if (!st)
        st = m->table_mode == 0 ? put(u, lut.data(), lut.size() * 4, 0, &tab)
                                : put(u, m->freq, size_t(m->rows) * m->G, 0, &tab);
    if (!st) st = put(u, m->payload, size_t(m->payload_size) * 2, kSlackBytes, &pay);
    // outliers: the records by row as stored, and by column for the backward
    void *ost = nullptr, *orec = nullptr, *ocst = nullptr, *ocrec = nullptr;
    if (!st && m->n_outliers) {
        if (!m->out_start || !m->out_rec) {
            matvans_upload_free(u);
            return fail(MATVANS_E_INVALID, "outliers without offsets or records");
        }
        std::vector<uint32_t> col_start(size_t(m->cols) + 1, 0), col_rec(m->n_outliers);
        for (uint32_t i = 0; i < m->n_outliers; i++) {
            const uint32_t c = m->out_rec[i] & 0xFFFFFFu;
            if (c >= m->cols) {
                matvans_upload_free(u);
                return fail(MATVANS_E_INVALID, "outlier %u: column %u >= %u", i, c, m->cols);
            }
            col_start[c + 1]++;
        }
        for (uint32_t c = 0; c < m->cols; c++) col_start[c + 1] += col_start[c];
        std::vector<uint32_t> fill(col_start.begin(), col_start.end() - 1);
        for (uint32_t r = 0; r < m->rows; r++)   // rows ascending within each column
            for (uint32_t i = m->out_start[r]; i < m->out_start[r + 1]; i++) {
                const uint32_t c = m->out_rec[i] & 0xFFFFFFu;
                col_rec[fill[c]++] = r | (m->out_rec[i] & 0xFF000000u);
            }
        st = put(u, m->out_start, (size_t(m->rows) + 1) * 4, 0, &ost);
        if (!st) st = put(u, m->out_rec, size_t(m->n_outliers) * 4, 0, &orec);
        if (!st) st = put(u, col_start.data(), col_start.size() * 4, 0, &ocst);
        if (!st) st = put(u, col_rec.data(), col_rec.size() * 4, 0, &ocrec);
    }
    if (st) {
        matvans_upload_free(u);
        return st;
    }
    *view = {m->rows, m->cols, m->streams, m->precision, m->table_mode, m->coder, m->min_value,
             m->G, m->payload_size, static_cast<const uint32_t *>(cur),
             m->table_mode == 0 ? static_cast<const uint32_t *>(tab) : nullptr,
             m->table_mode == 1 ? static_cast<const uint8_t *>(tab) : nullptr,
             static_cast<const uint16_t *>(pay), m->n_outliers,
             static_cast<const uint32_t *>(ost), static_cast<const uint32_t *>(orec),
             static_cast<const uint32_t *>(ocst), static_cast<const uint32_t *>(ocrec)};
    *out = u;
    mv::clear_error();
    return MATVANS_OK;
}
