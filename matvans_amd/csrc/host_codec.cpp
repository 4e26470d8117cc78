// Host side of the codec: error reporting, FxHash, the .bin v0 container, the 8-bit
// entropy model and the 32-stream row-interleaved rANS encoder.
//
// The format and the coder follow the reference's Python generator
// (matvansWarpPerRow/create_matrices_WarpPerRow.py) and C++ reader
// (matvansWarpPerRow/matvansWarpPerRow.cu); both are cited per function.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "matvans_internal.h"

namespace {
thread_local std::string g_last_error;
}

namespace mv {
int fail(int status, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}
void clear_error() { g_last_error.clear(); }
}  // namespace mv

using mv::fail;

extern "C" const char *matvans_last_error(void) { return g_last_error.c_str(); }
extern "C" int matvans_version(void) { return 100; }

// FxHash (matvansWarpPerRow.cu:31-42): rotate left 5, xor byte, multiply 0x27220A95.
extern "C" uint32_t matvans_fxhash(const int8_t *data, uint64_t n) {
    uint32_t h = 0;
    for (uint64_t i = 0; i < n; i++) {
        h = (h >> 27) | (h << 5);
        h = (h ^ static_cast<uint8_t>(data[i])) * 0x27220A95u;
    }
    return h;
}

// ------------------------------------------------------------------ .bin container
struct matvans_bin {
    uint32_t num_matrices = 0, result_hash = 0, max_word_count = 0;
    std::vector<int8_t> vector;
    std::vector<std::vector<uint8_t>> blobs;   // one owned byte blob per record
    std::vector<matvans_host_matrix> mats;
};

namespace {
struct Reader {
    const uint8_t *p, *end;
    bool take(void *dst, size_t n) {
        if (static_cast<size_t>(end - p) < n) return false;
        if (dst) std::memcpy(dst, p, n);
        p += n;
        return true;
    }
};

// Cursor checks: nondecreasing multiples of 256 words (one 32-lane x 16-byte granule)
// starting at 0 and ending inside the payload (create_matrices_WarpPerRow.py:405-422).
int check_cursors(const uint32_t *cur, uint32_t rows, uint32_t payload_size) {
    for (uint32_t r = 0; r < rows; r++) {
        uint32_t end = (r + 1 < rows) ? cur[r + 1] : payload_size;
        if (cur[r] % 256 != 0 || end < cur[r] || end > payload_size || (end - cur[r]) % 256)
            return fail(MATVANS_E_FORMAT, "row %u: bad cursor %u (next %u, payload %u)", r,
                        cur[r], end, payload_size);
    }
    if (rows && cur[0] != 0) return fail(MATVANS_E_FORMAT, "cursors[0] = %u != 0", cur[0]);
    return MATVANS_OK;
}
}  // namespace

extern "C" int matvans_bin_read(const char *path, matvans_bin **out) {
    if (!path || !out) return fail(MATVANS_E_INVALID, "null argument");
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(MATVANS_E_IO, "cannot open %s", path);
    std::vector<uint8_t> file;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0) { std::fclose(f); return fail(MATVANS_E_IO, "cannot size %s", path); }
    file.resize(static_cast<size_t>(sz));
    size_t got = sz ? std::fread(file.data(), 1, file.size(), f) : 0;
    std::fclose(f);
    if (got != file.size()) return fail(MATVANS_E_IO, "short read on %s", path);

    auto *bin = new matvans_bin();
    Reader rd{file.data(), file.data() + file.size()};
    uint32_t len_v = 0;
    // header '<LLL' then '<L' len_v (create_matrices_WarpPerRow.py:442,446)
    if (!rd.take(&bin->num_matrices, 4) || !rd.take(&bin->result_hash, 4) ||
        !rd.take(&bin->max_word_count, 4) || !rd.take(&len_v, 4)) {
        delete bin;
        return fail(MATVANS_E_FORMAT, "truncated header");
    }
    bin->vector.resize(len_v);
    // vector then pad to 4 bytes (:447-453).  The reference C++ reader does not skip the
    // pad (matvansWarpPerRow.cu:234-235), which only works for len_v % 4 == 0.
    if (!rd.take(bin->vector.data(), len_v) || !rd.take(nullptr, 3 - (len_v + 3) % 4)) {
        delete bin;
        return fail(MATVANS_E_FORMAT, "truncated vector");
    }
    for (uint32_t k = 0; k < bin->num_matrices; k++) {
        uint32_t rows, cols, psize;
        float gs;
        int8_t mn;
        uint8_t G;
        // '<LLf{rows}LLbB{G+1}B' (:292-302), cdf pad (:304-305), ppf (:308-310), payload
        // and its pad (:312-318); reader matvansWarpPerRow.cu:100-145.
        if (!rd.take(&rows, 4) || !rd.take(&cols, 4) || !rd.take(&gs, 4)) goto trunc;
        {
            const uint8_t *cur_p = rd.p;
            if (!rd.take(nullptr, size_t(rows) * 4)) goto trunc;
            if (!rd.take(&psize, 4) || !rd.take(&mn, 1) || !rd.take(&G, 1)) goto trunc;
            const uint8_t *cdf_p = rd.p;
            if (!rd.take(nullptr, size_t(G) + 1)) goto trunc;
            if ((G + 1) % 2 == 1 && !rd.take(nullptr, 1)) goto trunc;
            const uint8_t *ppf_p = rd.p;
            if (!rd.take(nullptr, 256)) goto trunc;
            const uint8_t *pay_p = rd.p;
            if (!rd.take(nullptr, size_t(psize) * 2)) goto trunc;
            if (psize % 2 == 1 && !rd.take(nullptr, 2)) goto trunc;
            // own a 16-byte aligned copy: [payload | cursors | cdf(257) | ppf(256)]
            size_t pay_bytes = (size_t(psize) * 2 + 15) / 16 * 16;
            size_t cur_bytes = (size_t(rows) * 4 + 15) / 16 * 16;
            std::vector<uint8_t> blob(pay_bytes + cur_bytes + 272 + 256 + 16);
            uint8_t *base = blob.data();
            uintptr_t mis = reinterpret_cast<uintptr_t>(base) % 16;
            uint8_t *a = base + (mis ? 16 - mis : 0);
            std::memcpy(a, pay_p, size_t(psize) * 2);
            std::memcpy(a + pay_bytes, cur_p, size_t(rows) * 4);
            std::memset(a + pay_bytes + cur_bytes, 0, 272);
            std::memcpy(a + pay_bytes + cur_bytes, cdf_p, size_t(G) + 1);
            std::memcpy(a + pay_bytes + cur_bytes + 272, ppf_p, 256);
            matvans_host_matrix m{};
            m.rows = rows;
            m.cols = cols;
            m.grid_spacing = gs;
            m.min_value = mn;
            m.G = G;
            m.payload_size = psize;
            m.payload = reinterpret_cast<const uint16_t *>(a);
            m.cursors = reinterpret_cast<const uint32_t *>(a + pay_bytes);
            m.cdf = a + pay_bytes + cur_bytes;
            m.ppf = a + pay_bytes + cur_bytes + 272;
            int st = check_cursors(m.cursors, rows, psize);
            if (st) { delete bin; return st; }
            bin->blobs.push_back(std::move(blob));
            // blob storage does not move when the outer vector grows (vector move keeps
            // the heap buffer), so the pointers above stay valid.
            bin->mats.push_back(m);
        }
        continue;
    trunc:
        delete bin;
        return fail(MATVANS_E_FORMAT, "matrix %u: truncated record", k);
    }
    *out = bin;
    mv::clear_error();
    return MATVANS_OK;
}

extern "C" void matvans_bin_free(matvans_bin *bin) { delete bin; }

// ------------------------------------------------------------------ raw container
// The uncompressed baseline file: header '<LL' (num_matrices, result_hash), '<L' len_v,
// the vector padded to 4 bytes, then per matrix '<LLf' (rows, cols, grid_spacing), the
// row-major int8 entries and one pad byte when rows * cols is odd
// (create_matrices_WarpPerRow.py:94-129; reader baseline_matVec_custom.cu:371-393,
// 480-508).  The reference reader does not skip the vector's pad either (:487-490), so it
// reads only files with len_v % 4 == 0; this reader skips it, as the writer emits it.
struct matvans_raw {
    uint32_t num_matrices = 0, result_hash = 0;
    std::vector<int8_t> vector;
    std::vector<std::vector<int8_t>> data;   // one owned row-major copy per matrix
    std::vector<uint32_t> rows, cols;
    std::vector<float> grid_spacing;
};

extern "C" int matvans_raw_read(const char *path, matvans_raw **out) {
    if (!path || !out) return fail(MATVANS_E_INVALID, "null argument");
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(MATVANS_E_IO, "cannot open %s", path);
    std::vector<uint8_t> file;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0) { std::fclose(f); return fail(MATVANS_E_IO, "cannot size %s", path); }
    file.resize(static_cast<size_t>(sz));
    size_t got = sz ? std::fread(file.data(), 1, file.size(), f) : 0;
    std::fclose(f);
    if (got != file.size()) return fail(MATVANS_E_IO, "short read on %s", path);

    auto *raw = new matvans_raw();
    Reader rd{file.data(), file.data() + file.size()};
    uint32_t len_v = 0;
    if (!rd.take(&raw->num_matrices, 4) || !rd.take(&raw->result_hash, 4) ||
        !rd.take(&len_v, 4)) {
        delete raw;
        return fail(MATVANS_E_FORMAT, "truncated header");
    }
    raw->vector.resize(len_v);
    if (!rd.take(raw->vector.data(), len_v) || !rd.take(nullptr, 3 - (len_v + 3) % 4)) {
        delete raw;
        return fail(MATVANS_E_FORMAT, "truncated vector");
    }
    for (uint32_t k = 0; k < raw->num_matrices; k++) {
        uint32_t rows = 0, cols = 0;
        float gs = 0.f;
        if (!rd.take(&rows, 4) || !rd.take(&cols, 4) || !rd.take(&gs, 4)) {
            delete raw;
            return fail(MATVANS_E_FORMAT, "matrix %u: truncated header", k);
        }
        const uint64_t n = uint64_t(rows) * cols;
        if (n > uint64_t(rd.end - rd.p)) {
            delete raw;
            return fail(MATVANS_E_FORMAT, "matrix %u: %ux%u entries past the end of the file",
                        k, rows, cols);
        }
        std::vector<int8_t> d(static_cast<size_t>(n));
        if (!rd.take(d.data(), static_cast<size_t>(n)) || (n % 2 == 1 && !rd.take(nullptr, 1))) {
            delete raw;
            return fail(MATVANS_E_FORMAT, "matrix %u: truncated entries", k);
        }
        raw->data.push_back(std::move(d));
        raw->rows.push_back(rows);
        raw->cols.push_back(cols);
        raw->grid_spacing.push_back(gs);
    }
    *out = raw;
    mv::clear_error();
    return MATVANS_OK;
}

extern "C" void matvans_raw_free(matvans_raw *raw) { delete raw; }

extern "C" int matvans_raw_header(const matvans_raw *raw, uint32_t *num_matrices,
                                  uint32_t *result_hash, const int8_t **vector,
                                  uint32_t *len_v) {
    if (!raw) return fail(MATVANS_E_INVALID, "null file");
    if (num_matrices) *num_matrices = raw->num_matrices;
    if (result_hash) *result_hash = raw->result_hash;
    if (vector) *vector = raw->vector.data();
    if (len_v) *len_v = static_cast<uint32_t>(raw->vector.size());
    return MATVANS_OK;
}

extern "C" int matvans_raw_matrix(const matvans_raw *raw, uint32_t k, uint32_t *rows,
                                  uint32_t *cols, float *grid_spacing, const int8_t **data) {
    if (!raw) return fail(MATVANS_E_INVALID, "null file");
    if (k >= raw->num_matrices)
        return fail(MATVANS_E_INVALID, "matrix %u of %u", k, raw->num_matrices);
    if (rows) *rows = raw->rows[k];
    if (cols) *cols = raw->cols[k];
    if (grid_spacing) *grid_spacing = raw->grid_spacing[k];
    if (data) *data = raw->data[k].data();
    return MATVANS_OK;
}

extern "C" int matvans_bin_header(const matvans_bin *bin, uint32_t *num_matrices,
                                  uint32_t *result_hash, uint32_t *max_word_count,
                                  const int8_t **vector, uint32_t *len_v) {
    if (!bin) return fail(MATVANS_E_INVALID, "null bin");
    if (num_matrices) *num_matrices = bin->num_matrices;
    if (result_hash) *result_hash = bin->result_hash;
    if (max_word_count) *max_word_count = bin->max_word_count;
    if (vector) *vector = bin->vector.data();
    if (len_v) *len_v = static_cast<uint32_t>(bin->vector.size());
    return MATVANS_OK;
}

extern "C" int matvans_bin_matrix(const matvans_bin *bin, uint32_t k, matvans_host_matrix *out) {
    if (!bin || !out) return fail(MATVANS_E_INVALID, "null argument");
    if (k >= bin->mats.size()) return fail(MATVANS_E_INVALID, "matrix %u of %zu", k,
                                           bin->mats.size());
    *out = bin->mats[k];
    return MATVANS_OK;
}

extern "C" int matvans_bin_write(const char *path, uint32_t result_hash, const int8_t *vector,
                                 uint32_t len_v, uint32_t num_matrices,
                                 const matvans_host_matrix *mats) {
    if (!path || (len_v && !vector) || (num_matrices && !mats))
        return fail(MATVANS_E_INVALID, "null argument");
    // max_word_count = max over matrices of (4+rows)*2 + (3+len(cdf))//2 + even payload
    // (create_matrices_WarpPerRow.py:286-288,455).
    uint64_t mwc = 0;
    for (uint32_t k = 0; k < num_matrices; k++) {
        const auto &m = mats[k];
        uint64_t w = (4ull + m.rows) * 2 + (3ull + m.G + 1) / 2 + m.payload_size +
                     m.payload_size % 2;
        mwc = std::max(mwc, w);
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(MATVANS_E_IO, "cannot create %s", path);
    auto put = [&](const void *p, size_t n) { return std::fwrite(p, 1, n, f) == n; };
    static const uint8_t zeros[4] = {0, 0, 0, 0};
    uint32_t mwc32 = static_cast<uint32_t>(mwc);
    bool ok = put(&num_matrices, 4) && put(&result_hash, 4) && put(&mwc32, 4) &&
              put(&len_v, 4) && put(vector, len_v) && put(zeros, 3 - (len_v + 3) % 4);
    for (uint32_t k = 0; ok && k < num_matrices; k++) {
        const auto &m = mats[k];
        uint8_t G = m.G;
        ok = put(&m.rows, 4) && put(&m.cols, 4) && put(&m.grid_spacing, 4) &&
             put(m.cursors, size_t(m.rows) * 4) && put(&m.payload_size, 4) &&
             put(&m.min_value, 1) && put(&G, 1) && put(m.cdf, size_t(G) + 1) &&
             ((G + 1) % 2 == 0 || put(zeros, 1)) && put(m.ppf, 256) &&
             put(m.payload, size_t(m.payload_size) * 2) &&
             (m.payload_size % 2 == 0 || put(zeros, 2));
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(MATVANS_E_IO, "write failed on %s", path);
    return MATVANS_OK;
}

// ------------------------------------------------------------------ entropy model
// create_matrices_WarpPerRow.py:131-187 with P = 8, then ppf (:189-208).
extern "C" int matvans_build_model(const int8_t *data, uint64_t n, int8_t *min_value,
                                   uint32_t *G, uint16_t *cdf16, uint8_t *ppf) {
    if (!data || !n || !min_value || !G || !cdf16 || !ppf)
        return fail(MATVANS_E_INVALID, "null argument or empty data");
    uint64_t hist[256] = {0};
    for (uint64_t i = 0; i < n; i++) hist[static_cast<uint8_t>(data[i]) ^ 0x80]++;
    int lo = 0, hi = 255;
    while (!hist[lo]) lo++;
    while (!hist[hi]) hi--;
    const int mn = lo - 128, mx = hi - 128;
    *min_value = static_cast<int8_t>(mn);
    if (mn == mx) {
        // dummy second symbol of probability 1/256 (:135-143)
        *G = 2;
        cdf16[0] = 0; cdf16[1] = 255; cdf16[2] = 256;
    } else {
        const uint32_t grid = static_cast<uint32_t>(mx - mn + 1);
        struct VC { int value; uint64_t count; };
        std::vector<VC> vc;
        for (int b = lo; b <= hi; b++)
            if (hist[b]) vc.push_back({b - 128, hist[b]});
        // descending counts; equal counts by descending value
        std::stable_sort(vc.begin(), vc.end(), [](const VC &a, const VC &b) {
            return a.count != b.count ? a.count > b.count : a.value > b.value;
        });
        const double scale = 256.0 / static_cast<double>(n);
        std::vector<int64_t> q(vc.size());
        int64_t sum = 0;
        for (size_t i = 0; i < vc.size(); i++) {
            double r = std::nearbyint(static_cast<double>(vc[i].count) * scale);  // half-even
            int64_t v = static_cast<int64_t>(static_cast<uint16_t>(static_cast<int64_t>(r)));
            q[i] = std::max<int64_t>(1, v);
            sum += q[i];
        }
        int64_t excess = sum - 256;
        if (excess > 0) {
            // shave from the most probable entries (:157-166)
            while (excess > 0) {
                int64_t nred = 0;
                for (auto v : q) nred += v > 1;
                int64_t nr = std::min(excess, nred);
                if (nr <= 0) return fail(MATVANS_E_RANGE, "model cannot be normalised");
                for (int64_t i = 0; i < nr; i++) q[i] -= 1;
                excess -= nr;
            }
        } else if (excess < 0) {
            // spread to the least probable entries (:167-170)
            if (-excess > static_cast<int64_t>(q.size()))
                return fail(MATVANS_E_RANGE, "model cannot be normalised");
            for (int64_t i = 0; i < -excess; i++) q[q.size() - 1 - i] += 1;
        }
        std::vector<uint32_t> pmf(grid + 1, 0);
        for (size_t i = 0; i < vc.size(); i++) pmf[vc[i].value - mn + 1] = static_cast<uint32_t>(q[i]);
        uint32_t acc = 0;
        for (uint32_t s = 0; s <= grid; s++) { acc += pmf[s]; cdf16[s] = static_cast<uint16_t>(acc); }
        if (cdf16[grid] != 256) return fail(MATVANS_E_RANGE, "model does not sum to 256");
        *G = grid;
    }
    // ppf[q] = largest s <= G-1 with (cdf[s] & 0xFF) <= q
    for (int qq = 0; qq < 256; qq++) {
        int best = 0;
        for (int s = static_cast<int>(*G) - 1; s >= 0; s--)
            if ((cdf16[s] & 0xFF) <= qq) { best = s; break; }
        ppf[qq] = static_cast<uint8_t>(best);
    }
    mv::clear_error();
    return MATVANS_OK;
}

// ------------------------------------------------------------------------ encoder
extern "C" uint64_t matvans_encode_bound(uint32_t rows, uint32_t cols) {
    // a lane emits at most one word per two symbols plus the state: ceil to granules
    uint64_t per_lane = (uint64_t(cols) / 32) / 2 + 3;
    return uint64_t(rows) * 256 * ((per_lane + 7) / 8);
}

namespace {
// Encode one lane of one row, in reverse column order; returns the stream in decode
// order [x_hi, x_lo, last flushed, ..., first flushed] (:371-379).
// push (:230-248): if (x >> 24) >= f flush x & 0xFFFF and x >>= 16; then
// x = ((x / f) << 8) | (x % f + cdf[s]).  interrupt (:267-272) emits lo then hi.
int encode_lane(const int8_t *row, uint32_t cols, uint32_t lane, int mn, uint32_t G,
                const uint16_t *cdf16, std::vector<uint16_t> &tmp, std::vector<uint16_t> &out) {
    const uint32_t steps = cols / 32;
    tmp.clear();
    uint32_t x = 1u << 16;
    for (int64_t k = int64_t(steps) - 1; k >= 0; k--) {
        uint32_t col = 512u * uint32_t(k / 16) + 16u * lane + uint32_t(k % 16);
        int s = int(row[col]) - mn;
        if (s < 0 || uint32_t(s) >= G) return -1;
        uint32_t c0 = cdf16[s], f = uint32_t(cdf16[s + 1]) - c0;
        if (f == 0 || f > 256) return -1;
        if ((x >> 24) >= f) { tmp.push_back(uint16_t(x & 0xFFFF)); x >>= 16; }
        x = ((x / f) << 8) | (x % f + c0);
    }
    tmp.push_back(uint16_t(x & 0xFFFF));
    tmp.push_back(uint16_t(x >> 16));
    out.assign(tmp.rbegin(), tmp.rend());
    return 0;
}
}  // namespace

extern "C" int matvans_encode_rows(const int8_t *mat, uint32_t rows, uint32_t cols,
                                   int8_t min_value, uint32_t G, const uint16_t *cdf16,
                                   uint32_t n_threads, uint32_t *cursors, uint16_t *payload,
                                   uint64_t payload_cap, uint64_t *payload_size) {
    if (!mat || !cdf16 || !cursors || !payload || !payload_size)
        return fail(MATVANS_E_INVALID, "null argument");
    if (cols == 0 || cols % 512) return fail(MATVANS_E_SHAPE, "cols %u not a multiple of 512", cols);
    if (G < 1 || G > 256 || cdf16[0] != 0 || cdf16[G] != 256)
        return fail(MATVANS_E_INVALID, "cdf must start at 0 and end at 256 (G=%u)", G);
    if (n_threads == 0) n_threads = std::max(1u, std::thread::hardware_concurrency());
    n_threads = std::min(n_threads, std::max(1u, rows));
    // pass 1: per-row streams (granule count per row), pass 2: interleave
    std::vector<uint32_t> kgr(rows);
    std::vector<std::vector<uint16_t>> streams(size_t(rows) * 32);
    std::atomic<int> bad{-1};
    auto pass1 = [&](uint32_t t) {
        std::vector<uint16_t> tmp;
        for (uint32_t r = t; r < rows; r += n_threads) {
            uint32_t mx = 0;
            for (uint32_t l = 0; l < 32; l++) {
                auto &s = streams[size_t(r) * 32 + l];
                if (encode_lane(mat + size_t(r) * cols, cols, l, min_value, G, cdf16, tmp, s)) {
                    bad = int(r);
                    return;
                }
                mx = std::max<uint32_t>(mx, uint32_t(s.size()));
            }
            kgr[r] = (mx + 7) / 8;
        }
    };
    std::vector<std::thread> th;
    for (uint32_t t = 1; t < n_threads; t++) th.emplace_back(pass1, t);
    pass1(0);
    for (auto &x : th) x.join();
    th.clear();
    if (bad >= 0) return fail(MATVANS_E_RANGE, "row %d holds a symbol outside the model", int(bad));
    uint64_t pos = 0;
    for (uint32_t r = 0; r < rows; r++) {
        if (pos > 0xFFFFFFFFull) return fail(MATVANS_E_CAPACITY, "payload exceeds 2^32 words");
        cursors[r] = uint32_t(pos);
        pos += uint64_t(kgr[r]) * 256;
    }
    if (pos > payload_cap) return fail(MATVANS_E_CAPACITY, "payload needs %llu words, cap %llu",
                                       (unsigned long long)pos, (unsigned long long)payload_cap);
    auto pass2 = [&](uint32_t t) {
        for (uint32_t r = t; r < rows; r += n_threads) {
            uint16_t *dst = payload + cursors[r];
            std::memset(dst, 0, size_t(kgr[r]) * 512);
            for (uint32_t l = 0; l < 32; l++) {
                const auto &s = streams[size_t(r) * 32 + l];
                for (size_t j = 0; j < s.size(); j++) dst[256 * (j / 8) + 8 * l + j % 8] = s[j];
            }
        }
    };
    for (uint32_t t = 1; t < n_threads; t++) th.emplace_back(pass2, t);
    pass2(0);
    for (auto &x : th) x.join();
    *payload_size = pos;
    mv::clear_error();
    return MATVANS_OK;
}
