// Host side of the MVX1 format extensions: S = 32, 64 or 128 interleaved streams per row,
// P = 8 or 12 bit probabilities, one table per matrix or one 8-bit table per row.
//
// These extend the reference's WarpPerRow codec (create_matrices_WarpPerRow.py:131-272,
// 353-430), which is the case S = 32, P = 8, one table: with those parameters the
// encoder below writes the reference's payload word for word.
//
// Stream l of a row decodes columns 16*S*(k/16) + 16*l + k%16.  A row is K granules of
// 512 bytes; every stream owns GW = 256/S consecutive u16 words of each granule, so word
// j of stream l is at cursor + 256*(j/GW) + GW*l + j%GW.  Coder 2 keeps the quad coder's
// streams without per-stream padding: per row the first 4 K words of every stream in K
// 16-byte-per-stream granules (K = the shortest stream's whole 4-word chunks), the rest
// of each stream back to back, with K and the tails' end offsets in row tables at the
// head of the payload.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "matvans_internal.h"

using mv::fail;

namespace {

// Integer frequencies summing to `total` from symbol counts: the reference's rounding,
// shaving and spreading (create_matrices_WarpPerRow.py:145-170) over entries ordered by
// count descending, equal counts by symbol index descending.  Absent symbols get 0.
int quantize_counts(const uint64_t *counts, uint32_t n_sym, uint32_t total, uint32_t *f) {
    std::vector<uint32_t> idx;
    uint64_t n = 0;
    for (uint32_t s = 0; s < n_sym; s++) {
        f[s] = 0;
        if (counts[s]) { idx.push_back(s); n += counts[s]; }
    }
    if (!n) return fail(MATVANS_E_INVALID, "no symbols");
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
        return counts[a] != counts[b] ? counts[a] > counts[b] : a > b;
    });
    const double scale = double(total) / double(n);
    std::vector<int64_t> q(idx.size());
    int64_t sum = 0;
    for (size_t i = 0; i < idx.size(); i++) {
        q[i] = std::max<int64_t>(1, int64_t(std::nearbyint(double(counts[idx[i]]) * scale)));
        sum += q[i];
    }
    int64_t excess = sum - int64_t(total);
    while (excess > 0) {
        int64_t nred = 0;
        for (auto v : q) nred += v > 1;
        int64_t nr = std::min(excess, nred);
        if (nr <= 0) return fail(MATVANS_E_RANGE, "model cannot be normalised");
        for (int64_t i = 0; i < nr; i++) q[i] -= 1;
        excess -= nr;
    }
    if (excess < 0) {
        if (-excess > int64_t(q.size())) return fail(MATVANS_E_RANGE, "model cannot be normalised");
        for (int64_t i = 0; i < -excess; i++) q[q.size() - 1 - i] += 1;
    }
    for (size_t i = 0; i < idx.size(); i++) f[idx[i]] = uint32_t(q[i]);
    return MATVANS_OK;
}

bool valid_params(uint32_t streams, uint32_t precision, uint32_t table_mode, uint32_t coder = 0) {
    return (streams == 32 || streams == 64 || streams == 128) && (precision == 8 || precision == 12) &&
           (table_mode == 0 || (table_mode == 1 && precision == 8)) &&
           (coder == 0 || (coder == 1 && precision == 8) ||
            (coder == 2 && precision == 8 && table_mode == 0));
}

// Coder 2: the row tables at the head of the payload (S + 2 u16 per row: K as two u16,
// then the S tail end offsets), zero-padded to a multiple of 8 u16.
uint64_t lanes_table_u16(uint32_t rows, uint32_t S) {
    return (uint64_t(rows) * (S + 2) + 7) / 8 * 8;
}
// A coder-2 row: K granules of S x 4 words, the tails, zero words to a multiple of four;
// its size in u16 words.
uint64_t lanes_row_u16(uint32_t S, uint64_t K, uint64_t tail_words) {
    return 2 * ((4 * K * S + tail_words + 3) / 4 * 4);
}

// Quad coder (coder 1): 64-bit state, 32-bit words, four 8-bit quantiles per step (the
// family-B quad coder of buffered/encoder.py:198-237 with P = 8 and one stream per
// coder).  Symbols in reverse decode order, starting from x = 2^32; words out in decode
// order [x_lo, x_hi, last flushed, ..., first flushed].
int encode_stream_quad(const int8_t *row, uint32_t cols, uint32_t S, uint32_t lane, int mn,
                       uint32_t G, const uint32_t *cdf, std::vector<uint32_t> &tmp,
                       std::vector<uint32_t> &out) {
    const uint32_t steps = cols / S;
    tmp.clear();
    uint64_t x = 1ull << 32;
    for (int64_t j = int64_t(steps / 4) - 1; j >= 0; j--) {
        uint32_t f[4], c[4];
        uint64_t F = 1;
        for (int i = 0; i < 4; i++) {
            uint32_t k = uint32_t(4 * j + i);
            uint32_t col = 16u * S * (k / 16) + 16u * lane + (k % 16);
            int s = int(row[col]) - mn;
            if (s < 0 || uint32_t(s) >= G) return -1;
            c[i] = cdf[s];
            f[i] = cdf[s + 1] - c[i];
            if (f[i] == 0 || f[i] > 255) return -1;
            F *= f[i];
        }
        if ((x >> 32) >= F) { tmp.push_back(uint32_t(x)); x >>= 32; }
        uint64_t concat = 0;
        for (int i = 0; i < 4; i++) {
            concat |= (uint64_t(c[i]) + x % f[i]) << (8 * i);
            x /= f[i];
        }
        x = (x << 32) | concat;
    }
    tmp.push_back(uint32_t(x >> 32));
    tmp.push_back(uint32_t(x));
    out.assign(tmp.rbegin(), tmp.rend());
    return 0;
}

// One stream, symbols in reverse decode order; output words in decode order.
int encode_stream(const int8_t *row, uint32_t cols, uint32_t S, uint32_t lane, uint32_t P,
                  int mn, uint32_t G, const uint32_t *cdf, std::vector<uint16_t> &tmp,
                  std::vector<uint16_t> &out) {
    const uint32_t steps = cols / S;
    tmp.clear();
    uint32_t x = 1u << 16;
    for (int64_t k = int64_t(steps) - 1; k >= 0; k--) {
        uint32_t col = 16u * S * uint32_t(k / 16) + 16u * lane + uint32_t(k % 16);
        int s = int(row[col]) - mn;
        if (s < 0 || uint32_t(s) >= G) return -1;
        uint32_t c0 = cdf[s], f = cdf[s + 1] - c0;
        if (f == 0) return -1;
        if ((x >> (32 - P)) >= f) { tmp.push_back(uint16_t(x & 0xFFFF)); x >>= 16; }
        x = ((x / f) << P) | (x % f + c0);
    }
    tmp.push_back(uint16_t(x & 0xFFFF));
    tmp.push_back(uint16_t(x >> 16));
    out.assign(tmp.rbegin(), tmp.rend());
    return 0;
}

}  // namespace

extern "C" int matvans_v1_model(const int8_t *mat, uint32_t rows, uint32_t cols,
                                uint32_t precision, uint32_t table_mode, int32_t *min_value,
                                uint32_t *G, uint16_t *cdf, uint8_t *freq) {
    if (!mat || !min_value || !G || !rows || !cols) return fail(MATVANS_E_INVALID, "null argument");
    if (!valid_params(32, precision, table_mode))
        return fail(MATVANS_E_UNSUPPORTED, "precision %u with table mode %u", precision, table_mode);
    if (table_mode == 0 && !cdf) return fail(MATVANS_E_INVALID, "null cdf");
    if (table_mode == 1 && !freq) return fail(MATVANS_E_INVALID, "null freq");
    const uint64_t n = uint64_t(rows) * cols;
    int mn = 127, mx = -128;
    for (uint64_t i = 0; i < n; i++) { mn = std::min<int>(mn, mat[i]); mx = std::max<int>(mx, mat[i]); }
    const uint32_t total = 1u << precision;
    *min_value = mn;
    if (table_mode == 0) {
        if (mn == mx) {   // dummy second symbol (create_matrices_WarpPerRow.py:135-143)
            *G = 2;
            cdf[0] = 0; cdf[1] = uint16_t(total - 1); cdf[2] = uint16_t(total);
            return MATVANS_OK;
        }
        const uint32_t g = uint32_t(mx - mn + 1);
        std::vector<uint64_t> counts(g, 0);
        for (uint64_t i = 0; i < n; i++) counts[mat[i] - mn]++;
        std::vector<uint32_t> f(g);
        if (int st = quantize_counts(counts.data(), g, total, f.data())) return st;
        cdf[0] = 0;
        for (uint32_t s = 0; s < g; s++) cdf[s + 1] = uint16_t(cdf[s] + f[s]);
        *G = g;
        mv::clear_error();
        return MATVANS_OK;
    }
    // one 8-bit table per row over the matrix alphabet; a row with one distinct symbol s
    // gets 255 / 1 with a dummy neighbour so that every frequency fits a byte
    const uint32_t g = std::max<uint32_t>(2, uint32_t(mx - mn + 1));
    std::vector<uint64_t> counts(g);
    std::vector<uint32_t> f(g);
    for (uint32_t r = 0; r < rows; r++) {
        std::fill(counts.begin(), counts.end(), 0);
        for (uint32_t c = 0; c < cols; c++) counts[mat[size_t(r) * cols + c] - mn]++;
        uint32_t present = 0, last = 0;
        for (uint32_t s = 0; s < g; s++) if (counts[s]) { present++; last = s; }
        uint8_t *fr = freq + size_t(r) * g;
        if (present == 1) {
            std::memset(fr, 0, g);
            fr[last] = 255;
            fr[last + 1 < g ? last + 1 : last - 1] = 1;
            continue;
        }
        if (int st = quantize_counts(counts.data(), g, 256, f.data()))
            return fail(st, "row %u: model cannot be normalised", r);
        for (uint32_t s = 0; s < g; s++) fr[s] = uint8_t(f[s]);
    }
    *G = g;
    mv::clear_error();
    return MATVANS_OK;
}

// ------------------------------------------------------------------ outlier split
// An 8-bit model gives every present value at least 1/256 of the code space, so a long
// tail of rare values costs every entry extra bits.  The split keeps the values within
// [-band, band] in the streams and codes every entry beyond as 0 (the most probable
// symbol), listing it as an outlier record col | (uint8)value << 24 per row; the decode
// adds value * x[col] for each record.
extern "C" int matvans_v1_outlier_band(const int8_t *mat, uint64_t n, int32_t *band) {
    if (!mat || !band || !n) return fail(MATVANS_E_INVALID, "null argument");
    uint64_t h[256] = {};
    for (uint64_t i = 0; i < n; i++) h[uint8_t(mat[i])]++;
    auto count = [&](int v) { return h[uint8_t(int8_t(v))]; };
    double best = -1.0;
    int best_t = 128;
    std::vector<uint64_t> c;
    std::vector<uint32_t> f;
    for (int t = 1; t <= 128; t++) {
        // band [-t, t] (t = 128: every value, no outliers)
        const int lo = t >= 128 ? -128 : -t, hi = t >= 128 ? 127 : t;
        c.assign(size_t(hi - lo + 1), 0);
        uint64_t outside = 0;
        for (int v = -128; v <= 127; v++) {
            if (v < lo || v > hi) outside += count(v);
            else c[size_t(v - lo)] += count(v);
        }
        c[size_t(-lo)] += outside;   // outliers are coded as 0
        uint32_t present = 0;
        for (auto x : c) present += x != 0;
        if (present < 2) continue;   // a one-symbol alphabet: the band below this one
        f.assign(c.size(), 0);
        if (quantize_counts(c.data(), uint32_t(c.size()), 256, f.data())) continue;
        double bits = 32.0 * double(outside);
        for (size_t s = 0; s < c.size(); s++)
            if (c[s]) bits -= double(c[s]) * std::log2(double(f[s]) / 256.0);
        if (best < 0 || bits < best) { best = bits; best_t = outside ? t : 128; }
    }
    *band = best_t >= 128 ? 127 : best_t;   // 127: the best band leaves no outliers
    mv::clear_error();
    return MATVANS_OK;
}

extern "C" int matvans_v1_split_outliers(const int8_t *mat, uint32_t rows, uint32_t cols,
                                         int32_t band, int8_t *coded, uint32_t *out_start,
                                         uint32_t *out_rec, uint64_t cap, uint64_t *n_out) {
    if (!mat || !coded || !out_start || !n_out) return fail(MATVANS_E_INVALID, "null argument");
    if (band < 0 || band > 127) return fail(MATVANS_E_INVALID, "band %d outside 0..127", band);
    if (cols >= (1u << 24)) return fail(MATVANS_E_SHAPE, "cols %u >= 2^24", cols);
    uint64_t n = 0;
    for (uint32_t r = 0; r < rows; r++) {
        if (n > 0xFFFFFFFFull) return fail(MATVANS_E_CAPACITY, "more than 2^32 outliers");
        out_start[r] = uint32_t(n);
        const int8_t *src = mat + size_t(r) * cols;
        int8_t *dst = coded + size_t(r) * cols;
        for (uint32_t c = 0; c < cols; c++) {
            const int v = src[c];
            if (v < -band || v > band) {
                if (n >= cap || !out_rec)
                    return fail(MATVANS_E_CAPACITY, "more than %llu outliers",
                                (unsigned long long)cap);
                out_rec[n++] = c | (uint32_t(uint8_t(int8_t(v))) << 24);
                dst[c] = 0;
            } else {
                dst[c] = int8_t(v);
            }
        }
    }
    if (n > 0xFFFFFFFFull) return fail(MATVANS_E_CAPACITY, "more than 2^32 outliers");
    out_start[rows] = uint32_t(n);
    *n_out = n;
    mv::clear_error();
    return MATVANS_OK;
}

extern "C" uint64_t matvans_v1_encode_bound(uint32_t rows, uint32_t cols, uint32_t streams,
                                            uint32_t precision) {
    if (!valid_params(streams, precision, 0) || !cols) return 0;
    // a stream's state gains at most P bits per symbol and loses 16 per word (the quad
    // coder's 32-bit words count as two)
    const uint64_t gw = 256 / streams;
    const uint64_t per_lane = (uint64_t(cols / streams) * precision + 15) / 16 + 8;
    return uint64_t(rows) * 256 * ((per_lane + gw - 1) / gw + 1);
}

extern "C" int matvans_v1_encode(const int8_t *mat, uint32_t rows, uint32_t cols,
                                 uint32_t streams, uint32_t precision, uint32_t table_mode,
                                 uint32_t coder, int32_t min_value, uint32_t G,
                                 const uint16_t *cdf, const uint8_t *freq, uint32_t n_threads,
                                 uint32_t *cursors, uint16_t *payload, uint64_t payload_cap,
                                 uint64_t *payload_size) {
    if (!mat || !cursors || !payload || !payload_size) return fail(MATVANS_E_INVALID, "null argument");
    if (!valid_params(streams, precision, table_mode, coder))
        return fail(MATVANS_E_UNSUPPORTED, "streams %u, precision %u, table mode %u, coder %u",
                    streams, precision, table_mode, coder);
    if (!cols || cols % (16 * streams))
        return fail(MATVANS_E_SHAPE, "cols %u not a multiple of %u", cols, 16 * streams);
    if (G < 2 || G > 256) return fail(MATVANS_E_INVALID, "G = %u", G);
    const uint32_t total = 1u << precision;
    if (table_mode == 0) {
        if (!cdf || cdf[0] != 0 || cdf[G] != total)
            return fail(MATVANS_E_INVALID, "cdf must run from 0 to %u", total);
        for (uint32_t s = 0; s < G; s++)
            if (cdf[s + 1] < cdf[s]) return fail(MATVANS_E_INVALID, "cdf decreases at %u", s);
    } else {
        if (!freq) return fail(MATVANS_E_INVALID, "null freq");
        for (uint32_t r = 0; r < rows; r++) {
            uint32_t sum = 0;
            for (uint32_t s = 0; s < G; s++) sum += freq[size_t(r) * G + s];
            if (sum != 256) return fail(MATVANS_E_INVALID, "row %u: frequencies sum to %u", r, sum);
        }
    }
    if (n_threads == 0) n_threads = std::max(1u, std::thread::hardware_concurrency());
    n_threads = std::min(n_threads, std::max(1u, rows));
    const uint32_t S = streams, GW = 256 / S, G4 = 128 / S;
    std::vector<uint32_t> kgr(rows);
    std::vector<std::vector<uint16_t>> st(coder == 0 ? size_t(rows) * S : 0);
    std::vector<std::vector<uint32_t>> st4(coder != 0 ? size_t(rows) * S : 0);
    std::atomic<int> bad{-1};
    auto pass1 = [&](uint32_t t) {
        std::vector<uint16_t> tmp;
        std::vector<uint32_t> tmp4;
        std::vector<uint32_t> c(G + 1);
        if (table_mode == 0) for (uint32_t s = 0; s <= G; s++) c[s] = cdf[s];
        for (uint32_t r = t; r < rows; r += n_threads) {
            if (table_mode == 1) {
                c[0] = 0;
                for (uint32_t s = 0; s < G; s++) c[s + 1] = c[s] + freq[size_t(r) * G + s];
            }
            uint32_t mx = 0, mn_len = 0xFFFFFFFFu;
            uint64_t sum = 0;
            for (uint32_t l = 0; l < S; l++) {
                int e;
                size_t len;
                if (coder == 0) {
                    auto &s = st[size_t(r) * S + l];
                    e = encode_stream(mat + size_t(r) * cols, cols, S, l, precision, min_value, G,
                                      c.data(), tmp, s);
                    len = s.size();
                } else {
                    auto &s = st4[size_t(r) * S + l];
                    e = encode_stream_quad(mat + size_t(r) * cols, cols, S, l, min_value, G,
                                           c.data(), tmp4, s);
                    len = s.size();
                }
                if (e) {
                    bad = int(r);
                    return;
                }
                mx = std::max<uint32_t>(mx, uint32_t(len));
                mn_len = std::min<uint32_t>(mn_len, uint32_t(len));
                sum += len;
            }
            // coder 2: the tails' words must fit the u16 end offsets
            const uint64_t K2 = mn_len / 4, tails = sum - 4 * K2 * S;
            if (coder == 2 && tails > 0xFFFFu) {
                bad = int(r);
                return;
            }
            // coder 2: the row's size in u16 words; otherwise its granules
            kgr[r] = coder == 0 ? (mx + GW - 1) / GW
                   : coder == 1 ? (mx + G4 - 1) / G4 : uint32_t(lanes_row_u16(S, K2, tails));
        }
    };
    std::vector<std::thread> th;
    for (uint32_t t = 1; t < n_threads; t++) th.emplace_back(pass1, t);
    pass1(0);
    for (auto &x : th) x.join();
    th.clear();
    if (bad >= 0) {
        if (coder == 2) {   // a symbol without probability, or tails beyond 65535 words
            size_t r = size_t(bad), total = 0, mn_len = ~size_t(0);
            for (uint32_t l = 0; l < S; l++) {
                total += st4[r * S + l].size();
                mn_len = std::min(mn_len, st4[r * S + l].size());
            }
            if (mn_len != ~size_t(0) && total - 4 * (mn_len / 4) * S > 0xFFFFu)
                return fail(MATVANS_E_SHAPE, "row %d: %zu tail words exceed the 16-bit "
                            "offsets of coder 2", bad.load(), total - 4 * (mn_len / 4) * S);
        }
        return fail(MATVANS_E_RANGE, "row %d holds a symbol the model gives no probability", int(bad));
    }
    const uint64_t head = coder == 2 ? lanes_table_u16(rows, S) : 0;
    uint64_t pos = head;
    for (uint32_t r = 0; r < rows; r++) {
        if (pos > 0xFFFFFFFFull) return fail(MATVANS_E_CAPACITY, "payload exceeds 2^32 words");
        cursors[r] = uint32_t(pos);
        pos += coder == 2 ? uint64_t(kgr[r]) : uint64_t(kgr[r]) * 256;
    }
    if (coder == 2 && pos <= payload_cap) std::memset(payload, 0, size_t(head) * 2);
    if (pos > 0xFFFFFFFFull) return fail(MATVANS_E_CAPACITY, "payload exceeds 2^32 words");
    if (pos > payload_cap) return fail(MATVANS_E_CAPACITY, "payload needs %llu words, cap %llu",
                                       (unsigned long long)pos, (unsigned long long)payload_cap);
    auto pass2 = [&](uint32_t t) {
        for (uint32_t r = t; r < rows; r += n_threads) {
            uint16_t *dst = payload + cursors[r];
            if (coder == 2) {
                std::memset(dst, 0, size_t(kgr[r]) * 2);
                size_t K = ~size_t(0);
                for (uint32_t l = 0; l < S; l++) K = std::min(K, st4[size_t(r) * S + l].size());
                K /= 4;
                uint16_t *tab = payload + size_t(r) * (S + 2);
                tab[0] = uint16_t(K);
                tab[1] = uint16_t(K >> 16);
                auto put = [](uint16_t *w, uint32_t v) { w[0] = uint16_t(v); w[1] = uint16_t(v >> 16); };
                uint16_t *tail = dst + 8 * K * S;   // tails after the K granules
                uint32_t end = 0;
                for (uint32_t l = 0; l < S; l++) {
                    const auto &s = st4[size_t(r) * S + l];
                    for (size_t j = 0; j < 4 * K; j++) put(dst + 2 * (4 * ((j / 4) * S + l) + j % 4), s[j]);
                    for (size_t j = 4 * K; j < s.size(); j++, end++) put(tail + 2 * size_t(end), s[j]);
                    tab[2 + l] = uint16_t(end);
                }
                continue;
            }
            std::memset(dst, 0, size_t(kgr[r]) * 512);
            for (uint32_t l = 0; l < S; l++) {
                if (coder == 0) {
                    const auto &s = st[size_t(r) * S + l];
                    for (size_t j = 0; j < s.size(); j++) dst[256 * (j / GW) + GW * l + j % GW] = s[j];
                } else {
                    const auto &s = st4[size_t(r) * S + l];
                    for (size_t j = 0; j < s.size(); j++) {
                        size_t w = 256 * (j / G4) + 2 * (G4 * l + j % G4);
                        dst[w] = uint16_t(s[j]);
                        dst[w + 1] = uint16_t(s[j] >> 16);
                    }
                }
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

// Decode table of a one-table matrix: A[q] = f | (uint8)(min + s) << 24 and
// B[q] = q - cdf[s] for every quantile q, with s the symbol cdf[s] <= q < cdf[s+1].
extern "C" int matvans_v1_lut(const uint16_t *cdf, uint32_t G, int32_t min_value,
                              uint32_t precision, uint32_t *lut) {
    if (!cdf || !lut) return fail(MATVANS_E_INVALID, "null argument");
    if (precision != 8 && precision != 12) return fail(MATVANS_E_UNSUPPORTED, "precision %u", precision);
    const uint32_t total = 1u << precision;
    if (G < 1 || G > 256 || cdf[0] != 0 || cdf[G] != total)
        return fail(MATVANS_E_INVALID, "cdf must run from 0 to %u", total);
    uint32_t s = 0;
    for (uint32_t q = 0; q < total; q++) {
        while (s + 1 < G && cdf[s + 1] <= q) s++;
        if (cdf[s] > q || cdf[s + 1] <= q) return fail(MATVANS_E_INVALID, "cdf decreases");
        uint32_t f = uint32_t(cdf[s + 1]) - cdf[s];
        lut[q] = f | (uint32_t(uint8_t(min_value + int(s))) << 24);
        lut[total + q] = q - cdf[s];
    }
    return MATVANS_OK;
}

// ---------------------------------------------------------------- MVX1 container
struct matvans_v1_file {
    uint32_t num_matrices = 0, result_hash = 0;
    std::vector<int8_t> vector;
    std::vector<std::vector<uint8_t>> blobs;
    std::vector<matvans_v1_host> mats;
};

namespace {
struct Rd {
    const uint8_t *p, *end;
    bool take(void *dst, size_t n) {
        if (size_t(end - p) < n) return false;
        if (dst) std::memcpy(dst, p, n);
        p += n;
        return true;
    }
};
size_t pad4(size_t n) { return (4 - n % 4) % 4; }
}  // namespace

extern "C" int matvans_v1_read(const char *path, matvans_v1_file **out) {
    if (!path || !out) return fail(MATVANS_E_INVALID, "null argument");
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(MATVANS_E_IO, "cannot open %s", path);
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0) { std::fclose(f); return fail(MATVANS_E_IO, "cannot size %s", path); }
    std::vector<uint8_t> file(static_cast<size_t>(sz));
    size_t got = sz ? std::fread(file.data(), 1, file.size(), f) : 0;
    std::fclose(f);
    if (got != file.size()) return fail(MATVANS_E_IO, "short read on %s", path);
    Rd rd{file.data(), file.data() + file.size()};
    char magic[4];
    uint32_t version = 0, len_v = 0;
    auto *v1 = new matvans_v1_file();
    auto bad = [&](const char *what, uint32_t k) {
        delete v1;
        return fail(MATVANS_E_FORMAT, "matrix %u: %s", k, what);
    };
    if (!rd.take(magic, 4) || std::memcmp(magic, "MVX1", 4) != 0) return bad("not an MVX1 file", 0);
    if (!rd.take(&version, 4) || (version != 1 && version != 2)) return bad("unsupported version", 0);
    if (!rd.take(&v1->num_matrices, 4) || !rd.take(&v1->result_hash, 4) || !rd.take(&len_v, 4))
        return bad("truncated header", 0);
    v1->vector.resize(len_v);
    if (!rd.take(v1->vector.data(), len_v) || !rd.take(nullptr, pad4(len_v)))
        return bad("truncated vector", 0);
    for (uint32_t k = 0; k < v1->num_matrices; k++) {
        matvans_v1_host m{};
        if (!rd.take(&m.rows, 4) || !rd.take(&m.cols, 4) || !rd.take(&m.grid_spacing, 4) ||
            !rd.take(&m.streams, 4) || !rd.take(&m.precision, 4) || !rd.take(&m.table_mode, 4) ||
            !rd.take(&m.coder, 4) || !rd.take(&m.min_value, 4) || !rd.take(&m.G, 4) ||
            !rd.take(&m.payload_size, 4))
            return bad("truncated record header", k);
        if (!valid_params(m.streams, m.precision, m.table_mode, m.coder)) return bad("bad parameters", k);
        if (m.G < 2 || m.G > 256 || m.min_value < -128 || m.min_value + int(m.G) - 1 > 127)
            return bad("bad alphabet", k);
        if (!m.cols || m.cols % (16 * m.streams)) return bad("cols do not fit the streams", k);
        const uint8_t *cur_p = rd.p;
        if (!rd.take(nullptr, size_t(m.rows) * 4)) return bad("truncated cursors", k);
        const uint8_t *tab_p = rd.p;
        size_t tab_n = m.table_mode == 0 ? size_t(m.G + 1) * 2 : size_t(m.rows) * m.G;
        if (!rd.take(nullptr, tab_n) || !rd.take(nullptr, pad4(tab_n))) return bad("truncated tables", k);
        const uint8_t *pay_p = rd.p;
        size_t pay_n = size_t(m.payload_size) * 2;
        if (!rd.take(nullptr, pay_n) || !rd.take(nullptr, pad4(pay_n))) return bad("truncated payload", k);
        // version 2: outlier count, then (when nonzero) out_start[rows + 1], out_rec[n]
        const uint8_t *ost_p = nullptr, *orec_p = nullptr;
        if (version == 2) {
            if (!rd.take(&m.n_outliers, 4)) return bad("truncated outlier count", k);
            if (m.n_outliers) {
                if (m.coder == 0 || m.table_mode != 0) return bad("outliers need the quad coder", k);
                ost_p = rd.p;
                if (!rd.take(nullptr, (size_t(m.rows) + 1) * 4)) return bad("truncated outliers", k);
                orec_p = rd.p;
                if (size_t(rd.end - rd.p) / 4 < m.n_outliers) return bad("truncated outliers", k);
                rd.take(nullptr, size_t(m.n_outliers) * 4);
            }
        }
        const size_t ob = m.n_outliers ? ((size_t(m.rows) + 1) * 4 + 15) / 16 * 16 : 0;
        const size_t orb = size_t(m.n_outliers) * 4;
        // own a 16-byte aligned copy: [payload | cursors | tables | out_start | out_rec]
        size_t pb = (pay_n + 15) / 16 * 16, cb = (size_t(m.rows) * 4 + 15) / 16 * 16;
        size_t tb = (tab_n + 15) / 16 * 16;
        std::vector<uint8_t> blob(pb + cb + tb + ob + orb + 32);
        uint8_t *a = blob.data();
        a += (16 - reinterpret_cast<uintptr_t>(a) % 16) % 16;
        std::memcpy(a, pay_p, pay_n);
        std::memcpy(a + pb, cur_p, size_t(m.rows) * 4);
        std::memcpy(a + pb + cb, tab_p, tab_n);
        if (m.n_outliers) {
            std::memcpy(a + pb + cb + tb, ost_p, (size_t(m.rows) + 1) * 4);
            std::memcpy(a + pb + cb + tb + ob, orec_p, orb);
            m.out_start = reinterpret_cast<const uint32_t *>(a + pb + cb + tb);
            m.out_rec = reinterpret_cast<const uint32_t *>(a + pb + cb + tb + ob);
            if (m.out_start[0] != 0 || m.out_start[m.rows] != m.n_outliers)
                return bad("bad outlier offsets", k);
            for (uint32_t r = 0; r < m.rows; r++)
                if (m.out_start[r + 1] < m.out_start[r]) return bad("outlier offsets decrease", k);
            for (uint32_t i = 0; i < m.n_outliers; i++)
                if ((m.out_rec[i] & 0xFFFFFFu) >= m.cols) return bad("outlier column", k);
        }
        m.payload = reinterpret_cast<const uint16_t *>(a);
        m.cursors = reinterpret_cast<const uint32_t *>(a + pb);
        if (m.table_mode == 0) {
            m.cdf = reinterpret_cast<const uint16_t *>(a + pb + cb);
            if (m.cdf[0] != 0 || m.cdf[m.G] != (1u << m.precision)) return bad("bad cdf", k);
            for (uint32_t s = 0; s < m.G; s++)
                if (m.cdf[s + 1] < m.cdf[s]) return bad("cdf decreases", k);
        } else {
            m.freq = a + pb + cb;
            for (uint32_t r = 0; r < m.rows; r++) {
                uint32_t sum = 0;
                for (uint32_t s = 0; s < m.G; s++) sum += m.freq[size_t(r) * m.G + s];
                if (sum != 256) return bad("row frequencies do not sum to 256", k);
            }
        }
        const uint32_t unit = m.coder == 2 ? 8u : 256u;   // 16-byte rows, or granules
        // coder 2: the row tables come first
        const uint64_t head = m.coder == 2 ? lanes_table_u16(m.rows, m.streams) : 0;
        if (head > m.payload_size) return bad("payload shorter than its row tables", k);
        for (uint32_t r = 0; r < m.rows; r++) {
            uint32_t end = r + 1 < m.rows ? m.cursors[r + 1] : m.payload_size;
            if (m.cursors[r] % unit || end < m.cursors[r] || end > m.payload_size ||
                (r == 0 && m.cursors[0] != head))
                return bad("bad cursors", k);
            if (m.coder == 2) {
                // K granules and the tails' end offsets: every stream holds at least its
                // two state words, and the granules and tails end within the row
                const uint16_t *t = m.payload + size_t(r) * (m.streams + 2);
                const uint64_t K = uint64_t(t[0]) | uint64_t(t[1]) << 16;
                const uint64_t room = (uint64_t(end) - m.cursors[r]) / 2;
                uint32_t prev = 0;
                for (uint32_t l = 0; l < m.streams; l++) {
                    const uint32_t e = t[2 + l];
                    if (e < prev || 4 * K + (e - prev) < 2) return bad("bad stream offsets", k);
                    prev = e;
                }
                if (4 * K * m.streams + uint64_t(prev) > room) return bad("streams overrun the row", k);
            }
        }
        if (m.rows && m.payload_size % unit) return bad("payload not a whole number of granules", k);
        v1->blobs.push_back(std::move(blob));
        v1->mats.push_back(m);
    }
    if (rd.p != rd.end) return bad("trailing bytes", v1->num_matrices);
    *out = v1;
    mv::clear_error();
    return MATVANS_OK;
}

extern "C" void matvans_v1_free(matvans_v1_file *f) { delete f; }

extern "C" int matvans_v1_header(const matvans_v1_file *f, uint32_t *num_matrices,
                                 uint32_t *result_hash, const int8_t **vector, uint32_t *len_v) {
    if (!f) return fail(MATVANS_E_INVALID, "null file");
    if (num_matrices) *num_matrices = f->num_matrices;
    if (result_hash) *result_hash = f->result_hash;
    if (vector) *vector = f->vector.data();
    if (len_v) *len_v = uint32_t(f->vector.size());
    return MATVANS_OK;
}

extern "C" int matvans_v1_matrix(const matvans_v1_file *f, uint32_t k, matvans_v1_host *out) {
    if (!f || !out) return fail(MATVANS_E_INVALID, "null argument");
    if (k >= f->mats.size()) return fail(MATVANS_E_INVALID, "matrix %u of %zu", k, f->mats.size());
    *out = f->mats[k];
    return MATVANS_OK;
}

extern "C" int matvans_v1_write(const char *path, uint32_t result_hash, const int8_t *vector,
                                uint32_t len_v, uint32_t num_matrices, const matvans_v1_host *mats) {
    if (!path || (len_v && !vector) || (num_matrices && !mats)) return fail(MATVANS_E_INVALID, "null argument");
    for (uint32_t k = 0; k < num_matrices; k++) {
        const auto &m = mats[k];
        if (!valid_params(m.streams, m.precision, m.table_mode, m.coder))
            return fail(MATVANS_E_UNSUPPORTED, "matrix %u: bad parameters", k);
        if ((m.table_mode == 0 && !m.cdf) || (m.table_mode == 1 && !m.freq) || !m.payload ||
            (m.rows && !m.cursors))
            return fail(MATVANS_E_INVALID, "matrix %u: null table, cursors or payload", k);
        if (m.n_outliers && (!m.out_start || !m.out_rec || m.coder == 0 || m.table_mode != 0))
            return fail(MATVANS_E_INVALID, "matrix %u: outliers need offsets, records and the "
                        "quad coder with one table", k);
    }
    bool any_outliers = false;
    for (uint32_t k = 0; k < num_matrices; k++) any_outliers |= mats[k].n_outliers != 0;
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(MATVANS_E_IO, "cannot create %s", path);
    static const uint8_t zeros[4] = {0, 0, 0, 0};
    auto put = [&](const void *p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; };
    const uint32_t version = any_outliers ? 2 : 1;
    bool ok = put("MVX1", 4) && put(&version, 4) && put(&num_matrices, 4) &&
              put(&result_hash, 4) && put(&len_v, 4) && put(vector, len_v) &&
              put(zeros, pad4(len_v));
    for (uint32_t k = 0; ok && k < num_matrices; k++) {
        const auto &m = mats[k];
        size_t tab_n = m.table_mode == 0 ? size_t(m.G + 1) * 2 : size_t(m.rows) * m.G;
        ok = put(&m.rows, 4) && put(&m.cols, 4) && put(&m.grid_spacing, 4) &&
             put(&m.streams, 4) && put(&m.precision, 4) && put(&m.table_mode, 4) &&
             put(&m.coder, 4) && put(&m.min_value, 4) && put(&m.G, 4) && put(&m.payload_size, 4) &&
             put(m.cursors, size_t(m.rows) * 4) &&
             put(m.table_mode == 0 ? static_cast<const void *>(m.cdf) : m.freq, tab_n) &&
             put(zeros, pad4(tab_n)) && put(m.payload, size_t(m.payload_size) * 2) &&
             put(zeros, pad4(size_t(m.payload_size) * 2));
        if (ok && version == 2) {
            ok = put(&m.n_outliers, 4);
            if (ok && m.n_outliers)
                ok = put(m.out_start, (size_t(m.rows) + 1) * 4) &&
                     put(m.out_rec, size_t(m.n_outliers) * 4);
        }
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(MATVANS_E_IO, "write failed on %s", path);
    mv::clear_error();
    return MATVANS_OK;
}
