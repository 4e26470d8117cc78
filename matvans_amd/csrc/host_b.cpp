// Host side of the reference's family-B codec ("buffered": quad rANS with 7-bit
// probabilities, 64-bit state and 32-bit words): the 7-bit model, the ans-warp-rows and
// ans-warp-stripes encoders (buffered/encoder.py:113-430, byte-exact with the
// reference), and a minimal reader of the safetensors container the reference stores
// them in (buffered/encoder.py:474-510; read by buffered/sane_safetensors.hpp:17-68).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "matvans_internal.h"

using mv::fail;

namespace {
constexpr uint32_t kW = 32;   // coders per warp in the reference files (encoder.py:441)

// DetachedAnsQuadCoder(7, 32, 64) encode (encoder.py:198-237).
struct QuadCoder {
    uint64_t state = 0;
    // quad: four symbol indices; cdf: 7-bit cdf (cdf[n] = 128)
    void encode(const int *s, const uint8_t *cdf, std::vector<uint32_t> &buf) {
        uint32_t left[4], prob[4];
        uint64_t total = 1;
        for (int i = 0; i < 4; i++) {
            left[i] = cdf[s[i]];
            prob[i] = uint32_t(cdf[s[i] + 1]) - left[i];
            total *= prob[i];
        }
        if ((state >> (64 - 28)) >= total) {
            buf.push_back(uint32_t(state & 0xFFFFFFFFu));
            state >>= 32;
        } else if (state == 0) {
            state = total << (64 - 32 - 28);
        }
        uint64_t concat = 0;
        for (int i = 0; i < 4; i++) {
            uint64_t rem = state % prob[i];
            state /= prob[i];
            concat |= (uint64_t(left[i]) + rem) << (7 * i);
        }
        state = (state << 28) | concat;
    }
};

// encoder.py:346-378: reverse the joint bulk, pad to 2W mod 4W words, insert the low then
// high state words after 2W bulk words, transpose chunks of 4 x W words.
void arrange(const std::vector<uint32_t> &buf, const QuadCoder *coders, uint32_t *out,
             size_t n_out) {
    std::vector<uint32_t> un;
    un.reserve(n_out);
    size_t nb = buf.size();
    auto bulk = [&](size_t i) { return i < nb ? buf[nb - 1 - i] : 0u; };
    for (size_t i = 0; i < 2 * kW; i++) un.push_back(bulk(i));
    for (uint32_t l = 0; l < kW; l++) un.push_back(uint32_t(coders[l].state & 0xFFFFFFFFu));
    for (uint32_t l = 0; l < kW; l++) un.push_back(uint32_t(coders[l].state >> 32));
    for (size_t i = 2 * kW; un.size() < n_out; i++) un.push_back(bulk(i));
    // chunk c: un[c*4W + k*W + l] -> out[c*4W + l*4 + k]
    for (size_t c = 0; c < n_out / (4 * kW); c++)
        for (uint32_t k = 0; k < 4; k++)
            for (uint32_t l = 0; l < kW; l++)
                out[c * 4 * kW + l * 4 + k] = un[c * 4 * kW + k * kW + l];
}

size_t arranged_words(size_t nb) {
    size_t pad = (2 * kW + 4 * kW - nb % (4 * kW)) % (4 * kW);   // (2W - nb) mod 4W
    return nb + pad + 2 * kW;
}
}  // namespace

extern "C" int matvans_b_model(const int8_t *data, uint64_t n, int32_t *min_value,
                               uint32_t *n_symbols, uint8_t *cdf, uint32_t *ppf,
                               uint32_t *ppf_cum, uint32_t *ppf_prob) {
    if (!data || !n || !min_value || !n_symbols || !cdf || !ppf || !ppf_cum || !ppf_prob)
        return fail(MATVANS_E_INVALID, "null argument or empty data");
    uint64_t hist[256] = {0};
    for (uint64_t i = 0; i < n; i++) hist[uint8_t(data[i]) ^ 0x80]++;
    // fit_128bit_encoder_model (encoder.py:268-306): float32 max(1, round(count*128/n)),
    // then the excess / deficit moved one unit at a time at the first maximum
    std::vector<int> values;
    std::vector<uint32_t> pmf;
    const float fn = float(n);
    for (int b = 0; b < 256; b++) {
        if (!hist[b]) continue;
        values.push_back(b - 128);
        float t = float(hist[b]) * 128.0f;
        t = t / fn;
        pmf.push_back(std::max<uint32_t>(1u, uint32_t(std::nearbyint(t))));
    }
    if (values.size() > 128)
        return fail(MATVANS_E_RANGE, "%zu distinct values: a 7-bit model holds at most 128",
                    values.size());
    int64_t sum = 0;
    for (auto p : pmf) sum += p;
    auto argmax = [&]() { return size_t(std::max_element(pmf.begin(), pmf.end()) - pmf.begin()); };
    for (; sum > 128; sum--) pmf[argmax()] -= 1;
    for (; sum < 128; sum++) pmf[argmax()] += 1;
    const int mn = values.front(), mx = values.back();
    const uint32_t ns = uint32_t(mx - mn + 1);
    uint32_t left = 0, last = 0;
    std::memset(cdf, 0, ns + 1);
    for (size_t i = 0; i < values.size(); i++) {
        uint32_t idx = uint32_t(values[i] - mn);
        for (uint32_t k = last + 1; k <= idx; k++) cdf[k] = uint8_t(left);
        last = idx;
        left += pmf[i];
    }
    cdf[ns] = uint8_t(left);
    // decoder_model_from_128bit_encoder_model (encoder.py:309-320), packed 4 per word
    uint8_t p8[128], c8[128], f8[128];
    for (uint32_t s = 0; s < ns; s++)
        for (uint32_t q = cdf[s]; q < cdf[s + 1]; q++) {
            p8[q] = uint8_t(int(s) + mn);
            c8[q] = cdf[s];
            f8[q] = uint8_t(cdf[s + 1] - cdf[s]);
        }
    for (int i = 0; i < 32; i++) {
        ppf[i] = ppf_cum[i] = ppf_prob[i] = 0;
        for (int k = 0; k < 4; k++) {
            ppf[i] |= uint32_t(p8[4 * i + k]) << (8 * k);
            ppf_cum[i] |= uint32_t(c8[4 * i + k]) << (8 * k);
            ppf_prob[i] |= uint32_t(f8[4 * i + k]) << (8 * k);
        }
    }
    *min_value = mn;
    *n_symbols = ns;
    mv::clear_error();
    return MATVANS_OK;
}

extern "C" uint64_t matvans_b_encode_bound(uint32_t rows, uint32_t cols, int layout) {
    // a coder emits at most one word per quad; per warp: states, padding, one chunk slack
    uint64_t quads = uint64_t(rows) * cols / 4;
    uint64_t warps = layout == 0 ? rows : (rows + kW - 1) / kW;
    return quads + warps * 6 * kW;
}

// compress_ans_warp_rows (encoder.py:381-430) for layout 0, compress_ans_warp_stripes
// (encoder.py:323-378) for layout 1.  The 32 coders carry their states from one warp to
// the next, as in the reference.  warp_offsets[i] = offset_base + (start of warp i in
// 16-byte units), as compress_all_matrices stores them (encoder.py:474-488).
extern "C" int matvans_b_encode(const int8_t *mat, uint32_t rows, uint32_t cols, int layout,
                                int32_t min_value, const uint8_t *cdf, uint32_t n_symbols,
                                uint32_t offset_base, uint32_t *compressed, uint64_t cap_words,
                                uint64_t *n_words, uint32_t *warp_offsets) {
    if (!mat || !cdf || !compressed || !n_words || !warp_offsets)
        return fail(MATVANS_E_INVALID, "null argument");
    if (layout == 0 && (cols == 0 || cols % (4 * kW)))
        return fail(MATVANS_E_SHAPE, "rows layout needs cols %% 128 == 0 (cols = %u)", cols);
    if (layout == 1 && (rows % kW || cols == 0 || cols % 4))
        return fail(MATVANS_E_SHAPE, "stripes layout needs rows %% 32 == 0 and cols %% 4 == 0");
    if (layout != 0 && layout != 1) return fail(MATVANS_E_INVALID, "layout %d", layout);
    if (n_symbols < 1 || n_symbols > 256 || cdf[0] != 0 || cdf[n_symbols] != 128)
        return fail(MATVANS_E_INVALID, "cdf must run from 0 to 128");
    QuadCoder coders[kW];
    std::vector<uint32_t> buf;
    uint64_t pos = 0;
    const uint32_t n_warps = layout == 0 ? rows : rows / kW;
    int s[4];
    for (uint32_t w = 0; w < n_warps; w++) {
        if (pos % 4 || offset_base + pos / 4 > 0xFFFFFFFFull)
            return fail(MATVANS_E_CAPACITY, "offsets exceed 32 bits");
        warp_offsets[w] = uint32_t(offset_base + pos / 4);
        buf.clear();
        if (layout == 0) {
            const int8_t *row = mat + size_t(w) * cols;
            for (int64_t c = int64_t(cols / (4 * kW)) - 1; c >= 0; c--)
                for (int64_t l = kW - 1; l >= 0; l--) {
                    const int8_t *q = row + (size_t(c) * kW + size_t(l)) * 4;
                    for (int k = 0; k < 4; k++) {
                        s[k] = int(q[k]) - min_value;
                        if (s[k] < 0 || uint32_t(s[k]) >= n_symbols || cdf[s[k] + 1] == cdf[s[k]])
                            return fail(MATVANS_E_RANGE, "row %u: symbol outside the model", w);
                    }
                    coders[l].encode(s, cdf, buf);
                }
        } else {
            for (int64_t c = int64_t(cols / 4) - 1; c >= 0; c--)
                for (int64_t l = kW - 1; l >= 0; l--) {
                    const int8_t *q = mat + (size_t(w) * kW + size_t(l)) * cols + size_t(c) * 4;
                    for (int k = 0; k < 4; k++) {
                        s[k] = int(q[k]) - min_value;
                        if (s[k] < 0 || uint32_t(s[k]) >= n_symbols || cdf[s[k] + 1] == cdf[s[k]])
                            return fail(MATVANS_E_RANGE, "stripe %u: symbol outside the model", w);
                    }
                    coders[l].encode(s, cdf, buf);
                }
        }
        size_t n_out = arranged_words(buf.size());
        if (pos + n_out > cap_words)
            return fail(MATVANS_E_CAPACITY, "compressed data exceeds %llu words",
                        (unsigned long long)cap_words);
        arrange(buf, coders, compressed + pos, n_out);
        pos += n_out;
    }
    *n_words = pos;
    mv::clear_error();
    return MATVANS_OK;
}

// ------------------------------------------------------------- safetensors reader
// Layout (buffered/safetensors.hh, read by sane_safetensors.hpp:17-68): u64 header
// length N, N bytes of JSON {"name": {"dtype": "U32", "shape": [..], "data_offsets":
// [begin, end]}, ..., "__metadata__": {...}}, then the data section.
struct matvans_st {
    std::vector<uint8_t> file;
    size_t data0 = 0;
    struct T {
        std::string dtype;
        std::vector<uint64_t> shape;
        uint64_t begin = 0, end = 0;
    };
    std::map<std::string, T> tensors;
};

namespace {
struct Json {
    const char *p, *e;
    bool ok = true;
    void ws() { while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++; }
    bool lit(char c) { ws(); if (p < e && *p == c) { p++; return true; } return false; }
    bool str(std::string &out) {
        ws();
        if (p >= e || *p != '"') return ok = false;
        p++;
        out.clear();
        while (p < e && *p != '"') {
            if (*p == '\\') { if (++p >= e) return ok = false; }
            out.push_back(*p++);
        }
        if (p >= e) return ok = false;
        p++;
        return true;
    }
    bool num(uint64_t &v) {
        ws();
        if (p >= e || *p < '0' || *p > '9') return ok = false;
        v = 0;
        while (p < e && *p >= '0' && *p <= '9') v = v * 10 + uint64_t(*p++ - '0');
        return true;
    }
    bool skip() {   // any value
        ws();
        if (p >= e) return ok = false;
        if (*p == '"') { std::string s; return str(s); }
        if (*p == '{' || *p == '[') {
            char close = *p == '{' ? '}' : ']';
            p++;
            if (lit(close)) return true;
            do {
                if (close == '}') { std::string k; if (!str(k) || !lit(':')) return ok = false; }
                if (!skip()) return false;
            } while (lit(','));
            return lit(close) || (ok = false);
        }
        while (p < e && *p != ',' && *p != '}' && *p != ']') p++;
        return true;
    }
};
}  // namespace

extern "C" int matvans_st_read(const char *path, matvans_st **out) {
    if (!path || !out) return fail(MATVANS_E_INVALID, "null argument");
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(MATVANS_E_IO, "cannot open %s", path);
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    auto *st = new matvans_st();
    st->file.resize(sz > 0 ? static_cast<size_t>(sz) : 0);
    size_t got = sz > 0 ? std::fread(st->file.data(), 1, st->file.size(), f) : 0;
    std::fclose(f);
    auto bad = [&](const char *what) { delete st; return fail(MATVANS_E_FORMAT, "%s: %s", path, what); };
    if (got != st->file.size() || st->file.size() < 8) return bad("truncated");
    uint64_t hn;
    std::memcpy(&hn, st->file.data(), 8);
    if (hn > st->file.size() - 8) return bad("header length past the end");
    st->data0 = 8 + size_t(hn);
    Json js{reinterpret_cast<const char *>(st->file.data()) + 8,
            reinterpret_cast<const char *>(st->file.data()) + 8 + hn};
    if (!js.lit('{')) return bad("header is not a JSON object");
    if (!js.lit('}')) {
        do {
            std::string name;
            if (!js.str(name) || !js.lit(':')) return bad("bad header");
            if (name == "__metadata__") { if (!js.skip()) return bad("bad metadata"); continue; }
            matvans_st::T t;
            if (!js.lit('{')) return bad("bad tensor entry");
            bool have_off = false;
            do {
                std::string key;
                if (!js.str(key) || !js.lit(':')) return bad("bad tensor entry");
                if (key == "dtype") {
                    if (!js.str(t.dtype)) return bad("bad dtype");
                } else if (key == "shape") {
                    if (!js.lit('[')) return bad("bad shape");
                    if (!js.lit(']')) {
                        do { uint64_t v; if (!js.num(v)) return bad("bad shape"); t.shape.push_back(v); }
                        while (js.lit(','));
                        if (!js.lit(']')) return bad("bad shape");
                    }
                } else if (key == "data_offsets") {
                    if (!js.lit('[') || !js.num(t.begin) || !js.lit(',') || !js.num(t.end) ||
                        !js.lit(']'))
                        return bad("bad data_offsets");
                    have_off = true;
                } else if (!js.skip()) {
                    return bad("bad tensor entry");
                }
            } while (js.lit(','));
            if (!js.lit('}') || !have_off) return bad("bad tensor entry");
            if (t.end < t.begin || t.end > st->file.size() - st->data0) return bad("data outside the file");
            st->tensors[name] = t;
        } while (js.lit(','));
        if (!js.lit('}')) return bad("bad header");
    }
    *out = st;
    mv::clear_error();
    return MATVANS_OK;
}

extern "C" void matvans_st_free(matvans_st *st) { delete st; }

extern "C" int matvans_st_tensor(const matvans_st *st, const char *name, const char **dtype,
                                 uint32_t *ndim, uint64_t *shape, const void **data,
                                 uint64_t *nbytes) {
    if (!st || !name) return fail(MATVANS_E_INVALID, "null argument");
    auto it = st->tensors.find(name);
    if (it == st->tensors.end()) return fail(MATVANS_E_FORMAT, "no tensor %s", name);
    const auto &t = it->second;
    if (dtype) *dtype = t.dtype.c_str();
    if (ndim) *ndim = uint32_t(t.shape.size());
    if (shape) for (size_t i = 0; i < t.shape.size() && i < 4; i++) shape[i] = t.shape[i];
    if (data) *data = st->file.data() + st->data0 + t.begin;
    if (nbytes) *nbytes = t.end - t.begin;
    return MATVANS_OK;
}
