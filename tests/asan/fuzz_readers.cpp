// Host-side sanitizer harness (built with -fsanitize=address,undefined by
// tests/test_host_asan.py; no GPU code): the library's file readers on the golden files,
// on every truncation of them (stepped) and on seeded single-byte corruptions, and the
// encoders on edge-case matrices.  A reader must reject damaged input with a status code
// and never read or write out of bounds; the sanitizers abort the run if it does.
//
//     fuzz_readers DIR [FILE...]    (FILE: .bin, .mvx or .safetensors; DIR for temporaries)
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "matvans.h"

namespace {

std::vector<uint8_t> slurp(const std::string &path) {
    std::vector<uint8_t> d;
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return d;
    uint8_t buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) d.insert(d.end(), buf, buf + n);
    std::fclose(f);
    return d;
}

void spit(const std::string &path, const uint8_t *p, size_t n) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror("write"); std::exit(3); }
    if (n) std::fwrite(p, 1, n, f);
    std::fclose(f);
}

// Reads the file with the reader for its kind and walks every record it returns.
int try_read(const std::string &path, const std::string &kind) {
    if (kind == "bin") {
        matvans_bin *b = nullptr;
        int st = matvans_bin_read(path.c_str(), &b);
        if (st) return st;
        uint32_t num = 0, rh = 0, mw = 0, lv = 0;
        const int8_t *v = nullptr;
        matvans_bin_header(b, &num, &rh, &mw, &v, &lv);
        volatile uint32_t sink = matvans_fxhash(v, lv);
        for (uint32_t k = 0; k < num; k++) {
            matvans_host_matrix m{};
            if (matvans_bin_matrix(b, k, &m)) continue;
            sink += m.cursors[m.rows - 1] + m.payload[m.payload_size - 1] + m.cdf[m.G] + m.ppf[255];
        }
        (void)sink;
        matvans_bin_free(b);
        return 0;
    }
    if (kind == "raw") {
        matvans_raw *r = nullptr;
        int st = matvans_raw_read(path.c_str(), &r);
        if (st) return st;
        uint32_t num = 0, rh = 0, lv = 0;
        const int8_t *v = nullptr;
        matvans_raw_header(r, &num, &rh, &v, &lv);
        volatile uint32_t sink = matvans_fxhash(v, lv);
        for (uint32_t k = 0; k < num; k++) {
            uint32_t rows = 0, cols = 0;
            float gs = 0.f;
            const int8_t *d = nullptr;
            if (matvans_raw_matrix(r, k, &rows, &cols, &gs, &d)) continue;
            if (size_t(rows) * cols) sink += matvans_fxhash(d, uint64_t(rows) * cols);
        }
        (void)sink;
        matvans_raw_free(r);
        return 0;
    }
    if (kind == "mvx") {
        matvans_v1_file *f = nullptr;
        int st = matvans_v1_read(path.c_str(), &f);
        if (st) return st;
        uint32_t num = 0, rh = 0, lv = 0;
        const int8_t *v = nullptr;
        matvans_v1_header(f, &num, &rh, &v, &lv);
        volatile uint32_t sink = matvans_fxhash(v, lv);
        for (uint32_t k = 0; k < num; k++) {
            matvans_v1_host m{};
            if (matvans_v1_matrix(f, k, &m)) continue;
            sink += m.cursors[m.rows - 1] + m.payload[m.payload_size - 1];
            sink += m.table_mode == 0 ? m.cdf[m.G] : m.freq[size_t(m.rows) * m.G - 1];
        }
        (void)sink;
        matvans_v1_free(f);
        return 0;
    }
    matvans_st *s = nullptr;
    int st = matvans_st_read(path.c_str(), &s);
    if (st) return st;
    const char *names[] = {"compressed", "all_warp_offsets", "all_ppfs", "all_ppfs_cum",
                           "all_ppfs_prob", "scales", "initial_vector", "hashes", "matrices",
                           "nonexistent"};
    volatile uint64_t sink = 0;
    for (const char *nm : names) {
        const char *dt = nullptr;
        uint32_t nd = 0;
        uint64_t shape[4] = {0, 0, 0, 0}, nb = 0;
        const void *data = nullptr;
        if (matvans_st_tensor(s, nm, &dt, &nd, shape, &data, &nb)) continue;
        if (nb) sink += static_cast<const uint8_t *>(data)[nb - 1];
    }
    (void)sink;
    matvans_st_free(s);
    return 0;
}

std::string kind_of(const std::string &f) {
    if (f.size() > 4 && f.compare(f.size() - 4, 4, ".bin") == 0) return "bin";
    if (f.size() > 4 && f.compare(f.size() - 4, 4, ".mvx") == 0) return "mvx";
    if (f.size() > 4 && f.compare(f.size() - 4, 4, ".raw") == 0) return "raw";
    return "st";
}

int g_cases = 0, g_rejected = 0;

// Truncations (every length below 64, then about 150 steps; each must be rejected) and
// seeded one-byte corruptions (half of them in the first 4 KiB: headers and tables; a
// flipped payload bit is a valid file) of one file.
bool fuzz_file(const std::string &path, const std::string &kind, const std::string &tmp) {
    std::vector<uint8_t> d = slurp(path);
    if (d.empty()) { std::fprintf(stderr, "cannot read %s\n", path.c_str()); return false; }
    if (try_read(path, kind)) { std::fprintf(stderr, "intact %s rejected\n", path.c_str()); return false; }
    const std::string p = tmp + "." + kind;
    const size_t step = d.size() / 150 + 1;
    for (size_t n = 0; n < d.size(); n += (n < 64 ? 1 : step)) {
        spit(p, d.data(), n);
        g_cases++;
        if (!try_read(p, kind)) {   // a truncated file must be rejected, not read short
            std::fprintf(stderr, "%s truncated to %zu bytes was accepted\n", path.c_str(), n);
            return false;
        }
        g_rejected++;
    }
    std::mt19937 rng(12345);
    for (int i = 0; i < 300; i++) {
        std::vector<uint8_t> c = d;
        size_t pos = (i % 2 ? rng() % std::min<size_t>(c.size(), 4096) : rng() % c.size());
        c[pos] ^= uint8_t(1u << (rng() % 8)) | uint8_t(i % 3 == 0 ? 0x80 : 0);
        spit(p, c.data(), c.size());
        g_cases++;
        if (try_read(p, kind)) g_rejected++;
    }
    std::remove(p.c_str());
    return true;
}

// Encoders on small edge-case matrices: constant rows, two values, the full int8 range,
// one row; every stream layout and coder the library writes.
bool encoders(const std::string &tmp) {
    bool fuzzed[2][2] = {{false, false}, {false, false}};
    std::mt19937 rng(7);
    const uint32_t shapes[][2] = {{1, 512}, {3, 1024}, {5, 2048}};
    for (auto &sh : shapes) {
        const uint32_t rows = sh[0], cols = sh[1];
        for (int kind = 0; kind < 4; kind++) {
            std::vector<int8_t> m(size_t(rows) * cols);
            for (auto &x : m) {
                int r = int(rng() % 255) - 127;
                x = int8_t(kind == 0 ? 5 : kind == 1 ? (r & 1) - 3 : kind == 2 ? r : r / 16);
            }
            int8_t mn = 0;
            uint32_t G = 0;
            uint16_t cdf16[257];
            uint8_t ppf[256];
            if (!matvans_build_model(m.data(), m.size(), &mn, &G, cdf16, ppf)) {
                std::vector<uint32_t> cur(rows);
                std::vector<uint16_t> pay(matvans_encode_bound(rows, cols));
                uint64_t size = 0;
                matvans_encode_rows(m.data(), rows, cols, mn, G, cdf16, 2, cur.data(), pay.data(),
                                    pay.size(), &size);
            }
            const uint32_t S_[] = {32, 64, 128};
            for (uint32_t S : S_) {
                if (cols % (16 * S)) continue;
                for (uint32_t P : {8u, 12u})
                    for (uint32_t mode : {0u, 1u})
                        for (uint32_t coder : {0u, 1u}) {
                            if ((mode == 1 && P != 8) || (coder == 1 && P != 8)) continue;
                            int32_t mv = 0;
                            uint32_t g = 0;
                            std::vector<uint16_t> cdf(257);
                            std::vector<uint8_t> freq(size_t(rows) * 256);
                            if (matvans_v1_model(m.data(), rows, cols, P, mode, &mv, &g, cdf.data(),
                                                 freq.data()))
                                continue;
                            std::vector<uint32_t> cur(rows);
                            std::vector<uint16_t> pay(matvans_v1_encode_bound(rows, cols, S, P));
                            uint64_t size = 0;
                            if (matvans_v1_encode(m.data(), rows, cols, S, P, mode, coder, mv, g,
                                                  cdf.data(), freq.data(), 2, cur.data(),
                                                  pay.data(), pay.size(), &size))
                                continue;
                            matvans_v1_host h{rows, cols, 1.0f, S, P, mode, coder, mv, g,
                                              uint32_t(size), cur.data(), cdf.data(),
                                              freq.data(), pay.data()};
                            int8_t vec[4] = {1, 2, 3, 4};
                            std::string p = tmp + "/enc.mvx";
                            if (matvans_v1_write(p.c_str(), 9, vec, 4, 1, &h)) continue;
                            // each table mode x coder: the first file written is fuzzed
                            if (!fuzzed[mode][coder] && kind == 3 && rows > 1) {
                                fuzzed[mode][coder] = true;
                                if (!fuzz_file(p, "mvx", tmp + "/_fuzz_tmp")) return false;
                            } else if (try_read(p, "mvx")) {
                                std::fprintf(stderr, "encoded MVX1 rejected: %s\n",
                                             matvans_last_error());
                                return false;
                            }
                        }
            }
            if (cols % 128 == 0 && kind != 2) {
                int32_t bm = 0;
                uint32_t ns = 0;
                uint8_t bcdf[257];   // the value range (-128..127) needs all 257
                uint32_t bp[32], bc[32], bf[32];
                if (!matvans_b_model(m.data(), m.size(), &bm, &ns, bcdf, bp, bc, bf)) {
                    for (int layout : {0, 1}) {
                        if (layout == 1 && rows % 32) continue;
                        std::vector<uint32_t> out(matvans_b_encode_bound(rows, cols, layout));
                        std::vector<uint32_t> offs(layout == 0 ? rows : rows / 32);
                        uint64_t nw = 0;
                        matvans_b_encode(m.data(), rows, cols, layout, bm, bcdf, ns, 0, out.data(),
                                         out.size(), &nw, offs.data());
                    }
                }
            }
        }
    }
    for (auto &m : fuzzed)
        for (bool f : m)
            if (!f) { std::fprintf(stderr, "an MVX1 variant was not fuzzed\n"); return false; }
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: fuzz_readers DIR [FILE...]\n");
        return 2;
    }
    const std::string dir = argv[1];
    for (int a = 2; a < argc; a++)
        if (!fuzz_file(argv[a], kind_of(argv[a]), dir + "/_fuzz_tmp")) return 1;
    if (!encoders(dir)) return 1;
    std::remove((dir + "/enc.mvx").c_str());
    std::printf("fuzz ok: %d damaged inputs, %d rejected\n", g_cases, g_rejected);
    return 0;
}
