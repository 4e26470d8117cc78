// One lane's rANS stream of the MVX1 layouts on gfx950: LDS word ring fed by buffer
// loads, and the decode steps for 8- and 12-bit probabilities.
//
// Layout: a row is K granules of 512 bytes; each of the row's S streams owns GW = 256/S
// consecutive u16 words of every granule (GW = 8 for S = 32, the reference layout of
// create_matrices_WarpPerRow.py:405-422; GW = 4 for S = 64; GW = 2 for S = 128).
// Word j of stream l is at u16 index cursor + 256*(j/GW) + GW*l + j%GW; words 0 and 1 are
// the initial state.
//
// Decode (P-bit probabilities, 32-bit state, 16-bit words):
//   q = x & (2^P - 1); x = (x >> P) * f[q] + bias[q]; if x < 2^16: x = x << 16 | word
// P = 8 keeps the state rotated (xr = rotr(x, 8), so x >> 8 is the low 24 bits that
// v_mad_u32_u24 reads and q is byte 3) and reads the ring once per two steps, because a
// renormalised state is >= 2^24 and the next step leaves it >= 2^16.  P = 12 keeps x
// plain and reads the ring after every step (two renormalisations can follow each other).
//
// Ring: R words per lane (a power of two), lane-interleaved u16 (word slot w of lane i
// at (w*64 + i)*2 bytes of the wave's ring), conflict-free for reads and writes.  Every
// 16-step group consumes at most M words (8 for P = 8, 16 for P = 12), and NG = M/GW
// granules are written per group, loaded two groups ahead.  A granule is loaded only when
// it will fit (the read index only grows); otherwise an out-of-range buffer offset
// returns zeros without touching memory and the write goes to dummy rows, so the number
// of loads in flight is fixed and the compiler's wait counters stay exact.
// Only the forward decodes these streams, each granule once: its loads are non-temporal
// (kLoadStream).
// No unwritten slot is ever read: a load blocked at group t-2 means the words written by
// group t exceed j(t-2) + R - GW >= j(t) - 2M + R - GW, i.e. reach j(t) + M + 1 whenever
// R >= 3M + GW (R = 32 for P = 8, R = 64 for P = 12); a group whose loads all went
// through adds M words, at least what the previous group consumed.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "alu.h"

namespace mv {
namespace v1 {

constexpr uint32_t kOOB = 0x80000000u;
constexpr uint32_t kSelRenorm8 = 0x00050401u;   // rotr(x1 << 16 | word, 8)
constexpr uint32_t kSelKeep8 = 0x04070605u;     // rotr(x1, 8)
constexpr uint32_t kSelRenorm12 = 0x05040100u;  // x1 << 16 | word
constexpr uint32_t kSelKeep12 = 0x07060504u;    // x1

template <int GW>
struct Chunk;
template <>
struct Chunk<8> {
    uint32_t d[4];
};
template <>
struct Chunk<4> {
    uint32_t d[2];
};
template <>
struct Chunk<2> {
    uint32_t d[1];
};

// CPOL: the loads' cache policy (kLoadCached or kLoadStream, alu.h)
template <int GW, int CPOL>
__device__ __forceinline__ Chunk<GW> load_chunk(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    Chunk<GW> c;
    if constexpr (GW == 8) {
        auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, CPOL);
        c.d[0] = v[0]; c.d[1] = v[1]; c.d[2] = v[2]; c.d[3] = v[3];
    } else if constexpr (GW == 4) {
        auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, CPOL);
        c.d[0] = v[0]; c.d[1] = v[1];
    } else {
        c.d[0] = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, CPOL);
    }
    return c;
}

template <int GW, int P, int R>
struct Lane {
    static constexpr int kRingWords = R;
    static constexpr uint32_t kRingWaveBytes = R * 64 * 2;
    static constexpr int M = (P == 8) ? 8 : 16;     // max words per 16-step group
    static constexpr int NG = M / GW;               // granules per set of M words
    // the ring is filled in sets of M words (NG granules) under one fit check, as the
    // quad coder's: no-unwritten-slot bound R >= 3M + M
    static_assert(R >= 4 * M && (R & (R - 1)) == 0, "ring too small");
    static constexpr int kPreSets = (R * 3 / 4) / M;   // sets loaded by the prologue

    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t lane_off;   // byte offset of this lane's slice in granule 0
    uint32_t K;          // granules of this lane's row (0 = inactive lane)
    uint32_t gs;         // next set to load
    uint32_t x;          // state (rotated for P = 8)
    uint32_t j;          // index of the next unread word
    uint32_t W;          // that word
    uint32_t waddr;      // LDS byte offset of word j within the ring array
    uint32_t wave_base;
    Chunk<GW> pa[NG], pb[NG];
    uint32_t sa, sb;     // their set index, or ~0 when not loaded
    uint16_t *ring, *dummy;   // dummy: M rows of 64 u16, written, never read
    uint32_t lane;

    // win_beg/win_end: wave-uniform u16 window holding every row of the wave;
    // row_cur/row_end: this lane's row; slot: stream index within the row.
    __device__ void init(const uint16_t *payload, uint32_t win_beg, uint32_t win_end, bool valid,
                         uint32_t row_cur, uint32_t row_end, uint32_t slot, uint16_t *ring_array,
                         uint16_t *dummy_rows, uint32_t wave, uint32_t lane_in_wave) {
        rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(payload) + win_beg, 0,
                                                 static_cast<int>((win_end - win_beg) * 2u),
                                                 0x00020000);
        K = 0;
        uint32_t roff = 0;
        if (valid) {
            if (row_end > win_end) row_end = win_end;
            if ((row_cur & 255u) == 0 && row_cur >= win_beg && row_end > row_cur) {
                K = (row_end - row_cur) >> 8;
                roff = (row_cur - win_beg) * 2u;
            }
        }
        lane_off = roff + slot * (GW * 2u);
        ring = ring_array;
        dummy = dummy_rows;
        lane = lane_in_wave;
        wave_base = wave * kRingWaveBytes;
        sa = sb = 0xFFFFFFFFu;
    }

    __device__ __forceinline__ uint16_t &ring_at(uint32_t w) {
        return ring[(wave_base >> 1) + w * 64 + lane];
    }

    // a set's NG granules from one byte offset, +512 apart; granules past the row read the
    // next row inside the window or zeros past it, into slots the fit check has freed, and
    // are never consumed
    __device__ __forceinline__ void load_set(Chunk<GW> (&p)[NG], uint32_t base) {
#pragma unroll
        for (int i = 0; i < NG; i++)
            p[i] = load_chunk<GW, kLoadStream>(rsrc, base + uint32_t(i) * 512u);
    }

    // set ~0 (nothing loaded) writes to the dummy rows, so the write is not behind a branch
    __device__ __forceinline__ void write_set(const Chunk<GW> (&p)[NG], uint32_t set) {
        uint16_t *q = set != 0xFFFFFFFFu ? &ring_at((set * M) & (kRingWords - 1)) : dummy + lane;
#pragma unroll
        for (int g = 0; g < NG; g++)
#pragma unroll
            for (int i = 0; i < GW / 2; i++) {
                q[(g * GW + 2 * i) * 64] = static_cast<uint16_t>(p[g].d[i]);
                q[(g * GW + 2 * i + 1) * 64] = static_cast<uint16_t>(p[g].d[i] >> 16);
            }
    }

    __device__ __forceinline__ void issue(Chunk<GW> (&p)[NG], uint32_t &sp) {
        const uint32_t g0 = gs * NG;
        const bool ok = g0 < K && (gs + 1) * M - j <= uint32_t(kRingWords);
        load_set(p, ok ? lane_off + g0 * 512u : kOOB);
        sp = ok ? gs : 0xFFFFFFFFu;
        gs += ok ? 1u : 0u;
    }

    __device__ void prologue() {
        Chunk<GW> pre[kPreSets][NG];
#pragma unroll
        for (uint32_t t = 0; t < uint32_t(kPreSets); t++)
            load_set(pre[t], t * NG < K ? lane_off + t * NG * 512u : kOOB);
#pragma unroll
        for (uint32_t t = 0; t < uint32_t(kPreSets); t++) write_set(pre[t], t);
        const uint32_t k_sets = (K + NG - 1) / NG;
        gs = k_sets < uint32_t(kPreSets) ? k_sets : uint32_t(kPreSets);
    }

    __device__ void start() {
        uint32_t x0 = (uint32_t(ring_at(0)) << 16) | ring_at(1);
        x = (P == 8) ? ((x0 >> 8) | (x0 << 24)) : x0;
        j = 2;
        W = ring_at(2);
        waddr = wave_base | (2u * 128u) | (lane * 2u);
        issue(pa, sa);
        issue(pb, sb);
    }

    __device__ __forceinline__ void refill_a() {
        write_set(pa, sa);
        issue(pa, sa);
    }
    __device__ __forceinline__ void refill_b() {
        write_set(pb, sb);
        issue(pb, sb);
    }

    // ring byte offset of the next word; the add may carry into the wave-base bits, the
    // mask restores them
    __device__ __forceinline__ void advance_word(uint32_t cc) {
        uint32_t t = waddr + (cc ? 128u : 0u);
        waddr = (t & (kRingWaveBytes - 1)) | (wave_base & ~(kRingWaveBytes - 1));
        W = *reinterpret_cast<const uint16_t *>(reinterpret_cast<const char *>(ring) + waddr);
    }

    __device__ __forceinline__ void end_group(uint32_t waddr_before) {
        j += ((waddr - waddr_before) >> 7) & uint32_t(kRingWords - 1);
    }

    // P = 8, two steps, dword tables A[q] = f | sym << 24 at lut[q], B[q] at lut[256 + q];
    // returns the two A entries (symbol in byte 3).
    __device__ __forceinline__ void pair8(const uint32_t *lut, uint32_t &e0, uint32_t &e1) {
        uint32_t cc = 0;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint32_t q = x >> 24;
            uint32_t ea = lut[q];
            uint32_t eb = lut[q + 256];
            uint32_t x1 = __umul24(x, ea) + eb;
            bool c = x1 < 0x10000u;
            x = __builtin_amdgcn_perm(x1, W, c ? kSelRenorm8 : kSelKeep8);
            cc += c ? 1u : 0u;
            (h == 0 ? e0 : e1) = ea;
        }
        advance_word(cc);
    }

    // P = 8, two steps, packed table T[q] = f | bias << 8 | sym << 24 (one LDS access per
    // step): x1 = (x >> 8) * f + bias with the byte fields taken by SDWA selects.
    // Returns the two table words (symbol in byte 3).
    __device__ __forceinline__ void pair8_packed(const uint32_t *tab, uint32_t &e0, uint32_t &e1) {
        uint32_t cc = 0;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint32_t e = tab[x >> 24];
            uint32_t x1 = add_b1(mul_d_b0(x, e), e);
            bool c = x1 < 0x10000u;
            x = __builtin_amdgcn_perm(x1, W, c ? kSelRenorm8 : kSelKeep8);
            cc += c ? 1u : 0u;
            (h == 0 ? e0 : e1) = e;
        }
        advance_word(cc);
    }

    // P = 8, two steps, byte tables f[q], bias[q], sym[q] at tab + 0 / 256 / 512 (tab may
    // differ per wave); returns the two symbols.
    __device__ __forceinline__ void pair8_u8(const uint8_t *tab, uint32_t &s0, uint32_t &s1) {
        uint32_t cc = 0;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint32_t q = x >> 24;
            uint32_t f = tab[q];
            uint32_t b = tab[256 + q];
            uint32_t sy = tab[512 + q];
            uint32_t x1 = __umul24(x, f) + b;
            bool c = x1 < 0x10000u;
            x = __builtin_amdgcn_perm(x1, W, c ? kSelRenorm8 : kSelKeep8);
            cc += c ? 1u : 0u;
            (h == 0 ? s0 : s1) = sy;
        }
        advance_word(cc);
    }

    // P = 12, one step, dword tables A[q] at lut[q], B[q] at lut[4096 + q]; returns A.
    __device__ __forceinline__ uint32_t step12(const uint32_t *lut) {
        uint32_t q = x & 0xFFFu;
        uint32_t ea = lut[q];
        uint32_t eb = lut[q + 4096];
        uint32_t x1 = __umul24(x >> 12, ea) + eb;
        bool c = x1 < 0x10000u;
        x = __builtin_amdgcn_perm(x1, W, c ? kSelRenorm12 : kSelKeep12);
        advance_word(c ? 1u : 0u);
        return ea;
    }
};

// Quad coder (coder 1): per-lane 64-bit state {lo, hi}, 32-bit words in a ring of
// R4 = 4 * SETW u32 words per lane (lane-interleaved: word slot w of lane i at
// (w*64 + i)*4 bytes of the wave's ring).  A 16-symbol group is four quads and takes at
// most 4 words.  The ring is filled in sets of M = SETW words per lane (SETW = 4: one set
// per group; SETW = 8: one set per two groups, which halves the per-set bookkeeping at
// twice the ring): NG = M / G4 consecutive granules (G4 = 128/S words per stream per
// granule), loaded two set periods ahead under one fit check, one ring position and one
// validity marker per set whatever S is.  A set period consumes at most M words, so the
// fit rule and the no-unwritten-slot bound are those of Lane with a set as the unit
// (R4 = 4M >= 3M + M).  A set may run past the row's last granule: those loads read the
// next row's words inside the window or return zeros past it (buffer range check), into
// ring slots the fit check has shown to be free, and the lane never consumes them.
//
// LAY 1 is coder 2's layout: a chunk is 4 consecutive words (16 bytes) of the lane's
// stream; the row's first Kb chunks of every stream sit in its body granules (chunk c of
// stream l at byte c * S * 16 + l * 16 of the body: one coalesced load per wave), the
// rest in the stream's tail, chunk Kb + t at 16 t bytes past the tail's first word (a
// dword-aligned 16-byte load).  The fit rule and the bounds above are unchanged.
//
// CPOL: kLoadStream for the forward's coder-1 granules (read once, whole); kLoadCached for
// coder 2's per-lane runs (a lane's next chunk shares the line) and for the backward,
// whose waves read neighbouring slots of the same lines.
template <int G4, int SETW = 4, int LAY = 0, int CPOL = kLoadCached>
struct QuadLane {
    static constexpr int CW = LAY ? 4 : G4;               // u32 words per chunk
    static constexpr uint32_t kStride = LAY ? 16u : 512u; // bytes between a lane's chunks
    static constexpr int M = SETW;
    static constexpr int R4 = 4 * SETW;
    static constexpr uint32_t kRingWaveBytes = R4 * 64 * 4;
    static constexpr int NG = M / CW;
    static_assert(R4 >= 3 * M + M, "ring too small");
    static_assert(NG >= 1 && (R4 & (R4 - 1)) == 0, "set shape");
    static constexpr int kPreSets = (R4 * 3 / 4) / M;

    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t lane_off, K, gs, j;   // K chunks of the lane (granules of the row); gs next set
    uint32_t Kb, gstride, tail_off;   // LAY 1: body chunks, body granule bytes, tail offset
    uint32_t lo, hi;       // state
    uint32_t W;            // next unread word
    uint32_t waddr;        // LDS byte offset of that word within the ring array
    uint32_t wave_base;
    Chunk<2 * CW> pa[NG], pb[NG];   // a set: NG chunks of 2*CW u16 = CW u32 words
    uint32_t sa, sb;                // their set index, or ~0 when not loaded
    uint32_t *ring, *dummy;         // dummy: M slots (M*64 words), written, never read
    uint32_t lane;

    __device__ void init(const uint16_t *payload, uint32_t win_beg, uint32_t win_end, bool valid,
                         uint32_t row_cur, uint32_t row_end, uint32_t slot, uint32_t *ring_array,
                         uint32_t *dummy_rows, uint32_t wave, uint32_t lane_in_wave) {
        rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(payload) + win_beg, 0,
                                                 static_cast<int>((win_end - win_beg) * 2u),
                                                 0x00020000);
        K = 0;
        uint32_t roff = 0;
        if (valid) {
            if (row_end > win_end) row_end = win_end;
            if ((row_cur & 255u) == 0 && row_cur >= win_beg && row_end > row_cur) {
                K = (row_end - row_cur) >> 8;
                roff = (row_cur - win_beg) * 2u;
            }
        }
        lane_off = roff + slot * (G4 * 4u);
        Kb = gstride = tail_off = 0;
        ring = ring_array;
        dummy = dummy_rows;
        lane = lane_in_wave;
        wave_base = wave * kRingWaveBytes;
        sa = sb = 0xFFFFFFFFu;
    }

    // LAY 1: stream `slot` of a row whose body starts at u32 index `body` of the payload
    // (kb chunks per stream, S streams) and whose own tail is tail_n words from u32 index
    // `tail`, inside the u16 window [win_beg, win_end) (win_beg a multiple of 8).
    __device__ void init_lanes(const uint16_t *payload, uint32_t win_beg, uint32_t win_end,
                               bool valid, uint32_t body, uint32_t kb, uint32_t S, uint32_t slot,
                               uint32_t tail, uint32_t tail_n, uint32_t *ring_array,
                               uint32_t *dummy_rows, uint32_t wave, uint32_t lane_in_wave) {
        static_assert(LAY == 1, "coder 2 layout");
        rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(payload) + win_beg, 0,
                                                 static_cast<int>((win_end - win_beg) * 2u),
                                                 0x00020000);
        K = Kb = 0;
        lane_off = tail_off = 0;
        gstride = S * 16u;
        const uint64_t body_end = uint64_t(body) + 4ull * kb * S;
        if (valid && 4u * kb + tail_n >= 2u && 2u * body >= win_beg && 2 * body_end <= win_end &&
            2u * tail >= win_beg && 2ull * (uint64_t(tail) + tail_n) <= win_end) {
            Kb = kb;
            K = kb + (tail_n + 3u) / 4u;
            lane_off = (2u * body - win_beg) * 2u + slot * 16u;
            tail_off = (2u * tail - win_beg) * 2u;
        }
        ring = ring_array;
        dummy = dummy_rows;
        lane = lane_in_wave;
        wave_base = wave * kRingWaveBytes;
        sa = sb = 0xFFFFFFFFu;
    }

    // byte offset of chunk c within the window.  LAY 1 computes both candidates and picks
    // one with a select: written as a conditional expression, the compiler put each arm
    // behind a branch whose address register was the destination of a load in flight, and
    // waited for every outstanding load (the prefetched sets too) at each set issue.
    __device__ __forceinline__ uint32_t chunk_off(uint32_t c) const {
        if constexpr (LAY == 1) {
            const uint32_t ob = lane_off + c * gstride;
            const uint32_t ot = tail_off + ((c - Kb) << 4);   // wraps for c < Kb: unused
            const uint32_t m = 0u - uint32_t(c < Kb);
            return (ob & m) | (ot & ~m);
        }
        return lane_off + c * kStride;
    }

    __device__ __forceinline__ uint32_t &ring_at(uint32_t w) {
        return ring[(wave_base >> 2) + w * 64 + lane];
    }

    // loads the NG chunks of set `set` (or, when !ok, out-of-range zeros)
    __device__ __forceinline__ void load_set(Chunk<2 * CW> (&p)[NG], uint32_t set, bool ok) {
#pragma unroll
        for (int i = 0; i < NG; i++) {
            const uint32_t off = chunk_off(set * NG + uint32_t(i));   // computed either way
            p[i] = load_chunk<2 * CW, CPOL>(rsrc, ok ? off : kOOB);
        }
    }

    __device__ __forceinline__ void write_set(const Chunk<2 * CW> (&p)[NG], uint32_t set) {
        uint32_t *q = set != 0xFFFFFFFFu ? &ring_at((set * M) & (R4 - 1)) : dummy + lane;
#pragma unroll
        for (int i = 0; i < NG; i++)
#pragma unroll
            for (int w = 0; w < CW; w++) q[(i * CW + w) * 64] = p[i].d[w];
    }

    __device__ __forceinline__ void issue(Chunk<2 * CW> (&p)[NG], uint32_t &sp) {
        const uint32_t g0 = gs * NG;
        const bool ok = g0 < K && (gs + 1) * M - j <= uint32_t(R4);
        load_set(p, gs, ok);
        sp = ok ? gs : 0xFFFFFFFFu;
        gs += ok ? 1u : 0u;
    }

    // The first sets straight into the ring; `between` runs after their loads are issued
    // and before their data is written (a block barrier there waits while they are in
    // flight instead of after every wave's arrived).
    template <class F>
    __device__ void prologue(F &&between) {
        Chunk<2 * CW> pre[kPreSets][NG];
#pragma unroll
        for (uint32_t t = 0; t < uint32_t(kPreSets); t++) load_set(pre[t], t, t * NG < K);
        between();
#pragma unroll
        for (uint32_t t = 0; t < uint32_t(kPreSets); t++) write_set(pre[t], t);
        const uint32_t k_sets = (K + NG - 1) / NG;
        gs = k_sets < uint32_t(kPreSets) ? k_sets : uint32_t(kPreSets);
    }
    __device__ void prologue() {
        prologue([] {});
    }

    __device__ void start() {
        lo = ring_at(0);
        hi = ring_at(1);
        j = 2;
        W = ring_at(2);
        waddr = wave_base | (2u * 256u) | (lane * 4u);
        issue(pa, sa);
        issue(pb, sb);
    }

    __device__ __forceinline__ void refill_a() {
        write_set(pa, sa);
        issue(pa, sa);
    }
    __device__ __forceinline__ void refill_b() {
        write_set(pb, sb);
        issue(pb, sb);
    }

    __device__ __forceinline__ void end_group(uint32_t waddr_before) {
        j += ((waddr - waddr_before) >> 8) & uint32_t(R4 - 1);
    }

    // One quad: refill if x < 2^32, then four independent lookups of the bytes q of lo
    // in the packed table T[q] = f | r << 8 | sym << 24 (r = q - cdf[s]; one dword per
    // quantile, so one LDS access per symbol) and
    // x = hi * f0 f1 f2 f3 + (r0 + f0 (r1 + f1 (r2 + f2 r3))), with the byte fields taken
    // by SDWA operand selects.  Returns the four symbols packed in column order.
    __device__ __forceinline__ uint32_t quad(const uint32_t *tab) {
        uint32_t e[4];
        quad_entries(tab, e);
        return __builtin_amdgcn_perm(e[1], e[0], 0x0c0c0703u) |
               __builtin_amdgcn_perm(e[3], e[2], 0x07030c0cu);
    }

    // The same step with the two-word table T2[q] = (f | sym << 24, r): one 8-byte lookup
    // per symbol, and the products from whole words with 24-bit multiplies (the symbol
    // byte sits above the 24 bits they read): F = (f0 f1)(f2 f3),
    // R = (r0 + f0 r1) + f0 f1 (r2 + f2 r3), six instructions instead of eight.
    __device__ __forceinline__ uint32_t quad(const uint2 *tab) {
        uint32_t e[4];
        quad_entries(tab, e);
        return __builtin_amdgcn_perm(e[1], e[0], 0x0c0c0703u) |
               __builtin_amdgcn_perm(e[3], e[2], 0x07030c0cu);
    }
    __device__ __forceinline__ void quad_entries(const uint2 *tab, uint32_t (&e)[4]) {
        bool need = hi == 0;
        hi = need ? lo : hi;
        lo = need ? W : lo;
        uint32_t t = waddr + (need ? 256u : 0u);
        waddr = (t & (kRingWaveBytes - 1)) | (waddr & ~(kRingWaveBytes - 1));
        W = *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(ring) + waddr);
        uint2 w[4];
#pragma unroll
        for (int i = 0; i < 4; i++) w[i] = tab[(lo >> (8 * i)) & 0xFFu];
        const uint32_t f01 = mul_u24(w[0].x, w[1].x);
        const uint32_t f23 = mul_u24(w[2].x, w[3].x);
        const uint32_t r01 = mad_u24(w[0].x, w[1].y, w[0].y);
        const uint32_t r23 = mad_u24(w[2].x, w[3].y, w[2].y);
        const uint32_t F = mul_u24(f01, f23);
        const uint32_t R = mad_u24(r23, f01, r01);
        uint64_t nx = uint64_t(hi) * F + R;
        lo = uint32_t(nx);
        hi = uint32_t(nx >> 32);
#pragma unroll
        for (int i = 0; i < 4; i++) e[i] = w[i].x;
    }

    // The same step, returning the four table entries (symbol in byte 3 of each, in
    // column order) instead of the packed symbols.
    __device__ __forceinline__ void quad_entries(const uint32_t *tab, uint32_t (&e)[4]) {
        bool need = hi == 0;
        hi = need ? lo : hi;
        lo = need ? W : lo;
        // the add may carry into the wave's base bits; the high bits are taken from the
        // old address (one bit-field insert)
        uint32_t t = waddr + (need ? 256u : 0u);
        waddr = (t & (kRingWaveBytes - 1)) | (waddr & ~(kRingWaveBytes - 1));
        W = *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(ring) + waddr);
#pragma unroll
        for (int i = 0; i < 4; i++) e[i] = tab[(lo >> (8 * i)) & 0xFFu];
        uint32_t f01 = mul_b0_b0(e[0], e[1]);
        uint32_t r01 = add_b1(mul_b1_b0(e[1], e[0]), e[0]);
        uint32_t f23 = mul_b0_b0(e[2], e[3]);
        uint32_t r23 = add_b1(mul_b1_b0(e[3], e[2]), e[2]);
        uint32_t F = mul_u24(f01, f23);
        uint32_t R = mad_u24(r23, f01, r01);
        uint64_t nx = uint64_t(hi) * F + R;
        lo = uint32_t(nx);
        hi = uint32_t(nx >> 32);
    }
};

}  // namespace v1
}  // namespace mv
