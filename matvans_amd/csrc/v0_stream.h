// One lane's rANS stream of the WarpPerRow ".bin v0" layout on gfx950: table build,
// LDS word ring fed by 128-bit buffer loads, and the decode step.  Shared by the
// forward (row-major: a wave decodes two rows) and backward (stream-major: a wave
// decodes one stream index of 64 rows) kernels.
//
// Codec (matvansWarpPerRow-kernel.cu:95-157, AnsCoder.pop in
// create_matrices_WarpPerRow.py:250-264): 32-bit state x, 8-bit quantiles, 16-bit words.
//   q = x & 0xFF; s = ppf[q]; f = (cdf[s+1] - cdf[s]) & 0xFF;
//   x = (x >> 8) * f + (q - cdf[s]);  if x < 2^16: x = x << 16 | next word.
// Stream layout (create_matrices_WarpPerRow.py:405-422): word j of stream l of row r at
// u16 index cursors[r] + 256*(j/8) + 8*l + j%8; words 0 and 1 are the initial state.
//
// Rotated state: the lane keeps xr = rotr(x, 8).  Then x >> 8 is the low 24 bits of xr
// (v_mad_u32_u24 reads only those) and q is byte 3 of xr.  With table A[q] = f | sym<<24
// (bits 8..23 zero) and B[q] = q - cdf[s]:
//   x1 = mad_u24(xr, A[q], B[q])                       (= (x >> 8) * f + bias)
//   xr = perm(x1, word, x1 < 2^16 ? rotr(x1<<16|word) : rotr(x1))   (one v_perm_b32)
// A renormalised state is >= 2^24, so the following step leaves it >= 2^16: two
// consecutive renormalisations cannot happen, and the ring is read once per two steps.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "alu.h"

namespace mv {

constexpr int kRingWords = 64;             // ring words per lane (power of two)
constexpr uint32_t kRingWaveBytes = kRingWords * 64 * 2;   // 8 KiB per wave
constexpr uint32_t kOOB = 0x80000000u;     // buffer offset past num_records: reads 0
constexpr uint32_t kDummyBytes = 8 * 64 * 2;   // 8 rows x 64 lanes of u16 per block

// Fused table entry {f, symbol, q - cdf[s]} for quantile q (reference kernel :120-136;
// the symbol is (int8)(min_value + s) as in :126).  Reads cdf only inside [0, G].
__device__ __forceinline__ void v0_table_entry(uint32_t q, int min_value, uint32_t G,
                                               const uint8_t *__restrict__ cdf,
                                               const uint8_t *__restrict__ ppf, uint32_t &f,
                                               uint32_t &sym, uint32_t &bias) {
    uint32_t s = ppf[q];
    uint32_t c0 = s <= G ? cdf[s] : 0u;
    uint32_t c1 = s + 1 <= G ? cdf[s + 1] : 0u;
    f = (c1 - c0) & 0xFFu;
    sym = static_cast<uint8_t>(min_value + static_cast<int>(s));
    bias = (q - c0) & 0xFFu;
}

template <int CPOL>
__device__ __forceinline__ uint4 buffer_load16(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    auto d = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, CPOL);
    return make_uint4(d[0], d[1], d[2], d[3]);
}

// Selectors of v_perm_b32(x1, word, sel): rotr(x1 << 16 | word, 8) and rotr(x1, 8).
constexpr uint32_t kSelRenorm = 0x00050401u;
constexpr uint32_t kSelKeep = 0x04070605u;

// CPOL: the granule loads' cache policy (alu.h): kLoadStream in the forward, which reads
// each granule once; kLoadCached in the backward, whose waves share the lines.
template <int CPOL = kLoadCached>
struct V0LaneT {
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t lane_off;   // byte offset of this lane's 16-byte slot in granule 0
    uint32_t K;          // granules of this lane's row (0 = inactive lane)
    uint32_t gl;         // next granule to load
    uint32_t xr;         // rotated state
    uint32_t j;          // index of the next unread word of the stream
    uint32_t W;          // that word
    uint32_t waddr;      // LDS byte offset (within the ring array) of word j
    uint32_t wave_base;  // LDS byte offset of this wave's ring
    uint4 pa, pb;        // granule loads in flight
    uint32_t ga, gb;     // their granule indices (~0 = nothing to write)
    uint16_t *ring;      // the block's ring array [waves][kRingWords][64]
    uint16_t *dummy;     // the block's dummy rows [8][64] (written, never read)
    uint32_t lane;

    // win_beg/win_end: u16 window [first word, end) that holds every row of the wave
    // (wave-uniform); row_cur: this lane's row start (u16), row_end its end.
    __device__ void init(const uint16_t *payload, uint32_t win_beg, uint32_t win_end,
                         bool valid, uint32_t row_cur, uint32_t row_end, uint32_t slot,
                         uint16_t *ring_array, uint16_t *dummy_rows, uint32_t wave,
                         uint32_t lane_in_wave) {
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
        lane_off = roff + slot * 16u;
        ring = ring_array;
        dummy = dummy_rows;
        lane = lane_in_wave;
        wave_base = wave * kRingWaveBytes;
        ga = gb = 0xFFFFFFFFu;
    }

    __device__ __forceinline__ uint16_t &ring_at(uint32_t word_slot) {
        return ring[(wave_base >> 1) + word_slot * 64 + lane];
    }

    // Writes granule g into its ring slots.  g == ~0 (nothing loaded) writes the zeros of
    // the out-of-range load to the block's dummy rows instead, so that the write is not
    // behind a branch: the loads in flight then stay visible to the compiler's counters
    // and only the granule being written is waited for.
    __device__ __forceinline__ void write_granule(uint4 d, uint32_t g) {
        uint32_t s = (g * 8u) & (kRingWords - 1);
        uint16_t *p = g != 0xFFFFFFFFu ? &ring_at(s) : dummy + lane;
        p[0 * 64] = static_cast<uint16_t>(d.x);
        p[1 * 64] = static_cast<uint16_t>(d.x >> 16);
        p[2 * 64] = static_cast<uint16_t>(d.y);
        p[3 * 64] = static_cast<uint16_t>(d.y >> 16);
        p[4 * 64] = static_cast<uint16_t>(d.z);
        p[5 * 64] = static_cast<uint16_t>(d.z >> 16);
        p[6 * 64] = static_cast<uint16_t>(d.w);
        p[7 * 64] = static_cast<uint16_t>(d.w >> 16);
    }

    // Load granule gl when it will fit into the ring at write time (two groups later;
    // j only grows); otherwise issue an out-of-range load so that the number of loads in
    // flight is the same on every path.
    __device__ __forceinline__ void issue(uint4 &p, uint32_t &gp) {
        bool ok = gl < K && (gl + 1) * 8u - j <= uint32_t(kRingWords);
        p = buffer_load16<CPOL>(rsrc, ok ? lane_off + gl * 512u : kOOB);
        gp = ok ? gl : 0xFFFFFFFFu;
        gl += ok ? 1u : 0u;
    }

    // First 6 granules straight into the ring, then two loads in flight.  Call before
    // the block barrier; start() after it.
    __device__ void prologue() {
        constexpr uint32_t kPre = 6;
        uint4 pre[kPre];
#pragma unroll
        for (uint32_t g = 0; g < kPre; g++)
            pre[g] = buffer_load16<CPOL>(rsrc, g < K ? lane_off + g * 512u : kOOB);
#pragma unroll
        for (uint32_t g = 0; g < kPre; g++) write_granule(pre[g], g < K ? g : 0xFFFFFFFFu);
        gl = K < kPre ? K : kPre;
    }

    __device__ void start() {
        uint32_t x = (uint32_t(ring_at(0)) << 16) | ring_at(1);
        xr = (x >> 8) | (x << 24);
        j = 2;
        W = ring_at(2);
        waddr = wave_base | (2u * 128u) | (lane * 2u);
        issue(pa, ga);
        issue(pb, gb);
    }

    // Group boundaries (every 16 steps), alternating A and B so that the registers of
    // the loads in flight stay fixed.
    __device__ __forceinline__ void refill_a() {
        write_granule(pa, ga);
        issue(pa, ga);
    }
    __device__ __forceinline__ void refill_b() {
        write_granule(pb, gb);
        issue(pb, gb);
    }

    // Two decode steps; returns the two table-A entries (symbol in byte 3).
    __device__ __forceinline__ void pair(const uint32_t *lut, uint32_t &e0, uint32_t &e1) {
        uint32_t cc = 0;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint32_t q = xr >> 24;
            uint32_t ea = lut[q];
            uint32_t eb = lut[q + 256];
            uint32_t x1 = __umul24(xr, ea) + eb;
            bool c = x1 < 0x10000u;
            xr = __builtin_amdgcn_perm(x1, W, c ? kSelRenorm : kSelKeep);
            cc += c ? 1u : 0u;
            (h == 0 ? e0 : e1) = ea;
        }
        advance_word(cc);
    }

    // Two decode steps with the packed table T[q] = f | bias << 8 | sym << 24: one LDS
    // access per step, x1 = (x >> 8) * f + bias from SDWA byte selects.  Returns the two
    // table words (symbol in byte 3).
    __device__ __forceinline__ void pair_packed(const uint32_t *tab, uint32_t &e0, uint32_t &e1) {
        uint32_t cc = 0;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint32_t e = tab[xr >> 24];
            uint32_t x1 = add_b1(mul_d_b0(xr, e), e);
            bool c = x1 < 0x10000u;
            xr = __builtin_amdgcn_perm(x1, W, c ? kSelRenorm : kSelKeep);
            cc += c ? 1u : 0u;
            (h == 0 ? e0 : e1) = e;
        }
        advance_word(cc);
    }

    // Two decode steps with three 256-byte tables f[q], bias[q], sym[q] (at +0, +256,
    // +512): each table is 64 dwords, one per LDS bank, so every read is conflict-free.
    __device__ __forceinline__ void pair_u8(const uint8_t *tab, uint32_t &s0, uint32_t &s1) {
        uint32_t cc = 0;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint32_t q = xr >> 24;
            uint32_t f = tab[q];
            uint32_t b = tab[256 + q];
            uint32_t sy = tab[512 + q];
            uint32_t x1 = __umul24(xr, f) + b;
            bool c = x1 < 0x10000u;
            xr = __builtin_amdgcn_perm(x1, W, c ? kSelRenorm : kSelKeep);
            cc += c ? 1u : 0u;
            (h == 0 ? s0 : s1) = sy;
        }
        advance_word(cc);
    }

    // ring byte offset of word j: wave base | (j % 64) * 128 | lane * 2; the add may carry
    // into the wave-base bits, the mask restores them
    __device__ __forceinline__ void advance_word(uint32_t cc) {
        uint32_t t = waddr + (cc ? 128u : 0u);
        waddr = (t & (kRingWaveBytes - 1)) | (wave_base & ~(kRingWaveBytes - 1));
        W = *reinterpret_cast<const uint16_t *>(reinterpret_cast<const char *>(ring) + waddr);
    }

    // After a 16-step group: advance j by the words consumed (< 64 per group).
    __device__ __forceinline__ void end_group(uint32_t waddr_before) {
        j += ((waddr - waddr_before) >> 7) & uint32_t(kRingWords - 1);
    }
};
using V0Lane = V0LaneT<kLoadCached>;

}  // namespace mv
