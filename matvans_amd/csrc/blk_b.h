// Block-level body of the family-B fused decode x int8 matvec on gfx950, shared by the
// one-matrix kernel (decode_b.hip) and the persistent batched kernel (batch.hip)
// (buffered/ans-warp-rows-kernel.cu:7-221, buffered/ans-warp-stripes-kernel.cu:21-227).
//
// Codec: each of the 32 coders of a warp holds a 64-bit state; a step decodes a quad of
// four 7-bit quantiles (bits 0..27 of the state), then
//   state = (state >> 28) * p0*p1*p2*p3 + (r0 + p0*(r1 + p1*(r2 + p2*r3)))
// and a coder whose state fell below 2^32 first takes the next 32-bit word of the warp's
// shared stream, the refilling coders taking words in lane order.  Layout ans-warp-rows:
// a warp is one row, coder l's quad at step t covers columns 4*(32t + l) .. +3.
// Layout ans-warp-stripes: a warp is 32 consecutive rows (coder l = row l), step t is
// columns 4t .. 4t+3 of every row.
//
// Mapping on CDNA4: a wave64 runs two warps, one per 32-lane half.  A warp's word stream
// goes through an LDS ring of 512 words per warp, filled one 512-byte chunk (a uint4
// per lane) per four steps, loaded two chunk-groups ahead.  The refill replaces the
// reference's register cache and shuffles: a ballot of the coders that need a word,
// v_mbcnt for each lane's rank in its half, and one LDS read.  Stream word n (state
// words excluded) lives in ring slot (n + 64) & 511, so a chunk never wraps.
// Table in LDS: one packed dword prob | (q - cum) << 8 | symbol << 24 per quantile, so a
// symbol costs one lookup and the products take the byte fields by SDWA selects (alu.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "alu.h"

namespace mv {
namespace blk {

constexpr int kBWaves = 4;
constexpr uint32_t kRing = 512;          // words per warp
constexpr uint32_t kOOB = 0x80000000u;

__device__ __forceinline__ uint32_t byte_of(uint32_t w, uint32_t k) { return (w >> (8 * k)) & 0xFFu; }

// Packed byte tables (ppf, cum, prob: entry q is byte q%4 of word q/4) to the fused
// dword table in LDS: T[q] = prob | (q - cum) << 8 | symbol << 24 (prob <= 128 and
// q - cum < prob, so both fit a byte).
__device__ __forceinline__ void build_tables(uint32_t *lut, const uint32_t *ppf,
                                             const uint32_t *cum, const uint32_t *prob,
                                             uint32_t tid) {
    if (tid < 128) {
        uint32_t w = tid >> 2, k = tid & 3;
        uint32_t sym = byte_of(ppf[w], k), c = byte_of(cum[w], k), f = byte_of(prob[w], k);
        lut[tid] = f | (((tid - c) & 0xFFu) << 8) | (sym << 24);
    }
}

// One warp's stream (one 32-lane half of the wave).
struct BStream {
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t lane_off;      // byte offset of this lane's uint4 in chunk 0 (window relative)
    uint32_t K;             // chunks of the warp
    uint32_t gl;            // next chunk to load
    uint32_t ring_base;     // LDS byte offset of this half's ring within the ring array
    uint32_t lane;          // lane within the half
    uint4 pa, pb;
    uint32_t ga, gb;
    uint32_t *ring;
    uint32_t *dummy;

    __device__ __forceinline__ void write_chunk(const uint4 &d, uint32_t g) {
        // chunk g >= 1: stream words 64 + 128(g-1) + 32k + lane -> slots 128g + 32k + lane
        uint32_t *p = g != 0xFFFFFFFFu ? ring + (ring_base >> 2) + ((g * 128u) & (kRing - 1)) + lane
                                       : dummy + lane;
        p[0] = d.x;
        p[32] = d.y;
        p[64] = d.z;
        p[96] = d.w;
    }
    // chunk g fits if its last stream word minus the read position is within the ring
    __device__ __forceinline__ void issue(uint4 &p, uint32_t &gp, uint32_t cur) {
        bool ok = gl < K && (64u + 128u * gl) - cur <= kRing;
        auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, ok ? lane_off + gl * 512u : kOOB, 0, 0);
        p = make_uint4(v[0], v[1], v[2], v[3]);
        gp = ok ? gl : 0xFFFFFFFFu;
        gl += ok ? 1u : 0u;
    }
};

// One family-B matrix with its vectors and outputs (device pointers).
struct BJob {
    int32_t *acc32;          // [n_vec][out_dim] or null
    int8_t *out8;            // [n_vec][out_dim] or null
    float scale;
    const uint32_t *compressed;
    uint32_t compressed_words;
    const uint32_t *warp_offsets;   // 16-byte units
    uint32_t end_offset;
    const uint32_t *ppf, *cum, *prob;
    const int8_t *vecs;
    uint32_t vec_stride, n_vec, in_dim, out_dim;
};

struct BSmem {
    uint32_t ring[kBWaves * 2 * kRing];
    uint32_t lut[256];
    uint32_t dummy[128];
};

template <int NV, int LAYOUT>
__device__ __forceinline__ void b_block(const BJob &J, uint32_t bid, BSmem &sm) {
    uint32_t *ring = sm.ring;
    uint32_t *lut = sm.lut;
    uint32_t *dummy = sm.dummy;
    int32_t *__restrict__ acc32 = J.acc32;
    int8_t *__restrict__ out8 = J.out8;
    const float scale = J.scale;
    const uint32_t *__restrict__ compressed = J.compressed;
    const uint32_t compressed_words = J.compressed_words;
    const uint32_t *__restrict__ warp_offsets = J.warp_offsets;
    const uint32_t end_offset = J.end_offset;
    const uint32_t *__restrict__ ppf = J.ppf;
    const uint32_t *__restrict__ cum = J.cum;
    const uint32_t *__restrict__ prob = J.prob;
    const int8_t *__restrict__ vecs = J.vecs;
    const uint32_t vec_stride = J.vec_stride, in_dim = J.in_dim, out_dim = J.out_dim;

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6, lane64 = tid & 63, half = lane64 >> 5, lane = lane64 & 31;
    build_tables(lut, ppf, cum, prob, tid);

    // warps of this wave: rows layout: warp = row; stripes: warp = 32 rows
    const uint32_t n_warps = LAYOUT == 0 ? out_dim : out_dim / 32u;
    const uint32_t warp = (bid * kBWaves + wave) * 2 + half;
    const uint32_t warp0 = (bid * kBWaves + wave) * 2;
    const bool valid = warp < n_warps;
    // u32-word window [wbeg, wend) holding both warps of the wave (offsets in uint4 units)
    uint32_t wbeg = 0, wend = 0;
    if (warp0 < n_warps) {
        wbeg = warp_offsets[warp0];
        wend = warp0 + 2 < n_warps ? warp_offsets[warp0 + 2] : end_offset;
    }
    wbeg = __builtin_amdgcn_readfirstlane(wbeg) * 4u;
    wend = __builtin_amdgcn_readfirstlane(wend) * 4u;
    if (wend > compressed_words) wend = compressed_words;
    if (wend < wbeg) wend = wbeg;
    BStream st;
    st.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(compressed) + wbeg, 0,
                                                int((wend - wbeg) * 4u), 0x00020000);
    uint32_t my_beg = valid ? warp_offsets[warp] * 4u : wbeg;
    uint32_t my_end = valid ? (warp + 1 < n_warps ? warp_offsets[warp + 1] : end_offset) * 4u : wbeg;
    if (my_end > wend) my_end = wend;
    if (my_beg < wbeg || my_beg > my_end) my_beg = my_end;
    st.K = valid ? (my_end - my_beg) / 128u : 0u;
    st.lane_off = (my_beg - wbeg) * 4u + lane * 16u;
    st.lane = lane;
    st.ring = ring;
    st.dummy = dummy;
    st.ring_base = (wave * 2 + half) * kRing * 4u;

    // chunk 0: words x, y -> stream 0..63 (slots 64..127), z, w -> initial state
    uint32_t lo = 0, hi = 0;
    {
        auto v = __builtin_amdgcn_raw_buffer_load_b128(st.rsrc, st.K ? st.lane_off : kOOB, 0, 0);
        ring[(st.ring_base >> 2) + 64 + lane] = v[0];
        ring[(st.ring_base >> 2) + 96 + lane] = v[1];
        lo = v[2];
        hi = v[3];
    }
    // prologue: chunks 1 and 2, then two loads in flight
    {
        uint4 c1, c2;
        uint32_t g1 = 0xFFFFFFFFu, g2 = 0xFFFFFFFFu;
        st.gl = 1;
        st.issue(c1, g1, 0);
        st.issue(c2, g2, 0);
        st.write_chunk(c1, g1);
        st.write_chunk(c2, g2);
    }
    __syncthreads();
    st.issue(st.pa, st.ga, 0);
    st.issue(st.pb, st.gb, 0);

    // per-half read positions (stream word index), wave-uniform
    uint32_t cur_a = 0, cur_b = 0;
    const uint32_t half_sel = half;

    int32_t acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = 0;

    const uint32_t n_iter = LAYOUT == 0 ? in_dim / 128u : in_dim / 4u;

    auto step = [&](uint32_t t, const uint32_t *vq) {
        // refill: coders whose state is below 2^32 take the next words in lane order
        bool need = hi == 0;
        uint64_t m = __ballot(need);
        uint32_t mlo = uint32_t(m), mhi = uint32_t(m >> 32);
        uint32_t plo = __popc(mlo), phi = __popc(mhi);
        uint32_t base = half_sel ? cur_b - plo : cur_a;
        uint32_t n = __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, base));
        uint32_t w = ring[(st.ring_base >> 2) + ((n + 64u) & (kRing - 1))];
        hi = need ? lo : hi;
        lo = need ? w : lo;
        cur_a += plo;
        cur_b += phi;
        // decode a quad: four lookups of the packed table, the byte fields taken by SDWA
        // operand selects: P = p0 p1 p2 p3, R = r0 + p0 (r1 + p1 (r2 + p2 r3))
        uint32_t e[4];
#pragma unroll
        for (int i = 0; i < 4; i++) e[i] = lut[(lo >> (7 * i)) & 0x7Fu];
        const uint32_t p01 = mul_b0_b0(e[0], e[1]);
        const uint32_t r01 = add_b1(mul_b1_b0(e[1], e[0]), e[0]);
        const uint32_t p23 = mul_b0_b0(e[2], e[3]);
        const uint32_t r23 = add_b1(mul_b1_b0(e[3], e[2]), e[2]);
        const uint32_t P = mul_u24(p01, p23);
        const uint32_t R = mad_u24(r23, p01, r01);
        uint64_t x = ((uint64_t(hi) << 32) | lo) >> 28;
        uint64_t ns = x * P + R;
        lo = uint32_t(ns);
        hi = uint32_t(ns >> 32);
        uint32_t packed = __builtin_amdgcn_perm(e[1], e[0], 0x0c0c0703u) |
                          __builtin_amdgcn_perm(e[3], e[2], 0x07030c0cu);
#pragma unroll
        for (int v = 0; v < NV; v++)
            acc[v] = __builtin_amdgcn_sdot4(int(packed), int(vq[v]), acc[v], false);
        (void)t;
    };

    // Vector words of a 4-step group gg, loaded with the chunk loads two groups ahead
    // and kept in registers: rows: words 32t + lane of each step t; stripes: word t (the
    // same for all lanes).  Buffer loads over each vector's in_dim bytes return zeros past
    // the end; they count on the vector-memory counter, so the LDS waits of the decode
    // never wait for them (a scalar load would share the LDS counter).
    __amdgpu_buffer_rsrc_t vr[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
        vr[v] = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<int8_t *>(vecs) + size_t(v < int(J.n_vec) ? v : 0) * vec_stride, 0,
            int(in_dim), 0x00020000);
    auto vload_group = [&](uint32_t gg, uint32_t (&w)[4][NV]) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t t = gg * 4 + k;
            const uint32_t off = LAYOUT == 0 ? (t * 32u + lane) * 4u : t * 4u;
#pragma unroll
            for (int v = 0; v < NV; v++)
                w[k][v] = __builtin_amdgcn_raw_buffer_load_b32(vr[v], t < n_iter ? off : kOOB, 0, 0);
        }
    };

    // groups of four steps; chunk writes and loads alternate A/B between groups, and so do
    // the vector-word registers
    const uint32_t groups = (n_iter + 3) / 4;
    uint32_t va[4][NV], vb[4][NV];
    vload_group(0, va);
    vload_group(1, vb);
    // full groups in pairs (A, B buffers), then at most one full group on A and one partial
    // group; the loop body has no conditional steps, so the read counters stay scalar
    const uint32_t full = n_iter / 4;
    uint32_t wk[4][NV];
    auto take = [&](uint32_t (&w)[4][NV]) {
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int v = 0; v < NV; v++) wk[k][v] = w[k][v];
    };
    uint32_t g = 0;
    for (; g + 1 < full; g += 2) {
        st.write_chunk(st.pa, st.ga);
        st.issue(st.pa, st.ga, half_sel ? cur_b : cur_a);
        take(va);
        vload_group(g + 2, va);
#pragma unroll
        for (int k = 0; k < 4; k++) step(g * 4 + k, wk[k]);
        st.write_chunk(st.pb, st.gb);
        st.issue(st.pb, st.gb, half_sel ? cur_b : cur_a);
        take(vb);
        vload_group(g + 3, vb);
#pragma unroll
        for (int k = 0; k < 4; k++) step(g * 4 + 4 + k, wk[k]);
    }
    if (g < groups) {   // group g on A (full or partial)
        st.write_chunk(st.pa, st.ga);
        st.issue(st.pa, st.ga, half_sel ? cur_b : cur_a);
        take(va);
        const uint32_t n = n_iter - g * 4 < 4 ? n_iter - g * 4 : 4;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (uint32_t(k) < n) step(g * 4 + k, wk[k]);
        g++;
    }
    if (g < groups) {   // group g on B (partial)
        st.write_chunk(st.pb, st.gb);
        take(vb);
        const uint32_t n = n_iter - g * 4;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (uint32_t(k) < n) step(g * 4 + k, wk[k]);
    }

    if constexpr (LAYOUT == 0) {
        // one row per half: sum the 32 coders
#pragma unroll
        for (int v = 0; v < NV; v++) {
            int32_t a = acc[v];
            a += __shfl_xor(a, 16, 64);
            a += __shfl_xor(a, 8, 64);
            a += __shfl_xor(a, 4, 64);
            a += __shfl_xor(a, 2, 64);
            a += __shfl_xor(a, 1, 64);
            acc[v] = a;
        }
        if (lane == 0 && valid) {
#pragma unroll
            for (int v = 0; v < NV && v < int(J.n_vec); v++) {
                if (acc32) acc32[size_t(v) * out_dim + warp] = acc[v];
                if (out8) {
                    float y = rintf(float(acc[v]) * scale);
                    y = y > 127.0f ? 127.0f : (y < -128.0f ? -128.0f : y);
                    out8[size_t(v) * out_dim + warp] = static_cast<int8_t>(y);
                }
            }
        }
    } else {
        const uint32_t row = warp * 32u + lane;
        if (valid) {
#pragma unroll
            for (int v = 0; v < NV && v < int(J.n_vec); v++) {
                if (acc32) acc32[size_t(v) * out_dim + row] = acc[v];
                if (out8) {
                    float y = rintf(float(acc[v]) * scale);
                    y = y > 127.0f ? 127.0f : (y < -128.0f ? -128.0f : y);
                    out8[size_t(v) * out_dim + row] = static_cast<int8_t>(y);
                }
            }
        }
    }
}

}  // namespace blk
}  // namespace mv
