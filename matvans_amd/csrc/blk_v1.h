// Block-level bodies of the MVX1 fused decode-matvec kernels (one block = 4 waves over
// a group of rows), shared by the one-matrix kernels (decode_v1.hip) and the persistent
// batched kernel (batch.hip).  The batched kernel calls them once per tile it takes from
// the work queue, with bid = the tile's block index within its matrix.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "v1_stream.h"

namespace mv {
namespace blk {

using v1::Lane;

constexpr int kWaves = 4;

__device__ __forceinline__ uint32_t dword_of(const uint4 &v, int d) {
    return d == 0 ? v.x : d == 1 ? v.y : d == 2 ? v.z : v.w;
}

// Exclusive scans over the 64 lanes of a wave (lane 0 gets 0).
__device__ __forceinline__ uint32_t wave_excl_sum(uint32_t v, uint32_t lane) {
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(inc, d, 64);
        inc += lane >= uint32_t(d) ? t : 0u;
    }
    return inc - v;
}
__device__ __forceinline__ uint32_t wave_excl_max(uint32_t v, uint32_t lane) {
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(inc, d, 64);
        inc = lane >= uint32_t(d) && t > inc ? t : inc;
    }
    uint32_t prev = __shfl_up(inc, 1, 64);
    return lane ? prev : 0u;
}

// One MVX1 matrix with its vectors and outputs (device pointers).  n_vec <= the kernel's
// NV: vectors past n_vec are neither read nor written.  dst and out8 may each be null.
struct V1Job {
    int32_t *dst;            // [n_vec][rows] int32 sums
    const int8_t *vecs;      // [n_vec][vec_stride]
    uint32_t vec_stride, n_vec;
    uint32_t rows, cols;
    const uint32_t *cursors;
    int32_t min_value;
    uint32_t G;
    const uint32_t *glut;    // A / B tables (2 * 2^P dwords), table mode 0
    const uint8_t *freq;     // per-row frequencies, table mode 1
    const uint16_t *payload;
    uint32_t payload_size;
    int8_t *out8;            // [n_vec][rows] clamp(rint(float(sum) * scale), -128, 127)
    float scale;
    // outlier split (quad coder; null: none): entries coded as 0 in the streams, listed
    // by row as col | (uint8)value << 24 at out_rec[out_start[row] ..]
    const uint32_t *out_start;
    const uint32_t *out_rec;
};

// Outlier records of one row, S lanes per row: lane sl takes records sl, sl + S, ...
// kSlots of them are loaded ahead of the decode (their offsets with the row cursors, the
// records with the first ring sets, the vector bytes just before the loop, so that no
// wait of the loop covers them); the rest, rare with the band the encoder picks, are read
// after the loop.  The corrections value * x[col] enter the lane's accumulators before the
// row reduction.
template <int NV, int S>
struct RowOutliers {
    static constexpr int kSlots = 2;
    uint32_t beg, end;
    uint32_t rec[kSlots];
    int32_t x[kSlots][NV];
    __device__ __forceinline__ void offsets(const V1Job &J, uint32_t row, bool valid) {
        beg = end = 0;
        if (J.out_start && valid) {
            beg = J.out_start[row];
            end = J.out_start[row + 1];
        }
    }
    __device__ __forceinline__ void set_offsets(uint32_t b, uint32_t e) {
        beg = b;
        end = e;
    }
    __device__ __forceinline__ void records(const V1Job &J, uint32_t sl) {
#pragma unroll
        for (int k = 0; k < kSlots; k++) {
            const uint32_t o = beg + sl + uint32_t(k) * S;
            rec[k] = o < end ? J.out_rec[o] : 0u;   // 0: column 0, value 0 (adds nothing)
        }
    }
    __device__ __forceinline__ void values(const V1Job &J, uint32_t nv) {
#pragma unroll
        for (int k = 0; k < kSlots; k++)
#pragma unroll
            for (int v = 0; v < NV; v++)
                x[k][v] = J.out_start ? int32_t(J.vecs[size_t(v < int(nv) ? v : 0) * J.vec_stride +
                                                      (rec[k] & 0xFFFFFFu)])
                                      : 0;
    }
    __device__ __forceinline__ void apply(const V1Job &J, uint32_t sl, uint32_t nv,
                                          int32_t (&acc)[NV]) const {
#pragma unroll
        for (int k = 0; k < kSlots; k++) {
            const int32_t val = int32_t(int8_t(rec[k] >> 24));
#pragma unroll
            for (int v = 0; v < NV; v++) acc[v] += val * x[k][v];
        }
        for (uint32_t o = beg + sl + uint32_t(kSlots) * S; o < end; o += S) {
            const uint32_t r = J.out_rec[o];
            const int32_t val = int32_t(int8_t(r >> 24));
#pragma unroll
            for (int v = 0; v < NV; v++)
                acc[v] += val * int32_t(J.vecs[size_t(v < int(nv) ? v : 0) * J.vec_stride +
                                               (r & 0xFFFFFFu)]);
        }
    }
};

// Output of row `row`, vector v: the int32 sum and / or the next layer's int8 input with
// the epilogue of the reference's family-B kernel (ans-warp-rows-kernel.cu:212-220).
__device__ __forceinline__ void put_row(const V1Job &J, int v, uint32_t row, int32_t a) {
    if (J.dst) J.dst[size_t(v) * J.rows + row] = a;
    if (J.out8) {
        float y = rintf(float(a) * J.scale);
        y = y > 127.0f ? 127.0f : (y < -128.0f ? -128.0f : y);
        J.out8[size_t(v) * J.rows + row] = static_cast<int8_t>(y);
    }
}

// Coder 2 (quad streams in body granules and tails): stream sl of a row, from the row's
// table at the head of the payload (K as two u16, then the tails' u16 end offsets).  The
// table's position depends on the row index only, so it loads with the row's cursor.
struct LaneTable {
    uint32_t kb, e0, e1;   // body chunks; the tail's start and end offsets (u32 words)
};
template <int S>
__device__ __forceinline__ LaneTable lane_table(const uint16_t *__restrict__ payload,
                                                uint32_t row, uint32_t sl) {
    const uint16_t *__restrict__ t = payload + size_t(row) * (S + 2);
    const uint32_t t0 = t[0], t1 = t[1], e1 = t[2 + sl], e0 = t[2 + (sl ? sl - 1u : 0u)];
    return LaneTable{t0 | t1 << 16, sl ? e0 : 0u, e1};
}
struct LaneStream {
    uint32_t body, kb, tail, tail_n;
};
// the stream's body (u32 index), body chunks, tail (u32 index) and tail words, for the row
// at u16 cursor rcur
template <int S>
__device__ __forceinline__ LaneStream lane_stream(const LaneTable &t, uint32_t rcur, bool valid) {
    if (!valid) return LaneStream{0, 0, 0, 0};
    return LaneStream{rcur / 2u, t.kb, rcur / 2u + 4u * t.kb * uint32_t(S) + t.e0,
                      t.e1 >= t.e0 ? t.e1 - t.e0 : 0u};
}

// Everything a lane of a quad-coder block reads before its payload: the wave's window of
// cursors, its row's cursor range, outlier offsets and (coder 2) row table.  The loads
// take clamped indices instead of branches, so they issue back to back and cost one
// memory round trip before the payload loads (behind branches each waited on its own).
template <int S, int RPW, int LAY>
struct RowIndex {
    uint32_t wbeg, wend, rcur, rend, obeg, oend;
    LaneStream ls;
    __device__ __forceinline__ void load(const V1Job &J, uint32_t row0, uint32_t row,
                                         uint32_t sl, bool valid) {
        const uint32_t rows = J.rows, rmax = rows - 1u, ps = J.payload_size;
        const uint32_t *__restrict__ cur = J.cursors;
        const uint32_t a = row0 < rows ? row0 : rmax;
        const uint32_t b = row0 + RPW < rows ? row0 + RPW : rmax;
        const uint32_t r = valid ? row : rmax;
        const uint32_t r1 = row + 1 < rows ? row + 1 : rmax;
        const bool outl = J.out_start != nullptr;
        const uint32_t *__restrict__ os = outl ? J.out_start : cur;   // out_start: rows + 1
        const uint32_t ca = cur[a], cb = cur[b], cr = cur[r], cr1 = cur[r1];
        const uint32_t o0 = os[r], o1 = os[outl ? r + 1 : r1];
        LaneTable lt{0, 0, 0};
        if constexpr (LAY == 1) lt = lane_table<S>(J.payload, r, sl);
        wbeg = row0 < rows ? ca : 0u;
        wend = row0 < rows ? (row0 + RPW < rows ? cb : ps) : 0u;
        if (wend > ps) wend = ps;
        if (wend < wbeg) wend = wbeg;
        wbeg = __builtin_amdgcn_readfirstlane(wbeg);
        wend = __builtin_amdgcn_readfirstlane(wend);
        rcur = valid ? cr : 0u;
        rend = valid ? (row + 1 < rows ? cr1 : ps) : 0u;
        obeg = valid && outl ? o0 : 0u;
        oend = valid && outl ? o1 : 0u;
        ls = LaneStream{0, 0, 0, 0};
        if constexpr (LAY == 1) ls = lane_stream<S>(lt, rcur, valid);
        // Consumed here: the compiler otherwise moves the loads that only lanes with a row
        // use into the caller's branch for those lanes, behind the wait for the others (a
        // second round trip before the payload).
        asm volatile("" ::"v"(rcur), "v"(rend), "v"(obeg), "v"(oend));
        if constexpr (LAY == 1) asm volatile("" ::"v"(lt.kb), "v"(lt.e0), "v"(lt.e1));
    }
};

// Rows a block covers for S streams per row.
template <int S>
constexpr uint32_t rows_per_block() { return S > 64 ? kWaves / (S / 64) : kWaves * (64 / S); }

template <int NV, int S, int P, int MODE>
struct V1Smem {
    static constexpr int R = P == 8 ? 32 : 64;
    static constexpr int WPR = S > 64 ? S / 64 : 1;
    uint16_t ring[kWaves * R * 64];
    uint32_t lut[MODE == 1 ? 1 : 2 << P];
    uint32_t rowtab[MODE == 1 ? kWaves * 384 : 1];   // per wave: f|bias|sym + scratch
    uint16_t dummy[(P == 8 ? 8 : 16) * 64];   // a ring set: M words per lane
    int32_t part[WPR > 1 ? kWaves * NV : 1];
};

template <int NV, int S, int MODE = 0, int SETW = 4>
struct QuadSmem {
    static constexpr int WPR = S > 64 ? S / 64 : 1;
    uint32_t ring[kWaves * 4 * SETW * 64];
    uint2 tab[MODE == 0 ? 256 : 1];                  // two-word table (f | sym << 24, r)
    uint32_t rowtab[MODE == 1 ? kWaves * 448 : 1];   // per wave: packed table + scratch
    uint32_t dummy[SETW * 64];
    int32_t part[WPR > 1 ? kWaves * NV : 1];
};

// One 8-bit table per row (table mode 1), built by the row's wave (S = 64: 64 lanes, four
// quantiles each): an exclusive scan of the row's frequencies gives cdf, every present
// symbol marks its first quantile, and an inclusive max-scan over the quantiles gives
// ppf.  scratch: 768 bytes of the wave's LDS.  put(q, f, q - cdf[s], (int8)(min + s)) is
// called for each of the lane's four quantiles.
template <class Put>
__device__ __forceinline__ void build_row_table(const uint8_t *__restrict__ freq, uint32_t row,
                                                uint32_t G, bool valid, int min_value,
                                                uint32_t lane, uint8_t *scratch, Put put) {
    uint8_t *mark = scratch, *sf = scratch + 256, *sc = scratch + 512;
    uint32_t f[4], c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint32_t s = lane * 4 + k;
        f[k] = valid && s < G ? freq[size_t(row) * G + s] : 0u;
    }
    c[0] = 0; c[1] = f[0]; c[2] = f[0] + f[1]; c[3] = c[2] + f[2];
    uint32_t e = wave_excl_sum(c[3] + f[3], lane);
    uint32_t fw = 0, cw = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        c[k] = (c[k] + e) & 0xFFu;
        fw |= f[k] << (8 * k);
        cw |= c[k] << (8 * k);
    }
    reinterpret_cast<uint32_t *>(mark)[lane] = 0;
    reinterpret_cast<uint32_t *>(sf)[lane] = fw;
    reinterpret_cast<uint32_t *>(sc)[lane] = cw;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (f[k]) mark[c[k]] = static_cast<uint8_t>(lane * 4 + k);
    __syncthreads();
    uint32_t m = reinterpret_cast<const uint32_t *>(mark)[lane];
    uint32_t a[4];
    a[0] = m & 0xFFu;
#pragma unroll
    for (int k = 1; k < 4; k++) {
        uint32_t b = (m >> (8 * k)) & 0xFFu;
        a[k] = b > a[k - 1] ? b : a[k - 1];
    }
    uint32_t pm = wave_excl_max(a[3], lane);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint32_t q = lane * 4 + k;
        uint32_t s = a[k] > pm ? a[k] : pm;
        put(q, uint32_t(sf[s]), (q - sc[s]) & 0xFFu,
            uint32_t(static_cast<uint8_t>(min_value + int(s))));
    }
}

// Single-symbol coder (MVX1 coder 0; the reference's .bin v0 layout is S = 32, P = 8).
template <int NV, int S, int P, int MODE>
__device__ __forceinline__ void v1_block(const V1Job &J, uint32_t bid, V1Smem<NV, S, P, MODE> &sm) {
    constexpr int RPW = S < 64 ? 64 / S : 1;   // rows per wave
    constexpr int WPR = S > 64 ? S / 64 : 1;   // waves per row
    constexpr int GW = 256 / S;                // words per lane per granule
    constexpr int R = P == 8 ? 32 : 64;        // ring words per lane
    constexpr int kLut = MODE == 1 ? 1 : 2 << P;   // dwords of the shared table
    uint16_t *ring = sm.ring;
    uint32_t *lut = sm.lut;
    uint32_t *rowtab = sm.rowtab;
    uint16_t *dummy = sm.dummy;
    int32_t *part = sm.part;
    const uint32_t rows = J.rows, cols = J.cols, G = J.G, vec_stride = J.vec_stride;
    const uint32_t payload_size = J.payload_size;
    const int min_value = J.min_value;
    const uint32_t *__restrict__ cursors = J.cursors;
    const uint32_t *__restrict__ glut = J.glut;
    const uint8_t *__restrict__ freq = J.freq;
    const uint16_t *__restrict__ payload = J.payload;
    const int8_t *__restrict__ vecs = J.vecs;

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6, lane = tid & 63;
    // stream slot within the row, and the rows of this wave
    const uint32_t sl = S >= 64 ? (wave % WPR) * 64 + lane : lane & (S - 1);
    const uint32_t row0 = S >= 64 ? bid * (kWaves / WPR) + wave / WPR
                                  : (bid * kWaves + wave) * RPW;
    const uint32_t row = row0 + (S >= 64 ? 0u : lane / S);
    const bool valid = row < rows;

    if constexpr (MODE == 0 && P == 8) {
        // packed table T[q] = f | bias << 8 | sym << 24 from A = f | sym << 24, B = bias
        // (f <= 255 whenever the alphabet has two or more symbols, as the model ensures)
        lut[tid] = (glut[tid] & 0xFF0000FFu) | ((glut[256 + tid] & 0xFFu) << 8);
    } else if constexpr (MODE == 0) {
        const uint4 *src = reinterpret_cast<const uint4 *>(glut);
        uint4 *d = reinterpret_cast<uint4 *>(lut);
        for (uint32_t i = tid; i < uint32_t(kLut / 4); i += 256) d[i] = src[i];
    } else {
        // one table per row; S == 64, so the wave owns exactly this row: three byte tables
        // f / bias / symbol (one dword per LDS bank each, conflict-free lookups)
        uint8_t *tab = reinterpret_cast<uint8_t *>(rowtab + wave * 384);
        build_row_table(freq, row, G, valid, min_value, lane, tab + 768,
                        [&](uint32_t q, uint32_t f, uint32_t b, uint32_t sy) {
                            tab[q] = static_cast<uint8_t>(f);
                            tab[256 + q] = static_cast<uint8_t>(b);
                            tab[512 + q] = static_cast<uint8_t>(sy);
                        });
    }

    // wave window: rows row0 .. row0 + RPW - 1
    uint32_t wbeg = 0, wend = 0;
    if (row0 < rows) {
        wbeg = cursors[row0];
        uint32_t last = row0 + RPW - 1 < rows ? row0 + RPW - 1 : rows - 1;
        wend = last + 1 < rows ? cursors[last + 1] : payload_size;
        if (wend > payload_size) wend = payload_size;
        if (wend < wbeg) wend = wbeg;
    }
    wbeg = __builtin_amdgcn_readfirstlane(wbeg);
    wend = __builtin_amdgcn_readfirstlane(wend);
    const uint32_t rcur = valid ? cursors[row] : 0u;
    const uint32_t rend = valid ? (row + 1 < rows ? cursors[row + 1] : payload_size) : 0u;

    Lane<GW, P, R> st;
    st.init(payload, wbeg, wend, valid, rcur, rend, sl, ring, dummy, wave, lane);
    st.prologue();

    const uint32_t groups = cols / (16u * S);
    auto vload = [&](uint32_t t, uint4 *c) {
        uint32_t tt = t < groups ? t : groups - 1;
#pragma unroll
        for (int v = 0; v < NV; v++)
            c[v] = *reinterpret_cast<const uint4 *>(vecs +
                                                    size_t(v < int(J.n_vec) ? v : 0) * vec_stride +
                                                    tt * (16u * S) + sl * 16u);
    };
    uint4 vc0[NV], vc1[NV];
    // load order vc0, granules A, granules B, vc1 as in the loop body (A, vc0, B, vc1), so
    // that the loop's wait before group A leaves the two newest loads (B, vc1) in flight
    vload(0, vc0);
    __syncthreads();
    st.start();
    vload(1, vc1);

    int32_t acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = 0;
    const uint8_t *tab = reinterpret_cast<const uint8_t *>(rowtab + wave * 384);

    auto group = [&](const uint4 *vc) {
        uint32_t wbefore = st.waddr;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            uint32_t e0, e1, e2, e3, packed;
            if constexpr (MODE == 1) {
                st.pair8_u8(tab, e0, e1);
                st.pair8_u8(tab, e2, e3);
                packed = __builtin_amdgcn_perm(e1, e0, 0x0c0c0400u) |
                         __builtin_amdgcn_perm(e3, e2, 0x04000c0cu);
            } else {
                if constexpr (P == 8) {
                    st.pair8_packed(lut, e0, e1);
                    st.pair8_packed(lut, e2, e3);
                } else {
                    e0 = st.step12(lut);
                    e1 = st.step12(lut);
                    e2 = st.step12(lut);
                    e3 = st.step12(lut);
                }
                packed = __builtin_amdgcn_perm(e1, e0, 0x0c0c0703u) |
                         __builtin_amdgcn_perm(e3, e2, 0x07030c0cu);
            }
#pragma unroll
            for (int v = 0; v < NV; v++)
                acc[v] = __builtin_amdgcn_sdot4(static_cast<int>(packed),
                                                static_cast<int>(dword_of(vc[v], d)), acc[v],
                                                false);
        }
        st.end_group(wbefore);
    };

    uint32_t t = 0;
    for (; t + 1 < groups; t += 2) {
        st.refill_a();
        group(vc0);
        vload(t + 2, vc0);
        st.refill_b();
        group(vc1);
        vload(t + 3, vc1);
    }
    if (t < groups) {
        st.refill_a();
        group(vc0);
    }

#pragma unroll
    for (int v = 0; v < NV; v++) {
        int32_t a = acc[v];
        if constexpr (S >= 64) a += __shfl_xor(a, 32, 64);
        a += __shfl_xor(a, 16, 64);
        a += __shfl_xor(a, 8, 64);
        a += __shfl_xor(a, 4, 64);
        a += __shfl_xor(a, 2, 64);
        a += __shfl_xor(a, 1, 64);
        acc[v] = a;
    }
    if constexpr (WPR > 1) {
        // the row's waves are consecutive in the block: add their totals in wave order
        if (lane == 0) {
#pragma unroll
            for (int v = 0; v < NV; v++) part[wave * NV + v] = acc[v];
        }
        __syncthreads();
        if (lane == 0 && wave % WPR == 0 && valid) {
#pragma unroll
            for (int v = 0; v < NV; v++) {
                int32_t a = 0;
#pragma unroll
                for (int w = 0; w < WPR; w++) a += part[(wave + w) * NV + v];
                if (v < int(J.n_vec)) put_row(J, v, row, a);
            }
        }
    } else if ((lane & (S - 1)) == 0 && valid) {
#pragma unroll
        for (int v = 0; v < NV; v++)
            if (v < int(J.n_vec)) put_row(J, v, row, acc[v]);
    }
}

// Quad coder (MVX1 coder 1, P = 8, one table per matrix): same mapping, per-lane 64-bit
// state; a 16-symbol group is four quads, each one v_dot4 per vector.  SETW words per
// lane per ring set: 4 (a set per group) or 8 (a set per two groups, twice the ring).
// LAY 1: coder 2, the same streams stored back to back (QuadLane LAY 1).
template <int NV, int S, int MODE = 0, int SETW = 4, int LAY = 0>
__device__ __forceinline__ void quad_block(const V1Job &J, uint32_t bid,
                                           QuadSmem<NV, S, MODE, SETW> &sm) {
    static_assert(MODE == 0 || S == 64, "one table per row needs one row per wave");
    static_assert(LAY == 0 || MODE == 0, "coder 2 has one table per matrix");
    constexpr int RPW = S < 64 ? 64 / S : 1;
    constexpr int WPR = S > 64 ? S / 64 : 1;
    constexpr int G4 = 128 / S;
    constexpr int GPS = SETW / 4;      // groups per set period
    // coder 1's granules are read once, whole: non-temporal loads
    using QL = v1::QuadLane<G4, SETW, LAY, LAY == 0 ? kLoadStream : kLoadCached>;
    uint32_t *ring = sm.ring;
    uint32_t *dummy = sm.dummy;
    int32_t *part = sm.part;
    const uint32_t rows = J.rows, cols = J.cols, vec_stride = J.vec_stride;
    const uint32_t payload_size = J.payload_size;
    const uint32_t *__restrict__ cursors = J.cursors;
    const uint32_t *__restrict__ glut = J.glut;
    const uint16_t *__restrict__ payload = J.payload;
    const int8_t *__restrict__ vecs = J.vecs;

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6, lane = tid & 63;
    const uint32_t sl = S >= 64 ? (wave % WPR) * 64 + lane : lane & (S - 1);
    const uint32_t row0 = S >= 64 ? bid * (kWaves / WPR) + wave / WPR
                                  : (bid * kWaves + wave) * RPW;
    const uint32_t row = row0 + (S >= 64 ? 0u : lane / S);
    const bool valid = row < rows;
    // the decode table's entries and every index, in one round trip; with several vectors
    // the first group's vector loads go first, to overlap it (the index values are consumed
    // where they load; in the batched launch that placement measured 2.5 us faster for the
    // decoder block at four vectors and 2 us slower at one)
    const uint32_t groups = cols / (16u * S);
    // the vectors through a buffer resource: the lane's byte offset is fixed, the group's
    // offset is wave-uniform (scalar), so a load costs no vector address arithmetic
    const uint32_t nv_eff = J.n_vec < uint32_t(NV) ? J.n_vec : uint32_t(NV);
    const __amdgpu_buffer_rsrc_t vrs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int8_t *>(vecs), 0,
        static_cast<int>((nv_eff - 1) * vec_stride + cols), 0x00020000);
    uint32_t voff[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) voff[v] = (v < int(nv_eff) ? uint32_t(v) : 0u) * vec_stride + sl * 16u;
    auto vload = [&](uint32_t t, uint32_t (&c)[NV][4]) {
        const uint32_t tt = t < groups ? t : groups - 1;
#pragma unroll
        for (int v = 0; v < NV; v++) {
            auto x = __builtin_amdgcn_raw_buffer_load_b128(vrs, voff[v], tt * (16u * S), 0);
            c[v][0] = x[0]; c[v][1] = x[1]; c[v][2] = x[2]; c[v][3] = x[3];
        }
    };
    uint32_t vc0[NV][4], vc1[NV][4];
    if constexpr (NV > 1) vload(0, vc0);
    uint32_t ga = 0, gb = 0;
    if constexpr (MODE == 0) {
        ga = glut[tid];
        gb = glut[256 + tid];
    }
    RowIndex<S, RPW, LAY> ix;
    ix.load(J, row0, row, sl, valid);
    // one table per matrix: the two-word entries (A, B) = (f | sym << 24, r) of
    // matvans_v1_lut; one per row: packed T[q] = f | r << 8 | sym << 24 (r = q - cdf[s];
    // f <= 255 and r < f for every table the encoder accepts)
    uint32_t *tab = nullptr;
    if constexpr (MODE == 0) {
        sm.tab[tid] = make_uint2(ga, gb);
    } else {                     // this wave's row's own table
        tab = sm.rowtab + wave * 448;
        build_row_table(J.freq, row, J.G, valid, J.min_value, lane,
                        reinterpret_cast<uint8_t *>(tab + 256),
                        [&](uint32_t q, uint32_t f, uint32_t b, uint32_t sy) {
                            tab[q] = f | (b << 8) | (sy << 24);
                        });
    }
    const uint32_t wbeg = ix.wbeg, wend = ix.wend, rcur = ix.rcur, rend = ix.rend;
    RowOutliers<NV, S> outl;
    outl.set_offsets(ix.obeg, ix.oend);

    QL st;
    if constexpr (LAY == 0) {
        st.init(payload, wbeg, wend, valid, rcur, rend, sl, ring, dummy, wave, lane);
    } else {
        st.init_lanes(payload, wbeg, wend, valid, ix.ls.body, ix.ls.kb, S, sl, ix.ls.tail,
                      ix.ls.tail_n, ring, dummy, wave, lane);
    }
    // load order vc0, granules A, granules B, vc1 as in the loop body (A, vc0, B, vc1), so
    // that the loop's wait before group A leaves the two newest loads (B, vc1) in flight;
    // the barrier (the decode table) waits while the first ring sets are in flight
    st.prologue([&] {
        outl.records(J, sl);
        if constexpr (NV == 1) vload(0, vc0);
        __syncthreads();
    });
    st.start();
    vload(1, vc1);
    outl.values(J, nv_eff);

    int32_t acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = 0;
    auto group = [&](const uint32_t (&vc)[NV][4]) {
#pragma unroll
        for (int d = 0; d < 4; d++) {
            uint32_t packed;
            if constexpr (MODE == 0) packed = st.quad(sm.tab);
            else packed = st.quad(tab);
#pragma unroll
            for (int v = 0; v < NV; v++)
                acc[v] = __builtin_amdgcn_sdot4(static_cast<int>(packed),
                                                static_cast<int>(vc[v][d]), acc[v], false);
        }
    };
    // a set period: GPS groups after one refill; the ring position advance is folded into
    // the consumed-word count once per period (at most 4 * GPS words < R4)
    // (par: the parity of t0, which picks the vector buffer of each group statically)
    auto one = [&](auto par, uint32_t tg) {
        if constexpr (decltype(par)::value == 0) {
            group(vc0);
            vload(tg + 2, vc0);
        } else {
            group(vc1);
            vload(tg + 2, vc1);
        }
    };
    static_assert(GPS == 1 || GPS == 2 || GPS == 4, "set period of one, two or four groups");
    auto period = [&](auto par, uint32_t t0, uint32_t ng) {
        constexpr int P = decltype(par)::value;
        uint32_t wbefore = st.waddr;
        one(std::integral_constant<int, P % 2>{}, t0);
        if constexpr (GPS >= 2) {
            if (ng > 1) one(std::integral_constant<int, (P + 1) % 2>{}, t0 + 1);
        }
        if constexpr (GPS == 4) {
            if (ng > 2) one(std::integral_constant<int, P % 2>{}, t0 + 2);
            if (ng > 3) one(std::integral_constant<int, (P + 1) % 2>{}, t0 + 3);
        }
        st.end_group(wbefore);
    };
    using Even = std::integral_constant<int, 0>;
    using Second = std::integral_constant<int, GPS % 2>;
    uint32_t t = 0;
    for (; t + 2 * GPS <= groups; t += 2 * GPS) {
        st.refill_a();
        period(Even{}, t, GPS);
        st.refill_b();
        period(Second{}, t + GPS, GPS);
    }
    if (t < groups) {
        st.refill_a();
        period(Even{}, t, groups - t < uint32_t(GPS) ? groups - t : uint32_t(GPS));
        if (t + GPS < groups) {
            st.refill_b();
            period(Second{}, t + GPS, groups - t - GPS);
        }
    }

    outl.apply(J, sl, nv_eff, acc);
#pragma unroll
    for (int v = 0; v < NV; v++) {
        int32_t a = acc[v];
        if constexpr (S >= 64) a += __shfl_xor(a, 32, 64);
        a += __shfl_xor(a, 16, 64);
        a += __shfl_xor(a, 8, 64);
        a += __shfl_xor(a, 4, 64);
        a += __shfl_xor(a, 2, 64);
        a += __shfl_xor(a, 1, 64);
        acc[v] = a;
    }
    if constexpr (WPR > 1) {
        if (lane == 0) {
#pragma unroll
            for (int v = 0; v < NV; v++) part[wave * NV + v] = acc[v];
        }
        __syncthreads();
        if (lane == 0 && wave % WPR == 0 && valid) {
#pragma unroll
            for (int v = 0; v < NV; v++) {
                int32_t a = 0;
#pragma unroll
                for (int w = 0; w < WPR; w++) a += part[(wave + w) * NV + v];
                if (v < int(J.n_vec)) put_row(J, v, row, a);
            }
        }
    } else if ((lane & (S - 1)) == 0 && valid) {
#pragma unroll
        for (int v = 0; v < NV; v++)
            if (v < int(J.n_vec)) put_row(J, v, row, acc[v]);
    }
}


// ---------------------------------------------------------------------------------------
// Quad coder, 32 streams per row, several vectors: the products on the matrix cores.
// A block of MW waves (4 or 8) decodes 2 MW rows (two per wave, one per half-wave, as
// quad_block).  At each 16-symbol group the block's rows x 512 columns of decoded int8
// are the A operand of v_mfma_i32_16x16x64_i8 (16 rows; with 4 waves rows 8..15 stay
// zero): every lane writes its four packed quads (16 columns) to an LDS stage, and after
// one block barrier wave w multiplies stage columns 512 w / MW .. + 512 / MW - 1, 64 at a
// time (A: lane l holds row l % 16, columns 16 (l / 16) .. + 15 of the slice) by the
// vectors' same columns (B: lane l holds vector l % 16; vectors past n_vec read as
// zeros).  Each wave accumulates its slices over all groups in one 16 x 16 int32 tile
// (lane l: rows 4 (l / 16) .. + 3 of vector l % 16); the tiles are added through LDS at
// the end.  The stage is double-buffered, so one barrier per group suffices: a wave writes
// stage (t + 1) % 2 only after the barrier of group t, which every wave passes after
// reading stage (t - 1) % 2.  Integer sums: bit-exact whatever the order.  Per group this
// replaces NV v_dot4 per quad and NV vector loads with 8 / MW MFMAs and as many LDS reads
// and vector loads, one LDS write and the barrier.
constexpr uint32_t kStageRow = 528;   // 512 columns + 16 bytes: rows in distinct banks

template <int MW>
struct QuadMfmaSmem {
    static constexpr int kRows = 2 * MW;
    uint32_t ring[MW * 16 * 64];
    uint2 tab[256];                          // two-word table (f | sym << 24, r)
    uint32_t dummy[4 * 64];
    uint32_t stage[2][16 * kStageRow / 4];   // 16 operand rows; also the tiles' sum at the end
    int32_t outl[kRows][8];                  // outlier corrections per row and vector
};

typedef int mv_i32x4 __attribute__((ext_vector_type(4)));

template <int NV, int MW, int LAY = 0>
__device__ __forceinline__ void quad_block_mfma(const V1Job &J, uint32_t bid,
                                                QuadMfmaSmem<MW> &sm) {
    static_assert(MW == 4 || MW == 8, "4 or 8 waves");
    constexpr int S = 32;
    constexpr uint32_t kRows = 2 * MW;
    constexpr int KS = 8 / MW;             // 64-column slices per wave and group
    using QL = v1::QuadLane<4, 4, LAY, LAY == 0 ? kLoadStream : kLoadCached>;
    const uint32_t rows = J.rows, cols = J.cols, vec_stride = J.vec_stride;
    const uint32_t payload_size = J.payload_size;
    const uint32_t *__restrict__ cursors = J.cursors;
    const uint32_t *__restrict__ glut = J.glut;
    const uint16_t *__restrict__ payload = J.payload;
    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6, lane = tid & 63;
    const uint32_t sl = lane & 31;
    const uint32_t rloc = 2 * wave + (lane >> 5);          // row within the block
    const uint32_t row0 = bid * kRows + 2 * wave;
    const uint32_t row = bid * kRows + rloc;
    const bool valid = row < rows;
    uint2 *tab = sm.tab;
    // the decode table's entries and every index, in one round trip
    const uint32_t ga = glut[tid & 255u], gb = glut[256u + (tid & 255u)];
    RowIndex<S, 2, LAY> ix;
    ix.load(J, row0, row, sl, valid);
    if (tid < 256) tab[tid] = make_uint2(ga, gb);
    if constexpr (MW == 4) {   // operand rows 8..15 of both stages stay zero
        uint32_t *z0 = sm.stage[0] + 8 * kStageRow / 4, *z1 = sm.stage[1] + 8 * kStageRow / 4;
        for (uint32_t k = tid; k < 8 * kStageRow / 4; k += 64 * MW) z0[k] = z1[k] = 0;
    }
    const uint32_t wbeg = ix.wbeg, wend = ix.wend, rcur = ix.rcur, rend = ix.rend;
    RowOutliers<NV, S> outl;
    outl.set_offsets(ix.obeg, ix.oend);

    QL st;
    if constexpr (LAY == 0) {
        st.init(payload, wbeg, wend, valid, rcur, rend, sl, sm.ring, sm.dummy, wave, lane);
    } else {
        st.init_lanes(payload, wbeg, wend, valid, ix.ls.body, ix.ls.kb, S, sl, ix.ls.tail,
                      ix.ls.tail_n, sm.ring, sm.dummy, wave, lane);
    }
    st.prologue();
    outl.records(J, sl);

    const uint32_t groups = cols / (16u * S);
    const uint32_t nv_eff = J.n_vec < uint32_t(NV) ? J.n_vec : uint32_t(NV);
    const __amdgpu_buffer_rsrc_t vrs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int8_t *>(J.vecs), 0,
        static_cast<int>((nv_eff - 1) * vec_stride + cols), 0x00020000);
    const uint32_t vj = lane & 15;
    const uint32_t col_base = (512u / MW) * wave + 16u * (lane >> 4);
    const uint32_t boff = vj < nv_eff ? vj * vec_stride + col_base : v1::kOOB;
    auto bload = [&](uint32_t t, mv_i32x4 (&b)[KS]) {
        const uint32_t tt = t < groups ? t : groups - 1;
#pragma unroll
        for (int k = 0; k < KS; k++) {
            auto x = __builtin_amdgcn_raw_buffer_load_b128(vrs, boff + 64u * k, tt * 512u, 0);
            b[k][0] = int(x[0]); b[k][1] = int(x[1]); b[k][2] = int(x[2]); b[k][3] = int(x[3]);
        }
    };
    uint8_t *stage0 = reinterpret_cast<uint8_t *>(sm.stage[0]);
    uint8_t *stage1 = reinterpret_cast<uint8_t *>(sm.stage[1]);
    const uint32_t wr_off = rloc * kStageRow + 16u * sl;
    const uint32_t rd_off = (lane & 15) * kStageRow + col_base;

    mv_i32x4 b0[KS], b1[KS];
    bload(0, b0);
    __syncthreads();
    st.start();
    bload(1, b1);
    outl.values(J, nv_eff);

    mv_i32x4 d = {0, 0, 0, 0};
    auto group = [&](uint8_t *stg, mv_i32x4 (&b)[KS], uint32_t t) {
        uint32_t P[4];
#pragma unroll
        for (int q = 0; q < 4; q++) P[q] = st.quad(tab);
        *reinterpret_cast<uint4 *>(stg + wr_off) = make_uint4(P[0], P[1], P[2], P[3]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KS; k++) {
            const uint4 a4 = *reinterpret_cast<const uint4 *>(stg + rd_off + 64u * k);
            mv_i32x4 a;
            a[0] = int(a4.x); a[1] = int(a4.y); a[2] = int(a4.z); a[3] = int(a4.w);
            d = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[k], d, 0, 0, 0);
        }
        bload(t + 2, b);
    };
    uint32_t t = 0;
    for (; t + 1 < groups; t += 2) {
        uint32_t wb = st.waddr;
        st.refill_a();
        group(stage0, b0, t);
        st.end_group(wb);
        wb = st.waddr;
        st.refill_b();
        group(stage1, b1, t + 1);
        st.end_group(wb);
    }
    if (t < groups) {
        uint32_t wb = st.waddr;
        st.refill_a();
        group(stage0, b0, t);
        st.end_group(wb);
    }

    // outlier corrections of this lane's row, summed over the row's 32 lanes
    int32_t oacc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) oacc[v] = 0;
    outl.apply(J, sl, nv_eff, oacc);
#pragma unroll
    for (int v = 0; v < NV; v++) {
        int32_t a = oacc[v];
        a += __shfl_xor(a, 16, 64);
        a += __shfl_xor(a, 8, 64);
        a += __shfl_xor(a, 4, 64);
        a += __shfl_xor(a, 2, 64);
        a += __shfl_xor(a, 1, 64);
        oacc[v] = a;
    }
    __syncthreads();   // every wave is done with the stages
    int32_t *tiles = reinterpret_cast<int32_t *>(sm.stage);   // [wave][16 rows][16 vectors]
#pragma unroll
    for (int r = 0; r < 4; r++) tiles[wave * 256 + (4 * (lane >> 4) + r) * 16 + (lane & 15)] = d[r];
    if (sl == 0) {
#pragma unroll
        for (int v = 0; v < NV; v++) sm.outl[rloc][v] = oacc[v];
    }
    __syncthreads();
    if (tid < kRows * uint32_t(NV)) {
        const uint32_t i = tid / NV, v = tid % NV;
        const uint32_t r = bid * kRows + i;
        int32_t sum = sm.outl[i][v];
#pragma unroll
        for (int w = 0; w < MW; w++) sum += tiles[w * 256 + i * 16 + v];
        if (r < rows && v < nv_eff) put_row(J, int(v), r, sum);
    }
}

}  // namespace blk
}  // namespace mv
