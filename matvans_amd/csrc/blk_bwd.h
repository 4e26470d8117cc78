// Block-level bodies of the backward dx = alpha * W^T (s_w * g) from the compressed
// streams, shared by the one-matrix kernels (backward.hip) and the persistent batched
// backward (batch.hip).  The reference has no backward pass.
//
// Mapping (stream-major): a unit is 8 waves = 8 consecutive streams of 64 consecutive
// rows; a wave decodes one stream of the 64 rows, one row per lane.  A stream covers the
// same columns in every row, so at each decode step all 64 lanes contribute to the same
// columns.  Every four columns (one decode quad) a fixed pair step (bwd_pair) folds the
// 64 rows to 16 lanes per column: the quad's symbols as two column pairs (two byte
// permutes of its table entries), one v_permlane32_swap, a multiply and an fma per pair
// of rows, one v_permlane16_swap and an add.  After a group's four quads, bwd_group_reduce sums each quad's 16 lanes for all
// four quads at once (a transposed reduction over DPP rows: row_mirror, row_half_mirror
// and two quad_perm adds), in the same order on every run.  The group's 16 totals are
// stored once, to partial[v][row_block][c]; reduce_tile sums the row blocks in a fixed
// order and applies alpha.  Both stages are deterministic bit for bit, so one-matrix and
// batched launches give identical results.
//
// With two or more gradient vectors the group sums run on the matrix cores instead
// (bwd_group_mfma), one MFMA sequence for all vectors.
//
// .bin v0 streams (create_matrices_WarpPerRow.py:361-366): stream l covers columns
// 512*(k/16) + 16*l + k%16.  Quad-coder streams (S = 32 or 64): stream l's group t covers
// columns 16*S*t + 16*l .. +15 and a quad step yields four consecutive ones.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "alu.h"
#include "blk_v1.h"
#include "v0_stream.h"
#include "v1_stream.h"

namespace mv {
namespace blk {

constexpr int kBwdWaves = 8;
constexpr int kRedParts = 16;

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f(float v) {
    int r = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, true);
    return __int_as_float(r);
}

// One matrix of a backward (device pointers).
struct BwdJob {
    float *partial;              // workspace [n_vec][nrb][cols], nrb = ceil(rows / 64)
    float *dx;                   // [n_vec][cols]
    const float *g;              // [n_vec][g_stride]
    const float *s_w;            // [rows]
    const float *alpha;          // [n_vec]
    uint32_t g_stride, n_vec, rows, cols;
    const uint32_t *cursors;
    const uint16_t *payload;
    uint32_t payload_size;
    int32_t min_value;           // .bin v0
    uint32_t G;
    const uint8_t *cdf, *ppf;
    const uint32_t *glut;        // quad: A / B decode tables (2 x 256 dwords)
    const uint8_t *freq;         // quad, one table per row: 8-bit frequencies [rows][G]
    // quad outlier split (null: none): entries coded as 0 in the streams, by column as
    // row | (uint8)value << 24 at out_col_rec[out_col_start[c] ..], rows ascending
    const uint32_t *out_col_start;
    const uint32_t *out_col_rec;
};

// From this many gradient vectors on, a group's sums are computed on the matrix cores
// (bwd_group_mfma) instead of pair steps and row reductions per vector.
constexpr int kMfmaNV = 2;

template <int NV>
struct V0BwdSmem {
    uint16_t ring[kBwdWaves * kRingWords * 64];
    uint32_t lut[256];
    uint16_t dummy[kDummyBytes / 2];
    uint32_t wstage[NV >= kMfmaNV ? kBwdWaves * 256 : 1];   // per wave: 64 x 16 int8
};

template <int NV, int S, int MODE = 0>
struct QuadBwdSmem {
    uint32_t ring[kBwdWaves * v1::QuadLane<128 / S>::R4 * 64];
    uint32_t tab[MODE == 0 ? 256 : 1];
    uint32_t rowtab[MODE == 1 ? 64 * 256 : 1];            // one packed table per row
    uint32_t rowscratch[MODE == 1 ? kBwdWaves * 192 : 1];  // 768 bytes per wave
    uint32_t dummy[v1::QuadLane<128 / S>::M * 64];
    uint32_t wstage[NV >= kMfmaNV ? kBwdWaves * 256 : 1];
};

// Per-lane setup shared by both codecs: the 64-row window, this lane's row extent and
// its scaled gradients (zero past the last row and past n_vec).
template <int NV>
struct BwdRows {
    uint32_t row, wbeg, wend, rcur, rend;
    bool valid;
    float gs[NV];
    __device__ __forceinline__ void init(const BwdJob &J, uint32_t rb, uint32_t lane) {
        const uint32_t row0 = rb * 64u;
        row = row0 + lane;
        valid = row < J.rows;
        const uint32_t last = row0 + 63u < J.rows ? row0 + 63u : J.rows - 1u;
        wbeg = J.cursors[row0];
        wend = last + 1 < J.rows ? J.cursors[last + 1] : J.payload_size;
        if (wend > J.payload_size) wend = J.payload_size;
        if (wend < wbeg) wend = wbeg;
        wbeg = __builtin_amdgcn_readfirstlane(wbeg);
        wend = __builtin_amdgcn_readfirstlane(wend);
        rcur = valid ? J.cursors[row] : 0u;
        rend = valid ? (row + 1 < J.rows ? J.cursors[row + 1] : J.payload_size) : 0u;
#pragma unroll
        for (int v = 0; v < NV; v++)
            gs[v] = valid && v < int(J.n_vec) ? J.s_w[row] * J.g[size_t(v) * J.g_stride + row]
                                              : 0.0f;
    }
};

// The symbols of a decode quad's four table entries e[0..3] (symbol in byte 3, column
// order) as the two inputs of the pair step: v1 = columns 0 and 2 in bytes 0 and 2,
// v2 = columns 1 and 3 in bytes 0 and 2.
__device__ __forceinline__ void quad_cols(const uint32_t (&e)[4], uint32_t &v1, uint32_t &v2) {
    v1 = __builtin_amdgcn_perm(e[2], e[0], 0x0c070c03u);
    v2 = __builtin_amdgcn_perm(e[3], e[1], 0x0c070c03u);
}

// Pair step of one decode quad of this lane's row (v1, v2 from quad_cols), gsA / gsB the
// scaled gradients of rows (lane & 31) and (lane & 31) + 32.  Swapping v1 against v2
// across the wave halves leaves, in bytes 0 and 2, columns (0, 2) in lanes 0..31 and
// (1, 3) in lanes 32..63, of row (lane & 31) in X and of row (lane & 31) + 32 in Y.  a, b
// are those two columns' pair sums; v_permlane16_swap + add leaves the 16-lane rows
// k = 0..3 of the wave with columns (0, 2, 1, 3), each lane holding the sum over four rows.
template <int NV>
__device__ __forceinline__ void bwd_pair(uint32_t v1, uint32_t v2, const float (&gsA)[NV],
                                         const float (&gsB)[NV], float (&s)[NV]) {
    auto sw = __builtin_amdgcn_permlane32_swap(v1, v2, false, false);
    const uint32_t X = sw[0], Y = sw[1];
    const float xa = cvt_f32_sbyte0(X), xb = cvt_f32_sbyte2(X);
    const float ya = cvt_f32_sbyte0(Y), yb = cvt_f32_sbyte2(Y);
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const float a = fmaf(ya, gsB[v], xa * gsA[v]);
        const float b = fmaf(yb, gsB[v], xb * gsA[v]);
        auto t = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false,
                                                  false);
        s[v] = __uint_as_float(t[0]) + __uint_as_float(t[1]);
    }
}

// Sums s[d] over the 16 lanes of each DPP row for the four quads d of a group at once:
// lanes i < 8 of a row keep quads 0, 1 and hand quads 2, 3 to lane 15 - i (row_mirror),
// then lanes with bit 2 clear keep the lower of their two quads and hand the other to
// lane i ^ 7 (row_half_mirror), then two quad_perm adds finish the row.  Lane i of row k
// ends with the total of quad i >> 2, column 4 (i >> 2) + (0, 2, 1, 3)[k].
template <int NV>
__device__ __forceinline__ void bwd_group_reduce(const float (&s)[4][NV], uint32_t lane,
                                                 float (&tot)[NV]) {
    const bool hi8 = (lane & 8u) != 0, hi4 = (lane & 4u) != 0;
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const float u0 = (hi8 ? s[2][v] : s[0][v]) + dpp_f<0x140, 0xF>(hi8 ? s[0][v] : s[2][v]);
        const float u1 = (hi8 ? s[3][v] : s[1][v]) + dpp_f<0x140, 0xF>(hi8 ? s[1][v] : s[3][v]);
        float w = (hi4 ? u1 : u0) + dpp_f<0x141, 0xF>(hi4 ? u0 : u1);
        w += dpp_f<0xB1, 0xF>(w);    // quad_perm [1,0,3,2]
        w += dpp_f<0x4E, 0xF>(w);    // quad_perm [2,3,0,1]
        tot[v] = w;
    }
}

// Per-lane constants of the pair step and the group store.
template <int NV>
struct BwdLane {
    float gsA[NV], gsB[NV];
    uint32_t col;     // column of this lane's group total within the 16-column group
    bool writer;      // lanes 4q of each 16-lane row hold quad q's total
    __device__ __forceinline__ void init(const float (&gs)[NV], uint32_t lane) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
            auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(gs[v]),
                                                       __float_as_uint(gs[v]), false, false);
            gsA[v] = __uint_as_float(sw[0]);
            gsB[v] = __uint_as_float(sw[1]);
        }
        const uint32_t i = lane & 15, k = lane >> 4;   // rows of 16 lanes: columns 0, 2, 1, 3
        col = 4 * (i >> 2) + ((k & 1) * 2 + (k >> 1));
        writer = (i & 3) == 0;
    }
    __device__ __forceinline__ void store(const float (&tot)[NV], float *out, size_t plane,
                                          uint32_t n_vec) const {
        if (writer) {
#pragma unroll
            for (int v = 0; v < NV; v++)
                if (v < int(n_vec)) out[size_t(v) * plane + col] = tot[v];
        }
    }
    // a group's four quads (next(v1, v2) decodes one, see quad_cols): pair steps, then
    // the transposed row reduction
    template <class Next>
    __device__ __forceinline__ void group(Next next, uint32_t lane, float (&tot)[NV]) const {
        float s[4][NV];
        // quad d's pair step follows quad d+1's decode: independent work between the
        // decode's dependent steps and the swaps' operand hazards
        uint32_t p1, p2;
        next(p1, p2);
#pragma unroll
        for (int d = 1; d < 4; d++) {
            uint32_t c1, c2;
            next(c1, c2);
            bwd_pair<NV>(p1, p2, gsA, gsB, s[d - 1]);
            p1 = c1;
            p2 = c2;
        }
        bwd_pair<NV>(p1, p2, gsA, gsB, s[3]);
        bwd_group_reduce<NV>(s, lane, tot);
    }
};

// Matrix-core path for several gradient vectors.  A group's 16 columns of the 64 rows
// form a 64 x 16 int8 block W_g (lane r holds row r: four packed quads).  The group sums
// out[c][v] = sum_r W_g[r][c] gs[r][v] are W_g^T gs, a 16 x 64 by 64 x 16 product:
// sixteen v_mfma_f32_16x16x4f32 steps over slabs of four rows.  In slab s, lane l supplies
// A[l%16][l/16] = W_g[4s + l/16][l%16] (read back from the wave's LDS stage) and
// B[l/16][l%16] = gs[4s + l/16][v = l%16] (MfmaGrad); the result D[i][j] = out[i][j] is
// held by lane 16 (i/4) + j, register i%4, so each lane stores four consecutive columns of
// one vector.  The hardware's summation order is fixed: deterministic like the pair steps.
struct MfmaGrad {
    float b[16];   // slab s: s_w[r] g[l%16][r], r = 4s + l/16 (0 past the rows / n_vec)
    __device__ __forceinline__ void init(const BwdJob &J, uint32_t rb, uint32_t lane) {
        const uint32_t v = lane & 15;
#pragma unroll
        for (int s = 0; s < 16; s++) {
            const uint32_t r = rb * 64u + 4u * s + (lane >> 4);
            b[s] = r < J.rows && v < J.n_vec ? J.s_w[r] * J.g[size_t(v) * J.g_stride + r] : 0.0f;
        }
    }
};

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void bwd_group_mfma(const uint32_t (&P)[4], const float (&b)[16],
                                               uint32_t lane, uint32_t *wst, float *out,
                                               size_t plane, uint32_t n_vec) {
    // stage row r = lane at bytes 16r .. 16r + 15 (the wave's LDS accesses run in order, so
    // the reads below see this write and the next group's write follows these reads)
    *reinterpret_cast<uint4 *>(wst + 4 * lane) = make_uint4(P[0], P[1], P[2], P[3]);
    const int8_t *w8 = reinterpret_cast<const int8_t *>(wst);
    const uint32_t i = lane & 15, k = lane >> 4;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 16; s++) {
        const float a = static_cast<float>(w8[(4 * s + k) * 16 + i]);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[s], acc, 0, 0, 0);
    }
    if (i < n_vec)   // lane l: columns 4 (l/16) .. +3 of vector l%16
        *reinterpret_cast<float4 *>(out + size_t(i) * plane + 4 * k) =
            make_float4(acc[0], acc[1], acc[2], acc[3]);
}

// .bin v0 unit: row block rb, streams 8 sg .. 8 sg + 7.  fresh (block-uniform): build the
// table and synchronise the block first; a unit that follows one of the same matrix in
// the same block reuses the table, and its waves go on without a barrier (each wave's
// ring and stage are its own).
template <int NV>
__device__ __forceinline__ void v0_bwd_unit(const BwdJob &J, uint32_t rb, uint32_t sg,
                                            uint32_t nrb, V0BwdSmem<NV> &sm, bool fresh = true) {
    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6, lane = tid & 63;
    if (fresh && tid < 256) {   // packed table f | bias << 8 | sym << 24
        uint32_t f, sym, bias;
        v0_table_entry(tid, J.min_value, J.G, J.cdf, J.ppf, f, sym, bias);
        sm.lut[tid] = f | (bias << 8) | (sym << 24);
    }
    const uint32_t stream = sg * kBwdWaves + wave;   // 0..31
    BwdRows<NV> R;
    R.init(J, rb, lane);

    V0Lane st;
    st.init(J.payload, R.wbeg, R.wend, R.valid, R.rcur, R.rend, stream, sm.ring, sm.dummy, wave,
            lane);
    st.prologue();
    if (fresh) __syncthreads();
    st.start();

    const uint32_t cols = J.cols;
    const uint32_t groups = cols / 512u;
    const size_t plane = size_t(nrb) * cols;
    float *base = J.partial + size_t(rb) * cols + stream * 16u;
    // four symbols (byte 3 of each table entry) packed in column order
    auto next4 = [&]() {
        uint32_t e[4];
        st.pair_packed(sm.lut, e[0], e[1]);
        st.pair_packed(sm.lut, e[2], e[3]);
        return __builtin_amdgcn_perm(e[1], e[0], 0x0c0c0703u) |
               __builtin_amdgcn_perm(e[3], e[2], 0x07030c0cu);
    };
    BwdLane<NV> L;
    MfmaGrad M;
    if constexpr (NV >= kMfmaNV) M.init(J, rb, lane);
    else L.init(R.gs, lane);
    auto group = [&](uint32_t t) {
        uint32_t wbefore = st.waddr;
        if constexpr (NV >= kMfmaNV) {
            uint32_t P[4];
#pragma unroll
            for (int d = 0; d < 4; d++) P[d] = next4();
            bwd_group_mfma(P, M.b, lane, sm.wstage + wave * 256, base + t * 512u, plane,
                           J.n_vec);
        } else {
            float tot[NV];
            L.group(
                [&](uint32_t &v1, uint32_t &v2) {
                    uint32_t e[4];
                    st.pair_packed(sm.lut, e[0], e[1]);
                    st.pair_packed(sm.lut, e[2], e[3]);
                    quad_cols(e, v1, v2);
                },
                lane, tot);
            L.store(tot, base + t * 512u, plane, J.n_vec);
        }
        st.end_group(wbefore);
    };
    uint32_t t = 0;
    for (; t + 1 < groups; t += 2) {
        st.refill_a();
        group(t);
        st.refill_b();
        group(t + 1);
    }
    if (t < groups) {
        st.refill_a();
        group(t);
    }
}

// Quad-coder unit (S = 32, 64 or 128 streams): row block rb, streams 8 sg .. 8 sg + 7;
// fresh as for v0_bwd_unit.
template <int NV, int S, int MODE = 0>
__device__ __forceinline__ void quad_bwd_unit(const BwdJob &J, uint32_t rb, uint32_t sg,
                                              uint32_t nrb, QuadBwdSmem<NV, S, MODE> &sm,
                                              bool fresh = true) {
    static_assert(MODE == 0 || S == 64, "one table per row: 64 streams");
    using QL = v1::QuadLane<128 / S>;
    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6, lane = tid & 63;
    // packed table T[q] = f | r << 8 | sym << 24: from the A / B tables (A = f | sym << 24,
    // B = r), or one per row of the unit, lane l reading row l's
    const uint32_t *tab;
    if constexpr (MODE == 0) {
        if (fresh && tid < 256)
            sm.tab[tid] = (J.glut[tid] & 0xFF0000FFu) | ((J.glut[256 + tid] & 0xFFu) << 8);
        tab = sm.tab;
    } else {
        __syncthreads();   // every wave of the previous unit is done with the tables
        uint8_t *scratch = reinterpret_cast<uint8_t *>(sm.rowscratch + wave * 192);
        for (uint32_t k = 0; k < 64u / kBwdWaves; k++) {   // the same count on every wave
            const uint32_t r = wave * (64u / kBwdWaves) + k;
            uint32_t *t = sm.rowtab + r * 256u;
            build_row_table(J.freq, rb * 64u + r, J.G, rb * 64u + r < J.rows, J.min_value,
                            lane, scratch,
                            [&](uint32_t q, uint32_t f, uint32_t b, uint32_t sy) {
                                t[q] = f | (b << 8) | (sy << 24);
                            });
        }
        tab = sm.rowtab + lane * 256u;
    }
    const uint32_t stream = sg * kBwdWaves + wave;   // 0..S-1
    BwdRows<NV> R;
    R.init(J, rb, lane);

    QL st;
    st.init(J.payload, R.wbeg, R.wend, R.valid, R.rcur, R.rend, stream, sm.ring, sm.dummy, wave,
            lane);
    st.prologue();
    if (fresh || MODE == 1) __syncthreads();
    st.start();

    const uint32_t cols = J.cols;
    const uint32_t groups = cols / (16u * S);
    const size_t plane = size_t(nrb) * cols;
    float *base = J.partial + size_t(rb) * cols + stream * 16u;
    BwdLane<NV> L;
    MfmaGrad M;
    if constexpr (NV >= kMfmaNV) M.init(J, rb, lane);
    else L.init(R.gs, lane);
    auto group = [&](uint32_t t) {
        uint32_t wbefore = st.waddr;
        if constexpr (NV >= kMfmaNV) {
            uint32_t P[4];
#pragma unroll
            for (int d = 0; d < 4; d++) P[d] = st.quad(tab);
            bwd_group_mfma(P, M.b, lane, sm.wstage + wave * 256, base + t * (16u * S), plane,
                           J.n_vec);
        } else {
            float tot[NV];
            L.group(
                [&](uint32_t &v1, uint32_t &v2) {
                    uint32_t e[4];
                    st.quad_entries(tab, e);
                    quad_cols(e, v1, v2);
                },
                lane, tot);
            L.store(tot, base + t * (16u * S), plane, J.n_vec);
        }
        st.end_group(wbefore);
    };
    uint32_t t = 0;
    for (; t + 1 < groups; t += 2) {
        st.refill_a();
        group(t);
        st.refill_b();
        group(t + 1);
    }
    if (t < groups) {
        st.refill_a();
        group(t);
    }
}

// dx[v][c] = alpha[v] * sum over row blocks of partial[v][b][c] for the 64 columns of
// column block cb, with 1024 threads (16 per column): thread part p sums the row blocks
// [p*chunk, (p+1)*chunk) in ascending order, then the column's outliers p, p + 16, ...
// (value * s_w[r] * g[v][r]), and the 16 parts are added in ascending order, so the
// result does not depend on scheduling.
__device__ __forceinline__ void reduce_tile(const BwdJob &J, uint32_t nrb, uint32_t v,
                                            uint32_t cb, float (&part)[kRedParts][64]) {
    const uint32_t cl = threadIdx.x & 63, p = threadIdx.x >> 6;
    const uint32_t cols = J.cols;
    const uint32_t c = cb * 64 + cl;
    const uint32_t chunk = (nrb + kRedParts - 1) / kRedParts;
    float s = 0.0f;
    if (c < cols) {
        const float *src = J.partial + size_t(v) * nrb * cols + c;
        uint32_t b0 = p * chunk, b1 = b0 + chunk < nrb ? b0 + chunk : nrb;
        // eight loads in flight, then the adds in ascending row-block order
        for (uint32_t b = b0; b < b1; b += 8) {
            float x[8];
#pragma unroll
            for (uint32_t k = 0; k < 8; k++)   // past b1: row block b1 - 1 again (not added)
                x[k] = src[size_t(b + k < b1 ? b + k : b1 - 1) * cols];
#pragma unroll
            for (uint32_t k = 0; k < 8; k++)
                if (b + k < b1) s += x[k];
        }
        if (J.out_col_start) {
            const uint32_t o1 = J.out_col_start[c + 1];
            for (uint32_t o = J.out_col_start[c] + p; o < o1; o += kRedParts) {
                const uint32_t r = J.out_col_rec[o];
                const uint32_t row = r & 0xFFFFFFu;
                const float val = float(int8_t(r >> 24));
                s = fmaf(val, J.s_w[row] * J.g[size_t(v) * J.g_stride + row], s);
            }
        }
    }
    part[p][cl] = s;
    __syncthreads();
    if (p == 0 && c < cols) {
        float tot = part[0][cl];
#pragma unroll
        for (int q = 1; q < kRedParts; q++) tot += part[q][cl];
        J.dx[size_t(v) * cols + c] = J.alpha[v] * tot;
    }
}

}  // namespace blk
}  // namespace mv
