// Block-level bodies of the backward dx = alpha * W^T (s_w * g) from the compressed
// streams, shared by the one-matrix kernels (backward.hip) and the persistent batched
// backward (batch.hip).  The reference has no backward pass.
//
// Mapping (stream-major): a unit is W waves (8, or 4) = W consecutive streams of 64
// consecutive rows; a wave decodes one stream of the 64 rows, one row per lane.  A stream covers the
// same columns in every row, so each 16-symbol group of the wave is a 64 x 16 int8 block
// of W whose column sums, weighted by the rows' scaled gradients, are one
// v_mfma_i32_16x16x64_i8 per four vectors (GroupPipe: the block and the gradient
// digits reach the matrix core through the wave's LDS stage and the transposed LDS
// loads).  The group's 16 sums per vector are stored once, to partial[v][row_block][c];
// reduce_tile sums the row blocks in a fixed order and applies alpha.  Both stages are
// deterministic bit for bit, so one-matrix and batched launches give identical results.
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
constexpr int kRedPartsMax = 16;   // threads per column of the reduce, at most
constexpr int kRedLoads = 16;      // loads in flight per reduce thread
// Threads per column for nrb row blocks and n_vec vectors: with one to three vectors,
// enough that each thread's row blocks are loaded at once (up to kRedPartsMax), and no
// more, since the reduce is short and the dispatch of its waves is most of its cost; from
// four vectors, where the reduce reads that many planes, at most four (fewer waves, more
// loads each: decoder block backward at 8 vectors 116 -> 110 us).  A function of the
// shape only: the summation order is the same for every launch of a matrix.
__host__ __device__ constexpr uint32_t red_parts(uint32_t nrb, uint32_t n_vec) {
    const uint32_t cap = n_vec >= 4 ? 4u : uint32_t(kRedPartsMax);
    const uint32_t p = (nrb + kRedLoads - 1) / kRedLoads;
    return p < 1 ? 1u : (p > cap ? cap : p);
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

// The decode table comes first, so that its lookups address it with the LDS instruction's
// immediate offset (below 64 KiB) and one shift per symbol; the stage next, 16-byte
// aligned for the transposed 8-byte loads and the 16-byte stage writes.
// W: waves (streams) per unit, 8 or 4.
template <int NV, int W = kBwdWaves>
struct V0BwdSmem {
    alignas(16) uint32_t lut[256];
    alignas(16) uint32_t wstage[W * 256];   // per wave: 64 rows x 16 int8
    uint16_t ring[W * kRingWords * 64];
    uint16_t dummy[kDummyBytes / 2];
    uint32_t cdfs[257];                      // the cdf, zero past G, while the table builds
};

// Ring words per lane per set of the quad backward.  4 (one set per group): 8-word sets
// save ~2 VALU per quad but double the ring, and the batched backward then fits two
// blocks per CU instead of three, which costs more than the VALU saved.
template <int S, int MODE>
constexpr int bwd_set_words() {
    return 4;
}

template <int NV, int S, int MODE = 0, int SETW = bwd_set_words<S, MODE>(), int W = kBwdWaves>
struct QuadBwdSmem {
    alignas(16) uint2 tab[MODE == 0 ? 256 : 2];   // two-word table (f | sym << 24, r)
    alignas(16) uint32_t wstage[W * 256];
    uint32_t ring[W * v1::QuadLane<128 / S, SETW>::R4 * 64];
    uint32_t rowtab[MODE == 1 ? 64 * 256 : 1];     // one packed table per row
    uint32_t rowscratch[MODE == 1 ? W * 192 : 1];  // 768 bytes per wave
    uint32_t dummy[v1::QuadLane<128 / S, SETW>::M * 64];
};

// Per-lane setup shared by both codecs: the 64-row window, this lane's row extent and
// its scaled gradients (zero past the last row and past n_vec).
template <int NV>
struct BwdRows {
    uint32_t row, wbeg, wend, rcur, rend;
    bool valid;
    float gs[NV];
    // Every load takes a clamped index instead of a branch, so that all of them issue
    // back to back and cost one memory round trip.
    __device__ __forceinline__ void init(const BwdJob &J, uint32_t rb, uint32_t lane) {
        const uint32_t rows = J.rows, rmax = rows - 1u, ps = J.payload_size;
        const uint32_t row0 = rb * 64u;
        row = row0 + lane;
        valid = row < rows;
        const uint32_t last = row0 + 63u < rows ? row0 + 63u : rmax;
        const uint32_t b = last + 1 < rows ? last + 1 : rmax;
        const uint32_t r = valid ? row : rmax, r1 = row + 1 < rows ? row + 1 : rmax;
        const uint32_t c0 = J.cursors[row0], cb = J.cursors[b], cr = J.cursors[r],
                       cr1 = J.cursors[r1];
        const float sw = J.s_w[r];
        float gv[NV];
#pragma unroll
        for (int v = 0; v < NV; v++)
            gv[v] = J.g[size_t(v < int(J.n_vec) ? v : 0) * J.g_stride + r];
        wbeg = c0;
        wend = last + 1 < rows ? cb : ps;
        if (wend > ps) wend = ps;
        if (wend < wbeg) wend = wbeg;
        wbeg = __builtin_amdgcn_readfirstlane(wbeg);
        wend = __builtin_amdgcn_readfirstlane(wend);
        rcur = valid ? cr : 0u;
        rend = valid ? (row + 1 < rows ? cr1 : ps) : 0u;
#pragma unroll
        for (int v = 0; v < NV; v++) gs[v] = valid && v < int(J.n_vec) ? sw * gv[v] : 0.0f;
        // consumed here: the compiler otherwise moves the row's cursor loads into the
        // caller's branch for lanes with a row (a second round trip before the payload)
        asm volatile("" ::"v"(rcur), "v"(rend));
    }
};


// ---- group sums on the matrix cores
//
// A group is 16 columns of the wave's 64 rows: a 64 x 16 int8 block W_g, lane r holding
// row r as four packed quads.  Its contribution out[v][c] = sum_r W_g[r][c] gs[r][v] is
// computed exactly in integers by v_mfma_i32_16x16x64_i8, with the scaled gradients
// gs = s_w * g written as four int8 digits per row (GradDigits):
//
//   A (16 x 64): row 4 (v % 4) + d holds digit d of vector v over the 64 rows (vectors
//                4..7 use a second A operand)
//   B (64 x 16): W_g
//   D (16 x 16): lane 16 k + i, register d = sum_r digit_d[r][v] W_g[r][i], v = k (+4)
//
// Both operands hold, in lane 16 k + i, the 16 values of K rows 16 k .. 16 k + 15 at M / N
// index i.  A wave stores its block row-major in its LDS stage (row r at bytes 16 r ..
// 16 r + 15, one 16-byte write per lane) and reads it back transposed with two
// ds_read_b64_tr_b8: within each 16 lanes, that load gives output lane 8 h + j, byte m,
// the byte j of the 8 bytes read by lane 2 m + h, so lane 2 m + h pointing at row
// 16 k + m (+ 8 for the second load), bytes 8 h .. 8 h + 7, leaves lane 16 k + i with byte
// i of rows 16 k .. 16 k + 7 (+ 8).  The digit operand is staged the same way once per
// unit.  The four digit sums are exact (|sum| <= 2^20) and are combined in a fixed order,
// so the result does not depend on scheduling.

typedef int bwd_i32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) bwd_i32x2 lds_i32x2;

// Byte offset, within a wave's 1 KiB stage, of the 8 bytes this lane reads in the first
// transposed load (the second reads 128 bytes further: rows + 8).
__device__ __forceinline__ uint32_t tr_offset(uint32_t lane) {
    return 16u * (16u * (lane >> 4) + ((lane & 15u) >> 1)) + 8u * (lane & 1u);
}

// The stage (64 rows x 16 bytes) as an MFMA operand: lane 16 k + i gets byte i of rows
// 16 k .. 16 k + 15.
__device__ __forceinline__ mv_i32x4 stage_operand(uint8_t *stage, uint32_t tr_off) {
    const bwd_i32x2 lo = __builtin_amdgcn_ds_read_tr8_b64_v2i32((lds_i32x2 *)(stage + tr_off));
    const bwd_i32x2 hi =
        __builtin_amdgcn_ds_read_tr8_b64_v2i32((lds_i32x2 *)(stage + tr_off + 128u));
    mv_i32x4 r;
    r[0] = lo[0]; r[1] = lo[1]; r[2] = hi[0]; r[3] = hi[1];
    return r;
}

// Scaled gradients of the wave's 64 rows as int8 digits.  For vector v, with m = f 2^e
// (f in [0.5, 1)) the largest |gs| over the 64 rows, q = rint(gs 2^-e * 127 * 2^24 / f) is
// an integer with |q| <= 127 * 2^24 + 256 < 2^31, written in balanced base 256,
// q = 2^24 d0 + 2^16 d1 + 2^8 d2 + d3, every digit in [-128, 127] (d3, d2, d1 are
// sign-extended low bytes and the remainders exact multiples of 256).  gs = u q with
// u = m / (127 * 2^24), up to the rounding of the product (2^-24 of |gs|) and of rint
// (u / 2 = 2^-32.99 m): as close as an fp32 product for all but the smallest gs.
// uh, ul = 2^16 u, u for the vector of this lane's D registers.
template <int NV>
struct GradDigits {
    static constexpr int NA = NV > 4 ? 2 : 1;
    mv_i32x4 a[NA];
    float uh[NA], ul[NA];   // 2^16 u and u
    __device__ __forceinline__ void init(const float (&gs)[NV], uint32_t lane, uint8_t *stage,
                                         uint32_t tr_off) {
        uint32_t pk[4 * NA];
        float top[4 * NA];   // m / 127 = 2^24 u
#pragma unroll
        for (int v = 0; v < 4 * NA; v++) {
            pk[v] = 0u;
            top[v] = 0.0f;
        }
#pragma unroll
        for (int v = 0; v < NV; v++) {
            float m = fabsf(gs[v]);
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
            int e = 0;
            const float f = frexpf(m, &e);   // 0 for m = 0
            const float inv = m > 0.0f ? 2130706432.0f / f : 0.0f;   // 127 * 2^24 / f
            top[v] = ldexpf(f * (1.0f / 127.0f), e);
            const int32_t q = int32_t(rintf(ldexpf(gs[v], -e) * inv));
            const int32_t d3 = int32_t(uint32_t(q) << 24) >> 24;
            const int32_t q1 = (q - d3) >> 8;
            const int32_t d2 = int32_t(uint32_t(q1) << 24) >> 24;
            const int32_t q2 = (q1 - d2) >> 8;
            const int32_t d1 = int32_t(uint32_t(q2) << 24) >> 24;
            const int32_t d0 = (q2 - d1) >> 8;
            pk[v] = (uint32_t(d0) & 0xFFu) | ((uint32_t(d1) & 0xFFu) << 8) |
                    ((uint32_t(d2) & 0xFFu) << 16) | (uint32_t(d3) << 24);
        }
        const uint32_t k = lane >> 4;
#pragma unroll
        for (int h = 0; h < NA; h++) {
            // the wave's own LDS accesses run in order: each read sees the write before it
            *reinterpret_cast<uint4 *>(stage + 16u * lane) =
                make_uint4(pk[4 * h], pk[4 * h + 1], pk[4 * h + 2], pk[4 * h + 3]);
            a[h] = stage_operand(stage, tr_off);
            float t = top[4 * h];
#pragma unroll
            for (int j = 1; j < 4; j++) t = k == uint32_t(j) ? top[4 * h + j] : t;
            uh[h] = t * 0x1p-8f;
            ul[h] = t * 0x1p-24f;
        }
    }
};

// The group step, software-pipelined by the units.  After group t's decode: mma() issues
// group t - 1's products (its B operand, read from the LDS stage before the decode, has
// long arrived: no wait on the LDS queue in the decode's instruction stream), stage()
// writes group t's packed quads P[0..3] (the lane's row, 16 columns in order) to the stage
// and issues its transposed reads, and finish() scales and stores group t - 1's sums.  A
// group's sums 2^24 D0 + 2^16 D1 + 2^8 D2 + D3 are scaled as
// 2^16 u (2^8 D0 + D1) + u (2^8 D2 + D3), both pairs exact in int32 (|D| <= 2^20).
// finish() stores vector v's 16 column sums to byte goff + 4 (v * plane + 0 .. 15) of the
// job's partial sums: goff is wave-uniform, the lane's offset fixed (loff).  Every lane
// issues the store, through a buffer resource over the partial sums (below 4 GiB, the
// entry points check); lanes without a sum (vector n_vec and up) give an offset past
// its end, which the hardware drops.  Behind a per-lane branch the store would make the
// compiler count it as possibly absent and wait for the other ring set's loads as well.
// The partial sums end at or below byte 2^32 - 32 (n_vec <= 8 planes, each below 2^29
// bytes), so a store at kDropOff is dropped.
constexpr uint32_t kDropOff = 0xFFFFFFFFu;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t partial_rsrc(const BwdJob &J, uint32_t nrb) {
    const uint64_t bytes = uint64_t(J.n_vec) * nrb * J.cols * sizeof(float);
    return __builtin_amdgcn_make_buffer_rsrc(J.partial, 0, static_cast<int>(uint32_t(bytes)),
                                             0x00020000);
}

template <int NV>
struct GroupPipe {
    static constexpr int NA = GradDigits<NV>::NA;
    mv_i32x4 d[NA];
    uint32_t loff[NA];   // byte offset 4 ((4 h + k) * plane + i) for lane 16 k + i
    uint32_t drop[NA];   // ~0 when 4 h + k >= n_vec, else 0
    __device__ __forceinline__ void init(uint32_t lane, size_t plane, uint32_t n_vec) {
        const uint32_t i = lane & 15u, k = lane >> 4;
#pragma unroll
        for (int h = 0; h < NA; h++) {
            d[h] = mv_i32x4{0, 0, 0, 0};
            loff[h] = uint32_t(((4u * h + k) * plane + i) * sizeof(float));
            drop[h] = 4u * h + k < n_vec ? 0u : kDropOff;
        }
    }
    __device__ __forceinline__ static mv_i32x4 stage(const uint32_t (&P)[4], uint32_t lane,
                                                      uint8_t *stg, uint32_t tr_off) {
        *reinterpret_cast<uint4 *>(stg + 16u * lane) = make_uint4(P[0], P[1], P[2], P[3]);
        return stage_operand(stg, tr_off);
    }
    __device__ __forceinline__ void mma(const GradDigits<NV> &gd, const mv_i32x4 &b) {
        const mv_i32x4 zero = {0, 0, 0, 0};
#pragma unroll
        for (int h = 0; h < NA; h++)
            d[h] = __builtin_amdgcn_mfma_i32_16x16x64_i8(gd.a[h], b, zero, 0, 0, 0);
    }
    // as many stores as finish() issues, into a buffer of zero bytes (the hardware drops
    // them), at offsets 64 bytes apart (the compiler neither drops nor merges any)
    __device__ __forceinline__ void drop_stores(const void *any, uint32_t j) const {
        const __amdgpu_buffer_rsrc_t r =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(any), 0, 0, 0x00020000);
#pragma unroll
        for (int h = 0; h < NA; h++)
            __builtin_amdgcn_raw_buffer_store_b32(
                0u, r, static_cast<int>(64u * (uint32_t(NA) * j + uint32_t(h))), 0, 0);
    }
    __device__ __forceinline__ void finish(const GradDigits<NV> &gd, __amdgpu_buffer_rsrc_t rs,
                                           uint32_t goff) const {
#pragma unroll
        for (int h = 0; h < NA; h++) {
            const int32_t hi = int32_t(uint32_t(d[h][0]) << 8) + d[h][1];
            const int32_t lo = int32_t(uint32_t(d[h][2]) << 8) + d[h][3];
            const float t = fmaf(float(hi), gd.uh[h], float(lo) * gd.ul[h]);
            const uint32_t off = goff == kDropOff ? kDropOff : (goff + loff[h]) | drop[h];
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(t), rs, static_cast<int>(off),
                                                  0, 0);
        }
    }
};

// .bin v0 unit: row block rb, streams W sg .. W sg + W - 1 (W waves).  fresh
// (block-uniform): build the table and synchronise the block first; a unit that follows
// one of the same matrix in the same block reuses the table, and its waves go on without
// a barrier (each wave's ring and stage are its own).
template <int NV, int W = kBwdWaves>
__device__ __forceinline__ void v0_bwd_unit(const BwdJob &J, uint32_t rb, uint32_t sg,
                                            uint32_t nrb, V0BwdSmem<NV, W> &sm,
                                            bool fresh = true) {
    const uint32_t tid = threadIdx.x;
    // the wave index is the same on every lane: scalar, so are the stream and the output
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    // fresh: quantile tid's symbol ppf[tid] and cdf entry tid load with the rows' indices
    // (one round trip); the table's dependent cdf[s] reads then come from LDS (zero past
    // G, as the reference reads them, matvansWarpPerRow-kernel.cu:86)
    const uint32_t G = J.G, q = tid & 255u;
    uint32_t s_q = 0, c_t = 0;
    if (fresh) {
        s_q = J.ppf[q];
        c_t = J.cdf[q <= G ? q : G];
    }
    const uint32_t stream = sg * W + wave;   // 0..31
    BwdRows<NV> R;
    R.init(J, rb, lane);
    if (fresh) {   // packed table f | bias << 8 | sym << 24
        if (tid < 256) sm.cdfs[tid] = tid <= G ? c_t : 0u;
        if (tid == 0) sm.cdfs[256] = 0u;   // G <= 255
        __syncthreads();
        if (tid < 256) {
            const uint32_t c0 = sm.cdfs[s_q], c1 = sm.cdfs[s_q + 1];
            const uint32_t f = (c1 - c0) & 0xFFu, bias = (tid - c0) & 0xFFu;
            const uint32_t sym = static_cast<uint8_t>(J.min_value + static_cast<int>(s_q));
            sm.lut[tid] = f | (bias << 8) | (sym << 24);
        }
    }

    V0Lane st;
    st.init(J.payload, R.wbeg, R.wend, R.valid, R.rcur, R.rend, stream, sm.ring, sm.dummy, wave,
            lane);
    st.prologue();
    if (fresh) __syncthreads();
    st.start();

    const uint32_t cols = J.cols;
    const uint32_t groups = cols / 512u;
    const size_t plane = size_t(nrb) * cols;
    // byte offsets into the partial sums (below 4 GiB): wave-uniform
    const uint32_t base = (rb * cols + stream * 16u) * uint32_t(sizeof(float));
    const __amdgpu_buffer_rsrc_t prs = partial_rsrc(J, nrb);
    // four symbols (byte 3 of each table entry) packed in column order
    auto next4 = [&]() {
        uint32_t e[4];
        st.pair_packed(sm.lut, e[0], e[1]);
        st.pair_packed(sm.lut, e[2], e[3]);
        return __builtin_amdgcn_perm(e[1], e[0], 0x0c0c0703u) |
               __builtin_amdgcn_perm(e[3], e[2], 0x07030c0cu);
    };
    uint8_t *stage = reinterpret_cast<uint8_t *>(sm.wstage + wave * 256);
    const uint32_t tr_off = tr_offset(lane);
    GradDigits<NV> gd;
    gd.init(R.gs, lane, stage, tr_off);
    GroupPipe<NV> pipe;
    pipe.init(lane, plane, J.n_vec);
    // group t - 1's output byte offset (wave-uniform; kDropOff before the first group),
    // its B operand in b.  The products and the store are issued for every group, the
    // first group's dropped, and two dropped sets of stores follow the ring's first loads:
    // the loop then issues the same memory operations between a ring set's load and its
    // use on every pass, and the compiler's wait for that set stops short of the other
    // set's load (with a pass or a prologue that had fewer, it waited for both).
    uint32_t prev = kDropOff;
    mv_i32x4 b = {0, 0, 0, 0};
    pipe.drop_stores(J.partial, 0);
    pipe.drop_stores(J.partial, 1);
    auto group = [&](uint32_t t) {
        uint32_t wbefore = st.waddr;
        uint32_t P[4];
#pragma unroll
        for (int d = 0; d < 4; d++) P[d] = next4();
        st.end_group(wbefore);
        pipe.mma(gd, b);
        b = GroupPipe<NV>::stage(P, lane, stage, tr_off);
        pipe.finish(gd, prs, prev);
        prev = base + t * 512u * uint32_t(sizeof(float));
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
    pipe.mma(gd, b);
    pipe.finish(gd, prs, prev);
}

// Quad-coder unit (S = 32, 64 or 128 streams): row block rb, streams W sg .. W sg + W - 1;
// fresh as for v0_bwd_unit.
// LAY 1: coder 2, the row's streams back to back (lane r finds its row's stream from the
// row's end offsets).
template <int NV, int S, int MODE = 0, int SETW = bwd_set_words<S, MODE>(), int W = kBwdWaves,
          int LAY = 0>
__device__ __forceinline__ void quad_bwd_unit(const BwdJob &J, uint32_t rb, uint32_t sg,
                                              uint32_t nrb,
                                              QuadBwdSmem<NV, S, MODE, SETW, W> &sm,
                                              bool fresh = true) {
    static_assert(MODE == 0 || S == 64, "one table per row: 64 streams");
    static_assert(LAY == 0 || MODE == 0, "coder 2 has one table per matrix");
    using QL = v1::QuadLane<128 / S, SETW, LAY>;
    constexpr int GPS = SETW / 4;   // groups per set period
    static_assert(GPS == 1 || GPS == 2, "set period of one or two groups");
    const uint32_t tid = threadIdx.x;
    // the wave index is the same on every lane: scalar, so are the stream and the output
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    // the table entries, the coder-2 row table and the rows' indices load together (one
    // memory round trip before the payload)
    const uint32_t stream = sg * W + wave;   // 0..S-1
    uint32_t ga = 0, gb = 0;
    if constexpr (MODE == 0) {
        ga = J.glut[tid & 255u];
        gb = J.glut[256u + (tid & 255u)];
    }
    LaneTable lt{0, 0, 0};
    if constexpr (LAY == 1) {
        const uint32_t r = rb * 64u + lane;
        lt = lane_table<S>(J.payload, r < J.rows ? r : J.rows - 1u, stream);
    }
    BwdRows<NV> R;
    R.init(J, rb, lane);
    // one table per matrix: the two-word entries (f | sym << 24, r) of the A / B tables;
    // one per row: packed T[q] = f | r << 8 | sym << 24, lane l reading row l's
    const uint32_t *tab = nullptr;
    if constexpr (MODE == 0) {
        if (fresh && tid < 256) sm.tab[tid] = make_uint2(ga, gb);
    } else {
        __syncthreads();   // every wave of the previous unit is done with the tables
        uint8_t *scratch = reinterpret_cast<uint8_t *>(sm.rowscratch + wave * 192);
        for (uint32_t k = 0; k < 64u / W; k++) {   // the same count on every wave
            const uint32_t r = wave * (64u / W) + k;
            uint32_t *t = sm.rowtab + r * 256u;
            build_row_table(J.freq, rb * 64u + r, J.G, rb * 64u + r < J.rows, J.min_value,
                            lane, scratch,
                            [&](uint32_t q, uint32_t f, uint32_t b, uint32_t sy) {
                                t[q] = f | (b << 8) | (sy << 24);
                            });
        }
        tab = sm.rowtab + lane * 256u;
    }

    QL st;
    if constexpr (LAY == 0) {
        st.init(J.payload, R.wbeg, R.wend, R.valid, R.rcur, R.rend, stream, sm.ring, sm.dummy,
                wave, lane);
    } else {
        const LaneStream ls = lane_stream<S>(lt, R.rcur, R.valid);
        st.init_lanes(J.payload, R.wbeg, R.wend, R.valid, ls.body, ls.kb, S, stream, ls.tail,
                      ls.tail_n, sm.ring, sm.dummy, wave, lane);
    }
    st.prologue();
    if (fresh || MODE == 1) __syncthreads();
    st.start();

    const uint32_t cols = J.cols;
    const uint32_t groups = cols / (16u * S);
    const size_t plane = size_t(nrb) * cols;
    // byte offsets into the partial sums (below 4 GiB): wave-uniform
    const uint32_t base = (rb * cols + stream * 16u) * uint32_t(sizeof(float));
    const __amdgpu_buffer_rsrc_t prs = partial_rsrc(J, nrb);
    uint8_t *stage = reinterpret_cast<uint8_t *>(sm.wstage + wave * 256);
    const uint32_t tr_off = tr_offset(lane);
    GradDigits<NV> gd;
    gd.init(R.gs, lane, stage, tr_off);
    GroupPipe<NV> pipe;
    pipe.init(lane, plane, J.n_vec);
    // group t - 1's output byte offset (wave-uniform; kDropOff before the first group),
    // its B operand in b.  The products and the store are issued for every group, the
    // first group's dropped, and a pass's worth of dropped stores (two set periods of GPS
    // groups) follows the ring's first loads: the loop then issues the same memory
    // operations between a ring set's load and its use on every pass, and the compiler's
    // wait for that set stops short of the other set's load (with a pass or a prologue
    // that had fewer, it waited for both).
    uint32_t prev = kDropOff;
    mv_i32x4 b = {0, 0, 0, 0};
#pragma unroll
    for (uint32_t j = 0; j < 2u * GPS; j++) pipe.drop_stores(J.partial, j);
    auto group = [&](uint32_t t) {
        uint32_t P[4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
            if constexpr (MODE == 0) P[d] = st.quad(sm.tab);
            else P[d] = st.quad(tab);
        }
        pipe.mma(gd, b);
        b = GroupPipe<NV>::stage(P, lane, stage, tr_off);
        pipe.finish(gd, prs, prev);
        prev = base + t * (16u * S) * uint32_t(sizeof(float));
    };
    // a set period: GPS groups after one refill, the consumed words counted once (at most
    // 4 * GPS < R4)
    auto period = [&](uint32_t t0, uint32_t ng) {
        uint32_t wbefore = st.waddr;
        group(t0);
        if constexpr (GPS == 2) {
            if (ng > 1) group(t0 + 1);
        }
        st.end_group(wbefore);
    };
    uint32_t t = 0;
    for (; t + 2 * GPS <= groups; t += 2 * GPS) {
        st.refill_a();
        period(t, GPS);
        st.refill_b();
        period(t + GPS, GPS);
    }
    if (t < groups) {
        st.refill_a();
        period(t, groups - t < uint32_t(GPS) ? groups - t : uint32_t(GPS));
        if (t + GPS < groups) {
            st.refill_b();
            period(t + GPS, groups - t - GPS);
        }
    }
    pipe.mma(gd, b);
    pipe.finish(gd, prs, prev);
}

// dx[v][c] = alpha[v] * sum over row blocks of partial[v][b][c] for the 64 columns of
// column block cb, with P = red_parts(nrb, n_vec) threads per column (a block of at least 64 P
// threads; the others add nothing): thread part p sums the row blocks [p*chunk,
// (p+1)*chunk) in ascending order, then the column's outliers p, p + P, ...
// (value * s_w[r] * g[v][r]), and the P parts are added in ascending order, so the
// result does not depend on scheduling.
__device__ __forceinline__ void reduce_tile(const BwdJob &J, uint32_t nrb, uint32_t v,
                                            uint32_t cb, float (&part)[kRedPartsMax][64]) {
    const uint32_t cl = threadIdx.x & 63, p = threadIdx.x >> 6;
    const uint32_t cols = J.cols;
    const uint32_t c = cb * 64 + cl;
    const uint32_t P = red_parts(nrb, J.n_vec);
    const uint32_t chunk = (nrb + P - 1) / P;
    float s = 0.0f;
    if (c < cols && p < P) {
        const float *src = J.partial + size_t(v) * nrb * cols + c;
        uint32_t b0 = p * chunk, b1 = b0 + chunk < nrb ? b0 + chunk : nrb;
        // kRedLoads loads in flight, then the adds in ascending row-block order
        for (uint32_t b = b0; b < b1; b += kRedLoads) {
            float x[kRedLoads];
#pragma unroll
            for (uint32_t k = 0; k < uint32_t(kRedLoads); k++)   // past b1: b1 - 1 (not added)
                x[k] = src[size_t(b + k < b1 ? b + k : b1 - 1) * cols];
#pragma unroll
            for (uint32_t k = 0; k < uint32_t(kRedLoads); k++)
                if (b + k < b1) s += x[k];
        }
        if (J.out_col_start) {
            const uint32_t o1 = J.out_col_start[c + 1];
            for (uint32_t o = J.out_col_start[c] + p; o < o1; o += P) {
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
        for (uint32_t q = 1; q < P; q++) tot += part[q][cl];
        J.dx[size_t(v) * cols + c] = J.alpha[v] * tot;
    }
}

}  // namespace blk
}  // namespace mv
