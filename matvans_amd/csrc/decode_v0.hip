// Fused rANS decode x int8 matvec for the WarpPerRow ".bin v0" layout, gfx950.
//
// Layout (create_matrices_WarpPerRow.py:353-430): row r owns K_r granules of 512 bytes
// starting at u16 word cursors[r]; granule g holds, for each of the 32 streams
// ("lanes") of the row, 8 consecutive u16 words of that stream at byte 16*lane.
// Stream l of row r decodes columns 512*(k/16) + 16*l + k%16 for k = 0 .. cols/32-1
// (matvansWarpPerRow-kernel.cu:111-133).
//
// Forward mapping on CDNA4 (row-major): one wave64 decodes two rows, one 32-lane half
// per row, so every lane owns exactly one stream (V0LaneT in v0_stream.h).  Every four
// symbols are packed into one dword with v_perm_b32 and multiplied into the int32
// accumulators with v_dot4_i32_i8, one per activation vector; the two halves are
// reduced separately at the end (the reference's __shfl_down tree, kernel.cu:163-174).
//
// v0_fwd_ref keeps the reference form of the decode step (x unrotated, one fused table
// entry, ring read every step); it is selectable with MATVANS_DECODE_VARIANT=0 for A/B
// measurements and computes the same bits.
#include <hip/hip_runtime.h>

#include "launch.h"
#include "v0_stream.h"

namespace mv {
namespace {

constexpr int kWaves = 4;                  // waves per block
constexpr int kRowsPerBlock = 2 * kWaves;  // two rows per wave

template <int NV>
struct VecChunk {
    uint4 v[NV];
};

__device__ __forceinline__ uint32_t dword_of(const uint4 &v, int d) {
    return d == 0 ? v.x : d == 1 ? v.y : d == 2 ? v.z : v.w;
}

// Wave window and lane row extents for the row-major mapping.
struct RowMajorSetup {
    uint32_t row, win_beg, win_end, row_cur, row_end;
    bool valid;
};

// The four cursor loads take clamped indices instead of branches, so they issue back to
// back (one memory round trip).
__device__ __forceinline__ RowMajorSetup row_major_setup(uint32_t row0, uint32_t lane,
                                                         uint32_t rows,
                                                         const uint32_t *__restrict__ cursors,
                                                         uint32_t payload_size) {
    RowMajorSetup s;
    const uint32_t rmax = rows - 1u;
    s.row = row0 + (lane >> 5);
    s.valid = s.row < rows;
    const uint32_t a = row0 < rows ? row0 : rmax, b = row0 + 2 < rows ? row0 + 2 : rmax;
    const uint32_t r = s.valid ? s.row : rmax, r1 = s.row + 1 < rows ? s.row + 1 : rmax;
    const uint32_t ca = cursors[a], cb = cursors[b], cr = cursors[r], cr1 = cursors[r1];
    uint32_t wbeg = row0 < rows ? ca : 0u;
    uint32_t wend = row0 < rows ? (row0 + 2 < rows ? cb : payload_size) : 0u;
    if (wend > payload_size) wend = payload_size;
    if (wend < wbeg) wend = wbeg;
    s.win_beg = __builtin_amdgcn_readfirstlane(wbeg);
    s.win_end = __builtin_amdgcn_readfirstlane(wend);
    s.row_cur = s.valid ? cr : 0u;
    s.row_end = s.valid ? (s.row + 1 < rows ? cr1 : payload_size) : 0u;
    // consumed here: the compiler otherwise moves the row's cursor loads into the caller's
    // branch for lanes with a row, behind the wait for the window's (a second round trip)
    asm volatile("" ::"v"(s.row_cur), "v"(s.row_end));
    return s;
}

// TABLE 0: dword tables A[q] = f | sym << 24 and B[q] = bias, read with one
// ds_read2st64_b32.  TABLE 2: three 256-byte tables f, bias, sym (64 dwords each, one per
// LDS bank, so every read is conflict-free).  TABLE 3: one packed dword
// T[q] = f | bias << 8 | sym << 24, one LDS access per step and one more dependent VALU
// op.  A random lookup costs the LDS about 7 CU-cycles per wave instruction (a
// two-address read2 about 13); at long rows this layout has two waves per SIMD and is
// bound by the dependent chain, where TABLE 0 measures faster.
template <int NV, int TABLE>
__global__ __launch_bounds__(256) void v0_fwd(int32_t *__restrict__ dst,
                                              const int8_t *__restrict__ vecs,
                                              uint32_t vec_stride, uint32_t rows, uint32_t cols,
                                              const uint32_t *__restrict__ cursors,
                                              int min_value, uint32_t G,
                                              const uint8_t *__restrict__ cdf,
                                              const uint8_t *__restrict__ ppf,
                                              const uint16_t *__restrict__ payload,
                                              uint32_t payload_size) {
    __shared__ uint16_t ring[kWaves * kRingWords * 64];
    __shared__ uint32_t lut[512];
    __shared__ uint16_t dummy[kDummyBytes / 2];
    __shared__ uint32_t cdfs[257];

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6, lane = tid & 63, sl = lane & 31;
    // every global load of the prologue at once: this thread's quantile symbol ppf[tid],
    // cdf entry tid, and the rows' cursors; the table's dependent cdf[s] reads then come
    // from LDS (zero past G, as the reference reads them, matvansWarpPerRow-kernel.cu:86)
    const uint32_t s_q = ppf[tid];
    const uint32_t c_t = cdf[tid <= G ? tid : G];
    RowMajorSetup rs = row_major_setup(blockIdx.x * kRowsPerBlock + wave * 2, lane, rows,
                                       cursors, payload_size);
    cdfs[tid] = tid <= G ? c_t : 0u;
    if (tid == 0) cdfs[256] = 0u;   // G <= 255
    __syncthreads();
    {
        const uint32_t c0 = cdfs[s_q], c1 = cdfs[s_q + 1];
        const uint32_t f = (c1 - c0) & 0xFFu, bias = (tid - c0) & 0xFFu;
        const uint32_t sym = static_cast<uint8_t>(min_value + static_cast<int>(s_q));
        if (TABLE == 0) {
            lut[tid] = f | (sym << 24);
            lut[tid + 256] = bias;
        } else if (TABLE == 3) {
            lut[tid] = f | (bias << 8) | (sym << 24);
        } else {
            uint8_t *tab = reinterpret_cast<uint8_t *>(lut);
            tab[tid] = static_cast<uint8_t>(f);
            tab[256 + tid] = static_cast<uint8_t>(bias);
            tab[512 + tid] = static_cast<uint8_t>(sym);
        }
    }
    V0LaneT<kLoadStream> st;   // each granule read once: non-temporal loads
    st.init(payload, rs.win_beg, rs.win_end, rs.valid, rs.row_cur, rs.row_end, sl, ring, dummy,
            wave, lane);
    st.prologue();

    // group t needs bytes [512t + 16*sl, +16) of each vector, loaded one group ahead
    const uint32_t groups = cols / 512u;
    auto vload = [&](uint32_t t, VecChunk<NV> &c) {
        uint32_t tt = t < groups ? t : groups - 1;
#pragma unroll
        for (int v = 0; v < NV; v++)
            c.v[v] = *reinterpret_cast<const uint4 *>(vecs + size_t(v) * vec_stride + tt * 512u +
                                                      sl * 16u);
    };
    VecChunk<NV> vc0, vc1;
    vload(0, vc0);
    vload(1, vc1);
    __syncthreads();
    st.start();

    int32_t acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = 0;

    auto group = [&](const VecChunk<NV> &vc) {
        uint32_t wbefore = st.waddr;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            uint32_t e0, e1, e2, e3, packed;
            if (TABLE == 0 || TABLE == 3) {
                if (TABLE == 0) {
                    st.pair(lut, e0, e1);
                    st.pair(lut, e2, e3);
                } else {
                    st.pair_packed(lut, e0, e1);
                    st.pair_packed(lut, e2, e3);
                }
                packed = __builtin_amdgcn_perm(e1, e0, 0x0c0c0703u) |
                         __builtin_amdgcn_perm(e3, e2, 0x07030c0cu);
            } else {
                const uint8_t *tab = reinterpret_cast<const uint8_t *>(lut);
                st.pair_u8(tab, e0, e1);
                st.pair_u8(tab, e2, e3);
                packed = __builtin_amdgcn_perm(e1, e0, 0x0c0c0400u) |
                         __builtin_amdgcn_perm(e3, e2, 0x04000c0cu);
            }
#pragma unroll
            for (int v = 0; v < NV; v++)
                acc[v] = __builtin_amdgcn_sdot4(static_cast<int>(packed),
                                                static_cast<int>(dword_of(vc.v[v], d)), acc[v],
                                                false);
        }
        st.end_group(wbefore);
    };

    // straight-line body (two groups, A/B loads) so that every wait counts exactly
    uint32_t t = 0;
    // (each vector chunk is reloaded in place right after its group has used it, so no
    // register copy of a load in flight is needed at the back edge)
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
        a += __shfl_xor(a, 16, 32);
        a += __shfl_xor(a, 8, 32);
        a += __shfl_xor(a, 4, 32);
        a += __shfl_xor(a, 2, 32);
        a += __shfl_xor(a, 1, 32);
        acc[v] = a;
    }
    if (sl == 0 && rs.valid) {
#pragma unroll
        for (int v = 0; v < NV; v++) dst[size_t(v) * rows + rs.row] = acc[v];
    }
}

// Reference form of the step (one fused table entry {f, sym, bias}, ring read every
// step); same ring and refill machinery.
template <int NV>
__global__ __launch_bounds__(256) void v0_fwd_ref(int32_t *__restrict__ dst,
                                                  const int8_t *__restrict__ vecs,
                                                  uint32_t vec_stride, uint32_t rows,
                                                  uint32_t cols,
                                                  const uint32_t *__restrict__ cursors,
                                                  int min_value, uint32_t G,
                                                  const uint8_t *__restrict__ cdf,
                                                  const uint8_t *__restrict__ ppf,
                                                  const uint16_t *__restrict__ payload,
                                                  uint32_t payload_size) {
    __shared__ uint16_t ring[kWaves * kRingWords * 64];
    __shared__ uint32_t lut[256];
    __shared__ uint16_t dummy[kDummyBytes / 2];

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6, lane = tid & 63, sl = lane & 31;
    {
        uint32_t f, sym, bias;
        v0_table_entry(tid, min_value, G, cdf, ppf, f, sym, bias);
        lut[tid] = f | (sym << 8) | (bias << 16);
    }
    RowMajorSetup rs = row_major_setup(blockIdx.x * kRowsPerBlock + wave * 2, lane, rows,
                                       cursors, payload_size);
    V0LaneT<kLoadStream> st;   // each granule read once: non-temporal loads
    st.init(payload, rs.win_beg, rs.win_end, rs.valid, rs.row_cur, rs.row_end, sl, ring, dummy,
            wave, lane);
    st.prologue();
    const uint32_t groups = cols / 512u;
    auto vload = [&](uint32_t t, VecChunk<NV> &c) {
        uint32_t tt = t < groups ? t : groups - 1;
#pragma unroll
        for (int v = 0; v < NV; v++)
            c.v[v] = *reinterpret_cast<const uint4 *>(vecs + size_t(v) * vec_stride + tt * 512u +
                                                      sl * 16u);
    };
    VecChunk<NV> vc0, vc1;
    vload(0, vc0);
    vload(1, vc1);
    __syncthreads();
    st.start();
    uint32_t x = (st.xr << 8) | (st.xr >> 24);
    int32_t acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = 0;

    auto group = [&](const VecChunk<NV> &vc) {
#pragma unroll
        for (int d = 0; d < 4; d++) {
            uint32_t e[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                uint32_t ent = lut[x & 0xFFu];
                uint32_t x1 = (x >> 8) * (ent & 0xFFu) + ((ent >> 16) & 0xFFu);
                bool c = x1 < 0x10000u;
                x = c ? ((x1 << 16) | st.W) : x1;
                st.j += c ? 1u : 0u;
                st.W = st.ring_at(st.j & (kRingWords - 1));
                e[u] = ent;
            }
            uint32_t packed = __builtin_amdgcn_perm(e[1], e[0], 0x0c0c0501u) |
                              __builtin_amdgcn_perm(e[3], e[2], 0x05010c0cu);
#pragma unroll
            for (int v = 0; v < NV; v++)
                acc[v] = __builtin_amdgcn_sdot4(static_cast<int>(packed),
                                                static_cast<int>(dword_of(vc.v[v], d)), acc[v],
                                                false);
        }
    };

    // straight-line body (two groups, A/B loads) so that every wait counts exactly
    uint32_t t = 0;
    // (each vector chunk is reloaded in place right after its group has used it, so no
    // register copy of a load in flight is needed at the back edge)
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
        a += __shfl_xor(a, 16, 32);
        a += __shfl_xor(a, 8, 32);
        a += __shfl_xor(a, 4, 32);
        a += __shfl_xor(a, 2, 32);
        a += __shfl_xor(a, 1, 32);
        acc[v] = a;
    }
    if (sl == 0 && rs.valid) {
#pragma unroll
        for (int v = 0; v < NV; v++) dst[size_t(v) * rows + rs.row] = acc[v];
    }
}

}  // namespace

template <int NV>
static hipError_t launch_nv(const V0Args &a, hipStream_t s) {
    dim3 grid((a.rows + kRowsPerBlock - 1) / kRowsPerBlock), block(256);
    if (a.variant == 0)
        hipLaunchKernelGGL((v0_fwd_ref<NV>), grid, block, 0, s, a.dst, a.vecs, a.vec_stride,
                           a.rows, a.cols, a.cursors, int(a.min_value), uint32_t(a.G), a.cdf,
                           a.ppf, a.payload, a.payload_size);
    else if (a.variant == 1)
        hipLaunchKernelGGL((v0_fwd<NV, 0>), grid, block, 0, s, a.dst, a.vecs, a.vec_stride,
                           a.rows, a.cols, a.cursors, int(a.min_value), uint32_t(a.G), a.cdf,
                           a.ppf, a.payload, a.payload_size);
    else if (a.variant == 3)
        hipLaunchKernelGGL((v0_fwd<NV, 3>), grid, block, 0, s, a.dst, a.vecs, a.vec_stride,
                           a.rows, a.cols, a.cursors, int(a.min_value), uint32_t(a.G), a.cdf,
                           a.ppf, a.payload, a.payload_size);
    else
        hipLaunchKernelGGL((v0_fwd<NV, 2>), grid, block, 0, s, a.dst, a.vecs, a.vec_stride,
                           a.rows, a.cols, a.cursors, int(a.min_value), uint32_t(a.G), a.cdf,
                           a.ppf, a.payload, a.payload_size);
    return hipGetLastError();
}

hipError_t launch_v0_decode_matvec(const V0Args &a, hipStream_t s) {
    switch (a.n_vec) {
        case 1: return launch_nv<1>(a, s);
        case 2: return launch_nv<2>(a, s);
        case 3: return launch_nv<3>(a, s);
        case 4: return launch_nv<4>(a, s);
        case 5: return launch_nv<5>(a, s);
        case 6: return launch_nv<6>(a, s);
        case 7: return launch_nv<7>(a, s);
        case 8: return launch_nv<8>(a, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace mv
