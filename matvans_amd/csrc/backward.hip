// Backward of the compressed linear layer on gfx950, from the ".bin v0" streams and from
// the MVX1 quad-coder streams (32 or 64 per row):
//
//   dx[v][c] = alpha[v] * sum_r W[r, c] * (s_w[r] * g[v][r])
//
// W is re-decoded from the same row-ordered rANS streams the forward reads; no dense
// copy is ever built.  The reference has no backward pass.
//
// Mapping (stream-major): a wave decodes stream index l of 64 consecutive rows, one row
// per lane.  Stream l covers the same columns 512*(k/16) + 16*l + k%16 in every row
// (create_matrices_WarpPerRow.py:361-366), so at step k all 64 lanes contribute to one
// column: each lane forms W[r,c] * gs[r] in fp32, and every four steps a fixed butterfly
// (two v_permlane32_swap, one v_permlane16_swap, four DPP row adds) sums the 64 values
// of four columns at once, in the same order on every run.  The 64-row totals are
// written to partial[v][row_block][c]; a second kernel sums the row blocks in a fixed
// order and applies alpha.  Both stages are deterministic bit for bit.
//
// A block is 8 waves = 8 consecutive streams of the same 64 rows, so one 128-byte line of
// a granule (8 lanes x 16 bytes) is read by the 8 waves of one CU.
//
// The quad coder's streams follow the same plan: a quad step yields four consecutive
// columns 16*S*t + 16*l + 4*d + (0..3) of stream l, which are summed over the 64 rows by
// the same butterfly; with 64 streams a granule holds 8 bytes per stream.
#include <hip/hip_runtime.h>

#include "launch.h"
#include "v0_stream.h"
#include "v1_stream.h"

namespace mv {
namespace {

constexpr int kBwdWaves = 8;

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f(float v) {
    int r = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, true);
    return __int_as_float(r);
}

// Sum over the 16 lanes of each DPP row in a fixed order (every lane gets its row sum).
__device__ __forceinline__ float row_sum_fixed(float v) {
    v += dpp_f<0xB1, 0xF>(v);    // quad_perm [1,0,3,2]
    v += dpp_f<0x4E, 0xF>(v);    // quad_perm [2,3,0,1]
    v += dpp_f<0x141, 0xF>(v);   // row_half_mirror
    v += dpp_f<0x140, 0xF>(v);   // row_mirror
    return v;
}

// Four column totals at once, each over the 64 lanes, in a fixed order.
// v_permlane32_swap(a, b) exchanges a[32..63] with b[0..31]: after it, a + b holds
// column a's pair sums in lanes 0..31 and column b's in lanes 32..63.
// v_permlane16_swap does the same between 16-lane halves of each 32-lane half.  The
// rows of 16 lanes then hold columns (c0, c2, c1, c3); row_sum_fixed finishes them.
__device__ __forceinline__ float sum4_fixed(float c0, float c1, float c2, float c3) {
    auto s01 = __builtin_amdgcn_permlane32_swap(__float_as_uint(c0), __float_as_uint(c1), false,
                                                false);
    auto s23 = __builtin_amdgcn_permlane32_swap(__float_as_uint(c2), __float_as_uint(c3), false,
                                                false);
    float a = __uint_as_float(s01[0]) + __uint_as_float(s01[1]);
    float b = __uint_as_float(s23[0]) + __uint_as_float(s23[1]);
    auto t = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false,
                                              false);
    return row_sum_fixed(__uint_as_float(t[0]) + __uint_as_float(t[1]));
}

template <int NV>
__global__ __launch_bounds__(512) void v0_bwd_partial(
    float *__restrict__ partial, const float *__restrict__ g, uint32_t g_stride,
    const float *__restrict__ s_w, uint32_t rows, uint32_t cols,
    const uint32_t *__restrict__ cursors, int min_value, uint32_t G,
    const uint8_t *__restrict__ cdf, const uint8_t *__restrict__ ppf,
    const uint16_t *__restrict__ payload, uint32_t payload_size) {
    __shared__ uint16_t ring[kBwdWaves * kRingWords * 64];
    __shared__ uint32_t lut[256];
    __shared__ uint16_t dummy[kDummyBytes / 2];

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6, lane = tid & 63;
    if (tid < 256) {   // packed table f | bias << 8 | sym << 24
        uint32_t f, sym, bias;
        v0_table_entry(tid, min_value, G, cdf, ppf, f, sym, bias);
        lut[tid] = f | (bias << 8) | (sym << 24);
    }
    const uint32_t stream = blockIdx.y * kBwdWaves + wave;   // 0..31
    const uint32_t row0 = blockIdx.x * 64u;
    const uint32_t row = row0 + lane;
    const bool valid = row < rows;

    // window: rows row0 .. min(row0+63, rows-1)
    const uint32_t last = (row0 + 63u < rows ? row0 + 63u : rows - 1u);
    uint32_t wbeg = cursors[row0];
    uint32_t wend = last + 1 < rows ? cursors[last + 1] : payload_size;
    if (wend > payload_size) wend = payload_size;
    if (wend < wbeg) wend = wbeg;
    wbeg = __builtin_amdgcn_readfirstlane(wbeg);
    wend = __builtin_amdgcn_readfirstlane(wend);
    const uint32_t rcur = valid ? cursors[row] : 0u;
    const uint32_t rend = valid ? (row + 1 < rows ? cursors[row + 1] : payload_size) : 0u;

    float gs[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
        gs[v] = valid ? s_w[row] * g[size_t(v) * g_stride + row] : 0.0f;

    V0Lane st;
    st.init(payload, wbeg, wend, valid, rcur, rend, stream, ring, dummy, wave, lane);
    st.prologue();
    __syncthreads();
    st.start();

    const uint32_t groups = cols / 512u;
    const size_t nrb = gridDim.x;   // row blocks
    auto out_ptr = [&](int v, uint32_t t) {
        return partial + (size_t(v) * nrb + blockIdx.x) * cols + t * 512u + stream * 16u;
    };

    auto group = [&](uint32_t t) {
        uint32_t wbefore = st.waddr;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            uint32_t e[4];
            st.pair_packed(lut, e[0], e[1]);
            st.pair_packed(lut, e[2], e[3]);
            float w[4];
#pragma unroll
            for (int u = 0; u < 4; u++) w[u] = static_cast<float>(static_cast<int8_t>(e[u] >> 24));
#pragma unroll
            for (int v = 0; v < NV; v++) {
                // lanes 0, 16, 32, 48 hold the totals of columns 4d+0, 4d+2, 4d+1, 4d+3
                float tot = sum4_fixed(w[0] * gs[v], w[1] * gs[v], w[2] * gs[v], w[3] * gs[v]);
                if ((lane & 15) == 0) {
                    uint32_t col = ((lane >> 4) & 1) * 2 + (lane >> 5);
                    out_ptr(v, t)[4 * d + col] = tot;
                }
            }
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

template <int NV, int S>
__global__ __launch_bounds__(512) void quad_bwd_partial(
    float *__restrict__ partial, const float *__restrict__ g, uint32_t g_stride,
    const float *__restrict__ s_w, uint32_t rows, uint32_t cols,
    const uint32_t *__restrict__ cursors, const uint32_t *__restrict__ glut,
    const uint16_t *__restrict__ payload, uint32_t payload_size) {
    constexpr int G4 = 128 / S;
    using QL = v1::QuadLane<G4>;
    __shared__ uint32_t ring[kBwdWaves * QL::R4 * 64];
    __shared__ uint32_t tab[256];
    __shared__ uint32_t dummy[G4 * 64];

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6, lane = tid & 63;
    // packed table T[q] = f | r << 8 | sym << 24 from the A / B tables (A = f | sym << 24,
    // B = r)
    if (tid < 256) tab[tid] = (glut[tid] & 0xFF0000FFu) | ((glut[256 + tid] & 0xFFu) << 8);
    const uint32_t stream = blockIdx.y * kBwdWaves + wave;   // 0..S-1
    const uint32_t row0 = blockIdx.x * 64u;
    const uint32_t row = row0 + lane;
    const bool valid = row < rows;

    const uint32_t last = (row0 + 63u < rows ? row0 + 63u : rows - 1u);
    uint32_t wbeg = cursors[row0];
    uint32_t wend = last + 1 < rows ? cursors[last + 1] : payload_size;
    if (wend > payload_size) wend = payload_size;
    if (wend < wbeg) wend = wbeg;
    wbeg = __builtin_amdgcn_readfirstlane(wbeg);
    wend = __builtin_amdgcn_readfirstlane(wend);
    const uint32_t rcur = valid ? cursors[row] : 0u;
    const uint32_t rend = valid ? (row + 1 < rows ? cursors[row + 1] : payload_size) : 0u;

    float gs[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
        gs[v] = valid ? s_w[row] * g[size_t(v) * g_stride + row] : 0.0f;

    QL st;
    st.init(payload, wbeg, wend, valid, rcur, rend, stream, ring, dummy, wave, lane);
    st.prologue();
    __syncthreads();
    st.start();

    const uint32_t groups = cols / (16u * S);
    const size_t nrb = gridDim.x;
    auto out_ptr = [&](int v, uint32_t t) {
        return partial + (size_t(v) * nrb + blockIdx.x) * cols + t * (16u * S) + stream * 16u;
    };
    auto group = [&](uint32_t t) {
        uint32_t wbefore = st.waddr;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const uint32_t packed = st.quad(tab);
            float w[4];
#pragma unroll
            for (int u = 0; u < 4; u++)
                w[u] = static_cast<float>(static_cast<int8_t>(packed >> (8 * u)));
#pragma unroll
            for (int v = 0; v < NV; v++) {
                float tot = sum4_fixed(w[0] * gs[v], w[1] * gs[v], w[2] * gs[v], w[3] * gs[v]);
                if ((lane & 15) == 0) {
                    uint32_t col = ((lane >> 4) & 1) * 2 + (lane >> 5);
                    out_ptr(v, t)[4 * d + col] = tot;
                }
            }
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

// dx[v][c] = alpha[v] * sum over row blocks of partial[v][b][c].  A block covers 64
// columns with 16 threads per column; thread part p sums the row blocks
// [p*chunk, (p+1)*chunk) in ascending order and the 16 parts are added in ascending
// order, so the result does not depend on scheduling.
constexpr int kRedParts = 16;

__global__ __launch_bounds__(1024) void v0_bwd_reduce(float *__restrict__ dx,
                                                      const float *__restrict__ partial,
                                                      uint32_t nrb, uint32_t cols,
                                                      const float *__restrict__ alpha) {
    __shared__ float part[kRedParts][64];
    const uint32_t v = blockIdx.y;
    const uint32_t cl = threadIdx.x & 63, p = threadIdx.x >> 6;
    const uint32_t c = blockIdx.x * 64 + cl;
    const uint32_t chunk = (nrb + kRedParts - 1) / kRedParts;
    float s = 0.0f;
    if (c < cols) {
        const float *src = partial + size_t(v) * nrb * cols + c;
        uint32_t b0 = p * chunk, b1 = b0 + chunk < nrb ? b0 + chunk : nrb;
        for (uint32_t b = b0; b < b1; b++) s += src[size_t(b) * cols];
    }
    part[p][cl] = s;
    __syncthreads();
    if (p == 0 && c < cols) {
        float tot = part[0][cl];
#pragma unroll
        for (int q = 1; q < kRedParts; q++) tot += part[q][cl];
        dx[size_t(v) * cols + c] = alpha[v] * tot;
    }
}

template <int NV>
hipError_t bwd_nv(const V0BwdArgs &a, hipStream_t s) {
    dim3 grid((a.rows + 63) / 64, 32 / kBwdWaves), block(512);
    hipLaunchKernelGGL((v0_bwd_partial<NV>), grid, block, 0, s, a.partial, a.g, a.g_stride,
                       a.s_w, a.rows, a.cols, a.cursors, int(a.min_value), uint32_t(a.G), a.cdf,
                       a.ppf, a.payload, a.payload_size);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(v0_bwd_reduce, dim3((a.cols + 63) / 64, a.n_vec), dim3(64 * kRedParts),
                       0, s, a.dx, a.partial, (a.rows + 63) / 64, a.cols, a.alpha);
    return hipGetLastError();
}

template <int NV, int S>
hipError_t quad_bwd_nv(const QuadBwdArgs &a, hipStream_t s) {
    dim3 grid((a.rows + 63) / 64, S / kBwdWaves), block(512);
    hipLaunchKernelGGL((quad_bwd_partial<NV, S>), grid, block, 0, s, a.partial, a.g, a.g_stride,
                       a.s_w, a.rows, a.cols, a.cursors, a.lut, a.payload, a.payload_size);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(v0_bwd_reduce, dim3((a.cols + 63) / 64, a.n_vec), dim3(64 * kRedParts),
                       0, s, a.dx, a.partial, (a.rows + 63) / 64, a.cols, a.alpha);
    return hipGetLastError();
}

template <int S>
hipError_t quad_bwd_s(const QuadBwdArgs &a, hipStream_t s) {
    switch (a.n_vec) {
        case 1: return quad_bwd_nv<1, S>(a, s);
        case 2: return quad_bwd_nv<2, S>(a, s);
        case 3: return quad_bwd_nv<3, S>(a, s);
        case 4: return quad_bwd_nv<4, S>(a, s);
        case 5: return quad_bwd_nv<5, S>(a, s);
        case 6: return quad_bwd_nv<6, S>(a, s);
        case 7: return quad_bwd_nv<7, S>(a, s);
        case 8: return quad_bwd_nv<8, S>(a, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

size_t v0_bwd_workspace_bytes(uint32_t rows, uint32_t cols, uint32_t n_vec) {
    return size_t((rows + 63) / 64) * cols * n_vec * sizeof(float);
}

hipError_t launch_v0_backward(const V0BwdArgs &a, hipStream_t s) {
    switch (a.n_vec) {
        case 1: return bwd_nv<1>(a, s);
        case 2: return bwd_nv<2>(a, s);
        case 3: return bwd_nv<3>(a, s);
        case 4: return bwd_nv<4>(a, s);
        case 5: return bwd_nv<5>(a, s);
        case 6: return bwd_nv<6>(a, s);
        case 7: return bwd_nv<7>(a, s);
        case 8: return bwd_nv<8>(a, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_quad_backward(const QuadBwdArgs &a, hipStream_t s) {
    if (a.streams == 32) return quad_bwd_s<32>(a, s);
    if (a.streams == 64) return quad_bwd_s<64>(a, s);
    return hipErrorInvalidValue;
}

}  // namespace mv
