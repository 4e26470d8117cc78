// Fused rANS decode x int8 matvec for the MVX1 layouts (S = 32, 64 or 128 streams per
// row, P = 8 or 12 bit probabilities, one table per matrix or one 8-bit table per row),
// gfx950.
//
// Mapping: every lane owns one stream.  S = 128: two waves of a block decode one row and
// add their sums through LDS; S = 64: a wave decodes one row; S = 32: a wave decodes two
// rows, one per 32-lane half (the reference layout, matvansWarpPerRow-kernel.cu:95-174).
// The word ring holds 32 words per lane for P = 8 and 64 for P = 12 (v1_stream.h), so a
// CU keeps up to 32 waves (P = 8) resident: the decode is a chain of dependent LDS
// lookups per stream, and throughput comes from the number of chains in flight.
// Stream l's 16-symbol group t covers columns 16*S*t + 16*l .. +15, so each lane
// multiplies its own 16 bytes of each vector per group:
// four symbols per v_dot4_i32_i8, one int32 accumulator per vector, then a DPP/shuffle
// reduction over the row's lanes.
//
// Tables in LDS: one table per matrix is copied from the device table prepared by
// matvans_v1_lut (2 KiB for P = 8, 32 KiB for P = 12) once per block; one table per row
// (table_mode 1, S = 64) is built by the row's wave from the row's 8-bit frequencies:
// an exclusive scan of the frequencies gives cdf, every present symbol marks its first
// quantile, and an inclusive max-scan over the quantiles gives ppf.  Byte tables
// f / bias / symbol (256 bytes each) keep every lookup conflict-free.
#include <hip/hip_runtime.h>

#include "launch.h"
#include "v1_stream.h"

namespace mv {
namespace {

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

template <int NV, int S, int P, int MODE>
__global__ __launch_bounds__(256) void v1_fwd(int32_t *__restrict__ dst,
                                              const int8_t *__restrict__ vecs,
                                              uint32_t vec_stride, uint32_t rows, uint32_t cols,
                                              const uint32_t *__restrict__ cursors,
                                              int min_value, uint32_t G,
                                              const uint32_t *__restrict__ glut,
                                              const uint8_t *__restrict__ freq,
                                              const uint16_t *__restrict__ payload,
                                              uint32_t payload_size) {
    constexpr int RPW = S < 64 ? 64 / S : 1;   // rows per wave
    constexpr int WPR = S > 64 ? S / 64 : 1;   // waves per row
    constexpr int GW = 256 / S;                // words per lane per granule
    constexpr int R = P == 8 ? 32 : 64;        // ring words per lane
    constexpr int kLut = MODE == 1 ? 1 : 2 << P;   // dwords of the shared table
    __shared__ uint16_t ring[kWaves * R * 64];
    __shared__ uint32_t lut[kLut];
    __shared__ uint32_t rowtab[MODE == 1 ? kWaves * 384 : 1];   // per wave: f|bias|sym + scratch
    __shared__ uint16_t dummy[8 * 64];
    __shared__ int32_t part[WPR > 1 ? kWaves * NV : 1];

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6, lane = tid & 63;
    // stream slot within the row, and the rows of this wave
    const uint32_t sl = S >= 64 ? (wave % WPR) * 64 + lane : lane & (S - 1);
    const uint32_t row0 = S >= 64 ? blockIdx.x * (kWaves / WPR) + wave / WPR
                                  : (blockIdx.x * kWaves + wave) * RPW;
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
        // one table per row; S == 64, so the wave owns exactly this row
        uint8_t *tab = reinterpret_cast<uint8_t *>(rowtab + wave * 384);
        uint8_t *mark = tab + 768, *sf = tab + 1024, *sc = tab + 1280;
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
            tab[q] = sf[s];
            tab[256 + q] = static_cast<uint8_t>(q - sc[s]);
            tab[512 + q] = static_cast<uint8_t>(min_value + int(s));
        }
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
            c[v] = *reinterpret_cast<const uint4 *>(vecs + size_t(v) * vec_stride +
                                                    tt * (16u * S) + sl * 16u);
    };
    uint4 vc0[NV], vc1[NV];
    vload(0, vc0);
    vload(1, vc1);
    __syncthreads();
    st.start();

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
                dst[size_t(v) * rows + row] = a;
            }
        }
    } else if ((lane & (S - 1)) == 0 && valid) {
#pragma unroll
        for (int v = 0; v < NV; v++) dst[size_t(v) * rows + row] = acc[v];
    }
}

// Quad coder (coder 1, P = 8, one table per matrix): same mapping, per-lane 64-bit
// state; a 16-symbol group is four quads, each one v_dot4 per vector.
template <int NV, int S>
__global__ __launch_bounds__(256) void v1q_fwd(int32_t *__restrict__ dst,
                                               const int8_t *__restrict__ vecs,
                                               uint32_t vec_stride, uint32_t rows, uint32_t cols,
                                               const uint32_t *__restrict__ cursors,
                                               const uint32_t *__restrict__ glut,
                                               const uint16_t *__restrict__ payload,
                                               uint32_t payload_size) {
    constexpr int RPW = S < 64 ? 64 / S : 1;
    constexpr int WPR = S > 64 ? S / 64 : 1;
    constexpr int G4 = 128 / S;
    using QL = v1::QuadLane<G4>;
    __shared__ uint32_t ring[kWaves * QL::R4 * 64];
    __shared__ uint32_t tab[256];
    __shared__ uint32_t dummy[4 * 64];
    __shared__ int32_t part[WPR > 1 ? kWaves * NV : 1];

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6, lane = tid & 63;
    const uint32_t sl = S >= 64 ? (wave % WPR) * 64 + lane : lane & (S - 1);
    const uint32_t row0 = S >= 64 ? blockIdx.x * (kWaves / WPR) + wave / WPR
                                  : (blockIdx.x * kWaves + wave) * RPW;
    const uint32_t row = row0 + (S >= 64 ? 0u : lane / S);
    const bool valid = row < rows;
    // packed table T[q] = f | r << 8 | sym << 24 from the A / B tables (A = f | sym << 24,
    // B = r); f <= 255 and r < f for every table the encoder accepts
    tab[tid] = (glut[tid] & 0xFF0000FFu) | ((glut[256 + tid] & 0xFFu) << 8);
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

    QL st;
    st.init(payload, wbeg, wend, valid, rcur, rend, sl, ring, dummy, wave, lane);
    st.prologue();

    const uint32_t groups = cols / (16u * S);
    auto vload = [&](uint32_t t, uint4 *c) {
        uint32_t tt = t < groups ? t : groups - 1;
#pragma unroll
        for (int v = 0; v < NV; v++)
            c[v] = *reinterpret_cast<const uint4 *>(vecs + size_t(v) * vec_stride +
                                                    tt * (16u * S) + sl * 16u);
    };
    uint4 vc0[NV], vc1[NV];
    vload(0, vc0);
    vload(1, vc1);
    __syncthreads();
    st.start();

    int32_t acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = 0;
    auto group = [&](const uint4 *vc) {
        uint32_t wbefore = st.waddr;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            uint32_t packed = st.quad(tab);
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
                dst[size_t(v) * rows + row] = a;
            }
        }
    } else if ((lane & (S - 1)) == 0 && valid) {
#pragma unroll
        for (int v = 0; v < NV; v++) dst[size_t(v) * rows + row] = acc[v];
    }
}

template <int NV, int S>
hipError_t launch_quad(const V1Args &a, hipStream_t s) {
    constexpr uint32_t rows_per_block = S > 64 ? kWaves / (S / 64) : kWaves * (64 / S);
    dim3 grid((a.rows + rows_per_block - 1) / rows_per_block), block(256);
    hipLaunchKernelGGL((v1q_fwd<NV, S>), grid, block, 0, s, a.dst, a.vecs, a.vec_stride, a.rows,
                       a.cols, a.cursors, a.lut, a.payload, a.payload_size);
    return hipGetLastError();
}

template <int NV, int S, int P, int MODE>
hipError_t launch_cfg(const V1Args &a, hipStream_t s) {
    constexpr uint32_t rows_per_block = S > 64 ? kWaves / (S / 64) : kWaves * (64 / S);
    dim3 grid((a.rows + rows_per_block - 1) / rows_per_block), block(256);
    hipLaunchKernelGGL((v1_fwd<NV, S, P, MODE>), grid, block, 0, s, a.dst, a.vecs, a.vec_stride,
                       a.rows, a.cols, a.cursors, int(a.min_value), a.G, a.lut, a.freq,
                       a.payload, a.payload_size);
    return hipGetLastError();
}

template <int NV>
hipError_t launch_v1_nv(const V1Args &a, hipStream_t s) {
    if (a.coder == 1) {
        if (a.precision != 8 || a.table_mode != 0) return hipErrorInvalidValue;
        if (a.streams == 128) return launch_quad<NV, 128>(a, s);
        if (a.streams == 64) return launch_quad<NV, 64>(a, s);
        return launch_quad<NV, 32>(a, s);
    }
    if (a.table_mode == 1) {
        if (a.streams == 64 && a.precision == 8) return launch_cfg<NV, 64, 8, 1>(a, s);
        return hipErrorInvalidValue;
    }
    if (a.streams == 128)
        return a.precision == 8 ? launch_cfg<NV, 128, 8, 0>(a, s) : launch_cfg<NV, 128, 12, 0>(a, s);
    if (a.streams == 64)
        return a.precision == 8 ? launch_cfg<NV, 64, 8, 0>(a, s) : launch_cfg<NV, 64, 12, 0>(a, s);
    return a.precision == 8 ? launch_cfg<NV, 32, 8, 0>(a, s) : launch_cfg<NV, 32, 12, 0>(a, s);
}

}  // namespace

hipError_t launch_v1_decode_matvec(const V1Args &a, hipStream_t s) {
    switch (a.n_vec) {
        case 1: return launch_v1_nv<1>(a, s);
        case 2: return launch_v1_nv<2>(a, s);
        case 3: return launch_v1_nv<3>(a, s);
        case 4: return launch_v1_nv<4>(a, s);
        case 5: return launch_v1_nv<5>(a, s);
        case 6: return launch_v1_nv<6>(a, s);
        case 7: return launch_v1_nv<7>(a, s);
        case 8: return launch_v1_nv<8>(a, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace mv
