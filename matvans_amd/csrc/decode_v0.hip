// Fused rANS decode x int8 matvec for the WarpPerRow ".bin v0" layout, gfx950.
//
// Layout (create_matrices_WarpPerRow.py:353-430): row r owns K_r granules of 512 bytes
// starting at u16 word cursors[r]; granule g holds, for each of the 32 streams
// ("lanes") of the row, 8 consecutive u16 words of that stream at byte 16*lane.
// Stream l of row r decodes columns 512*(k/16) + 16*l + k%16 for k = 0 .. cols/32-1
// (matvansWarpPerRow-kernel.cu:111-133).
//
// Mapping on CDNA4: one wave64 decodes two rows, one 32-lane half per row, so every
// lane owns exactly one stream.  Per lane:
//   * decode step (kernel.cu:120-136): one LDS read of a fused 32-bit table entry
//     {f, symbol, q - cdf[s]} indexed by q = x & 0xFF, then x = (x >> 8) * f + bias;
//   * renormalisation (kernel.cu:154-157): the next word of the stream is already in
//     a register (read from an LDS ring one step ahead); x = x<<16 | word when
//     x < 2^16, and the ring read pointer advances;
//   * every 4 symbols are packed into one dword with v_perm and multiplied into the
//     int32 accumulators with v_dot4_i32_i8 (one per activation vector).
// The ring (64 words per lane, u16, lane-interleaved so that the 64 lanes of a wave hit
// 64 different 2-byte slots of one 128-byte row) is filled from HBM with 128-bit
// buffer loads, two 16-step groups ahead of use.  A load is issued only when the ring
// will have room for the granule when it is written (every read pointer only grows),
// otherwise an out-of-range buffer offset is used, which returns zeros without
// touching memory, so that every group issues the same number of loads.
#include <hip/hip_runtime.h>

#include "launch.h"

namespace mv {
namespace {

constexpr int kWaves = 4;                 // waves per block
constexpr int kRowsPerBlock = 2 * kWaves; // two rows per wave
constexpr int kRing = 64;                 // ring words per lane (power of two, >= 48)
constexpr uint32_t kOOB = 0x80000000u;    // buffer offset beyond num_records

__device__ __forceinline__ uint32_t lut_entry(uint32_t q, int min_value, uint32_t G,
                                              const uint8_t *__restrict__ cdf,
                                              const uint8_t *__restrict__ ppf) {
    uint32_t s = ppf[q];
    uint32_t c0 = s <= G ? cdf[s] : 0u;
    uint32_t c1 = s + 1 <= G ? cdf[s + 1] : 0u;
    uint32_t f = (c1 - c0) & 0xFFu;                       // kernel.cu:135
    uint32_t sym = static_cast<uint8_t>(min_value + static_cast<int>(s));  // :126 (int8)
    uint32_t bias = (q - c0) & 0xFFu;                     // :136
    return f | (sym << 8) | (bias << 16);
}

template <int NV>
struct VecChunk {
    uint4 v[NV];
};

__device__ __forceinline__ uint4 bload(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    auto d = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    return make_uint4(d[0], d[1], d[2], d[3]);
}

template <int NV>
__global__ __launch_bounds__(256) void v0_decode_matvec(
    int32_t *__restrict__ dst, const int8_t *__restrict__ vecs, uint32_t vec_stride,
    uint32_t rows, uint32_t cols, const uint32_t *__restrict__ cursors, int min_value,
    uint32_t G, const uint8_t *__restrict__ cdf, const uint8_t *__restrict__ ppf,
    const uint16_t *__restrict__ payload, uint32_t payload_size) {
    __shared__ uint32_t lut[256];
    __shared__ uint16_t ring[kWaves][kRing][64];

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6, lane = tid & 63, sl = lane & 31;
    const uint32_t row0 = blockIdx.x * kRowsPerBlock + wave * 2;   // wave-uniform
    const uint32_t row = row0 + (lane >> 5);

    lut[tid] = lut_entry(tid, min_value, G, cdf, ppf);

    // Byte window of the wave's two rows: [cursors[row0], end of row0+1).
    uint32_t wbeg = 0, wend = 0;
    if (row0 < rows) {
        wbeg = cursors[row0];
        uint32_t last = row0 + 1 < rows ? row0 + 1 : row0;
        wend = last + 1 < rows ? cursors[last + 1] : payload_size;
        if (wend > payload_size) wend = payload_size;
        if (wend < wbeg) wend = wbeg;
    }
    wbeg = __builtin_amdgcn_readfirstlane(wbeg);
    wend = __builtin_amdgcn_readfirstlane(wend);
    // this lane's row extent in granules
    uint32_t K = 0, roff = 0;
    if (row < rows) {
        uint32_t c = cursors[row];
        uint32_t e = row + 1 < rows ? cursors[row + 1] : payload_size;
        if (e > wend) e = wend;
        if ((c & 255u) == 0 && c >= wbeg && e > c) {
            K = (e - c) >> 8;
            roff = (c - wbeg) * 2u;
        }
    }
    __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint16_t *>(payload) + wbeg, 0, static_cast<int>((wend - wbeg) * 2u),
        0x00020000);
    const uint32_t lane_off = roff + sl * 16u;   // + 512*g for granule g
    auto goff = [&](uint32_t g) { return g < K ? lane_off + g * 512u : kOOB; };

    uint16_t *myring = &ring[wave][0][lane];
    auto write_granule = [&](uint4 d, uint32_t g) {
        uint32_t s = (g * 8u) & (kRing - 1);
        myring[(s + 0) * 64] = static_cast<uint16_t>(d.x);
        myring[(s + 1) * 64] = static_cast<uint16_t>(d.x >> 16);
        myring[(s + 2) * 64] = static_cast<uint16_t>(d.y);
        myring[(s + 3) * 64] = static_cast<uint16_t>(d.y >> 16);
        myring[(s + 4) * 64] = static_cast<uint16_t>(d.z);
        myring[(s + 5) * 64] = static_cast<uint16_t>(d.z >> 16);
        myring[(s + 6) * 64] = static_cast<uint16_t>(d.w);
        myring[(s + 7) * 64] = static_cast<uint16_t>(d.w >> 16);
    };

    // Prologue: the first 6 granules (48 words) go straight into the ring.
    constexpr uint32_t kPre = 6;
    uint4 pre[kPre];
#pragma unroll
    for (uint32_t g = 0; g < kPre; g++) pre[g] = bload(rsrc, goff(g));
#pragma unroll
    for (uint32_t g = 0; g < kPre; g++)
        if (g < K) write_granule(pre[g], g);
    uint32_t gl = K < kPre ? K : kPre;   // next granule to load

    // vector chunks: group t needs bytes [512t + 16*sl, +16) of each vector
    const uint32_t groups = cols / 512u;
    auto vload = [&](uint32_t t, VecChunk<NV> &c) {
        uint32_t tt = t < groups ? t : groups - 1;
#pragma unroll
        for (int v = 0; v < NV; v++)
            c.v[v] = *reinterpret_cast<const uint4 *>(vecs + size_t(v) * vec_stride +
                                                      tt * 512u + sl * 16u);
    };
    VecChunk<NV> vc0, vc1;
    vload(0, vc0);
    vload(1, vc1);

    __syncthreads();

    uint32_t x = (uint32_t(myring[0]) << 16) | myring[64];
    uint32_t j = 2;                       // next unread word of the stream
    uint32_t W = myring[2 * 64];          // that word
    int32_t acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = 0;

    // two pending granule loads in flight, alternating A / B
    uint4 pa, pb;
    uint32_t ga = 0xFFFFFFFFu, gb = 0xFFFFFFFFu;
    auto issue = [&](uint4 &p, uint32_t &gp) {
        bool ok = gl < K && (gl + 1) * 8u - j <= uint32_t(kRing);
        p = bload(rsrc, ok ? lane_off + gl * 512u : kOOB);
        gp = ok ? gl : 0xFFFFFFFFu;
        gl += ok ? 1u : 0u;
    };
    issue(pa, ga);
    issue(pb, gb);

    auto group = [&](const VecChunk<NV> &vc) {
#pragma unroll
        for (int d = 0; d < 4; d++) {
            uint32_t e[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                uint32_t ent = lut[x & 0xFFu];
                uint32_t x1 = (x >> 8) * (ent & 0xFFu) + ((ent >> 16) & 0xFFu);
                bool c = x1 < 0x10000u;
                x = c ? ((x1 << 16) | W) : x1;
                j += c ? 1u : 0u;
                W = myring[(j & (kRing - 1)) * 64];
                e[u] = ent;
            }
            uint32_t packed = __builtin_amdgcn_perm(e[1], e[0], 0x0c0c0501u) |
                              __builtin_amdgcn_perm(e[3], e[2], 0x05010c0cu);
#pragma unroll
            for (int v = 0; v < NV; v++) {
                uint32_t xv = d == 0 ? vc.v[v].x : d == 1 ? vc.v[v].y : d == 2 ? vc.v[v].z
                                                                             : vc.v[v].w;
                acc[v] = __builtin_amdgcn_sdot4(static_cast<int>(packed), static_cast<int>(xv),
                                                acc[v], false);
            }
        }
    };

    // Main loop, two groups per iteration so that the A/B registers stay static.
    for (uint32_t t = 0; t < groups; t += 2) {
        // -- group t
        if (ga != 0xFFFFFFFFu) write_granule(pa, ga);
        issue(pa, ga);
        {
            VecChunk<NV> cur = vc0;
            vload(t + 2, vc0);
            group(cur);
        }
        if (t + 1 >= groups) break;
        // -- group t + 1
        if (gb != 0xFFFFFFFFu) write_granule(pb, gb);
        issue(pb, gb);
        {
            VecChunk<NV> cur = vc1;
            vload(t + 3, vc1);
            group(cur);
        }
    }

    // reduce each 32-lane half
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
    if (sl == 0 && row < rows) {
#pragma unroll
        for (int v = 0; v < NV; v++) dst[size_t(v) * rows + row] = acc[v];
    }
}

}  // namespace

template <int NV>
static hipError_t launch_nv(const V0Args &a, hipStream_t s) {
    dim3 grid((a.rows + kRowsPerBlock - 1) / kRowsPerBlock), block(256);
    hipLaunchKernelGGL(v0_decode_matvec<NV>, grid, block, 0, s, a.dst, a.vecs, a.vec_stride,
                       a.rows, a.cols, a.cursors, int(a.min_value), uint32_t(a.G), a.cdf, a.ppf,
                       a.payload, a.payload_size);
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
