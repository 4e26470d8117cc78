// Fused decode x int8 matvec for the reference's family-B layouts on gfx950
// (buffered/ans-warp-rows-kernel.cu:7-221, buffered/ans-warp-stripes-kernel.cu:21-227).
//
// The block body (codec, mapping, LDS ring and refill) is blk::b_block in blk_b.h.
#include <hip/hip_runtime.h>

#include "blk_b.h"
#include "launch.h"

namespace mv {
namespace {

template <int NV, int LAYOUT>
__global__ __launch_bounds__(256) void b_fwd(blk::BJob job) {
    __shared__ blk::BSmem sm;
    blk::b_block<NV, LAYOUT>(job, blockIdx.x, sm);
}

template <int NV, int LAYOUT>
hipError_t launch_b(const BArgs &a, hipStream_t s) {
    const uint32_t n_warps = LAYOUT == 0 ? a.out_dim : a.out_dim / 32u;
    const uint32_t warps_per_block = 2 * blk::kBWaves;
    dim3 grid((n_warps + warps_per_block - 1) / warps_per_block), block(256);
    blk::BJob j{a.acc32, a.out8, a.scale, a.compressed, a.compressed_words, a.warp_offsets,
                a.end_offset, a.ppf, a.ppf_cum, a.ppf_prob, a.vecs, a.vec_stride, a.n_vec,
                a.in_dim, a.out_dim};
    hipLaunchKernelGGL((b_fwd<NV, LAYOUT>), grid, block, 0, s, j);
    return hipGetLastError();
}

template <int LAYOUT>
hipError_t launch_layout(const BArgs &a, hipStream_t s) {
    switch (a.n_vec) {
        case 1: return launch_b<1, LAYOUT>(a, s);
        case 2: return launch_b<2, LAYOUT>(a, s);
        case 3: return launch_b<3, LAYOUT>(a, s);
        case 4: return launch_b<4, LAYOUT>(a, s);
        case 5: return launch_b<5, LAYOUT>(a, s);
        case 6: return launch_b<6, LAYOUT>(a, s);
        case 7: return launch_b<7, LAYOUT>(a, s);
        case 8: return launch_b<8, LAYOUT>(a, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_b_decode_matvec(const BArgs &a, hipStream_t s) {
    return a.layout == 0 ? launch_layout<0>(a, s) : launch_layout<1>(a, s);
}

}  // namespace mv
