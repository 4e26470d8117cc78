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
#include "blk_v1.h"

#ifndef MV_MFMA_MIN_VEC
#define MV_MFMA_MIN_VEC 4
#endif
#ifndef MV_MFMA_WAVES
#define MV_MFMA_WAVES 8
#endif

namespace mv {
namespace {

using blk::kWaves;
using blk::V1Job;

template <int NV, int S, int P, int MODE>
__global__ __launch_bounds__(256) void v1_fwd(V1Job job) {
    __shared__ blk::V1Smem<NV, S, P, MODE> sm;
    blk::v1_block<NV, S, P, MODE>(job, blockIdx.x, sm);
}

// ring words per lane per set of the one-matrix quad kernels: two groups per set for 32
// streams per row (halves the ring bookkeeping; 2 waves per SIMD at 4096 rows leave LDS
// to spare), one group per set for 64 and 128 streams (more resident waves per CU)
// Coder 2 reads each lane's own run of words, 8 per set whatever the stream count (16-word
// sets, a 64-byte line per lane and set, measured slower for 32 streams: 26.5 vs 24.6 us
// at 4096x14336).
template <int S, int LAY = 0>
constexpr int quad_set_words() { return LAY ? 8 : (S == 32 ? 8 : 4); }

// LAY: 0 for coder 1 (granule layout), 1 for coder 2 (streams back to back)
template <int NV, int S, int MODE, int LAY = 0>
__global__ __launch_bounds__(256) void v1q_fwd(V1Job job) {
    constexpr int W = quad_set_words<S, LAY>();
    __shared__ blk::QuadSmem<NV, S, MODE, W> sm;
    blk::quad_block<NV, S, MODE, W, LAY>(job, blockIdx.x, sm);
}

template <int NV, int MW, int LAY = 0>
__global__ __launch_bounds__(64 * MW) void v1q_mfma(V1Job job) {
    __shared__ blk::QuadMfmaSmem<MW> sm;
    blk::quad_block_mfma<NV, MW, LAY>(job, blockIdx.x, sm);
}

V1Job job_of(const V1Args &a) {
    return V1Job{a.dst,       a.vecs,    a.vec_stride, a.n_vec,        a.rows,    a.cols,
                 a.cursors,   a.min_value, a.G,        a.lut,          a.freq,    a.payload,
                 a.payload_size, nullptr, 0.0f,        a.out_start,    a.out_rec};
}

template <int NV, int S, int MODE = 0, int LAY = 0>
hipError_t launch_quad(const V1Args &a, hipStream_t s) {
    constexpr uint32_t rows_per_block = S > 64 ? kWaves / (S / 64) : kWaves * (64 / S);
    dim3 grid((a.rows + rows_per_block - 1) / rows_per_block), block(256);
    hipLaunchKernelGGL((v1q_fwd<NV, S, MODE, LAY>), grid, block, 0, s, job_of(a));
    return hipGetLastError();
}

template <int NV, int LAY>
hipError_t launch_mfma(const V1Args &a, hipStream_t s) {
    constexpr int MW = MV_MFMA_WAVES;
    dim3 grid((a.rows + 2 * MW - 1) / (2 * MW)), block(64 * MW);
    hipLaunchKernelGGL((v1q_mfma<NV, MW, LAY>), grid, block, 0, s, job_of(a));
    return hipGetLastError();
}

template <int NV, int S, int P, int MODE>
hipError_t launch_cfg(const V1Args &a, hipStream_t s) {
    constexpr uint32_t rows_per_block = S > 64 ? kWaves / (S / 64) : kWaves * (64 / S);
    dim3 grid((a.rows + rows_per_block - 1) / rows_per_block), block(256);
    hipLaunchKernelGGL((v1_fwd<NV, S, P, MODE>), grid, block, 0, s, job_of(a));
    return hipGetLastError();
}

// From this many vectors on, the quad coder with 32 streams per row and one table takes
// its products on the matrix cores (quad_block_mfma).
constexpr uint32_t kMfmaMinVec = MV_MFMA_MIN_VEC;

template <int NV>
hipError_t launch_v1_nv(const V1Args &a, hipStream_t s) {
    if (a.coder == 2) {   // quad streams back to back: one table per matrix
        if (a.precision != 8 || a.table_mode != 0) return hipErrorInvalidValue;
        if (a.streams == 32 && a.n_vec >= kMfmaMinVec) return launch_mfma<NV, 1>(a, s);
        if (a.streams == 128) return launch_quad<NV, 128, 0, 1>(a, s);
        if (a.streams == 64) return launch_quad<NV, 64, 0, 1>(a, s);
        return launch_quad<NV, 32, 0, 1>(a, s);
    }
    if (a.coder == 1) {
        if (a.precision != 8) return hipErrorInvalidValue;
        if (a.table_mode == 0 && a.streams == 32 && a.n_vec >= kMfmaMinVec)
            return launch_mfma<NV, 0>(a, s);
        if (a.table_mode == 1)
            return a.streams == 64 ? launch_quad<NV, 64, 1>(a, s) : hipErrorInvalidValue;
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
