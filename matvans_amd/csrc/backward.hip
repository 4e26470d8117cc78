// Backward of the compressed linear layer on gfx950, from the ".bin v0" streams and from
// the MVX1 quad-coder streams (32, 64 or 128 per row):
//
//   dx[v][c] = alpha[v] * sum_r W[r, c] * (s_w[r] * g[v][r])
//
// W is re-decoded from the same row-ordered rANS streams the forward reads; no dense
// copy is ever built.  The reference has no backward pass.  The block bodies (stream-major
// decode, fixed-order 64-row pair steps and row reductions, fixed-order row-block reduce)
// are in blk_bwd.h; this file holds the one-matrix launches: a grid of (row blocks,
// stream groups) units, then one reduce kernel.

#include <hip/hip_runtime.h>

#include <mutex>

#include "blk_bwd.h"
#include "launch.h"
#include "matvans.h"

// The partial sums of matvans_v1_decompress_backward_noalloc: a file-scope device array
// (one per GPU the library runs on, held until the process ends).  Host code never passes
// this name to a kernel or a copy: it is a host-side stand-in, not a device address.  The
// device address comes from hipGetSymbolAddress once per device (backward_scratch).
static __device__ __align__(256) float scratch[MATVANS_NOALLOC_SCRATCH_BYTES / sizeof(float)];

namespace mv {
namespace {

using blk::BwdJob;

// W waves (streams) per unit: 8, or 4 when the 8-wave grid would not fit on the GPU at
// once (see fits_at_once)
template <int NV, int W>
__global__ __launch_bounds__(64 * W) void v0_bwd_partial(BwdJob job) {
    __shared__ blk::V0BwdSmem<NV, W> sm;
    blk::v0_bwd_unit<NV, W>(job, blockIdx.x, blockIdx.y, gridDim.x, sm);
}

// LAY: 0 for coder 1 (granule layout), 1 for coder 2 (streams back to back)
template <int NV, int S, int MODE, int W, int LAY = 0>
__global__ __launch_bounds__(64 * W) void quad_bwd_partial(BwdJob job) {
    __shared__ blk::QuadBwdSmem<NV, S, MODE, blk::bwd_set_words<S, MODE>(), W> sm;
    blk::quad_bwd_unit<NV, S, MODE, blk::bwd_set_words<S, MODE>(), W, LAY>(
        job, blockIdx.x, blockIdx.y, gridDim.x, sm);
}

__global__ __launch_bounds__(64 * blk::kRedPartsMax) void bwd_reduce(BwdJob job, uint32_t nrb) {
    __shared__ float part[blk::kRedPartsMax][64];
    blk::reduce_tile(job, nrb, blockIdx.y, blockIdx.x, part);
}

hipError_t launch_reduce(const BwdJob &j, hipStream_t s) {
    const uint32_t nrb = (j.rows + 63) / 64;
    hipLaunchKernelGGL(bwd_reduce, dim3((j.cols + 63) / 64, j.n_vec), dim3(64 * blk::red_parts(nrb, j.n_vec)),
                       0, s, j, nrb);
    return hipGetLastError();
}

// Whether `units` blocks of `kernel` (`threads` each) are resident on the current GPU at
// once.  A grid that overflows by a fraction runs its last units in a nearly idle second
// round; units of 4 waves (half the LDS, twice the units) then fit in one round.  Every
// wave computes the same either way: the results are identical.
template <class K>
bool fits_at_once(K kernel, int threads, uint32_t units) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess)
        return true;
    return units <= uint32_t(cus) * uint32_t(per);
}

template <int NV>
hipError_t bwd_nv(const BwdJob &j, hipStream_t s) {
    const uint32_t nrb = (j.rows + 63) / 64;
    if (fits_at_once(v0_bwd_partial<NV, 8>, 512, nrb * 4u))
        hipLaunchKernelGGL((v0_bwd_partial<NV, 8>), dim3(nrb, 4), dim3(512), 0, s, j);
    else
        hipLaunchKernelGGL((v0_bwd_partial<NV, 4>), dim3(nrb, 8), dim3(256), 0, s, j);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_reduce(j, s);
}

// One table per row (MODE 1) keeps 8-wave units: a unit builds its 64 rows' tables
// whatever its wave count, so 4-wave units would build every table twice and hold as
// much LDS per block (4096x14336: 53 us with 8-wave units in two rounds, 118 us with
// 4-wave units).
template <int NV, int S, int MODE, int LAY>
hipError_t quad_bwd_nv(const BwdJob &j, hipStream_t s) {
    const uint32_t nrb = (j.rows + 63) / 64;
    if (MODE == 1 ||
        fits_at_once(quad_bwd_partial<NV, S, MODE, 8, LAY>, 512, nrb * uint32_t(S / 8)))
        hipLaunchKernelGGL((quad_bwd_partial<NV, S, MODE, 8, LAY>), dim3(nrb, S / 8), dim3(512),
                           0, s, j);
    else
        hipLaunchKernelGGL((quad_bwd_partial<NV, S, MODE, 4, LAY>), dim3(nrb, S / 4), dim3(256),
                           0, s, j);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_reduce(j, s);
}

template <int S, int MODE = 0, int LAY = 0>
hipError_t quad_bwd_s(const BwdJob &j, hipStream_t s) {
    switch (j.n_vec) {
        case 1: return quad_bwd_nv<1, S, MODE, LAY>(j, s);
        case 2: return quad_bwd_nv<2, S, MODE, LAY>(j, s);
        case 3: return quad_bwd_nv<3, S, MODE, LAY>(j, s);
        case 4: return quad_bwd_nv<4, S, MODE, LAY>(j, s);
        case 5: return quad_bwd_nv<5, S, MODE, LAY>(j, s);
        case 6: return quad_bwd_nv<6, S, MODE, LAY>(j, s);
        case 7: return quad_bwd_nv<7, S, MODE, LAY>(j, s);
        case 8: return quad_bwd_nv<8, S, MODE, LAY>(j, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t backward_scratch(float **ptr) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void *addr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = p;
    }
    *ptr = static_cast<float *>(addr[dev]);
    return hipSuccess;
}

size_t v0_bwd_workspace_bytes(uint32_t rows, uint32_t cols, uint32_t n_vec) {
    return size_t((rows + 63) / 64) * cols * n_vec * sizeof(float);
}

hipError_t launch_v0_backward(const V0BwdArgs &a, hipStream_t s) {
    BwdJob j{a.partial, a.dx, a.g, a.s_w, a.alpha, a.g_stride, a.n_vec, a.rows, a.cols,
             a.cursors, a.payload, a.payload_size, int32_t(a.min_value), uint32_t(a.G), a.cdf,
             a.ppf, nullptr};
    if (!bwd_partials_fit(a.rows, a.cols)) return hipErrorInvalidValue;
    switch (a.n_vec) {
        case 1: return bwd_nv<1>(j, s);
        case 2: return bwd_nv<2>(j, s);
        case 3: return bwd_nv<3>(j, s);
        case 4: return bwd_nv<4>(j, s);
        case 5: return bwd_nv<5>(j, s);
        case 6: return bwd_nv<6>(j, s);
        case 7: return bwd_nv<7>(j, s);
        case 8: return bwd_nv<8>(j, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_quad_backward(const QuadBwdArgs &a, hipStream_t s) {
    BwdJob j{a.partial, a.dx,       a.g,       a.s_w,           a.alpha,
             a.g_stride, a.n_vec,   a.rows,    a.cols,          a.cursors,
             a.payload,  a.payload_size, a.min_value, a.G,      nullptr,
             nullptr,    a.lut,     a.freq,    a.out_col_start, a.out_col_rec};
    if (!bwd_partials_fit(a.rows, a.cols)) return hipErrorInvalidValue;
    if (a.coder == 2) {   // streams back to back, one table per matrix
        if (a.table_mode != 0) return hipErrorInvalidValue;
        if (a.streams == 32) return quad_bwd_s<32, 0, 1>(j, s);
        if (a.streams == 64) return quad_bwd_s<64, 0, 1>(j, s);
        if (a.streams == 128) return quad_bwd_s<128, 0, 1>(j, s);
        return hipErrorInvalidValue;
    }
    if (a.table_mode == 1) return a.streams == 64 ? quad_bwd_s<64, 1>(j, s) : hipErrorInvalidValue;
    if (a.streams == 32) return quad_bwd_s<32>(j, s);
    if (a.streams == 64) return quad_bwd_s<64>(j, s);
    if (a.streams == 128) return quad_bwd_s<128>(j, s);
    return hipErrorInvalidValue;
}

}  // namespace mv
