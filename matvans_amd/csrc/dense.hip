// Same-GPU comparators and epilogues.
//
//  int8_matvec   replaces vecMat_int8_warpRows (baselines/baseline_int8/
//                baseline_matVec_custom.cu:99-155): uncompressed row-major int8 W, int32
//                dot with v_dot4_i32_i8, 1..8 vectors per pass over W.
//  f32_gemv      replaces cublasSgemv (baselines/cuBLAS_baseline_float/
//                baseline_matVec_cuBLAS.cu:111-144) on the row-major fp32 W.
//  absmax_requant  absMaxWithThrustDevice + normalizeAndRoundtoInt8
//                (matvansWarpPerRow/matvansWarpPerRow.cu:51-72), in one stream-ordered
//                launch that leaves delta on the device.
//  f32_normalize_round  absMaxWithThrustDevice + normalizeAndRound of the fp32 chain
//                (baselines/cuBLAS_baseline_float/baseline_matVec_cuBLAS.cu:48-64,129-135):
//                y = roundf(y / (max|y| / 127)) in place, delta left on the device.
//  scale_output  y = s_x * s_w[r] * float(acc).
//
// Both matvecs are HBM-bound: each wave streams whole rows with 16-byte nontemporal
// loads (four in flight per lane and row) while the activation vectors sit in LDS (in
// the caches beyond 64 KiB; the int8 matvec stages up to 160 KiB in persistent 16-wave
// blocks); the int8 matvec streams one or two rows per wave.
#include <hip/hip_runtime.h>

#include "launch.h"

namespace mv {
namespace {

constexpr int kThreads = 256;
constexpr int kRowsPerBlock = 8;   // 2 rows per wave, 4 waves
constexpr int kInt8Loads = 4;      // int8 matvec: 16-byte loads in flight per lane and row
constexpr uint32_t kMaxLdsBytes = 64 * 1024;
constexpr uint32_t kMaxBigLdsBytes = 160 * 1024;   // the LDS of one gfx950 CU
constexpr uint32_t kBigMinLds = 16 * 1024;   // int8 matvec: persistent 16-wave blocks above

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// 16-byte streaming loads (the matrix is read exactly once)
__device__ __forceinline__ uint4 ntload(const uint4 *p) {
    u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
    return make_uint4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ float4 ntload(const float4 *p) {
    f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(p));
    return make_float4(v[0], v[1], v[2], v[3]);
}

__device__ __forceinline__ int32_t wave_sum(int32_t a) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    return a;
}
__device__ __forceinline__ float wave_sumf(float a) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    return a;
}

// R rows per wave in lockstep: each 16-byte chunk of the vectors is read once for R rows
// (vector traffic per matrix byte NV / R); each lane keeps kInt8Loads chunks of each row
// in flight.
template <int NV, int R, bool LDS, int WPB = 4>
__global__ __launch_bounds__(64 * WPB) void int8_matvec_kernel(
    int32_t *__restrict__ dst, const int8_t *__restrict__ mat, const int8_t *__restrict__ vecs,
    uint32_t vec_stride, uint32_t rows, uint32_t cols) {
    extern __shared__ uint4 svec[];   // [NV][cols/16] when LDS
    const uint32_t nch = cols / 16;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (LDS) {
        for (uint32_t v = 0; v < NV; v++)
            for (uint32_t c = tid; c < nch; c += 64 * WPB)
                svec[v * nch + c] =
                    *reinterpret_cast<const uint4 *>(vecs + size_t(v) * vec_stride + c * 16);
        __syncthreads();
    }
    auto vchunk = [&](uint32_t v, uint32_t c) -> uint4 {
        if (LDS) return svec[v * nch + c];
        return *reinterpret_cast<const uint4 *>(vecs + size_t(v) * vec_stride + c * 16);
    };
    for (uint32_t row0 = (blockIdx.x * WPB + wave) * R; row0 < rows;
         row0 += gridDim.x * WPB * R) {
        const uint4 *m[R];
        bool ok[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            ok[r] = row0 + r < rows;
            m[r] = reinterpret_cast<const uint4 *>(mat + size_t(ok[r] ? row0 + r : row0) * cols);
        }
        int32_t acc[R][NV];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int v = 0; v < NV; v++) acc[r][v] = 0;
        for (uint32_t c0 = lane; c0 < nch; c0 += 64 * kInt8Loads) {
            uint4 a[kInt8Loads][R];
#pragma unroll
            for (int u = 0; u < kInt8Loads; u++) {
                const uint32_t c = c0 + 64 * u;
#pragma unroll
                for (int r = 0; r < R; r++)
                    a[u][r] = c < nch && ok[r] ? ntload(m[r] + c) : make_uint4(0, 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < kInt8Loads; u++) {
                const uint32_t c = c0 + 64 * u;
                if (c >= nch) break;
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    const uint4 b = vchunk(v, c);
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        int32_t t = acc[r][v];
                        t = __builtin_amdgcn_sdot4(int(a[u][r].x), int(b.x), t, false);
                        t = __builtin_amdgcn_sdot4(int(a[u][r].y), int(b.y), t, false);
                        t = __builtin_amdgcn_sdot4(int(a[u][r].z), int(b.z), t, false);
                        acc[r][v] = __builtin_amdgcn_sdot4(int(a[u][r].w), int(b.w), t, false);
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int v = 0; v < NV; v++) {
                const int32_t t = wave_sum(acc[r][v]);
                if (lane == 0 && ok[r]) dst[size_t(v) * rows + row0 + r] = t;
            }
    }
}

template <bool LDS>
__global__ __launch_bounds__(kThreads) void f32_gemv_kernel(float *__restrict__ y,
                                                            const float *__restrict__ mat,
                                                            const float *__restrict__ x,
                                                            uint32_t rows, uint32_t cols,
                                                            float alpha) {
    extern __shared__ float4 sx[];
    const uint32_t nch = cols / 4;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (LDS) {
        for (uint32_t c = tid; c < nch; c += kThreads)
            sx[c] = reinterpret_cast<const float4 *>(x)[c];
        __syncthreads();
    }
    const uint32_t rbeg = blockIdx.x * kRowsPerBlock;
    for (uint32_t row = rbeg + wave; row < rows && row < rbeg + kRowsPerBlock; row += 4) {
        const float4 *m = reinterpret_cast<const float4 *>(mat + size_t(row) * cols);
        float acc = 0.f;
        for (uint32_t c0 = lane; c0 < nch; c0 += 256) {
            float4 a[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                uint32_t c = c0 + 64 * u;
                a[u] = c < nch ? ntload(m + c) : make_float4(0, 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                uint32_t c = c0 + 64 * u;
                if (c >= nch) break;
                float4 b = LDS ? sx[c] : reinterpret_cast<const float4 *>(x)[c];
                acc = fmaf(a[u].x, b.x, acc);
                acc = fmaf(a[u].y, b.y, acc);
                acc = fmaf(a[u].z, b.z, acc);
                acc = fmaf(a[u].w, b.w, acc);
            }
        }
        float s = wave_sumf(acc);
        if (lane == 0) y[row] = alpha * s;
    }
}

// One block: block-wide max of |acc|, then the requantisation of every element.
__global__ __launch_bounds__(1024) void absmax_requant_kernel(const int32_t *__restrict__ acc,
                                                              uint32_t n, int8_t *__restrict__ out8,
                                                              float *__restrict__ delta,
                                                              int clamp) {
    __shared__ uint32_t red[16];
    uint32_t m = 0;
    // four loads in flight per thread; an index past n re-reads element i (max unchanged)
    for (uint32_t i = threadIdx.x; i < n; i += 4 * blockDim.x) {
        int32_t a[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t j = i + k * blockDim.x;
            a[k] = acc[j < n ? j : i];
        }
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t u = a[k] < 0 ? 0u - uint32_t(a[k]) : uint32_t(a[k]);
            m = u > m ? u : m;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t t = __shfl_xor(m, o, 64);
        m = t > m ? t : m;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t t = threadIdx.x < (blockDim.x >> 6) ? red[threadIdx.x] : 0u;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            uint32_t s = __shfl_xor(t, o, 64);
            t = s > t ? s : t;
        }
        if (threadIdx.x == 0) red[0] = t;
    }
    __syncthreads();
    // thrust::transform_reduce of float(abs(x)) with an integer max: float(max |acc|)
    const float a = static_cast<float>(red[0]);
    const float d = a / 127.0f;   // IEEE fp32 division (v_delta = abs_max / 127)
    if (threadIdx.x == 0) *delta = d;
    for (uint32_t i = threadIdx.x; i < n; i += 4 * blockDim.x) {
        int32_t a[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t j = i + k * blockDim.x;
            a[k] = acc[j < n ? j : i];
        }
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t j = i + k * blockDim.x;
            float q = rintf(static_cast<float>(a[k]) / d);   // __float2int_rn(afl / scalar)
            int32_t v = (q == q) ? static_cast<int32_t>(fminf(fmaxf(q, -2147483648.0f),
                                                              2147483520.0f))
                                 : 0;
            if (clamp) v = v < -128 ? -128 : (v > 127 ? 127 : v);
            if (j < n) out8[j] = static_cast<int8_t>(v);
        }
    }
}

// One block: block-wide max of |y| (fp32; the order of a max does not matter), delta =
// max / 127, then y[i] = roundf(y[i] / delta) (IEEE division, half away from zero).
__global__ __launch_bounds__(1024) void f32_normalize_round_kernel(float *__restrict__ y,
                                                                   uint32_t n,
                                                                   float *__restrict__ delta) {
    __shared__ float red[16];
    float m = 0.0f;
    // four loads in flight per thread; an index past n re-reads element i (max unchanged)
    for (uint32_t i = threadIdx.x; i < n; i += 4 * blockDim.x) {
        float a[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t j = i + k * blockDim.x;
            a[k] = fabsf(y[j < n ? j : i]);
        }
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) m = m < a[k] ? a[k] : m;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float t = __shfl_xor(m, o, 64);
        m = m < t ? t : m;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x < 64) {
        float t = threadIdx.x < (blockDim.x >> 6) ? red[threadIdx.x] : 0.0f;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float u = __shfl_xor(t, o, 64);
            t = t < u ? u : t;
        }
        if (threadIdx.x == 0) red[0] = t;
    }
    __syncthreads();
    const float d = red[0] / 127.0f;
    if (threadIdx.x == 0) *delta = d;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) y[i] = roundf(y[i] / d);
}

__global__ void scale_output_kernel(const int32_t *__restrict__ acc, uint32_t rows,
                                    uint32_t n_vec, const float *__restrict__ s_w,
                                    const float *__restrict__ s_x, float *__restrict__ y) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * n_vec) return;
    uint32_t v = i / rows, r = i - v * rows;
    y[i] = (s_x[v] * s_w[r]) * static_cast<float>(acc[i]);
}

// One block per vector: s_x = max|x| / 127 (1 when x == 0), xq = clamp(rint(x / s_x)).
__global__ __launch_bounds__(1024) void quantize_vectors_kernel(const float *__restrict__ x,
                                                                uint32_t x_stride, uint32_t cols,
                                                                int8_t *__restrict__ xq,
                                                                uint32_t xq_stride,
                                                                float *__restrict__ s_x) {
    __shared__ float red[16];
    const float *xv = x + size_t(blockIdx.x) * x_stride;
    float m = 0.0f;
    for (uint32_t i = threadIdx.x; i < cols; i += blockDim.x) m = fmaxf(m, fabsf(xv[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x < 64) {
        float t = threadIdx.x < (blockDim.x >> 6) ? red[threadIdx.x] : 0.0f;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t = fmaxf(t, __shfl_xor(t, o, 64));
        if (threadIdx.x == 0) red[0] = t;
    }
    __syncthreads();
    const float sx = red[0] > 0.0f ? red[0] / 127.0f : 1.0f;
    if (threadIdx.x == 0) s_x[blockIdx.x] = sx;
    int8_t *q = xq + size_t(blockIdx.x) * xq_stride;
    for (uint32_t i = threadIdx.x; i < cols; i += blockDim.x)
        q[i] = static_cast<int8_t>(fminf(fmaxf(rintf(xv[i] / sx), -127.0f), 127.0f));
}

template <int NV, int R>
hipError_t int8_nv_r(int32_t *dst, const int8_t *mat, const int8_t *vecs, uint32_t vec_stride,
                     uint32_t rows, uint32_t cols, hipStream_t s) {
    dim3 grid((rows + 4 * R - 1) / (4 * R)), block(kThreads);
    uint32_t lds = NV * cols;
    if (lds <= kMaxLdsBytes)
        hipLaunchKernelGGL((int8_matvec_kernel<NV, R, true>), grid, block, lds, s, dst, mat,
                           vecs, vec_stride, rows, cols);
    else
        hipLaunchKernelGGL((int8_matvec_kernel<NV, R, false>), grid, block, 0, s, dst, mat,
                           vecs, vec_stride, rows, cols);
    return hipGetLastError();
}

// Several vectors of 16-160 KiB: persistent 16-wave blocks, as many as the CUs hold at
// once (two per CU while two stages fit in the LDS), stage the vectors once and loop
// over the rows, one row per wave, so the stage is not re-read for every few rows.
// Otherwise 4-wave blocks: two rows per wave once there are enough rows to keep the CUs
// busy with half the waves (>= 8192) or once the vectors outgrow 32 KiB (past the LDS
// their traffic per matrix byte halves), one row per wave otherwise.  Measured at
// 4096 x 4096, 14336 x 4096 and 4096 x 14336 with 1, 4 and 8 vectors, this picks the
// fastest of the variants tried in every case.
template <int NV>
hipError_t int8_nv(int32_t *dst, const int8_t *mat, const int8_t *vecs, uint32_t vec_stride,
                   uint32_t rows, uint32_t cols, uint32_t cus, hipStream_t s) {
    const uint32_t lds = NV * cols;
    if (NV >= 2 && lds > kBigMinLds && lds <= kMaxBigLdsBytes) {
        const uint32_t per_cu = lds * 2 <= kMaxBigLdsBytes ? 2u : 1u;
        uint32_t g = (rows + 15) / 16;
        g = g < cus * per_cu ? g : cus * per_cu;
        hipLaunchKernelGGL((int8_matvec_kernel<NV, 1, true, 16>), dim3(g), dim3(1024), lds, s,
                           dst, mat, vecs, vec_stride, rows, cols);
        return hipGetLastError();
    }
    if (rows >= 8192 || uint64_t(NV) * cols > 32768u)
        return int8_nv_r<NV, 2>(dst, mat, vecs, vec_stride, rows, cols, s);
    return int8_nv_r<NV, 1>(dst, mat, vecs, vec_stride, rows, cols, s);
}

}  // namespace

hipError_t launch_int8_matvec(int32_t *dst, const int8_t *mat, const int8_t *vecs,
                              uint32_t n_vec, uint32_t vec_stride, uint32_t rows, uint32_t cols,
                              uint32_t cus, hipStream_t s) {
    switch (n_vec) {
        case 1: return int8_nv<1>(dst, mat, vecs, vec_stride, rows, cols, cus, s);
        case 2: return int8_nv<2>(dst, mat, vecs, vec_stride, rows, cols, cus, s);
        case 3: return int8_nv<3>(dst, mat, vecs, vec_stride, rows, cols, cus, s);
        case 4: return int8_nv<4>(dst, mat, vecs, vec_stride, rows, cols, cus, s);
        case 5: return int8_nv<5>(dst, mat, vecs, vec_stride, rows, cols, cus, s);
        case 6: return int8_nv<6>(dst, mat, vecs, vec_stride, rows, cols, cus, s);
        case 7: return int8_nv<7>(dst, mat, vecs, vec_stride, rows, cols, cus, s);
        case 8: return int8_nv<8>(dst, mat, vecs, vec_stride, rows, cols, cus, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_f32_gemv(float *y, const float *mat, const float *x, uint32_t rows,
                           uint32_t cols, float alpha, hipStream_t s) {
    dim3 grid((rows + kRowsPerBlock - 1) / kRowsPerBlock), block(kThreads);
    uint32_t lds = cols * 4u;
    if (lds <= kMaxLdsBytes)
        hipLaunchKernelGGL((f32_gemv_kernel<true>), grid, block, lds, s, y, mat, x, rows, cols,
                           alpha);
    else
        hipLaunchKernelGGL((f32_gemv_kernel<false>), grid, block, 0, s, y, mat, x, rows, cols,
                           alpha);
    return hipGetLastError();
}

hipError_t launch_absmax_requant(const int32_t *acc, uint32_t n, int8_t *out8, float *delta,
                                 int clamp, hipStream_t s) {
    hipLaunchKernelGGL(absmax_requant_kernel, dim3(1), dim3(1024), 0, s, acc, n, out8, delta,
                       clamp);
    return hipGetLastError();
}

hipError_t launch_f32_normalize_round(float *y, uint32_t n, float *delta, hipStream_t s) {
    hipLaunchKernelGGL(f32_normalize_round_kernel, dim3(1), dim3(1024), 0, s, y, n, delta);
    return hipGetLastError();
}

hipError_t launch_quantize_vectors(const float *x, uint32_t n_vec, uint32_t x_stride,
                                   uint32_t cols, int8_t *xq, uint32_t xq_stride, float *s_x,
                                   hipStream_t s) {
    hipLaunchKernelGGL(quantize_vectors_kernel, dim3(n_vec), dim3(1024), 0, s, x, x_stride, cols,
                       xq, xq_stride, s_x);
    return hipGetLastError();
}

hipError_t launch_scale_output(const int32_t *acc, uint32_t rows, uint32_t n_vec,
                               const float *s_w, const float *s_x, float *y, hipStream_t s) {
    uint32_t n = rows * n_vec;
    hipLaunchKernelGGL(scale_output_kernel, dim3((n + 255) / 256), dim3(256), 0, s, acc, rows,
                       n_vec, s_w, s_x, y);
    return hipGetLastError();
}

}  // namespace mv
