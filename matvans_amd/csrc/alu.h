// VALU helpers for the decode steps on gfx950.  The table words carry several byte
// fields; these take them with SDWA operand selects (one instruction each) and keep the
// 24-bit multiplies that the fields need, which the compiler would otherwise rebuild
// from masks and full 32-bit multiplies.
#pragma once
#include <cstdint>

namespace mv {

// (a & 0xFFFFFF) * (b & 0xFFFFFF)
__device__ __forceinline__ uint32_t mul_u24(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// (a & 0xFFFFFF) * (b & 0xFFFFFF) + c
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
// (a & 0xFF) * (b & 0xFF)
__device__ __forceinline__ uint32_t mul_b0_b0(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_mul_u32_u24_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_0 "
        "src1_sel:BYTE_0" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// ((a >> 8) & 0xFF) * (b & 0xFF)
__device__ __forceinline__ uint32_t mul_b1_b0(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_mul_u32_u24_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_1 "
        "src1_sel:BYTE_0" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// (a & 0xFFFFFF) * (b & 0xFF)
__device__ __forceinline__ uint32_t mul_d_b0(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_mul_u32_u24_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD "
        "src1_sel:BYTE_0" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// a + ((b >> 8) & 0xFF)
__device__ __forceinline__ uint32_t add_b1(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD "
        "src1_sel:BYTE_1" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// float of the signed byte 0 / byte 2 of a (sign-extending SDWA source select)
__device__ __forceinline__ float cvt_f32_sbyte0(uint32_t a) {
    float r;
    asm("v_cvt_f32_i32_sdwa %0, sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_0"
        : "=v"(r) : "v"(a));
    return r;
}
__device__ __forceinline__ float cvt_f32_sbyte2(uint32_t a) {
    float r;
    asm("v_cvt_f32_i32_sdwa %0, sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_2"
        : "=v"(r) : "v"(a));
    return r;
}

// Cache-policy operand of the payload's buffer loads.  kLoadCached (0): the default.
// kLoadStream (2, the non-temporal bit): for payload bytes that one wave reads once, as
// the forward decoders' granule loads, whose lines no other wave wants (4096x14336 q32
// forward 19.36 -> 18.46 us, v0 32.6 -> 31.8 us, decoder block q32 at one vector 63.7 ->
// 61.1 us).  The backward keeps kLoadCached: its eight waves read neighbouring 16-byte
// slots of the same 128-byte lines, shared through the caches (with kLoadStream the q32
// backward took 36.3 us instead of 24.6).
constexpr int kLoadCached = 0;
constexpr int kLoadStream = 2;

}  // namespace mv
