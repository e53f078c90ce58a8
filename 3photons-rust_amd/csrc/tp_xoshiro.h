// tp_xoshiro.h -- xoshiro256+ on the device (src/random/standard.rs via
// rand_xoshiro 0.6.0 and rand 0.8.5's Standard f64), and GF(2) skip-ahead by
// a 256x256 bit matrix (256 columns of 4 words, uniform across the wavefront).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tpd {

// rotl(x, 45) = rotr(x, 19) and x << 17 on the two 32-bit halves with
// v_alignbit_b32 (a 64-bit rotate would otherwise take a 64-bit shift plus
// two 32-bit operations).
__device__ __forceinline__ uint64_t pack64(uint32_t lo, uint32_t hi) { return ((uint64_t)hi << 32) | lo; }
__device__ __forceinline__ uint64_t rotl45(uint64_t x)
{
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    return pack64(__builtin_amdgcn_alignbit(hi, lo, 19), __builtin_amdgcn_alignbit(lo, hi, 19));
}
__device__ __forceinline__ uint64_t shl17(uint64_t x)
{
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    return pack64(lo << 17, __builtin_amdgcn_alignbit(hi, lo, 15));
}

// next_u64: output s0 + s3 before the state update.
__device__ __forceinline__ uint64_t xo_next(uint64_t s[4])
{
    const uint64_t r = s[0] + s[3];
    const uint64_t t = shl17(s[1]);
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl45(s[3]);
    return r;
}

// rand's Standard f64: (u64 >> 11) * 2^-53, evaluated as
// hi * 2^-32 + (lo >> 11) * 2^-53: both terms and their sum are exact, so the
// fma returns the same double as the reference's conversion.
__device__ __forceinline__ double xo_uniform(uint64_t s[4])
{
    const uint64_t x = xo_next(s);
    const uint32_t hi = (uint32_t)(x >> 32), lo = (uint32_t)x;
    return fma((double)hi, 0x1.0p-32, (double)(lo >> 11) * 0x1.0p-53);
}

// The same uniform times 2^53, i.e. the integer u64 >> 11 as an exact double
// (hi * 2^21 + (lo >> 11) < 2^53), for code that folds the 2^-53 into its
// own operations.
__device__ __forceinline__ double xo_uniform53(uint64_t s[4])
{
    const uint64_t x = xo_next(s);
    uint32_t hi = (uint32_t)(x >> 32);
    const uint32_t lo = (uint32_t)x;
    // (an empty asm on hi: without it the compiler matches the two 32-bit
    // conversions as a 64-bit one and adds a +0.0 term)
    asm("" : "+v"(hi));
    return fma((double)hi, 0x1.0p21, (double)(lo >> 11));
}

// s <- M s.  Read through the constant address space so that the uniform
// columns come in with scalar loads into SGPRs.
typedef const __attribute__((address_space(4))) uint64_t cu64;

__device__ __forceinline__ void gf2_apply(const uint64_t *__restrict__ mg, uint64_t s[4])
{
    cu64 *m = (cu64 *)mg;
    uint64_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint64_t x = s[w];
#pragma unroll 4
        for (int b = 0; b < 64; ++b) {
            uint64_t mask = 0ull - ((x >> b) & 1ull);
            cu64 *c = m + 4 * (64 * w + b);
            r0 ^= c[0] & mask;
            r1 ^= c[1] & mask;
            r2 ^= c[2] & mask;
            r3 ^= c[3] & mask;
        }
    }
    s[0] = r0;
    s[1] = r1;
    s[2] = r2;
    s[3] = r3;
}

// ---- xoshiro128+ (the f32 generator of src/random/standard.rs) ---------------
__device__ __forceinline__ uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

__device__ __forceinline__ uint32_t xo128_next(uint32_t s[4])
{
    const uint32_t r = s[0] + s[3];
    const uint32_t t = s[1] << 9;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl32(s[3], 11);
    return r;
}

// rand's Standard f32: (u32 >> 8) * 2^-24 (exact).
__device__ __forceinline__ float xo128_uniform(uint32_t s[4]) { return (float)(xo128_next(s) >> 8) * 0x1.0p-24f; }

// s <- M s for a 128x128 GF(2) matrix stored as 128 columns of 4 words.
typedef const __attribute__((address_space(4))) uint32_t cu32;

__device__ __forceinline__ void gf2_apply128(const uint32_t *__restrict__ mg, uint32_t s[4])
{
    cu32 *m = (cu32 *)mg;
    uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint32_t x = s[w];
#pragma unroll 4
        for (int b = 0; b < 32; ++b) {
            const uint32_t mask = 0u - ((x >> b) & 1u);
            cu32 *c = m + 4 * (32 * w + b);
            r0 ^= c[0] & mask;
            r1 ^= c[1] & mask;
            r2 ^= c[2] & mask;
            r3 ^= c[3] & mask;
        }
    }
    s[0] = r0;
    s[1] = r1;
    s[2] = r2;
    s[3] = r3;
}

} // namespace tpd
