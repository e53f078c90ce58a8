// tp_xoshiro.h -- xoshiro256+ on the device (src/random/standard.rs via
// rand_xoshiro 0.6.0 and rand 0.8.5's Standard f64), and GF(2) skip-ahead by
// a 256x256 bit matrix (256 columns of 4 words, uniform across the wavefront).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tpd {

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// next_u64: output s0 + s3 before the state update.
__device__ __forceinline__ uint64_t xo_next(uint64_t s[4])
{
    uint64_t r = s[0] + s[3];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
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
