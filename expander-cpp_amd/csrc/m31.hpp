// M31 = GF(2^31 - 1) for host and gfx950 device code.
//
// Semantics follow the reference's field (include/field/M31.hpp:35-142 scalar,
// include/field/M31_avx.tcc:97-176 packed lanes): every operation takes canonical
// values in [0, p) and returns canonical values.  On the device a product is formed
// with v_mul_lo_u32 / v_mul_hi_u32 and reduced by the Mersenne shift-add fold; sums
// use a wrap-free add followed by an unsigned min, which is the conditional subtract.
#pragma once

#include <cstdint>

#ifndef GKRX_HD
#if defined(__HIP__)
#define GKRX_HD __host__ __device__ inline __attribute__((always_inline))
#else
#define GKRX_HD inline
#endif
#endif

namespace gkrx {

constexpr uint32_t kP = 0x7fffffffu;

GKRX_HD uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }

// a + b for a, b < p: a + b < 2^32, and min(s, s - p) is s mod p.
GKRX_HD uint32_t m31_add(uint32_t a, uint32_t b)
{
    uint32_t s = a + b;
    return umin32(s, s - kP);
}

// a - b for a, b < p: on borrow the difference wraps above p, and d + p is the answer.
GKRX_HD uint32_t m31_sub(uint32_t a, uint32_t b)
{
    uint32_t d = a - b;
    return umin32(d, d + kP);
}

GKRX_HD uint32_t m31_neg(uint32_t a) { return a == 0 ? 0 : kP - a; }

// a * b for a, b < p.  prod < 2^62, so (prod >> 31) < 2^31 - 3 and
// (prod & p) + (prod >> 31) < 2p: one conditional subtract is enough.
GKRX_HD uint32_t m31_mul(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t lo = a * b;
    uint32_t hi = __umulhi(a, b);
    uint32_t t = (lo & kP) + ((hi << 1) | (lo >> 31));
#else
    uint64_t x = (uint64_t)a * b;
    uint32_t t = (uint32_t)((x & kP) + (x >> 31));
#endif
    return umin32(t, t - kP);
}

// A value below 2^32 congruent to a * b (a, b < p): the product folded once, without the
// final conditional subtract; for sums that are reduced once at the end (m31_reduce64).
GKRX_HD uint32_t m31_mul_wide(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t lo = a * b;
    uint32_t hi = __umulhi(a, b);
    return (lo & kP) + ((hi << 1) | (lo >> 31));
#else
    uint64_t x = (uint64_t)a * b;
    return (uint32_t)((x & kP) + (x >> 31));
#endif
}

// M31::from_bytes on a little-endian word (M31.hpp:111-116): full reduction.
GKRX_HD uint32_t m31_from_u32_full(uint32_t x)
{
    x = (x & kP) + (x >> 31);
    return umin32(x, x - kP);
}

// Reduce an unsigned 64-bit sum of canonical values to canonical form.
GKRX_HD uint32_t m31_reduce64(uint64_t x)
{
    x = (x & kP) + (x >> 31);  // < 2^31 + 2^33
    x = (x & kP) + (x >> 31);  // < 2^31 + 5
    uint32_t t = (uint32_t)x;
    return umin32(t, t - kP);
}

}  // namespace gkrx
