// BN254 scalar field Fr, r = 0x30644e72e131a029b85045b68181585d2833e84879b9709143e1f593f0000001,
// for host and gfx950 device code.
//
// The reference names BN254 only as a Field_type (include/configuration/config.hpp:12-16,
// field_size 254 at :54); its field header is absent, so the representation here is the
// build's own: elements are 8 little-endian 32-bit limbs; on the device they are kept in
// Montgomery form (R = 2^256) and multiplied with CIOS, each 32 x 32 -> 64-bit step a
// v_mad_u64_u32 and the carries one v_addc_co_u32 chain per row.  Serialization
// (transcript, proof, claims) is the canonical value as 32 little-endian bytes; a challenge
// is the 32 digest bytes read as a little-endian integer reduced mod r.
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

struct Fr {
    uint32_t w[8];
};

// modulus, -r^-1 mod 2^32, R mod r, R^2 mod r
#define GKRX_BN_R {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u, 0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u}
#define GKRX_BN_R2 {0xae216da7u, 0x1bb8e645u, 0xe35c59e3u, 0x53fe3ab1u, 0x53bb8085u, 0x8c49833du, 0x7f4e44a5u, 0x0216d0b1u}
#define GKRX_BN_ONE {0x4ffffffbu, 0xac96341cu, 0x9f60cd29u, 0x36fc7695u, 0x7879462eu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u}
constexpr uint32_t kBnN0Inv = 0xefffffffu;

GKRX_HD Fr fr_zero()
{
    Fr z;
    for (int i = 0; i < 8; i++) z.w[i] = 0;
    return z;
}

// a >= b as 256-bit integers
GKRX_HD bool fr_geq(const uint32_t a[8], const uint32_t b[8])
{
    for (int i = 7; i >= 0; i--) {
        if (a[i] != b[i]) return a[i] > b[i];
    }
    return true;
}

// One word of a carry chain: a + b + carry (carry 0 or 1 in and out), and a - b - borrow.
// On gfx950 these are single v_addc_co_u32 / v_subb_co_u32 instructions.
GKRX_HD uint32_t adc32(uint32_t a, uint32_t b, uint32_t &carry)
{
    uint32_t out;
    const uint32_t r = __builtin_addc(a, b, carry, &out);
    carry = out;
    return r;
}
GKRX_HD uint32_t sbb32(uint32_t a, uint32_t b, uint32_t &borrow)
{
    uint32_t out;
    const uint32_t r = __builtin_subc(a, b, borrow, &out);
    borrow = out;
    return r;
}

// a - b, returns the borrow
GKRX_HD uint32_t sub256(uint32_t out[8], const uint32_t a[8], const uint32_t b[8])
{
    uint32_t borrow = 0;
    for (int i = 0; i < 8; i++) out[i] = sbb32(a[i], b[i], borrow);
    return borrow;
}

// a + b, returns the carry
GKRX_HD uint32_t add256(uint32_t out[8], const uint32_t a[8], const uint32_t b[8])
{
    uint32_t carry = 0;
    for (int i = 0; i < 8; i++) out[i] = adc32(a[i], b[i], carry);
    return carry;
}

// (a + b) mod r for a, b < r (r < 2^254, so a + b has no carry out of 256 bits)
GKRX_HD Fr fr_add(const Fr &a, const Fr &b)
{
    const uint32_t R[8] = GKRX_BN_R;
    Fr s, t;
    add256(s.w, a.w, b.w);
    const uint32_t borrow = sub256(t.w, s.w, R);
    for (int i = 0; i < 8; i++) t.w[i] = borrow ? s.w[i] : t.w[i];
    return t;
}

// (a - b) mod r for a, b < r
GKRX_HD Fr fr_sub(const Fr &a, const Fr &b)
{
    const uint32_t R[8] = GKRX_BN_R;
    Fr d, e;
    const uint32_t mask = 0u - sub256(d.w, a.w, b.w);  // all ones on borrow: add r back
    uint32_t carry = 0;
    for (int i = 0; i < 8; i++) e.w[i] = adc32(d.w[i], R[i] & mask, carry);
    return e;
}

// Montgomery product a * b * 2^-256 mod r (CIOS, 32-bit words); a, b < r give a result < r.
// Each row of a word product adds the eight products a_j b_i + t_j (each below 2^64, and
// independent of one another) and then joins their halves with one carry chain, so that
// the 64-bit multiply-adds of a row do not wait on each other's carries.
GKRX_HD Fr fr_mont_mul(const Fr &a, const Fr &b)
{
    const uint32_t R[8] = GKRX_BN_R;
    uint32_t t[10];
    for (int i = 0; i < 10; i++) t[i] = 0;
    for (int i = 0; i < 8; i++) {
        // t += a * b_i
        uint64_t p[8];
        for (int j = 0; j < 8; j++) p[j] = (uint64_t)a.w[j] * b.w[i] + t[j];
        uint32_t c = 0;
        t[0] = (uint32_t)p[0];
        for (int j = 1; j < 8; j++) t[j] = adc32((uint32_t)p[j], (uint32_t)(p[j - 1] >> 32), c);
        t[8] = adc32(t[8], (uint32_t)(p[7] >> 32), c);
        t[9] += c;
        // t = (t + m r) / 2^32 with m = -t_0 / r mod 2^32 (the low word of t + m r is zero)
        const uint32_t m = t[0] * kBnN0Inv;
        uint64_t q[8];
        for (int j = 0; j < 8; j++) q[j] = (uint64_t)m * R[j] + t[j];
        c = 0;
        for (int j = 1; j < 8; j++) t[j - 1] = adc32((uint32_t)q[j], (uint32_t)(q[j - 1] >> 32), c);
        t[7] = adc32(t[8], (uint32_t)(q[7] >> 32), c);
        t[8] = t[9] + c;
        t[9] = 0;
    }
    Fr out, d;
    for (int i = 0; i < 8; i++) out.w[i] = t[i];
    const uint32_t borrow = sub256(d.w, out.w, R);
    const bool keep = t[8] == 0 && borrow;  // t < r: no subtraction
    for (int i = 0; i < 8; i++) d.w[i] = keep ? out.w[i] : d.w[i];
    return d;
}

GKRX_HD Fr fr_to_mont(const Fr &a)
{
    const Fr r2 = {GKRX_BN_R2};
    return fr_mont_mul(a, r2);
}

GKRX_HD Fr fr_from_mont(const Fr &a)
{
    Fr one = fr_zero();
    one.w[0] = 1;
    return fr_mont_mul(a, one);
}

// canonical value of a 256-bit little-endian integer mod r (2^256 < 6 r)
GKRX_HD Fr fr_reduce256(const uint32_t x[8])
{
    const uint32_t R[8] = GKRX_BN_R;
    Fr v;
    for (int i = 0; i < 8; i++) v.w[i] = x[i];
    for (int k = 0; k < 5; k++)
        if (fr_geq(v.w, R)) sub256(v.w, v.w, R);
    return v;
}

// Montgomery form of the small integer c
GKRX_HD Fr fr_mont_of_u32(uint32_t c)
{
    Fr v = fr_zero();
    v.w[0] = c;
    return fr_to_mont(v);
}

}  // namespace gkrx
