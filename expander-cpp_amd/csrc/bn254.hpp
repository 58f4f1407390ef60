// BN254 scalar field Fr, r = 0x30644e72e131a029b85045b68181585d2833e84879b9709143e1f593f0000001,
// for host and gfx950 device code.
//
// The reference names BN254 only as a Field_type (include/configuration/config.hpp:12-16,
// field_size 254 at :54); its field header is absent, so the representation here is the
// build's own: elements are 8 little-endian 32-bit limbs; on the device they are kept in
// Montgomery form (R = 2^256) and multiplied with CIOS, each 32 x 32 -> 64-bit step a
// v_mad_u64_u32.  Serialization (transcript, proof, claims) is the canonical value as 32
// little-endian bytes; a challenge is the 32 digest bytes read as a little-endian integer
// reduced mod r.
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

// a - b, returns the borrow
GKRX_HD uint32_t sub256(uint32_t out[8], const uint32_t a[8], const uint32_t b[8])
{
    uint64_t borrow = 0;
    for (int i = 0; i < 8; i++) {
        const uint64_t d = (uint64_t)a[i] - b[i] - borrow;
        out[i] = (uint32_t)d;
        borrow = (d >> 63) & 1u;
    }
    return (uint32_t)borrow;
}

// a + b, returns the carry
GKRX_HD uint32_t add256(uint32_t out[8], const uint32_t a[8], const uint32_t b[8])
{
    uint64_t carry = 0;
    for (int i = 0; i < 8; i++) {
        const uint64_t s = (uint64_t)a[i] + b[i] + carry;
        out[i] = (uint32_t)s;
        carry = s >> 32;
    }
    return (uint32_t)carry;
}

// (a + b) mod r for a, b < r (r < 2^254, so a + b has no carry out of 256 bits)
GKRX_HD Fr fr_add(const Fr &a, const Fr &b)
{
    const uint32_t R[8] = GKRX_BN_R;
    Fr s, t;
    add256(s.w, a.w, b.w);
    const uint32_t borrow = sub256(t.w, s.w, R);
    return borrow ? s : t;
}

// (a - b) mod r for a, b < r
GKRX_HD Fr fr_sub(const Fr &a, const Fr &b)
{
    const uint32_t R[8] = GKRX_BN_R;
    Fr d;
    if (sub256(d.w, a.w, b.w)) add256(d.w, d.w, R);
    return d;
}

// Montgomery product a * b * 2^-256 mod r (CIOS, 32-bit words); a, b < r give a result < r.
GKRX_HD Fr fr_mont_mul(const Fr &a, const Fr &b)
{
    const uint32_t R[8] = GKRX_BN_R;
    uint32_t t[10];
    for (int i = 0; i < 10; i++) t[i] = 0;
    for (int i = 0; i < 8; i++) {
        uint64_t c = 0;
        for (int j = 0; j < 8; j++) {
            const uint64_t p = (uint64_t)a.w[j] * b.w[i] + t[j] + c;
            t[j] = (uint32_t)p;
            c = p >> 32;
        }
        uint64_t s = (uint64_t)t[8] + c;
        t[8] = (uint32_t)s;
        t[9] = (uint32_t)(s >> 32);
        const uint32_t m = t[0] * kBnN0Inv;
        uint64_t p = (uint64_t)m * R[0] + t[0];
        c = p >> 32;
        for (int j = 1; j < 8; j++) {
            p = (uint64_t)m * R[j] + t[j] + c;
            t[j - 1] = (uint32_t)p;
            c = p >> 32;
        }
        s = (uint64_t)t[8] + c;
        t[7] = (uint32_t)s;
        t[8] = t[9] + (uint32_t)(s >> 32);
    }
    Fr out, d;
    for (int i = 0; i < 8; i++) out.w[i] = t[i];
    if (t[8] != 0 || fr_geq(out.w, R)) {
        sub256(d.w, out.w, R);
        return d;
    }
    return out;
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
