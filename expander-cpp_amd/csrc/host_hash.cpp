// Host transcript hashes.  The transcript prefix of a proof -- the hash of the RAW
// commitment (the whole input layer, include/LinearGKR/LinearGKR.hpp:62-69), grinding
// (:13-32) and the output-layer challenges (include/LinearGKR/gkr.hpp:42-49) -- is a
// serial chain that no amount of device parallelism shortens, so it runs on a host core
// (SHA-256 with the x86 SHA extensions when present) while the device evaluates the
// circuit.  Every later challenge is drawn on the device.
#include <immintrin.h>

#include <cpuid.h>
#include <cstdint>
#include <cstring>

#include "bn254.hpp"
#include "host_hash.hpp"

namespace gkrx {

namespace {

const uint32_t kK[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

inline uint32_t ror(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

void sha256_blocks_portable(uint32_t s[8], const uint8_t *p, size_t nblocks)
{
    for (size_t b = 0; b < nblocks; b++, p += 64) {
        uint32_t w[64];
        for (int t = 0; t < 16; t++)
            w[t] = ((uint32_t)p[4 * t] << 24) | ((uint32_t)p[4 * t + 1] << 16) | ((uint32_t)p[4 * t + 2] << 8) |
                   p[4 * t + 3];
        for (int t = 16; t < 64; t++)
            w[t] = w[t - 16] + (ror(w[t - 15], 7) ^ ror(w[t - 15], 18) ^ (w[t - 15] >> 3)) + w[t - 7] +
                   (ror(w[t - 2], 17) ^ ror(w[t - 2], 19) ^ (w[t - 2] >> 10));
        uint32_t a = s[0], bb = s[1], c = s[2], d = s[3], e = s[4], f = s[5], g = s[6], h = s[7];
        for (int t = 0; t < 64; t++) {
            uint32_t t1 = h + (ror(e, 6) ^ ror(e, 11) ^ ror(e, 25)) + ((e & f) ^ (~e & g)) + kK[t] + w[t];
            uint32_t t2 = (ror(a, 2) ^ ror(a, 13) ^ ror(a, 22)) + ((a & bb) ^ (a & c) ^ (bb & c));
            h = g; g = f; f = e; e = d + t1; d = c; c = bb; bb = a; a = t1 + t2;
        }
        s[0] += a; s[1] += bb; s[2] += c; s[3] += d; s[4] += e; s[5] += f; s[6] += g; s[7] += h;
    }
}

// SHA-256 compression with the x86 SHA extensions (sha256rnds2 / msg1 / msg2), following
// the register arrangement Intel documents for them (state as ABEF / CDGH halves).
__attribute__((target("sha,sse4.1,ssse3"))) void sha256_blocks_shani(uint32_t s[8], const uint8_t *p,
                                                                       size_t nblocks)
{
    const __m128i kBswap = _mm_set_epi64x(0x0c0d0e0f08090a0bULL, 0x0405060700010203ULL);
    __m128i tmp = _mm_loadu_si128(reinterpret_cast<const __m128i *>(&s[0]));
    __m128i st1 = _mm_loadu_si128(reinterpret_cast<const __m128i *>(&s[4]));
    tmp = _mm_shuffle_epi32(tmp, 0xB1);
    st1 = _mm_shuffle_epi32(st1, 0x1B);
    __m128i st0 = _mm_alignr_epi8(tmp, st1, 8);  // ABEF
    st1 = _mm_blend_epi16(st1, tmp, 0xF0);       // CDGH
    for (size_t b = 0; b < nblocks; b++, p += 64) {
        const __m128i save0 = st0, save1 = st1;
        __m128i w[4];
        for (int i = 0; i < 4; i++)
            w[i] = _mm_shuffle_epi8(_mm_loadu_si128(reinterpret_cast<const __m128i *>(p + 16 * i)), kBswap);
        for (int g = 0; g < 16; g++) {
            if (g >= 4) {
                __m128i x = _mm_sha256msg1_epu32(w[g & 3], w[(g + 1) & 3]);
                x = _mm_add_epi32(x, _mm_alignr_epi8(w[(g + 3) & 3], w[(g + 2) & 3], 4));
                w[g & 3] = _mm_sha256msg2_epu32(x, w[(g + 3) & 3]);
            }
            __m128i msg = _mm_add_epi32(w[g & 3], _mm_loadu_si128(reinterpret_cast<const __m128i *>(&kK[4 * g])));
            st1 = _mm_sha256rnds2_epu32(st1, st0, msg);
            msg = _mm_shuffle_epi32(msg, 0x0E);
            st0 = _mm_sha256rnds2_epu32(st0, st1, msg);
        }
        st0 = _mm_add_epi32(st0, save0);
        st1 = _mm_add_epi32(st1, save1);
    }
    tmp = _mm_shuffle_epi32(st0, 0x1B);
    st1 = _mm_shuffle_epi32(st1, 0xB1);
    st0 = _mm_blend_epi16(tmp, st1, 0xF0);
    st1 = _mm_alignr_epi8(st1, tmp, 8);
    _mm_storeu_si128(reinterpret_cast<__m128i *>(&s[0]), st0);
    _mm_storeu_si128(reinterpret_cast<__m128i *>(&s[4]), st1);
}

bool cpu_has_sha()
{
    unsigned a, b, c, d;
    if (!__get_cpuid_count(7, 0, &a, &b, &c, &d)) return false;
    return (b >> 29) & 1u;
}

const bool kHasSha = cpu_has_sha();

void sha256_blocks(uint32_t s[8], const uint8_t *p, size_t nblocks)
{
    if (kHasSha) sha256_blocks_shani(s, p, nblocks);
    else sha256_blocks_portable(s, p, nblocks);
}

const uint64_t kRC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
    0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

inline uint64_t rol(uint64_t x, int n) { return (x << n) | (x >> ((64 - n) & 63)); }

// Keccak-f[1600] on 25 named lanes (sXY = lane (x, y) of state s), two rounds per trip
// (a -> e -> a) so that the state stays in registers.  KF_THETA forms the column parities
// c and t[x] = c[x-1] ^ rol(c[x+1], 1).  KF_ROW builds output row Y: its lane X is input
// lane (x, y) = ((3Y + X) mod 5, X) (pi moves (x, y) to (y, 2x + 3y)) xor t[x], rotated by
// rho(x, y) (the triangular numbers 1, 3, 6, ... mod 64 along pi's orbit of (1, 0)), then
// chi along the row.
#define KF_THETA(s)                                                                                           \
    const uint64_t c0 = s##00 ^ s##01 ^ s##02 ^ s##03 ^ s##04, c1 = s##10 ^ s##11 ^ s##12 ^ s##13 ^ s##14,    \
                   c2 = s##20 ^ s##21 ^ s##22 ^ s##23 ^ s##24, c3 = s##30 ^ s##31 ^ s##32 ^ s##33 ^ s##34,    \
                   c4 = s##40 ^ s##41 ^ s##42 ^ s##43 ^ s##44;                                                \
    const uint64_t t0 = c4 ^ rol(c1, 1), t1 = c0 ^ rol(c2, 1), t2 = c1 ^ rol(c3, 1), t3 = c2 ^ rol(c4, 1),     \
                   t4 = c3 ^ rol(c0, 1);
#define KF_ROW(o, s, Y, l0, x0, r0, l1, x1, r1, l2, x2, r2, l3, x3, r3, l4, x4, r4)                          \
    {                                                                                                         \
        const uint64_t b0 = rol(s##l0 ^ t##x0, r0), b1 = rol(s##l1 ^ t##x1, r1), b2 = rol(s##l2 ^ t##x2, r2), \
                       b3 = rol(s##l3 ^ t##x3, r3), b4 = rol(s##l4 ^ t##x4, r4);                              \
        o##0##Y = b0 ^ (~b1 & b2);                                                                            \
        o##1##Y = b1 ^ (~b2 & b3);                                                                            \
        o##2##Y = b2 ^ (~b3 & b4);                                                                            \
        o##3##Y = b3 ^ (~b4 & b0);                                                                            \
        o##4##Y = b4 ^ (~b0 & b1);                                                                            \
    }
#define KF_ROUND(o, s, rc)                                                   \
    {                                                                        \
        KF_THETA(s)                                                          \
        KF_ROW(o, s, 0, 00, 0, 0, 11, 1, 44, 22, 2, 43, 33, 3, 21, 44, 4, 14) \
        KF_ROW(o, s, 1, 30, 3, 28, 41, 4, 20, 02, 0, 3, 13, 1, 45, 24, 2, 61) \
        KF_ROW(o, s, 2, 10, 1, 1, 21, 2, 6, 32, 3, 25, 43, 4, 8, 04, 0, 18)   \
        KF_ROW(o, s, 3, 40, 4, 27, 01, 0, 36, 12, 1, 10, 23, 2, 15, 34, 3, 56) \
        KF_ROW(o, s, 4, 20, 2, 62, 31, 3, 55, 42, 4, 39, 03, 0, 41, 14, 1, 2) \
        o##00 ^= (rc);                                                       \
    }

__attribute__((always_inline)) inline void keccak_f_lanes(uint64_t A[25])
{
    uint64_t a00 = A[0], a10 = A[1], a20 = A[2], a30 = A[3], a40 = A[4], a01 = A[5], a11 = A[6], a21 = A[7],
             a31 = A[8], a41 = A[9], a02 = A[10], a12 = A[11], a22 = A[12], a32 = A[13], a42 = A[14], a03 = A[15],
             a13 = A[16], a23 = A[17], a33 = A[18], a43 = A[19], a04 = A[20], a14 = A[21], a24 = A[22], a34 = A[23],
             a44 = A[24];
    uint64_t e00, e10, e20, e30, e40, e01, e11, e21, e31, e41, e02, e12, e22, e32, e42, e03, e13, e23, e33, e43,
        e04, e14, e24, e34, e44;
    for (int round = 0; round < 24; round += 2) {
        KF_ROUND(e, a, kRC[round])
        KF_ROUND(a, e, kRC[round + 1])
    }
    A[0] = a00, A[1] = a10, A[2] = a20, A[3] = a30, A[4] = a40, A[5] = a01, A[6] = a11, A[7] = a21, A[8] = a31;
    A[9] = a41, A[10] = a02, A[11] = a12, A[12] = a22, A[13] = a32, A[14] = a42, A[15] = a03, A[16] = a13;
    A[17] = a23, A[18] = a33, A[19] = a43, A[20] = a04, A[21] = a14, A[22] = a24, A[23] = a34, A[24] = a44;
}
#undef KF_ROUND
#undef KF_ROW
#undef KF_THETA

// the same code twice: with BMI1/BMI2 (andn for ~b & c, rorx for the rotations) where the
// CPU has them, plain x86-64 otherwise
void keccak_f_plain(uint64_t A[25]) { keccak_f_lanes(A); }
__attribute__((target("bmi,bmi2"))) void keccak_f_bmi(uint64_t A[25]) { keccak_f_lanes(A); }

const bool kHasBmi2 = __builtin_cpu_supports("bmi") && __builtin_cpu_supports("bmi2");

void keccak_f(uint64_t A[25])
{
    if (kHasBmi2) keccak_f_bmi(A);
    else keccak_f_plain(A);
}

// ---- eight Keccak-256 sponges at once: lane l of every state word belongs to message l
// (AVX-512; the permutation is the one above, word for word, on eight states).
__attribute__((target("avx512f"))) void keccak_f_x8(__m512i A[25])
{
    for (int round = 0; round < 24; round++) {
        __m512i C[5], D[5], B[25];
        for (int x = 0; x < 5; x++)
            C[x] = _mm512_ternarylogic_epi64(_mm512_ternarylogic_epi64(A[x], A[x + 5], A[x + 10], 0x96), A[x + 15],
                                             A[x + 20], 0x96);  // a ^ b ^ c
        for (int x = 0; x < 5; x++) D[x] = _mm512_xor_si512(C[(x + 4) % 5], _mm512_rol_epi64(C[(x + 1) % 5], 1));
        B[0] = _mm512_xor_si512(A[0], D[0]);
        B[10] = _mm512_rol_epi64(_mm512_xor_si512(A[1], D[1]), 1);
        B[20] = _mm512_rol_epi64(_mm512_xor_si512(A[2], D[2]), 62);
        B[5] = _mm512_rol_epi64(_mm512_xor_si512(A[3], D[3]), 28);
        B[15] = _mm512_rol_epi64(_mm512_xor_si512(A[4], D[4]), 27);
        B[16] = _mm512_rol_epi64(_mm512_xor_si512(A[5], D[0]), 36);
        B[1] = _mm512_rol_epi64(_mm512_xor_si512(A[6], D[1]), 44);
        B[11] = _mm512_rol_epi64(_mm512_xor_si512(A[7], D[2]), 6);
        B[21] = _mm512_rol_epi64(_mm512_xor_si512(A[8], D[3]), 55);
        B[6] = _mm512_rol_epi64(_mm512_xor_si512(A[9], D[4]), 20);
        B[7] = _mm512_rol_epi64(_mm512_xor_si512(A[10], D[0]), 3);
        B[17] = _mm512_rol_epi64(_mm512_xor_si512(A[11], D[1]), 10);
        B[2] = _mm512_rol_epi64(_mm512_xor_si512(A[12], D[2]), 43);
        B[12] = _mm512_rol_epi64(_mm512_xor_si512(A[13], D[3]), 25);
        B[22] = _mm512_rol_epi64(_mm512_xor_si512(A[14], D[4]), 39);
        B[23] = _mm512_rol_epi64(_mm512_xor_si512(A[15], D[0]), 41);
        B[8] = _mm512_rol_epi64(_mm512_xor_si512(A[16], D[1]), 45);
        B[18] = _mm512_rol_epi64(_mm512_xor_si512(A[17], D[2]), 15);
        B[3] = _mm512_rol_epi64(_mm512_xor_si512(A[18], D[3]), 21);
        B[13] = _mm512_rol_epi64(_mm512_xor_si512(A[19], D[4]), 8);
        B[14] = _mm512_rol_epi64(_mm512_xor_si512(A[20], D[0]), 18);
        B[24] = _mm512_rol_epi64(_mm512_xor_si512(A[21], D[1]), 2);
        B[9] = _mm512_rol_epi64(_mm512_xor_si512(A[22], D[2]), 61);
        B[19] = _mm512_rol_epi64(_mm512_xor_si512(A[23], D[3]), 56);
        B[4] = _mm512_rol_epi64(_mm512_xor_si512(A[24], D[4]), 14);
        for (int y = 0; y < 25; y += 5)
            for (int x = 0; x < 5; x++)  // a ^ (~b & c)
                A[y + x] = _mm512_ternarylogic_epi64(B[y + x], B[y + (x + 1) % 5], B[y + (x + 2) % 5], 0xD2);
        A[0] = _mm512_xor_si512(A[0], _mm512_set1_epi64((long long)kRC[round]));
    }
}

// Keccak-256 of eight messages of the same length; lane l reads in[l].
__attribute__((target("avx512f"))) void keccak256_x8(const uint8_t *const in[8], size_t len, uint64_t out[4][8])
{
    constexpr size_t kRate = 136;
    __m512i A[25];
    for (auto &a : A) a = _mm512_setzero_si512();
    const __m512i base = _mm512_set_epi64((long long)(uintptr_t)in[7], (long long)(uintptr_t)in[6],
                                          (long long)(uintptr_t)in[5], (long long)(uintptr_t)in[4],
                                          (long long)(uintptr_t)in[3], (long long)(uintptr_t)in[2],
                                          (long long)(uintptr_t)in[1], (long long)(uintptr_t)in[0]);
    size_t off = 0;
    for (; len - off >= kRate; off += kRate) {
        for (int i = 0; i < 17; i++) {  // word i of the block of every lane (absolute addresses)
            const __m512i at = _mm512_add_epi64(base, _mm512_set1_epi64((long long)(off + 8 * i)));
            A[i] = _mm512_xor_si512(A[i], _mm512_i64gather_epi64(at, nullptr, 1));
        }
        keccak_f_x8(A);
    }
    alignas(64) uint8_t blk[8][kRate];
    const size_t rem = len - off;
    for (int l = 0; l < 8; l++) {
        memset(blk[l], 0, kRate);
        memcpy(blk[l], in[l] + off, rem);
        blk[l][rem] ^= 0x01;
        blk[l][kRate - 1] ^= 0x80;
    }
    const __m512i lanes = _mm512_set_epi64(7 * kRate, 6 * kRate, 5 * kRate, 4 * kRate, 3 * kRate, 2 * kRate, kRate, 0);
    for (int i = 0; i < 17; i++) {
        const __m512i at = _mm512_add_epi64(lanes, _mm512_set1_epi64(8 * i));
        A[i] = _mm512_xor_si512(A[i], _mm512_i64gather_epi64(at, &blk[0][0], 1));
    }
    keccak_f_x8(A);
    for (int i = 0; i < 4; i++) _mm512_storeu_si512(out[i], A[i]);
}

// Keccak-256 of eight 32-byte messages held as 4 words per lane, in place.
__attribute__((target("avx512f"))) void keccak256_32_x8(__m512i m[4])
{
    __m512i A[25];
    for (auto &a : A) a = _mm512_setzero_si512();
    for (int i = 0; i < 4; i++) A[i] = m[i];
    A[4] = _mm512_set1_epi64(0x01);                      // pad byte after the 32 message bytes
    A[16] = _mm512_set1_epi64((long long)0x8000000000000000ULL);  // last rate byte
    keccak_f_x8(A);
    for (int i = 0; i < 4; i++) m[i] = A[i];
}

bool cpu_has_avx512() { return __builtin_cpu_supports("avx512f"); }

const bool kHasAvx512 = cpu_has_avx512();

}  // namespace

void host_sha256(const uint8_t *in, size_t len, uint8_t out[32])
{
    uint32_t s[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    const size_t full = len / 64;
    sha256_blocks(s, in, full);
    uint8_t tail[128] = {0};
    const size_t rem = len - 64 * full;
    memcpy(tail, in + 64 * full, rem);
    tail[rem] = 0x80;
    const size_t tl = rem + 9 <= 64 ? 64 : 128;
    const uint64_t bits = (uint64_t)len * 8;
    for (int i = 0; i < 8; i++) tail[tl - 1 - i] = (uint8_t)(bits >> (8 * i));
    sha256_blocks(s, tail, tl / 64);
    for (int i = 0; i < 8; i++) {
        out[4 * i] = (uint8_t)(s[i] >> 24);
        out[4 * i + 1] = (uint8_t)(s[i] >> 16);
        out[4 * i + 2] = (uint8_t)(s[i] >> 8);
        out[4 * i + 3] = (uint8_t)s[i];
    }
}

void host_keccak256(const uint8_t *in, size_t len, uint8_t out[32])
{
    constexpr size_t kRate = 136;
    uint64_t A[25] = {0};
    size_t off = 0;
    while (len - off >= kRate) {
        for (size_t i = 0; i < kRate / 8; i++) {
            uint64_t w;
            memcpy(&w, in + off + 8 * i, 8);
            A[i] ^= w;
        }
        keccak_f(A);
        off += kRate;
    }
    uint8_t blk[kRate] = {0};
    memcpy(blk, in + off, len - off);
    blk[len - off] ^= 0x01;
    blk[kRate - 1] ^= 0x80;
    for (size_t i = 0; i < kRate / 8; i++) {
        uint64_t w;
        memcpy(&w, blk + 8 * i, 8);
        A[i] ^= w;
    }
    keccak_f(A);
    memcpy(out, A, 32);
}

void host_hash(int kind, const uint8_t *in, size_t len, uint8_t out[32])
{
    uint8_t tmp[32];
    if (kind == 1) host_keccak256(in, len, tmp);
    else host_sha256(in, len, tmp);
    memcpy(out, tmp, 32);
}

bool host_sha_extensions() { return kHasSha; }

// M31::from_bytes of the first four digest bytes (include/field/M31.hpp:111-116)
static uint32_t digest_to_m31(const uint8_t d[32])
{
    uint32_t x;
    memcpy(&x, d, 4);
    x = (x & 0x7fffffffu) + (x >> 31);
    if (x >= 0x7fffffffu) x -= 0x7fffffffu;
    return x;
}

void host_transcript_prefix(int kind, const uint8_t *commit, size_t commit_len, uint32_t R, uint32_t mlast,
                            uint32_t grinding_bits, uint8_t grind_out[32], uint32_t *rz1, uint32_t tstate[10])
{
    uint8_t digest[32];
    // grind: challenge_fs(256 / 31 = 8); the first hashes the commitment bytes, the
    // others re-hash the digest; each is written as 4 LE bytes (ceil(31, 8) = 4).
    uint8_t hb[32] = {0};
    host_hash(kind, commit, commit_len, digest);
    for (int i = 0; i < 8; i++) {
        if (i > 0) host_hash(kind, digest, 32, digest);
        uint32_t h = digest_to_m31(digest);
        memcpy(hb + 4 * i, &h, 4);
    }
    for (uint32_t i = 0; i < (1u << grinding_bits); i++) host_hash(kind, hb, 32, hb);
    memcpy(grind_out, hb, 32);
    // output-layer challenges rz1[j][i], var-outer / rep-inner; the first hashes the
    // 32 grind bytes, the rest re-hash the digest.
    bool pending = true;
    for (uint32_t i = 0; i < mlast; i++)
        for (uint32_t j = 0; j < R; j++) {
            if (pending) {
                host_hash(kind, hb, 32, digest);
                pending = false;
            } else {
                host_hash(kind, digest, 32, digest);
            }
            rz1[(size_t)j * mlast + i] = digest_to_m31(digest);
        }
    // transcript state relative to the tail that starts at the grind bytes
    tstate[0] = 32;
    tstate[1] = pending ? 0 : 32;
    memcpy(&tstate[2], digest, 32);
}

bool host_keccak_x8() { return kHasAvx512; }

void host_keccak256_many(uint32_t n, const uint8_t *const *in, size_t len, uint8_t *out)
{
    for (uint32_t g = 0; g < n; g += 8) {
        const uint32_t cnt = n - g < 8 ? n - g : 8;
        if (!kHasAvx512) {
            for (uint32_t l = 0; l < cnt; l++) host_keccak256(in[g + l], len, out + 32 * (size_t)(g + l));
            continue;
        }
        const uint8_t *p[8];
        for (uint32_t l = 0; l < 8; l++) p[l] = in[g + (l < cnt ? l : 0)];
        uint64_t o[4][8];
        keccak256_x8(p, len, o);
        for (uint32_t l = 0; l < cnt; l++)
            for (int i = 0; i < 4; i++) memcpy(out + 32 * (size_t)(g + l) + 8 * i, &o[i][l], 8);
    }
}

__attribute__((target("avx512f"))) static void prefix_x8_body(const uint8_t *const commit[8], size_t commit_len,
                                                              uint32_t grinding_bits, uint8_t grind_out[8][32],
                                                              uint8_t digest_out[8][32])
{
    // challenge_fs(8) after the commitment: the first hashes the commitment, the others the
    // previous digest; each contributes 4 LE bytes of M31 to the grind buffer
    uint64_t o[4][8];
    keccak256_x8(commit, commit_len, o);
    __m512i d[4], hb[4];
    for (int i = 0; i < 4; i++) d[i] = _mm512_loadu_si512(o[i]);
    alignas(64) uint32_t h[8][8];  // h[i][lane]
    for (int i = 0; i < 8; i++) {
        if (i > 0) keccak256_32_x8(d);
        alignas(64) uint64_t w0[8];
        _mm512_store_si512(w0, d[0]);
        for (int l = 0; l < 8; l++) {
            uint32_t x = (uint32_t)w0[l];
            x = (x & 0x7fffffffu) + (x >> 31);
            if (x >= 0x7fffffffu) x -= 0x7fffffffu;
            h[i][l] = x;
        }
    }
    alignas(64) uint64_t hw[4][8];
    for (int i = 0; i < 4; i++)
        for (int l = 0; l < 8; l++) hw[i][l] = (uint64_t)h[2 * i][l] | ((uint64_t)h[2 * i + 1][l] << 32);
    for (int i = 0; i < 4; i++) hb[i] = _mm512_load_si512(hw[i]);
    for (uint32_t r = 0; r < (1u << grinding_bits); r++) keccak256_32_x8(hb);
    for (int i = 0; i < 4; i++) {
        _mm512_store_si512(hw[i], hb[i]);
        _mm512_storeu_si512(o[i], d[i]);
    }
    for (int l = 0; l < 8; l++)
        for (int i = 0; i < 4; i++) {
            memcpy(grind_out[l] + 8 * i, &hw[i][l], 8);
            memcpy(digest_out[l] + 8 * i, &o[i][l], 8);
        }
}

void host_transcript_prefix_x8(uint32_t n, const uint8_t *const *commit, size_t commit_len, uint32_t R,
                               const uint32_t *mlast, uint32_t grinding_bits, uint8_t (*grind_out)[32],
                               uint32_t *const *rz1, uint32_t (*tstate)[10])
{
    if (!kHasAvx512) {
        for (uint32_t l = 0; l < n; l++)
            host_transcript_prefix(1, commit[l], commit_len, R, mlast[l], grinding_bits, grind_out[l], rz1[l],
                                   tstate[l]);
        return;
    }
    const uint8_t *p[8];
    for (uint32_t l = 0; l < 8; l++) p[l] = commit[l < n ? l : 0];
    uint8_t grind[8][32], digest[8][32];
    prefix_x8_body(p, commit_len, grinding_bits, grind, digest);
    for (uint32_t l = 0; l < n; l++) {
        memcpy(grind_out[l], grind[l], 32);
        // the output challenges, as in host_transcript_prefix (a few dozen short hashes)
        uint8_t dg[32];
        bool pending = true;
        for (uint32_t i = 0; i < mlast[l]; i++)
            for (uint32_t j = 0; j < R; j++) {
                if (pending) {
                    host_keccak256(grind[l], 32, dg);
                    pending = false;
                } else {
                    host_keccak256(dg, 32, dg);
                }
                rz1[l][(size_t)j * mlast[l] + i] = digest_to_m31(dg);
            }
        tstate[l][0] = 32;
        tstate[l][1] = pending ? 0 : 32;
        memcpy(&tstate[l][2], pending ? digest[l] : dg, 32);
    }
}

// the digest as a 256-bit little-endian integer reduced mod r
static void digest_to_fr(const uint8_t d[32], uint32_t out[8])
{
    uint32_t w[8];
    memcpy(w, d, 32);
    const Fr v = fr_reduce256(w);
    memcpy(out, v.w, 32);
}

void host_transcript_prefix_bn254(int kind, const uint8_t *commit, size_t commit_len, uint32_t mlast,
                                  uint32_t grinding_bits, uint8_t grind_out[32], uint32_t *rz1, uint32_t tstate[10])
{
    uint8_t digest[32];
    uint8_t hb[32];
    host_hash(kind, commit, commit_len, digest);
    uint32_t h[8];
    digest_to_fr(digest, h);
    memcpy(hb, h, 32);
    for (uint32_t i = 0; i < (1u << grinding_bits); i++) host_hash(kind, hb, 32, hb);
    memcpy(grind_out, hb, 32);
    bool pending = true;
    for (uint32_t i = 0; i < mlast; i++) {
        if (pending) {
            host_hash(kind, hb, 32, digest);
            pending = false;
        } else {
            host_hash(kind, digest, 32, digest);
        }
        digest_to_fr(digest, rz1 + 8 * (size_t)i);
    }
    tstate[0] = 32;
    tstate[1] = pending ? 0 : 32;
    memcpy(&tstate[2], digest, 32);
}

}  // namespace gkrx
