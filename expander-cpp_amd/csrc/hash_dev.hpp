// Transcript hashes for one device thread: SHA-256 (FIPS 180-4) and Keccak-256
// (Keccak[c=512], rate 136 B, domain pad 0x01).
//
// The reference's transcript hashes with SHA-256 (include/hash/hashes.hpp:22-31 via
// lib/btc_sha256/crypto/sha256.cpp); Keccak-256 is the alternative FiatShamir_hash_type
// (include/configuration/config.hpp:19-20).  Messages here are always whole 32-bit
// words (field elements and digests), read little-endian from memory; digests are
// returned as 8 little-endian words, i.e. the digest bytes in memory order.
#pragma once

#include <cstdint>

namespace gkrx {

enum HashKind : int { kHashSha256 = 0, kHashKeccak256 = 1 };

__device__ __forceinline__ uint32_t rotr32d(uint32_t x, int n) { return __builtin_amdgcn_alignbit(x, x, n); }
__device__ __forceinline__ uint32_t bswap32d(uint32_t x) { return __builtin_bswap32(x); }

__constant__ static const uint32_t kSha256K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

// a ^ b ^ c in one gfx950 v_bitop3_b32 (truth table 0x96)
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// One SHA-256 compression of a block given as 16 big-endian words.
__device__ __forceinline__ void sha256_compress(uint32_t st[8], uint32_t w[16])
{
    uint32_t a = st[0], b = st[1], c = st[2], d = st[3], e = st[4], f = st[5], g = st[6], h = st[7];
#pragma unroll
    for (int t = 0; t < 64; t++) {
        uint32_t wt;
        if (t < 16) {
            wt = w[t];
        } else {
            uint32_t w15 = w[(t - 15) & 15], w2 = w[(t - 2) & 15];
            uint32_t s0 = xor3(rotr32d(w15, 7), rotr32d(w15, 18), w15 >> 3);
            uint32_t s1 = xor3(rotr32d(w2, 17), rotr32d(w2, 19), w2 >> 10);
            wt = w[t & 15] + s0 + w[(t - 7) & 15] + s1;
            w[t & 15] = wt;
        }
        uint32_t S1 = xor3(rotr32d(e, 6), rotr32d(e, 11), rotr32d(e, 25));
        uint32_t ch = (e & f) ^ (~e & g);
        uint32_t t1 = h + S1 + ch + kSha256K[t] + wt;
        uint32_t S0 = xor3(rotr32d(a, 2), rotr32d(a, 13), rotr32d(a, 22));
        uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + S0 + mj;
    }
    st[0] += a; st[1] += b; st[2] += c; st[3] += d;
    st[4] += e; st[5] += f; st[6] += g; st[7] += h;
}

// SHA-256 of nwords little-endian words at msg (byte length 4 * nwords).
__device__ inline void sha256_words(const uint32_t *msg, uint32_t nwords, uint32_t out[8])
{
    uint32_t st[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                      0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    uint32_t w[16];
    const uint64_t bitlen = (uint64_t)nwords * 32u;
    // total words after padding: message, 0x80000000, zeros, 2 length words
    const uint32_t total = ((nwords + 1 + 2 + 15) / 16) * 16;
    for (uint32_t base = 0; base < total; base += 16) {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            uint32_t idx = base + i;
            uint32_t v;
            if (idx < nwords) v = bswap32d(msg[idx]);
            else if (idx == nwords) v = 0x80000000u;
            else if (idx == total - 2) v = (uint32_t)(bitlen >> 32);
            else if (idx == total - 1) v = (uint32_t)bitlen;
            else v = 0;
            w[i] = v;
        }
        sha256_compress(st, w);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) out[i] = bswap32d(st[i]);
}

__constant__ static const uint64_t kKeccakRC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
    0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

__device__ __forceinline__ uint64_t rotl64d(uint64_t x, int n) { return (x << n) | (x >> ((64 - n) & 63)); }

// Keccak-f[1600] on 25 lanes indexed x + 5y.  Rho offsets and the pi lane walk are
// the standard ones (the (x, y) -> (y, 2x + 3y) orbit from (1, 0) with triangular offsets).
__device__ inline void keccak_f1600(uint64_t A[25])
{
#pragma unroll 1
    for (int round = 0; round < 24; round++) {
        uint64_t C0 = A[0] ^ A[5] ^ A[10] ^ A[15] ^ A[20];
        uint64_t C1 = A[1] ^ A[6] ^ A[11] ^ A[16] ^ A[21];
        uint64_t C2 = A[2] ^ A[7] ^ A[12] ^ A[17] ^ A[22];
        uint64_t C3 = A[3] ^ A[8] ^ A[13] ^ A[18] ^ A[23];
        uint64_t C4 = A[4] ^ A[9] ^ A[14] ^ A[19] ^ A[24];
        uint64_t D0 = C4 ^ rotl64d(C1, 1), D1 = C0 ^ rotl64d(C2, 1), D2 = C1 ^ rotl64d(C3, 1);
        uint64_t D3 = C2 ^ rotl64d(C4, 1), D4 = C3 ^ rotl64d(C0, 1);
        // theta + rho + pi: B[y + 5((2x+3y)%5)] = rotl(A[x+5y] ^ D[x], r[x][y])
        uint64_t B0 = A[0] ^ D0;
        uint64_t B10 = rotl64d(A[1] ^ D1, 1);
        uint64_t B7 = rotl64d(A[10] ^ D0, 3);
        uint64_t B11 = rotl64d(A[7] ^ D2, 6);
        uint64_t B17 = rotl64d(A[11] ^ D1, 10);
        uint64_t B18 = rotl64d(A[17] ^ D2, 15);
        uint64_t B3 = rotl64d(A[18] ^ D3, 21);
        uint64_t B5 = rotl64d(A[3] ^ D3, 28);
        uint64_t B16 = rotl64d(A[5] ^ D0, 36);
        uint64_t B8 = rotl64d(A[16] ^ D1, 45);
        uint64_t B21 = rotl64d(A[8] ^ D3, 55);
        uint64_t B24 = rotl64d(A[21] ^ D1, 2);
        uint64_t B4 = rotl64d(A[24] ^ D4, 14);
        uint64_t B15 = rotl64d(A[4] ^ D4, 27);
        uint64_t B23 = rotl64d(A[15] ^ D0, 41);
        uint64_t B19 = rotl64d(A[23] ^ D3, 56);
        uint64_t B13 = rotl64d(A[19] ^ D4, 8);
        uint64_t B12 = rotl64d(A[13] ^ D3, 25);
        uint64_t B2 = rotl64d(A[12] ^ D2, 43);
        uint64_t B20 = rotl64d(A[2] ^ D2, 62);
        uint64_t B14 = rotl64d(A[20] ^ D0, 18);
        uint64_t B22 = rotl64d(A[14] ^ D4, 39);
        uint64_t B9 = rotl64d(A[22] ^ D2, 61);
        uint64_t B6 = rotl64d(A[9] ^ D4, 20);
        uint64_t B1 = rotl64d(A[6] ^ D1, 44);
        // chi
        A[0] = B0 ^ (~B1 & B2); A[1] = B1 ^ (~B2 & B3); A[2] = B2 ^ (~B3 & B4);
        A[3] = B3 ^ (~B4 & B0); A[4] = B4 ^ (~B0 & B1);
        A[5] = B5 ^ (~B6 & B7); A[6] = B6 ^ (~B7 & B8); A[7] = B7 ^ (~B8 & B9);
        A[8] = B8 ^ (~B9 & B5); A[9] = B9 ^ (~B5 & B6);
        A[10] = B10 ^ (~B11 & B12); A[11] = B11 ^ (~B12 & B13); A[12] = B12 ^ (~B13 & B14);
        A[13] = B13 ^ (~B14 & B10); A[14] = B14 ^ (~B10 & B11);
        A[15] = B15 ^ (~B16 & B17); A[16] = B16 ^ (~B17 & B18); A[17] = B17 ^ (~B18 & B19);
        A[18] = B18 ^ (~B19 & B15); A[19] = B19 ^ (~B15 & B16);
        A[20] = B20 ^ (~B21 & B22); A[21] = B21 ^ (~B22 & B23); A[22] = B22 ^ (~B23 & B24);
        A[23] = B23 ^ (~B24 & B20); A[24] = B24 ^ (~B20 & B21);
        A[0] ^= kKeccakRC[round];
    }
}

// Keccak-256 of nwords little-endian words.
__device__ inline void keccak256_words(const uint32_t *msg, uint32_t nwords, uint32_t out[8])
{
    uint64_t A[25];
#pragma unroll
    for (int i = 0; i < 25; i++) A[i] = 0;
    const uint32_t nbytes = nwords * 4u;
    // blocks of 136 bytes = 34 words; the pad byte sits at nbytes, 0x80 at the block end
    const uint32_t nblocks = nbytes / 136 + 1;
    for (uint32_t b = 0; b < nblocks; b++) {
        const uint32_t w0 = b * 34;
#pragma unroll
        for (int i = 0; i < 17; i++) {
            uint32_t idx = w0 + 2 * i;
            uint32_t lo = idx < nwords ? msg[idx] : 0u;
            uint32_t hi = idx + 1 < nwords ? msg[idx + 1] : 0u;
            if (b == nblocks - 1) {
                // pad byte 0x01 at byte offset nbytes (word aligned: low byte of word nwords)
                if (idx == nwords) lo ^= 0x01u;
                if (idx + 1 == nwords) hi ^= 0x01u;
                if (i == 16) hi ^= 0x80000000u;
            }
            A[i] ^= (uint64_t)lo | ((uint64_t)hi << 32);
        }
        keccak_f1600(A);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        out[2 * i] = (uint32_t)A[i];
        out[2 * i + 1] = (uint32_t)(A[i] >> 32);
    }
}

__device__ __forceinline__ void hash_words(int kind, const uint32_t *msg, uint32_t nwords, uint32_t out[8])
{
    if (kind == kHashKeccak256) keccak256_words(msg, nwords, out);
    else sha256_words(msg, nwords, out);
}

// ------------------------------------------------------------ wave-parallel Keccak
// One wavefront computes one Keccak-f[1600] permutation with the 25 state words spread
// over lanes, so one VALU instruction advances every word: lane l belongs to row
// y = l / 8 (rows 0..4 use lanes 0..39) and holds column x = (l % 8 + 4) % 5.  Lanes with
// l % 8 in 1..5 are the state proper; l % 8 = 0, 6, 7 carry copies of columns 4, 0, 1, so
// that the x - 1, x + 1 and x + 2 neighbours of every state lane sit in the same
// 8-lane group and a DPP row shift reaches them.  Lanes 40..63 stay zero.
//   theta : column parity = XOR of the 5 rows, formed with a DPP row rotate by 8 (rows
//           2r, 2r+1 share a 16-lane DPP row) and three gfx950 permlane16 / permlane32 swaps
//           that serve both 32-bit halves of the words at once;
//   rho   : per-lane rotate (v_alignbit with a per-lane shift);
//   pi    : one ds_bpermute (the only LDS-latency step of a round);
//   chi   : DPP row shifts by 1 and 2 inside the group;
//   iota  : lane 1 (x = 0, y = 0).
struct KeccakWaveLane {
    uint32_t valid;   // all ones for lanes 0..39, zero above
    uint32_t src4;    // pi: byte address of the source lane
    uint32_t rsh;     // rho: alignbit shift (32 - r % 32) & 31
    bool rswap;       // rho: pre-swap halves ((r >= 32) xor (r % 32 == 0))
    bool copy67;      // lanes l % 8 in {6, 7}: refreshed from l - 5 after chi
    bool is00;        // the lane of word (0, 0)
    int word;         // absorb: rate word x + 5y (< 17) or -1
};

// rho offsets r[x + 5y]
__constant__ static const uint8_t kKeccakRho[25] = {0,  1,  62, 28, 27, 36, 44, 6,  55, 20, 3,  10, 43,
                                                    25, 39, 41, 45, 15, 21, 8,  18, 2,  61, 56, 14};

__device__ __forceinline__ KeccakWaveLane keccak_wave_lane()
{
    KeccakWaveLane k;
    const uint32_t lane = __lane_id();
    const uint32_t y = lane >> 3, j = lane & 7, x = (j + 4) % 5;
    const bool ok = lane < 40;
    k.valid = ok ? 0xffffffffu : 0u;
    // pi: destination (x', y') reads source (3 (y' - 3 x') mod 5, x')
    const uint32_t xs = (3u * ((y + 15u - 3u * x) % 5u)) % 5u, ys = x;
    k.src4 = ok ? 4u * (8u * ys + xs + 1u) : 4u * lane;
    const uint32_t r = ok ? kKeccakRho[x + 5 * y] : 0u;
    k.rsh = (32u - (r & 31u)) & 31u;
    k.rswap = (r >= 32u) != ((r & 31u) == 0u);
    k.copy67 = j >= 6;
    k.is00 = lane == 1;
    k.word = (ok && x + 5 * y < 17) ? (int)(x + 5 * y) : -1;
    return k;
}

__device__ __forceinline__ uint32_t dpp_from_next(uint32_t v) { return __builtin_amdgcn_update_dpp(0u, v, 0x101, 0xf, 0xf, true); }
__device__ __forceinline__ uint32_t dpp_from_next2(uint32_t v) { return __builtin_amdgcn_update_dpp(0u, v, 0x102, 0xf, 0xf, true); }
__device__ __forceinline__ uint32_t dpp_from_prev(uint32_t v) { return __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xf, 0xf, true); }
__device__ __forceinline__ uint32_t dpp_from_prev5(uint32_t v) { return __builtin_amdgcn_update_dpp(0u, v, 0x115, 0xf, 0xf, true); }
__device__ __forceinline__ uint32_t dpp_ror8(uint32_t v) { return __builtin_amdgcn_update_dpp(0u, v, 0x128, 0xf, 0xf, true); }

// XOR of the 4 rows of 16 lanes of tlo and of thi at once (every lane gets the XOR over its
// lane % 16 position, of tlo in clo and of thi in chi).  permlane16_swap exchanges the odd
// rows of its first operand with the even rows of its second, permlane32_swap the upper half
// of the first with the lower half of the second:
//   swap16(tlo, thi) -> [lo0 hi0 lo2 hi2], [lo1 hi1 lo3 hi3];  x = their XOR
//   swap32(x, x)     -> [x0 x1 x0 x1], [x2 x3 x2 x3];          y = [LO HI LO HI]
//   swap16(y, y)     -> [LO LO LO LO], [HI HI HI HI]
__device__ __forceinline__ void xor_rows2(uint32_t tlo, uint32_t thi, uint32_t &clo, uint32_t &chi)
{
    auto a = __builtin_amdgcn_permlane16_swap(tlo, thi, false, false);
    const uint32_t x = a[0] ^ a[1];
    auto b = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    const uint32_t y = b[0] ^ b[1];
    auto c = __builtin_amdgcn_permlane16_swap(y, y, false, false);
    clo = c[0];
    chi = c[1];
}

__device__ __forceinline__ void keccak_f1600_wave(uint32_t &lo, uint32_t &hi, const KeccakWaveLane &k)
{
    // round constants as literals: the 24 rounds are unrolled, so iota is one masked XOR per
    // half on the lane of word (0, 0) with no constant-memory load
    constexpr uint64_t kRC[24] = {
        0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
        0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
        0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
        0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
        0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
        0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};
    const uint32_t m00 = k.is00 ? 0xffffffffu : 0u;
#pragma unroll
    for (int round = 0; round < 24; round++) {
        // theta
        uint32_t clo, chi;
        xor_rows2(lo ^ dpp_ror8(lo), hi ^ dpp_ror8(hi), clo, chi);
        const uint32_t nlo = dpp_from_next(clo), nhi = dpp_from_next(chi);
        lo ^= dpp_from_prev(clo) ^ ((nlo << 1) | (nhi >> 31));
        hi ^= dpp_from_prev(chi) ^ ((nhi << 1) | (nlo >> 31));
        // rho
        const uint32_t sl = k.rswap ? hi : lo, sh = k.rswap ? lo : hi;
        const uint32_t rhi = __builtin_amdgcn_alignbit(sh, sl, k.rsh);
        const uint32_t rlo = __builtin_amdgcn_alignbit(sl, sh, k.rsh);
        // pi
        lo = (uint32_t)__builtin_amdgcn_ds_bpermute((int)k.src4, (int)rlo) & k.valid;
        hi = (uint32_t)__builtin_amdgcn_ds_bpermute((int)k.src4, (int)rhi) & k.valid;
        // chi, then iota
        const uint32_t b1lo = dpp_from_next(lo), b1hi = dpp_from_next(hi);
        const uint32_t b2lo = dpp_from_next2(lo), b2hi = dpp_from_next2(hi);
        lo ^= (~b1lo & b2lo) ^ (m00 & (uint32_t)kRC[round]);
        hi ^= (~b1hi & b2hi) ^ (m00 & (uint32_t)(kRC[round] >> 32));
        // the copies of columns 0 and 1 in slots 6 and 7 come from slots 1 and 2
        const uint32_t clo6 = dpp_from_prev5(lo), chi6 = dpp_from_prev5(hi);
        if (k.copy67) {
            lo = clo6;
            hi = chi6;
        }
    }
}

// Keccak-256 of nwords little-endian words at msg, computed by a whole wavefront (every
// lane must be active); every lane receives the digest.
__device__ inline void keccak256_wave(const uint32_t *msg, uint32_t nwords, uint32_t out[8])
{
    const KeccakWaveLane k = keccak_wave_lane();
    uint32_t lo = 0, hi = 0;
    const uint32_t nbytes = nwords * 4u;
    const uint32_t nblocks = nbytes / 136 + 1;
    for (uint32_t b = 0; b < nblocks; b++) {
        if (k.word >= 0) {
            const uint32_t idx = b * 34 + 2 * (uint32_t)k.word;
            uint32_t wl = idx < nwords ? msg[idx] : 0u;
            uint32_t wh = idx + 1 < nwords ? msg[idx + 1] : 0u;
            if (b == nblocks - 1) {
                if (idx == nwords) wl ^= 0x01u;
                if (idx + 1 == nwords) wh ^= 0x01u;
                if (k.word == 16) wh ^= 0x80000000u;
            }
            lo ^= wl;
            hi ^= wh;
        }
        keccak_f1600_wave(lo, hi, k);
    }
    // digest = words (0,0) .. (3,0) = lanes 1 .. 4
#pragma unroll
    for (int w = 0; w < 4; w++) {
        out[2 * w] = __builtin_amdgcn_readlane(lo, w + 1);
        out[2 * w + 1] = __builtin_amdgcn_readlane(hi, w + 1);
    }
}

// SHA-256 of nwords words by a whole wavefront.  The compression is a serial chain, so
// every lane runs it on the same data; the pointer is moved to vector registers first,
// which keeps the arithmetic on the VALU (faster here than the scalar unit, where the
// compiler would otherwise place uniform work).
__device__ inline void sha256_wave(const uint32_t *msg, uint32_t nwords, uint32_t out[8])
{
    uint64_t p = reinterpret_cast<uint64_t>(msg);
    asm volatile("" : "+v"(p));
    sha256_words(reinterpret_cast<const uint32_t *>(p), nwords, out);
}

// Transcript hash computed by a whole wavefront; every lane receives the digest.
__device__ __forceinline__ void hash_words_wave(int kind, const uint32_t *msg, uint32_t nwords, uint32_t out[8])
{
    if (kind == kHashKeccak256) keccak256_wave(msg, nwords, out);
    else sha256_wave(msg, nwords, out);
}

}  // namespace gkrx
