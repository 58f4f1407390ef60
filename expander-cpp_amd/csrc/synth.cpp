// Seeded synthetic layered circuits and witnesses.
//
// The reference has no generator of its own (its tests draw wiring from
// SparseCircuitConnection::random, include/circuit/circuit.hpp:40-65, and witnesses from
// rand(), circuit.hpp:287-295).  These are the build's benchmark shapes: one gate per
// output index, add with a given probability, coefficient 1, optional locality wiring,
// the largest input index forced to 2^k - 1 so the reference's _compute_nb_vars
// (circuit.hpp:124-149) sees exactly k input variables.
#include <cstdint>
#include <cstring>

#include "gkrx.h"

namespace {

struct Xoshiro256ss {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x)
    {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    explicit Xoshiro256ss(uint64_t seed)
    {
        for (int i = 0; i < 4; i++) s[i] = splitmix(seed);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    uint32_t hi32() { return (uint32_t)(next() >> 32); }
};

constexpr uint64_t kLayerMix = 0xd1b54a32d192ed03ULL;
constexpr uint64_t kWitnessMix = 0x5eed5eed5eed5eedULL;

}  // namespace

extern "C" int gkrx_synth_layer(uint64_t seed, uint32_t layer, uint32_t log_width, uint32_t add_permille,
                                uint32_t locality, uint32_t *mul_out, uint32_t *n_mul, uint32_t *add_out,
                                uint32_t *n_add)
{
    if (log_width < 1 || log_width > 30 || add_permille > 1000 || !mul_out || !add_out || !n_mul || !n_add)
        return GKRX_ERR_ARG;
    Xoshiro256ss rng(seed ^ ((uint64_t)(layer + 1) * kLayerMix));
    const uint32_t W = 1u << log_width;
    const uint64_t add_thresh = ((uint64_t)add_permille << 32) / 1000;
    const uint64_t local_thresh = (8ULL << 32) / 10;
    auto pick = [&](uint32_t o) -> uint32_t {
        if (locality && rng.hi32() < local_thresh) {
            int64_t off = (int64_t)(rng.hi32() % 2049u) - 1024;
            int64_t idx = (int64_t)o + off;
            if (idx < 0) idx = 0;
            if (idx > (int64_t)W - 1) idx = W - 1;
            return (uint32_t)idx;
        }
        return (uint32_t)(rng.next() >> (64 - log_width));
    };
    uint32_t nm = 0, na = 0;
    for (uint32_t o = 0; o < W; o++) {
        bool is_add = rng.hi32() < add_thresh;
        uint32_t i0 = pick(o);
        uint32_t i1 = is_add ? 0 : pick(o);
        if (o == W - 1) i0 = W - 1;
        if (is_add) {
            add_out[3 * na] = i0;
            add_out[3 * na + 1] = o;
            add_out[3 * na + 2] = 1;
            na++;
        } else {
            mul_out[4 * nm] = i0;
            mul_out[4 * nm + 1] = i1;
            mul_out[4 * nm + 2] = o;
            mul_out[4 * nm + 3] = 1;
            nm++;
        }
    }
    *n_mul = nm;
    *n_add = na;
    return GKRX_OK;
}

// BN254 scalars: four 64-bit draws masked to 254 bits (little-endian limbs), rejected
// while >= r; each element is 8 little-endian u32 words.
extern "C" int gkrx_synth_witness_bn254(uint64_t seed, uint64_t n_elems, uint32_t *out)
{
    if (!out) return GKRX_ERR_ARG;
    static const uint64_t kR[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL,
                                   0x30644e72e131a029ULL};
    Xoshiro256ss rng(seed ^ kWitnessMix);
    for (uint64_t i = 0; i < n_elems; i++) {
        uint64_t v[4];
        for (;;) {
            for (int k = 0; k < 4; k++) v[k] = rng.next();
            v[3] &= (1ULL << 62) - 1;
            bool below = false;
            for (int k = 3; k >= 0; k--) {
                if (v[k] != kR[k]) {
                    below = v[k] < kR[k];
                    break;
                }
            }
            if (below) break;
        }
        for (int k = 0; k < 4; k++) {
            out[8 * i + 2 * k] = (uint32_t)v[k];
            out[8 * i + 2 * k + 1] = (uint32_t)(v[k] >> 32);
        }
    }
    return GKRX_OK;
}

extern "C" int gkrx_synth_witness_m31(uint64_t seed, uint64_t n_elems, uint32_t lanes, uint32_t *out)
{
    if (!out || lanes == 0) return GKRX_ERR_ARG;
    Xoshiro256ss rng(seed ^ kWitnessMix);
    const uint64_t total = n_elems * lanes;
    for (uint64_t i = 0; i < total; i++) {
        uint32_t x;
        do {
            x = rng.hi32() & 0x7fffffffu;
        } while (x == 0x7fffffffu);
        out[i] = x;
    }
    return GKRX_OK;
}
