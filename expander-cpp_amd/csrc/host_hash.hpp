// Host transcript hashes and the host part of the transcript (see host_hash.cpp).
#pragma once

#include <cstddef>
#include <cstdint>

namespace gkrx {

void host_sha256(const uint8_t *in, size_t len, uint8_t out[32]);
void host_keccak256(const uint8_t *in, size_t len, uint8_t out[32]);
// kind: 0 = SHA-256, 1 = Keccak-256; out may alias in.
void host_hash(int kind, const uint8_t *in, size_t len, uint8_t out[32]);
bool host_sha_extensions();

// Prover::prove up to the first layer sumcheck (include/LinearGKR/LinearGKR.hpp:62-69,
// grind :13-32, include/LinearGKR/gkr.hpp:42-49): hash of the commitment, grinding,
// then R * mlast output challenges rz1[j * mlast + i].  tstate receives the transcript
// state relative to the proof tail that begins with the 32 grind bytes:
// [0] bytes in the tail, [1] start of the unhashed suffix, [2..9] digest.
void host_transcript_prefix(int kind, const uint8_t *commit, size_t commit_len, uint32_t R, uint32_t mlast,
                            uint32_t grinding_bits, uint8_t grind_out[32], uint32_t *rz1, uint32_t tstate[10]);
// Keccak-256 transcript prefixes of n <= 8 proofs whose commitments have the same length,
// at once (eight sponges in the AVX-512 lanes of one core when host_keccak_x8()); the
// results equal host_transcript_prefix(1, ...) of each.
bool host_keccak_x8();
void host_transcript_prefix_x8(uint32_t n, const uint8_t *const *commit, size_t commit_len, uint32_t R,
                               const uint32_t *mlast, uint32_t grinding_bits, uint8_t (*grind_out)[32],
                               uint32_t *const *rz1, uint32_t (*tstate)[10]);
// Keccak-256 of n messages of one length (eight at a time where the CPU allows);
// out receives 32 bytes per message.
void host_keccak256_many(uint32_t n, const uint8_t *const *in, size_t len, uint8_t *out);
// The same over the BN254 scalar field (one repetition): grinding draws 256 / 254 = 1
// challenge, written as its 32 canonical bytes; rz1 receives mlast challenges as canonical
// values, 8 LE u32 words each.
void host_transcript_prefix_bn254(int kind, const uint8_t *commit, size_t commit_len, uint32_t mlast,
                                  uint32_t grinding_bits, uint8_t grind_out[32], uint32_t *rz1, uint32_t tstate[10]);

}  // namespace gkrx
