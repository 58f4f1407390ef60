/*
 * CPU oracle for the GKR prove path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference prover (PolyhedraZK/Expander-cpp,
 * header-only C++).  It exists so that tests, the smoke check and the CPU-baseline
 * leg of the benchmark have an independent checker.  Nothing in the product
 * (libgkrx) links, loads or calls it.
 *
 * Every function cites the reference file:line it follows.  Pinning (what fixes
 * each function independently of this restatement):
 *   - m31_*            : Python-int closed forms and brute force (tests/test_oracle_field.py),
 *                        tests/ff.cpp properties.
 *   - oracle_sha256    : FIPS 180-4 known answers + hashlib.sha256 + the vectors of
 *                        lib/btc_sha256/crypto/sha256.cpp:489-575.
 *   - oracle_keccak256 : published Keccak-256 known answers; the permutation is also
 *                        pinned against hashlib.sha3_256 (same sponge, pad byte 0x06).
 *   - transcript/prove : byte-identical proofs against the reference prover built from
 *                        its own sources (oracle/Makefile -> oracle/_ref), golden
 *                        fixtures in tests/golden/, and the reference verifier.
 *   - verify           : accepts every reference proof, rejects tampered ones.
 *   - BN254 parts      : the reference has no BN254 field (parity unpinned against the
 *                        reference); pinned by Python big-int closed forms only.
 */
#ifndef GKRX_ORACLE_H
#define GKRX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_P31 0x7fffffffu

enum { ORACLE_HASH_SHA256 = 0, ORACLE_HASH_KECCAK256 = 1 };

/* --- M31 scalar field (include/field/M31.hpp:35-142) --------------------------- */
uint32_t oracle_m31_add(uint32_t a, uint32_t b);
uint32_t oracle_m31_sub(uint32_t a, uint32_t b);
uint32_t oracle_m31_neg(uint32_t a);
uint32_t oracle_m31_mul(uint32_t a, uint32_t b);
uint32_t oracle_m31_from_u32(uint32_t v);          /* M31(uint32) ctor: one fold only */
uint32_t oracle_m31_from_bytes(const uint8_t *in); /* full reduction */

/* --- hashes ------------------------------------------------------------------- */
void oracle_sha256(const uint8_t *in, size_t len, uint8_t out[32]);
/* Raw compression of nblocks 64-byte blocks into state (no padding), the unit the
 * reference's SHA-256 self-test pins (lib/btc_sha256/crypto/sha256.cpp:489-575). */
void oracle_sha256_blocks(uint32_t state[8], const uint8_t *data, size_t nblocks);
/* Keccak sponge with rate 136 and a caller-chosen domain pad byte:
 * pad 0x01 = Keccak-256, pad 0x06 = SHA3-256. */
void oracle_keccak256_pad(const uint8_t *in, size_t len, uint8_t pad, uint8_t out[32]);
void oracle_keccak256(const uint8_t *in, size_t len, uint8_t out[32]);

/* --- transcript (include/fiat_shamir/transcript.hpp:87-175) ------------------ */
void *oracle_transcript_new(int hash_kind);
void oracle_transcript_append(void *t, const uint8_t *bytes, size_t len);
uint32_t oracle_transcript_challenge(void *t);
void oracle_transcript_free(void *t);

/* --- circuits ----------------------------------------------------------------- */
typedef struct {
    uint32_t nb_layers;
    /* layer 0 is the input layer, as in the reference's Circuit::layers */
    uint32_t *n_mul, *n_add;
    uint32_t **mul; /* per layer: n_mul * (in0, in1, out, coef) */
    uint32_t **add; /* per layer: n_add * (in0, out, coef) */
    uint32_t *nb_in_vars, *nb_out_vars;
} oracle_circuit;

int oracle_circuit_load_text(const char *mul_path, const char *add_path, oracle_circuit **out);
void oracle_circuit_free(oracle_circuit *c);

/* Forward pass (include/circuit/circuit.hpp:92-105, 273-280).
 * witness: 2^n_in(0) * L lanes; values: caller buffer receiving every layer's
 * input values followed by the last layer's output, element-major/lane-minor.
 * Returns the number of u32 written, or needed when values == NULL. */
size_t oracle_evaluate(const oracle_circuit *c, uint32_t L, const uint32_t *witness, uint32_t *values);

/* eq table (include/LinearGKR/sumcheck_common.hpp:16-53) */
void oracle_eq_evals(const uint32_t *r, uint32_t n, uint32_t mul_factor, uint32_t *out);
/* eval_multilinear (include/poly_commit/poly.hpp:88-104), L lanes */
void oracle_eval_multilinear(const uint32_t *evals, uint32_t L, const uint32_t *r, uint32_t nvars,
                             uint32_t *out);

/* Prover::prove (include/LinearGKR/LinearGKR.hpp:58-90).  *proof is malloc'ed.
 * claimed: R * L u32.  Returns 0 on success. */
int oracle_prove(const oracle_circuit *c, uint32_t L, int hash_kind, const uint32_t *witness,
                 uint8_t **proof, size_t *proof_len, uint32_t *claimed);
/* Verifier::verify (include/LinearGKR/LinearGKR.hpp:104-149). 1 = accept, 0 = reject,
 * negative = malformed input. */
int oracle_verify(const oracle_circuit *c, uint32_t L, int hash_kind, const uint32_t *claimed,
                  const uint8_t *proof, size_t proof_len);
/* Number of soundness repetitions for M31 (include/configuration/config.hpp:61-65). */
uint32_t oracle_num_repetitions(void);

void oracle_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
