/*
 * gkrx -- C ABI of the MI355X GKR prover.
 *
 * The reference (PolyhedraZK/Expander-cpp) is header-only C++ with no ABI.  Each entry
 * point below replaces one piece of its C++ surface; the citation names the reference
 * interface it stands in for (paths relative to the reference root).  The C++ drop-in
 * classes (gkr::Config, gkr::Circuit, gkr::Prover, gkr::Verifier, gkr::Proof) in
 * include/gkrx/expander.hpp are thin wrappers over these calls.
 *
 * Conventions: plain pointers and sizes; every call returns an int status (GKRX_OK = 0,
 * negative on error) unless noted; gkrx_last_error() gives a message for the calling
 * thread.  Field elements of M31 with L lanes are L little-endian u32 values (canonical,
 * < 2^31 - 1), lane 0 first -- the bytes F::to_bytes writes (include/field/M31.hpp:420-426).
 */
#ifndef GKRX_H
#define GKRX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GKRX_OK 0
#define GKRX_ERR_ARG (-1)     /* bad argument or misuse (the reference asserts) */
#define GKRX_ERR_CONFIG (-2)  /* unsupported Config (reference: assert in Prover ctor, LinearGKR.hpp:42-47) */
#define GKRX_ERR_CIRCUIT (-3) /* malformed circuit file or unsupported shape */
#define GKRX_ERR_IO (-4)      /* file could not be opened */
#define GKRX_ERR_NOMEM (-5)   /* host or device allocation failed */
#define GKRX_ERR_DEVICE (-6)  /* HIP runtime error or no usable GPU */
#define GKRX_ERR_STATE (-7)   /* call order (e.g. prove before a witness is set) */

/* Enumerations mirror include/configuration/config.hpp:5-24. */
enum { GKRX_PC_RAW = 0, GKRX_PC_KZG = 1, GKRX_PC_ORION = 2, GKRX_PC_FRI = 3 };
enum { GKRX_FIELD_M31 = 0, GKRX_FIELD_BABYBEAR = 1, GKRX_FIELD_BN254 = 2 };
enum { GKRX_FS_SHA256 = 0, GKRX_FS_KECCAK256 = 1, GKRX_FS_POSEIDON = 2, GKRX_FS_ANIMOE = 3, GKRX_FS_MIMC7 = 4 };

/* Config fields (include/configuration/config.hpp:26-86).  nb_parallel is the lane
 * count L of the data-parallel M31 field (8 or 16). */
typedef struct gkrx_config {
    int32_t field_type;
    int32_t fs_hash;
    int32_t pc_type;
    int32_t security_bits;
    int32_t grinding_bits;
    int32_t nb_parallel;
} gkrx_config;

typedef struct gkrx_ctx gkrx_ctx;
typedef struct gkrx_circuit gkrx_circuit;

/* Config::Config (config.hpp:71-81): M31, SHA-256, Raw, 100-bit security, 10 grinding
 * bits; nb_parallel = 8, the lane count the shipped x86 reference uses (M31.hpp:234-238). */
void gkrx_config_default(gkrx_config *cfg);
/* Config::get_num_repetitions (config.hpp:61-65, 82-85). */
int gkrx_config_num_repetitions(const gkrx_config *cfg);

/* Prover::Prover(const Config&) (include/LinearGKR/LinearGKR.hpp:42-47) plus the device
 * resources the prover needs: one context per GPU.  device = HIP device ordinal. */
int gkrx_ctx_create(const gkrx_config *cfg, int device, gkrx_ctx **out);
/* Fails with GKRX_ERR_STATE while circuits created on the context are still alive. */
int gkrx_ctx_destroy(gkrx_ctx *ctx);

/* Circuit::load_extracted_gates (include/circuit/circuit.hpp:152-215): the text
 * format (mul file one line per layer, output layer first). */
int gkrx_circuit_load_text(gkrx_ctx *ctx, const char *mul_path, const char *add_path, gkrx_circuit **out);
/* The same reader and checks with no context and no device: reads and validates the two
 * files on the host and reports the shape (Circuit::_compute_nb_vars, circuit.hpp:124-149),
 * from which proof size and device memory follow.  *nb_layers: capacity of nb_in_vars /
 * nb_out_vars on entry (either may be NULL), the layer count on return.  field_type is
 * GKRX_FIELD_M31 (coefficients must fold into the field) or GKRX_FIELD_BN254. */
int gkrx_circuit_parse_text(const char *mul_path, const char *add_path, int field_type, uint32_t *nb_layers,
                            uint32_t *nb_in_vars, uint32_t *nb_out_vars);
/* Circuit built from gate arrays, layer 0 = input layer as in Circuit::layers.
 * mul[l]: n_mul[l] x (in0, in1, out, coef); add[l]: n_add[l] x (in0, out, coef).
 * Variable counts follow Circuit::_compute_nb_vars (circuit.hpp:124-149). */
int gkrx_circuit_from_arrays(gkrx_ctx *ctx, uint32_t nb_layers, const uint32_t *n_mul, const uint32_t *const *mul,
                             const uint32_t *n_add, const uint32_t *const *add, gkrx_circuit **out);
/* CircuitLayer::nb_input_vars / nb_output_vars, per layer (arrays of nb_layers). */
int gkrx_circuit_shape(const gkrx_circuit *c, uint32_t *nb_layers, uint32_t *nb_in_vars, uint32_t *nb_out_vars);
/* The CIRCUIT2 binary format: CircuitRaw operator>> (include/circuit/circuit_raw.hpp:539-592)
 * followed by Circuit::from_circuit_raw (include/circuit/circuit.hpp:217-251), which flattens
 * the segment tree into layers and drops constant gates.  coef_mode 0 reproduces the
 * reference, whose element reader sets every coefficient read from such a file to 1
 * (include/field/M31.hpp:124-135); coef_mode 1 reads each 32-byte coefficient as a
 * little-endian integer reduced into the field. */
int gkrx_circuit_load_circuit2(gkrx_ctx *ctx, const char *path, int coef_mode, gkrx_circuit **out);
/* The same flattening with no context and no device: writes the circuit in the text format
 * (Circuit::load_extracted_gates, circuit.hpp:152-215). */
int gkrx_circuit2_to_text(const char *path, int field_type, int coef_mode, const char *mul_path,
                          const char *add_path);
/* Writes the circuit in the text format read by load_extracted_gates. */
int gkrx_circuit_write_text(const gkrx_circuit *c, const char *mul_path, const char *add_path);
/* layers[0].input_layer_vals.evals = witness: 2^nb_input_vars(0) elements x L lanes (host memory). */
int gkrx_circuit_set_witness(gkrx_circuit *c, const uint32_t *lanes, size_t n_u32);
/* Circuit::evaluate (circuit.hpp:273-280), on the device; every layer stays in HBM. */
int gkrx_circuit_evaluate(gkrx_circuit *c);
/* layers.back().output_layer_vals.evals (2^nb_output_vars(D-1) x L u32). */
int gkrx_circuit_get_output(gkrx_circuit *c, uint32_t *out, size_t n_u32);
/* Input values of layer l (2^nb_input_vars(l) x L u32); l == nb_layers gives the output. */
int gkrx_circuit_get_layer_values(gkrx_circuit *c, uint32_t layer, uint32_t *out, size_t n_u32);
int gkrx_circuit_destroy(gkrx_circuit *c);

/* Bytes of the proof Prover::prove returns for this circuit. */
size_t gkrx_proof_size(const gkrx_ctx *ctx, const gkrx_circuit *c);
/* Prover::prove (LinearGKR.hpp:58-90): evaluates the circuit if needed and proves it.
 * proof: caller buffer of at least gkrx_proof_size bytes; *proof_len receives the
 * length.  claimed_v: R * L u32 (R = gkrx_config_num_repetitions), the first element
 * of the tuple prove returns. */
int gkrx_prove(gkrx_ctx *ctx, gkrx_circuit *c, uint8_t *proof, size_t proof_cap, size_t *proof_len,
               uint32_t *claimed_v);
/* gkrx_prove without allocating: no counterpart in the reference.  For callers that cannot
 * allocate device memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc -- this variant calls no hipMalloc, hipFree, hipHostMalloc or other
 * allocating routine: the prover tables live in a file-scope device array of 64 MiB (each
 * GPU the library runs on holds it until the process ends; the context's creation looks up
 * its address), and the transcript seed goes to the device as a kernel argument.  Same
 * arguments, proof bytes and claims as gkrx_prove (M31 only).  Circuits whose tables need
 * more than 64 MiB (about 2^16 rows per layer at 16 lanes) fail with GKRX_ERR_ARG.  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (from two contexts)
 * race; calls on one context are serialized by it. */
int gkrx_prove_noalloc(gkrx_ctx *ctx, gkrx_circuit *c, uint8_t *proof, size_t proof_cap, size_t *proof_len,
                       uint32_t *claimed_v);
/* n independent proofs at once (no reference counterpart: the reference proves one
 * proof per thread, benchmarks/bench_keccak.cpp:32-42).  Proof i is byte-identical to
 * gkrx_prove on circuits[i].  The transcript prefixes run on a pool of host threads, then
 * the device proves every circuit with the context's batch engine (circuits must be at
 * most 2^20 wide per layer).  proofs[i] holds proof_caps[i] >= gkrx_proof_size bytes;
 * claimed_v[i] receives R * L u32. */
int gkrx_prove_batch(gkrx_ctx *ctx, uint32_t n, gkrx_circuit *const *circuits, uint8_t *const *proofs,
                     const size_t *proof_caps, size_t *proof_lens, uint32_t *const *claimed_v);

/* Batch engines (no reference counterpart).  PER_PROOF: one launch, one workgroup runs a
 * whole proof.  LOCKSTEP (the default): the batch advances one op of every proof per
 * step; a step is one launch of the wide ops (each proof's op on as many workgroups of
 * `block` threads as its tables need), one of the round pairs, one of the serial chains
 * and the launches of the transcript ends, the three kinds on three streams.
 * LOCKSTEP_GROUPS: the same steps for three groups of proofs, each group on a stream of
 * its own, so that one group's transcript ends and chains overlap the other groups'
 * passes.  Batches of more than 2048 proofs run on the PER_PROOF engine.  block is 64,
 * 128, 256 or 512.  Results do not depend on the choice. */
#define GKRX_BATCH_PER_PROOF 1
#define GKRX_BATCH_LOCKSTEP 2
#define GKRX_BATCH_LOCKSTEP_GROUPS 4
int gkrx_ctx_set_batch_engine(gkrx_ctx *ctx, int engine, uint32_t block);
int gkrx_ctx_get_batch_engine(const gkrx_ctx *ctx, int *engine, uint32_t *block);
/* HIP graphs for gkrx_prove (0 off, the default; 1 on): the device part of a prove (the
 * seed copies and every launch) is captured once per circuit and replayed on later proves of
 * the same circuit (new witnesses); the proof bytes are the same either way.  Ignored while
 * profiling.  No counterpart in the reference (its prover runs on the CPU,
 * include/LinearGKR/LinearGKR.hpp:58-90). */
int gkrx_ctx_set_graphs(gkrx_ctx *ctx, int on);
int gkrx_ctx_get_graphs(const gkrx_ctx *ctx, int *on);
/* Lanes of a table entry per thread in the lockstep engine's round-pair kernel: 1, 2 or 4
 * (one 4-, 8- or 16-byte access per thread; L / lanes threads per entry).  Default 4. */
int gkrx_ctx_set_batch_lanes(gkrx_ctx *ctx, uint32_t lanes);
int gkrx_ctx_get_batch_lanes(const gkrx_ctx *ctx, uint32_t *lanes);
/* Autotuner: times a batched prove of these circuits (witnesses set) as gkrx_prove_batch
 * runs it -- the host transcript prefixes on host threads while the device works -- under
 * every engine, block size and lanes per thread (best of two timed runs after a warmup),
 * keeps the fastest in the context -- among settings within 1% of the fastest, the first
 * in the order below -- and reports it (lanes is 4 for PER_PROOF).  ms_each (optional, 25
 * entries): PER_PROOF, then LOCKSTEP with block 64, 128, 256, 512, each with lanes 1, 2, 4
 * (block-major), then LOCKSTEP_GROUPS in the same order. */
int gkrx_batch_autotune(gkrx_ctx *ctx, uint32_t n, gkrx_circuit *const *circuits, int *engine, uint32_t *block,
                        uint32_t *lanes, double *ms_each);
/* Verifier::verify (LinearGKR.hpp:104-149), on the host.  Returns 1 = accepted,
 * 0 = rejected, negative = error. */
int gkrx_verify(gkrx_ctx *ctx, const gkrx_circuit *c, const uint32_t *claimed_v, const uint8_t *proof,
                size_t proof_len);
/* The same Verifier::verify with no context and no device: the circuit comes as gate
 * arrays (as in gkrx_circuit_from_arrays) and the check runs on the calling host thread,
 * as the reference verifier does.  Returns 1 accepted, 0 rejected, < 0 on error. */
int gkrx_verify_arrays(const gkrx_config *cfg, uint32_t nb_layers, const uint32_t *n_mul, const uint32_t *const *mul,
                       const uint32_t *n_add, const uint32_t *const *add, const uint32_t *claimed_v,
                       const uint8_t *proof, size_t proof_len);

/* ---- Proof files: the layout Proof::write_and_compress writes before compressing and
 * Proof::read_proof reads (include/fiat_shamir/transcript.hpp:54-85): a little-endian u32
 * byte count, then the proof bytes.  gkrx_proof_read with proof = NULL reports the size in
 * *proof_len.  (The reference then runs `tar -czf` on the file; compression is left to the
 * caller.) */
int gkrx_proof_write(const char *path, const uint8_t *proof, size_t proof_len);
int gkrx_proof_read(const char *path, uint8_t *proof, size_t proof_cap, size_t *proof_len);

/* ---- Serving.  No reference counterpart (the reference proves one circuit per
 * Prover::prove call): a continuous-batching server for callers that prove a stream of
 * requests.  Requests queue in arrival order and are admitted while the HBM they need
 * (the circuit's resident values and wiring plus the prover tables, computed from the
 * circuit shape) fits hbm_budget_bytes; everything admitted is launched as one batch on
 * the first free slot (a stream with its own scratch; `slots` <= 16), so requests that
 * arrive while a batch runs start on the next free slot and overlap with it on the GPU.
 * Each proof is byte-identical to gkrx_prove of the same circuit. */
typedef struct gkrx_server gkrx_server;
int gkrx_server_create(gkrx_ctx *ctx, uint64_t hbm_budget_bytes, uint32_t slots, gkrx_server **out);
/* Queues one prove of c (witness set).  proof / claimed_v must stay valid, and c must not
 * be proved, changed or destroyed, until the request completes. */
int gkrx_server_submit(gkrx_server *s, gkrx_circuit *c, uint8_t *proof, size_t proof_cap, uint32_t *claimed_v,
                       uint64_t *ticket);
/* Waits up to timeout_us (< 0: without limit) for a request: GKRX_OK when it completed
 * (proof_len and latency_ms = submit to completion are set; the ticket is released),
 * 1 if it is still pending, < 0 if it failed. */
int gkrx_server_wait(gkrx_server *s, uint64_t ticket, int64_t timeout_us, size_t *proof_len, double *latency_ms);
/* gkrx_server_wait with the request's time split (ms): [0] waiting for a prefix thread,
 * [1] host transcript prefix, [2] waiting for admission and a free slot, [3] on the GPU
 * (launch until the slot's completion is seen), [4] results handed out, [5] total
 * (submit to completion, the latency gkrx_server_wait reports). */
int gkrx_server_wait_split(gkrx_server *s, uint64_t ticket, int64_t timeout_us, size_t *proof_len,
                           double split_ms[6]);
/* out: [0] submitted, [1] completed, [2] batches launched, [3] largest batch,
 * [4] peak admitted bytes, [5] admitted bytes now. */
int gkrx_server_stats(gkrx_server *s, uint64_t out[6]);
/* Finishes every queued request, then frees the server. */
int gkrx_server_destroy(gkrx_server *s);

/* Millisecond split of the last gkrx_prove or M31 gkrx_prove_batch on this context:
 * [0] host transcript prefix (commitment hash + grinding + output challenges), [1] device
 * work, [2] total.  gkrx_prove runs them in turn.  gkrx_prove_batch overlaps them: [0] is
 * the wall time of the prefix threads, [1] the time from the first launch to the proofs
 * on the host (it includes waits for prefixes the device needs before they are done). */
int gkrx_last_timing(const gkrx_ctx *ctx, double *ms3);

/* Per-kernel-class timing with HIP events on the context's stream (benchmark support; no
 * reference counterpart -- the reference's Timing class, sumcheck_helper.hpp:18-46,
 * prints wall-clock phases).  enable = 1 resets the counters and starts recording;
 * gkrx_ctx_kernel_stat reads class idx (0 .. classes-1; GKRX_ERR_ARG past the end). */
int gkrx_ctx_profile(gkrx_ctx *ctx, int enable);
/* The same, timing only the classes whose bit is set in mask (bit idx = class idx; 0 turns
 * timing off).  Every timed launch adds two event packets to the stream, so a benchmark
 * that times one kernel inside its timed loop selects just that class. */
int gkrx_ctx_profile_classes(gkrx_ctx *ctx, uint64_t mask);
int gkrx_ctx_kernel_stat(gkrx_ctx *ctx, int idx, const char **name, uint64_t *launches, double *ms);

/* Thread-local message of the last failed call. */
const char *gkrx_last_error(void);

/* Synthetic benchmark inputs (no reference counterpart: the reference draws wiring from
 * SparseCircuitConnection::random, circuit.hpp:40-65, and witnesses from rand(),
 * circuit.hpp:287-295).  One gate per output index of a 2^log_width layer; add with
 * probability add_permille / 1000; coefficient 1; with locality, 80% of input indices
 * fall within +-1024 of the output index; the largest input index is forced to
 * 2^log_width - 1.  mul_out holds up to 2^log_width x 4 u32, add_out up to 2^log_width x 3. */
int gkrx_synth_layer(uint64_t seed, uint32_t layer, uint32_t log_width, uint32_t add_permille, uint32_t locality,
                     uint32_t *mul_out, uint32_t *n_mul, uint32_t *add_out, uint32_t *n_add);
/* n_elems x lanes M31 values, uniform in [0, p) by rejection. */
int gkrx_synth_witness_m31(uint64_t seed, uint64_t n_elems, uint32_t lanes, uint32_t *out);
/* n_elems BN254 scalars (8 LE u32 words each), uniform in [0, r): four 64-bit draws
 * masked to 254 bits, rejected while >= r.  No reference counterpart (the reference has
 * no BN254 field). */
int gkrx_synth_witness_bn254(uint64_t seed, uint64_t n_elems, uint32_t *out);

/* Transcript hashes of the library (host implementations), for tests. */
int gkrx_hash(int fs_hash, const uint8_t *in, size_t len, uint8_t out[32]);
/* n messages of one length hashed at once (Keccak-256: eight sponges per core in AVX-512
 * lanes where the CPU has them -- the path the batched prover's transcript prefixes
 * take); out receives 32 bytes per message.  Returns 1 when the eight-lane path ran. */
int gkrx_hash_many(int fs_hash, uint32_t n, const uint8_t *const *in, size_t len, uint8_t *out);

#ifdef __cplusplus
}
#endif

#endif /* GKRX_H */
