// Internal structures shared by the C-ABI layer (capi.cpp) and the device prover
// (prover_m31.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "gkrx.h"
#include "profiler.hpp"

namespace gkrx {

void set_error(const std::string &msg);

#define GKRX_HIP_CHECK(expr)                                                                  \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            ::gkrx::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));             \
            return GKRX_ERR_DEVICE;                                                           \
        }                                                                                     \
    } while (0)

constexpr uint32_t kAddMark = 0xffffffffu;  // "no second input" in a CSR entry
// Words per CSR entry.  The g_x and h_y entries carry their own row, so that a tile of rows
// can spread its entries evenly over a workgroup; 16 bytes load in one instruction.
constexpr uint32_t kFwWords = 3, kGxWords = 4, kHyWords = 4;

// Host CPUs this process may run on (its affinity mask; hardware_concurrency counts every
// online CPU, which can be many more than a container or a job scheduler grants).
inline unsigned usable_cpus()
{
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof set, &set) == 0) {
        const int c = CPU_COUNT(&set);
        if (c > 0) return (unsigned)c;
    }
    return std::max(1u, std::thread::hardware_concurrency());
}

// Runs fn(i) for i < n on up to max_threads host threads (the caller is one of them).
template <typename F>
void parallel_for(size_t n, unsigned max_threads, F &&fn)
{
    const unsigned hw = std::max(1u, std::min(max_threads, usable_cpus()));
    const unsigned nt = (unsigned)std::min<size_t>(hw, n);
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t i; (i = next++) < n;) fn(i);
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
}
constexpr int kMaxVars = 30;

// One layer in device form.  Every table is in the reference's natural index order
// (entry i of a multilinear table is the hypercube point with little-endian bits i).
struct LayerDev {
    uint32_t n = 0, m = 0;       // nb_input_vars, nb_output_vars
    uint64_t v_off = 0;          // u32 offset of the layer input values in the value buffer
    uint64_t out_off = 0;        // u32 offset of the layer output values
    // forward CSR by output index: entry {in0, in1 or kAddMark, coef} (kFwWords)
    uint32_t *fw_ptr = nullptr, *fw_ent = nullptr;
    // g_x CSR by first input x: entry {in1 or kAddMark, z, coef, x} (kGxWords)
    uint32_t *gx_ptr = nullptr, *gx_ent = nullptr;
    // h_y CSR by second input y (mul gates only): entry {x, z, coef, y} (kHyWords)
    uint32_t *hy_ptr = nullptr, *hy_ent = nullptr;
};

struct DevCircuit {
    uint32_t D = 0;
    uint32_t L = 0;                  // lanes
    std::vector<LayerDev> layers;
    uint32_t *csr_pool = nullptr;    // one allocation holding every CSR
    uint32_t *values = nullptr;      // all layer values, D + 1 tables of L lanes
    uint64_t values_u32 = 0;
    uint64_t csr_u32 = 0;            // size of csr_pool
    uint32_t max_n = 0, max_m = 0;
    LayerDev *d_layers = nullptr;    // device copy of `layers` (read by the batch kernel)
    uint64_t serial = 0;             // unique per upload (identifies the circuit to a cached graph)
};

// One proof of a batch: its circuit, the host transcript prefix, and where the proof
// bytes after the commitment go.
struct BatchItem {
    const DevCircuit *dc;
    const uint32_t *rz1;     // R * m_last output challenges
    const uint32_t *tstate;  // 10 words (length, hash start, digest)
    const uint8_t *grind;    // 32 grind bytes (the first bytes of the tail)
    uint64_t tail_bytes;
    uint8_t *tail_out;       // host destination of the tail
};

// Batched prover engines: the whole proof in one workgroup (one launch per batch), or
// the lockstep engine (one launch per step of every proof, each proof's step on as many
// workgroups as its tables need).
enum BatchEngine : int { kBatchPerProof = 1, kBatchLockstep = 2, kBatchLockstepGrouped = 4 };

// Device scratch of the batched prover, grown to the largest batch seen.
struct BatchScratch {
    int engine = kBatchLockstep;
    uint32_t block = 256;           // workgroup size of the lockstep work kernel: 64, 128, 256 or 512
    uint32_t lanes = 4;             // lockstep round pairs: lanes of an entry per thread (1, 2 or 4)
    // lockstep: run the serial chains and the round pairs on streams of their own, beside the
    // other wide ops (false: everything on the caller's stream)
    bool chain_stream = true;
    hipStream_t aux = nullptr, aux2 = nullptr;
    hipEvent_t ev_main = nullptr, ev_aux = nullptr, ev_aux2 = nullptr;
    uint32_t *pool = nullptr;
    uint64_t pool_u32 = 0;
    uint8_t *meta = nullptr;        // job table, order or step schedule
    uint32_t *queue = nullptr;      // per-proof engine: the work queue's head (next order slot)
    uint64_t meta_bytes = 0;
    uint8_t *host_meta = nullptr;   // pinned staging of the metadata and seeds
    uint64_t host_meta_bytes = 0;
    // what `meta` holds: the key of the batch it was built for (circuit addresses and
    // shapes, transcript positions, pool), its size, section offsets and the step records
    std::vector<uint64_t> meta_key;
    uint64_t meta_used = 0, o_order = 0, o_tasks = 0, o_starts = 0;
    std::vector<uint32_t> meta_steps;
    std::vector<uint32_t> meta_seed_step;   // lockstep: the step (of its group) whose work seeds job i
    std::vector<uint32_t> meta_seed_order;  // the jobs in the order of their seed slots
    std::vector<uint32_t> meta_need_order;  // the jobs in the order their seeds are needed
    std::vector<uint32_t> meta_group_off;   // lockstep: group g's steps are [off[g], off[g + 1])
    std::vector<uint32_t> meta_group_slot;  // group g's seed slots are [slot[g], slot[g + 1])
    uint32_t *claims_host = nullptr;  // claims (item-major, R * L each), then each item's transcript length
    uint64_t claims_cap = 0;
    uint64_t claims_words = 0;        // claim words of the last batch
    // pinned staging of the proof tails: the device copies land here (a copy into pageable
    // memory would block the launching thread), dev_batch_collect hands them out
    uint8_t *tails_host = nullptr;
    uint64_t tails_cap = 0;
    std::vector<uint64_t> tail_stage_off;
    // pinned host buffer for n claim words (nullptr when allocation fails)
    uint32_t *host_claims(uint64_t n)
    {
        if (n > claims_cap) {
            if (claims_host) (void)hipHostFree(claims_host);
            claims_host = nullptr;
            claims_cap = 0;
            if (hipHostMalloc(&claims_host, n * 4, hipHostMallocDefault) != hipSuccess) return nullptr;
            claims_cap = n;
        }
        return claims_host;
    }
};

// Device scratch of one prover context, grown to the largest circuit seen.
struct DevScratch {
    uint64_t bytes = 0;
    void *pool = nullptr;
    uint32_t *eqz = nullptr;      // R * maxM
    uint32_t *eqx = nullptr;      // R * maxN
    uint32_t *hg0 = nullptr;      // R * maxN * L    (phase-1 round-0 hg)
    uint32_t *hgA = nullptr;      // R * maxN/2 * L
    uint32_t *hgB = nullptr;      // R * maxN/4 * L
    uint32_t *fA = nullptr;       // R * maxN/2 * L
    uint32_t *fB = nullptr;       // R * maxN/4 * L
    uint32_t *s0 = nullptr;       // R * maxN        (phase-2 scalar h table)
    uint32_t *sA = nullptr;       // R * maxN/2
    uint32_t *sB = nullptr;       // R * maxN/4
    uint32_t *partials = nullptr; // R * kPairVals L * maxBlocks (block-major block partials)
    uint32_t *chal = nullptr;     // D * R * 2 * 32 (rx then ry per layer and rep) + rz1 seed
    uint32_t *ab = nullptr;       // D * 2 (alpha, beta after each layer)
    uint32_t *claims = nullptr;   // R * L
    uint32_t *vxy = nullptr;      // R * L (vx) + R * L (vy)
    uint32_t *tstate = nullptr;   // transcript state: len, hash_start, digest[8]
    uint32_t *tail = nullptr;     // transcript bytes after the commitment
    uint64_t tail_cap = 0;        // bytes
    uint32_t max_n = 0, max_m = 0, L = 0, D = 0;
    // pinned host staging for the small uploads of a prove (outlives the async copies)
    uint32_t *stage = nullptr;
    uint64_t stage_u32 = 0;
    // the device part of a prove captured as a HIP graph (the seed copies from `stage` and
    // every launch), replayed while the circuit, R, hash and transcript offsets are the same
    hipGraphExec_t graph = nullptr;
    uint64_t graph_serial = 0, graph_len = 0, graph_hs = 0;
    uint32_t graph_R = 0, graph_pre = 0;  // graph_pre: bytes of the tail prefix it copies
    int graph_kind = -1;
};

// Device scratch of the BN254 prover (one lane, one repetition; every table entry is one
// Fr of 8 u32 limbs in Montgomery form).
struct BnScratch {
    void *pool = nullptr;
    uint32_t *eqz = nullptr;      // maxM Fr
    uint32_t *eqx = nullptr;      // maxN Fr
    uint32_t *h0 = nullptr;       // maxN Fr: phase-1 hg, then the phase-2 table s
    uint32_t *hA = nullptr;       // maxN/2 Fr (tables of odd rounds)
    uint32_t *hB = nullptr;       // maxN/4 Fr (tables of even rounds >= 2)
    uint32_t *fA = nullptr;
    uint32_t *fB = nullptr;
    uint32_t *partials = nullptr; // [block][3] Fr
    uint32_t *eq_parts = nullptr; // the halves of two eq tables (2 x (2^ceil(m/2) + 2^floor(m/2)) Fr)
    uint32_t *chal = nullptr;     // (D + 1) slots of 64 Fr: rx then ry
    uint32_t *ab = nullptr;       // (D + 1) x (alpha, beta)
    uint32_t *claims = nullptr;   // canonical claimed_v
    uint32_t *vxy = nullptr;      // vx, vy
    uint32_t *tstate = nullptr;   // transcript state: len, hash_start, digest[8]
    uint32_t *tail = nullptr;     // transcript bytes after the commitment
    uint64_t tail_cap = 0;
    uint32_t max_n = 0, max_m = 0, D = 0;
    uint32_t *stage = nullptr;    // pinned staging of the seeds
    uint64_t stage_u32 = 0;
};

struct ProveTimings {
    double host_prefix_ms = 0, device_ms = 0, total_ms = 0;
};

// Device entry points implemented in prover_m31.hip
int dev_upload_circuit(DevCircuit &dc, uint32_t L, const std::vector<uint32_t> &n, const std::vector<uint32_t> &m,
                       const std::vector<std::vector<uint32_t>> &mul, const std::vector<std::vector<uint32_t>> &add);
void dev_free_circuit(DevCircuit &dc);
int dev_ensure_scratch(DevScratch &s, const DevCircuit &dc, uint32_t R, uint64_t tail_bytes);
// The allocation-free prove (gkrx_prove_noalloc): its prover tables live in a file-scope
// device array of kNoAllocScratchBytes.  dev_noalloc_init (once per device, with the device
// current; the context's creation calls it) looks up that array's device address.
constexpr uint64_t kNoAllocScratchBytes = 64ull << 20;
int dev_noalloc_init(int device);
// Bytes of prover tables a prove of dc needs (the allocation-free prove's size check).
uint64_t dev_noalloc_bytes(const DevCircuit &dc, uint32_t R, uint64_t tail_bytes);
// dev_prove on the device array: the seed goes in by a kernel argument, nothing is
// allocated or copied from host staging; view receives the table addresses (tail, claims,
// tstate) to copy back.  GKRX_ERR_ARG when dc needs more than kNoAllocScratchBytes.
int dev_prove_noalloc(const DevCircuit &dc, uint32_t R, int hash_kind, const uint32_t *rz1_host,
                      const uint32_t *tstate_host, const uint8_t *grind, uint64_t tail_bytes, int device,
                      hipStream_t st, DevScratch &view);
void dev_free_scratch(DevScratch &s);
void dev_drop_graph(DevScratch &s);  // releases the cached prove graph, if any
int dev_forward(const DevCircuit &dc, hipStream_t st, Profiler *pf);
// Runs the on-device part of the prove: output claims, every layer sumcheck and the
// transcript from the state in `tstate_host` (len, hash_start, digest) and the tail
// prefix already in s.tail.  rz1 (R * m_last) seeds the last layer.
// graph: run the device part as a captured HIP graph (captured on first use, replayed after);
// ignored while profiling.
int dev_prove(const DevCircuit &dc, DevScratch &s, uint32_t R, int hash_kind, const uint32_t *rz1_host,
              const uint32_t *tstate_host, const uint8_t *tail_prefix, uint32_t tail_prefix_len, hipStream_t st,
              Profiler *pf, bool graph = false);
// The transcript prefixes of a batch, computed while the batch is enqueued: start() hands
// over the order in which the items' prefixes are needed, wait(i) returns once item i's
// rz1, tstate and grind (which BatchItem points at) are written.
struct BatchPrefixes {
    virtual void start(const std::vector<uint32_t> &order) = 0;
    virtual void wait(uint32_t i) = 0;
    virtual ~BatchPrefixes() = default;
};

// Proves every item with the engine bs.engine selects.  Enqueues the copies of the
// tails, of the claims (item-major, R * L each, to bs.host_claims()) and of each item's
// final transcript length into pinned staging; once the stream has reached them,
// dev_batch_collect checks the lengths against the proof layout (GKRX_ERR_DEVICE if one
// differs) and copies each tail to items[i].tail_out.  With `pre`, the items' prefixes may still be in progress: each is
// awaited just before the first step that needs it (the lockstep engine seeds a proof's
// transcript after its forward pass), otherwise they must be ready.
int dev_prove_batch(const std::vector<BatchItem> &items, uint32_t R, int hash_kind, BatchScratch &bs, hipStream_t st,
                    Profiler *pf, BatchPrefixes *pre = nullptr);
int dev_batch_collect(const std::vector<BatchItem> &items, const BatchScratch &bs);
// Device bytes the batch prover takes for one proof of dc (its tables and transcript).
uint64_t dev_batch_job_bytes(const DevCircuit &dc, uint32_t R, uint64_t tail_bytes);
// Device bytes of a circuit's resident values and wiring.
inline uint64_t dev_circuit_bytes(const DevCircuit &dc) { return 4 * (dc.values_u32 + dc.csr_u32); }
// Whether a circuit fits the batch kernel (every layer at most 2^20 wide).
int dev_batch_supported(const DevCircuit &dc);
void dev_free_batch(BatchScratch &bs);

// CIRCUIT2 binary circuit -> per-layer gate lists (circuit2.cpp); mul entries
// (in0, in1, out, coef), add entries (in0, out, coef), layer 0 first.
int circuit2_flatten(const char *path, int field_type, int coef_mode, std::vector<std::vector<uint32_t>> &mul,
                     std::vector<std::vector<uint32_t>> &add);

// BN254 entry points (prover_bn254.hip).  A BN254 DevCircuit is uploaded with L = 8: one
// entry is 8 u32 words.
int dev_bn_to_mont(uint32_t *values, uint64_t n, hipStream_t st);
int dev_bn_forward(const DevCircuit &dc, hipStream_t st, Profiler *pf);
int dev_bn_ensure_scratch(BnScratch &s, const DevCircuit &dc, uint64_t tail_bytes);
void dev_bn_free_scratch(BnScratch &s);
uint64_t dev_bn_scratch_bytes(const DevCircuit &dc, uint64_t tail_bytes);
// rz1_host: m_last canonical challenges (8 words each); claims land in s.claims.
int dev_bn_prove(const DevCircuit &dc, BnScratch &s, int hash_kind, const uint32_t *rz1_host,
                 const uint32_t *tstate_host, const uint8_t *tail_prefix, uint32_t tail_prefix_len, hipStream_t st,
                 Profiler *pf);

}  // namespace gkrx
