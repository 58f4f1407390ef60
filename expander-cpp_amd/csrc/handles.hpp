// The opaque handles of the C ABI (gkrx_ctx, gkrx_circuit), shared by capi.cpp and
// server.cpp.  Not part of the public interface.
#pragma once

#include <atomic>
#include <cstdint>
#include <mutex>
#include <vector>

#include "gkrx.h"
#include "internal.hpp"

// One stream of a BN254 batch (gkrx_prove_batch proves up to kBnBatchLanes BN254 proofs at
// once, each on a lane with its own tables and pinned staging of its results).
struct BnLane {
    hipStream_t st = nullptr;
    gkrx::BnScratch bn;
    uint8_t *tail_host = nullptr;    // pinned: the proof tail of the lane's current proof
    uint64_t tail_cap = 0;
    uint32_t *small_host = nullptr;  // pinned: claims (8 words), then the transcript state (10 words)
    int64_t item = -1;               // batch item in flight on the lane
};
constexpr uint32_t kBnBatchLanes = 4;

struct gkrx_ctx {
    gkrx_config cfg;
    int device = 0;
    uint32_t R = 3;
    uint32_t L = 8;       // lanes of the field element
    uint32_t W = 8;       // u32 words per element: L for M31, 8 for BN254
    bool bn254 = false;
    hipStream_t stream = nullptr;
    gkrx::DevScratch scratch;
    gkrx::BnScratch bn;
    gkrx::ProveTimings last;
    gkrx::Profiler prof;
    gkrx::BatchScratch batch;
    std::vector<BnLane> bn_lanes;  // created on the first BN254 batch
    bool graphs = false;  // gkrx_prove replays its device part as a captured HIP graph
    std::mutex mu;
    std::atomic<int> live_circuits{0};  // circuits (and servers) bound to this context
};

struct gkrx_circuit {
    gkrx_ctx *ctx = nullptr;
    uint32_t D = 0;
    std::vector<uint32_t> n, m;
    std::vector<std::vector<uint32_t>> mul, add;  // canonical coefficients
    gkrx::DevCircuit dev;
    std::vector<uint32_t> witness;
    bool has_witness = false;
    bool evaluated = false;
    std::atomic<bool> in_server{false};  // a server request on it has not completed
};

namespace gkrx {

// Proof bytes after the commitment: 32 grind bytes, then (6 n R + 2 R) S per layer
// (LinearGKR.hpp:58-90, sumcheck.hpp:72-103).
inline size_t tail_bytes(const gkrx_ctx *ctx, const gkrx_circuit *c)
{
    const size_t S = (size_t)ctx->W * 4;
    size_t t = 32;
    for (uint32_t l = 0; l < c->D; l++) t += ((size_t)6 * c->n[l] * ctx->R + 2 * ctx->R) * S;
    return t;
}

inline size_t commit_bytes(const gkrx_ctx *ctx, const gkrx_circuit *c)
{
    return ((size_t)1 << c->n[0]) * ctx->W * 4;
}

}  // namespace gkrx
